// Batched cluster bootstrap refits of the two-way (worker, firm) FE regression
// on the wide solver's launch path.
//
// A cluster (worker) resample is a vector of multiplicities m_w; the refit is
// the least-squares fit of y on X with worker and firm effects, each row
// weighted by its worker's multiplicity.  Within a worker the weight is
// constant, so the worker projection is the plain worker mean, and with
// u = v - (worker mean of v) the firm effects of column v solve
//     S_m psi = sum over matches (w, f) of m_w s_(w,f) at firm f,
//     S_m = sum_w m_w L_w   (L_w: the worker's firm-pair Laplacian, c_f c_g / Dw),
// grounded at the last firm.  The transformed column is u - (psi_f - psibar_w)
// with psibar_w = sum_a (c_a / Dw) psi_fa, and the Gram of the transformed
// columns follows from the match moments
//     G = sum_matches m_w (Q - d s' - s d' + n d d'),  d = psi_f - psibar_w,
// s = sum u, Q = sum u u', n = rows of the match.  A lane is one (resample,
// column) pair, four lanes per resample (columns past C are zero), 64
// resamples per wide solve: the resamples of a batch share every operator
// pass, each with its own weights.  The Gram is accumulated in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "absorb_kernels.h"
#include "akm_wide.h"
#include "common.h"
#include "xfe_hip.h"

namespace xfe {
namespace {

constexpr int kBootCols = 4;                           // lanes per resample
constexpr int kBootBatch = kWideMaxLanes / kBootCols;  // resamples per wide solve
constexpr int kGram = kBootCols * (kBootCols + 1) / 2; // upper triangle of the Gram
constexpr int kGramGrid = 512;

// multT (N x kBootBatch) from mult (Rb x N, resample-major); missing resamples 0
__global__ void __launch_bounds__(kBlock) boot_transpose_kernel(const int* mult, int Rb, int N, int* multT) {
    const long long total = static_cast<long long>(N) * kBootBatch;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int w = static_cast<int>(i / kBootBatch), b = static_cast<int>(i % kBootBatch);
        multT[i] = b < Rb ? mult[static_cast<size_t>(b) * N + w] : 0;
    }
}

// Per (firm, lane): the right-hand side sum_m m_w s[m, c] and the Jacobi
// inverse of the diagonal sum_m m_w c_m (Dw - c_m) / Dw, over the firm's
// matches (CSR f_ptr / f_m).  A firm whose sampled workers are all stayers
// has a zero row of S: its diagonal is exactly zero in this form (and its
// right-hand side too, s being zero for single-match workers), so it gets no
// Jacobi weight; the ground firm and diagonals below 1e-12 of the firm's
// sampled count give 0 as well.
__global__ void __launch_bounds__(kBlock)
boot_rhs_kernel(const int* f_ptr, const int* f_m, const int* m_w, const double* m_c, const double* dw,
                const double* s, const int* multT, int J, int ground, double* b, float* minv) {
    const long long total = static_cast<long long>(J) * kWideMaxLanes;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int f = static_cast<int>(i / kWideMaxLanes), l = static_cast<int>(i % kWideMaxLanes);
        const int rs = l / kBootCols, c = l % kBootCols;
        double rhs = 0.0, d = 0.0, cnt = 0.0;
        for (int k = f_ptr[f]; k < f_ptr[f + 1]; ++k) {
            const int m = f_m[k];
            const int w = m_w[m];
            const double mw = static_cast<double>(multT[static_cast<size_t>(w) * kBootBatch + rs]);
            if (mw == 0.0) continue;
            const double cm = m_c[m];
            rhs += mw * s[static_cast<size_t>(m) * kBootCols + c];
            d += mw * cm * (dw[w] - cm) / dw[w];
            cnt += mw * cm;
        }
        b[i] = f == ground ? 0.0 : rhs;
        minv[i] = (f == ground || !(d > 1e-12 * cnt)) ? 0.0f : static_cast<float>(1.0 / d);
    }
}

// Gram partials: thread (worker slot, resample rs) over workers strided by
// the grid; per worker psibar = sum_a (c_a / Dw) psi[f_a], then per match
// d = psi[f] - psibar and m_w (Q - d s' - s d' + n d d'); the block sums the
// worker slots per resample in slot order into part[block][rs][kGram].
__global__ void __launch_bounds__(kBlock)
boot_gram_kernel(const int* w_ptr, const int* m_f, const double* m_c, const double* idw, const double* s,
                 const double* Q, const double* psi, const int* multT, int N, double* part) {
    constexpr int kSlots = kBlock / kBootBatch;
    __shared__ double red[kSlots][kBootBatch][kGram];
    const int rs = threadIdx.x % kBootBatch, slot = threadIdx.x / kBootBatch;
    double g[kGram];
#pragma unroll
    for (int k = 0; k < kGram; ++k) g[k] = 0.0;
    for (int w = blockIdx.x * kSlots + slot; w < N; w += gridDim.x * kSlots) {
        const double mw = static_cast<double>(multT[static_cast<size_t>(w) * kBootBatch + rs]);
        if (mw == 0.0) continue;
        const int a0 = w_ptr[w], a1 = w_ptr[w + 1];
        double pb[kBootCols] = {0.0, 0.0, 0.0, 0.0};
        for (int a = a0; a < a1; ++a) {
            const double* ps = psi + static_cast<size_t>(m_f[a]) * kWideMaxLanes + rs * kBootCols;
            const double wt = m_c[a] * idw[w];
#pragma unroll
            for (int c = 0; c < kBootCols; ++c) pb[c] += wt * ps[c];
        }
        for (int a = a0; a < a1; ++a) {
            const double* ps = psi + static_cast<size_t>(m_f[a]) * kWideMaxLanes + rs * kBootCols;
            const double* sa = s + static_cast<size_t>(a) * kBootCols;
            const double* qa = Q + static_cast<size_t>(a) * kGram;
            const double n = m_c[a];
            double d[kBootCols];
#pragma unroll
            for (int c = 0; c < kBootCols; ++c) d[c] = ps[c] - pb[c];
            int k = 0;
#pragma unroll
            for (int c = 0; c < kBootCols; ++c)
#pragma unroll
                for (int e = c; e < kBootCols; ++e, ++k)
                    g[k] += mw * (qa[k] - d[c] * sa[e] - sa[c] * d[e] + n * d[c] * d[e]);
        }
    }
#pragma unroll
    for (int k = 0; k < kGram; ++k) red[slot][rs][k] = g[k];
    __syncthreads();
    if (slot == 0) {
        for (int k = 0; k < kGram; ++k) {
            double t = 0.0;
            for (int j = 0; j < kSlots; ++j) t += red[j][rs][k];
            part[(static_cast<size_t>(blockIdx.x) * kBootBatch + rs) * kGram + k] = t;
        }
    }
}

// out[rs][k] = sum over blocks (in block order)
__global__ void __launch_bounds__(kBlock) boot_gram_reduce_kernel(const double* part, int nblk, double* out) {
    for (int i = threadIdx.x; i < kBootBatch * kGram; i += kBlock) {
        double t = 0.0;
        for (int b = 0; b < nblk; ++b) t += part[static_cast<size_t>(b) * kBootBatch * kGram + i];
        out[i] = t;
    }
}

int grid_ew(long long total) {
    const long long g = (total + kBlock - 1) / kBlock;
    return static_cast<int>(std::max(1LL, std::min(g, 8192LL)));
}

}  // namespace
}  // namespace xfe

using namespace xfe;

struct xfe_boot_context {
    int n = 0, N = 0, J = 0, M = 0, C = 0;
    hipStream_t s = nullptr;
    DevBuf<int> w_ptr, m_f, m_w, f_ptr, f_m, mult, multT;
    DevBuf<double> m_c, idw, dw, ms, mQ, part, gram;
    DevBuf<float> minv;
    AkmWide wide;
    double* h_gram = nullptr;
    int last_passes = 0, last_updates = 0;

    ~xfe_boot_context() {
        wide.release();
        for (DevBuf<int>* d : {&w_ptr, &m_f, &m_w, &f_ptr, &f_m, &mult, &multT}) d->release();
        for (DevBuf<double>* d : {&m_c, &idw, &dw, &ms, &mQ, &part, &gram}) d->release();
        minv.release();
        if (h_gram) (void)hipHostFree(h_gram);
        if (s) (void)hipStreamDestroy(s);
    }
};

extern "C" {

xfe_boot_context* xfe_boot_create(int n, int n_workers, int n_firms, const int* worker, const int* firm,
                                  const double* V, int ldv, int C) {
    clear_error();
    if (n <= 0 || n_workers <= 0 || n_firms < 2 || !worker || !firm || !V || C < 1 || C > kBootCols || ldv < n) {
        set_error("xfe_boot_create: invalid arguments");
        return nullptr;
    }
    const int N = n_workers, J = n_firms;
    for (int i = 0; i < n; ++i) {
        if (worker[i] < 0 || worker[i] >= N || firm[i] < 0 || firm[i] >= J) {
            set_error("xfe_boot_create: ids out of range");
            return nullptr;
        }
    }
    // rows by (worker, firm); worker means of each column; matches with their
    // moment sums of the worker-demeaned columns
    std::vector<int> rows(n);
    for (int i = 0; i < n; ++i) rows[i] = i;
    std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) {
        return worker[a] < worker[b] || (worker[a] == worker[b] && firm[a] < firm[b]);
    });
    std::vector<double> wmean(static_cast<size_t>(N) * kBootCols, 0.0), wcnt(N, 0.0);
    for (int i = 0; i < n; ++i) {
        wcnt[worker[i]] += 1.0;
        for (int c = 0; c < C; ++c) wmean[static_cast<size_t>(worker[i]) * kBootCols + c] += V[static_cast<size_t>(c) * ldv + i];
    }
    for (int w = 0; w < N; ++w)
        for (int c = 0; c < C; ++c)
            if (wcnt[w] > 0.0) wmean[static_cast<size_t>(w) * kBootCols + c] /= wcnt[w];
    std::vector<int> h_wptr(N + 1, 0), h_mf, h_mw;
    std::vector<double> h_mc, h_s, h_Q;
    for (int k = 0; k < n; ++k) {
        const int i = rows[k];
        const int w = worker[i], f = firm[i];
        if (h_mw.empty() || h_mw.back() != w || h_mf.back() != f) {
            h_mw.push_back(w);
            h_mf.push_back(f);
            h_mc.push_back(0.0);
            h_s.insert(h_s.end(), kBootCols, 0.0);
            h_Q.insert(h_Q.end(), kGram, 0.0);
            h_wptr[w + 1] += 1;
        }
        double u[kBootCols] = {0.0, 0.0, 0.0, 0.0};
        for (int c = 0; c < C; ++c) u[c] = V[static_cast<size_t>(c) * ldv + i] - wmean[static_cast<size_t>(w) * kBootCols + c];
        h_mc.back() += 1.0;
        double* sm = &h_s[h_s.size() - kBootCols];
        double* qm = &h_Q[h_Q.size() - kGram];
        int q = 0;
        for (int c = 0; c < kBootCols; ++c) {
            sm[c] += u[c];
            for (int e = c; e < kBootCols; ++e, ++q) qm[q] += u[c] * u[e];
        }
    }
    for (int w = 0; w < N; ++w) h_wptr[w + 1] += h_wptr[w];
    const int M = static_cast<int>(h_mw.size());
    // a worker with one match: its match holds all its rows, so s = sum u is
    // zero (exactly, not the rounding of the worker-mean subtraction)
    for (int w = 0; w < N; ++w)
        if (h_wptr[w + 1] - h_wptr[w] == 1)
            for (int c = 0; c < kBootCols; ++c) h_s[static_cast<size_t>(h_wptr[w]) * kBootCols + c] = 0.0;
    std::vector<double> h_idw(N, 0.0);
    for (int w = 0; w < N; ++w)
        if (wcnt[w] > 0.0) h_idw[w] = 1.0 / wcnt[w];
    // firm CSR over the matches
    std::vector<int> h_fptr(J + 1, 0), h_fm(M);
    for (int m = 0; m < M; ++m) h_fptr[h_mf[m] + 1] += 1;
    for (int f = 0; f < J; ++f) h_fptr[f + 1] += h_fptr[f];
    {
        std::vector<int> pos(h_fptr.begin(), h_fptr.end() - 1);
        for (int m = 0; m < M; ++m) h_fm[pos[h_mf[m]]++] = m;
    }
    // per-worker firm pairs (both directions), weight c_f c_g / Dw
    std::vector<int> pf, pg, pw;
    std::vector<double> pb;
    for (int w = 0; w < N; ++w) {
        for (int a = h_wptr[w]; a < h_wptr[w + 1]; ++a)
            for (int b = h_wptr[w]; b < h_wptr[w + 1]; ++b)
                if (a != b) {
                    pf.push_back(h_mf[a]);
                    pg.push_back(h_mf[b]);
                    pw.push_back(w);
                    pb.push_back(h_mc[a] * h_mc[b] * h_idw[w]);
                }
    }
    xfe_boot_context* c = nullptr;
    try {
        (void)hipGetLastError();
        c = new xfe_boot_context();
        c->n = n; c->N = N; c->J = J; c->M = M; c->C = C;
        XFE_CHECK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
        auto up = [&](auto& d, const auto& h) {
            d.ensure(h.size());
            if (!h.empty())
                XFE_CHECK(hipMemcpyAsync(d.p, h.data(), sizeof(h[0]) * h.size(), hipMemcpyHostToDevice, c->s));
        };
        up(c->w_ptr, h_wptr); up(c->m_f, h_mf); up(c->m_w, h_mw); up(c->f_ptr, h_fptr); up(c->f_m, h_fm);
        up(c->m_c, h_mc); up(c->idw, h_idw); up(c->dw, wcnt); up(c->ms, h_s); up(c->mQ, h_Q);
        c->wide.build_worker_pairs(J, J - 1, pf, pg, pw, pb, c->s);
        c->wide.ensure(kWideMaxLanes);
        c->mult.ensure(static_cast<size_t>(kBootBatch) * N);
        c->multT.ensure(static_cast<size_t>(kBootBatch) * N);
        c->minv.ensure(static_cast<size_t>(J) * kWideMaxLanes);
        c->part.ensure(static_cast<size_t>(kGramGrid) * kBootBatch * kGram);
        c->gram.ensure(static_cast<size_t>(kBootBatch) * kGram);
        XFE_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c->h_gram), sizeof(double) * kBootBatch * kGram,
                                hipHostMallocDefault));
        c->wide.mult = c->multT.p;
        c->wide.mstride = kBootBatch;
        c->wide.mcols = kBootCols;
        c->wide.minv_lane = c->minv.p;
        XFE_CHECK(hipStreamSynchronize(c->s));
        return c;
    } catch (const std::exception& e) {
        set_error(e.what());
        delete c;
        return nullptr;
    }
}

void xfe_boot_destroy(xfe_boot_context* c) { delete c; }

int xfe_boot_refits(xfe_boot_context* c, int R, const int* mult, double tol, int max_iter, double* gram_out,
                    int* iters_out) {
    clear_error();
    if (!c || R < 1 || !mult || !gram_out || !iters_out || max_iter < 1 || !(tol >= 0.0)) {
        set_error("xfe_boot_refits: invalid arguments");
        return -1;
    }
    for (long long i = 0; i < static_cast<long long>(R) * c->N; ++i) {
        if (mult[i] < 0) {
            set_error("xfe_boot_refits: negative multiplicity");
            return -1;
        }
    }
    try {
        (void)hipGetLastError();
        const int N = c->N, J = c->J, C = c->C;
        hipStream_t s = c->s;
        bool failed = false;
        c->last_passes = 0;
        c->last_updates = 0;
        for (int r0 = 0; r0 < R; r0 += kBootBatch) {
            const int Rb = std::min(kBootBatch, R - r0);
            XFE_CHECK(hipMemcpyAsync(c->mult.p, mult + static_cast<size_t>(r0) * N, sizeof(int) * Rb * N,
                                     hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(boot_transpose_kernel, dim3(grid_ew(static_cast<long long>(N) * kBootBatch)),
                               dim3(kBlock), 0, s, c->mult.p, Rb, N, c->multT.p);
            hipLaunchKernelGGL(boot_rhs_kernel, dim3(grid_ew(static_cast<long long>(J) * kWideMaxLanes)),
                               dim3(kBlock), 0, s, c->f_ptr.p, c->f_m.p, c->m_w.p, c->m_c.p, c->dw.p, c->ms.p,
                               c->multT.p, J, J - 1, c->wide.b64.p, c->minv.p);
            XFE_CHECK(hipGetLastError());
            c->wide.trace = std::getenv("XFE_BOOT_TRACE") ? "boot" : nullptr;
            c->wide.solve_device(kWideMaxLanes, tol, max_iter);
            c->last_passes += c->wide.last_passes;
            c->last_updates += c->wide.last_updates;
            c->wide.unscale_x64();
            const int gg = std::max(1, std::min(kGramGrid, (N + 3) / 4));
            hipLaunchKernelGGL(boot_gram_kernel, dim3(gg), dim3(kBlock), 0, s, c->w_ptr.p, c->m_f.p, c->m_c.p,
                               c->idw.p, c->ms.p, c->mQ.p, c->wide.x64.p, c->multT.p, N, c->part.p);
            hipLaunchKernelGGL(boot_gram_reduce_kernel, dim3(1), dim3(kBlock), 0, s, c->part.p, gg, c->gram.p);
            XFE_CHECK(hipGetLastError());
            XFE_CHECK(hipMemcpyAsync(c->h_gram, c->gram.p, sizeof(double) * kBootBatch * kGram,
                                     hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
            for (int b = 0; b < Rb; ++b) {
                int it = 0;
                for (int col = 0; col < C; ++col) {
                    const int li = c->wide.lane_iterations(b * kBootCols + col);
                    if (li < 0) {
                        it = -1;
                        break;
                    }
                    it = std::max(it, li);
                }
                if (it < 0) failed = true;
                if (it < 0 && c->wide.trace)
                    for (int col = 0; col < C; ++col)
                        std::fprintf(stderr, "[boot] refit %d column %d: %s\n", r0 + b, col,
                                     c->wide.lane_report(b * kBootCols + col).c_str());
                iters_out[r0 + b] = it;
                double* g = gram_out + static_cast<size_t>(r0 + b) * C * C;
                int k = 0;
                for (int a = 0; a < kBootCols; ++a)
                    for (int e = a; e < kBootCols; ++e, ++k)
                        if (a < C && e < C) {
                            const double v = c->h_gram[b * kGram + k];
                            g[a * C + e] = v;
                            g[e * C + a] = v;
                        }
            }
        }
        if (failed) set_error("xfe_boot_refits: some refit solves did not converge");
        return failed ? 1 : 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int xfe_boot_last(xfe_boot_context* c, int* fp32_passes, int* fp64_replacements) {
    clear_error();
    if (!c || !fp32_passes || !fp64_replacements) {
        set_error("xfe_boot_last: invalid arguments");
        return -1;
    }
    *fp32_passes = c->last_passes;
    *fp64_replacements = c->last_updates;
    return 0;
}

}  // extern "C"
