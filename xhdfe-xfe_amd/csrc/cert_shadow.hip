// Shadow-mode device certificate with rigorous bounds (cuda_cert_shadow_begin /
// _finish / _abort, src/fe_absorption_cuda.hpp:74-103, impl
// src/fe_absorption_cuda_certificate.cu).  Its own kernels, scratch and
// reductions; contraction of a*b+c into FMA is off in this file, so every
// product and sum below is one correctly rounded (to nearest) FP64 operation.
//
// Quantities (per column c of [y | X], D standard dimensions, weights w):
//   Orig_sq(c) = sum_i v_ic^2,  Op_sq = sum_d sum_i w_i^2 = D sum_i w_i^2,
//   Res_sq(c)  = sum_d sum_g (sum_{i in g} w_i r_ic)^2   (r: the residual).
// Bounds (Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed.,
// Lemma 3.1 and section 4.2): a sum of k terms, each a product rounded once,
// computed in any order, is within gamma_{k+1} sum |terms| of the exact sum,
// gamma_k = k u / (1 - k u), u = 2^-53.  Each level sum S_g is bracketed by
// |S_g| -+ gamma_{n_g+1} T_g with T_g the computed sum of |w_i r_ic|; the
// squares and their total get another gamma factor; every bound is pushed one
// ulp outward where it is formed, and an absolute term of 2^-1021 per
// operation covers underflow (including subnormals flushed to zero by the
// hardware FP64 atomics).  The decision needs no sqrt or division:
//   T_lower = L^2 A_lower B_lower, T_upper = L^2 A_upper B_upper (rounded
//   outward), PASS iff R_upper <= T_lower, FAIL iff R_lower > T_upper,
//   INDETERMINATE otherwise, ERROR for any non-finite input.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "absorb_kernels.h"
#include "common.h"
#include "xfe_hip.h"

#pragma clang fp contract(off)

namespace xfe {
namespace {

constexpr int kCsBlock = 256;
constexpr double kU = 1.1102230246251565e-16;       // 2^-53
constexpr double kEta = 4.4501477170144028e-308;    // 2^-1021, per operation (underflow)

__device__ __host__ inline double gamma_k(double k) {
    // k u / (1 - k u), pushed up by a relative 4u to cover its own roundings
    const double ku = k * kU;
    return ku / (1.0 - ku) * (1.0 + 4.0 * kU);
}

// counts[g] += 1, wsum[g] += w_i (1 without weights); an id outside [0, G)
// sets *bad and is skipped.
__global__ void cs_count_kernel(const int* __restrict__ gid, const double* __restrict__ w, int n, int G,
                                int* __restrict__ counts, double* __restrict__ wsum, int* __restrict__ bad) {
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int g = gid[i];
        if (g < 0 || g >= G) {
            atomicOr(bad, 1);
            continue;
        }
        atomicAdd(&counts[g], 1);
        atomicAdd(&wsum[g], w ? w[i] : 1.0);
    }
}

__device__ inline double block_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = kCsBlock / 2; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    const double out = red[0];
    __syncthreads();
    return out;
}

__device__ inline double block_max(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = kCsBlock / 2; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    const double out = red[0];
    __syncthreads();
    return out;
}

// Per block: sum_i v_ic^2 for every column (ncols) and sum_i w_i^2 (slot ncols).
__global__ void cs_sumsq_kernel(const double* __restrict__ y, const double* __restrict__ X, int ld, int n,
                                int cols, const double* __restrict__ w, double* __restrict__ part) {
    __shared__ double red[kCsBlock];
    const int stride = gridDim.x * blockDim.x;
    for (int c = 0; c <= cols + 1; ++c) {
        double acc = 0.0;
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
            double v;
            if (c == 0) v = y[i];
            else if (c <= cols) v = X[static_cast<size_t>(c - 1) * ld + i];
            else v = w ? w[i] : 1.0;
            acc += v * v;
        }
        const double s = block_sum(acc, red);
        if (threadIdx.x == 0) part[static_cast<size_t>(blockIdx.x) * (cols + 2) + c] = s;
    }
}

// S[g * C1 + c] += w_i r_ic, T[g * C1 + c] += |w_i r_ic| (C1 = cols + 1).
// With lds (both G x C1 tables fit), each block sums into LDS copies and adds
// their non-zero entries to S and T once; any order of the additions is
// covered by the bound.
__global__ void cs_level_kernel(const double* __restrict__ y, const double* __restrict__ X, int ld, int n,
                                int cols, const int* __restrict__ gid, int G, const double* __restrict__ w,
                                double* __restrict__ S, double* __restrict__ T, int lds) {
    extern __shared__ __align__(16) double ltab[];
    const int stride = gridDim.x * blockDim.x;
    const int C1 = cols + 1;
    const int GC = G * C1;
    if (lds) {
        for (int q = threadIdx.x; q < 2 * GC; q += blockDim.x) ltab[q] = 0.0;
        __syncthreads();
    }
    double* s_tab = lds ? ltab : S;
    double* t_tab = lds ? ltab + GC : T;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (gid[i] < 0 || gid[i] >= G) continue;  // rejected at begin
        const double wi = w ? w[i] : 1.0;
        const size_t base = static_cast<size_t>(gid[i]) * C1;
        for (int c = 0; c < C1; ++c) {
            const double v = c == 0 ? y[i] : X[static_cast<size_t>(c - 1) * ld + i];
            const double x = wi * v;
            atomicAdd(&s_tab[base + c], x);
            atomicAdd(&t_tab[base + c], fabs(x));
        }
    }
    if (lds) {
        __syncthreads();
        for (int q = threadIdx.x; q < GC; q += blockDim.x) {
            if (ltab[q] != 0.0) atomicAdd(&S[q], ltab[q]);
            if (ltab[GC + q] != 0.0) atomicAdd(&T[q], ltab[GC + q]);
        }
    }
}

// Per block and column: sum of lo_g^2, hi_g^2 and S_g^2 over the levels
// (part[(block * C1 + c) * 3 + {0,1,2}]).
__global__ void cs_bracket_kernel(const double* __restrict__ S, const double* __restrict__ T,
                                  const int* __restrict__ counts, int G, int C1, double* __restrict__ part) {
    __shared__ double red[kCsBlock];
    const int stride = gridDim.x * blockDim.x;
    for (int c = 0; c < C1; ++c) {
        double lo = 0.0, hi = 0.0, mid = 0.0;
        for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += stride) {
            const int ng = counts[g];
            if (ng == 0) continue;
            const double s = fabs(S[static_cast<size_t>(g) * C1 + c]);
            const double t = T[static_cast<size_t>(g) * C1 + c];
            const double e = gamma_k(static_cast<double>(ng) + 1.0) * t * (1.0 + 2.0 * kU) +
                             (static_cast<double>(ng) + 2.0) * kEta;
            const double l = fmax(0.0, (s - e) * (1.0 - 2.0 * kU));
            const double h = (s + e) * (1.0 + 2.0 * kU);
            lo += l * l;
            hi += h * h;
            mid += s * s;
        }
        const double a = block_sum(lo, red);
        const double b = block_sum(hi, red);
        const double m = block_sum(mid, red);
        if (threadIdx.x == 0) {
            double* p = part + (static_cast<size_t>(blockIdx.x) * C1 + c) * 3;
            p[0] = a;
            p[1] = b;
            p[2] = m;
        }
    }
}

// Weight-sum cross-check per level: part[block * 3 + {0: max |diff|, 1: max
// (diff - tol), 2: empty levels}] with tol = gamma_{n_g} * recomputed sum.
__global__ void cs_wsum_kernel(const double* __restrict__ re, const double* __restrict__ given,
                               const int* __restrict__ counts, int G, double* __restrict__ part) {
    __shared__ double red[kCsBlock];
    const int stride = gridDim.x * blockDim.x;
    double md = 0.0, mx = -1.0, empty = 0.0;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += stride) {
        const int ng = counts[g];
        if (ng == 0) empty += 1.0;
        const double d = given ? fabs(re[g] - given[g]) : 0.0;
        const double tol = 2.0 * gamma_k(static_cast<double>(ng) + 1.0) * fabs(re[g]) + (ng + 2.0) * kEta;
        md = fmax(md, d);
        mx = fmax(mx, d - tol);
    }
    const double a = block_max(md, red);
    const double b = block_max(mx, red);
    const double e = block_sum(empty, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x * 3 + 0] = a;
        part[blockIdx.x * 3 + 1] = b;
        part[blockIdx.x * 3 + 2] = e;
    }
}

int cs_grid(long long work) {
    long long g = (work + kCsBlock - 1) / kCsBlock;
    return static_cast<int>(std::max(1LL, std::min(g, 1024LL)));
}

double up(double x) { return std::nextafter(x, INFINITY); }
double down(double x) { return std::max(0.0, std::nextafter(x, -INFINITY)); }

// Sum of k non-negative terms computed in some order: bracket of the exact sum.
void bracket_sum(double computed_lo, double computed_hi, double k, double& lo, double& hi) {
    const double g = gamma_k(k + 2.0);
    lo = down(down(computed_lo * down(1.0 - g)) - (k + 2.0) * kEta);
    hi = up(up(computed_hi * up(1.0 + g)) + (k + 2.0) * kEta);
}

}  // namespace
}  // namespace xfe

using namespace xfe;

struct xfe_cert_shadow {
    int n = 0, cols = 0, ld = 0, D = 0;
    const double* w = nullptr;
    std::vector<const int*> gid;
    std::vector<int> G;
    std::vector<DevBuf<int>> counts;
    std::vector<double> B_lo, B_hi, B_mid;  // Orig_sq(c)
    double A_lo = 0.0, A_hi = 0.0, A_mid = 0.0, Q_mid = 0.0;
    double wsum_maxdiff = 0.0, wsum_tol = 0.0;
    int wsum_ok = 1;
    long long empty = 0;
    double begin_ms = 0.0;
    hipStream_t s = nullptr;
    DevBuf<double> part, S, T, wre;
    DevBuf<int> bad;

};

extern "C" {

xfe_cert_shadow* xfe_cert_shadow_begin(const double* d_y, const double* d_x, int n, int cols, int ld,
                                       const double* d_w, const int* const* d_group_ids,
                                       const double* const* d_weight_sums, const int* num_groups, int num_dims,
                                       const double* const* d_slope_values,
                                       const unsigned char* slope_has_intercept) {
    clear_error();
    if (!d_y || (cols > 0 && !d_x) || n <= 0 || cols < 0 || cols + 1 > XFE_MAX_COLS || ld < n || num_dims <= 0 ||
        num_dims > XFE_MAX_DIMS || !d_group_ids || !num_groups) {
        set_error("xfe_cert_shadow_begin: invalid arguments");
        return nullptr;
    }
    if (d_slope_values) {
        for (int d = 0; d < num_dims; ++d)
            if (d_slope_values[d]) {
                set_error("xfe_cert_shadow_begin: slope dimensions are not supported");
                return nullptr;
            }
    }
    (void)slope_has_intercept;
    auto* h = new xfe_cert_shadow();
    try {
        const auto t0 = std::chrono::steady_clock::now();
        (void)hipGetLastError();
        XFE_CHECK(hipDeviceSynchronize());  // the caller's buffers are complete
        XFE_CHECK(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
        h->n = n; h->cols = cols; h->ld = ld; h->D = num_dims; h->w = d_w;
        const int C1 = cols + 1;
        const int grid = cs_grid(n);
        // Orig_sq and sum w^2
        h->part.ensure(static_cast<size_t>(grid) * (C1 + 1));
        hipLaunchKernelGGL(cs_sumsq_kernel, dim3(grid), dim3(kCsBlock), 0, h->s, d_y, d_x, ld, n, cols, d_w,
                           h->part.p);
        XFE_CHECK(hipGetLastError());
        std::vector<double> hp(static_cast<size_t>(grid) * (C1 + 1));
        XFE_CHECK(hipMemcpyAsync(hp.data(), h->part.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, h->s));
        XFE_CHECK(hipStreamSynchronize(h->s));
        h->B_lo.resize(C1); h->B_hi.resize(C1); h->B_mid.resize(C1);
        for (int c = 0; c <= C1; ++c) {
            double t = 0.0;
            for (int b = 0; b < grid; ++b) t += hp[static_cast<size_t>(b) * (C1 + 1) + c];
            if (c < C1) {
                h->B_mid[c] = t;
                bracket_sum(t, t, static_cast<double>(n), h->B_lo[c], h->B_hi[c]);
            } else {
                h->Q_mid = t;
                double qlo, qhi;
                if (d_w) {
                    bracket_sum(t, t, static_cast<double>(n), qlo, qhi);
                } else {
                    qlo = qhi = t;  // n ones: exact below 2^53
                }
                h->A_mid = num_dims * t;
                h->A_lo = d_w ? down(num_dims * qlo) : num_dims * qlo;
                h->A_hi = d_w ? up(num_dims * qhi) : num_dims * qhi;
            }
        }
        // per-level counts and the weight-sum cross-check
        h->counts.resize(num_dims);
        h->bad.ensure(1);
        XFE_CHECK(hipMemsetAsync(h->bad.p, 0, sizeof(int), h->s));
        double md = 0.0, mx = -1.0;
        for (int d = 0; d < num_dims; ++d) {
            const int G = num_groups[d];
            if (G <= 0 || !d_group_ids[d]) throw Error("invalid FE input");
            h->gid.push_back(d_group_ids[d]);
            h->G.push_back(G);
            h->counts[d].ensure(G);
            h->wre.ensure(G);
            XFE_CHECK(hipMemsetAsync(h->counts[d].p, 0, sizeof(int) * G, h->s));
            XFE_CHECK(hipMemsetAsync(h->wre.p, 0, sizeof(double) * G, h->s));
            hipLaunchKernelGGL(cs_count_kernel, dim3(grid), dim3(kCsBlock), 0, h->s, d_group_ids[d], d_w, n, G,
                               h->counts[d].p, h->wre.p, h->bad.p);
            XFE_CHECK(hipGetLastError());
            const int gg = cs_grid(G);
            h->part.ensure(static_cast<size_t>(gg) * 3);
            hipLaunchKernelGGL(cs_wsum_kernel, dim3(gg), dim3(kCsBlock), 0, h->s, h->wre.p,
                               d_weight_sums ? d_weight_sums[d] : nullptr, h->counts[d].p, G, h->part.p);
            XFE_CHECK(hipGetLastError());
            std::vector<double> wp(static_cast<size_t>(gg) * 3);
            XFE_CHECK(hipMemcpyAsync(wp.data(), h->part.p, sizeof(double) * wp.size(), hipMemcpyDeviceToHost, h->s));
            XFE_CHECK(hipStreamSynchronize(h->s));
            for (int b = 0; b < gg; ++b) {
                md = std::max(md, wp[b * 3]);
                mx = std::max(mx, wp[b * 3 + 1]);
                h->empty += static_cast<long long>(wp[b * 3 + 2]);
            }
        }
        int bad = 0;
        XFE_CHECK(hipMemcpyAsync(&bad, h->bad.p, sizeof(int), hipMemcpyDeviceToHost, h->s));
        XFE_CHECK(hipStreamSynchronize(h->s));
        if (bad) throw Error("xfe_cert_shadow_begin: a group id is outside [0, num_groups)");
        h->wsum_maxdiff = md;
        h->wsum_tol = md - mx;  // the tolerance at the worst level
        h->wsum_ok = mx <= 0.0 ? 1 : 0;
        h->begin_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return h;
    } catch (const std::exception& e) {
        set_error(e.what());
        if (h->s) (void)hipStreamDestroy(h->s);
        delete h;
        return nullptr;
    }
}

static void cs_release(xfe_cert_shadow* h) {
    for (auto& b : h->counts) b.release();
    h->part.release(); h->S.release(); h->T.release(); h->wre.release(); h->bad.release();
    if (h->s) (void)hipStreamDestroy(h->s);
    delete h;
}

int xfe_cert_shadow_finish(xfe_cert_shadow* h, const double* d_y_res, const double* d_x_res, double limit,
                           xfe_cert_shadow_result* out) {
    clear_error();
    if (!out) {
        set_error("xfe_cert_shadow_finish: invalid arguments");
        if (h) cs_release(h);
        return 0;
    }
    std::memset(out, 0, sizeof(*out));
    out->verdict = 3;
    out->host_like_pass = -1;
    if (!h || !d_y_res || (h->cols > 0 && !d_x_res)) {
        set_error("xfe_cert_shadow_finish: invalid arguments");
        if (h) cs_release(h);
        return 0;
    }
    try {
        const auto t0 = std::chrono::steady_clock::now();
        (void)hipGetLastError();
        XFE_CHECK(hipDeviceSynchronize());
        const int C1 = h->cols + 1, n = h->n;
        std::vector<double> R_lo(C1, 0.0), R_hi(C1, 0.0), R_mid(C1, 0.0);
        double K = 0.0;  // levels over all dimensions (terms of the outer sums)
        double rel_perdim = 0.0;
        for (int d = 0; d < h->D; ++d) {
            const int G = h->G[d];
            const size_t GC = static_cast<size_t>(G) * C1;
            h->S.ensure(GC);
            h->T.ensure(GC);
            XFE_CHECK(hipMemsetAsync(h->S.p, 0, sizeof(double) * GC, h->s));
            XFE_CHECK(hipMemsetAsync(h->T.p, 0, sizeof(double) * GC, h->s));
            const size_t lds_bytes = 2 * GC * sizeof(double);
            const int lds = lds_bytes <= 48 * 1024 ? 1 : 0;
            hipLaunchKernelGGL(cs_level_kernel, dim3(cs_grid(n)), dim3(kCsBlock), lds ? lds_bytes : 0, h->s, d_y_res,
                               d_x_res, h->ld, n, h->cols, h->gid[d], G, h->w, h->S.p, h->T.p, lds);
            XFE_CHECK(hipGetLastError());
            const int gg = cs_grid(G);
            h->part.ensure(static_cast<size_t>(gg) * C1 * 3);
            hipLaunchKernelGGL(cs_bracket_kernel, dim3(gg), dim3(kCsBlock), 0, h->s, h->S.p, h->T.p, h->counts[d].p,
                               G, C1, h->part.p);
            XFE_CHECK(hipGetLastError());
            std::vector<double> bp(static_cast<size_t>(gg) * C1 * 3);
            XFE_CHECK(hipMemcpyAsync(bp.data(), h->part.p, sizeof(double) * bp.size(), hipMemcpyDeviceToHost, h->s));
            XFE_CHECK(hipStreamSynchronize(h->s));
            for (int c = 0; c < C1; ++c) {
                double lo = 0.0, hi = 0.0, mid = 0.0;
                for (int b = 0; b < gg; ++b) {
                    const double* p = &bp[(static_cast<size_t>(b) * C1 + c) * 3];
                    lo += p[0]; hi += p[1]; mid += p[2];
                }
                R_lo[c] += lo; R_hi[c] += hi; R_mid[c] += mid;
                const double den = std::sqrt(h->Q_mid) * std::sqrt(h->B_mid[c]);
                if (den > 0.0) rel_perdim = std::max(rel_perdim, std::sqrt(mid) / den);
            }
            K += G;
        }
        const double L = limit;
        int verdict = 0;  // worst over columns: ERROR > FAIL > INDETERMINATE > PASS
        auto rank = [](int v) { return v == 3 ? 3 : (v == 1 ? 2 : (v == 2 ? 1 : 0)); };
        double rel_max = 0.0;
        int deciding = 0;
        for (int c = 0; c < C1; ++c) {
            double rlo, rhi;
            bracket_sum(R_lo[c], R_hi[c], K, rlo, rhi);
            const double llo = down(L * L), lhi = up(L * L);
            const double t_lo = down(down(llo * h->A_lo) * h->B_lo[c]);
            const double t_hi = up(up(lhi * h->A_hi) * h->B_hi[c]);
            int v;
            const bool finite = std::isfinite(rlo) && std::isfinite(rhi) && std::isfinite(h->A_lo) &&
                                std::isfinite(h->A_hi) && std::isfinite(h->B_lo[c]) && std::isfinite(h->B_hi[c]) &&
                                std::isfinite(L) && L >= 0.0;
            if (!finite) v = 3;
            else if (rhi <= t_lo) v = 0;
            else if (rlo > t_hi) v = 1;
            else v = 2;
            const double den = std::sqrt(h->A_mid) * std::sqrt(h->B_mid[c]);
            const double rel = den > 0.0 ? std::sqrt(R_mid[c]) / den : (R_mid[c] > 0.0 ? INFINITY : 0.0);
            if (rank(v) > rank(verdict) || (rank(v) == rank(verdict) && rel > rel_max)) {
                deciding = c;
                out->r_lower = rlo;
                out->r_upper = rhi;
            }
            if (rank(v) > rank(verdict)) verdict = v;
            rel_max = std::max(rel_max, rel);
        }
        (void)deciding;
        out->verdict = verdict;
        out->rel_max = rel_max;
        out->rel_max_perdim = rel_perdim;
        out->host_like_pass = rel_max <= L ? 1 : 0;
        out->wsum_maxdiff = h->wsum_maxdiff;
        out->wsum_tol = h->wsum_tol;
        out->wsum_ok = h->wsum_ok;
        out->empty_groups = h->empty;
        out->shadow_begin_ms = h->begin_ms;
        out->shadow_finish_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        out->dims = h->D;
        out->rhs = C1;
        cs_release(h);
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        out->verdict = 3;
        cs_release(h);
        return 0;
    }
}

void xfe_cert_shadow_abort(xfe_cert_shadow* h) {
    if (h) cs_release(h);
}

}  // extern "C"
