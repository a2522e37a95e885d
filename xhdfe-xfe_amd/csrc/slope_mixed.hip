// Mixed-precision absorption with heterogeneous slopes (xfe_absorb with a
// slope dimension and precision 1): the worker-like dimension p (the standard
// dimension with the most levels) is eliminated exactly, and PCG runs on the
// Schur system of the other unknowns,
//     S x = D_R' M_p D_R x = D_R' M_p v = b,
// where the columns of D_R are, per remaining dimension, its level indicators
// (standard dimensions and slope intercepts) or its level indicators times z
// (slope terms), and M_p subtracts the mean over p's levels.  A slope's
// intercept that a finer standard dimension already spans (each of that
// dimension's levels inside one of the slope's levels, as firm x year cells
// inside firms) adds nothing to the span and is left out, so S stays
// nonsingular on the cells.
//
// One application of S, with the rows sorted by p: per (p level, column) the
// rows' u = sum_r m_r x_r[g_r] (m the slope value or 1), their mean, and
// u - mean into a row-major n x 32 block R; then per remaining dimension the
// level sums of m R over fixed chunks of at most kChunk rows of one level
// (rows in level order through a permutation; each gather of a row's 32
// columns is one 128-byte line), and the chunk totals per level in chunk
// order (deterministic, no atomics).  The same pass in FP64 with the data
// (u = v - sum m x) gives the right-hand side and the final within transform.
// The PCG is the wide solver's (FP32 Jacobi-PCG, FP64 residual replacement,
// one lane per column).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "absorb_kernels.h"
#include "akm_wide.h"
#include "common.h"
#include "slope_mixed.h"

namespace xfe {
namespace {

constexpr int kCols = 32;       // row-major stride of R (columns <= 32)
constexpr int kChunk = 128;     // rows per reduction chunk
constexpr int kRowCache = 32;   // rows of a p level kept in registers between the two loops
constexpr int kSlotsPerBlock = kBlock / kCols;

template <typename T>
struct SmArgs {
    int n = 0, C = 0, Np = 0, NR = 0, L = 0;
    const int* pofs = nullptr;                 // Np + 1 row offsets of p's levels (rows sorted by p)
    const int* gid[kSmMaxR] = {};              // level of each row (rows sorted by p)
    const T* mult[kSmMaxR] = {};               // slope values (rows sorted by p), or null
    int off[kSmMaxR] = {};                     // first unknown of each dimension
    const T* x = nullptr;                      // N x L coefficients (null: zero)
    const double* V = nullptr;                 // data, n x C column-major (ld), rows sorted by p
    int ld = 0;
    T sign = T(1);                             // u = v + sign * D_R x
    T* R = nullptr;                            // out: n x kCols row-major (u - mean)
    double* Vout = nullptr;                    // STORE: the transform, column-major (ld)
    double* ssq = nullptr;                     // STORE: per block and column sum over p levels of (sum u - mean)^2
};

template <typename T, bool HAS_V, bool STORE>
__global__ void __launch_bounds__(kBlock) sm_apply_kernel(SmArgs<T> a) {
    const int col = threadIdx.x % kCols, slot = threadIdx.x / kCols;
    const bool live = col < a.C;
    double ss = 0.0;
    for (int w = blockIdx.x * kSlotsPerBlock + slot; w < a.Np; w += gridDim.x * kSlotsPerBlock) {
        const int i0 = a.pofs[w], i1 = a.pofs[w + 1];
        if (i1 <= i0 || !live) continue;
        T cache[kRowCache];
        T sum = T(0);
        auto row_value = [&](int i) {
            T u = T(0);
            if constexpr (HAS_V) u = static_cast<T>(a.V[static_cast<size_t>(col) * a.ld + i]);
            if (a.x) {
                for (int r = 0; r < a.NR; ++r) {
                    const T m = a.mult[r] ? a.mult[r][i] : T(1);
                    u += a.sign * m * a.x[static_cast<size_t>(a.off[r] + a.gid[r][i]) * a.L + col];
                }
            }
            return u;
        };
        for (int i = i0; i < i1; ++i) {
            const T u = row_value(i);
            if (i - i0 < kRowCache) cache[i - i0] = u;
            sum += u;
        }
        const T mean = sum / static_cast<T>(i1 - i0);
        double tot = 0.0;
        for (int i = i0; i < i1; ++i) {
            const T u = i - i0 < kRowCache ? cache[i - i0] : row_value(i);
            const T r = u - mean;
            a.R[static_cast<size_t>(i) * kCols + col] = r;
            if constexpr (STORE) {
                a.Vout[static_cast<size_t>(col) * a.ld + i] = static_cast<double>(r);
                tot += static_cast<double>(r);
            }
        }
        if constexpr (STORE) ss += tot * tot;
    }
    if constexpr (STORE) {
        __shared__ double red[kBlock];
        red[threadIdx.x] = ss;
        __syncthreads();
        if (slot == 0) {
            double t = 0.0;
            for (int k = 0; k < kSlotsPerBlock; ++k) t += red[k * kCols + col];
            a.ssq[static_cast<size_t>(blockIdx.x) * kCols + col] = t;
        }
    }
}

// part[k][col] = sum over chunk k's rows (level order through perm) of m R[row][col]
template <typename T>
__global__ void __launch_bounds__(kBlock)
sm_chunk_kernel(const int* __restrict__ perm, const int* __restrict__ cstart, int nchunks, const int* __restrict__ cend,
                const T* __restrict__ mult, const T* __restrict__ R, int stride, int C, T* __restrict__ part) {
    const int col = threadIdx.x % kCols, slot = threadIdx.x / kCols;
    if (col >= C) return;
    for (int k = blockIdx.x * kSlotsPerBlock + slot; k < nchunks; k += gridDim.x * kSlotsPerBlock) {
        T acc = T(0);
        for (int q = cstart[k]; q < cend[k]; ++q) {
            const int i = perm[q];
            const T m = mult ? mult[i] : T(1);
            acc += m * R[static_cast<size_t>(i) * stride + col];
        }
        part[static_cast<size_t>(k) * kCols + col] = acc;
    }
}

// out[(off + g) L + l] = sum of the level's chunk totals (l < C), 0 for l >= C
template <typename T>
__global__ void __launch_bounds__(kBlock)
sm_level_kernel(const int* __restrict__ cptr, const T* __restrict__ part, int G, int off, int L, int C,
                T* __restrict__ out) {
    const long long total = static_cast<long long>(G) * L;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int g = static_cast<int>(j / L), l = static_cast<int>(j % L);
        T acc = T(0);
        if (l < C)
            for (int k = cptr[g]; k < cptr[g + 1]; ++k) acc += part[static_cast<size_t>(k) * kCols + l];
        out[static_cast<size_t>(off + g) * L + l] = acc;
    }
}

// Per row of a p level, the Jacobi diagonal contributions of dimension r:
// d1 = m^2 minus, at the first row of each level within the p level,
// (sum of the level's m within the p level)^2 / rows; d2 = m^2.
__global__ void __launch_bounds__(kBlock)
sm_diag_rows_kernel(const int* pofs, int Np, const int* gid, const double* mult, double* d1, double* d2) {
    for (int w = blockIdx.x * kBlock + threadIdx.x; w < Np; w += gridDim.x * kBlock) {
        const int i0 = pofs[w], i1 = pofs[w + 1];
        const double nr = static_cast<double>(i1 - i0);
        for (int i = i0; i < i1; ++i) {
            const double m = mult ? mult[i] : 1.0;
            double c = m * m;
            bool first = true;
            for (int j = i0; j < i && first; ++j) first = gid[j] != gid[i];
            if (first) {
                double s = 0.0;
                for (int j = i; j < i1; ++j)
                    if (gid[j] == gid[i]) s += mult ? mult[j] : 1.0;
                c -= s * s / nr;
            }
            d1[i] = c;
            d2[i] = m * m;
        }
    }
}

// minv = 1 / diag where diag > 1e-12 of the level's sum of m^2, else 0
__global__ void __launch_bounds__(kBlock) sm_minv_kernel(const double* d1, const double* d2, int N, float* minv) {
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < N; g += gridDim.x * kBlock)
        minv[g] = d1[g] > 1e-12 * d2[g] ? static_cast<float>(1.0 / d1[g]) : 0.0f;
}

// right-hand sides of levels without a Jacobi weight (no coupling left after
// the elimination: their column of M_p D_R is zero) are zero
__global__ void __launch_bounds__(kBlock) sm_zero_b_kernel(double* b, const float* minv, int N, int L) {
    const long long total = static_cast<long long>(N) * L;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock)
        if (minv[j / L] == 0.0f) b[j] = 0.0;
}

// per block and column partial sums of squares of a rows x L table (l < C)
__global__ void __launch_bounds__(kBlock)
sm_sumsq_kernel(const double* a, int rows, int L, int C, double* part) {
    __shared__ double red[kBlock];
    const int col = threadIdx.x % kCols, slot = threadIdx.x / kCols;
    double acc = 0.0;
    if (col < C)
        for (int g = blockIdx.x * kSlotsPerBlock + slot; g < rows; g += gridDim.x * kSlotsPerBlock) {
            const double v = a[static_cast<size_t>(g) * L + col];
            acc += v * v;
        }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (slot == 0) {
        double t = 0.0;
        for (int k = 0; k < kSlotsPerBlock; ++k) t += red[k * kCols + col];
        part[static_cast<size_t>(blockIdx.x) * kCols + col] = t;
    }
}

__global__ void __launch_bounds__(kBlock) sm_iota_kernel(int* v, int n) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v[i] = i;
}

__global__ void __launch_bounds__(kBlock) sm_count_kernel(const int* key, int n, int* cnt) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) atomicAdd(&cnt[key[i]], 1);
}

__global__ void __launch_bounds__(kBlock) sm_nchunk_kernel(const int* cnt, int G, int* nch) {
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < G; g += gridDim.x * kBlock)
        nch[g] = (cnt[g] + kChunk - 1) / kChunk;
}

// chunk bounds of each level: [ptr[g] + j kChunk, min(ptr[g + 1], ...)) in level order
__global__ void __launch_bounds__(kBlock)
sm_chunk_fill_kernel(const int* ptr, const int* cptr, int G, int* cstart, int* cend) {
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < G; g += gridDim.x * kBlock) {
        int q = ptr[g];
        for (int k = cptr[g]; k < cptr[g + 1]; ++k) {
            cstart[k] = q;
            q = min(q + kChunk, ptr[g + 1]);
            cend[k] = q;
        }
    }
}

template <typename S, typename D>
__global__ void __launch_bounds__(kBlock) sm_gather_kernel(const S* src, const int* perm, int n, D* dst) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        dst[i] = static_cast<D>(src[perm[i]]);
}

__global__ void __launch_bounds__(kBlock)
sm_gather_cols_kernel(const double* V, int ld_in, const int* perm, int n, int C, double* out, int ld_out) {
    const long long total = static_cast<long long>(n) * C;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int c = static_cast<int>(j / n), i = static_cast<int>(j % n);
        out[static_cast<size_t>(c) * ld_out + i] = V[static_cast<size_t>(c) * ld_in + perm[i]];
    }
}

__global__ void __launch_bounds__(kBlock)
sm_scatter_cols_kernel(const double* Vs, int ld_in, const int* perm, int n, int C, double* V, int ld_out) {
    const long long total = static_cast<long long>(n) * C;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int c = static_cast<int>(j / n), i = static_cast<int>(j % n);
        V[static_cast<size_t>(c) * ld_out + perm[i]] = Vs[static_cast<size_t>(c) * ld_in + i];
    }
}

// nested check: lo[a] / hi[a] = min / max of b over the rows of level a
__global__ void __launch_bounds__(kBlock) sm_nest_kernel(const int* a, const int* b, int n, int* lo, int* hi) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        atomicMin(&lo[a[i]], b[i]);
        atomicMax(&hi[a[i]], b[i]);
    }
}

__global__ void __launch_bounds__(kBlock) sm_nest_flag_kernel(const int* lo, const int* hi, int G, int* bad) {
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < G; g += gridDim.x * kBlock)
        if (hi[g] >= 0 && lo[g] != hi[g]) atomicOr(bad, 1);
}

__global__ void __launch_bounds__(kBlock) sm_fill_int_kernel(int* v, int n, int x) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v[i] = x;
}

int sm_grid(long long work) {
    const long long g = (work + kBlock - 1) / kBlock;
    return static_cast<int>(std::max(1LL, std::min(g, 16384LL)));
}

// One remaining dimension: its unknowns' offset, slope values (rows sorted by
// p) and the level-order permutation with its chunks (shared by the intercept
// and slope parts of one grouping).
struct RDim {
    int d = 0;          // the caller's dimension
    int G = 0;
    int off = 0;
    bool slope = false;
    int grouping = 0;   // index into the groupings
};

struct Grouping {
    int d = 0;
    int G = 0;
    int nchunks = 0;
    DevBuf<int> gid, perm, cnt, ptr, nch, cptr, cstart, cend;
};

}  // namespace

struct SlopeMixed final : WideOp {
    hipStream_t s = nullptr;
    std::vector<int> dropped;  // dimensions whose intercept the last run left out
    int n = 0, C = 0, Np = 0, NR = 0, N = 0, ld = 0;
    std::vector<RDim> rd;
    std::vector<Grouping> gr;
    DevBuf<int> pperm, pkey, pcnt, pofs, tmp_int, flag;
    DevBuf<unsigned char> tmp;
    DevBuf<double> Vs, z64[kSmMaxR], R64, part64, d1, d2, dl1, dl2, ssq;
    DevBuf<float> z32[kSmMaxR], R32, part32, minv;
    AkmWide wide;
    int gridw = 1;

    template <typename T>
    SmArgs<T> args(const T* x, T* R) {
        SmArgs<T> a;
        a.n = n; a.C = C; a.Np = Np; a.NR = NR; a.L = wide.L; a.pofs = pofs.p; a.x = x; a.R = R;
        for (int r = 0; r < NR; ++r) {
            a.gid[r] = gr[rd[r].grouping].gid.p;
            a.off[r] = rd[r].off;
            if constexpr (sizeof(T) == 4) a.mult[r] = rd[r].slope ? z32[r].p : nullptr;
            else a.mult[r] = rd[r].slope ? z64[r].p : nullptr;
        }
        a.V = Vs.p;
        a.ld = ld;
        return a;
    }

    // level sums of m R per remaining dimension into out (N x L)
    template <typename T>
    void reduce(const T* R, T* part, T* out, int stride, int cols) {
        for (int r = 0; r < NR; ++r) {
            Grouping& g = gr[rd[r].grouping];
            const T* m = nullptr;
            if (rd[r].slope) {
                if constexpr (sizeof(T) == 4) m = z32[r].p;
                else m = z64[r].p;
            }
            hipLaunchKernelGGL((sm_chunk_kernel<T>), dim3(sm_grid(static_cast<long long>(g.nchunks) * kCols)),
                               dim3(kBlock), 0, s, g.perm.p, g.cstart.p, g.nchunks, g.cend.p, m, R, stride, cols,
                               part);
            hipLaunchKernelGGL((sm_level_kernel<T>), dim3(sm_grid(static_cast<long long>(g.G) * wide.L)),
                               dim3(kBlock), 0, s, g.cptr.p, part, g.G, rd[r].off, wide.L, cols, out);
            XFE_CHECK(hipGetLastError());
        }
    }

    void apply(const float* p, float* q) override {
        SmArgs<float> a = args<float>(p, R32.p);
        a.sign = 1.0f;
        hipLaunchKernelGGL((sm_apply_kernel<float, false, false>), dim3(gridw), dim3(kBlock), 0, s, a);
        XFE_CHECK(hipGetLastError());
        reduce<float>(R32.p, part32.p, q, kCols, C);
    }

    void apply(const double* p, double* q) override {
        SmArgs<double> a = args<double>(p, R64.p);
        a.sign = 1.0;
        hipLaunchKernelGGL((sm_apply_kernel<double, false, false>), dim3(gridw), dim3(kBlock), 0, s, a);
        XFE_CHECK(hipGetLastError());
        reduce<double>(R64.p, part64.p, q, kCols, C);
    }

    void release() {
        wide.release();
        for (auto& g : gr)
            for (DevBuf<int>* b : {&g.gid, &g.perm, &g.cnt, &g.ptr, &g.nch, &g.cptr, &g.cstart, &g.cend}) b->release();
        for (DevBuf<int>* b : {&pperm, &pkey, &pcnt, &pofs, &tmp_int, &flag}) b->release();
        tmp.release();
        for (DevBuf<double>* b : {&Vs, &R64, &part64, &d1, &d2, &dl1, &dl2, &ssq}) b->release();
        for (int r = 0; r < kSmMaxR; ++r) {
            z64[r].release();
            z32[r].release();
        }
        R32.release(); part32.release(); minv.release();
    }
};

namespace {

template <typename K, typename V>
void radix_pairs(DevBuf<unsigned char>& tmp, const K* kin, K* kout, const V* vin, V* vout, int m, int bits,
                 hipStream_t s) {
    size_t bytes = 0;
    XFE_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, static_cast<size_t>(m), 0, bits, s));
    tmp.ensure(bytes);
    XFE_CHECK(rocprim::radix_sort_pairs(tmp.p, bytes, kin, kout, vin, vout, static_cast<size_t>(m), 0, bits, s));
}

void excl_scan(DevBuf<unsigned char>& tmp, const int* in, int* out, int m, hipStream_t s) {
    size_t bytes = 0;
    XFE_CHECK(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, static_cast<size_t>(m), rocprim::plus<int>(), s));
    tmp.ensure(bytes);
    XFE_CHECK(rocprim::exclusive_scan(tmp.p, bytes, in, out, 0, static_cast<size_t>(m), rocprim::plus<int>(), s));
}

int bits_of(long long m) {
    int b = 1;
    while ((1LL << b) <= m) ++b;
    return b;
}

// level offsets (G + 1) of sorted keys: counts, exclusive scan, last entry
void offsets(DevBuf<unsigned char>& tmp, const int* keys, int n, int G, DevBuf<int>& cnt, DevBuf<int>& ptr,
             hipStream_t s) {
    cnt.ensure(static_cast<size_t>(G) + 1);
    ptr.ensure(static_cast<size_t>(G) + 1);
    XFE_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int) * (static_cast<size_t>(G) + 1), s));
    hipLaunchKernelGGL(sm_count_kernel, dim3(sm_grid(n)), dim3(kBlock), 0, s, keys, n, cnt.p);
    XFE_CHECK(hipGetLastError());
    excl_scan(tmp, cnt.p, ptr.p, G + 1, s);  // cnt[G] = 0: ptr[G] = n
}

}  // namespace

// per level and column sums of a column-major block over one grouping
__global__ void __launch_bounds__(kBlock)
sm_level_atomic_kernel(const double* V, int n, int C, int ld, const int* gid, double* out) {
    const long long total = static_cast<long long>(n) * C;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int c = static_cast<int>(j / n), i = static_cast<int>(j % n);
        atomicAdd(&out[static_cast<size_t>(gid[i]) * kCols + c], V[static_cast<size_t>(c) * ld + i]);
    }
}

void slope_mixed_level_sumsq(SlopeMixed& sm, const double* V, int n, int C, int ld, const int* gid, int G,
                             hipStream_t s, std::vector<double>& acc) {
    sm.dl1.ensure(static_cast<size_t>(G) * kCols);
    XFE_CHECK(hipMemsetAsync(sm.dl1.p, 0, sizeof(double) * static_cast<size_t>(G) * kCols, s));
    hipLaunchKernelGGL(sm_level_atomic_kernel, dim3(sm_grid(static_cast<long long>(n) * C)), dim3(kBlock), 0, s, V, n,
                       C, ld, gid, sm.dl1.p);
    const int gs = std::min(1024, sm_grid(static_cast<long long>(G) * kCols));
    sm.ssq.ensure(static_cast<size_t>(gs) * kCols);
    hipLaunchKernelGGL(sm_sumsq_kernel, dim3(gs), dim3(kBlock), 0, s, sm.dl1.p, G, kCols, C, sm.ssq.p);
    XFE_CHECK(hipGetLastError());
    std::vector<double> hp(static_cast<size_t>(gs) * kCols);
    XFE_CHECK(hipMemcpyAsync(hp.data(), sm.ssq.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    for (int b = 0; b < gs; ++b)
        for (int c = 0; c < C; ++c) acc[c] += hp[static_cast<size_t>(b) * kCols + c];
}

bool slope_mixed_intercept_dropped(const SlopeMixed& sm, int d) {
    return std::find(sm.dropped.begin(), sm.dropped.end(), d) != sm.dropped.end();
}

SlopeMixed* slope_mixed_new() { return new SlopeMixed(); }
void slope_mixed_delete(SlopeMixed* sm) {
    if (sm) {
        sm->release();
        delete sm;
    }
}

bool slope_mixed_eligible(int C, int D, const int* mode, const double* w) {
    if (w || C < 1 || C > kCols || D < 2) return false;
    int standard = 0, r = 0;
    for (int d = 0; d < D; ++d) {
        standard += mode[d] == 0;
        r += mode[d] == 1 ? 2 : 1;
    }
    return standard >= 1 && r - 1 <= kSmMaxR;
}

void slope_mixed_run(SlopeMixed& sm, double* V, int n, int C, int ld, const std::vector<const int*>& gid,
                     const std::vector<int>& G, const std::vector<const double*>& z, const std::vector<int>& mode,
                     double tol, int max_iter, hipStream_t s, SlopeMixedResult& out) {
    sm.s = s;
    sm.n = n;
    sm.C = C;
    const int D = static_cast<int>(G.size());
    // p: the standard dimension with the most levels
    int p = -1;
    for (int d = 0; d < D; ++d)
        if (mode[d] == 0 && (p < 0 || G[d] > G[p])) p = d;
    sm.Np = G[p];
    const int grid = sm_grid(n);
    // rows sorted by p (stable)
    sm.pkey.ensure(n); sm.pperm.ensure(n); sm.tmp_int.ensure(n);
    hipLaunchKernelGGL(sm_iota_kernel, dim3(grid), dim3(kBlock), 0, s, sm.tmp_int.p, n);
    radix_pairs<int, int>(sm.tmp, gid[p], sm.pkey.p, sm.tmp_int.p, sm.pperm.p, n, bits_of(G[p]), s);
    offsets(sm.tmp, sm.pkey.p, n, G[p], sm.pcnt, sm.pofs, s);
    sm.ld = ((n + 63) / 64) * 64;
    sm.Vs.ensure(static_cast<size_t>(sm.ld) * C);
    hipLaunchKernelGGL(sm_gather_cols_kernel, dim3(sm_grid(static_cast<long long>(n) * C)), dim3(kBlock), 0, s, V, ld,
                       sm.pperm.p, n, C, sm.Vs.p, sm.ld);
    XFE_CHECK(hipGetLastError());
    // the remaining dimensions' groupings (rows sorted by p, then level order);
    // the vector only grows, so the device buffers of earlier calls are reused
    if (static_cast<int>(sm.gr.size()) < D) sm.gr.resize(D);
    std::vector<int> gidx(D, -1);
    int ng = 0;
    for (int d = 0; d < D; ++d) {
        if (d == p) continue;
        Grouping& g = sm.gr[ng];
        g.d = d;
        g.G = G[d];
        g.gid.ensure(n); g.perm.ensure(n);
        hipLaunchKernelGGL((sm_gather_kernel<int, int>), dim3(grid), dim3(kBlock), 0, s, gid[d], sm.pperm.p, n,
                           g.gid.p);
        hipLaunchKernelGGL(sm_iota_kernel, dim3(grid), dim3(kBlock), 0, s, sm.tmp_int.p, n);
        sm.pkey.ensure(n);
        radix_pairs<int, int>(sm.tmp, g.gid.p, sm.pkey.p, sm.tmp_int.p, g.perm.p, n, bits_of(G[d]), s);
        offsets(sm.tmp, sm.pkey.p, n, G[d], g.cnt, g.ptr, s);
        g.nch.ensure(static_cast<size_t>(G[d]) + 1);
        g.cptr.ensure(static_cast<size_t>(G[d]) + 1);
        hipLaunchKernelGGL(sm_nchunk_kernel, dim3(sm_grid(G[d])), dim3(kBlock), 0, s, g.cnt.p, G[d], g.nch.p);
        XFE_CHECK(hipMemsetAsync(g.nch.p + G[d], 0, sizeof(int), s));
        excl_scan(sm.tmp, g.nch.p, g.cptr.p, G[d] + 1, s);
        int nch = 0;
        XFE_CHECK(hipMemcpyAsync(&nch, g.cptr.p + G[d], sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        g.nchunks = nch;
        g.cstart.ensure(std::max(nch, 1));
        g.cend.ensure(std::max(nch, 1));
        hipLaunchKernelGGL(sm_chunk_fill_kernel, dim3(sm_grid(G[d])), dim3(kBlock), 0, s, g.ptr.p, g.cptr.p, G[d],
                           g.cstart.p, g.cend.p);
        XFE_CHECK(hipGetLastError());
        gidx[d] = ng++;
    }
    // a slope's intercept is left out when a finer standard dimension (or p)
    // nests in the slope's grouping
    auto nested_in = [&](int fine, int coarse) {
        sm.tmp_int.ensure(std::max(n, 2 * G[fine]));
        int* lo = sm.tmp_int.p;
        int* hi = sm.tmp_int.p + G[fine];
        hipLaunchKernelGGL(sm_fill_int_kernel, dim3(sm_grid(G[fine])), dim3(kBlock), 0, s, lo, G[fine], INT_MAX);
        hipLaunchKernelGGL(sm_fill_int_kernel, dim3(sm_grid(G[fine])), dim3(kBlock), 0, s, hi, G[fine], -1);
        hipLaunchKernelGGL(sm_nest_kernel, dim3(grid), dim3(kBlock), 0, s, gid[fine], gid[coarse], n, lo, hi);
        sm.flag.ensure(1);
        XFE_CHECK(hipMemsetAsync(sm.flag.p, 0, sizeof(int), s));
        hipLaunchKernelGGL(sm_nest_flag_kernel, dim3(sm_grid(G[fine])), dim3(kBlock), 0, s, lo, hi, G[fine],
                           sm.flag.p);
        int bad = 1;
        XFE_CHECK(hipMemcpyAsync(&bad, sm.flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        return bad == 0;
    };
    sm.rd.clear();
    sm.dropped.clear();
    int N = 0;
    out.intercepts_dropped = 0;
    for (int d = 0; d < D; ++d) {
        if (d == p) continue;
        if (mode[d] == 0 || mode[d] == 1) {
            bool redundant = false;
            if (mode[d] == 1)
                for (int e = 0; e < D && !redundant; ++e)
                    if (e != d && mode[e] == 0 && G[e] >= G[d] && nested_in(e, d)) redundant = true;
            if (redundant) {
                ++out.intercepts_dropped;
                sm.dropped.push_back(d);
            } else {
                RDim r;
                r.d = d; r.G = G[d]; r.off = N; r.slope = false; r.grouping = gidx[d];
                sm.rd.push_back(r);
                N += G[d];
            }
        }
        if (mode[d] == 1 || mode[d] == 2) {
            RDim r;
            r.d = d; r.G = G[d]; r.off = N; r.slope = true; r.grouping = gidx[d];
            sm.rd.push_back(r);
            N += G[d];
        }
    }
    sm.NR = static_cast<int>(sm.rd.size());
    sm.N = N;
    for (int r = 0; r < sm.NR; ++r) {
        if (!sm.rd[r].slope) continue;
        sm.z64[r].ensure(n);
        sm.z32[r].ensure(n);
        hipLaunchKernelGGL((sm_gather_kernel<double, double>), dim3(grid), dim3(kBlock), 0, s, z[sm.rd[r].d],
                           sm.pperm.p, n, sm.z64[r].p);
        hipLaunchKernelGGL((sm_gather_kernel<double, float>), dim3(grid), dim3(kBlock), 0, s, z[sm.rd[r].d],
                           sm.pperm.p, n, sm.z32[r].p);
    }
    XFE_CHECK(hipGetLastError());
    // buffers
    size_t maxch = 1;
    for (int k = 0; k < ng; ++k) maxch = std::max(maxch, static_cast<size_t>(sm.gr[k].nchunks));
    sm.R32.ensure(static_cast<size_t>(n) * kCols);
    sm.R64.ensure(static_cast<size_t>(n) * kCols);
    sm.part32.ensure(maxch * kCols);
    sm.part64.ensure(maxch * kCols);
    sm.gridw = sm_grid(static_cast<long long>(sm.Np) * kCols);
    // Jacobi diagonal of S: per row contributions, reduced per level
    sm.d1.ensure(n); sm.d2.ensure(n); sm.dl1.ensure(static_cast<size_t>(N) * kCols);
    sm.dl2.ensure(static_cast<size_t>(N) * kCols);
    sm.minv.ensure(N);
    for (int r = 0; r < sm.NR; ++r) {
        Grouping& g = sm.gr[sm.rd[r].grouping];
        hipLaunchKernelGGL(sm_diag_rows_kernel, dim3(sm_grid(sm.Np)), dim3(kBlock), 0, s, sm.pofs.p, sm.Np, g.gid.p,
                           sm.rd[r].slope ? sm.z64[r].p : nullptr, sm.d1.p, sm.d2.p);
        for (int k = 0; k < 2; ++k) {
            hipLaunchKernelGGL((sm_chunk_kernel<double>), dim3(sm_grid(static_cast<long long>(g.nchunks) * kCols)),
                               dim3(kBlock), 0, s, g.perm.p, g.cstart.p, g.nchunks, g.cend.p, nullptr,
                               k == 0 ? sm.d1.p : sm.d2.p, 1, 1, sm.part64.p);
            hipLaunchKernelGGL((sm_level_kernel<double>), dim3(sm_grid(g.G)), dim3(kBlock), 0, s, g.cptr.p,
                               sm.part64.p, g.G, sm.rd[r].off, 1, 1, k == 0 ? sm.dl1.p : sm.dl2.p);
        }
        XFE_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(sm_minv_kernel, dim3(sm_grid(N)), dim3(kBlock), 0, s, sm.dl1.p, sm.dl2.p, N, sm.minv.p);
    XFE_CHECK(hipGetLastError());
    // the wide solver on this operator; b = D_R' M_p v in FP64
    sm.wide.set_operator(&sm, N, s, sm.minv.p);
    sm.wide.ensure(C);
    {
        SmArgs<double> a = sm.args<double>(nullptr, sm.R64.p);
        hipLaunchKernelGGL((sm_apply_kernel<double, true, false>), dim3(sm.gridw), dim3(kBlock), 0, s, a);
        XFE_CHECK(hipGetLastError());
        sm.reduce<double>(sm.R64.p, sm.part64.p, sm.wide.b64.p, kCols, C);
        hipLaunchKernelGGL(sm_zero_b_kernel, dim3(sm_grid(static_cast<long long>(N) * sm.wide.L)), dim3(kBlock), 0, s,
                           sm.wide.b64.p, sm.minv.p, N, sm.wide.L);
        XFE_CHECK(hipGetLastError());
    }
    sm.wide.solve_device(C, tol, max_iter);
    sm.wide.unscale_x64();
    out.iterations = 0;
    out.converged = true;
    for (int c = 0; c < C; ++c) {
        const int it = sm.wide.lane_iterations(c);
        if (it < 0) out.converged = false;
        out.iterations = std::max(out.iterations, it);
    }
    out.passes = sm.wide.last_passes;
    out.replacements = sm.wide.last_updates;
    if (std::getenv("XFE_SLOPE_TRACE")) {
        std::fprintf(stderr, "[slope mixed] N %d NR %d p %d dropped %d passes %d replacements %d\n", N, sm.NR, p,
                     out.intercepts_dropped, out.passes, out.replacements);
        for (int r = 0; r < sm.NR; ++r)
            std::fprintf(stderr, "[slope mixed]   block %d: dim %d slope %d levels %d off %d chunks %d\n", r,
                         sm.rd[r].d, sm.rd[r].slope ? 1 : 0, sm.rd[r].G, sm.rd[r].off,
                         sm.gr[sm.rd[r].grouping].nchunks);
        for (int c = 0; c < C; ++c) std::fprintf(stderr, "[slope mixed]   lane %d: %s\n", c, sm.wide.lane_report(c).c_str());
    }
    // the transform v - D_R beta - (its p means), in FP64, into Vs and R64;
    // the certificate's level sums: p from the pass, the others from R64
    const int gw = sm.gridw;
    sm.ssq.ensure(static_cast<size_t>(std::max(gw, 1024)) * kCols);
    {
        SmArgs<double> a = sm.args<double>(sm.wide.x64.p, sm.R64.p);
        a.sign = -1.0;
        a.Vout = sm.Vs.p;
        a.ssq = sm.ssq.p;
        hipLaunchKernelGGL((sm_apply_kernel<double, true, true>), dim3(gw), dim3(kBlock), 0, s, a);
        XFE_CHECK(hipGetLastError());
    }
    std::vector<double> resid(C, 0.0), hp(static_cast<size_t>(gw) * kCols);
    XFE_CHECK(hipMemcpyAsync(hp.data(), sm.ssq.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    for (int b = 0; b < gw; ++b)
        for (int c = 0; c < C; ++c) resid[c] += hp[static_cast<size_t>(b) * kCols + c];
    // level sums of the transform in q64 (free after the solve), then squares
    sm.reduce<double>(sm.R64.p, sm.part64.p, sm.wide.q64.p, kCols, C);
    const int gs = std::min(1024, sm_grid(static_cast<long long>(N) * kCols));
    hipLaunchKernelGGL(sm_sumsq_kernel, dim3(gs), dim3(kBlock), 0, s, sm.wide.q64.p, N, sm.wide.L, C, sm.ssq.p);
    XFE_CHECK(hipGetLastError());
    hp.assign(static_cast<size_t>(gs) * kCols, 0.0);
    XFE_CHECK(hipMemcpyAsync(hp.data(), sm.ssq.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    for (int b = 0; b < gs; ++b)
        for (int c = 0; c < C; ++c) resid[c] += hp[static_cast<size_t>(b) * kCols + c];
    out.resid_sq = resid;
    // back to the caller's row order
    hipLaunchKernelGGL(sm_scatter_cols_kernel, dim3(sm_grid(static_cast<long long>(n) * C)), dim3(kBlock), 0, s,
                       sm.Vs.p, sm.ld, sm.pperm.p, n, C, V, ld);
    XFE_CHECK(hipGetLastError());
    XFE_CHECK(hipStreamSynchronize(s));
    out.unknowns = N;
    out.primary = p;
}

}  // namespace xfe
