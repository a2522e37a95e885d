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
        // FP32 rows are kept in registers between the two loops; FP64 rows are
        // formed again (a 32-entry FP64 array goes to scratch memory)
        constexpr int kCacheRows = sizeof(T) == 4 ? kRowCache : 0;
        T cache[kCacheRows > 0 ? kCacheRows : 1];
        T sum = T(0);
        // u_i - u_i0, formed coefficient by coefficient (differences of the
        // level values first; a slope term of one level as (z_i - z_i0) x), so
        // the FP32 result keeps the accuracy of small differences on the
        // slow, smooth modes instead of rounding each u to |u|
        // the level's first data value, read once: the STORE pass writes the
        // transform over V in place (row i0 first), and rows past the register
        // cache are formed again after that
        double v0 = 0.0;
        if constexpr (HAS_V) v0 = a.V[static_cast<size_t>(col) * a.ld + i0];
        auto row_delta = [&](int i) {
            T d = T(0);
            if constexpr (HAS_V) d = static_cast<T>(a.V[static_cast<size_t>(col) * a.ld + i] - v0);
            if (a.x && i != i0) {
                for (int r = 0; r < a.NR; ++r) {
                    const int gi = a.gid[r][i], g0 = a.gid[r][i0];
                    const T xi = a.x[static_cast<size_t>(a.off[r] + gi) * a.L + col];
                    const T x0 = a.x[static_cast<size_t>(a.off[r] + g0) * a.L + col];
                    T t;
                    if (a.mult[r]) {
                        const T mi = a.mult[r][i], m0 = a.mult[r][i0];
                        t = gi == g0 ? (mi - m0) * xi : mi * xi - m0 * x0;
                    } else {
                        t = xi - x0;
                    }
                    d += a.sign * t;
                }
            }
            return d;
        };
        for (int i = i0; i < i1; ++i) {
            const T u = row_delta(i);
            if (i - i0 < kCacheRows) cache[i - i0] = u;
            sum += u;
        }
        const T mean = sum / static_cast<T>(i1 - i0);
        double tot = 0.0;
        for (int i = i0; i < i1; ++i) {
            const T u = i - i0 < kCacheRows ? cache[i - i0] : row_delta(i);
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
// (accumulated in FP64 for either storage type: a slope's level sums run over
// every row of a firm, and an FP32 sum of that many terms of both signs loses
// the digits the solve needs)
template <typename T>
__global__ void __launch_bounds__(kBlock)
sm_chunk_kernel(const int* __restrict__ perm, const int* __restrict__ cstart, int nchunks, const int* __restrict__ cend,
                const T* __restrict__ mult, const T* __restrict__ R, int stride, int C, T* __restrict__ part) {
    const int col = threadIdx.x % kCols, slot = threadIdx.x / kCols;
    if (col >= C) return;
    for (int k = blockIdx.x * kSlotsPerBlock + slot; k < nchunks; k += gridDim.x * kSlotsPerBlock) {
        double acc = 0.0;
        for (int q = cstart[k]; q < cend[k]; ++q) {
            const int i = perm[q];
            const double m = mult ? static_cast<double>(mult[i]) : 1.0;
            acc += m * static_cast<double>(R[static_cast<size_t>(i) * stride + col]);
        }
        part[static_cast<size_t>(k) * kCols + col] = static_cast<T>(acc);
    }
}

// out[(off + g) L + l] = sum of the level's chunk totals (l < C), 0 for l >= C;
// levels of more than big chunks are left to sm_level_big_kernel
template <typename T>
__global__ void __launch_bounds__(kBlock)
sm_level_kernel(const int* __restrict__ cptr, const T* __restrict__ part, int G, int off, int L, int C,
                T* __restrict__ out, int big) {
    const long long total = static_cast<long long>(G) * L;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int g = static_cast<int>(j / L), l = static_cast<int>(j % L);
        const int k0 = cptr[g], k1 = cptr[g + 1];
        if (k1 - k0 > big) continue;
        double acc = 0.0;
        if (l < C)
            for (int k = k0; k < k1; ++k) acc += static_cast<double>(part[static_cast<size_t>(k) * kCols + l]);
        out[static_cast<size_t>(off + g) * L + l] = static_cast<T>(acc);
    }
}

// The levels with many chunks (large firms), one block per level: 8 groups of
// kCols threads each sum every 8th chunk, then the 8 group sums are added in
// order (the same order on every run).
constexpr int kLevelGroups = kBlock / kCols;
template <typename T>
__global__ void __launch_bounds__(kBlock)
sm_level_big_kernel(const int* __restrict__ list, int nbig, const int* __restrict__ cptr, const T* __restrict__ part,
                    int off, int L, int C, T* __restrict__ out) {
    __shared__ double red[kBlock];
    const int l = threadIdx.x % kCols, grp = threadIdx.x / kCols;
    for (int b = blockIdx.x; b < nbig; b += gridDim.x) {
        const int g = list[b];
        const int k0 = cptr[g], k1 = cptr[g + 1];
        double acc = 0.0;
        if (l < C)
            for (int k = k0 + grp; k < k1; k += kLevelGroups)
                acc += static_cast<double>(part[static_cast<size_t>(k) * kCols + l]);
        red[threadIdx.x] = acc;
        __syncthreads();
        if (grp == 0) {
            double t = 0.0;
            for (int q = 0; q < kLevelGroups; ++q) t += red[q * kCols + l];
            if (l < L) out[static_cast<size_t>(off + g) * L + l] = static_cast<T>(l < C ? t : 0.0);
        }
        for (int l2 = kCols + threadIdx.x; l2 < L; l2 += kBlock) out[static_cast<size_t>(off + g) * L + l2] = T(0);
        __syncthreads();
    }
}

// the levels of more than big chunks, in any order (each is summed on its own)
__global__ void __launch_bounds__(kBlock)
sm_big_list_kernel(const int* cptr, int G, int big, int* list, int* count) {
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < G; g += gridDim.x * kBlock)
        if (cptr[g + 1] - cptr[g] > big) list[atomicAdd(count, 1)] = g;
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

// connected components of the bipartite graph (p level, block level) of the
// rows: parent over Gp + Gb nodes, hooking the larger root under the smaller
// until no edge changes a root, then full compression
__global__ void __launch_bounds__(kBlock)
sm_hook_kernel(const int* a, const int* b, int n, int Gp, int* parent, int* changed) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        int u = a[i], v = Gp + b[i];
        while (parent[u] != u) u = parent[u];
        while (parent[v] != v) v = parent[v];
        if (u == v) continue;
        // another hook of the same root may win the race: the round is
        // repeated while any edge still joins two roots
        atomicMin(&parent[max(u, v)], min(u, v));
        *changed = 1;
    }
}

__global__ void __launch_bounds__(kBlock) sm_compress_kernel(int* parent, int m) {
    for (int x = blockIdx.x * kBlock + threadIdx.x; x < m; x += gridDim.x * kBlock) {
        int r = x;
        while (parent[r] != r) r = parent[r];
        parent[x] = r;
    }
}

// the pinned level of each component: the smallest block level in it
__global__ void __launch_bounds__(kBlock) sm_pin_min_kernel(const int* parent, int Gp, int Gb, int* pin_of_root) {
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < Gb; g += gridDim.x * kBlock)
        atomicMin(&pin_of_root[parent[Gp + g]], g);
}

__global__ void __launch_bounds__(kBlock)
sm_pin_flag_kernel(const int* parent, const int* pin_of_root, const int* cnt, int Gp, int Gb, int* flag) {
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < Gb; g += gridDim.x * kBlock)
        flag[g] = cnt[g] > 0 && pin_of_root[parent[Gp + g]] == g ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock)
sm_pin_list_kernel(const int* flag, const int* pos, int Gb, int off, int base, int* list) {
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < Gb; g += gridDim.x * kBlock)
        if (flag[g]) list[base + pos[g]] = off + g;
}

// q (and b, minv) at the pinned unknowns: zero in every lane
template <typename T>
__global__ void __launch_bounds__(kBlock) sm_pin_zero_kernel(const int* list, int np, int L, T* q) {
    const long long total = static_cast<long long>(np) * L;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock)
        q[static_cast<size_t>(list[j / L]) * L + j % L] = T(0);
}

__global__ void __launch_bounds__(kBlock) sm_pin_minv_kernel(const int* list, int np, float* minv) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < np; j += gridDim.x * kBlock) minv[list[j]] = 0.0f;
}

// ---- per-group block preconditioner: the unknowns whose levels nest in one
// level of the slope's grouping (a firm's year cells and its slope) form a
// dense block of S, assembled exactly in FP64 and inverted

constexpr int kBlkMax = 64;  // largest block inverted; larger groups stay on Jacobi

// block of each unknown: the grouping level of one of its rows (every row of
// a nested level has the same); unknowns without rows keep -1
__global__ void __launch_bounds__(kBlock)
sm_blk_of_kernel(const int* gid_r, const int* gid_b, int n, int off, int* blk) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) blk[off + gid_r[i]] = gid_b[i];
}

// sort keys: the block, or nb (past the last block) for unknowns outside the
// blocks (no rows, no Jacobi weight: pinned or without coupling)
__global__ void __launch_bounds__(kBlock)
sm_blk_key_kernel(const int* blk, const float* minv, int N, int nb, int* key, int* val) {
    for (int u = blockIdx.x * kBlock + threadIdx.x; u < N; u += gridDim.x * kBlock) {
        key[u] = (blk[u] >= 0 && minv[u] != 0.0f) ? blk[u] : nb;
        val[u] = u;
    }
}

// per sorted position: the unknown's local index inside its block
__global__ void __launch_bounds__(kBlock)
sm_blk_local_kernel(const int* skey, const int* sval, const int* bptr, int N, int nb, int* loc) {
    for (int q = blockIdx.x * kBlock + threadIdx.x; q < N; q += gridDim.x * kBlock) {
        const int b = skey[q];
        loc[sval[q]] = b < nb ? q - bptr[b] : -1;
    }
}

// matrix offsets: k^2 for blocks of 2..kBlkMax unknowns, 0 otherwise
__global__ void __launch_bounds__(kBlock) sm_blk_sq_kernel(const int* bcnt, int nb, int* sq) {
    for (int b = blockIdx.x * kBlock + threadIdx.x; b < nb; b += gridDim.x * kBlock) {
        const int k = bcnt[b];
        sq[b] = (k >= 2 && k <= kBlkMax) ? k * k : 0;
    }
}

struct BlkArgs {
    const int* pofs = nullptr;
    int Np = 0, NR = 0;
    const int* gid[kSmMaxR] = {};
    const double* mult[kSmMaxR] = {};
    int off[kSmMaxR] = {};
    const int* blk = nullptr;    // block of each unknown
    const int* loc = nullptr;    // local index (-1: outside the blocks)
    const int* bcnt = nullptr;   // unknowns per block
    const int* moff = nullptr;   // matrix offset per block
    double* mat = nullptr;
};

// S restricted to the blocks: per p level (thread), the rows' own terms
// m_u m_v and the elimination term -a_u a_v / n (a = sums over the p level's
// rows of m per unknown), for pairs of unknowns of one block
__global__ void __launch_bounds__(kBlock) sm_blk_assemble_kernel(BlkArgs a) {
    constexpr int kMaxE = 96;
    for (int w = blockIdx.x * kBlock + threadIdx.x; w < a.Np; w += gridDim.x * kBlock) {
        const int i0 = a.pofs[w], i1 = a.pofs[w + 1];
        if (i1 <= i0) continue;
        const double inv_n = 1.0 / static_cast<double>(i1 - i0);
        int eu[kMaxE];
        double ea[kMaxE];
        int ne = 0;
        bool overflow = false;
        for (int i = i0; i < i1; ++i) {
            // the row's own terms (its unknowns: one per remaining block)
            int ru[kSmMaxR];
            double rm[kSmMaxR];
            int rn = 0;
            for (int r = 0; r < a.NR; ++r) {
                const int u = a.off[r] + a.gid[r][i];
                if (a.loc[u] < 0) continue;
                ru[rn] = u;
                rm[rn] = a.mult[r] ? a.mult[r][i] : 1.0;
                ++rn;
                // accumulate a_u
                int k = 0;
                while (k < ne && eu[k] != u) ++k;
                if (k == ne) {
                    if (ne == kMaxE) {
                        overflow = true;
                        continue;
                    }
                    eu[ne] = u;
                    ea[ne] = 0.0;
                    ++ne;
                }
                ea[k] += rm[rn - 1];
            }
            for (int x = 0; x < rn; ++x)
                for (int y = 0; y < rn; ++y) {
                    if (a.blk[ru[x]] != a.blk[ru[y]]) continue;
                    const int b = a.blk[ru[x]];
                    const int k = a.bcnt[b];
                    if (k < 2 || k > kBlkMax) continue;
                    atomicAdd(&a.mat[a.moff[b] + a.loc[ru[x]] * k + a.loc[ru[y]]], rm[x] * rm[y]);
                }
        }
        if (overflow) continue;  // a p level touching more unknowns than kMaxE: its correction is skipped
        for (int x = 0; x < ne; ++x)
            for (int y = 0; y < ne; ++y) {
                if (a.blk[eu[x]] != a.blk[eu[y]]) continue;
                const int b = a.blk[eu[x]];
                const int k = a.bcnt[b];
                if (k < 2 || k > kBlkMax) continue;
                atomicAdd(&a.mat[a.moff[b] + a.loc[eu[x]] * k + a.loc[eu[y]]], -ea[x] * ea[y] * inv_n);
            }
    }
}

// Gauss-Jordan inverse of each block (symmetric positive semi-definite: no
// pivoting), one wave per block in LDS.  A pivot at or below kBlkDrop of the
// largest diagonal is a zero: the leading unknowns up to it carry a null
// vector of S (a slope collinear with its firm's cells), so the unknown is
// left out of the block (its row and column of the inverse zero: the solve
// keeps it at zero, which leaves the attainable residual unchanged).  A pivot
// between that and kBlkFail marks the block too ill-conditioned for an FP32
// inverse: Jacobi keeps its unknowns.
constexpr double kBlkDrop = 1e-10;
constexpr double kBlkFail = 1e-6;
__global__ void __launch_bounds__(64)
sm_blk_invert_kernel(const int* bcnt, const int* moff, int nb, const double* mat, float* inv, double* inv64, int* ok,
                     double fail) {
    __shared__ double A[kBlkMax][kBlkMax + 1];
    __shared__ double Y[kBlkMax][kBlkMax + 1];
    __shared__ int bad;
    const int t = threadIdx.x;
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int k = bcnt[b];
        if (k < 2 || k > kBlkMax) {
            if (t == 0) ok[b] = 0;
            continue;
        }
        const double* M = mat + moff[b];
        for (int e = t; e < k * k; e += 64) {
            A[e / k][e % k] = M[e];
            Y[e / k][e % k] = (e / k == e % k) ? 1.0 : 0.0;
        }
        if (t == 0) bad = 0;
        __syncthreads();
        double dmax = 0.0;
        for (int j = 0; j < k; ++j) dmax = fmax(dmax, fabs(A[j][j]));
        for (int c = 0; c < k; ++c) {
            const double piv = A[c][c];
            if (!(piv > kBlkDrop * dmax)) {
                __syncthreads();
                if (t < k) {
                    A[c][t] = 0.0;
                    A[t][c] = 0.0;
                    Y[c][t] = 0.0;
                }
                __syncthreads();
                continue;
            }
            if (!(piv > fail * dmax)) {
                if (t == 0) bad = 1;
                break;
            }
            __syncthreads();
            // scale the pivot row
            if (t < k) {
                A[c][t] /= piv;
                Y[c][t] /= piv;
            }
            __syncthreads();
            // eliminate the other rows (thread t owns row t)
            if (t < k && t != c) {
                const double f = A[t][c];
                for (int j = 0; j < k; ++j) {
                    A[t][j] -= f * A[c][j];
                    Y[t][j] -= f * Y[c][j];
                }
            }
            __syncthreads();
        }
        __syncthreads();
        const int good = bad == 0;
        if (good)
            for (int e = t; e < k * k; e += 64) {
                inv[moff[b] + e] = static_cast<float>(Y[e / k][e % k]);
                inv64[moff[b] + e] = Y[e / k][e % k];
            }
        if (t == 0) ok[b] = good;
        __syncthreads();
    }
}

// ---- FP64 block-PCG continuation (level-major N x L vectors, lanes < C)

// part[b * L + l] = sum over block b's rows of a[u, l] b[u, l] (kBlock % L == 0)
__global__ void __launch_bounds__(kBlock)
sm_dot64_kernel(const double* a, const double* b, int N, int L, double* part) {
    __shared__ double red[kBlock];
    const int l = threadIdx.x % L, rows = kBlock / L;
    double acc = 0.0;
    for (long long u = static_cast<long long>(blockIdx.x) * rows + threadIdx.x / L; u < N;
         u += static_cast<long long>(gridDim.x) * rows) {
        const size_t i = static_cast<size_t>(u) * L + l;
        acc += a[i] * b[i];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < L) {
        double t = 0.0;
        for (int k = 0; k < rows; ++k) t += red[k * L + threadIdx.x];
        part[static_cast<size_t>(blockIdx.x) * L + threadIdx.x] = t;
    }
}

// x += alpha p, r -= alpha q (per lane alpha)
__global__ void __launch_bounds__(kBlock)
sm_update64_kernel(double* x, double* r, const double* p, const double* q, const double* alpha, long long total,
                   int L) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const double a = alpha[i % L];
        if (a == 0.0) continue;
        x[i] += a * p[i];
        r[i] -= a * q[i];
    }
}

// z = minv r (Jacobi; the block rows are overwritten by the block solve)
__global__ void __launch_bounds__(kBlock)
sm_jacobi64_kernel(const double* r, const float* minv, double* z, long long total, int L) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock)
        z[i] = static_cast<double>(minv[i / L]) * r[i];
}

// p = z + beta p (per lane beta)
__global__ void __launch_bounds__(kBlock)
sm_dir64_kernel(double* p, const double* z, const double* beta, long long total, int L) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock)
        p[i] = z[i] + beta[i % L] * p[i];
}

// r = b - q
__global__ void __launch_bounds__(kBlock)
sm_res64_kernel(double* r, const double* b, const double* q, long long total) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock)
        r[i] = b[i] - q[i];
}

// z = B^-1 r on the blocks, one thread per (blocked unknown, lane): the
// unknown at sorted position q (block skey[q], row q - bptr[b] of its inverse)
// gets the dot of that row with the block's r values
template <typename T>
__global__ void __launch_bounds__(kBlock)
sm_blk_rows_kernel(const int* skey, const int* bptr, const int* bcnt, const int* moff, const int* ok, const int* sval,
                   int nin, int L, int C, const T* inv, const T* r, T* z) {
    const long long total = static_cast<long long>(nin) * L;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int q = static_cast<int>(j / L), l = static_cast<int>(j % L);
        const int b = skey[q];
        if (l >= C || !ok[b]) continue;
        const int k = bcnt[b], q0 = bptr[b];
        const T* row = inv + moff[b] + static_cast<size_t>(q - q0) * k;
        const int* u = sval + q0;
        T acc = T(0);
        for (int e = 0; e < k; ++e) acc += row[e] * r[static_cast<size_t>(u[e]) * L + l];
        z[static_cast<size_t>(sval[q]) * L + l] = acc;
    }
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

constexpr int kBigChunks = 16;  // levels of more chunks are summed by a block each

struct Grouping {
    int d = 0;
    int G = 0;
    int nchunks = 0;
    int nbig = 0;
    DevBuf<int> gid, perm, cnt, ptr, nch, cptr, cstart, cend, big, nbig_d;
};

}  // namespace

struct SlopeMixed final : WideOp {
    hipStream_t s = nullptr;
    std::vector<int> dropped;  // dimensions whose intercept the last run left out
    int n = 0, C = 0, Np = 0, NR = 0, N = 0, ld = 0;
    std::vector<RDim> rd;
    std::vector<Grouping> gr;
    DevBuf<int> pperm, pkey, pcnt, pofs, tmp_int, flag, uf, pin_root, pin_flag, pin_pos, pins;
    int npins = 0;
    std::vector<const int*> src_gid;  // the caller's id arrays of the last run, per dimension
    bool transform_in_r64 = false;    // R64 holds the last run's transform rows (rows sorted by p)
    // block preconditioner
    bool blk_on = false;
    int nblk = 0;
    int nin = 0;  // unknowns inside the blocks (sorted positions 0 .. nin - 1)
    DevBuf<int> blk, bkey, bval, skey, sval, bcnt, bptr, bsq, moff, loc, bok;
    DevBuf<double> bmat, binv64;
    DevBuf<float> binv;
    // FP64 continuation vectors (N x L) and per-lane scalars
    DevBuf<double> c_r, c_z, c_p, c_q, c_part, c_scal;
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
                               dim3(kBlock), 0, s, g.cptr.p, part, g.G, rd[r].off, wide.L, cols, out, kBigChunks);
            if (g.nbig > 0)
                hipLaunchKernelGGL((sm_level_big_kernel<T>), dim3(std::min(g.nbig, 65535)), dim3(kBlock), 0, s,
                                   g.big.p, g.nbig, g.cptr.p, part, rd[r].off, wide.L, cols, out);
            XFE_CHECK(hipGetLastError());
        }
    }

    bool has_blocks() const override { return blk_on; }

    void precondition(const float* r, float* z) override {
        hipLaunchKernelGGL((sm_blk_rows_kernel<float>), dim3(sm_grid(static_cast<long long>(nin) * wide.L)),
                           dim3(kBlock), 0, s, skey.p, bptr.p, bcnt.p, moff.p, bok.p, sval.p, nin, wide.L, C, binv.p, r,
                           z);
        XFE_CHECK(hipGetLastError());
    }

    void apply(const float* p, float* q) override {
        SmArgs<float> a = args<float>(p, R32.p);
        a.sign = 1.0f;
        const int slot = profile_begin(s);
        hipLaunchKernelGGL((sm_apply_kernel<float, false, false>), dim3(gridw), dim3(kBlock), 0, s, a);
        XFE_CHECK(hipGetLastError());
        // per row: the ids and slope values of the remaining blocks, one C-float
        // coefficient row per block, the C-float row of R; the offsets of p
        int nslope = 0;
        for (int r = 0; r < NR; ++r) nslope += rd[r].slope ? 1 : 0;
        profile_end(slot, s, static_cast<double>(n) * (4.0 * NR + 4.0 * nslope + 4.0 * C * (NR + 1)) + 4.0 * Np);
        reduce<float>(R32.p, part32.p, q, kCols, C);
        pin_zero(q);
    }

    template <typename T>
    void pin_zero(T* q) {
        if (npins <= 0) return;
        hipLaunchKernelGGL((sm_pin_zero_kernel<T>), dim3(sm_grid(static_cast<long long>(npins) * wide.L)),
                           dim3(kBlock), 0, s, pins.p, npins, wide.L, q);
        XFE_CHECK(hipGetLastError());
    }

    void apply(const double* p, double* q) override {
        SmArgs<double> a = args<double>(p, R64.p);
        a.sign = 1.0;
        hipLaunchKernelGGL((sm_apply_kernel<double, false, false>), dim3(gridw), dim3(kBlock), 0, s, a);
        XFE_CHECK(hipGetLastError());
        reduce<double>(R64.p, part64.p, q, kCols, C);
        pin_zero(q);
    }

    // Per-lane sums of a . b over the N levels (lanes < C).
    void dots(const double* a, const double* b, std::vector<double>& out) {
        const int L = wide.L;
        const int g = std::min(1024, std::max(1, (N + kBlock / L - 1) / (kBlock / L)));
        c_part.ensure(static_cast<size_t>(g) * L);
        hipLaunchKernelGGL(sm_dot64_kernel, dim3(g), dim3(kBlock), 0, s, a, b, N, L, c_part.p);
        XFE_CHECK(hipGetLastError());
        std::vector<double> h(static_cast<size_t>(g) * L);
        XFE_CHECK(hipMemcpyAsync(h.data(), c_part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        out.assign(L, 0.0);
        for (int b2 = 0; b2 < g; ++b2)
            for (int l = 0; l < L; ++l) out[l] += h[static_cast<size_t>(b2) * L + l];
    }

    void precondition64(const double* r, double* z) {
        const long long total = static_cast<long long>(N) * wide.L;
        hipLaunchKernelGGL(sm_jacobi64_kernel, dim3(sm_grid(total)), dim3(kBlock), 0, s, r, minv.p, z, total, wide.L);
        if (blk_on)
            hipLaunchKernelGGL((sm_blk_rows_kernel<double>), dim3(sm_grid(static_cast<long long>(nin) * wide.L)),
                               dim3(kBlock), 0, s, skey.p, bptr.p, bcnt.p, moff.p, bok.p, sval.p, nin, wide.L, C,
                               binv64.p, r, z);
        XFE_CHECK(hipGetLastError());
    }

    // FP64 PCG (Jacobi + the firm blocks) on the wide solver's scaled system
    // (b64 of unit norm per lane) from its iterate x64, for the lanes the FP32
    // runs left unconverged; the residual is recomputed from x every 128
    // iterations.  A lane converges at |b - S x| <= tol, or at the end within
    // accept.  Marks the converged lanes in the wide solver; returns the
    // iterations run.
    int fp64_continue(double tol, double accept, int max_iter) {
        const int L = wide.L;
        const long long total = static_cast<long long>(N) * L;
        for (DevBuf<double>* b : {&c_r, &c_z, &c_p, &c_q}) b->ensure(static_cast<size_t>(total));
        c_scal.ensure(2 * static_cast<size_t>(L));
        std::vector<char> act(L, 0), ok(L, 0);
        std::vector<int> extra(L, 0);
        int nact = 0;
        for (int c = 0; c < C; ++c)
            if (wide.lane_iterations(c) < 0) {
                act[c] = 1;
                ++nact;
            }
        if (nact == 0) return 0;
        auto residual = [&]() {
            apply(wide.x64.p, c_q.p);
            hipLaunchKernelGGL(sm_res64_kernel, dim3(sm_grid(total)), dim3(kBlock), 0, s, c_r.p, wide.b64.p, c_q.p,
                               total);
            XFE_CHECK(hipGetLastError());
        };
        const bool trace = std::getenv("XFE_SLOPE_TRACE") != nullptr;
        residual();
        precondition64(c_r.p, c_z.p);
        XFE_CHECK(hipMemcpyAsync(c_p.p, c_z.p, sizeof(double) * total, hipMemcpyDeviceToDevice, s));
        std::vector<double> rz, rr, pq, coef(2 * static_cast<size_t>(L), 0.0);
        dots(c_r.p, c_z.p, rz);
        dots(c_r.p, c_r.p, rr);
        int it = 0;
        for (it = 1; it <= max_iter && nact > 0; ++it) {
            apply(c_p.p, c_q.p);
            dots(c_p.p, c_q.p, pq);
            for (int l = 0; l < L; ++l) {
                double a = 0.0;
                if (act[l] && pq[l] > 0.0 && rz[l] > 0.0) a = rz[l] / pq[l];
                coef[l] = a;
            }
            XFE_CHECK(hipMemcpyAsync(c_scal.p, coef.data(), sizeof(double) * L, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(sm_update64_kernel, dim3(sm_grid(total)), dim3(kBlock), 0, s, wide.x64.p, c_r.p, c_p.p,
                               c_q.p, c_scal.p, total, L);
            XFE_CHECK(hipGetLastError());
            if (it % 128 == 0) residual();
            precondition64(c_r.p, c_z.p);
            std::vector<double> rz_new;
            dots(c_r.p, c_z.p, rz_new);
            dots(c_r.p, c_r.p, rr);
            for (int l = 0; l < L; ++l) {
                if (!act[l]) {
                    coef[L + l] = 0.0;
                    continue;
                }
                ++extra[l];
                if (std::sqrt(std::max(0.0, rr[l])) <= tol || !(pq[l] > 0.0)) {
                    act[l] = 0;
                    --nact;
                    ok[l] = std::sqrt(std::max(0.0, rr[l])) <= accept;
                }
                coef[L + l] = rz[l] > 0.0 ? rz_new[l] / rz[l] : 0.0;
            }
            rz = rz_new;
            XFE_CHECK(hipMemcpyAsync(c_scal.p + L, coef.data() + L, sizeof(double) * L, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(sm_dir64_kernel, dim3(sm_grid(total)), dim3(kBlock), 0, s, c_p.p, c_z.p, c_scal.p + L,
                               total, L);
            XFE_CHECK(hipGetLastError());
            if (trace && it % 100 == 0) {
                double worst = 0.0;
                for (int l = 0; l < L; ++l)
                    if (act[l]) worst = std::max(worst, std::sqrt(std::max(0.0, rr[l])));
                std::fprintf(stderr, "[slope fp64 continuation] iteration %d: %d lanes active, worst |r| %.3e\n", it,
                             nact, worst);
            }
        }
        // the lanes still running: their true residual decides (accept)
        if (nact > 0) {
            residual();
            dots(c_r.p, c_r.p, rr);
            for (int l = 0; l < L; ++l)
                if (act[l]) ok[l] = std::sqrt(std::max(0.0, rr[l])) <= accept;
        }
        wide.mark_converged(L, ok, extra);
        return it - 1;
    }

    void release() {
        wide.release();
        for (auto& g : gr)
            for (DevBuf<int>* b : {&g.gid, &g.perm, &g.cnt, &g.ptr, &g.nch, &g.cptr, &g.cstart, &g.cend, &g.big,
                                   &g.nbig_d})
                b->release();
        for (DevBuf<int>* b : {&pperm, &pkey, &pcnt, &pofs, &tmp_int, &flag, &uf, &pin_root, &pin_flag, &pin_pos, &pins,
                               &blk, &bkey, &bval, &skey, &sval, &bcnt, &bptr, &bsq, &moff, &loc, &bok})
            b->release();
        bmat.release();
        binv.release();
        binv64.release();
        for (DevBuf<double>* b : {&c_r, &c_z, &c_p, &c_q, &c_part, &c_scal}) b->release();
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
    // the dimension's grouping from the run (its slope's), when the run's R64
    // still holds the transform rows: chunked sums in a fixed order; else
    // atomics over V
    int gi = -1;
    for (int r = 0; r < sm.NR && gi < 0; ++r)
        if (sm.gr[sm.rd[r].grouping].gid.p && sm.gr[sm.rd[r].grouping].G == G && gid == sm.src_gid[sm.rd[r].d])
            gi = sm.rd[r].grouping;
    if (gi >= 0 && sm.transform_in_r64) {
        Grouping& g = sm.gr[gi];
        hipLaunchKernelGGL((sm_chunk_kernel<double>), dim3(sm_grid(static_cast<long long>(g.nchunks) * kCols)),
                           dim3(kBlock), 0, s, g.perm.p, g.cstart.p, g.nchunks, g.cend.p, nullptr, sm.R64.p, kCols, C,
                           sm.part64.p);
        hipLaunchKernelGGL((sm_level_kernel<double>), dim3(sm_grid(static_cast<long long>(G) * kCols)), dim3(kBlock), 0,
                           s, g.cptr.p, sm.part64.p, G, 0, kCols, C, sm.dl1.p, kBigChunks);
        if (g.nbig > 0)
            hipLaunchKernelGGL((sm_level_big_kernel<double>), dim3(std::min(g.nbig, 65535)), dim3(kBlock), 0, s,
                               g.big.p, g.nbig, g.cptr.p, sm.part64.p, 0, kCols, C, sm.dl1.p);
        XFE_CHECK(hipGetLastError());
    } else {
        XFE_CHECK(hipMemsetAsync(sm.dl1.p, 0, sizeof(double) * static_cast<size_t>(G) * kCols, s));
        hipLaunchKernelGGL(sm_level_atomic_kernel, dim3(sm_grid(static_cast<long long>(n) * C)), dim3(kBlock), 0, s,
                           V, n, C, ld, gid, sm.dl1.p);
    }
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
                     double tol, int max_iter, int fp64_max_iter, hipStream_t s, SlopeMixedResult& out) {
    sm.s = s;
    sm.n = n;
    sm.C = C;
    sm.src_gid = gid;
    sm.transform_in_r64 = false;
    const int D = static_cast<int>(G.size());
    const int grid = sm_grid(n);
    // nested: every level of `fine` inside one level of `coarse`
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
    // p: the standard dimension with the most levels among those that do not
    // nest inside a slope's grouping (firm x year cells inside firms): with
    // such a dimension eliminated the slope varies inside p's levels and the
    // FP32 recurrence stalls; if every standard dimension nests, the one with
    // the most levels
    int p = -1, p_any = -1;
    for (int d = 0; d < D; ++d) {
        if (mode[d] != 0) continue;
        if (p_any < 0 || G[d] > G[p_any]) p_any = d;
        bool in_slope = false;
        for (int e = 0; e < D && !in_slope; ++e)
            if (mode[e] != 0 && G[e] <= G[d] && nested_in(d, e)) in_slope = true;
        if (!in_slope && (p < 0 || G[d] > G[p])) p = d;
    }
    if (p < 0) p = p_any;
    sm.Np = G[p];
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
        // the levels with more than kBigChunks chunks
        g.big.ensure(std::max(G[d], 1));
        g.nbig_d.ensure(1);
        XFE_CHECK(hipMemsetAsync(g.nbig_d.p, 0, sizeof(int), s));
        hipLaunchKernelGGL(sm_big_list_kernel, dim3(sm_grid(G[d])), dim3(kBlock), 0, s, g.cptr.p, G[d], kBigChunks,
                           g.big.p, g.nbig_d.p);
        XFE_CHECK(hipGetLastError());
        XFE_CHECK(hipMemcpyAsync(&g.nbig, g.nbig_d.p, sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        gidx[d] = ng++;
    }
    // a slope's intercept is left out when a finer standard dimension nests in
    // the slope's grouping
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
                               sm.part64.p, g.G, sm.rd[r].off, 1, 1, k == 0 ? sm.dl1.p : sm.dl2.p, INT_MAX);
        }
        XFE_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(sm_minv_kernel, dim3(sm_grid(N)), dim3(kBlock), 0, s, sm.dl1.p, sm.dl2.p, N, sm.minv.p);
    XFE_CHECK(hipGetLastError());
    // ground: a constant over a connected component of (p, a standard block)
    // lies in p's span, so S is singular along it; one level per component of
    // each standard block is pinned to zero (its row and column of S dropped)
    sm.npins = 0;
    {
        std::vector<int> counts;
        int total_pins = 0;
        sm.pins.ensure(static_cast<size_t>(N) + 1);
        for (int r = 0; r < sm.NR; ++r) {
            if (sm.rd[r].slope) continue;
            Grouping& g = sm.gr[sm.rd[r].grouping];
            const int Gb = g.G, Gp = sm.Np, m = Gp + Gb;
            sm.uf.ensure(m);
            hipLaunchKernelGGL(sm_iota_kernel, dim3(sm_grid(m)), dim3(kBlock), 0, s, sm.uf.p, m);
            sm.flag.ensure(1);
            for (int round = 0; round < 1024; ++round) {
                XFE_CHECK(hipMemsetAsync(sm.flag.p, 0, sizeof(int), s));
                hipLaunchKernelGGL(sm_hook_kernel, dim3(grid), dim3(kBlock), 0, s, gid[p], gid[sm.rd[r].d], n, Gp,
                                   sm.uf.p, sm.flag.p);
                hipLaunchKernelGGL(sm_compress_kernel, dim3(sm_grid(m)), dim3(kBlock), 0, s, sm.uf.p, m);
                int changed = 0;
                XFE_CHECK(hipMemcpyAsync(&changed, sm.flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
                XFE_CHECK(hipStreamSynchronize(s));
                if (!changed) break;
                if (round == 1023) throw Error("slope_mixed: components did not settle");
            }
            sm.pin_root.ensure(m);
            sm.pin_flag.ensure(static_cast<size_t>(Gb) + 1);
            sm.pin_pos.ensure(static_cast<size_t>(Gb) + 1);
            hipLaunchKernelGGL(sm_fill_int_kernel, dim3(sm_grid(m)), dim3(kBlock), 0, s, sm.pin_root.p, m, INT_MAX);
            hipLaunchKernelGGL(sm_pin_min_kernel, dim3(sm_grid(Gb)), dim3(kBlock), 0, s, sm.uf.p, Gp, Gb,
                               sm.pin_root.p);
            hipLaunchKernelGGL(sm_pin_flag_kernel, dim3(sm_grid(Gb)), dim3(kBlock), 0, s, sm.uf.p, sm.pin_root.p,
                               g.cnt.p, Gp, Gb, sm.pin_flag.p);
            XFE_CHECK(hipMemsetAsync(sm.pin_flag.p + Gb, 0, sizeof(int), s));
            excl_scan(sm.tmp, sm.pin_flag.p, sm.pin_pos.p, Gb + 1, s);
            int np = 0;
            XFE_CHECK(hipMemcpyAsync(&np, sm.pin_pos.p + Gb, sizeof(int), hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
            hipLaunchKernelGGL(sm_pin_list_kernel, dim3(sm_grid(Gb)), dim3(kBlock), 0, s, sm.pin_flag.p,
                               sm.pin_pos.p, Gb, sm.rd[r].off, total_pins, sm.pins.p);
            XFE_CHECK(hipGetLastError());
            total_pins += np;
        }
        sm.npins = total_pins;
        if (total_pins > 0)
            hipLaunchKernelGGL(sm_pin_minv_kernel, dim3(sm_grid(total_pins)), dim3(kBlock), 0, s, sm.pins.p,
                               total_pins, sm.minv.p);
        XFE_CHECK(hipGetLastError());
        out.pinned = total_pins;
    }
    // blocks: the unknowns nesting in one level of a slope's grouping (its
    // slope, the cells of a firm) form one exactly inverted block of S
    sm.blk_on = false;
    sm.nblk = 0;
    int ds = -1;
    for (int r = 0; r < sm.NR && ds < 0; ++r)
        if (sm.rd[r].slope) ds = sm.rd[r].d;
    // (XFE_SLOPE_JACOBI set: Jacobi only, a diagnostic and the fallback test's
    // way to make a run stop unconverged)
    if (ds >= 0 && !std::getenv("XFE_SLOPE_JACOBI")) {
        std::vector<int> members;
        for (int r = 0; r < sm.NR; ++r) {
            const int d = sm.rd[r].d;
            if (d == ds || (G[d] >= G[ds] && nested_in(d, ds))) members.push_back(r);
        }
        if (members.size() >= 2) {
            const int nb = G[ds];
            sm.blk.ensure(N);
            hipLaunchKernelGGL(sm_fill_int_kernel, dim3(sm_grid(N)), dim3(kBlock), 0, s, sm.blk.p, N, -1);
            for (int r : members)
                hipLaunchKernelGGL(sm_blk_of_kernel, dim3(grid), dim3(kBlock), 0, s, gid[sm.rd[r].d], gid[ds], n,
                                   sm.rd[r].off, sm.blk.p);
            sm.bkey.ensure(N); sm.bval.ensure(N); sm.skey.ensure(N); sm.sval.ensure(N); sm.loc.ensure(N);
            hipLaunchKernelGGL(sm_blk_key_kernel, dim3(sm_grid(N)), dim3(kBlock), 0, s, sm.blk.p, sm.minv.p, N, nb,
                               sm.bkey.p, sm.bval.p);
            XFE_CHECK(hipGetLastError());
            radix_pairs<int, int>(sm.tmp, sm.bkey.p, sm.skey.p, sm.bval.p, sm.sval.p, N, bits_of(nb + 1), s);
            offsets(sm.tmp, sm.skey.p, N, nb + 1, sm.bcnt, sm.bptr, s);
            hipLaunchKernelGGL(sm_blk_local_kernel, dim3(sm_grid(N)), dim3(kBlock), 0, s, sm.skey.p, sm.sval.p,
                               sm.bptr.p, N, nb, sm.loc.p);
            sm.bsq.ensure(static_cast<size_t>(nb) + 1);
            sm.moff.ensure(static_cast<size_t>(nb) + 1);
            hipLaunchKernelGGL(sm_blk_sq_kernel, dim3(sm_grid(nb)), dim3(kBlock), 0, s, sm.bcnt.p, nb, sm.bsq.p);
            XFE_CHECK(hipMemsetAsync(sm.bsq.p + nb, 0, sizeof(int), s));
            XFE_CHECK(hipGetLastError());
            excl_scan(sm.tmp, sm.bsq.p, sm.moff.p, nb + 1, s);
            int total = 0;
            XFE_CHECK(hipMemcpyAsync(&total, sm.moff.p + nb, sizeof(int), hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipMemcpyAsync(&sm.nin, sm.bptr.p + nb, sizeof(int), hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
            if (total > 0) {
                sm.bmat.ensure(total);
                sm.binv.ensure(total);
                sm.binv64.ensure(total);
                sm.bok.ensure(nb);
                XFE_CHECK(hipMemsetAsync(sm.bmat.p, 0, sizeof(double) * total, s));
                BlkArgs a;
                a.pofs = sm.pofs.p; a.Np = sm.Np; a.NR = sm.NR;
                for (int r = 0; r < sm.NR; ++r) {
                    a.gid[r] = sm.gr[sm.rd[r].grouping].gid.p;
                    a.mult[r] = sm.rd[r].slope ? sm.z64[r].p : nullptr;
                    a.off[r] = sm.rd[r].off;
                }
                a.blk = sm.blk.p; a.loc = sm.loc.p; a.bcnt = sm.bcnt.p; a.moff = sm.moff.p; a.mat = sm.bmat.p;
                hipLaunchKernelGGL(sm_blk_assemble_kernel, dim3(sm_grid(sm.Np)), dim3(kBlock), 0, s, a);
                hipLaunchKernelGGL(sm_blk_invert_kernel, dim3(std::min(nb, 8192)), dim3(64), 0, s, sm.bcnt.p,
                                   sm.moff.p, nb, sm.bmat.p, sm.binv.p, sm.binv64.p, sm.bok.p, kBlkFail);
                XFE_CHECK(hipGetLastError());
                sm.blk_on = true;
                sm.nblk = nb;
            }
        }
    }
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
        sm.pin_zero(sm.wide.b64.p);
    }
    sm.wide.trace = std::getenv("XFE_SLOPE_TRACE") ? "slope mixed" : nullptr;
    sm.wide.restart_each_replacement = true;
    sm.wide.solve_device(C, tol, max_iter, std::max(tol, 1e-9));
    // lanes the FP32 runs left unconverged continue in FP64 from their iterate
    // (XFE_SLOPE_NO_CONTINUATION set: none -- the fallback test's way to reach
    // the FP64 slope path)
    const bool cont = fp64_max_iter > 0 && !std::getenv("XFE_SLOPE_NO_CONTINUATION");
    out.fp64_iterations = cont ? sm.fp64_continue(tol, std::max(tol, 1e-9), fp64_max_iter) : 0;
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
        std::fprintf(stderr,
                     "[slope mixed] N %d NR %d p %d dropped %d pinned %d blocks %d passes %d replacements %d "
                     "fp64 iterations %d\n",
                     N, sm.NR, p, out.intercepts_dropped, out.pinned, sm.blk_on ? sm.nblk : 0, out.passes,
                     out.replacements, out.fp64_iterations);
        for (int r = 0; r < sm.NR; ++r)
            std::fprintf(stderr, "[slope mixed]   block %d: dim %d slope %d levels %d off %d chunks %d\n", r,
                         sm.rd[r].d, sm.rd[r].slope ? 1 : 0, sm.rd[r].G, sm.rd[r].off,
                         sm.gr[sm.rd[r].grouping].nchunks);
        for (int c = 0; c < C; ++c) std::fprintf(stderr, "[slope mixed]   lane %d: %s\n", c, sm.wide.lane_report(c).c_str());
    }
    out.unknowns = N;
    out.primary = p;
    // a lane that did not converge leaves the caller's columns untouched (the
    // caller falls back to the FP64 path on them)
    if (!out.converged) {
        out.resid_sq.assign(C, 0.0);
        XFE_CHECK(hipStreamSynchronize(s));
        return;
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
        sm.transform_in_r64 = true;
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
}

}  // namespace xfe
