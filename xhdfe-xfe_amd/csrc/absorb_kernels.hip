// Device kernels of the MI355X fixed-effects absorber (gfx950, wave64).
//
// Data layout in HBM (all FP64 here):
//   V        n x C column-major, leading dimension ld (ld % 64 == 0): column 0
//            is y, columns 1..k are X.  Rows are permuted once so that the
//            dimension with the most levels is sorted (its levels are runs).
//   gid[d]   int32 compact level id per (permuted) row, one array per dim.
//   sums[d]  G_d x C level-major accumulators (row of a level is contiguous,
//            so one gather of a row's C means touches one or two cache lines).
//   means[d] G_d x C level-major means, inv[d] G_d inverse weight sums.
//
// One FE projection v <- v - P_d v is split into (a) a reduce of v into
// sums[d] and (b) the subtraction of means[d].  The reduce of the next
// dimension in a sweep is fused into the subtraction pass of the previous
// one, so a sweep over D dimensions costs D streaming passes over V (plus the
// first reduce), each reading and writing n*C values once.
//
// Reduction inside a pass: rows of a wave with equal consecutive level ids
// (runs; every level of the sorted dimension is one run, and the
// worker-sorted order also makes firm ids come in runs) are combined with a
// segmented suffix scan over wave64 lanes; only the run heads issue one FP64
// atomic per column.  Dimensions whose accumulator fits a block's LDS budget
// (e.g. 36 years x C) are privatised in LDS and flushed once per block.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "absorb_kernels.h"
#include "wave_ops.h"

namespace xfe {


// ---------------------------------------------------------------------------
// The fused streaming pass.  For each row i and column c in [c0, c0+CC):
//   v = SRC==2 ? 1 : SRC==1 ? P - coef (P - Q) : src   (IT extrapolation fused in)
//   SRC==1: xout = v
//   if gid_a: v -= means_a[gid_a[i], c]                  (finish projection a)
//   if dst:  dst = v
//   reductions: sum v^2 (NQ 1), or IT stats against (it_x, it_gx) (NQ 3)
//   if gid_r: sums_r[gid_r[i], c] += w_i v               (start projection r)
//
// One row per thread per grid-stride step: loads of the whole block row are
// issued back to back and coalesce into 512-byte wave accesses per column.
// RM selects how the level sums of projection r are accumulated:
//   kRedLds    -- LDS-privatised, flushed once per block (small dimensions);
//   kRedSorted -- the level-sorted dimension: a run that neither touches lane
//                 0 nor lane 63 is a whole level owned by this wave and is
//                 written with a plain store; other runs use an atomic;
//   kRedGlobal -- FP64 atomics from run heads, into the private copy of the
//                 XCD the wave runs on when xcd_copies > 1.
// ---------------------------------------------------------------------------
template <int CC, int RM, int NQ, int SRC>
__global__ void __launch_bounds__(kBlock) pass_kernel(PassArgs a) {
    extern __shared__ __align__(16) double lsum_all[];
    const int n = a.n;
    const size_t ld = static_cast<size_t>(a.ld);
    // kRedLds: one table per wave when lds_copies = 4 (the few levels of a small
    // dimension otherwise serialize the LDS atomics of the whole block)
    double* lsum = lsum_all;
    if constexpr (RM == kRedLds) {
        for (int j = threadIdx.x; j < a.lds_copies * a.G_r * CC; j += kBlock) lsum_all[j] = 0.0;
        lsum = lsum_all + static_cast<size_t>((threadIdx.x / kWave) % a.lds_copies) * a.G_r * CC;
        __syncthreads();
    }
    constexpr int NQA = NQ > 0 ? NQ : 1;
    double q[NQA];
#pragma unroll
    for (int j = 0; j < NQA; ++j) q[j] = 0.0;
    double* sums_r = a.sums_r;
    if constexpr (RM == kRedGlobal) {
        if (a.xcd_copies > 1) {
            // hardware XCC id of this wave (gfx94x/gfx950 HW_REG_XCC_ID, 4 bits)
            const unsigned xcc = __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11));
            sums_r += static_cast<size_t>(xcc % static_cast<unsigned>(a.xcd_copies)) * a.copy_stride;
        }
    }
    const int lane = lane_id();
    const double* __restrict__ src = SRC == 0 ? a.src + static_cast<size_t>(a.c0) * ld : nullptr;
    const double* __restrict__ P = SRC == 1 ? a.P + static_cast<size_t>(a.c0) * ld : nullptr;
    const double* __restrict__ Q = SRC == 1 ? a.Q + static_cast<size_t>(a.c0) * ld : nullptr;
    const bool apply = a.gid_a != nullptr;
    const bool store = a.dst != nullptr;
    const bool weighted = a.w != nullptr;

    for (int base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const int i = base + threadIdx.x;
        const bool valid = i < n;
        double v[CC];
        int ga = 0;
        int key = -1 - lane;
        if (valid) {
            if (apply) ga = a.gid_a[i];
            if (RM != 0) key = a.gid_r[i];
#pragma unroll
            for (int c = 0; c < CC; ++c) {
                if constexpr (SRC == 2) {
                    v[c] = 1.0;
                } else if constexpr (SRC == 1) {
                    const double p = P[c * ld + i];
                    v[c] = p - a.coef * (p - Q[c * ld + i]);
                } else {
                    v[c] = src[c * ld + i];
                }
            }
            if constexpr (SRC == 1) {
#pragma unroll
                for (int c = 0; c < CC; ++c) a.xout[(a.c0 + c) * ld + i] = v[c];
            }
            if (apply) {
                const double* m = a.means_a + static_cast<size_t>(ga) * a.C + a.c0;
#pragma unroll
                for (int c = 0; c < CC; ++c) v[c] -= m[c];
            }
            if (store) {
#pragma unroll
                for (int c = 0; c < CC; ++c) a.dst[(a.c0 + c) * ld + i] = v[c];
            }
            if constexpr (NQ == 1) {
#pragma unroll
                for (int c = 0; c < CC; ++c) q[0] += v[c] * v[c];
            } else if constexpr (NQ == 3) {
#pragma unroll
                for (int c = 0; c < CC; ++c) {
                    const size_t off = (a.c0 + c) * ld + i;
                    const double gx = a.it_gx[off];
                    const double dg = v[c] - gx;
                    const double d2 = dg - gx + a.it_x[off];
                    q[0] += dg * d2;
                    q[1] += d2 * d2;
                    q[2] += dg * dg;
                }
            }
            if (weighted && RM != 0) {
                const double wi = a.w[i];
#pragma unroll
                for (int c = 0; c < CC; ++c) v[c] *= wi;
            }
        } else {
#pragma unroll
            for (int c = 0; c < CC; ++c) v[c] = 0.0;
        }
        if constexpr (RM != 0) {
            const int end = run_end_lane(key);
            const bool head = run_head(key) && valid;
            const bool owned = RM == kRedSorted && lane > 0 && end < kWave - 1;
#pragma unroll
            for (int c = 0; c < CC; ++c) {
                const double s = run_sum(v[c], end);
                if (head) {
                    if constexpr (RM == kRedLds) {
                        atomicAdd(&lsum[key * CC + c], s);
                    } else {
                        double* t = &sums_r[static_cast<size_t>(key) * a.C + a.c0 + c];
                        if (owned) {
                            *t = s;
                        } else {
                            atomicAdd(t, s);
                        }
                    }
                }
            }
        }
    }
    if constexpr (RM == kRedLds) {
        __syncthreads();
        for (int j = threadIdx.x; j < a.G_r * CC; j += kBlock) {
            double s = 0.0;
            for (int k = 0; k < a.lds_copies; ++k) s += lsum_all[static_cast<size_t>(k) * a.G_r * CC + j];
            if (s != 0.0) {
                atomicAdd(&a.sums_r[static_cast<size_t>(j / CC) * a.C + a.c0 + (j % CC)], s);
            }
        }
    }
    if constexpr (NQ > 0) {
        block_sum_store<NQA>(q, a.partials + static_cast<size_t>(blockIdx.x) * NQ);
    }
}

// means = (sum of the `copies` accumulator copies) * inv (inv == nullptr:
// the raw sums), every copy reset to 0; optional block max |mean| into partials.
__global__ void __launch_bounds__(kBlock)
finalize_kernel(double* sums, double* means, const double* inv, int G, int C, int copies,
                double* partials) {
    double mx = 0.0;
    const size_t total = static_cast<size_t>(G) * C;
    for (size_t j = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x; j < total;
         j += static_cast<size_t>(gridDim.x) * kBlock) {
        double acc = 0.0;
        for (int x = 0; x < copies; ++x) {
            acc += sums[static_cast<size_t>(x) * total + j];
            sums[static_cast<size_t>(x) * total + j] = 0.0;
        }
        const double m = inv ? acc * inv[j / C] : acc;
        means[j] = m;
        mx = fmax(mx, fabs(m));
    }
    if (partials) {
        __shared__ double red[kBlock / kWave];
        const double wm = wave_max(mx);
        if (lane_id() == 0) red[threadIdx.x / kWave] = wm;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 0.0;
            for (int wv = 0; wv < kBlock / kWave; ++wv) r = fmax(r, red[wv]);
            partials[blockIdx.x] = r;
        }
    }
}

// Deterministic reduction of `count` partial records of NQ doubles each
// (fixed tree order); mode 0 = sum, 1 = max.  One block.
__global__ void __launch_bounds__(kBlock)
reduce_partials_kernel(const double* partials, int count, int nq, int mode, double* out) {
    __shared__ double red[kBlock];
    for (int j = 0; j < nq; ++j) {
        double acc = 0.0;
        for (int t = threadIdx.x; t < count; t += kBlock) {
            const double x = partials[static_cast<size_t>(t) * nq + j];
            acc = mode ? fmax(acc, x) : acc + x;
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                red[threadIdx.x] = mode ? fmax(red[threadIdx.x], red[threadIdx.x + s])
                                        : red[threadIdx.x] + red[threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) out[j] = red[0];
        __syncthreads();
    }
}

// IT statistics over three stored vectors (grand acceleration), n x C block.
__global__ void __launch_bounds__(kBlock)
it_stats_kernel(const double* x, const double* gx, const double* ggx, int n, int C, int ld,
                double* partials) {
    double q[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        for (int c = 0; c < C; ++c) {
            const size_t off = static_cast<size_t>(c) * ld + i;
            const double g = gx[off];
            const double dg = ggx[off] - g;
            const double d2 = dg - g + x[off];
            q[0] += dg * d2;
            q[1] += d2 * d2;
            q[2] += dg * dg;
        }
    }
    block_sum_store<3>(q, partials + static_cast<size_t>(blockIdx.x) * 3);
}

// Per-column sums over an n x C block: mode 0 sum v^2, mode 1 sum a*b.
// partials: per block C values.
__global__ void __launch_bounds__(kBlock)
col_dot_kernel(const double* a, const double* b, int n, int C, int ld, double* partials) {
    __shared__ double red[kBlock / kWave];
    for (int c = 0; c < C; ++c) {
        double acc = 0.0;
        const double* ac = a + static_cast<size_t>(c) * ld;
        const double* bc = b + static_cast<size_t>(c) * ld;
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
            acc += ac[i] * bc[i];
        const double s = wave_sum(acc);
        if (lane_id() == 0) red[threadIdx.x / kWave] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 0.0;
            for (int wv = 0; wv < kBlock / kWave; ++wv) r += red[wv];
            partials[static_cast<size_t>(blockIdx.x) * C + c] = r;
        }
        __syncthreads();
    }
}

// part[block * C + c] = sum of a[j]^2 over the block's elements of column c of
// a level-major table (j = level * C + c).  The grid's thread stride must be a
// multiple of C (every thread keeps one column); C <= kMaxTableCols.
__global__ void __launch_bounds__(kBlock)
table_sumsq_kernel(const double* a, long long total, int C, double* part) {
    __shared__ double red[kMaxTableCols];
    const long long g0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    const int c = static_cast<int>(g0 % C);
    double acc = 0.0;
    for (long long j = g0; j < total; j += stride) acc += a[j] * a[j];
    for (int k = threadIdx.x; k < C; k += kBlock) red[k] = 0.0;
    __syncthreads();
    atomicAdd(&red[c], acc);
    __syncthreads();
    for (int k = threadIdx.x; k < C; k += kBlock) part[static_cast<size_t>(blockIdx.x) * C + k] = red[k];
}

// y = a - alpha_c * b per column (alpha on device, C entries); in place allowed.
__global__ void __launch_bounds__(kBlock)
col_axpy_kernel(double* y, const double* a, const double* b, const double* alpha, int n, int C,
                int ld, double sign) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        for (int c = 0; c < C; ++c) {
            const size_t off = static_cast<size_t>(c) * ld + i;
            y[off] = a[off] + sign * alpha[c] * b[off];
        }
    }
}

// y = a - b (elementwise over the n x C block).
__global__ void __launch_bounds__(kBlock)
sub_kernel(double* y, const double* a, const double* b, int n, int C, int ld) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        for (int c = 0; c < C; ++c) {
            const size_t off = static_cast<size_t>(c) * ld + i;
            y[off] = a[off] - b[off];
        }
    }
}

// Inverse weight sums (0 for empty levels).
__global__ void __launch_bounds__(kBlock) inverse_kernel(const double* ws, double* inv, int G) {
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < G; g += gridDim.x * kBlock) {
        const double s = ws[g];
        inv[g] = s > 0.0 ? 1.0 / s : 0.0;
    }
}

// Permutation helpers.
__global__ void __launch_bounds__(kBlock) iota_kernel(int* v, int n) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) v[i] = i;
}

__global__ void __launch_bounds__(kBlock)
gather_ids_kernel(const int* src, const int* perm, int* dst, int n) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        dst[i] = perm ? src[perm[i]] : src[i];
}

__global__ void __launch_bounds__(kBlock)
gather_cols_kernel(const double* src, int ld_src, const int* perm, double* dst, int ld_dst, int n,
                   int C) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const int s = perm ? perm[i] : i;
        for (int c = 0; c < C; ++c)
            dst[static_cast<size_t>(c) * ld_dst + i] = src[static_cast<size_t>(c) * ld_src + s];
    }
}

__global__ void __launch_bounds__(kBlock)
scatter_cols_kernel(const double* src, int ld_src, const int* perm, double* dst, int ld_dst, int n,
                    int C) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const int d = perm ? perm[i] : i;
        for (int c = 0; c < C; ++c)
            dst[static_cast<size_t>(c) * ld_dst + d] = src[static_cast<size_t>(c) * ld_src + i];
    }
}

__global__ void __launch_bounds__(kBlock)
gather_vec_kernel(const double* src, const int* perm, double* dst, int n) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        dst[i] = perm ? src[perm[i]] : src[i];
}

// 1 if the ids are non-decreasing (checked by the host through a flag).
__global__ void __launch_bounds__(kBlock) sorted_check_kernel(const int* gid, int n, int* flag) {
    for (int i = blockIdx.x * kBlock + threadIdx.x + 1; i < n; i += gridDim.x * kBlock)
        if (gid[i] < gid[i - 1]) atomicOr(flag, 1);
}

// ---------------------------------------------------------------------------
// Launchers (host side, same translation unit so the templates instantiate).
// ---------------------------------------------------------------------------

template <int CC, int RM, int NQ, int SRC>
static hipError_t launch_pass_t(const PassArgs& a_in, int grid, hipStream_t s) {
    PassArgs a = a_in;
    const size_t table = static_cast<size_t>(a.G_r) * CC * sizeof(double);
    a.lds_copies = (RM == kRedLds && (kBlock / kWave) * table <= 64 * 1024) ? kBlock / kWave : 1;
    const size_t shmem = RM == kRedLds ? a.lds_copies * table : 0;
    hipLaunchKernelGGL((pass_kernel<CC, RM, NQ, SRC>), dim3(grid), dim3(kBlock), shmem, s, a);
    return hipGetLastError();
}

// Instantiated variants: plain source with any reduce mode and NQ in {0,1}
// (NQ 3 only without a reduce), the IT extrapolation source with a reduce,
// and the constant source (weight sums) with one column.
template <int CC>
static hipError_t launch_pass_cc(const PassArgs& a, int src, int rmode, int nq, int grid,
                                 hipStream_t s) {
    if (src == 0) {
        if (nq == 3) {
            if (rmode != 0) return hipErrorInvalidValue;
            return launch_pass_t<CC, 0, 3, 0>(a, grid, s);
        }
        switch (rmode * 2 + (nq == 1)) {
            case 0: return launch_pass_t<CC, 0, 0, 0>(a, grid, s);
            case 1: return launch_pass_t<CC, 0, 1, 0>(a, grid, s);
            case 2: return launch_pass_t<CC, kRedLds, 0, 0>(a, grid, s);
            case 3: return launch_pass_t<CC, kRedLds, 1, 0>(a, grid, s);
            case 4: return launch_pass_t<CC, kRedSorted, 0, 0>(a, grid, s);
            case 5: return launch_pass_t<CC, kRedSorted, 1, 0>(a, grid, s);
            case 6: return launch_pass_t<CC, kRedGlobal, 0, 0>(a, grid, s);
            case 7: return launch_pass_t<CC, kRedGlobal, 1, 0>(a, grid, s);
            default: return hipErrorInvalidValue;
        }
    }
    if (src == 1) {
        if (nq != 0) return hipErrorInvalidValue;
        switch (rmode) {
            case 0: return launch_pass_t<CC, 0, 0, 1>(a, grid, s);
            case kRedLds: return launch_pass_t<CC, kRedLds, 0, 1>(a, grid, s);
            case kRedSorted: return launch_pass_t<CC, kRedSorted, 0, 1>(a, grid, s);
            default: return launch_pass_t<CC, kRedGlobal, 0, 1>(a, grid, s);
        }
    }
    if constexpr (CC == 1) {
        if (nq != 0) return hipErrorInvalidValue;
        switch (rmode) {
            case kRedLds: return launch_pass_t<1, kRedLds, 0, 2>(a, grid, s);
            case kRedSorted: return launch_pass_t<1, kRedSorted, 0, 2>(a, grid, s);
            case kRedGlobal: return launch_pass_t<1, kRedGlobal, 0, 2>(a, grid, s);
            default: return hipErrorInvalidValue;
        }
    }
    return hipErrorInvalidValue;
}

hipError_t launch_pass(const PassArgs& a, int cc, int rmode, int nq, int grid, hipStream_t s) {
    const int src = a.one ? 2 : (a.P ? 1 : 0);
    if (rmode != 0 && a.gid_r == nullptr) return hipErrorInvalidValue;
    if (src == 1 && a.xout == nullptr) return hipErrorInvalidValue;
    switch (cc) {
        case 1: return launch_pass_cc<1>(a, src, rmode, nq, grid, s);
        case 2: return launch_pass_cc<2>(a, src, rmode, nq, grid, s);
        case 3: return launch_pass_cc<3>(a, src, rmode, nq, grid, s);
        case 4: return launch_pass_cc<4>(a, src, rmode, nq, grid, s);
        case 5: return launch_pass_cc<5>(a, src, rmode, nq, grid, s);
        case 6: return launch_pass_cc<6>(a, src, rmode, nq, grid, s);
        case 7: return launch_pass_cc<7>(a, src, rmode, nq, grid, s);
        case 8: return launch_pass_cc<8>(a, src, rmode, nq, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_finalize(double* sums, double* means, const double* inv, int G, int C,
                           int copies, double* partials, int grid, hipStream_t s) {
    hipLaunchKernelGGL(finalize_kernel, dim3(grid), dim3(kBlock), 0, s, sums, means, inv, G, C,
                       copies, partials);
    return hipGetLastError();
}

hipError_t launch_reduce_partials(const double* partials, int count, int nq, int mode, double* out,
                                  hipStream_t s) {
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(kBlock), 0, s, partials, count, nq,
                       mode, out);
    return hipGetLastError();
}

hipError_t launch_it_stats(const double* x, const double* gx, const double* ggx, int n, int C,
                           int ld, double* partials, int grid, hipStream_t s) {
    hipLaunchKernelGGL(it_stats_kernel, dim3(grid), dim3(kBlock), 0, s, x, gx, ggx, n, C, ld,
                       partials);
    return hipGetLastError();
}

namespace {
// Profile of a level table (sum, sum of squares, max, entries <= 1 + 1e-12):
// per block partials in thread order, then one block sums them in block order
__global__ void __launch_bounds__(kBlock) level_profile_kernel(const double* ws, int G, double* part) {
    __shared__ double r[4][kBlock];
    double s = 0.0, q = 0.0, m = -1e308, one = 0.0;
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < G; g += gridDim.x * kBlock) {
        const double v = ws[g];
        s += v;
        q += v * v;
        m = fmax(m, v);
        one += v <= 1.0 + 1e-12 ? 1.0 : 0.0;
    }
    r[0][threadIdx.x] = s;
    r[1][threadIdx.x] = q;
    r[2][threadIdx.x] = m;
    r[3][threadIdx.x] = one;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            r[0][threadIdx.x] += r[0][threadIdx.x + w];
            r[1][threadIdx.x] += r[1][threadIdx.x + w];
            r[2][threadIdx.x] = fmax(r[2][threadIdx.x], r[2][threadIdx.x + w]);
            r[3][threadIdx.x] += r[3][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) part[static_cast<size_t>(blockIdx.x) * 4 + threadIdx.x] = r[threadIdx.x][0];
}

__global__ void __launch_bounds__(64) level_profile_final_kernel(const double* part, int nblk, double* out) {
    if (threadIdx.x != 0) return;
    double s = 0.0, q = 0.0, m = -1e308, one = 0.0;
    for (int b = 0; b < nblk; ++b) {
        s += part[4 * b];
        q += part[4 * b + 1];
        m = fmax(m, part[4 * b + 2]);
        one += part[4 * b + 3];
    }
    out[0] = s;
    out[1] = q;
    out[2] = m;
    out[3] = one;
}
}  // namespace

hipError_t launch_level_profile(const double* ws, int G, double* part, int grid, double* out, hipStream_t s) {
    hipLaunchKernelGGL(level_profile_kernel, dim3(grid), dim3(kBlock), 0, s, ws, G, part);
    hipLaunchKernelGGL(level_profile_final_kernel, dim3(1), dim3(64), 0, s, part, grid, out);
    return hipGetLastError();
}

hipError_t launch_col_dot(const double* a, const double* b, int n, int C, int ld, double* partials,
                          int grid, hipStream_t s) {
    hipLaunchKernelGGL(col_dot_kernel, dim3(grid), dim3(kBlock), 0, s, a, b, n, C, ld, partials);
    return hipGetLastError();
}

hipError_t launch_table_sumsq(const double* a, long long total, int C, double* part, int grid, hipStream_t s) {
    if (C <= 0 || C > kMaxTableCols) return hipErrorInvalidValue;
    hipLaunchKernelGGL(table_sumsq_kernel, dim3(grid), dim3(kBlock), 0, s, a, total, C, part);
    return hipGetLastError();
}

hipError_t launch_col_axpy(double* y, const double* a, const double* b, const double* alpha, int n,
                           int C, int ld, double sign, int grid, hipStream_t s) {
    hipLaunchKernelGGL(col_axpy_kernel, dim3(grid), dim3(kBlock), 0, s, y, a, b, alpha, n, C, ld,
                       sign);
    return hipGetLastError();
}

hipError_t launch_sub(double* y, const double* a, const double* b, int n, int C, int ld, int grid,
                      hipStream_t s) {
    hipLaunchKernelGGL(sub_kernel, dim3(grid), dim3(kBlock), 0, s, y, a, b, n, C, ld);
    return hipGetLastError();
}

// ---- heterogeneous slopes (slope_project_inplace, src/fe_absorption.cpp:4096-4300)

// mom[3 g + {0, 1, 2}] += w, w z, w z^2 over the level's rows (z may be null)
__global__ void __launch_bounds__(kBlock)
slope_moments_kernel(const int* gid, const double* z, const double* w, int n, double* mom) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double wi = w ? w[i] : 1.0;
        double* m = mom + 3 * static_cast<size_t>(gid[i]);
        atomicAdd(&m[0], wi);
        if (z) {
            const double zi = z[i];
            atomicAdd(&m[1], wi * zi);
            atomicAdd(&m[2], wi * zi * zi);
        }
    }
}

// sums[k][2 (g C + c) + {0, 1}] += w v, w z v (the second only with z), into
// copy k = block % copies of the G x C x 2 table.  Consecutive rows of equal
// level are summed in the wave first (segmented sums over lanes), so a run
// costs one atomic per value: with the rows sorted by this dimension a level
// of r rows takes about r / 64 atomics per column instead of r.
__global__ void __launch_bounds__(kBlock)
slope_reduce_kernel(const double* V, int ld, int C, int n, const int* gid, const double* z, const double* w,
                    double* sums, int G, int copies) {
    double* tab = sums + static_cast<size_t>(blockIdx.x % copies) * 2 * static_cast<size_t>(G) * C;
    const int lane = lane_id();
    for (int base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const int i = base + threadIdx.x;
        const bool valid = i < n;
        const int key = valid ? gid[i] : -1 - lane;
        const int end = run_end_lane(key);
        const bool head = run_head(key) && valid;
        const double wi = valid ? (w ? w[i] : 1.0) : 0.0;
        const double zi = (valid && z) ? z[i] : 0.0;
        double* sg = tab + 2 * static_cast<size_t>(valid ? key : 0) * C;
        for (int c = 0; c < C; ++c) {
            const double v = valid ? wi * V[static_cast<size_t>(c) * ld + i] : 0.0;
            const double s1 = run_sum(v, end);
            if (head) atomicAdd(&sg[2 * c], s1);
            if (z) {
                const double s2 = run_sum(zi * v, end);
                if (head) atomicAdd(&sg[2 * c + 1], s2);
            }
        }
    }
}

// Per level and column: coef[2 (g C + c) + {0, 1}] = (alpha, gamma).  mode 0:
// the level mean; 1: intercept and slope, the 2 x 2 solve when det > 1e-12
// max(1, sw szz), else the mean; 2: slope only, szy / szz when
// szz > 1e-12 max(1, szz), else 0 (:4186-4240).
__global__ void __launch_bounds__(kBlock)
slope_coef_kernel(const double* sums, const double* mom, int G, int C, int mode, int copies, double* coef) {
    const long long total = static_cast<long long>(G) * C;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int g = static_cast<int>(j / C);
        const double sw = mom[3 * g], sz = mom[3 * g + 1], szz = mom[3 * g + 2];
        double sv = 0.0, szv = 0.0;
        for (int k = 0; k < copies; ++k) {
            sv += sums[static_cast<size_t>(k) * 2 * total + 2 * j];
            szv += sums[static_cast<size_t>(k) * 2 * total + 2 * j + 1];
        }
        double alpha = 0.0, gamma = 0.0;
        if (mode == 0) {
            if (sw > 0.0) alpha = sv / sw;
        } else if (mode == 1) {
            if (sw > 0.0) {
                const double det = sw * szz - sz * sz;
                if (det > 1e-12 * fmax(1.0, sw * szz)) {
                    alpha = (szz * sv - sz * szv) / det;
                    gamma = (-sz * sv + sw * szv) / det;
                } else {
                    alpha = sv / sw;
                }
            }
        } else if (szz > 1e-12 * fmax(1.0, szz)) {
            gamma = szv / szz;
        }
        coef[2 * j] = alpha;
        coef[2 * j + 1] = gamma;
    }
}

// v -= alpha[g] + gamma[g] z (z may be null)
__global__ void __launch_bounds__(kBlock)
slope_apply_kernel(double* V, int ld, int C, int n, const int* gid, const double* z, const double* coef) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double zi = z ? z[i] : 0.0;
        const double* cg = coef + 2 * static_cast<size_t>(gid[i]) * C;
        for (int c = 0; c < C; ++c) V[static_cast<size_t>(c) * ld + i] -= cg[2 * c] + cg[2 * c + 1] * zi;
    }
}

// partials[block * C + c] = sum_i w_i a_ic b_ic over the block's rows (w may be null)
__global__ void __launch_bounds__(kBlock)
slope_wdot_kernel(const double* a, const double* b, const double* w, int n, int C, int ld, double* partials) {
    __shared__ double red[kBlock / kWave];
    for (int c = 0; c < C; ++c) {
        double acc = 0.0;
        const double* ac = a + static_cast<size_t>(c) * ld;
        const double* bc = b + static_cast<size_t>(c) * ld;
        for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
            acc += (w ? w[i] : 1.0) * ac[i] * bc[i];
        const double s = wave_sum(acc);
        if (lane_id() == 0) red[threadIdx.x / kWave] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 0.0;
            for (int wv = 0; wv < kBlock / kWave; ++wv) r += red[wv];
            partials[static_cast<size_t>(blockIdx.x) * C + c] = r;
        }
        __syncthreads();
    }
}

hipError_t launch_slope_wdot(const double* a, const double* b, const double* w, int n, int C, int ld,
                             double* partials, int grid, hipStream_t s) {
    hipLaunchKernelGGL(slope_wdot_kernel, dim3(grid), dim3(kBlock), 0, s, a, b, w, n, C, ld, partials);
    return hipGetLastError();
}

hipError_t launch_slope_moments(const int* gid, const double* z, const double* w, int n, double* mom, int grid,
                                hipStream_t s) {
    hipLaunchKernelGGL(slope_moments_kernel, dim3(grid), dim3(kBlock), 0, s, gid, z, w, n, mom);
    return hipGetLastError();
}

hipError_t launch_slope_reduce(const double* V, int ld, int C, int n, const int* gid, const double* z,
                               const double* w, double* sums, int G, int copies, int grid, hipStream_t s) {
    hipLaunchKernelGGL(slope_reduce_kernel, dim3(grid), dim3(kBlock), 0, s, V, ld, C, n, gid, z, w, sums, G,
                       copies);
    return hipGetLastError();
}

hipError_t launch_slope_coef(const double* sums, const double* mom, int G, int C, int mode, int copies,
                             double* coef, int grid, hipStream_t s) {
    hipLaunchKernelGGL(slope_coef_kernel, dim3(grid), dim3(kBlock), 0, s, sums, mom, G, C, mode, copies, coef);
    return hipGetLastError();
}

hipError_t launch_slope_apply(double* V, int ld, int C, int n, const int* gid, const double* z, const double* coef,
                              int grid, hipStream_t s) {
    hipLaunchKernelGGL(slope_apply_kernel, dim3(grid), dim3(kBlock), 0, s, V, ld, C, n, gid, z, coef);
    return hipGetLastError();
}

hipError_t launch_inverse(const double* ws, double* inv, int G, int grid, hipStream_t s) {
    hipLaunchKernelGGL(inverse_kernel, dim3(grid), dim3(kBlock), 0, s, ws, inv, G);
    return hipGetLastError();
}

hipError_t launch_iota(int* v, int n, int grid, hipStream_t s) {
    hipLaunchKernelGGL(iota_kernel, dim3(grid), dim3(kBlock), 0, s, v, n);
    return hipGetLastError();
}

hipError_t launch_gather_ids(const int* src, const int* perm, int* dst, int n, int grid,
                             hipStream_t s) {
    hipLaunchKernelGGL(gather_ids_kernel, dim3(grid), dim3(kBlock), 0, s, src, perm, dst, n);
    return hipGetLastError();
}

hipError_t launch_gather_cols(const double* src, int ld_src, const int* perm, double* dst,
                              int ld_dst, int n, int C, int grid, hipStream_t s) {
    hipLaunchKernelGGL(gather_cols_kernel, dim3(grid), dim3(kBlock), 0, s, src, ld_src, perm, dst,
                       ld_dst, n, C);
    return hipGetLastError();
}

hipError_t launch_scatter_cols(const double* src, int ld_src, const int* perm, double* dst,
                               int ld_dst, int n, int C, int grid, hipStream_t s) {
    hipLaunchKernelGGL(scatter_cols_kernel, dim3(grid), dim3(kBlock), 0, s, src, ld_src, perm, dst,
                       ld_dst, n, C);
    return hipGetLastError();
}

hipError_t launch_gather_vec(const double* src, const int* perm, double* dst, int n, int grid,
                             hipStream_t s) {
    hipLaunchKernelGGL(gather_vec_kernel, dim3(grid), dim3(kBlock), 0, s, src, perm, dst, n);
    return hipGetLastError();
}

hipError_t launch_sorted_check(const int* gid, int n, int* flag, int grid, hipStream_t s) {
    hipLaunchKernelGGL(sorted_check_kernel, dim3(grid), dim3(kBlock), 0, s, gid, n, flag);
    return hipGetLastError();
}

// FE recovery (savefe).  After the within transform vt = M_D partial by the
// Schur path, the coefficients of the non-primary dimensions are known; the
// primary level g gets alpha_p[g] = weighted mean over its rows of
// partial - vt - sum_k coef_k[gid_k]: the rows of one level carry the same
// value up to rounding, and the mean absorbs the rounding.
__global__ void recover_primary_kernel(RecoverArgs a, double* __restrict__ sums, int n) {
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double t = a.partial[i] - a.vt[i];
        for (int k = 0; k < a.D; ++k) {
            if (k == a.p) continue;
            t -= a.alpha[k][a.gid[k][i]];
        }
        if (a.w) t *= a.w[i];
        atomicAdd(&sums[a.gid[a.p][i]], t);
    }
}

// alpha[g] = sums[g] / wsum[g] in place (0 for a level without weight).
__global__ void recover_divide_kernel(double* __restrict__ alpha, const double* __restrict__ wsum, int G) {
    const int stride = gridDim.x * blockDim.x;
    for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += stride)
        alpha[g] = wsum[g] > 0.0 ? alpha[g] / wsum[g] : 0.0;
}

// res_i = partial_i - sum_d alpha_d[gid_d[i]] (the residual fe_recovery_max_delta checks).
__global__ void recover_residual_kernel(RecoverArgs a, double* __restrict__ res, int n) {
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double t = a.partial[i];
        for (int k = 0; k < a.D; ++k) t -= a.alpha[k][a.gid[k][i]];
        res[i] = t;
    }
}

// contrib[d * n + i] = alpha_d[gid_d[i]] (the per-row contributions, dimension-major).
__global__ void recover_expand_kernel(RecoverArgs a, double* __restrict__ contrib, int n) {
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        for (int k = 0; k < a.D; ++k) contrib[static_cast<size_t>(k) * n + i] = a.alpha[k][a.gid[k][i]];
}

hipError_t launch_recover_expand(const RecoverArgs& a, double* contrib, int n, int grid, hipStream_t s) {
    hipLaunchKernelGGL(recover_expand_kernel, dim3(grid), dim3(kBlock), 0, s, a, contrib, n);
    return hipGetLastError();
}

hipError_t launch_recover_primary(const RecoverArgs& a, double* sums, int n, int grid, hipStream_t s) {
    hipLaunchKernelGGL(recover_primary_kernel, dim3(grid), dim3(kBlock), 0, s, a, sums, n);
    return hipGetLastError();
}

hipError_t launch_recover_divide(double* alpha, const double* wsum, int G, int grid, hipStream_t s) {
    hipLaunchKernelGGL(recover_divide_kernel, dim3(grid), dim3(kBlock), 0, s, alpha, wsum, G);
    return hipGetLastError();
}

hipError_t launch_recover_residual(const RecoverArgs& a, double* res, int n, int grid, hipStream_t s) {
    hipLaunchKernelGGL(recover_residual_kernel, dim3(grid), dim3(kBlock), 0, s, a, res, n);
    return hipGetLastError();
}

}  // namespace xfe

