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

namespace xfe {

namespace {

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Sum over the lanes of a wave (result valid in lane 0).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, kWave));
    return v;
}

// Block-level sum of NQ scalars; thread 0 writes them to out[0..NQ).
template <int NQ>
__device__ __forceinline__ void block_sum_store(double (&q)[NQ], double* out) {
    __shared__ double red[NQ][kBlock / kWave];
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
        const double s = wave_sum(q[j]);
        if (lane == 0) red[j][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
            double s = 0.0;
            for (int wv = 0; wv < kBlock / kWave; ++wv) s += red[j][wv];
            out[j] = s;
        }
    }
}

// End lane (inclusive) of the run of equal keys that contains this lane.
__device__ __forceinline__ int run_end_lane(int key) {
    const int lane = lane_id();
    const int prev = __shfl_up(key, 1, kWave);
    const bool head = (lane == 0) || (prev != key);
    const uint64_t heads = __ballot(head);
    // heads strictly above this lane
    const uint64_t above = (lane == 63) ? 0ull : (heads & ~((2ull << lane) - 1ull));
    return above ? (__builtin_ctzll(above) - 1) : 63;
}

__device__ __forceinline__ bool run_head(int key) {
    const int lane = lane_id();
    const int prev = __shfl_up(key, 1, kWave);
    return (lane == 0) || (prev != key);
}

// Segmented suffix sum: after the loop a run's head lane holds the run total.
__device__ __forceinline__ double run_sum(double v, int end) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const double o = __shfl_down(v, off, kWave);
        if (lane + off <= end) v += o;
    }
    return v;
}

}  // namespace

// ---------------------------------------------------------------------------
// The fused streaming pass.  For each row i and column c in [c0, c0+CC):
//   v = one ? 1 : P ? P - coef (P - Q) : src      (IT extrapolation fused in)
//   if xout: xout = v
//   if gid_a: v -= means_a[gid_a[i], c]           (finish projection a)
//   if dst:  dst = v
//   reductions: sum v^2, or IT stats against (it_x, it_gx)
//   if gid_r: sums_r[gid_r[i], c] += w_i v        (start projection r)
//
// A block handles tiles of RS x 256 consecutive rows; row slot r of thread t
// is row tile + r*256 + t, so every slot is a coalesced wave64 access and the
// loads of all RS slots are issued before any reduction (memory-level
// parallelism).  RM selects how level sums are accumulated:
//   kRedLds    -- LDS-privatised, flushed once per block (small dimensions);
//   kRedSorted -- the level-sorted dimension: a run that is the whole level
//                 inside the wave's 64 rows is stored with a plain store, only
//                 levels cut by a wave boundary use an atomic;
//   kRedGlobal -- FP64 atomics on run heads, into the private copy of the
//                 XCD the wave runs on when xcd_copies > 1.
// ---------------------------------------------------------------------------
template <int CC, int RM, int NQ, int RS>
__global__ void __launch_bounds__(kBlock) pass_kernel(PassArgs a) {
    extern __shared__ double lsum[];
    const int n = a.n;
    const size_t ld = static_cast<size_t>(a.ld);
    if constexpr (RM == kRedLds) {
        for (int j = threadIdx.x; j < a.G_r * CC; j += kBlock) lsum[j] = 0.0;
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
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

    constexpr int TILE = RS * kBlock;
    for (int tile = blockIdx.x * TILE; tile < n; tile += gridDim.x * TILE) {
        double v[RS][CC];
        int ga[RS], gr[RS];
        bool valid[RS];
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            const int i = tile + r * kBlock + threadIdx.x;
            valid[r] = i < n;
            ga[r] = (a.gid_a && valid[r]) ? a.gid_a[i] : 0;
            gr[r] = (a.gid_r && valid[r]) ? a.gid_r[i] : -1 - lane;
#pragma unroll
            for (int c = 0; c < CC; ++c) {
                const size_t off = static_cast<size_t>(a.c0 + c) * ld + i;
                if (a.one) {
                    v[r][c] = 1.0;
                } else if (a.P) {
                    const double p = valid[r] ? a.P[off] : 0.0;
                    const double qq = valid[r] ? a.Q[off] : 0.0;
                    v[r][c] = p - a.coef * (p - qq);
                } else {
                    v[r][c] = valid[r] ? a.src[off] : 0.0;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RS; ++r) {
            const int i = tile + r * kBlock + threadIdx.x;
            if (a.xout && valid[r]) {
#pragma unroll
                for (int c = 0; c < CC; ++c) a.xout[static_cast<size_t>(a.c0 + c) * ld + i] = v[r][c];
            }
            if (a.gid_a && valid[r]) {
                const double* m = a.means_a + static_cast<size_t>(ga[r]) * a.C + a.c0;
#pragma unroll
                for (int c = 0; c < CC; ++c) v[r][c] -= m[c];
            }
            if (a.dst && valid[r]) {
#pragma unroll
                for (int c = 0; c < CC; ++c) a.dst[static_cast<size_t>(a.c0 + c) * ld + i] = v[r][c];
            }
            if constexpr (NQ == 1) {
#pragma unroll
                for (int c = 0; c < CC; ++c) q[0] += v[r][c] * v[r][c];
            } else if constexpr (NQ == 3) {
                if (valid[r]) {
#pragma unroll
                    for (int c = 0; c < CC; ++c) {
                        const size_t off = static_cast<size_t>(a.c0 + c) * ld + i;
                        const double gx = a.it_gx[off];
                        const double dg = v[r][c] - gx;
                        const double d2 = dg - gx + a.it_x[off];
                        q[0] += dg * d2;
                        q[1] += d2 * d2;
                        q[2] += dg * dg;
                    }
                }
            }
        }
        if (a.gid_r) {
#pragma unroll
            for (int r = 0; r < RS; ++r) {
                const int i = tile + r * kBlock + threadIdx.x;
                const int key = gr[r];
                const double wi = (a.w && valid[r]) ? a.w[i] : 1.0;
                const int end = run_end_lane(key);
                const bool head = run_head(key) && valid[r];
                bool owned = false;
                if constexpr (RM == kRedSorted) {
                    // the 64 rows of this wave are [wbase, wbase + 64)
                    const int wbase = tile + r * kBlock + wave * kWave;
                    const int before = (wbase > 0 && wbase <= n) ? a.gid_r[wbase - 1] : INT_MIN;
                    const int after = wbase + kWave < n ? a.gid_r[wbase + kWave] : INT_MIN;
                    owned = (lane > 0 || before != key) && (end < kWave - 1 || after != key);
                }
#pragma unroll
                for (int c = 0; c < CC; ++c) {
                    const double s = run_sum(a.w ? wi * v[r][c] : v[r][c], end);
                    if (head) {
                        if constexpr (RM == kRedLds) {
                            atomicAdd(&lsum[key * CC + c], s);
                        } else {
                            double* t = &sums_r[static_cast<size_t>(key) * a.C + a.c0 + c];
                            if (RM == kRedSorted && owned) {
                                *t = s;
                            } else {
                                atomicAdd(t, s);
                            }
                        }
                    }
                }
            }
        }
    }
    if constexpr (RM == kRedLds) {
        __syncthreads();
        for (int j = threadIdx.x; j < a.G_r * CC; j += kBlock) {
            const double s = lsum[j];
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

template <int CC, int RM, int NQ>
static hipError_t launch_pass_t(const PassArgs& a, int grid, hipStream_t s) {
    const size_t shmem = RM == kRedLds ? static_cast<size_t>(a.G_r) * CC * sizeof(double) : 0;
    hipLaunchKernelGGL((pass_kernel<CC, RM, NQ, kRowSlots>), dim3(grid), dim3(kBlock), shmem, s, a);
    return hipGetLastError();
}

template <int CC, int RM>
static hipError_t launch_pass_rm(const PassArgs& a, int nq, int grid, hipStream_t s) {
    if (nq == 0) return launch_pass_t<CC, RM, 0>(a, grid, s);
    if (nq == 1) return launch_pass_t<CC, RM, 1>(a, grid, s);
    return launch_pass_t<CC, RM, 3>(a, grid, s);
}

template <int CC>
static hipError_t launch_pass_cc(const PassArgs& a, int rmode, int nq, int grid, hipStream_t s) {
    switch (rmode) {
        case kRedLds: return launch_pass_rm<CC, kRedLds>(a, nq, grid, s);
        case kRedSorted: return launch_pass_rm<CC, kRedSorted>(a, nq, grid, s);
        default: return launch_pass_rm<CC, kRedGlobal>(a, nq, grid, s);
    }
}

hipError_t launch_pass(const PassArgs& a, int cc, int rmode, int nq, int grid, hipStream_t s) {
    switch (cc) {
        case 1: return launch_pass_cc<1>(a, rmode, nq, grid, s);
        case 2: return launch_pass_cc<2>(a, rmode, nq, grid, s);
        case 3: return launch_pass_cc<3>(a, rmode, nq, grid, s);
        case 4: return launch_pass_cc<4>(a, rmode, nq, grid, s);
        case 5: return launch_pass_cc<5>(a, rmode, nq, grid, s);
        case 6: return launch_pass_cc<6>(a, rmode, nq, grid, s);
        case 7: return launch_pass_cc<7>(a, rmode, nq, grid, s);
        case 8: return launch_pass_cc<8>(a, rmode, nq, grid, s);
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

hipError_t launch_col_dot(const double* a, const double* b, int n, int C, int ld, double* partials,
                          int grid, hipStream_t s) {
    hipLaunchKernelGGL(col_dot_kernel, dim3(grid), dim3(kBlock), 0, s, a, b, n, C, ld, partials);
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

}  // namespace xfe
