// Kernels of the mixed-precision solver (precision 1): preconditioned CG in
// coefficient space on the Schur complement of the level-sorted dimension p.
//
// With p the dimension with the most levels (workers) and R the others
// (firms, years), eliminating p exactly leaves the system
//     S x = D_R' M_p D_R x = b,   M_p = I - D_p (D_p' D_p)^-1 D_p'
// over the coefficients x of R only.  One application of S is one pass over
// the row ids with no n x C data: per row gather the coefficients of R and add
// them, subtract the mean over the row's level of p (inside the level-aligned
// tile), reduce into the levels of R (spells for firms, LDS for years).  The
// same kernel in FP64 with the data block V and sign -1 is the refinement
// residual  u = M_p (v - D_R beta)  and its reduce b = D_R' u, and (STORE)
// the final within transform.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "absorb_kernels.h"
#include "lp_precond.h"
#include "wave_ops.h"

namespace xfe {

namespace {

constexpr int kSlots = kTileRows / kBlock;
constexpr int kChunks = kSlots * (kBlock / kWave);

__device__ __forceinline__ int tile_rank2(const int* cnt, int chunk, uint64_t mask) {
    int base = 0;
    for (int k = 0; k < chunk; ++k) base += cnt[k];
    const int lane = lane_id();
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    return base + __popcll(mask & upto) - 1;
}

template <typename T>
__device__ __forceinline__ T shfl_down_t(T v, int off) { return __shfl_down(v, off, kWave); }

template <typename T>
__device__ __forceinline__ T run_sum_t(T v, int end) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const T o = shfl_down_t(v, off);
        if (lane + off <= end) v += o;
    }
    return v;
}

}  // namespace

template <typename T, int CC, int NR, bool HAS_V, bool STORE>
__global__ void __launch_bounds__(kBlock) schur_kernel(SchurArgs<T> a) {
    // 16-byte aligned: the dynamic block follows the static LDS, and the FP64
    // LDS atomics / 128-bit LDS loads below need naturally aligned addresses
    extern __shared__ __align__(16) unsigned char dyn[];
    T* sm = reinterpret_cast<T*>(dyn);                 // kTileRows x CC
    T* lr = sm + kTileRows * CC;                       // LDS-mode sums of the R dims
    __shared__ int cnt[kChunks];
    __shared__ int cnt_total;
    const size_t ld = static_cast<size_t>(a.ld);
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    for (int k = 0; k < NR; ++k) {
        if (a.mode_r[k] == kRedLds) {
            for (int j = threadIdx.x; j < a.G_r[k] * CC * a.lds_copies[k]; j += kBlock) lr[a.lds_off[k] + j] = T(0);
        }
    }
    double q = 0.0;
    for (int t = blockIdx.x; t < a.num_tiles; t += gridDim.x) {
        const int t0 = a.tile_start[t];
        const int t1 = a.tile_start[t + 1];
        if (t0 >= t1) continue;
        T u[kSlots][CC];
        int kp[kSlots], kr[NR][kSlots];
        bool valid[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
            const int i = t0 + j * kBlock + threadIdx.x;
            valid[j] = i < t1;
            kp[j] = -1 - lane;
#pragma unroll
            for (int k = 0; k < NR; ++k) kr[k][j] = -1 - lane;
#pragma unroll
            for (int c = 0; c < CC; ++c) u[j][c] = T(0);
            if (valid[j]) {
                kp[j] = a.gid_p[i];
                if constexpr (HAS_V) {
#pragma unroll
                    for (int c = 0; c < CC; ++c) u[j][c] = a.V[(a.c0 + c) * ld + i];
                }
#pragma unroll
                for (int k = 0; k < NR; ++k) {
                    kr[k][j] = a.gid_r[k][i];
                    const T* cf = a.coef_r[k] + static_cast<size_t>(kr[k][j]) * a.C + a.c0;
#pragma unroll
                    for (int c = 0; c < CC; ++c) u[j][c] += a.sign * cf[c];
                }
            }
        }
        int pp[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
            pp[j] = __shfl_up(kp[j], 1, kWave);
            const int i = t0 + j * kBlock + threadIdx.x;
            if (lane == 0 && valid[j]) pp[j] = i > t0 ? a.gid_p[i - 1] : INT_MIN;
        }
        // ---- demean over the levels of p inside the tile
        for (int k = threadIdx.x; k < kTileRows * CC; k += kBlock) sm[k] = T(0);
        uint64_t mask[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
            mask[j] = __ballot(valid[j] && pp[j] != kp[j]);
            if (lane == 0) cnt[j * (kBlock / kWave) + wave] = __popcll(mask[j]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
            const int rank = tile_rank2(cnt, j * (kBlock / kWave) + wave, mask[j]);
            const int end = run_end_lane(kp[j]);
            const bool seg = run_head(kp[j]) && valid[j];
#pragma unroll
            for (int c = 0; c < CC; ++c) {
                const T s = run_sum_t(u[j][c], end);
                if (seg) atomicAdd(&sm[rank * CC + c], s);
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
            if (valid[j]) {
                const int rank = tile_rank2(cnt, j * (kBlock / kWave) + wave, mask[j]);
                const T inv = a.inv_p[kp[j]];
#pragma unroll
                for (int c = 0; c < CC; ++c) u[j][c] -= sm[rank * CC + c] * inv;
            }
        }
        __syncthreads();
        if constexpr (STORE) {
#pragma unroll
            for (int j = 0; j < kSlots; ++j) {
                if (!valid[j]) continue;
                const int i = t0 + j * kBlock + threadIdx.x;
#pragma unroll
                for (int c = 0; c < CC; ++c) {
                    a.dst[(a.c0 + c) * ld + i] = u[j][c];
                    q += static_cast<double>(u[j][c]) * static_cast<double>(u[j][c]);
                }
            }
            if (a.p_sums) {
                // the sums of the stored rows over each level of p (every level
                // lies inside one tile, so each is complete here and written
                // with plain stores by its run heads)
                for (int k = threadIdx.x; k < kTileRows * CC; k += kBlock) sm[k] = T(0);
                __syncthreads();
#pragma unroll
                for (int j = 0; j < kSlots; ++j) {
                    const int rank = tile_rank2(cnt, j * (kBlock / kWave) + wave, mask[j]);
                    const int end = run_end_lane(kp[j]);
                    const bool seg = run_head(kp[j]) && valid[j];
#pragma unroll
                    for (int c = 0; c < CC; ++c) {
                        const T s = run_sum_t(u[j][c], end);
                        if (seg) atomicAdd(&sm[rank * CC + c], s);
                    }
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < kSlots; ++j) {
                    const bool seg = run_head(kp[j]) && valid[j];
                    const int rank = tile_rank2(cnt, j * (kBlock / kWave) + wave, mask[j]);
                    if (!seg) continue;
                    T* ps = a.p_sums + static_cast<size_t>(kp[j]) * a.C + a.c0;
#pragma unroll
                    for (int c = 0; c < CC; ++c) ps[c] = sm[rank * CC + c];
                }
                __syncthreads();
            }
        }
        // ---- reduce into the levels of each R dimension
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (a.mode_r[k] == kRedSpell) {
                for (int m = threadIdx.x; m < kTileRows * CC; m += kBlock) sm[m] = T(0);
                int pr[kSlots];
                uint64_t smask[kSlots];
#pragma unroll
                for (int j = 0; j < kSlots; ++j) {
                    pr[j] = __shfl_up(kr[k][j], 1, kWave);
                    const int i = t0 + j * kBlock + threadIdx.x;
                    if (lane == 0 && valid[j]) pr[j] = i > t0 ? a.gid_r[k][i - 1] : INT_MIN;
                    smask[j] = __ballot(valid[j] && (pp[j] != kp[j] || pr[j] != kr[k][j]));
                    if (lane == 0) cnt[j * (kBlock / kWave) + wave] = __popcll(smask[j]);
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    int tot = 0;
                    for (int m = 0; m < kChunks; ++m) tot += cnt[m];
                    cnt_total = tot;
                }
#pragma unroll
                for (int j = 0; j < kSlots; ++j) {
                    const int rank = tile_rank2(cnt, j * (kBlock / kWave) + wave, smask[j]);
                    const int end = run_end_lane2(kp[j], kr[k][j]);
                    const bool seg = run_head2(kp[j], kr[k][j]) && valid[j];
#pragma unroll
                    for (int c = 0; c < CC; ++c) {
                        const T s = run_sum_t(u[j][c], end);
                        if (seg) atomicAdd(&sm[rank * CC + c], s);
                    }
                }
                __syncthreads();
                const int nsp = cnt_total;
                T* out = a.spell_sums[k] + static_cast<size_t>(a.spell_base[k][t]) * a.C + a.c0;
                for (int m = threadIdx.x; m < nsp * CC; m += kBlock) {
                    out[static_cast<size_t>(m / CC) * a.C + (m % CC)] = sm[m];
                }
                __syncthreads();
            } else {
                // a small dimension (years): each row adds to its wave's copy of
                // the level table (rows of one worker are consecutive years, so
                // runs of one level are a row long and a segmented sum gains
                // nothing)
                T* l = lr + a.lds_off[k] + static_cast<size_t>(wave % a.lds_copies[k]) * a.G_r[k] * CC;
#pragma unroll
                for (int j = 0; j < kSlots; ++j) {
                    if (!valid[j]) continue;
#pragma unroll
                    for (int c = 0; c < CC; ++c) atomicAdd(&l[kr[k][j] * CC + c], u[j][c]);
                }
            }
        }
    }
    __syncthreads();
    for (int k = 0; k < NR; ++k) {
        if (a.mode_r[k] != kRedLds) continue;
        const T* l = lr + a.lds_off[k];
        const int GC = a.G_r[k] * CC;
        for (int j = threadIdx.x; j < GC; j += kBlock) {
            T s = T(0);
            for (int q = 0; q < a.lds_copies[k]; ++q) s += l[static_cast<size_t>(q) * GC + j];
            if (s != T(0)) atomicAdd(&a.lds_target[k][static_cast<size_t>(j / CC) * a.C + a.c0 + (j % CC)], s);
        }
    }
    if constexpr (STORE) {
        double qq[1] = {q};
        block_sum_store<1>(qq, a.partials + blockIdx.x);
    }
}

// Level sums of one R dimension from its spell sums (spells in order of level).
template <typename T, int CC>
__global__ void __launch_bounds__(kBlock)
pull_kernel(const T* S, int C, int c0, const int* order, const int* skey, int M, T* sums) {
    const int lane = lane_id();
    for (int base = blockIdx.x * kBlock; base < M; base += gridDim.x * kBlock) {
        const int j = base + threadIdx.x;
        const bool valid = j < M;
        const int key = valid ? skey[j] : -1 - lane;
        T v[CC];
        if (valid) {
            const T* s = S + static_cast<size_t>(order[j]) * C + c0;
#pragma unroll
            for (int c = 0; c < CC; ++c) v[c] = s[c];
        } else {
#pragma unroll
            for (int c = 0; c < CC; ++c) v[c] = T(0);
        }
        const int end = run_end_lane(key);
        const bool head = run_head(key) && valid;
        const bool owned = lane > 0 && end < kWave - 1;
#pragma unroll
        for (int c = 0; c < CC; ++c) {
            const T s = run_sum_t(v[c], end);
            if (head) {
                T* t = &sums[static_cast<size_t>(key) * C + c0 + c];
                if (owned) *t = s;
                else atomicAdd(t, s);
            }
        }
    }
}

// ------------------------------------------------------------ CG vector ops
// Vectors are level-major N x C blocks (N levels of all R dims, C columns);
// scalars per column live in device arrays of C doubles.

// partials[block*C + c] = sum_j a[j,c] * b[j,c] (FP64 accumulation)
template <typename T>
__global__ void __launch_bounds__(kBlock)
vdot_kernel(const T* a, const T* b, int N, int C, double* partials) {
    __shared__ double red[kBlock / kWave][kMaxCgCols];
    double acc[kMaxCgCols];
    for (int c = 0; c < C; ++c) acc[c] = 0.0;
    const long long total = static_cast<long long>(N) * C;
    // thread-strided over levels so that a thread always owns the same columns
    for (long long g = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; g < N;
         g += static_cast<long long>(gridDim.x) * kBlock) {
        for (int c = 0; c < C; ++c)
            acc[c] += static_cast<double>(a[g * C + c]) * static_cast<double>(b[g * C + c]);
    }
    (void)total;
    for (int c = 0; c < C; ++c) {
        const double s = wave_sum(acc[c]);
        if (lane_id() == 0) red[threadIdx.x / kWave][c] = s;
    }
    __syncthreads();
    if (threadIdx.x < C) {
        double s = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) s += red[w][threadIdx.x];
        partials[static_cast<size_t>(blockIdx.x) * C + threadIdx.x] = s;
    }
}

// Device-side CG over level-major N x C blocks.  The vector kernels use a
// grid whose thread stride is a multiple of C, so every thread always meets
// the same column c = (global thread id) % C and keeps one FP64 accumulator;
// each block stores its column sums in a partials array (plain stores), and
// single-block scalar kernels reduce the partials and update the scalar block
// `cg` (layout kCg* in absorb_kernels.h).  (A last-block reduction inside the
// vector kernels needs a device-scope release per block, which writes back
// the XCD's L2 and costs more than the extra launch.)

namespace {
// part[block * C + c] = sum of v over the block's threads of column c, added
// in thread order (thread t holds column (block offset + t) % C): the same
// inputs give the same bits on every run
__device__ __forceinline__ void cg_block_store(double v, int c, int C, double* part) {
    __shared__ double red[kBlock];
    red[threadIdx.x] = v;
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < C) {
        // threads k = t, t + C, t + 2C, ... hold this thread's column c (= column of thread t)
        double t = 0.0;
        for (int k = threadIdx.x; k < kBlock; k += C) t += red[k];
        part[static_cast<size_t>(blockIdx.x) * C + c] = t;
    }
    __syncthreads();
}

// Threads of the single-block scalar kernels: the partials (up to 2048 blocks
// x 32 columns per array) are read by 16 waves, each thread with 8 loads in
// flight (a 256-thread block is latency-bound on them).
constexpr int kRed = 1024;

// out[c] = sum over nblk blocks of part[b * C + c], c < C (one block of kRed):
// T = floor(kRed / C) * C threads read the array linearly with stride T, so
// thread t always meets column t % C, with 8 independent accumulators to keep
// 8 loads in flight; then one thread per column adds the T / C thread sums.
// Two arrays (part2 may be null) are reduced in the same pass.
__device__ __forceinline__ void cg_reduce_parts(const double* part, const double* part2, int nblk, int C,
                                                double* out, double* out2) {
    __shared__ double red[2][kRed];
    const int T = (kRed / C) * C;
    const int total = nblk * C;
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double b[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (static_cast<int>(threadIdx.x) < T) {
        int e = threadIdx.x;
        for (; e + 7 * T < total; e += 8 * T) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                a[u] += part[e + u * T];
                if (part2) b[u] += part2[e + u * T];
            }
        }
        for (; e < total; e += T) {
            a[0] += part[e];
            if (part2) b[0] += part2[e];
        }
    }
    red[0][threadIdx.x] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    red[1][threadIdx.x] = ((b[0] + b[1]) + (b[2] + b[3])) + ((b[4] + b[5]) + (b[6] + b[7]));
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < C) {
        double t = 0.0, t2 = 0.0;
        for (int k = threadIdx.x; k < T; k += C) {
            t += red[0][k];
            t2 += red[1][k];
        }
        out[threadIdx.x] = t;
        if (out2) out2[threadIdx.x] = t2;
    }
    __syncthreads();
}
}  // namespace

// part1[block, c] = sum_j a[j, c] b[j, c]
template <typename T>
__global__ void __launch_bounds__(kBlock)
cg_dot_kernel(const T* a, const T* b, long long total, int C, double* part1) {
    const long long g0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    const int c = static_cast<int>(g0 % C);
    double acc = 0.0;
    for (long long j = g0; j < total; j += stride) acc += static_cast<double>(a[j]) * static_cast<double>(b[j]);
    cg_block_store(acc, c, C, part1);
}

// alpha = rz / pq for active columns; a column with pq not in (0, 1e300) or
// rz <= 0 stops (without converging)
__global__ void __launch_bounds__(kRed) cg_alpha_kernel(double* cg, const double* part1, int nblk, int C) {
    __shared__ double pqs[kMaxCgCols];
    cg_reduce_parts(part1, nullptr, nblk, C, pqs, nullptr);
    const int c = threadIdx.x;
    if (c >= C) return;
    const double pq = pqs[c];
    const double rz = cg[kCgRz + c];
    double al = 0.0;
    if (cg[kCgActive + c] != 0.0) {
        if (pq > 0.0 && pq < 1e300 && rz > 0.0) al = rz / pq;
        else cg[kCgActive + c] = 0.0;
    }
    cg[kCgAlpha + c] = al;
}

// x += alpha p; r -= alpha q; z = minv r; part1 = r.r, part2 = r.z per block and column
template <typename T>
__global__ void __launch_bounds__(kBlock)
cg_step_kernel(T* x, T* r, T* z, const T* p, const T* qv, const T* minv, long long total, int C,
               const double* cg, double* part1, double* part2) {
    const long long g0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    const int c = static_cast<int>(g0 % C);
    const T al = static_cast<T>(cg[kCgAlpha + c]);
    double rr = 0.0, rz = 0.0;
    for (long long j = g0; j < total; j += stride) {
        x[j] += al * p[j];
        const T rn = r[j] - al * qv[j];
        r[j] = rn;
        const T zn = minv[j / C] * rn;
        z[j] = zn;
        rr += static_cast<double>(rn) * static_cast<double>(rn);
        rz += static_cast<double>(rn) * static_cast<double>(zn);
    }
    cg_block_store(rr, c, C, part1);
    cg_block_store(rz, c, C, part2);
}

// per active column: one more iteration; r.r <= tol2 converges it, reaching
// max_iter stops it; otherwise beta = rz_new / rz.  cg[kCgCount] = active columns
__global__ void __launch_bounds__(kRed)
cg_beta_kernel(double* cg, const double* part1, const double* part2, int nblk, int C, double tol2,
               int max_iter) {
    __shared__ double rrs[kMaxCgCols], rzs[kMaxCgCols];
    __shared__ int cnt;
    cg_reduce_parts(part1, part2, nblk, C, rrs, rzs);
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int c = threadIdx.x;
    if (c < C) {
        double be = 0.0;
        if (cg[kCgActive + c] != 0.0) {
            const double its = cg[kCgIters + c] + 1.0;
            cg[kCgIters + c] = its;
            if (rrs[c] <= tol2) {
                cg[kCgActive + c] = 0.0;
                cg[kCgOk + c] = 1.0;
            } else if (its >= max_iter) {
                cg[kCgActive + c] = 0.0;
            } else {
                be = rzs[c] / cg[kCgRz + c];
                cg[kCgRz + c] = rzs[c];
                atomicAdd(&cnt, 1);
            }
        }
        cg[kCgBeta + c] = be;
    }
    __syncthreads();
    if (threadIdx.x == 0) cg[kCgCount] = cnt;
}

// p = z + beta p
template <typename T>
__global__ void __launch_bounds__(kBlock)
cg_dir_kernel(T* p, const T* z, const double* cg, long long total, int C) {
    const long long g0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    const T be = static_cast<T>(cg[kCgBeta + static_cast<int>(g0 % C)]);
    for (long long j = g0; j < total; j += stride) p[j] = z[j] + be * p[j];
}

// CG start: r = b * scale_c (rounded to T), x = 0, z = minv r, p = z; part1 = r.z
template <typename T>
__global__ void __launch_bounds__(kBlock)
cg_init_kernel(const double* b, const double* scale, const T* minv, T* x, T* r, T* z, T* p, long long total,
               int C, double* part1) {
    const long long g0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    const int c = static_cast<int>(g0 % C);
    const double sc = scale[c];
    double rz = 0.0;
    for (long long j = g0; j < total; j += stride) {
        const T rv = static_cast<T>(b[j] * sc);
        x[j] = T(0);
        r[j] = rv;
        const T zv = minv[j / C] * rv;
        z[j] = zv;
        p[j] = zv;
        rz += static_cast<double>(rv) * static_cast<double>(zv);
    }
    cg_block_store(rz, c, C, part1);
}

// rz = r.z; a column with a zero right-hand side is converged at once (x = 0),
// one with r.z <= 0 starts inactive
__global__ void __launch_bounds__(kRed)
cg_start_kernel(double* cg, const double* part1, int nblk, const double* scale, int C) {
    __shared__ double rzs[kMaxCgCols];
    __shared__ int cnt;
    cg_reduce_parts(part1, nullptr, nblk, C, rzs, nullptr);
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int c = threadIdx.x;
    if (c < C) {
        cg[kCgRz + c] = rzs[c];
        const bool zero = !(scale[c] > 0.0);
        const bool act = !zero && rzs[c] > 0.0;
        cg[kCgActive + c] = act ? 1.0 : 0.0;
        cg[kCgOk + c] = zero ? 1.0 : 0.0;
        cg[kCgIters + c] = 0.0;
        cg[kCgAlpha + c] = 0.0;
        cg[kCgBeta + c] = 0.0;
        if (act) atomicAdd(&cnt, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) cg[kCgCount] = cnt;
}

// beta64 += x32 / scale_c
__global__ void __launch_bounds__(kBlock)
refine_update_kernel(double* beta, const float* x, const double* inv_scale, int N, int C) {
    const long long total = static_cast<long long>(N) * C;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        beta[j] += static_cast<double>(x[j]) * inv_scale[j % C];
    }
}

// Reliable-update CG: one more iteration per active column; the iterated r.r
// at or below cg[kCgTol2 + c] asks the host for an FP64 residual (cg[kCgUpd]);
// reaching max_iter stops the column; beta = rz_new / rz.  The stop on the
// tolerance is decided on the FP64 residual only.
__global__ void __launch_bounds__(kRed)
cg_beta_rel_kernel(double* cg, const double* part1, const double* part2, int nblk, int C, int max_iter) {
    __shared__ double rrs[kMaxCgCols], rzs[kMaxCgCols];
    __shared__ int cnt, upd;
    cg_reduce_parts(part1, part2, nblk, C, rrs, rzs);
    if (threadIdx.x == 0) {
        cnt = 0;
        upd = 0;
    }
    __syncthreads();
    const int c = threadIdx.x;
    if (c < C) {
        double be = 0.0;
        if (cg[kCgActive + c] != 0.0) {
            const double its = cg[kCgIters + c] + 1.0;
            cg[kCgIters + c] = its;
            if (its >= max_iter) {
                cg[kCgActive + c] = 0.0;
            } else {
                if (rrs[c] <= cg[kCgTol2 + c]) atomicOr(&upd, 1);
                be = rzs[c] > 0.0 && cg[kCgRz + c] > 0.0 ? rzs[c] / cg[kCgRz + c] : 0.0;
                cg[kCgRz + c] = rzs[c];
                atomicAdd(&cnt, 1);
            }
        }
        cg[kCgBeta + c] = be;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        cg[kCgCount] = cnt;
        if (upd) cg[kCgUpd] = 1.0;
    }
}

// Residual replacement, vector part: r = r64 * inv_t_c (rounded to FP32),
// z = minv r.  The last step already formed p = z_iter + beta p_prev from the
// iterated residual; the direction keeps its part beta p_prev (rescaled into the
// new residual's units by f_c) and takes the new z: p = f (p - z_iter) + z.  A
// column restarted after a stop (f_c = 0) takes p = z.  x = 0; part1 = r.z per
// block and column.
__global__ void __launch_bounds__(kBlock)
cg_reset_kernel(const double* r64, const double* rel, const float* minv, float* x, float* r, float* z, float* p,
                long long total, int C, double* part1) {
    const long long g0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    const int c = static_cast<int>(g0 % C);
    const double it = rel[c];
    const float f = static_cast<float>(rel[kMaxCgCols + c]);
    double rz = 0.0;
    for (long long j = g0; j < total; j += stride) {
        const float rv = static_cast<float>(r64[j] * it);
        r[j] = rv;
        const float zv = minv[j / C] * rv;
        p[j] = f == 0.0f ? zv : f * (p[j] - z[j]) + zv;
        z[j] = zv;
        x[j] = 0.0f;
        rz += static_cast<double>(rv) * static_cast<double>(zv);
    }
    cg_block_store(rz, c, C, part1);
}

// Residual replacement, scalar part: rz of the new residual, the activity, the
// update thresholds; the update flag is cleared.
__global__ void __launch_bounds__(kRed)
cg_reset_scalars_kernel(double* cg, const double* part1, int nblk, const double* rel, int C) {
    __shared__ double rzs[kMaxCgCols];
    __shared__ int cnt;
    cg_reduce_parts(part1, nullptr, nblk, C, rzs, nullptr);
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    const int c = threadIdx.x;
    if (c < C) {
        const bool act = rel[3 * kMaxCgCols + c] != 0.0;
        cg[kCgRz + c] = rzs[c];
        cg[kCgActive + c] = (act && rzs[c] > 0.0) ? 1.0 : 0.0;
        cg[kCgAlpha + c] = 0.0;
        cg[kCgBeta + c] = 0.0;
        cg[kCgTol2 + c] = rel[2 * kMaxCgCols + c];
        if (cg[kCgActive + c] != 0.0) atomicAdd(&cnt, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        cg[kCgCount] = cnt;
        cg[kCgUpd] = 0.0;
    }
}

// Diagonal of S for the Jacobi preconditioner, from runs of equal level of r
// inside a level of p: sum over runs of n - n^2 / N_p (exact when every
// (p-level, r-level) pair forms one run).  Result is inverted in place later.
__global__ void __launch_bounds__(kBlock)
schur_diag_kernel(const int* tile_start, int num_tiles, const int* gid_p, const int* gid_r,
                  const double* inv_p, double* diag) {
    __shared__ int cnt[kChunks];
    __shared__ int cnt_total;
    __shared__ int run_n[kTileRows];
    __shared__ int run_key[kTileRows];
    __shared__ double run_ip[kTileRows];
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    for (int t = blockIdx.x; t < num_tiles; t += gridDim.x) {
        const int t0 = tile_start[t], t1 = tile_start[t + 1];
        if (t0 >= t1) continue;
        for (int k = threadIdx.x; k < kTileRows; k += kBlock) run_n[k] = 0;
        int kp[kSlots], kr[kSlots], rank[kSlots];
        bool valid[kSlots];
        uint64_t mask[kSlots];
        for (int j = 0; j < kSlots; ++j) {
            const int i = t0 + j * kBlock + threadIdx.x;
            valid[j] = i < t1;
            kp[j] = valid[j] ? gid_p[i] : -1 - lane;
            kr[j] = valid[j] ? gid_r[i] : -1 - lane;
            int pp = __shfl_up(kp[j], 1, kWave), pr = __shfl_up(kr[j], 1, kWave);
            if (lane == 0 && valid[j]) {
                pp = i > t0 ? gid_p[i - 1] : INT_MIN;
                pr = i > t0 ? gid_r[i - 1] : INT_MIN;
            }
            mask[j] = __ballot(valid[j] && (pp != kp[j] || pr != kr[j]));
            if (lane == 0) cnt[j * (kBlock / kWave) + wave] = __popcll(mask[j]);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int m = 0; m < kChunks; ++m) tot += cnt[m];
            cnt_total = tot;
        }
        for (int j = 0; j < kSlots; ++j) {
            rank[j] = tile_rank2(cnt, j * (kBlock / kWave) + wave, mask[j]);
            if (valid[j]) {
                atomicAdd(&run_n[rank[j]], 1);
                run_key[rank[j]] = kr[j];
                run_ip[rank[j]] = inv_p[kp[j]];
            }
        }
        __syncthreads();
        for (int k = threadIdx.x; k < cnt_total; k += kBlock) {
            const double nn = run_n[k];
            atomicAdd(&diag[run_key[k]], nn - nn * nn * run_ip[k]);
        }
        __syncthreads();
    }
}

// minv = 1 / max(diag, floor * count) (levels without any variation keep a finite entry)
__global__ void __launch_bounds__(kBlock)
diag_invert_kernel(const double* diag, const double* counts, float* minv, int N) {
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < N; g += gridDim.x * kBlock) {
        const double c = counts[g];
        double d = diag[g];
        const double floor = 1e-6 * (c > 0.0 ? c : 1.0);
        if (!(d > floor)) d = c > 0.0 ? c : 1.0;
        minv[g] = static_cast<float>(1.0 / d);
    }
}

__global__ void __launch_bounds__(kBlock) to_float_kernel(const double* a, float* b, long long m) {
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < m;
         j += static_cast<long long>(gridDim.x) * kBlock)
        b[j] = static_cast<float>(a[j]);
}

// ------------------------------------------------------------- launchers

template <typename T, int CC, int NR, bool HAS_V, bool STORE>
static hipError_t launch_schur_t(const SchurArgs<T>& a, size_t lds_extra, int grid, hipStream_t s) {
    const size_t shmem = static_cast<size_t>(kTileRows) * CC * sizeof(T) + lds_extra;
    hipLaunchKernelGGL((schur_kernel<T, CC, NR, HAS_V, STORE>), dim3(grid), dim3(kBlock), shmem, s, a);
    return hipGetLastError();
}

template <typename T, int CC>
static hipError_t launch_schur_cc(const SchurArgs<T>& a, int nr, bool has_v, bool store,
                                  size_t lds_extra, int grid, hipStream_t s) {
    if (nr == 1) {
        if (!has_v) return launch_schur_t<T, CC, 1, false, false>(a, lds_extra, grid, s);
        if (!store) return launch_schur_t<T, CC, 1, true, false>(a, lds_extra, grid, s);
        return launch_schur_t<T, CC, 1, true, true>(a, lds_extra, grid, s);
    }
    if (!has_v) return launch_schur_t<T, CC, 2, false, false>(a, lds_extra, grid, s);
    if (!store) return launch_schur_t<T, CC, 2, true, false>(a, lds_extra, grid, s);
    return launch_schur_t<T, CC, 2, true, true>(a, lds_extra, grid, s);
}

template <typename T>
hipError_t launch_schur(const SchurArgs<T>& a, int cc, int nr, bool has_v, bool store,
                        size_t lds_extra, int grid, hipStream_t s) {
    if (nr < 1 || nr > 2) return hipErrorInvalidValue;
    switch (cc) {
        case 1: return launch_schur_cc<T, 1>(a, nr, has_v, store, lds_extra, grid, s);
        case 2: return launch_schur_cc<T, 2>(a, nr, has_v, store, lds_extra, grid, s);
        case 3: return launch_schur_cc<T, 3>(a, nr, has_v, store, lds_extra, grid, s);
        case 4: return launch_schur_cc<T, 4>(a, nr, has_v, store, lds_extra, grid, s);
        case 5: return launch_schur_cc<T, 5>(a, nr, has_v, store, lds_extra, grid, s);
        case 6: return launch_schur_cc<T, 6>(a, nr, has_v, store, lds_extra, grid, s);
        case 7: return launch_schur_cc<T, 7>(a, nr, has_v, store, lds_extra, grid, s);
        case 8: return launch_schur_cc<T, 8>(a, nr, has_v, store, lds_extra, grid, s);
        default: return hipErrorInvalidValue;
    }
}
template hipError_t launch_schur<float>(const SchurArgs<float>&, int, int, bool, bool, size_t, int, hipStream_t);
template hipError_t launch_schur<double>(const SchurArgs<double>&, int, int, bool, bool, size_t, int, hipStream_t);

template <typename T>
hipError_t launch_pull(const T* S, int C, int c0, int cc, const int* order, const int* skey, int M,
                       T* sums, int grid, hipStream_t s) {
    switch (cc) {
#define XFE_PULLT(K) case K: hipLaunchKernelGGL((pull_kernel<T, K>), dim3(grid), dim3(kBlock), 0, s, S, C, c0, order, skey, M, sums); break;
        XFE_PULLT(1) XFE_PULLT(2) XFE_PULLT(3) XFE_PULLT(4) XFE_PULLT(5) XFE_PULLT(6) XFE_PULLT(7) XFE_PULLT(8)
#undef XFE_PULLT
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_pull<float>(const float*, int, int, int, const int*, const int*, int, float*, int, hipStream_t);
template hipError_t launch_pull<double>(const double*, int, int, int, const int*, const int*, int, double*, int, hipStream_t);

// Residual replacement with the block preconditioner, first part: r = r64 *
// inv_t_c, the kept direction p = f (p - z_iter) (0 for a restart), z = minv r
// (the blocked firms are overwritten next), x = 0.
__global__ void __launch_bounds__(kBlock)
cg_reset_pre_kernel(const double* r64, const double* rel, const float* minv, float* x, float* r, float* z, float* p,
                    long long total, int C) {
    const long long g0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    const int c = static_cast<int>(g0 % C);
    const double it = rel[c];
    const float f = static_cast<float>(rel[kMaxCgCols + c]);
    for (long long j = g0; j < total; j += stride) {
        const float rv = static_cast<float>(r64[j] * it);
        r[j] = rv;
        p[j] = f == 0.0f ? 0.0f : f * (p[j] - z[j]);
        z[j] = minv[j / C] * rv;
        x[j] = 0.0f;
    }
}

// p += z; part1 = r.z per block and column
__global__ void __launch_bounds__(kBlock)
cg_padd_dot_kernel(const float* r, const float* z, float* p, long long total, int C, double* part1) {
    const long long g0 = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    const int c = static_cast<int>(g0 % C);
    double rz = 0.0;
    for (long long j = g0; j < total; j += stride) {
        const float zv = z[j];
        p[j] += zv;
        rz += static_cast<double>(r[j]) * static_cast<double>(zv);
    }
    cg_block_store(rz, c, C, part1);
}

// ---- the Schur-complement PCG in FP64 (precision 2): level-major N x C vectors,
// per-column scalars coef[c]

// y += sign coef[c] x
__global__ void __launch_bounds__(kBlock)
s64_axpy_kernel(double* y, const double* x, const double* coef, long long total, int C, double sign) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const double a = coef[i % C];
        if (a != 0.0) y[i] += sign * a * x[i];
    }
}

// z = minv[level] r
__global__ void __launch_bounds__(kBlock)
s64_jacobi_kernel(const double* r, const float* minv, double* z, long long total, int C) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock)
        z[i] = static_cast<double>(minv[i / C]) * r[i];
}

// p = z + beta[c] p
__global__ void __launch_bounds__(kBlock)
s64_dir_kernel(double* p, const double* z, const double* beta, long long total, int C) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock)
        p[i] = z[i] + beta[i % C] * p[i];
}

hipError_t launch_s64_axpy(double* y, const double* x, const double* coef, long long total, int C, double sign,
                           int grid, hipStream_t s) {
    hipLaunchKernelGGL(s64_axpy_kernel, dim3(grid), dim3(kBlock), 0, s, y, x, coef, total, C, sign);
    return hipGetLastError();
}

hipError_t launch_s64_jacobi(const double* r, const float* minv, double* z, long long total, int C, int grid,
                             hipStream_t s) {
    hipLaunchKernelGGL(s64_jacobi_kernel, dim3(grid), dim3(kBlock), 0, s, r, minv, z, total, C);
    return hipGetLastError();
}

hipError_t launch_s64_dir(double* p, const double* z, const double* beta, long long total, int C, int grid,
                          hipStream_t s) {
    hipLaunchKernelGGL(s64_dir_kernel, dim3(grid), dim3(kBlock), 0, s, p, z, beta, total, C);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_vdot(const T* a, const T* b, int N, int C, double* partials, int grid, hipStream_t s) {
    hipLaunchKernelGGL((vdot_kernel<T>), dim3(grid), dim3(kBlock), 0, s, a, b, N, C, partials);
    return hipGetLastError();
}
template hipError_t launch_vdot<float>(const float*, const float*, int, int, double*, int, hipStream_t);
template hipError_t launch_vdot<double>(const double*, const double*, int, int, double*, int, hipStream_t);

template <typename T>
hipError_t launch_cg_dot(const T* a, const T* b, long long total, int C, double* cg, double* part, int grid,
                         hipStream_t s) {
    hipLaunchKernelGGL((cg_dot_kernel<T>), dim3(grid), dim3(kBlock), 0, s, a, b, total, C, part);
    hipLaunchKernelGGL(cg_alpha_kernel, dim3(1), dim3(kRed), 0, s, cg, part, grid, C);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_step(T* x, T* r, T* z, T* p, const T* q, const T* minv, long long total, int C, double* cg,
                          double* part, double tol2, int max_iter, int grid, hipStream_t s) {
    double* part2 = part + static_cast<size_t>(grid) * C;
    hipLaunchKernelGGL((cg_step_kernel<T>), dim3(grid), dim3(kBlock), 0, s, x, r, z, p, q, minv, total, C, cg,
                       part, part2);
    hipLaunchKernelGGL(cg_beta_kernel, dim3(1), dim3(kRed), 0, s, cg, part, part2, grid, C, tol2, max_iter);
    hipLaunchKernelGGL((cg_dir_kernel<T>), dim3(grid), dim3(kBlock), 0, s, p, z, cg, total, C);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_init(const double* b, const double* scale, const T* minv, T* x, T* r, T* z, T* p,
                          long long total, int C, double* cg, double* part, int grid, hipStream_t s) {
    hipLaunchKernelGGL((cg_init_kernel<T>), dim3(grid), dim3(kBlock), 0, s, b, scale, minv, x, r, z, p, total, C,
                       part);
    hipLaunchKernelGGL(cg_start_kernel, dim3(1), dim3(kRed), 0, s, cg, part, grid, scale, C);
    return hipGetLastError();
}

#define XFE_CG_INST(T)                                                                                         \
    template hipError_t launch_cg_dot<T>(const T*, const T*, long long, int, double*, double*, int, hipStream_t); \
    template hipError_t launch_cg_step<T>(T*, T*, T*, T*, const T*, const T*, long long, int, double*, double*,   \
                                          double, int, int, hipStream_t);                                      \
    template hipError_t launch_cg_init<T>(const double*, const double*, const T*, T*, T*, T*, T*, long long, int, \
                                          double*, double*, int, hipStream_t);
XFE_CG_INST(float)
XFE_CG_INST(double)
#undef XFE_CG_INST

hipError_t launch_refine_update(double* beta, const float* x, const double* inv_scale, int N, int C,
                                int grid, hipStream_t s) {
    hipLaunchKernelGGL(refine_update_kernel, dim3(grid), dim3(kBlock), 0, s, beta, x, inv_scale, N, C);
    return hipGetLastError();
}

hipError_t launch_cg_step_rel(float* x, float* r, float* z, float* p, const float* q, const float* minv,
                              long long total, int C, double* cg, double* part, int max_iter, int grid,
                              const lp::BlockView* blk, int firm_off, hipStream_t s) {
    double* part2 = part + static_cast<size_t>(grid) * C;
    hipLaunchKernelGGL((cg_step_kernel<float>), dim3(grid), dim3(kBlock), 0, s, x, r, z, p, q, minv, total, C, cg,
                       part, part2);
    if (blk && blk->nb > 0) {
        // the step's r.z used minv on the blocked firms too: recompute it after the block solves
        const size_t o = static_cast<size_t>(firm_off) * C;
        hipError_t e = lp::launch_block_apply(*blk, r + o, z + o, C, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((cg_dot_kernel<float>), dim3(grid), dim3(kBlock), 0, s, r, z, total, C, part2);
    }
    hipLaunchKernelGGL(cg_beta_rel_kernel, dim3(1), dim3(kRed), 0, s, cg, part, part2, grid, C, max_iter);
    hipLaunchKernelGGL((cg_dir_kernel<float>), dim3(grid), dim3(kBlock), 0, s, p, z, cg, total, C);
    return hipGetLastError();
}

hipError_t launch_cg_reliable_reset(const double* r64, const double* rel, const float* minv, float* x, float* r,
                                    float* z, float* p, long long total, int C, double* cg, double* part, int grid,
                                    const lp::BlockView* blk, int firm_off, hipStream_t s) {
    if (blk && blk->nb > 0) {
        hipLaunchKernelGGL(cg_reset_pre_kernel, dim3(grid), dim3(kBlock), 0, s, r64, rel, minv, x, r, z, p, total, C);
        const size_t o = static_cast<size_t>(firm_off) * C;
        hipError_t e = lp::launch_block_apply(*blk, r + o, z + o, C, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(cg_padd_dot_kernel, dim3(grid), dim3(kBlock), 0, s, r, z, p, total, C, part);
    } else {
        hipLaunchKernelGGL(cg_reset_kernel, dim3(grid), dim3(kBlock), 0, s, r64, rel, minv, x, r, z, p, total, C,
                           part);
    }
    hipLaunchKernelGGL(cg_reset_scalars_kernel, dim3(1), dim3(kRed), 0, s, cg, part, grid, rel, C);
    return hipGetLastError();
}

hipError_t launch_schur_diag(const int* tile_start, int num_tiles, const int* gid_p,
                             const int* gid_r, const double* inv_p, double* diag, int grid,
                             hipStream_t s) {
    hipLaunchKernelGGL(schur_diag_kernel, dim3(grid), dim3(kBlock), 0, s, tile_start, num_tiles,
                       gid_p, gid_r, inv_p, diag);
    return hipGetLastError();
}

hipError_t launch_diag_invert(const double* diag, const double* counts, float* minv, int N,
                              int grid, hipStream_t s) {
    hipLaunchKernelGGL(diag_invert_kernel, dim3(grid), dim3(kBlock), 0, s, diag, counts, minv, N);
    return hipGetLastError();
}

hipError_t launch_to_float(const double* a, float* b, long long m, int grid, hipStream_t s) {
    hipLaunchKernelGGL(to_float_kernel, dim3(grid), dim3(kBlock), 0, s, a, b, m);
    return hipGetLastError();
}

}  // namespace xfe
