// Atomic-free streaming passes over level-aligned row tiles.
//
// Rows are permuted so that one FE dimension p (the one with the most levels,
// workers in AKM panels) is level-sorted.  The rows are cut into tiles of at
// most kTileRows rows whose boundaries are level starts of p, so every level
// of p lies inside one tile and one block sees all of it:
//
//  * the projection of p is done entirely inside the pass (level sums in LDS,
//    block barrier, subtract) -- no global accumulator, no atomics, no extra
//    pass;
//  * a large unsorted dimension r (firms) is reduced through "spells": maximal
//    runs of equal r inside a level of p (a worker's stay at one firm).  A
//    spell never crosses a tile, so the block writes every spell sum with a
//    plain store into a spell array in row order; spell_pull_kernel then forms
//    the level sums of r from the spells sorted by r.  Per pass this moves
//    ~n_spells x C values instead of one atomic per row run;
//  * a small dimension (years) is privatised in LDS as in pass_kernel.
//
// Inside a tile, row j*256 + t is handled by thread t in slot j; the rank of a
// row's level (or spell) within the tile comes from ballots of level heads and
// a prefix over the (slot, wave) chunks, and the partial sums of each wave run
// are combined in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdint>

#include "absorb_kernels.h"
#include "wave_ops.h"

namespace xfe {

namespace {

constexpr int kSlots = kTileRows / kBlock;
constexpr int kChunks = kSlots * (kBlock / kWave);

// Rank (0-based) of this row's run among the runs of the tile, from the head
// flags of all rows: chunk counts in `cnt` (kChunks entries) must be complete.
__device__ __forceinline__ int tile_rank(const int* cnt, int chunk, uint64_t mask) {
    int base = 0;
    for (int k = 0; k < chunk; ++k) base += cnt[k];
    const int lane = lane_id();
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    return base + __popcll(mask & upto) - 1;
}

}  // namespace

template <int CC, int NQ, int SRC, int RR, int PP>
__global__ void __launch_bounds__(kBlock) tile_pass_kernel(TileArgs a) {
    extern __shared__ __align__(16) double dsm[];
    double* sm = dsm;                       // kTileRows x CC run sums (levels, then spells)
    double* lr = dsm + kTileRows * CC;      // G_r x CC (RR == kRedLds)
    __shared__ int cnt[kChunks];
    __shared__ int cnt_total;
    const size_t ld = static_cast<size_t>(a.ld);
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    if constexpr (RR == kRedLds) {
        for (int j = threadIdx.x; j < a.G_r * CC; j += kBlock) lr[j] = 0.0;
    }
    constexpr int NQA = NQ > 0 ? NQ : 1;
    double q[NQA];
#pragma unroll
    for (int j = 0; j < NQA; ++j) q[j] = 0.0;
    double maxp = 0.0;
    const double* __restrict__ src = SRC == 0 ? a.src + static_cast<size_t>(a.c0) * ld : nullptr;
    const double* __restrict__ P = SRC == 1 ? a.P + static_cast<size_t>(a.c0) * ld : nullptr;
    const double* __restrict__ Q = SRC == 1 ? a.Q + static_cast<size_t>(a.c0) * ld : nullptr;
    const bool apply = a.gid_a != nullptr;
    const bool weighted = a.w != nullptr;

    for (int t = blockIdx.x; t < a.num_tiles; t += gridDim.x) {
        const int t0 = a.tile_start[t];
        const int t1 = a.tile_start[t + 1];
        if (t0 >= t1) continue;  // uniform over the block
        double v[kSlots][CC];
        int kp[kSlots], kr[kSlots];
        bool valid[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
            const int i = t0 + j * kBlock + threadIdx.x;
            valid[j] = i < t1;
            kp[j] = -1 - lane;
            kr[j] = -1 - lane;
            if (valid[j]) {
                if (PP || RR == kRedSpell) kp[j] = a.gid_p[i];
                if (RR != 0) kr[j] = a.gid_r[i];
#pragma unroll
                for (int c = 0; c < CC; ++c) {
                    if constexpr (SRC == 1) {
                        const double p = P[c * ld + i];
                        v[j][c] = p - a.coef * (p - Q[c * ld + i]);
                    } else {
                        v[j][c] = src[c * ld + i];
                    }
                }
                if constexpr (SRC == 1) {
#pragma unroll
                    for (int c = 0; c < CC; ++c) a.xout[(a.c0 + c) * ld + i] = v[j][c];
                }
                if (apply) {
                    const double* m = a.means_a + static_cast<size_t>(a.gid_a[i]) * a.C + a.c0;
#pragma unroll
                    for (int c = 0; c < CC; ++c) v[j][c] -= m[c];
                }
            } else {
#pragma unroll
                for (int c = 0; c < CC; ++c) v[j][c] = 0.0;
            }
        }
        // previous row's keys (row t0 - 1 belongs to another tile: a head)
        int pp[kSlots], pr[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
            pp[j] = __shfl_up(kp[j], 1, kWave);
            pr[j] = __shfl_up(kr[j], 1, kWave);
            const int i = t0 + j * kBlock + threadIdx.x;
            if (lane == 0 && valid[j]) {
                pp[j] = i > t0 ? a.gid_p[i - 1] : INT_MIN;
                if (RR == kRedSpell) pr[j] = i > t0 ? a.gid_r[i - 1] : INT_MIN;
            }
        }

        if constexpr (PP) {
            // ---- projection of p inside the tile
            for (int k = threadIdx.x; k < kTileRows * CC; k += kBlock) sm[k] = 0.0;
            uint64_t mask[kSlots];
#pragma unroll
            for (int j = 0; j < kSlots; ++j) {
                const bool head = valid[j] && (pp[j] != kp[j]);
                mask[j] = __ballot(head);
                if (lane == 0) cnt[j * (kBlock / kWave) + wave] = __popcll(mask[j]);
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kSlots; ++j) {
                const int rank = tile_rank(cnt, j * (kBlock / kWave) + wave, mask[j]);
                const int end = run_end_lane(kp[j]);
                const bool seg = run_head(kp[j]) && valid[j];
                const double wi = (weighted && valid[j]) ? a.w[t0 + j * kBlock + threadIdx.x] : 1.0;
#pragma unroll
                for (int c = 0; c < CC; ++c) {
                    const double s = run_sum(weighted ? wi * v[j][c] : v[j][c], end);
                    if (seg) atomicAdd(&sm[rank * CC + c], s);
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kSlots; ++j) {
                if (valid[j]) {
                    const int rank = tile_rank(cnt, j * (kBlock / kWave) + wave, mask[j]);
                    const double inv = a.inv_p[kp[j]];
#pragma unroll
                    for (int c = 0; c < CC; ++c) {
                        const double m = sm[rank * CC + c] * inv;
                        v[j][c] -= m;
                        maxp = fmax(maxp, fabs(m));
                    }
                }
            }
            __syncthreads();
        }

        if constexpr (RR == kRedSpell) {
            // ---- spell sums of r (runs of equal r inside a level of p)
            for (int k = threadIdx.x; k < kTileRows * CC; k += kBlock) sm[k] = 0.0;
            uint64_t mask[kSlots];
#pragma unroll
            for (int j = 0; j < kSlots; ++j) {
                const bool head = valid[j] && (pp[j] != kp[j] || pr[j] != kr[j]);
                mask[j] = __ballot(head);
                if (lane == 0) cnt[j * (kBlock / kWave) + wave] = __popcll(mask[j]);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                int tot = 0;
                for (int k = 0; k < kChunks; ++k) tot += cnt[k];
                cnt_total = tot;
            }
#pragma unroll
            for (int j = 0; j < kSlots; ++j) {
                const int rank = tile_rank(cnt, j * (kBlock / kWave) + wave, mask[j]);
                const int end = run_end_lane2(kp[j], kr[j]);
                const bool seg = run_head2(kp[j], kr[j]) && valid[j];
                const double wi = (weighted && valid[j]) ? a.w[t0 + j * kBlock + threadIdx.x] : 1.0;
#pragma unroll
                for (int c = 0; c < CC; ++c) {
                    const double s = run_sum(weighted ? wi * v[j][c] : v[j][c], end);
                    if (seg) atomicAdd(&sm[rank * CC + c], s);
                }
            }
            __syncthreads();
            const int nsp = cnt_total;
            double* out = a.spell_sums + static_cast<size_t>(a.tile_spell_base[t]) * a.C + a.c0;
            for (int k = threadIdx.x; k < nsp * CC; k += kBlock) {
                const int sp = k / CC, c = k % CC;
                out[static_cast<size_t>(sp) * a.C + c] = sm[k];
            }
            __syncthreads();
        } else if constexpr (RR == kRedLds) {
#pragma unroll
            for (int j = 0; j < kSlots; ++j) {
                const int end = run_end_lane(kr[j]);
                const bool seg = run_head(kr[j]) && valid[j];
                const double wi = (weighted && valid[j]) ? a.w[t0 + j * kBlock + threadIdx.x] : 1.0;
#pragma unroll
                for (int c = 0; c < CC; ++c) {
                    const double s = run_sum(weighted ? wi * v[j][c] : v[j][c], end);
                    if (seg) atomicAdd(&lr[kr[j] * CC + c], s);
                }
            }
        }

        // ---- outputs
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
            if (!valid[j]) continue;
            const int i = t0 + j * kBlock + threadIdx.x;
            if (a.dst) {
#pragma unroll
                for (int c = 0; c < CC; ++c) a.dst[(a.c0 + c) * ld + i] = v[j][c];
            }
            if constexpr (NQ == 1) {
#pragma unroll
                for (int c = 0; c < CC; ++c) q[0] += v[j][c] * v[j][c];
            } else if constexpr (NQ == 3) {
#pragma unroll
                for (int c = 0; c < CC; ++c) {
                    const size_t off = (a.c0 + c) * ld + i;
                    const double gx = a.it_gx[off];
                    const double dg = v[j][c] - gx;
                    const double d2 = dg - gx + a.it_x[off];
                    q[0] += dg * d2;
                    q[1] += d2 * d2;
                    q[2] += dg * dg;
                }
            }
        }
    }
    if constexpr (RR == kRedLds) {
        __syncthreads();
        for (int j = threadIdx.x; j < a.G_r * CC; j += kBlock) {
            const double s = lr[j];
            if (s != 0.0) atomicAdd(&a.sums_r[static_cast<size_t>(j / CC) * a.C + a.c0 + (j % CC)], s);
        }
    }
    if constexpr (NQ > 0) {
        block_sum_store<NQA>(q, a.partials + static_cast<size_t>(blockIdx.x) * NQ);
    }
    if constexpr (PP) {
        if (a.maxp_partials) {
            __shared__ double red[kBlock / kWave];
            const double wm = wave_max(maxp);
            if (lane == 0) red[wave] = wm;
            __syncthreads();
            if (threadIdx.x == 0) {
                double r = 0.0;
                for (int k = 0; k < kBlock / kWave; ++k) r = fmax(r, red[k]);
                a.maxp_partials[blockIdx.x] = r;
            }
        }
    }
}

// Level sums of r from the spell sums, spells visited in order of r: runs of
// equal r inside a wave are combined; a run that neither touches lane 0 nor
// lane 63 is a whole level and is stored plainly, other runs add atomically.
template <int CC>
__global__ void __launch_bounds__(kBlock)
spell_pull_kernel(const double* S, int C, int c0, const int* order, const int* skey, int M,
                  double* sums) {
    const int lane = lane_id();
    for (int base = blockIdx.x * kBlock; base < M; base += gridDim.x * kBlock) {
        const int j = base + threadIdx.x;
        const bool valid = j < M;
        const int key = valid ? skey[j] : -1 - lane;
        double v[CC];
        if (valid) {
            const double* s = S + static_cast<size_t>(order[j]) * C + c0;
#pragma unroll
            for (int c = 0; c < CC; ++c) v[c] = s[c];
        } else {
#pragma unroll
            for (int c = 0; c < CC; ++c) v[c] = 0.0;
        }
        const int end = run_end_lane(key);
        const bool head = run_head(key) && valid;
        const bool owned = lane > 0 && end < kWave - 1;
#pragma unroll
        for (int c = 0; c < CC; ++c) {
            const double s = run_sum(v[c], end);
            if (head) {
                double* t = &sums[static_cast<size_t>(key) * C + c0 + c];
                if (owned) {
                    *t = s;
                } else {
                    atomicAdd(t, s);
                }
            }
        }
    }
}

// ---------------------------------------------------------------- set-up

// Level sizes of the sorted dimension and their maximum.
__global__ void __launch_bounds__(kBlock) level_count_kernel(const int* gid, int n, int* cnt) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) atomicAdd(&cnt[gid[i]], 1);
}

// Rows per level (unweighted weight sums) for a dimension with few levels:
// a block-private LDS histogram, flushed with one atomic per non-empty bin
__global__ void __launch_bounds__(kBlock) level_count_lds_kernel(const int* gid, int n, int G, int* cnt) {
    extern __shared__ int hist[];
    for (int j = threadIdx.x; j < G; j += kBlock) hist[j] = 0;
    __syncthreads();
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) atomicAdd(&hist[gid[i]], 1);
    __syncthreads();
    for (int j = threadIdx.x; j < G; j += kBlock)
        if (hist[j]) atomicAdd(&cnt[j], hist[j]);
}

// The same for the level-sorted dimension: a wave's run of equal ids that
// neither starts at lane 0 nor ends at lane 63 is its level's every row, so
// its head stores the count; the two edge runs of a wave add theirs with an
// atomic (a level split over waves)
__global__ void __launch_bounds__(kBlock) level_count_sorted_kernel(const int* gid, int n, int* cnt) {
    const int lane = lane_id();
    for (long long base = static_cast<long long>(blockIdx.x) * kBlock; base < n;
         base += static_cast<long long>(gridDim.x) * kBlock) {
        const long long i = base + threadIdx.x;
        const bool valid = i < n;
        const int key = valid ? gid[i] : -1 - lane;
        const int end = run_end_lane(key);
        if (valid && run_head(key)) {
            if (lane > 0 && end < kWave - 1) cnt[key] = end - lane + 1;
            else atomicAdd(&cnt[key], end - lane + 1);
        }
    }
}

__global__ void __launch_bounds__(kBlock) count_to_double_kernel(const int* cnt, int G, double* out) {
    for (int g = blockIdx.x * kBlock + threadIdx.x; g < G; g += gridDim.x * kBlock)
        out[g] = static_cast<double>(cnt[g]);
}

__global__ void __launch_bounds__(kBlock) max_int_kernel(const int* v, int m, int* out) {
    int mx = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) mx = max(mx, v[i]);
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, mx);
}

__global__ void __launch_bounds__(kBlock) max_count_kernel(const double* v, int m, int* out) {
    int mx = 0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock)
        mx = max(mx, static_cast<int>(v[i] + 0.5));
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_down(mx, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, mx);
}

// tile_start[k] = first level start of p at or after row k * step (n past the end).
__global__ void __launch_bounds__(kBlock)
tile_start_kernel(const int* gid, int n, int step, int num_tiles, int* tile_start) {
    for (int k = blockIdx.x * kBlock + threadIdx.x; k <= num_tiles; k += gridDim.x * kBlock) {
        long long r = static_cast<long long>(k) * step;
        if (k == num_tiles || r >= n) {
            tile_start[k] = n;
            continue;
        }
        int i = static_cast<int>(r);
        if (i > 0) {
            const int key = gid[i];
            while (i < n && gid[i] == key && gid[i - 1] == key) ++i;
        }
        tile_start[k] = i;
    }
}

// Spell heads: a level start of p or a change of r.
__global__ void __launch_bounds__(kBlock)
spell_head_kernel(const int* gp, const int* gr, int n, int* flags) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        flags[i] = (i == 0 || gp[i] != gp[i - 1] || gr[i] != gr[i - 1]) ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock)
spell_key_kernel(const int* flags, const int* inc, const int* gr, int n, int* skey) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        if (flags[i]) skey[inc[i] - 1] = gr[i];
}

__global__ void __launch_bounds__(kBlock)
tile_spell_base_kernel(const int* tile_start, int num_tiles, const int* inc, int n, int* base) {
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < num_tiles; k += gridDim.x * kBlock) {
        const int r = tile_start[k];
        base[k] = r < n ? inc[r] - 1 : 0;
    }
}

// ---------------------------------------------------------------- launchers

template <int CC, int NQ, int SRC, int RR, int PP>
static hipError_t launch_tile_t(const TileArgs& a, int grid, hipStream_t s) {
    size_t shmem = static_cast<size_t>(kTileRows) * CC * sizeof(double);
    if (RR == kRedLds) shmem += static_cast<size_t>(a.G_r) * CC * sizeof(double);
    hipLaunchKernelGGL((tile_pass_kernel<CC, NQ, SRC, RR, PP>), dim3(grid), dim3(kBlock), shmem, s, a);
    return hipGetLastError();
}

template <int CC, int SRC, int RR, int PP>
static hipError_t launch_tile_nq(const TileArgs& a, int nq, int grid, hipStream_t s) {
    if constexpr (SRC == 0) {
        if (nq == 1) return launch_tile_t<CC, 1, SRC, RR, PP>(a, grid, s);
        if (nq == 3) {
            if constexpr (RR == 0) return launch_tile_t<CC, 3, SRC, RR, PP>(a, grid, s);
            return hipErrorInvalidValue;
        }
    } else {
        if (nq != 0) return hipErrorInvalidValue;
    }
    return launch_tile_t<CC, 0, SRC, RR, PP>(a, grid, s);
}

template <int CC, int SRC>
static hipError_t launch_tile_rr(const TileArgs& a, int rr, bool pp, int nq, int grid, hipStream_t s) {
    switch (rr * 2 + (pp ? 1 : 0)) {
        case 0: return launch_tile_nq<CC, SRC, 0, 0>(a, nq, grid, s);
        case 1: return launch_tile_nq<CC, SRC, 0, 1>(a, nq, grid, s);
        case 2 * kRedLds: return launch_tile_nq<CC, SRC, kRedLds, 0>(a, nq, grid, s);
        case 2 * kRedLds + 1: return launch_tile_nq<CC, SRC, kRedLds, 1>(a, nq, grid, s);
        case 2 * kRedSpell: return launch_tile_nq<CC, SRC, kRedSpell, 0>(a, nq, grid, s);
        case 2 * kRedSpell + 1: return launch_tile_nq<CC, SRC, kRedSpell, 1>(a, nq, grid, s);
        default: return hipErrorInvalidValue;
    }
}

template <int CC>
static hipError_t launch_tile_cc(const TileArgs& a, int rr, bool pp, int nq, int grid, hipStream_t s) {
    if (a.P) return launch_tile_rr<CC, 1>(a, rr, pp, nq, grid, s);
    return launch_tile_rr<CC, 0>(a, rr, pp, nq, grid, s);
}

hipError_t launch_tile_pass(const TileArgs& a, int cc, int rr, bool pp, int nq, int grid,
                            hipStream_t s) {
    if (rr != 0 && a.gid_r == nullptr) return hipErrorInvalidValue;
    if ((pp || rr == kRedSpell) && a.gid_p == nullptr) return hipErrorInvalidValue;
    if (a.P && a.xout == nullptr) return hipErrorInvalidValue;
    switch (cc) {
        case 1: return launch_tile_cc<1>(a, rr, pp, nq, grid, s);
        case 2: return launch_tile_cc<2>(a, rr, pp, nq, grid, s);
        case 3: return launch_tile_cc<3>(a, rr, pp, nq, grid, s);
        case 4: return launch_tile_cc<4>(a, rr, pp, nq, grid, s);
        case 5: return launch_tile_cc<5>(a, rr, pp, nq, grid, s);
        case 6: return launch_tile_cc<6>(a, rr, pp, nq, grid, s);
        case 7: return launch_tile_cc<7>(a, rr, pp, nq, grid, s);
        case 8: return launch_tile_cc<8>(a, rr, pp, nq, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_spell_pull(const double* S, int C, int c0, int cc, const int* order,
                             const int* skey, int M, double* sums, int grid, hipStream_t s) {
    switch (cc) {
#define XFE_PULL(K) case K: hipLaunchKernelGGL(spell_pull_kernel<K>, dim3(grid), dim3(kBlock), 0, s, S, C, c0, order, skey, M, sums); break;
        XFE_PULL(1) XFE_PULL(2) XFE_PULL(3) XFE_PULL(4) XFE_PULL(5) XFE_PULL(6) XFE_PULL(7) XFE_PULL(8)
#undef XFE_PULL
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_level_count(const int* gid, int n, int* cnt, int grid, hipStream_t s) {
    hipLaunchKernelGGL(level_count_kernel, dim3(grid), dim3(kBlock), 0, s, gid, n, cnt);
    return hipGetLastError();
}

hipError_t launch_level_weights(const int* gid, int n, int G, bool sorted, int* cnt, double* wsum, int grid,
                                hipStream_t s) {
    if (sorted)
        hipLaunchKernelGGL(level_count_sorted_kernel, dim3(grid), dim3(kBlock), 0, s, gid, n, cnt);
    else if (G <= kLevelCountLds)
        hipLaunchKernelGGL(level_count_lds_kernel, dim3(grid), dim3(kBlock), sizeof(int) * G, s, gid, n, G, cnt);
    else
        return hipErrorInvalidValue;
    const int gg = static_cast<int>(std::min<long long>((G + kBlock - 1) / kBlock, 4096));
    hipLaunchKernelGGL(count_to_double_kernel, dim3(std::max(gg, 1)), dim3(kBlock), 0, s, cnt, G, wsum);
    return hipGetLastError();
}

hipError_t launch_max_count(const double* v, int m, int* out, int grid, hipStream_t s) {
    hipLaunchKernelGGL(max_count_kernel, dim3(grid), dim3(kBlock), 0, s, v, m, out);
    return hipGetLastError();
}

hipError_t launch_max_int(const int* v, int m, int* out, int grid, hipStream_t s) {
    hipLaunchKernelGGL(max_int_kernel, dim3(grid), dim3(kBlock), 0, s, v, m, out);
    return hipGetLastError();
}

hipError_t launch_tile_start(const int* gid, int n, int step, int num_tiles, int* tile_start,
                             int grid, hipStream_t s) {
    hipLaunchKernelGGL(tile_start_kernel, dim3(grid), dim3(kBlock), 0, s, gid, n, step, num_tiles,
                       tile_start);
    return hipGetLastError();
}

hipError_t launch_spell_heads(const int* gp, const int* gr, int n, int* flags, int grid,
                              hipStream_t s) {
    hipLaunchKernelGGL(spell_head_kernel, dim3(grid), dim3(kBlock), 0, s, gp, gr, n, flags);
    return hipGetLastError();
}

hipError_t launch_spell_keys(const int* flags, const int* inc, const int* gr, int n, int* skey,
                             int grid, hipStream_t s) {
    hipLaunchKernelGGL(spell_key_kernel, dim3(grid), dim3(kBlock), 0, s, flags, inc, gr, n, skey);
    return hipGetLastError();
}

hipError_t launch_tile_spell_base(const int* tile_start, int num_tiles, const int* inc, int n,
                                  int* base, int grid, hipStream_t s) {
    hipLaunchKernelGGL(tile_spell_base_kernel, dim3(grid), dim3(kBlock), 0, s, tile_start,
                       num_tiles, inc, n, base);
    return hipGetLastError();
}

}  // namespace xfe
