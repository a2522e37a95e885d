// Wave64 building blocks shared by the streaming kernels: lane reductions,
// run (segment) detection over consecutive lanes and segmented sums.
#ifndef XFE_WAVE_OPS_H
#define XFE_WAVE_OPS_H

#include <hip/hip_runtime.h>

#include <cstdint>

#include "absorb_kernels.h"

namespace xfe {

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


// Run end / head when a run is a stretch of equal (k1, k2) pairs.
__device__ __forceinline__ int run_end_lane2(int k1, int k2) {
    const int lane = lane_id();
    const int p1 = __shfl_up(k1, 1, kWave);
    const int p2 = __shfl_up(k2, 1, kWave);
    const bool head = (lane == 0) || (p1 != k1) || (p2 != k2);
    const uint64_t heads = __ballot(head);
    const uint64_t above = (lane == 63) ? 0ull : (heads & ~((2ull << lane) - 1ull));
    return above ? (__builtin_ctzll(above) - 1) : 63;
}

__device__ __forceinline__ bool run_head2(int k1, int k2) {
    const int lane = lane_id();
    const int p1 = __shfl_up(k1, 1, kWave);
    const int p2 = __shfl_up(k2, 1, kWave);
    return (lane == 0) || (p1 != k1) || (p2 != k2);
}

}  // namespace xfe

#endif
