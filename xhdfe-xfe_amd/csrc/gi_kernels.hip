// Kernels of the group()/individual() absorber (absorb_fixed_effects_group_individual_cuda,
// src/fe_absorption_cuda.cu:4740-5219).  Rows are the groups of the
// individual dimension (one observation per group); V is n x C column-major
// (ld n).  The standard dimensions are projected by level sums and means; the
// individual dimension by a Gauss-Seidel update per individual, one launch per
// colour class (individuals of one colour share no row, so their reads and
// writes touch disjoint rows).
#include <hip/hip_runtime.h>

#include "absorb_kernels.h"
#include "wave_ops.h"

namespace xfe {

// sums[g * C + c] += w_i v[i, c] for every row i of level g.  With lds (the
// G x C table fits in the dynamic LDS), each block accumulates a private copy
// in LDS and adds its non-zero entries to sums once.
__global__ void gi_std_accumulate_kernel(const double* __restrict__ V, int n, int C, const int* __restrict__ gid,
                                         const double* __restrict__ w, double* __restrict__ sums, int GC, int lds) {
    extern __shared__ __align__(16) double lsum[];
    const int stride = gridDim.x * blockDim.x;
    if (lds) {
        for (int q = threadIdx.x; q < GC; q += blockDim.x) lsum[q] = 0.0;
        __syncthreads();
    }
    double* tab = lds ? lsum : sums;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double wi = w ? w[i] : 1.0;
        double* s = tab + static_cast<size_t>(gid[i]) * C;
        for (int c = 0; c < C; ++c) atomicAdd(&s[c], wi * V[static_cast<size_t>(c) * n + i]);
    }
    if (lds) {
        __syncthreads();
        for (int q = threadIdx.x; q < GC; q += blockDim.x)
            if (lsum[q] != 0.0) atomicAdd(&sums[q], lsum[q]);
    }
}

// v[i, c] -= sums[g * C + c] / wsum[g] (the weighted level mean).
__global__ void gi_std_apply_kernel(double* __restrict__ V, int n, int C, const int* __restrict__ gid,
                                    const double* __restrict__ sums, const double* __restrict__ wsum) {
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int g = gid[i];
        const double inv = 1.0 / wsum[g];
        const double* s = sums + static_cast<size_t>(g) * C;
        for (int c = 0; c < C; ++c) V[static_cast<size_t>(c) * n + i] -= s[c] * inv;
    }
}

// wsum[g] += w_i (counts without weights).
__global__ void gi_weight_sums_kernel(int n, const int* __restrict__ gid, const double* __restrict__ w,
                                      double* __restrict__ wsum) {
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        atomicAdd(&wsum[gid[i]], w ? w[i] : 1.0);
}

// denom[j] = sum over the groups of individual j of w_g s_g^2 (1 when not positive).
__global__ void gi_denom_kernel(const int* __restrict__ ind_ptr, const int* __restrict__ ind_group,
                                const double* __restrict__ scale, const double* __restrict__ w, int m,
                                double* __restrict__ denom) {
    const int stride = gridDim.x * blockDim.x;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) {
        double acc = 0.0;
        for (int q = ind_ptr[j]; q < ind_ptr[j + 1]; ++q) {
            const int g = ind_group[q];
            const double s = scale[g];
            acc += (w ? w[g] : 1.0) * s * s;
        }
        denom[j] = acc > 0.0 ? acc : 1.0;
    }
}

// One colour class, one wavefront per individual: for each column c the
// lanes stride over the individual's groups, a = relax * sum_g w_g s_g v[g, c]
// / denom[j] by a wave sum, then v[g, c] -= s_g a (atomic, so an incidence
// listed twice is still subtracted twice).  Individuals with thousands of
// groups no longer hold a launch back on one thread.
__global__ void gi_color_sweep_kernel(double* __restrict__ V, int n, int C, const int* __restrict__ ind_ptr,
                                      const int* __restrict__ ind_group, const double* __restrict__ scale,
                                      const double* __restrict__ w, const double* __restrict__ denom,
                                      const int* __restrict__ members, int count, double relax) {
    const int lane = lane_id();
    const int waves = gridDim.x * (blockDim.x / kWave);
    for (int t = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; t < count; t += waves) {
        const int j = members[t];
        const int b = ind_ptr[j], e = ind_ptr[j + 1];
        const double inv = relax / denom[j];
        for (int c = 0; c < C; ++c) {
            double* v = V + static_cast<size_t>(c) * n;
            double num = 0.0;
            for (int q = b + lane; q < e; q += kWave) {
                const int g = ind_group[q];
                num += (w ? w[g] : 1.0) * scale[g] * v[g];
            }
            const double a = __shfl(wave_sum(num), 0, kWave) * inv;  // the total, from lane 0
            for (int q = b + lane; q < e; q += kWave) {
                const int g = ind_group[q];
                atomicAdd(&v[g], -(scale[g] * a));
            }
        }
    }
}

// Per-block sums of squares of the first `total` values (one partial per block).
__global__ void gi_sumsq_kernel(const double* __restrict__ V, long long total, double* __restrict__ partials) {
    __shared__ double red[kBlock];
    double acc = 0.0;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
        const double v = V[i];
        acc += v * v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

hipError_t launch_gi_std_accumulate(const double* V, int n, int C, const int* gid, const double* w, double* sums,
                                    int G, int grid, hipStream_t s) {
    const int GC = G * C;
    const size_t bytes = static_cast<size_t>(GC) * sizeof(double);
    const int lds = bytes <= 48 * 1024 ? 1 : 0;
    hipLaunchKernelGGL(gi_std_accumulate_kernel, dim3(grid), dim3(kBlock), lds ? bytes : 0, s, V, n, C, gid, w,
                       sums, GC, lds);
    return hipGetLastError();
}

hipError_t launch_gi_std_apply(double* V, int n, int C, const int* gid, const double* sums, const double* wsum,
                               int grid, hipStream_t s) {
    hipLaunchKernelGGL(gi_std_apply_kernel, dim3(grid), dim3(kBlock), 0, s, V, n, C, gid, sums, wsum);
    return hipGetLastError();
}

hipError_t launch_gi_weight_sums(int n, const int* gid, const double* w, double* wsum, int grid, hipStream_t s) {
    hipLaunchKernelGGL(gi_weight_sums_kernel, dim3(grid), dim3(kBlock), 0, s, n, gid, w, wsum);
    return hipGetLastError();
}

hipError_t launch_gi_denom(const int* ind_ptr, const int* ind_group, const double* scale, const double* w, int m,
                           double* denom, int grid, hipStream_t s) {
    hipLaunchKernelGGL(gi_denom_kernel, dim3(grid), dim3(kBlock), 0, s, ind_ptr, ind_group, scale, w, m, denom);
    return hipGetLastError();
}

hipError_t launch_gi_color_sweep(double* V, int n, int C, const int* ind_ptr, const int* ind_group,
                                 const double* scale, const double* w, const double* denom, const int* members,
                                 int count, double relax, int grid, hipStream_t s) {
    hipLaunchKernelGGL(gi_color_sweep_kernel, dim3(grid), dim3(kBlock), 0, s, V, n, C, ind_ptr, ind_group, scale, w,
                       denom, members, count, relax);
    return hipGetLastError();
}

hipError_t launch_gi_sumsq(const double* V, long long total, double* partials, int grid, hipStream_t s) {
    hipLaunchKernelGGL(gi_sumsq_kernel, dim3(grid), dim3(kBlock), 0, s, V, total, partials);
    return hipGetLastError();
}

}  // namespace xfe
