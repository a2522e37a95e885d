// The learned preconditioner's graph network on the device: forward, loss and
// backward over the worker-firm mobility graph, and the Adam step.  The dense
// transforms run on BF16 MFMA (v_mfma_f32_32x32x8_bf16, FP32 accumulation); the
// message passing is gathers over CSR lists (one wavefront per node, one lane
// per feature), both directions, so the backward pass needs no atomics except
// for the firm-pair head.
//
// Network (width 64, three message-passing layers; the parameter vector is
// packed in the order of lp_gnn.h's offsets):
//   hf = relu(xf Wf0 + bf0), hw = relu(xw Ww0 + bw0)                (8 input features each)
//   for l = 1..3:  hw = relu([hw | mean_{f in N(w)} hf] Wwl + bwl)
//                  hf = relu([hf | mean_{w in N(f)} hw] Wfl + bfl)
//   u_f = hf . wa + ba  (per-level score),  t_fg = sum_k hf_k hg_k we_k + be + ws q_fg
//   (q_fg: the pair's own feature, its coupling strength)
// Loss: mean over firms with a target of (u_f - u*_f)^2 plus the mean binary
// cross-entropy of sigmoid(t_fg) against y*_fg over the firm pairs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "common.h"
#include "lp_gnn.h"
#include "xfe_hip.h"

namespace xfe {
namespace lp {

namespace {

constexpr int kB = 256;   // threads per block (4 waves)
constexpr int kW = 64;    // wave size
constexpr int kPart = 67; // per-block partials of a head: 64 weight gradients, loss, bias gradient, ws gradient
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

int grid_of(long long work, int per_block, int cap = 4096) {
    const long long g = (work + per_block - 1) / per_block;
    return static_cast<int>(std::max(1LL, std::min(g, static_cast<long long>(cap))));
}

// float -> bfloat16 bits, round to nearest, ties to even (finite inputs)
__device__ __forceinline__ short to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<short>(u >> 16);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 v;
    v[0] = to_bf16(a);
    v[1] = to_bf16(b);
    v[2] = to_bf16(c);
    v[3] = to_bf16(d);
    return v;
}

// ---------------------------------------------------------------- GEMMs
// MFMA 32x32x8 operand layout (one wavefront): lane l, i = l % 32, h = l / 32.
// A (32 x 8): A[i][4h + t], t = 0..3.  B (8 x 32): B[4h + t][i].
// D (32 x 32): acc[v] = D[8 (v / 4) + 4 h + (v % 4)][i].

// Y (M x 64) = act([X1 | X2] W + b); X1: M x K1 (ld K1), X2: M x 64 or null
// (K = K1 + 64 when present), W: K x 64 row-major.  K1 is 8 or 64.
template <int K1, bool TWO>
__global__ void __launch_bounds__(kB) gemm_fwd_kernel(const float* __restrict__ X1, const float* __restrict__ X2,
                                                      int M, const float* __restrict__ W,
                                                      const float* __restrict__ bias, float* __restrict__ Y) {
    constexpr int K = K1 + (TWO ? 64 : 0);
    __shared__ __align__(16) short wt[64 * K];  // W transposed: wt[j * K + k]
    for (int e = threadIdx.x; e < K * 64; e += kB) {
        const int k = e / 64, j = e % 64;
        wt[j * K + k] = to_bf16(W[e]);
    }
    __syncthreads();
    const int lane = threadIdx.x & (kW - 1);
    const int i = lane & 31, h = lane >> 5;
    const int wave = threadIdx.x / kW;
    for (int r0 = (blockIdx.x * 4 + wave) * 32; r0 < M; r0 += gridDim.x * 4 * 32) {
        f32x16 acc0 = {}, acc1 = {};
        const int row = r0 + i;
        const bool ok = row < M;
#pragma unroll
        for (int kk = 0; kk < K; kk += 8) {
            const int k = kk + 4 * h;
            float4 xa = make_float4(0.f, 0.f, 0.f, 0.f);
            if (ok) {
                if (k < K1) xa = *reinterpret_cast<const float4*>(X1 + static_cast<size_t>(row) * K1 + k);
                else xa = *reinterpret_cast<const float4*>(X2 + static_cast<size_t>(row) * 64 + (k - K1));
            }
            const bf16x4 a = pack4(xa.x, xa.y, xa.z, xa.w);
            const bf16x4 b0 = *reinterpret_cast<const bf16x4*>(&wt[i * K + k]);
            const bf16x4 b1 = *reinterpret_cast<const bf16x4*>(&wt[(32 + i) * K + k]);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b1, acc1, 0, 0, 0);
        }
        const float bi0 = bias[i], bi1 = bias[32 + i];
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int rr = r0 + 8 * (v / 4) + 4 * h + (v % 4);
            if (rr < M) {
                Y[static_cast<size_t>(rr) * 64 + i] = fmaxf(acc0[v] + bi0, 0.f);
                Y[static_cast<size_t>(rr) * 64 + 32 + i] = fmaxf(acc1[v] + bi1, 0.f);
            }
        }
    }
}

// [D1 | D2] (M x 128) = dZ (M x 64) W^T, W: 128 x 64 row-major; D2 may be null.
__global__ void __launch_bounds__(kB) gemm_dx_kernel(const float* __restrict__ dZ, int M,
                                                     const float* __restrict__ W, float* __restrict__ D1,
                                                     float* __restrict__ D2) {
    __shared__ __align__(16) short ws[128 * 64];  // W as bf16, row j (input feature), 64 contiguous k
    for (int e = threadIdx.x; e < 128 * 64; e += kB) ws[e] = to_bf16(W[e]);
    __syncthreads();
    const int lane = threadIdx.x & (kW - 1);
    const int i = lane & 31, h = lane >> 5;
    const int wave = threadIdx.x / kW;
    for (int r0 = (blockIdx.x * 4 + wave) * 32; r0 < M; r0 += gridDim.x * 4 * 32) {
        f32x16 acc[4] = {{}, {}, {}, {}};
        const int row = r0 + i;
        const bool ok = row < M;
#pragma unroll
        for (int kk = 0; kk < 64; kk += 8) {
            const int k = kk + 4 * h;
            float4 xa = make_float4(0.f, 0.f, 0.f, 0.f);
            if (ok) xa = *reinterpret_cast<const float4*>(dZ + static_cast<size_t>(row) * 64 + k);
            const bf16x4 a = pack4(xa.x, xa.y, xa.z, xa.w);
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const bf16x4 b = *reinterpret_cast<const bf16x4*>(&ws[(cb * 32 + i) * 64 + k]);
                acc[cb] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[cb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int rr = r0 + 8 * (v / 4) + 4 * h + (v % 4);
            if (rr < M) {
                D1[static_cast<size_t>(rr) * 64 + i] = acc[0][v];
                D1[static_cast<size_t>(rr) * 64 + 32 + i] = acc[1][v];
                if (D2) {
                    D2[static_cast<size_t>(rr) * 64 + i] = acc[2][v];
                    D2[static_cast<size_t>(rr) * 64 + 32 + i] = acc[3][v];
                }
            }
        }
    }
}

// Partial dW = [X1 | X2]^T dZ over the rows of this block (K x 64 per block,
// summed over its waves in LDS): X1 M x K1 (K1 = 8 or 64), X2 M x 64 or null.
template <int K1, bool TWO>
__global__ void __launch_bounds__(kB) gemm_dw_kernel(const float* __restrict__ X1, const float* __restrict__ X2,
                                                     const float* __restrict__ dZ, int M, float* __restrict__ part) {
    constexpr int K = K1 + (TWO ? 64 : 0);
    constexpr int NKB = (K + 31) / 32;  // 32-row blocks of the K x 64 result
    __shared__ float red[K * 64];
    for (int e = threadIdx.x; e < K * 64; e += kB) red[e] = 0.f;
    __syncthreads();
    const int lane = threadIdx.x & (kW - 1);
    const int i = lane & 31, h = lane >> 5;
    const int wave = threadIdx.x / kW;
    f32x16 acc[NKB][2];
#pragma unroll
    for (int a = 0; a < NKB; ++a) {
        acc[a][0] = f32x16{};
        acc[a][1] = f32x16{};
    }
    for (int m0 = (blockIdx.x * 4 + wave) * 8; m0 < M; m0 += gridDim.x * 4 * 8) {
        float dz0[4], dz1[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int m = m0 + 4 * h + t;
            const bool ok = m < M;
            dz0[t] = ok ? dZ[static_cast<size_t>(m) * 64 + i] : 0.f;
            dz1[t] = ok ? dZ[static_cast<size_t>(m) * 64 + 32 + i] : 0.f;
        }
        const bf16x4 b0 = pack4(dz0[0], dz0[1], dz0[2], dz0[3]);
        const bf16x4 b1 = pack4(dz1[0], dz1[1], dz1[2], dz1[3]);
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
            const int kcol = kb * 32 + i;  // row of the result = column of [X1 | X2]
            float xv[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int m = m0 + 4 * h + t;
                float v = 0.f;
                if (m < M && kcol < K) {
                    v = kcol < K1 ? X1[static_cast<size_t>(m) * K1 + kcol]
                                  : X2[static_cast<size_t>(m) * 64 + (kcol - K1)];
                }
                xv[t] = v;
            }
            const bf16x4 a = pack4(xv[0], xv[1], xv[2], xv[3]);
            acc[kb][0] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b0, acc[kb][0], 0, 0, 0);
            acc[kb][1] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b1, acc[kb][1], 0, 0, 0);
        }
    }
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int r = kb * 32 + 8 * (v / 4) + 4 * h + (v % 4);
            if (r < K) {
                atomicAdd(&red[r * 64 + i], acc[kb][0][v]);
                atomicAdd(&red[r * 64 + 32 + i], acc[kb][1][v]);
            }
        }
    __syncthreads();
    for (int e = threadIdx.x; e < K * 64; e += kB) part[static_cast<size_t>(blockIdx.x) * K * 64 + e] = red[e];
}

// out[e] = sum over blocks of part[b][e] (double accumulation), e < n
__global__ void __launch_bounds__(kB) reduce_parts_kernel(const float* __restrict__ part, int nblk, int n,
                                                          float* __restrict__ out) {
    for (int e = blockIdx.x * kB + threadIdx.x; e < n; e += gridDim.x * kB) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += part[static_cast<size_t>(b) * n + e];
        out[e] = static_cast<float>(s);
    }
}

// column sums of dZ (M x 64) times the relu mask of H: partial per block
__global__ void __launch_bounds__(kB) colsum_kernel(const float* __restrict__ dZ, int M, float* __restrict__ part) {
    __shared__ float red[kB];
    const int c = threadIdx.x & 63, r = threadIdx.x >> 6;
    float s = 0.f;
    for (int m = blockIdx.x * 4 + r; m < M; m += gridDim.x * 4) s += dZ[static_cast<size_t>(m) * 64 + c];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < 64)
        part[static_cast<size_t>(blockIdx.x) * 64 + c] = red[c] + red[64 + c] + red[128 + c] + red[192 + c];
}

// ---------------------------------------------------------------- message passing
// out[v] = mean over src[idx[ptr[v] .. ptr[v] + min(deg, cap))] (0 when
// empty); one wave per v: a row of 64 features is 16 lanes x float4, so the
// wave's four 16-lane groups take rows q, q + 1, q + 2, q + 3 of the list and
// each group keeps two rows in flight (eight per wave); the groups' sums meet
// by two lane exchanges (xor 16, xor 32) in a fixed order
__global__ void __launch_bounds__(kB) mean_gather_kernel(const int* __restrict__ ptr, const int* __restrict__ idx,
                                                         const float* __restrict__ src, int n, int cap,
                                                         float* __restrict__ out) {
    const int lane = threadIdx.x & (kW - 1);
    const int grp = lane >> 4, c4 = lane & 15;
    const float4* __restrict__ src4 = reinterpret_cast<const float4*>(src);
    for (int v = blockIdx.x * 4 + threadIdx.x / kW; v < n; v += gridDim.x * 4) {
        const int b = ptr[v], e = ptr[v + 1] - b > cap ? b + cap : ptr[v + 1];  // no b + cap overflow
        float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0;
        int q = b + grp;
        for (; q + 4 < e; q += 8) {
            const float4 x0 = src4[static_cast<size_t>(idx[q]) * 16 + c4];
            const float4 x1 = src4[static_cast<size_t>(idx[q + 4]) * 16 + c4];
            s0.x += x0.x; s0.y += x0.y; s0.z += x0.z; s0.w += x0.w;
            s1.x += x1.x; s1.y += x1.y; s1.z += x1.z; s1.w += x1.w;
        }
        if (q < e) {
            const float4 x0 = src4[static_cast<size_t>(idx[q]) * 16 + c4];
            s0.x += x0.x; s0.y += x0.y; s0.z += x0.z; s0.w += x0.w;
        }
        float4 t = make_float4(s0.x + s1.x, s0.y + s1.y, s0.z + s1.z, s0.w + s1.w);
        t.x += __shfl_xor(t.x, 16); t.y += __shfl_xor(t.y, 16); t.z += __shfl_xor(t.z, 16); t.w += __shfl_xor(t.w, 16);
        t.x += __shfl_xor(t.x, 32); t.y += __shfl_xor(t.y, 32); t.z += __shfl_xor(t.z, 32); t.w += __shfl_xor(t.w, 32);
        if (grp == 0) {
            const float d = static_cast<float>(e - b);
            reinterpret_cast<float4*>(out)[static_cast<size_t>(v) * 16 + c4] =
                e > b ? make_float4(t.x / d, t.y / d, t.z / d, t.w / d) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

// mean_gather for short lists (the movers' firms): each 16-lane group of a
// wave takes a node of its own (four per wave) and walks its list two rows
// at a time, so four independent index-then-row load chains run per wave
__global__ void __launch_bounds__(kB) mean_gather_short_kernel(const int* __restrict__ ptr,
                                                               const int* __restrict__ idx,
                                                               const float* __restrict__ src, int n, int cap,
                                                               float* __restrict__ out) {
    const int lane = threadIdx.x & (kW - 1);
    const int grp = lane >> 4, c4 = lane & 15;
    const float4* __restrict__ src4 = reinterpret_cast<const float4*>(src);
    for (int v = (blockIdx.x * 4 + threadIdx.x / kW) * 4 + grp; v < n; v += gridDim.x * 16) {
        const int b = ptr[v], e = ptr[v + 1] - b > cap ? b + cap : ptr[v + 1];
        float4 s0 = make_float4(0.f, 0.f, 0.f, 0.f), s1 = s0;
        int q = b;
        for (; q + 1 < e; q += 2) {
            const float4 x0 = src4[static_cast<size_t>(idx[q]) * 16 + c4];
            const float4 x1 = src4[static_cast<size_t>(idx[q + 1]) * 16 + c4];
            s0.x += x0.x; s0.y += x0.y; s0.z += x0.z; s0.w += x0.w;
            s1.x += x1.x; s1.y += x1.y; s1.z += x1.z; s1.w += x1.w;
        }
        if (q < e) {
            const float4 x0 = src4[static_cast<size_t>(idx[q]) * 16 + c4];
            s0.x += x0.x; s0.y += x0.y; s0.z += x0.z; s0.w += x0.w;
        }
        const float d = static_cast<float>(e - b);
        reinterpret_cast<float4*>(out)[static_cast<size_t>(v) * 16 + c4] =
            e > b ? make_float4((s0.x + s1.x) / d, (s0.y + s1.y) / d, (s0.z + s1.z) / d, (s0.w + s1.w) / d)
                  : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// transpose of mean_gather seen from the sources: acc[u] += sum over the
// destinations v listed for u (rptr/ridx) of g[v] / min(deg(v), cap), over the
// edges whose position in v's list (rank; null: all) is below cap
__global__ void __launch_bounds__(kB) mean_scatter_kernel(const int* __restrict__ rptr, const int* __restrict__ ridx,
                                                          const int* __restrict__ rank, const int* __restrict__ dptr,
                                                          const float* __restrict__ g, int n, int cap,
                                                          float* __restrict__ acc) {
    const int lane = threadIdx.x & (kW - 1);
    for (int u = blockIdx.x * 4 + threadIdx.x / kW; u < n; u += gridDim.x * 4) {
        float s = 0.f;
        for (int q = rptr[u]; q < rptr[u + 1]; ++q) {
            if (rank && rank[q] >= cap) continue;
            const int v = ridx[q];
            const int d = min(dptr[v + 1] - dptr[v], cap);
            s += g[static_cast<size_t>(v) * 64 + lane] / static_cast<float>(d);
        }
        acc[static_cast<size_t>(u) * 64 + lane] += s;
    }
}

// dZ = dH * (H > 0) in place
__global__ void __launch_bounds__(kB) relu_mask_kernel(float* __restrict__ dH, const float* __restrict__ Hv,
                                                       long long m) {
    for (long long e = blockIdx.x * static_cast<long long>(kB) + threadIdx.x; e < m;
         e += static_cast<long long>(gridDim.x) * kB)
        if (!(Hv[e] > 0.f)) dH[e] = 0.f;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kW);
    return v;
}

// ---------------------------------------------------------------- heads and loss
// u_f = hf . wa + ba
__global__ void __launch_bounds__(kB) head_u_kernel(const float* __restrict__ hf, int F, const float* __restrict__ th,
                                                    float* __restrict__ u) {
    const int lane = threadIdx.x & (kW - 1);
    const float w = th[kOffWa + lane], b = th[kOffBa];
    for (int f = blockIdx.x * 4 + threadIdx.x / kW; f < F; f += gridDim.x * 4) {
        const float s = wave_sum_f(hf[static_cast<size_t>(f) * 64 + lane] * w);
        if (lane == 0) u[f] = s + b;
    }
}

// t_p = sum_k hf[pf]_k hf[pg]_k we_k + be + ws q_p
__global__ void __launch_bounds__(kB) head_t_kernel(const float* __restrict__ hf, const int* __restrict__ pf,
                                                    const int* __restrict__ pg, const float* __restrict__ pq, int np,
                                                    const float* __restrict__ th, float* __restrict__ t) {
    const int lane = threadIdx.x & (kW - 1);
    const float w = th[kOffWe + lane], b = th[kOffBe], ws = th[kOffWs];
    for (int p = blockIdx.x * 4 + threadIdx.x / kW; p < np; p += gridDim.x * 4) {
        const float s = wave_sum_f(hf[static_cast<size_t>(pf[p]) * 64 + lane] * hf[static_cast<size_t>(pg[p]) * 64 + lane] * w);
        if (lane == 0) t[p] = s + b + (pq ? ws * pq[p] : 0.0f);
    }
}

// firm part of the loss: du_f, sum of m (u - u*)^2 and of m (per block), and the
// head gradients wa (per block, 64) and ba; dhf = du wa
__global__ void __launch_bounds__(kB) loss_u_kernel(const float* __restrict__ hf, const float* __restrict__ u,
                                                    const float* __restrict__ ut, const float* __restrict__ um, int F,
                                                    float inv_nm, const float* __restrict__ th, float* __restrict__ dhf,
                                                    double* __restrict__ part) {
    __shared__ float gw[4][64];
    __shared__ double ls[4];
    const int lane = threadIdx.x & (kW - 1), wave = threadIdx.x / kW;
    const float w = th[kOffWa + lane];
    float g = 0.f, gb = 0.f;
    double l = 0.0;
    for (int f = blockIdx.x * 4 + wave; f < F; f += gridDim.x * 4) {
        const float d = um[f] * (u[f] - ut[f]);
        const float du = 2.f * d * inv_nm;
        const float hv = hf[static_cast<size_t>(f) * 64 + lane];
        g += du * hv;
        gb += du;
        l += static_cast<double>(d) * (u[f] - ut[f]);
        dhf[static_cast<size_t>(f) * 64 + lane] = du * w;
    }
    gw[wave][lane] = g;
    if (lane == 0) ls[wave] = l;
    __syncthreads();
    if (threadIdx.x < 64) {
        double* o = part + static_cast<size_t>(blockIdx.x) * kPart;
        o[threadIdx.x] = static_cast<double>(gw[0][threadIdx.x]) + gw[1][threadIdx.x] + gw[2][threadIdx.x] + gw[3][threadIdx.x];
    }
    if (threadIdx.x == 0) {
        part[static_cast<size_t>(blockIdx.x) * kPart + 64] = ls[0] + ls[1] + ls[2] + ls[3];
        part[static_cast<size_t>(blockIdx.x) * kPart + 66] = 0.0;
    }
    // ba gradient: the lane-0 sums of du across waves (every lane holds the same gb)
    __shared__ float gbs[4];
    if (lane == 0) gbs[wave] = gb;
    __syncthreads();
    if (threadIdx.x == 0) part[static_cast<size_t>(blockIdx.x) * kPart + 65] = static_cast<double>(gbs[0]) + gbs[1] + gbs[2] + gbs[3];
}

// pair part: dt = (sigmoid(t) - y) / np; BCE sum; we / be gradients (per
// block); dhf[pf] += dt we hf[pg] and symmetric (FP32 atomics)
__global__ void __launch_bounds__(kB) loss_t_kernel(const float* __restrict__ hf, const float* __restrict__ t,
                                                    const float* __restrict__ y, const int* __restrict__ pf,
                                                    const int* __restrict__ pg, const float* __restrict__ pq, int np,
                                                    float inv_np, const float* __restrict__ th,
                                                    float* __restrict__ dhf, double* __restrict__ part) {
    __shared__ float gw[4][64];
    __shared__ double ls[4];
    __shared__ float gbs[4], gss[4];
    const int lane = threadIdx.x & (kW - 1), wave = threadIdx.x / kW;
    const float w = th[kOffWe + lane];
    float g = 0.f, gb = 0.f, gs = 0.f;
    double l = 0.0;
    for (int p = blockIdx.x * 4 + wave; p < np; p += gridDim.x * 4) {
        const float tv = t[p], yv = y[p];
        const float sg = 1.f / (1.f + expf(-tv));
        const float dt = (sg - yv) * inv_np;
        const size_t a = static_cast<size_t>(pf[p]) * 64 + lane, b = static_cast<size_t>(pg[p]) * 64 + lane;
        const float ha = hf[a], hb = hf[b];
        g += dt * ha * hb;
        gb += dt;
        gs += pq ? dt * pq[p] : 0.f;
        l += static_cast<double>(fmaxf(tv, 0.f) - tv * yv + log1pf(expf(-fabsf(tv))));
        atomicAdd(&dhf[a], dt * w * hb);
        atomicAdd(&dhf[b], dt * w * ha);
    }
    gw[wave][lane] = g;
    if (lane == 0) {
        ls[wave] = l;
        gbs[wave] = gb;
        gss[wave] = gs;
    }
    __syncthreads();
    double* o = part + static_cast<size_t>(blockIdx.x) * kPart;
    if (threadIdx.x < 64)
        o[threadIdx.x] = static_cast<double>(gw[0][threadIdx.x]) + gw[1][threadIdx.x] + gw[2][threadIdx.x] + gw[3][threadIdx.x];
    if (threadIdx.x == 0) {
        o[64] = ls[0] + ls[1] + ls[2] + ls[3];
        o[65] = static_cast<double>(gbs[0]) + gbs[1] + gbs[2] + gbs[3];
        o[66] = static_cast<double>(gss[0]) + gss[1] + gss[2] + gss[3];
    }
}

// Adam (Kingma & Ba 2015) with bias correction; the step count enters as the
// correction factors c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
__global__ void __launch_bounds__(kB) adam_kernel(float* th, float* m, float* v, const float* g, int n, float lr,
                                                  float b1, float b2, float c1, float c2, float eps) {
    for (int e = blockIdx.x * kB + threadIdx.x; e < n; e += gridDim.x * kB) {
        const float gg = g[e];
        const float mm = b1 * m[e] + (1.f - b1) * gg;
        const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
        m[e] = mm;
        v[e] = vv;
        th[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    }
}

}  // namespace

// ---------------------------------------------------------------- host side

void Gnn::ensure(int F_, int Wm_, int np_, bool keep_) {
    F = F_;
    Wm = Wm_;
    np = np_;
    keep = keep_;
    const int layers = keep ? kLayers + 1 : 2;
    hf.ensure(static_cast<size_t>(std::max(F, 1)) * 64 * layers);
    hw.ensure(static_cast<size_t>(std::max(Wm, 1)) * 64 * layers);
    af.ensure(static_cast<size_t>(std::max(F, 1)) * 64 * (keep ? kLayers : 1));
    aw.ensure(static_cast<size_t>(std::max(Wm, 1)) * 64 * (keep ? kLayers : 1));
    u.ensure(std::max(F, 1));
    t.ensure(std::max(np, 1));
}

float* Gnn::hf_at(int l) { return hf.p + static_cast<size_t>(keep ? l : (l & 1)) * F * 64; }
float* Gnn::hw_at(int l) { return hw.p + static_cast<size_t>(keep ? l : (l & 1)) * Wm * 64; }
float* Gnn::af_at(int l) { return af.p + static_cast<size_t>(keep ? l - 1 : 0) * F * 64; }
float* Gnn::aw_at(int l) { return aw.p + static_cast<size_t>(keep ? l - 1 : 0) * Wm * 64; }

void Gnn::forward(const Graph& g, const float* th, hipStream_t s) {
    const int gf = grid_of(F, 128), gw = grid_of(Wm, 128);
    if (F > 0)
        hipLaunchKernelGGL((gemm_fwd_kernel<8, false>), dim3(gf), dim3(kB), 0, s, g.xf, nullptr, F, th + kOffWf0,
                           th + kOffBf0, hf_at(0));
    if (Wm > 0)
        hipLaunchKernelGGL((gemm_fwd_kernel<8, false>), dim3(gw), dim3(kB), 0, s, g.xw, nullptr, Wm, th + kOffWw0,
                           th + kOffBw0, hw_at(0));
    for (int l = 1; l <= kLayers; ++l) {
        if (Wm > 0) {
            hipLaunchKernelGGL(mean_gather_short_kernel, dim3(grid_of(Wm, 16)), dim3(kB), 0, s, g.w_ptr, g.w_idx,
                               hf_at(l - 1), Wm, INT_MAX, aw_at(l));
            hipLaunchKernelGGL((gemm_fwd_kernel<64, true>), dim3(gw), dim3(kB), 0, s, hw_at(l - 1), aw_at(l), Wm,
                               th + off_Ww(l), th + off_bw(l), hw_at(l));
        }
        if (F > 0) {
            hipLaunchKernelGGL(mean_gather_kernel, dim3(grid_of(F, 4)), dim3(kB), 0, s, g.f_ptr, g.f_idx, hw_at(l), F,
                               kFirmCap, af_at(l));
            hipLaunchKernelGGL((gemm_fwd_kernel<64, true>), dim3(gf), dim3(kB), 0, s, hf_at(l - 1), af_at(l), F,
                               th + off_Wf(l), th + off_bf(l), hf_at(l));
        }
    }
    if (F > 0) hipLaunchKernelGGL(head_u_kernel, dim3(grid_of(F, 4)), dim3(kB), 0, s, hf_at(kLayers), F, th, u.p);
    if (np > 0)
        hipLaunchKernelGGL(head_t_kernel, dim3(grid_of(np, 4)), dim3(kB), 0, s, hf_at(kLayers), g.pf, g.pg, g.pq, np,
                           th, t.p);
    XFE_CHECK(hipGetLastError());
}

// dW partial product and its reduction into grad + off (K x 64)
template <int K1, bool TWO>
static void dw_into(const float* X1, const float* X2, const float* dZ, int M, float* part, float* dst, hipStream_t s) {
    constexpr int K = K1 + (TWO ? 64 : 0);
    const int nb = std::min(grid_of(M, 32), 512);
    hipLaunchKernelGGL((gemm_dw_kernel<K1, TWO>), dim3(nb), dim3(kB), 0, s, X1, X2, dZ, M, part);
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(grid_of(K * 64, kB)), dim3(kB), 0, s, part, nb, K * 64, dst);
}

static void bias_into(const float* dZ, int M, float* part, float* dst, hipStream_t s) {
    const int nb = std::min(grid_of(M, 4), 512);
    hipLaunchKernelGGL(colsum_kernel, dim3(nb), dim3(kB), 0, s, dZ, M, part);
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(kB), 0, s, part, nb, 64, dst);
}

double Gnn::loss_backward(const Graph& g, const float* th, const Targets& tg, float* grad, hipStream_t s) {
    dhf.ensure(static_cast<size_t>(std::max(F, 1)) * 64);
    dhw.ensure(static_cast<size_t>(std::max(Wm, 1)) * 64);
    dtf.ensure(static_cast<size_t>(std::max(F, 1)) * 64);
    dtw.ensure(static_cast<size_t>(std::max(Wm, 1)) * 64);
    dagg.ensure(static_cast<size_t>(std::max(std::max(F, Wm), 1)) * 64);
    float* gf = dhf.p;  // d hf of the current layer
    float* gw = dhw.p;  // d hw of the current layer
    float* tf = dtf.p;
    float* tw = dtw.p;
    part.ensure(static_cast<size_t>(512) * 128 * 64);
    const int nbl = 512;
    dpart.ensure(static_cast<size_t>(2) * nbl * kPart);
    hdpart.resize(static_cast<size_t>(2) * nbl * kPart);
    XFE_CHECK(hipMemsetAsync(grad, 0, sizeof(float) * kNumParams, s));
    XFE_CHECK(hipMemsetAsync(gw, 0, sizeof(float) * static_cast<size_t>(std::max(Wm, 1)) * 64, s));
    double nm = tg.nmask > 0 ? tg.nmask : 1.0;
    const int gu = std::min(grid_of(F, 4), nbl), gt = std::min(grid_of(np, 4), nbl);
    XFE_CHECK(hipMemsetAsync(dpart.p, 0, sizeof(double) * 2 * nbl * kPart, s));
    if (F > 0)
        hipLaunchKernelGGL(loss_u_kernel, dim3(gu), dim3(kB), 0, s, hf_at(kLayers), u.p, tg.u_tgt, tg.u_mask, F,
                           static_cast<float>(1.0 / nm), th, gf, dpart.p);
    if (np > 0)
        hipLaunchKernelGGL(loss_t_kernel, dim3(gt), dim3(kB), 0, s, hf_at(kLayers), t.p, tg.y_tgt, g.pf, g.pg, g.pq,
                           np, static_cast<float>(1.0 / np), th, gf, dpart.p + static_cast<size_t>(nbl) * kPart);
    XFE_CHECK(hipGetLastError());
    XFE_CHECK(hipMemcpyAsync(hdpart.data(), dpart.p, sizeof(double) * 2 * nbl * kPart, hipMemcpyDeviceToHost, s));
    // layers, top down: gf = d hf_l, gw = d hw_l (from layer l + 1's use of hw_l)
    for (int l = kLayers; l >= 1; --l) {
        if (F > 0) {
            hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_of(static_cast<long long>(F) * 64, kB)), dim3(kB), 0, s, gf,
                               hf_at(l), static_cast<long long>(F) * 64);
            dw_into<64, true>(hf_at(l - 1), af_at(l), gf, F, part.p, grad + off_Wf(l), s);
            bias_into(gf, F, part.p, grad + off_bf(l), s);
            // [d hf_{l-1} | d af_l] = dzf Wf^T
            hipLaunchKernelGGL(gemm_dx_kernel, dim3(grid_of(F, 128)), dim3(kB), 0, s, gf, F, th + off_Wf(l), tf,
                               dagg.p);
            // d hw_l += transpose of the firm mean (over the firms listed for each mover)
            if (Wm > 0)
                hipLaunchKernelGGL(mean_scatter_kernel, dim3(grid_of(Wm, 4)), dim3(kB), 0, s, g.w_ptr, g.w_idx,
                                   g.w_rank, g.f_ptr, dagg.p, Wm, kFirmCap, gw);
            std::swap(gf, tf);  // gf = d hf_{l-1}
        }
        if (Wm > 0) {
            hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_of(static_cast<long long>(Wm) * 64, kB)), dim3(kB), 0, s,
                               gw, hw_at(l), static_cast<long long>(Wm) * 64);
            dw_into<64, true>(hw_at(l - 1), aw_at(l), gw, Wm, part.p, grad + off_Ww(l), s);
            bias_into(gw, Wm, part.p, grad + off_bw(l), s);
            hipLaunchKernelGGL(gemm_dx_kernel, dim3(grid_of(Wm, 128)), dim3(kB), 0, s, gw, Wm, th + off_Ww(l), tw,
                               dagg.p);
            if (F > 0)
                hipLaunchKernelGGL(mean_scatter_kernel, dim3(grid_of(F, 4)), dim3(kB), 0, s, g.f_ptr, g.f_idx,
                                   static_cast<const int*>(nullptr), g.w_ptr, dagg.p, F, INT_MAX, gf);
            std::swap(gw, tw);  // gw = d hw_{l-1}
        }
    }
    if (F > 0) {
        hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_of(static_cast<long long>(F) * 64, kB)), dim3(kB), 0, s, gf,
                           hf_at(0), static_cast<long long>(F) * 64);
        dw_into<8, false>(g.xf, nullptr, gf, F, part.p, grad + kOffWf0, s);
        bias_into(gf, F, part.p, grad + kOffBf0, s);
    }
    if (Wm > 0) {
        hipLaunchKernelGGL(relu_mask_kernel, dim3(grid_of(static_cast<long long>(Wm) * 64, kB)), dim3(kB), 0, s, gw,
                           hw_at(0), static_cast<long long>(Wm) * 64);
        dw_into<8, false>(g.xw, nullptr, gw, Wm, part.p, grad + kOffWw0, s);
        bias_into(gw, Wm, part.p, grad + kOffBw0, s);
    }
    XFE_CHECK(hipGetLastError());
    XFE_CHECK(hipStreamSynchronize(s));
    // head gradients and the loss from the per-block partials
    double lu = 0.0, le = 0.0, gba = 0.0, gbe = 0.0, gws = 0.0;
    std::vector<double> gwa(64, 0.0), gwe(64, 0.0);
    for (int b = 0; b < nbl; ++b) {
        const double* pu = hdpart.data() + static_cast<size_t>(b) * kPart;
        const double* pt = hdpart.data() + static_cast<size_t>(nbl + b) * kPart;
        for (int k = 0; k < 64; ++k) {
            gwa[k] += pu[k];
            gwe[k] += pt[k];
        }
        lu += pu[64];
        gba += pu[65];
        le += pt[64];
        gbe += pt[65];
        gws += pt[66];
    }
    float hg[2 * 64 + 3];
    for (int k = 0; k < 64; ++k) hg[k] = static_cast<float>(gwa[k]);
    hg[64] = static_cast<float>(gba);
    for (int k = 0; k < 64; ++k) hg[65 + k] = static_cast<float>(gwe[k]);
    hg[129] = static_cast<float>(gbe);
    hg[130] = static_cast<float>(gws);
    // wa, ba, we, be, ws are contiguous at the end of the vector
    XFE_CHECK(hipMemcpyAsync(grad + kOffWa, hg, sizeof(hg), hipMemcpyHostToDevice, s));
    XFE_CHECK(hipStreamSynchronize(s));
    last_loss_u = lu / nm;
    last_loss_e = np > 0 ? le / np : 0.0;
    return last_loss_u + last_loss_e;
}

void adam_step(float* th, float* m, float* v, const float* g, int step, float lr, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = static_cast<float>(1.0 / (1.0 - std::pow(0.9, step)));
    const float c2 = static_cast<float>(1.0 / (1.0 - std::pow(0.999, step)));
    hipLaunchKernelGGL(adam_kernel, dim3(grid_of(kNumParams, kB)), dim3(kB), 0, s, th, m, v, g, kNumParams, lr, b1,
                       b2, c1, c2, 1e-8f);
    XFE_CHECK(hipGetLastError());
}

}  // namespace lp
}  // namespace xfe

using namespace xfe;

namespace xfe {
namespace lp {
// w_rank of host edge lists: the position of each mover -> firm edge in the
// firm's list (-1 when the edge is missing there)
std::vector<int> host_w_rank(int F, int Wm, const int* f_ptr, const int* f_idx, const int* w_ptr, const int* w_idx) {
    const int E = Wm > 0 ? w_ptr[Wm] : 0;
    std::vector<int> rank(static_cast<size_t>(std::max(E, 1)), -1);
    for (int f = 0; f < F; ++f)
        for (int q = f_ptr[f]; q < f_ptr[f + 1]; ++q) {
            const int u = f_idx[q];
            for (int e = w_ptr[u]; e < w_ptr[u + 1]; ++e)
                if (w_idx[e] == f && rank[e] < 0) {
                    rank[e] = q - f_ptr[f];
                    break;
                }
        }
    return rank;
}
}  // namespace lp
}  // namespace xfe

extern "C" {

int xfe_lp_param_count(void) { return lp::kNumParams; }

int xfe_lp_gnn_eval(int F, int Wm, const float* xf, const float* xw, const int* f_ptr, const int* f_idx,
                    const int* w_ptr, const int* w_idx, int npairs, const int* pf, const int* pg, const float* pq,
                    const float* u_tgt,
                    const float* u_mask, const float* y_tgt, const float* theta, double* loss, float* grad,
                    float* u_out, float* t_out) {
    clear_error();
    if (F < 0 || Wm < 0 || npairs < 0 || !theta || !loss || (F > 0 && (!xf || !f_ptr || !u_tgt || !u_mask)) ||
        (Wm > 0 && (!xw || !w_ptr)) || (npairs > 0 && (!pf || !pg || !y_tgt))) {
        set_error("xfe_lp_gnn_eval: invalid arguments");
        return 0;
    }
    const int Ef = F > 0 ? f_ptr[F] : 0, Ew = Wm > 0 ? w_ptr[Wm] : 0;
    if (Ef != Ew || (Ef > 0 && (!f_idx || !w_idx))) {
        set_error("xfe_lp_gnn_eval: the two edge lists differ in length");
        return 0;
    }
    for (int e = 0; e < Ef; ++e)
        if (f_idx[e] < 0 || f_idx[e] >= Wm || w_idx[e] < 0 || w_idx[e] >= F) {
            set_error("xfe_lp_gnn_eval: edge endpoint out of range");
            return 0;
        }
    for (int p = 0; p < npairs; ++p)
        if (pf[p] < 0 || pf[p] >= F || pg[p] < 0 || pg[p] >= F) {
            set_error("xfe_lp_gnn_eval: pair endpoint out of range");
            return 0;
        }
    try {
        (void)hipGetLastError();
        hipStream_t s = nullptr;
        XFE_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{s};
        DevBuf<float> d_xf, d_xw, d_ut, d_um, d_y, d_th, d_g, d_pq;
        DevBuf<int> d_fp, d_fi, d_wp, d_wi, d_pf, d_pg, d_rk;
        auto up = [&](auto& buf, const auto* src, size_t n) {
            buf.ensure(n);
            if (n) XFE_CHECK(hipMemcpyAsync(buf.p, src, sizeof(*src) * n, hipMemcpyHostToDevice, s));
        };
        up(d_xf, xf, static_cast<size_t>(F) * lp::kInF);
        up(d_xw, xw, static_cast<size_t>(Wm) * lp::kInW);
        up(d_ut, u_tgt, static_cast<size_t>(F));
        up(d_um, u_mask, static_cast<size_t>(F));
        up(d_y, y_tgt, static_cast<size_t>(npairs));
        up(d_th, theta, static_cast<size_t>(lp::kNumParams));
        d_g.ensure(lp::kNumParams);
        up(d_fp, f_ptr, static_cast<size_t>(F + 1));
        up(d_fi, f_idx, static_cast<size_t>(Ef));
        up(d_wp, w_ptr, static_cast<size_t>(Wm + 1));
        up(d_wi, w_idx, static_cast<size_t>(Ew));
        up(d_pf, pf, static_cast<size_t>(npairs));
        up(d_pg, pg, static_cast<size_t>(npairs));
        if (pq) up(d_pq, pq, static_cast<size_t>(npairs));
        const std::vector<int> rk = lp::host_w_rank(F, Wm, f_ptr, f_idx, w_ptr, w_idx);
        up(d_rk, rk.data(), static_cast<size_t>(Ew));
        lp::Graph g{d_xf.p, d_xw.p, d_fp.p, d_fi.p, d_wp.p, d_wi.p, d_pf.p, d_pg.p, pq ? d_pq.p : nullptr, d_rk.p};
        double nm = 0.0;
        for (int f = 0; f < F; ++f) nm += u_mask[f];
        lp::Targets tg{d_ut.p, d_um.p, d_y.p, nm};
        lp::Gnn net;
        net.ensure(F, Wm, npairs, true);
        net.forward(g, d_th.p, s);
        *loss = net.loss_backward(g, d_th.p, tg, d_g.p, s);
        if (grad) XFE_CHECK(hipMemcpyAsync(grad, d_g.p, sizeof(float) * lp::kNumParams, hipMemcpyDeviceToHost, s));
        if (u_out && F) XFE_CHECK(hipMemcpyAsync(u_out, net.u.p, sizeof(float) * F, hipMemcpyDeviceToHost, s));
        if (t_out && npairs) XFE_CHECK(hipMemcpyAsync(t_out, net.t.p, sizeof(float) * npairs, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        net.release();
        for (auto* b : {&d_xf, &d_xw, &d_ut, &d_um, &d_y, &d_th, &d_g, &d_pq}) b->release();
        for (auto* b : {&d_fp, &d_fi, &d_wp, &d_wi, &d_pf, &d_pg, &d_rk}) b->release();
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

}  // extern "C"
