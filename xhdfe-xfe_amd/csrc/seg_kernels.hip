// Segment form of the Schur-complement operator of the mixed-precision solver.
//
// Rows are level-sorted by the eliminated dimension p (workers).  A segment is
// a maximal run of rows with the same level of p, the same level of r1 (firms)
// and, when a second dimension r2 (years) is kept, consecutive levels of r2
// (r2 of the next row = r2 of this row + 1).  On an AKM panel a segment is a
// job spell: about 7.5M segments for 54M rows.
//
// For coefficients x = (x1, x2) and u_i = x1[r1_i] + x2[r2_i], the operator
// S x = D_R' M_p D_R x is, with P the prefix sums of x2 over the levels of r2,
//     U_s  = L_s x1[f_s] + P[a_s + L_s] - P[a_s]        (sum of u over segment s)
//     m_w  = (1 / n_w) sum_{s in w} U_s                 (mean of u over worker w)
//     (S x)_1[f] = sum_{s: f_s = f} L_s (x1[f] - m_w(s)) + P[a_s + L_s] - P[a_s]
//     (S x)_2[y] = n_y x2[y] + sum_{s covering y} (x1[f_s] - m_w(s))
// The last sum is a difference array over the segment's range [a_s, a_s + L_s)
// followed by a prefix sum.  One application is two passes over the segments
// (worker-major for m, firm-major for (S x)_1 and the year differences) with
// no per-row work.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "absorb_kernels.h"
#include "wave_ops.h"

namespace xfe {

namespace {

template <typename T>
__device__ __forceinline__ T seg_run_sum(T v, int end) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const T o = __shfl_down(v, off, kWave);
        if (lane + off <= end) v += o;
    }
    return v;
}

// Segmented sum over a wave of sorted keys, without atomics.  A run that lies
// strictly inside the wave (head lane > 0, end lane < 63, or ending at the
// last entry of the list) holds every entry of its key: its sum is stored in
// out.  The run at lane 0 and the run reaching lane 63 may continue in the
// neighbouring waves: they go to the boundary list at entries 2w and 2w + 1
// (a wave that is one run, or ends inside the list's tail, stores a zero
// second entry under its first key), which stays sorted and is reduced the
// same way (seg_bnd_kernel).  With `final` (the list is one wave) every run is
// complete and stored in out.
template <typename T, int CC>
__device__ __forceinline__ void seg_wave_reduce(int key, bool valid, const T (&v)[CC], long long wv,
                                                bool lane63_valid, bool final, T* out, int C, int c0,
                                                int* bkey, T* bval) {
    const int lane = lane_id();
    const int prev = __shfl_up(key, 1, kWave);
    const bool is_head = lane == 0 || prev != key;
    const uint64_t heads = __ballot(is_head);
    const uint64_t above = lane == 63 ? 0ull : (heads & ~((2ull << lane) - 1ull));
    const int end = above ? (__builtin_ctzll(above) - 1) : 63;
    const int h63 = 63 - __builtin_clzll(heads);
    const bool head = is_head && valid;
    const bool inner = final || (lane > 0 && end < kWave - 1);
#pragma unroll
    for (int c = 0; c < CC; ++c) {
        const T s = seg_run_sum<T>(v[c], end);
        if (head) {
            if (inner) out[static_cast<size_t>(key) * C + c0 + c] = s;
            else if (lane == 0) bval[(2 * wv) * CC + c] = s;
            else bval[(2 * wv + 1) * CC + c] = s;
        }
        if (!final && lane == 0 && valid && (h63 == 0 || !lane63_valid)) bval[(2 * wv + 1) * CC + c] = T(0);
    }
    if (!final && head && !inner) bkey[2 * wv + (lane == 0 ? 0 : 1)] = key;
    if (!final && lane == 0 && valid && (h63 == 0 || !lane63_valid)) bkey[2 * wv + 1] = key;
}

// bytes of an LDS table of e values of T, rounded up to 16
template <typename T>
__host__ __device__ constexpr size_t seg_tab_bytes(int e) {
    return (static_cast<size_t>(e) * sizeof(T) + 15) / 16 * 16;
}

// P[k][c] = sum_{y < k} x2[y][c0 + c] for k = 0..G2, built in LDS: the block
// stages x2 with coalesced loads, then one thread per column scans in LDS
// (G2 <= kSegMaxRange).  Ends with a barrier.
template <typename T, int CC>
__device__ __forceinline__ void build_prefix(const T* x2, int G2, int C, int c0, T* P) {
    for (int k = threadIdx.x; k < G2 * CC; k += kBlock) {
        const int y = k / CC;
        const int c = k - y * CC;
        P[k + CC] = x2[static_cast<size_t>(y) * C + c0 + c];
    }
    __syncthreads();
    if (threadIdx.x < CC) {
        const int c = threadIdx.x;
        T acc = T(0);
        P[c] = acc;
        for (int y = 1; y <= G2; ++y) {
            acc += P[y * CC + c];
            P[y * CC + c] = acc;
        }
    }
    __syncthreads();
}

}  // namespace

// ------------------------------------------------------------------ set-up

// 1 where a row starts a segment.
__global__ void __launch_bounds__(kBlock)
seg_head_kernel(const int* gp, const int* g1, const int* g2, int n, int* flags) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        int h = 1;
        if (i > 0) {
            h = gp[i] != gp[i - 1] || g1[i] != g1[i - 1] || (g2 != nullptr && g2[i] != g2[i - 1] + 1);
        }
        flags[i] = h;
    }
}

// Segment records from the head rows (inc = inclusive scan of the flags).
__global__ void __launch_bounds__(kBlock)
seg_fill_kernel(const int* flags, const int* inc, const int* gp, const int* g1, const int* g2, int n,
                int* seg_start, int* seg_w, int* seg_r1, int* seg_a) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        if (!flags[i]) continue;
        const int s = inc[i] - 1;
        seg_start[s] = i;
        seg_w[s] = gp[i];
        seg_r1[s] = g1[i];
        if (g2 != nullptr) seg_a[s] = g2[i];
    }
}

// Segment lengths from their start rows.
__global__ void __launch_bounds__(kBlock) seg_len_kernel(const int* seg_start, int M, int n, int* seg_len) {
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < M; s += gridDim.x * kBlock) {
        const int e = s + 1 < M ? seg_start[s + 1] : n;
        seg_len[s] = e - seg_start[s];
    }
}

// off[w] = first segment whose (ascending) key is >= w, for w = 0..Nw
// (off[Nw] = M): segment offsets per worker, empty workers included.
__global__ void __launch_bounds__(kBlock) seg_offsets_kernel(const int* key, int M, int Nw, int* off) {
    for (int s = blockIdx.x * kBlock + threadIdx.x; s < M; s += gridDim.x * kBlock) {
        const int w = key[s];
        const int prev = s > 0 ? key[s - 1] : -1;
        if (w != prev) {
            for (int v = prev + 1; v <= w; ++v) off[v] = s;
        }
        if (s == M - 1) {
            for (int v = w + 1; v <= Nw; ++v) off[v] = M;
        }
    }
}

// Primary firm of each worker (the firm of its first segment; G1 when the
// worker has no rows, so that it sorts last).
__global__ void __launch_bounds__(kBlock)
seg_primary_kernel(const int* wseg, const int* seg_r1, int Nw, int G1, int* pf) {
    for (int w = blockIdx.x * kBlock + threadIdx.x; w < Nw; w += gridDim.x * kBlock)
        pf[w] = wseg[w] < wseg[w + 1] ? seg_r1[wseg[w]] : G1;
}

// flag += number of levels f < G with off[f] == off[f + 1] (no segment)
__global__ void __launch_bounds__(kBlock) seg_empty_kernel(const int* off, int G, int* flag) {
    for (int f = blockIdx.x * kBlock + threadIdx.x; f < G; f += gridDim.x * kBlock)
        if (off[f] == off[f + 1]) atomicAdd(flag, 1);
}

// movers (two or more segments) among the worker ranks
__global__ void __launch_bounds__(kBlock) seg_mover_flag_kernel(const int* wseg, int Nw, int* flag) {
    for (int w = blockIdx.x * kBlock + threadIdx.x; w < Nw; w += gridDim.x * kBlock)
        flag[w] = wseg[w + 1] - wseg[w] >= 2 ? 1 : 0;
}

// mover list, and the firm-major records relabelled: a mover's segments carry
// its mover index, a stayer's segment a negative length
__global__ void __launch_bounds__(kBlock) seg_mover_list_kernel(const int* flag, const int* idx, int Nw, int* mlist) {
    for (int w = blockIdx.x * kBlock + threadIdx.x; w < Nw; w += gridDim.x * kBlock)
        if (flag[w]) mlist[idx[w]] = w;
}

__global__ void __launch_bounds__(kBlock) seg_mover_relabel_kernel(const int* flag, const int* idx, int M, int* f_w,
                                                                   int* f_len) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < M; j += gridDim.x * kBlock) {
        const int w = f_w[j];
        if (flag[w]) f_w[j] = idx[w];
        else f_len[j] = -f_len[j];
    }
}

// rank[perm[t]] = t
__global__ void __launch_bounds__(kBlock) seg_rank_kernel(const int* perm, int m, int* rank) {
    for (int t = blockIdx.x * kBlock + threadIdx.x; t < m; t += gridDim.x * kBlock) rank[perm[t]] = t;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) seg_gather_kernel(const T* src, const int* perm, T* dst, int m) {
    for (int t = blockIdx.x * kBlock + threadIdx.x; t < m; t += gridDim.x * kBlock) dst[t] = src[perm[t]];
}

// ------------------------------------------------------------------ apply

// Pass 1 (one thread per mover): m_w = (1 / n_w) sum_{s in w} U_s.  A stayer
// (one segment s, n_w = L) needs no mean: x1[f] - m_w = -(P[a + L] - P[a]) / L,
// so its segment adds L (x1[f] - m_w) + (P[a + L] - P[a]) = 0 to the firm and
// -(1 / L) sum_{y' in s} x2[y'] to each year y in s: over all stayers the
// fixed matrix K[y][y'] = -sum_{stayer segments s containing y and y'} 1 / L_s
// (seg_year_kernel).
template <typename T, int CC, bool HAS_Y>
__global__ void __launch_bounds__(kBlock) seg_worker_kernel(SegArgs<T> a) {
    extern __shared__ __align__(16) unsigned char seg_dyn[];
    T* P = reinterpret_cast<T*>(seg_dyn);
    if constexpr (HAS_Y) build_prefix<T, CC>(a.x2, a.G2, a.C, a.c0, P);
    for (int mi = blockIdx.x * kBlock + threadIdx.x; mi < a.Nm; mi += gridDim.x * kBlock) {
        const int w = a.mlist[mi];
        const int s0 = a.wseg[w], s1 = a.wseg[w + 1];
        T acc[CC];
#pragma unroll
        for (int c = 0; c < CC; ++c) acc[c] = T(0);
        for (int s = s0; s < s1; ++s) {
            const int f = a.seg_r1[s];
            const int L = a.seg_len[s];
            const T* xf = a.x1 + static_cast<size_t>(f) * a.C + a.c0;
            const T tl = static_cast<T>(L);
#pragma unroll
            for (int c = 0; c < CC; ++c) acc[c] += tl * xf[c];
            if constexpr (HAS_Y) {
                const int y0 = a.seg_a[s];
#pragma unroll
                for (int c = 0; c < CC; ++c) acc[c] += P[(y0 + L) * CC + c] - P[y0 * CC + c];
            }
        }
        const T iw = a.inv_w[w];
        T* m = a.wmean + static_cast<size_t>(mi) * CC;
#pragma unroll
        for (int c = 0; c < CC; ++c) m[c] = acc[c] * iw;
    }
}

// Pass 2 (one thread per mover segment, segments in firm order): (S x)_1 by a
// segmented sum over equal firms, and the year differences into an FP64 LDS
// histogram with one copy per wave (FP32 LDS atomics on the few year slots
// stall the LDS pipe), summed into hpart[block].  DIAG: the Jacobi diagonal
// instead (T = double, CC = 1): L - L^2 / n_w per segment, and -1 / n_w over
// its years.  Stayers' segments are not in the list: they add nothing to the
// firms, and to the years a fixed linear map of x2 (seg_year_kernel's K).
template <typename T, int CC, bool HAS_Y, bool DIAG>
__global__ void __launch_bounds__(kBlock) seg_firm_kernel(SegArgs<T> a) {
    extern __shared__ __align__(16) unsigned char seg_dyn[];
    T* P = reinterpret_cast<T*>(seg_dyn);          // (G2 + 1) x CC prefix sums
    const int E = (a.G2 + 1) * CC;
    double* H = reinterpret_cast<double*>(seg_dyn + seg_tab_bytes<T>(E));  // kBlock / kWave copies of E
    const int lane = lane_id();
    double* Hw = H + (threadIdx.x / kWave) * E;
    if constexpr (HAS_Y) {
        if constexpr (!DIAG) build_prefix<T, CC>(a.x2, a.G2, a.C, a.c0, P);
        for (int k = threadIdx.x; k < (kBlock / kWave) * E; k += kBlock) H[k] = 0.0;
        __syncthreads();
    }
    for (int base = blockIdx.x * kBlock; base < a.M; base += gridDim.x * kBlock) {
        const int j = base + threadIdx.x;
        const bool valid = j < a.M;
        const int key = valid ? a.f_key[j] : -1 - lane;
        T v[CC];
#pragma unroll
        for (int c = 0; c < CC; ++c) v[c] = T(0);
        if (valid) {
            const int w = a.f_w[j];
            const int L = a.f_len[j];
            const T tl = static_cast<T>(L);
            if constexpr (DIAG) {
                const T iw = a.inv_wm[w];
                v[0] = tl - tl * tl * iw;
                if constexpr (HAS_Y) {
                    const int y0 = a.f_a[j];
                    atomicAdd(&Hw[y0], -static_cast<double>(iw));
                    atomicAdd(&Hw[y0 + L], static_cast<double>(iw));
                }
            } else {
                const T* xf = a.x1 + static_cast<size_t>(key) * a.C + a.c0;
                const T* m = a.wmean + static_cast<size_t>(w) * CC;
                T d[CC];
#pragma unroll
                for (int c = 0; c < CC; ++c) {
                    d[c] = xf[c] - m[c];
                    v[c] = tl * d[c];
                }
                if constexpr (HAS_Y) {
                    const int y0 = a.f_a[j];
#pragma unroll
                    for (int c = 0; c < CC; ++c) {
                        v[c] += P[(y0 + L) * CC + c] - P[y0 * CC + c];
                        atomicAdd(&Hw[y0 * CC + c], static_cast<double>(d[c]));
                        atomicAdd(&Hw[(y0 + L) * CC + c], -static_cast<double>(d[c]));
                    }
                }
            }
        }
        const long long wbase = base + (threadIdx.x & ~(kWave - 1));
        if (wbase < a.M) {
            seg_wave_reduce<T, CC>(key, valid, v, wbase / kWave, wbase + kWave - 1 < a.M, false, a.out1, a.C,
                                   a.c0, a.bkey, a.bval);
        }
    }
    if constexpr (HAS_Y) {
        __syncthreads();
        double* hp = a.hpart + static_cast<size_t>(blockIdx.x) * E;
        for (int k = threadIdx.x; k < E; k += kBlock) {
            double t = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / kWave; ++w) t += H[w * E + k];
            hp[k] = t;
        }
    }
}

// One level of the boundary reduction: the sorted list (kin, vin) of m
// entries -> complete runs into out, boundary runs into (kout, vout).
template <typename T, int CC>
__global__ void __launch_bounds__(kBlock)
seg_bnd_kernel(const int* kin, const T* vin, int m, T* out, int C, int c0, int* kout, T* vout, int final) {
    const int lane = lane_id();
    for (int base = blockIdx.x * kBlock; base < m; base += gridDim.x * kBlock) {
        const int j = base + threadIdx.x;
        const bool valid = j < m;
        const int key = valid ? kin[j] : -1 - lane;
        T v[CC];
#pragma unroll
        for (int c = 0; c < CC; ++c) v[c] = valid ? vin[static_cast<size_t>(j) * CC + c] : T(0);
        const long long wbase = base + (threadIdx.x & ~(kWave - 1));
        if (wbase < m) {
            seg_wave_reduce<T, CC>(key, valid, v, wbase / kWave, wbase + kWave - 1 < m, final != 0, out, C, c0,
                                   kout, vout);
        }
    }
}

// The last levels of the boundary reduction in one block of kBndTail threads:
// levels are reduced in place of two ping-pong lists until one wave holds the
// list, which is then stored completely.  (kin, vin) is the first list.
template <typename T, int CC>
__global__ void __launch_bounds__(kBndTail)
seg_bnd_tail_kernel(int* k0, T* v0, int* k1, T* v1, int m, T* out, int C, int c0) {
    const int lane = lane_id();
    int* kin = k0;
    T* vin = v0;
    int* kout = k1;
    T* vout = v1;
    for (;;) {
        const bool final = m <= kWave;
        for (int base = 0; base < m; base += kBndTail) {
            const int j = base + threadIdx.x;
            const bool valid = j < m;
            const int key = valid ? kin[j] : -1 - lane;
            T v[CC];
#pragma unroll
            for (int c = 0; c < CC; ++c) v[c] = valid ? vin[static_cast<size_t>(j) * CC + c] : T(0);
            const long long wbase = base + (threadIdx.x & ~(kWave - 1));
            if (wbase < m) {
                seg_wave_reduce<T, CC>(key, valid, v, wbase / kWave, wbase + kWave - 1 < m, final, out, C, c0,
                                       kout, vout);
            }
        }
        if (final) break;
        __syncthreads();  // the next level reads what other waves wrote
        m = static_cast<int>(2 * ((static_cast<long long>(m) + kWave - 1) / kWave));
        int* tk = kin; kin = kout; kout = tk;
        T* tv = vin; vin = vout; vout = tv;
    }
}

// tot[e] = sum over blocks of hpart[b][e], one block per entry e.
__global__ void __launch_bounds__(kBlock) seg_hsum_kernel(const double* hpart, int nblk, int E, double* tot) {
    __shared__ double red[kBlock / kWave];
    const int e = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kBlock) acc += hpart[static_cast<size_t>(b) * E + e];
    acc = wave_sum(acc);
    if (lane_id() == 0) red[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < kBlock / kWave; ++k) s += red[k];
        tot[e] = s;
    }
}

// (S x)_2[y] = n_y x2[y] + prefix_y(tot); with x2 == nullptr the result is
// n_y + prefix_y(tot) (the Jacobi diagonal of the year block).  The block
// stages tot, n_y and x2 in LDS, one thread per column scans.
// K (G2 x G2, may be null): the stayers' year block, added as K x2 (or its
// diagonal when x2 == nullptr).
template <typename T>
__global__ void __launch_bounds__(kBlock)
seg_year_kernel(const double* tot, int G2, int cc, const double* counts, const T* x2, int C, int c0,
                const double* K, T* out2) {
    __shared__ double st[(kSegMaxRange + 1) * 8];
    __shared__ double sx[kSegMaxRange * 8];
    __shared__ double sn[kSegMaxRange];
    for (int k = threadIdx.x; k < G2 * cc; k += kBlock) {
        st[k] = tot[k];
        const int y = k / cc;
        sx[k] = x2 ? static_cast<double>(x2[static_cast<size_t>(y) * C + c0 + (k - y * cc)]) : 1.0;
    }
    for (int y = threadIdx.x; y < G2; y += kBlock) sn[y] = counts[y];
    __syncthreads();
    for (int k = threadIdx.x; k < G2 * cc; k += kBlock) {
        // the prefix of column c up to level y, by one thread per entry (G2 <= kSegMaxRange)
        const int y = k / cc;
        const int c = k - y * cc;
        double run = 0.0;
        for (int t = 0; t <= y; ++t) run += st[t * cc + c];
        double ks = 0.0;
        if (K) {
            if (x2) {
                for (int t = 0; t < G2; ++t) ks += K[static_cast<size_t>(y) * G2 + t] * sx[t * cc + c];
            } else {
                ks = K[static_cast<size_t>(y) * G2 + y];
            }
        }
        out2[static_cast<size_t>(y) * C + c0 + c] = static_cast<T>(sn[y] * sx[k] + run + ks);
    }
}

// Stayer segments (negative length in the firm-major records) counted by
// (first year, length): H[a * (G2 + 1) + L]; LDS copy per block when it fits.
__global__ void __launch_bounds__(kBlock) seg_stayer_hist_kernel(const int* f_len, const int* f_a, int M, int G2,
                                                                 int* H) {
    extern __shared__ __align__(16) int hl[];
    const int E = G2 * (G2 + 1);
    const bool lds = E <= 16384;
    if (lds) {
        for (int k = threadIdx.x; k < E; k += kBlock) hl[k] = 0;
        __syncthreads();
    }
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < M; j += gridDim.x * kBlock) {
        const int Ls = f_len[j];
        if (Ls >= 0) continue;
        const int idx = f_a[j] * (G2 + 1) - Ls;
        if (lds) atomicAdd(&hl[idx], 1);
        else atomicAdd(&H[idx], 1);
    }
    if (lds) {
        __syncthreads();
        for (int k = threadIdx.x; k < E; k += kBlock)
            if (hl[k]) atomicAdd(&H[k], hl[k]);
    }
}

// mover segments of the firm-major records, compacted in order
__global__ void __launch_bounds__(kBlock) seg_mover_flag_seg_kernel(const int* f_len, int M, int* flag) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < M; j += gridDim.x * kBlock) flag[j] = f_len[j] > 0 ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) seg_compact_kernel(const int* flag, const int* idx, int M, const int* k_in,
                                                             const int* w_in, const int* l_in, const int* a_in,
                                                             int* k_out, int* w_out, int* l_out, int* a_out) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < M; j += gridDim.x * kBlock) {
        if (!flag[j]) continue;
        const int o = idx[j];
        k_out[o] = k_in[j];
        w_out[o] = w_in[j];
        l_out[o] = l_in[j];
        if (a_in) a_out[o] = a_in[j];
    }
}

// ------------------------------------------------------------- launchers

hipError_t launch_seg_heads(const int* gp, const int* g1, const int* g2, int n, int* flags, int grid,
                            hipStream_t s) {
    hipLaunchKernelGGL(seg_head_kernel, dim3(grid), dim3(kBlock), 0, s, gp, g1, g2, n, flags);
    return hipGetLastError();
}

hipError_t launch_seg_fill(const int* flags, const int* inc, const int* gp, const int* g1, const int* g2,
                           int n, int* seg_start, int* seg_w, int* seg_r1, int* seg_a, int grid,
                           hipStream_t s) {
    hipLaunchKernelGGL(seg_fill_kernel, dim3(grid), dim3(kBlock), 0, s, flags, inc, gp, g1, g2, n,
                       seg_start, seg_w, seg_r1, seg_a);
    return hipGetLastError();
}

hipError_t launch_seg_len(const int* seg_start, int M, int n, int* seg_len, int grid, hipStream_t s) {
    hipLaunchKernelGGL(seg_len_kernel, dim3(grid), dim3(kBlock), 0, s, seg_start, M, n, seg_len);
    return hipGetLastError();
}

hipError_t launch_seg_offsets(const int* key, int M, int Nw, int* off, int grid, hipStream_t s) {
    hipLaunchKernelGGL(seg_offsets_kernel, dim3(grid), dim3(kBlock), 0, s, key, M, Nw, off);
    return hipGetLastError();
}

hipError_t launch_seg_primary(const int* wseg, const int* seg_r1, int Nw, int G1, int* pf, int grid,
                              hipStream_t s) {
    hipLaunchKernelGGL(seg_primary_kernel, dim3(grid), dim3(kBlock), 0, s, wseg, seg_r1, Nw, G1, pf);
    return hipGetLastError();
}

hipError_t launch_seg_movers(const int* wseg, int Nw, int* flag, int grid, hipStream_t s) {
    hipLaunchKernelGGL(seg_mover_flag_kernel, dim3(grid), dim3(kBlock), 0, s, wseg, Nw, flag);
    return hipGetLastError();
}

hipError_t launch_seg_mover_relabel(const int* flag, const int* idx, int Nw, int M, int* mlist, int* f_w, int* f_len,
                                    int grid, hipStream_t s) {
    hipLaunchKernelGGL(seg_mover_list_kernel, dim3(grid), dim3(kBlock), 0, s, flag, idx, Nw, mlist);
    hipLaunchKernelGGL(seg_mover_relabel_kernel, dim3(grid), dim3(kBlock), 0, s, flag, idx, M, f_w, f_len);
    return hipGetLastError();
}

hipError_t launch_seg_empty(const int* off, int G, int* flag, int grid, hipStream_t s) {
    hipLaunchKernelGGL(seg_empty_kernel, dim3(grid), dim3(kBlock), 0, s, off, G, flag);
    return hipGetLastError();
}

hipError_t launch_seg_rank(const int* perm, int m, int* rank, int grid, hipStream_t s) {
    hipLaunchKernelGGL(seg_rank_kernel, dim3(grid), dim3(kBlock), 0, s, perm, m, rank);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_seg_gather(const T* src, const int* perm, T* dst, int m, int grid, hipStream_t s) {
    hipLaunchKernelGGL((seg_gather_kernel<T>), dim3(grid), dim3(kBlock), 0, s, src, perm, dst, m);
    return hipGetLastError();
}
template hipError_t launch_seg_gather<float>(const float*, const int*, float*, int, int, hipStream_t);
template hipError_t launch_seg_gather<double>(const double*, const int*, double*, int, int, hipStream_t);
template hipError_t launch_seg_gather<int>(const int*, const int*, int*, int, int, hipStream_t);

template <typename T, int CC>
static hipError_t seg_worker_cc(const SegArgs<T>& a, bool has_y, int grid, hipStream_t s) {
    const size_t tab = has_y ? seg_tab_bytes<T>((a.G2 + 1) * CC) : 0;
    if (has_y) hipLaunchKernelGGL((seg_worker_kernel<T, CC, true>), dim3(grid), dim3(kBlock), tab, s, a);
    else hipLaunchKernelGGL((seg_worker_kernel<T, CC, false>), dim3(grid), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

template <typename T, int CC>
static hipError_t seg_firm_cc(const SegArgs<T>& a, bool has_y, bool diag, int grid, hipStream_t s) {
    const size_t tab = has_y ? seg_tab_bytes<T>((a.G2 + 1) * CC) : 0;
    const size_t ftab = has_y ? tab + static_cast<size_t>(kBlock / kWave) * (a.G2 + 1) * CC * sizeof(double) : 0;
    if (diag) {
        if constexpr (CC == 1) {
            if (has_y) hipLaunchKernelGGL((seg_firm_kernel<T, 1, true, true>), dim3(grid), dim3(kBlock), ftab, s, a);
            else hipLaunchKernelGGL((seg_firm_kernel<T, 1, false, true>), dim3(grid), dim3(kBlock), 0, s, a);
            return hipGetLastError();
        }
        return hipErrorInvalidValue;
    }
    if (has_y) hipLaunchKernelGGL((seg_firm_kernel<T, CC, true, false>), dim3(grid), dim3(kBlock), ftab, s, a);
    else hipLaunchKernelGGL((seg_firm_kernel<T, CC, false, false>), dim3(grid), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

#define XFE_SEG_SWITCH(CALL)                                \
    switch (cc) {                                           \
        case 1: return CALL(1);                             \
        case 2: return CALL(2);                             \
        case 3: return CALL(3);                             \
        case 4: return CALL(4);                             \
        case 5: return CALL(5);                             \
        case 6: return CALL(6);                             \
        case 7: return CALL(7);                             \
        case 8: return CALL(8);                             \
        default: return hipErrorInvalidValue;               \
    }

template <typename T>
hipError_t launch_seg_worker(const SegArgs<T>& a, int cc, bool has_y, int grid, hipStream_t s) {
#define XFE_W(K) seg_worker_cc<T, K>(a, has_y, grid, s)
    XFE_SEG_SWITCH(XFE_W)
#undef XFE_W
}

template <typename T>
hipError_t launch_seg_firm(const SegArgs<T>& a, int cc, bool has_y, bool diag, int grid, hipStream_t s) {
#define XFE_F(K) seg_firm_cc<T, K>(a, has_y, diag, grid, s)
    XFE_SEG_SWITCH(XFE_F)
#undef XFE_F
}
#undef XFE_SEG_SWITCH
template hipError_t launch_seg_worker<float>(const SegArgs<float>&, int, bool, int, hipStream_t);
template hipError_t launch_seg_worker<double>(const SegArgs<double>&, int, bool, int, hipStream_t);
template hipError_t launch_seg_firm<float>(const SegArgs<float>&, int, bool, bool, int, hipStream_t);
template hipError_t launch_seg_firm<double>(const SegArgs<double>&, int, bool, bool, int, hipStream_t);

template <typename T, int CC>
static void seg_bnd_cc(const int* kin, const T* vin, int m, T* out, int C, int c0, int* kout, T* vout, int final,
                       int grid, hipStream_t s) {
    hipLaunchKernelGGL((seg_bnd_kernel<T, CC>), dim3(grid), dim3(kBlock), 0, s, kin, vin, m, out, C, c0, kout,
                       vout, final);
}

template <typename T>
hipError_t launch_seg_bnd(const int* kin, const T* vin, int m, T* out, int C, int c0, int cc, int* kout, T* vout,
                          int final, int grid, hipStream_t s) {
    switch (cc) {
        case 1: seg_bnd_cc<T, 1>(kin, vin, m, out, C, c0, kout, vout, final, grid, s); break;
        case 2: seg_bnd_cc<T, 2>(kin, vin, m, out, C, c0, kout, vout, final, grid, s); break;
        case 3: seg_bnd_cc<T, 3>(kin, vin, m, out, C, c0, kout, vout, final, grid, s); break;
        case 4: seg_bnd_cc<T, 4>(kin, vin, m, out, C, c0, kout, vout, final, grid, s); break;
        case 5: seg_bnd_cc<T, 5>(kin, vin, m, out, C, c0, kout, vout, final, grid, s); break;
        case 6: seg_bnd_cc<T, 6>(kin, vin, m, out, C, c0, kout, vout, final, grid, s); break;
        case 7: seg_bnd_cc<T, 7>(kin, vin, m, out, C, c0, kout, vout, final, grid, s); break;
        case 8: seg_bnd_cc<T, 8>(kin, vin, m, out, C, c0, kout, vout, final, grid, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_seg_bnd<float>(const int*, const float*, int, float*, int, int, int, int*, float*, int,
                                          int, hipStream_t);
template hipError_t launch_seg_bnd<double>(const int*, const double*, int, double*, int, int, int, int*, double*,
                                           int, int, hipStream_t);

template <typename T>
hipError_t launch_seg_bnd_tail(int* k0, T* v0, int* k1, T* v1, int m, T* out, int C, int c0, int cc,
                               hipStream_t s) {
    switch (cc) {
#define XFE_TAIL(K) case K: hipLaunchKernelGGL((seg_bnd_tail_kernel<T, K>), dim3(1), dim3(kBndTail), 0, s, k0, v0, k1, v1, m, out, C, c0); break;
        XFE_TAIL(1) XFE_TAIL(2) XFE_TAIL(3) XFE_TAIL(4) XFE_TAIL(5) XFE_TAIL(6) XFE_TAIL(7) XFE_TAIL(8)
#undef XFE_TAIL
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
template hipError_t launch_seg_bnd_tail<float>(int*, float*, int*, float*, int, float*, int, int, int, hipStream_t);
template hipError_t launch_seg_bnd_tail<double>(int*, double*, int*, double*, int, double*, int, int, int,
                                                hipStream_t);

template <typename T>
hipError_t launch_seg_year(const double* hpart, int nblk, int G2, int cc, double* tot, const double* counts,
                           const T* x2, int C, int c0, const double* K, T* out2, hipStream_t s) {
    const int E = (G2 + 1) * cc;
    hipLaunchKernelGGL(seg_hsum_kernel, dim3(E), dim3(kBlock), 0, s, hpart, nblk, E, tot);
    hipLaunchKernelGGL((seg_year_kernel<T>), dim3(1), dim3(kBlock), 0, s, tot, G2, cc, counts, x2, C, c0, K,
                       out2);
    return hipGetLastError();
}
template hipError_t launch_seg_year<float>(const double*, int, int, int, double*, const double*, const float*,
                                           int, int, const double*, float*, hipStream_t);
template hipError_t launch_seg_year<double>(const double*, int, int, int, double*, const double*,
                                            const double*, int, int, const double*, double*, hipStream_t);

hipError_t launch_seg_stayer_hist(const int* f_len, const int* f_a, int M, int G2, int* H, int grid, hipStream_t s) {
    const int E = G2 * (G2 + 1);
    const size_t shm = E <= 16384 ? static_cast<size_t>(E) * sizeof(int) : 0;
    hipLaunchKernelGGL(seg_stayer_hist_kernel, dim3(grid), dim3(kBlock), shm, s, f_len, f_a, M, G2, H);
    return hipGetLastError();
}

hipError_t launch_seg_mover_segments(const int* f_len, int M, int* flag, int grid, hipStream_t s) {
    hipLaunchKernelGGL(seg_mover_flag_seg_kernel, dim3(grid), dim3(kBlock), 0, s, f_len, M, flag);
    return hipGetLastError();
}

hipError_t launch_seg_compact(const int* flag, const int* idx, int M, const int* k_in, const int* w_in,
                              const int* l_in, const int* a_in, int* k_out, int* w_out, int* l_out, int* a_out,
                              int grid, hipStream_t s) {
    hipLaunchKernelGGL(seg_compact_kernel, dim3(grid), dim3(kBlock), 0, s, flag, idx, M, k_in, w_in, l_in, a_in,
                       k_out, w_out, l_out, a_out);
    return hipGetLastError();
}

}  // namespace xfe
