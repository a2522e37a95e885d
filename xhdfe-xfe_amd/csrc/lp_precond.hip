// The aggregate-block preconditioner of the mixed-precision Schur PCG: firm
// pairs coupled by movers (weights and strengths), the handshake aggregation
// over pair scores, the aggregates' blocks of S and their inverses, and the
// apply kernel of the FP32 inner CG.
//
// Everything is integer work or FP64 sums taken in a fixed order (one thread
// per run of a sorted key list), so the same input gives the same bits: the
// pair list comes from a stable radix sort of the (firm, firm) keys each
// worker's segments emit, the aggregation from 64-bit atomic maxima of keys
// that carry the score and the pair index (no ties), the blocks from a stable
// sort of the firms by aggregate.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>

#include "common.h"
#include "lp_precond.h"

namespace xfe {
namespace lp {

namespace {

constexpr int kB = 256;
constexpr int kCapMax = 16;  // largest aggregate the block inverse handles

int grid_of(long long work, int per_block = kB, int cap = 8192) {
    const long long g = (work + per_block - 1) / per_block;
    return static_cast<int>(std::max(1LL, std::min(g, static_cast<long long>(cap))));
}

int bits_for(unsigned long long v) {
    int b = 1;
    while (b < 64 && (1ULL << b) <= v) ++b;
    return b;
}

// ------------------------------------------------------------------ pairs

__global__ void __launch_bounds__(kB) pair_count_kernel(const int* wseg, int Nw, long long* cnt) {
    for (int w = blockIdx.x * kB + threadIdx.x; w < Nw; w += gridDim.x * kB) {
        const long long k = wseg[w + 1] - wseg[w];
        cnt[w] = k >= 2 ? k * (k - 1) / 2 : 0;
    }
}

// Every pair of segments of one worker: key (min firm) * G1 + max firm, value
// L_i L_j / n_w (twice that for two segments at the same firm: the cross term
// of (L_1 + L_2)^2 that the per-segment diagonal misses).
__global__ void __launch_bounds__(kB) pair_emit_kernel(const int* wseg, const int* r1, const int* len,
                                                       const double* inv_w, const long long* off, int Nw, int G1,
                                                       unsigned long long* key, double* val) {
    for (int w = blockIdx.x * kB + threadIdx.x; w < Nw; w += gridDim.x * kB) {
        const int b = wseg[w], e = wseg[w + 1];
        if (e - b < 2) continue;
        long long o = off[w];
        const double iw = inv_w[w];
        for (int i = b; i < e; ++i) {
            const int fi = r1[i];
            const double li = static_cast<double>(len[i]) * iw;
            for (int j = i + 1; j < e; ++j) {
                const int fj = r1[j];
                const int a = min(fi, fj), c = max(fi, fj);
                key[o] = static_cast<unsigned long long>(a) * static_cast<unsigned long long>(G1) + c;
                val[o] = (a == c ? 2.0 : 1.0) * li * static_cast<double>(len[j]);
                ++o;
            }
        }
    }
}

__global__ void __launch_bounds__(kB) head_kernel(const unsigned long long* key, long long m, int* head) {
    for (long long i = blockIdx.x * static_cast<long long>(kB) + threadIdx.x; i < m;
         i += static_cast<long long>(gridDim.x) * kB)
        head[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

// run[i] is the inclusive count of heads: the first entry of run u (1-based) sits at i
__global__ void __launch_bounds__(kB) run_start_kernel(const int* head, const int* run, long long m, int* pos) {
    for (long long i = blockIdx.x * static_cast<long long>(kB) + threadIdx.x; i < m;
         i += static_cast<long long>(gridDim.x) * kB)
        if (head[i]) pos[run[i] - 1] = static_cast<int>(i);
}

// Sum of each run in index order; diagonal runs correct the diagonal, the
// others are flagged for compaction.
__global__ void __launch_bounds__(kB) run_sum_kernel(const unsigned long long* key, const double* val, const int* pos,
                                                     int U, int G1, double* wsum, double* dex, int* offd) {
    for (int u = blockIdx.x * kB + threadIdx.x; u < U; u += gridDim.x * kB) {
        const int b = pos[u], e = pos[u + 1];
        double s = 0.0;
        for (int i = b; i < e; ++i) s += val[i];
        const unsigned long long k = key[b];
        const int a = static_cast<int>(k / static_cast<unsigned long long>(G1));
        const int c = static_cast<int>(k % static_cast<unsigned long long>(G1));
        wsum[u] = s;
        if (a == c) dex[a] -= s;  // one run per firm: no race
        offd[u] = a != c ? 1 : 0;
    }
}

__global__ void __launch_bounds__(kB) compact_pairs_kernel(const unsigned long long* key, const int* pos,
                                                           const double* wsum, const int* offd, const int* idx, int U,
                                                           int G1, int* pf, int* pg, double* W) {
    for (int u = blockIdx.x * kB + threadIdx.x; u < U; u += gridDim.x * kB) {
        if (!offd[u]) continue;
        const unsigned long long k = key[pos[u]];
        const int o = idx[u];
        pf[o] = static_cast<int>(k / static_cast<unsigned long long>(G1));
        pg[o] = static_cast<int>(k % static_cast<unsigned long long>(G1));
        W[o] = wsum[u];
    }
}

__global__ void __launch_bounds__(kB) strength_kernel(const int* pf, const int* pg, const double* W, const double* dex,
                                                      int U, float* st) {
    for (int p = blockIdx.x * kB + threadIdx.x; p < U; p += gridDim.x * kB) {
        const double a = dex[pf[p]], b = dex[pg[p]];
        st[p] = (a > 0.0 && b > 0.0) ? static_cast<float>(W[p] / std::sqrt(a * b)) : 0.0f;
    }
}

// ------------------------------------------------------------------ aggregation

__global__ void __launch_bounds__(kB) agg_init_kernel(int* agg, int* asize, int G1) {
    for (int f = blockIdx.x * kB + threadIdx.x; f < G1; f += gridDim.x * kB) {
        agg[f] = f;
        asize[f] = 1;
    }
}

// order-preserving map of a float to 32 bits (larger float -> larger value); finite inputs
__device__ __forceinline__ unsigned int ord_float(float x) {
    const unsigned int u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void __launch_bounds__(kB) agg_propose_kernel(const int* pf, const int* pg, const float* score, int U,
                                                         const int* agg, const int* asize, int cap,
                                                         unsigned long long* best) {
    for (int p = blockIdx.x * kB + threadIdx.x; p < U; p += gridDim.x * kB) {
        const int a = agg[pf[p]], b = agg[pg[p]];
        const float sc = score[p];
        if (a == b || asize[a] + asize[b] > cap || !(std::fabs(sc) <= 3.0e38f)) continue;
        const unsigned long long k = (static_cast<unsigned long long>(ord_float(sc)) << 32) |
                                     static_cast<unsigned long long>(0xFFFFFFFFu - static_cast<unsigned int>(p));
        // a plain read first: most pairs of a large firm lose to the best so far,
        // and skipping their atomics keeps one address from serializing the launch
        if (__hip_atomic_load(&best[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k) atomicMax(&best[a], k);
        if (__hip_atomic_load(&best[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k) atomicMax(&best[b], k);
    }
}

// A root whose best pair is also the best pair of the other end merges the
// larger root into the smaller one (each root is in at most one mutual pair).
__global__ void __launch_bounds__(kB) agg_match_kernel(const int* pf, const int* pg, const int* agg, int* asize,
                                                       const unsigned long long* best, int G1, int* newroot) {
    for (int a = blockIdx.x * kB + threadIdx.x; a < G1; a += gridDim.x * kB) {
        const unsigned long long k = best[a];
        if (k == 0ULL || agg[a] != a) continue;
        const int p = static_cast<int>(0xFFFFFFFFu - static_cast<unsigned int>(k & 0xFFFFFFFFULL));
        const int x = agg[pf[p]], y = agg[pg[p]];
        const int other = x == a ? y : x;
        if (other > a && best[other] == k) {
            newroot[other] = a;
            asize[a] += asize[other];
        }
    }
}

__global__ void __launch_bounds__(kB) agg_relabel_kernel(int* agg, const int* newroot, int G1) {
    for (int f = blockIdx.x * kB + threadIdx.x; f < G1; f += gridDim.x * kB) {
        const int r = newroot[agg[f]];
        if (r >= 0) agg[f] = r;
    }
}

// ------------------------------------------------------------------ blocks

__global__ void __launch_bounds__(kB) agg_count_kernel(const int* agg, int G1, int* cnt) {
    for (int f = blockIdx.x * kB + threadIdx.x; f < G1; f += gridDim.x * kB) atomicAdd(&cnt[agg[f]], 1);
}

__global__ void __launch_bounds__(kB) block_sizes_kernel(const int* cnt, int G1, int* bsz, int* bsq) {
    for (int r = blockIdx.x * kB + threadIdx.x; r < G1; r += gridDim.x * kB) {
        const int c = cnt[r] >= 2 ? cnt[r] : 0;
        bsz[r] = c;
        bsq[r] = c * c;
    }
}

__global__ void __launch_bounds__(kB) block_keys_kernel(const int* agg, const int* bsz, int G1, int* key, int* iota) {
    for (int f = blockIdx.x * kB + threadIdx.x; f < G1; f += gridDim.x * kB) {
        const int r = agg[f];
        key[f] = bsz[r] > 0 ? r : G1;
        iota[f] = f;
    }
}

__global__ void __launch_bounds__(kB) block_meta_kernel(const int* sorted, int nb, const int* agg, const int* bsz,
                                                        const int* bstart, const int* boff2, int* firm, int* start,
                                                        int* size, int* row, int* loc) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < nb; i += gridDim.x * kB) {
        const int f = sorted[i];
        const int r = agg[f];
        const int st = bstart[r], s = bsz[r], l = i - st;
        firm[i] = f;
        start[i] = st;
        size[i] = s;
        row[i] = boff2[r] + l * s;
        loc[f] = l;
    }
}

__global__ void __launch_bounds__(kB) block_diag_kernel(const int* firm, const int* start, const int* row, int nb,
                                                        const double* dex, double* bmat) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < nb; i += gridDim.x * kB)
        bmat[row[i] + (i - start[i])] = dex[firm[i]];
}

__global__ void __launch_bounds__(kB) block_off_kernel(const int* pf, const int* pg, const double* W, int U,
                                                       const int* agg, const int* bsz, const int* boff2,
                                                       const int* loc, double* bmat) {
    for (int p = blockIdx.x * kB + threadIdx.x; p < U; p += gridDim.x * kB) {
        const int f = pf[p], g = pg[p];
        const int r = agg[f];
        if (r != agg[g] || bsz[r] == 0) continue;
        const int s = bsz[r], o = boff2[r], lf = loc[f], lg = loc[g];
        bmat[o + lf * s + lg] = -W[p];
        bmat[o + lg * s + lf] = -W[p];
    }
}

// Generalized inverse of each block (symmetric positive semi-definite):
// L D L' with a pivot dropped when it falls to 1e-10 of its diagonal entry
// (an isolated group of firms: its constant vector is in the null space of
// S), then inv = L^-T D^+ L^-1.  One wavefront (the whole 64-thread
// workgroup) per block, the block in LDS, FP64; stored in FP32.
__global__ void __launch_bounds__(64) block_invert_kernel(const int* heads, int nblocks, const int* size,
                                                          const int* row, const double* bmat, float* inv) {
    __shared__ double A[kCapMax][kCapMax + 1];
    __shared__ double Y[kCapMax][kCapMax + 1];
    __shared__ double D[kCapMax], A0[kCapMax];
    const int t = threadIdx.x;
    for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const int i = heads[b];
        const int s = size[i];
        const int o = row[i];
        for (int e = t; e < s * s; e += 64) A[e / s][e % s] = bmat[o + e];
        __syncthreads();
        if (t < s) A0[t] = A[t][t];
        __syncthreads();
        for (int k = 0; k < s; ++k) {
            // A[k][k] holds the pivot after the updates of columns < k
            const double d = A[k][k];
            const bool keep = d > 1e-10 * A0[k] && A0[k] > 0.0;
            __syncthreads();
            if (t == 0) D[k] = keep ? d : 0.0;
            if (t > k && t < s) A[t][k] = keep ? A[t][k] / d : 0.0;  // L[t][k]
            __syncthreads();
            // trailing update: A[m][n] -= L[m][k] D[k] L[n][k], m, n > k
            const int r = s - k - 1;
            for (int e = t; e < r * r; e += 64) {
                const int m = k + 1 + e / r, n = k + 1 + e % r;
                if (n <= m) A[m][n] -= A[m][k] * D[k] * A[n][k];
            }
            __syncthreads();
        }
        // Y = L^-1, one lane per column
        if (t < s) {
            for (int a = 0; a < s; ++a) Y[a][t] = a == t ? 1.0 : 0.0;
            for (int a = t + 1; a < s; ++a) {
                double v = 0.0;
                for (int k = t; k < a; ++k) v -= A[a][k] * Y[k][t];
                Y[a][t] = v;
            }
        }
        __syncthreads();
        for (int e = t; e < s * s; e += 64) {
            const int a = e / s, c = e % s;
            double v = 0.0;
            for (int k = max(a, c); k < s; ++k)
                if (D[k] > 0.0) v += Y[k][a] * Y[k][c] / D[k];
            inv[o + e] = static_cast<float>(v);
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kB) head_flag_kernel(const int* start, int nb, int* flag) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < nb; i += gridDim.x * kB) flag[i] = start[i] == i ? 1 : 0;
}

__global__ void __launch_bounds__(kB) head_compact_kernel(const int* flag, const int* idx, int nb, int* heads) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < nb; i += gridDim.x * kB)
        if (flag[i]) heads[idx[i]] = i;
}

// One thread per (block, column): the block's r values once, then every row
// of its inverse.
__global__ void __launch_bounds__(kB) block_apply_kernel(BlockView b, const float* __restrict__ r,
                                                         float* __restrict__ z, int C) {
    const long long total = static_cast<long long>(b.nblocks) * C;
    for (long long j = blockIdx.x * static_cast<long long>(kB) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kB) {
        const int h = b.heads[static_cast<int>(j / C)], c = static_cast<int>(j % C);
        const int s = b.size[h];
        const float* inv = b.inv + b.row[h];
        float rv[kCapMax];
        int fv[kCapMax];
#pragma unroll
        for (int k = 0; k < kCapMax; ++k) {
            fv[k] = k < s ? b.firm[h + k] : 0;
            rv[k] = k < s ? r[static_cast<size_t>(fv[k]) * C + c] : 0.0f;
        }
        for (int a = 0; a < s; ++a) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < kCapMax; ++k) acc += (k < s ? inv[a * s + k] : 0.0f) * rv[k];
            z[static_cast<size_t>(b.firm[h + a]) * C + c] = acc;
        }
    }
}

template <typename K, typename V>
void sort_pairs(DevBuf<unsigned char>& tmp, const K* kin, K* kout, const V* vin, V* vout, size_t m, int end_bit,
                hipStream_t s) {
    size_t bytes = 0;
    XFE_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, m, 0, end_bit, s));
    tmp.ensure(bytes);
    XFE_CHECK(rocprim::radix_sort_pairs(tmp.p, bytes, kin, kout, vin, vout, m, 0, end_bit, s));
}

template <typename T>
void exclusive_scan(DevBuf<unsigned char>& tmp, const T* in, T* out, size_t m, hipStream_t s) {
    size_t bytes = 0;
    XFE_CHECK(rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), m, rocprim::plus<T>(), s));
    tmp.ensure(bytes);
    XFE_CHECK(rocprim::exclusive_scan(tmp.p, bytes, in, out, T(0), m, rocprim::plus<T>(), s));
}

template <typename T>
void inclusive_scan(DevBuf<unsigned char>& tmp, const T* in, T* out, size_t m, hipStream_t s) {
    size_t bytes = 0;
    XFE_CHECK(rocprim::inclusive_scan(nullptr, bytes, in, out, m, rocprim::plus<T>(), s));
    tmp.ensure(bytes);
    XFE_CHECK(rocprim::inclusive_scan(tmp.p, bytes, in, out, m, rocprim::plus<T>(), s));
}

}  // namespace

hipError_t launch_block_apply(const BlockView& b, const float* r, float* z, int C, hipStream_t s) {
    if (b.nb <= 0 || b.nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(block_apply_kernel, dim3(grid_of(static_cast<long long>(b.nblocks) * C)), dim3(kB), 0, s, b,
                       r, z, C);
    return hipGetLastError();
}

bool AggPrecond::build_pairs(const SegView& sv, hipStream_t s) {
    G1 = sv.G1;
    U = 0;
    nb = 0;
    nblocks = 0;
    if (sv.Nw <= 0 || G1 <= 0) return false;
    if (!host_int) XFE_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_int), 64 * sizeof(long long)));
    long long* hl = reinterpret_cast<long long*>(host_int);
    wcnt.ensure(sv.Nw);
    woff.ensure(sv.Nw);
    const int gw = grid_of(sv.Nw);
    hipLaunchKernelGGL(pair_count_kernel, dim3(gw), dim3(kB), 0, s, sv.wseg, sv.Nw, wcnt.p);
    exclusive_scan<long long>(tmp, wcnt.p, woff.p, static_cast<size_t>(sv.Nw), s);
    XFE_CHECK(hipMemcpyAsync(hl, woff.p + (sv.Nw - 1), sizeof(long long), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipMemcpyAsync(hl + 1, wcnt.p + (sv.Nw - 1), sizeof(long long), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    const long long E = hl[0] + hl[1];
    emitted = E;
    if (E <= 0 || E >= INT_MAX) return false;
    key_a.ensure(E); key_b.ensure(E); val_a.ensure(E); val_b.ensure(E);
    head.ensure(E); run.ensure(E);
    hipLaunchKernelGGL(pair_emit_kernel, dim3(gw), dim3(kB), 0, s, sv.wseg, sv.r1, sv.len, sv.inv_w, woff.p, sv.Nw,
                       G1, key_a.p, val_a.p);
    XFE_CHECK(hipGetLastError());
    const unsigned long long maxkey = static_cast<unsigned long long>(G1) * static_cast<unsigned long long>(G1);
    sort_pairs<unsigned long long, double>(tmp, key_a.p, key_b.p, val_a.p, val_b.p, static_cast<size_t>(E),
                                           bits_for(maxkey), s);
    const int ge = grid_of(E);
    hipLaunchKernelGGL(head_kernel, dim3(ge), dim3(kB), 0, s, key_b.p, E, head.p);
    inclusive_scan<int>(tmp, head.p, run.p, static_cast<size_t>(E), s);
    XFE_CHECK(hipMemcpyAsync(host_int, run.p + (E - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    const int Uall = host_int[0];
    pos.ensure(static_cast<size_t>(Uall) + 1);
    hipLaunchKernelGGL(run_start_kernel, dim3(ge), dim3(kB), 0, s, head.p, run.p, E, pos.p);
    const int Ei = static_cast<int>(E);
    XFE_CHECK(hipMemcpyAsync(pos.p + Uall, &Ei, sizeof(int), hipMemcpyHostToDevice, s));
    dex.ensure(G1);
    XFE_CHECK(hipMemcpyAsync(dex.p, sv.diag, sizeof(double) * G1, hipMemcpyDeviceToDevice, s));
    // run sums into val_a (free after the sort), off-diagonal flags into head
    hipLaunchKernelGGL(run_sum_kernel, dim3(grid_of(Uall)), dim3(kB), 0, s, key_b.p, val_b.p, pos.p, Uall, G1,
                       val_a.p, dex.p, head.p);
    exclusive_scan<int>(tmp, head.p, run.p, static_cast<size_t>(Uall), s);
    XFE_CHECK(hipMemcpyAsync(host_int, run.p + (Uall - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipMemcpyAsync(host_int + 1, head.p + (Uall - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    U = host_int[0] + host_int[1];
    if (U <= 0) return false;
    pf.ensure(U); pg.ensure(U); W.ensure(U); st.ensure(U);
    hipLaunchKernelGGL(compact_pairs_kernel, dim3(grid_of(Uall)), dim3(kB), 0, s, key_b.p, pos.p, val_a.p, head.p,
                       run.p, Uall, G1, pf.p, pg.p, W.p);
    hipLaunchKernelGGL(strength_kernel, dim3(grid_of(U)), dim3(kB), 0, s, pf.p, pg.p, W.p, dex.p, U, st.p);
    XFE_CHECK(hipGetLastError());
    return true;
}

bool AggPrecond::set_pairs(int G1_, const std::vector<int>& f, const std::vector<int>& g,
                           const std::vector<double>& w, const std::vector<double>& d, hipStream_t s) {
    G1 = G1_;
    U = static_cast<int>(f.size());
    nb = 0;
    nblocks = 0;
    if (U <= 0 || G1 <= 0) return false;
    if (!host_int) XFE_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_int), 64 * sizeof(long long)));
    pf.ensure(U); pg.ensure(U); W.ensure(U); st.ensure(U); dex.ensure(G1);
    XFE_CHECK(hipMemcpyAsync(pf.p, f.data(), sizeof(int) * U, hipMemcpyHostToDevice, s));
    XFE_CHECK(hipMemcpyAsync(pg.p, g.data(), sizeof(int) * U, hipMemcpyHostToDevice, s));
    XFE_CHECK(hipMemcpyAsync(W.p, w.data(), sizeof(double) * U, hipMemcpyHostToDevice, s));
    XFE_CHECK(hipMemcpyAsync(dex.p, d.data(), sizeof(double) * G1, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(strength_kernel, dim3(grid_of(U)), dim3(kB), 0, s, pf.p, pg.p, W.p, dex.p, U, st.p);
    XFE_CHECK(hipGetLastError());
    XFE_CHECK(hipStreamSynchronize(s));  // the host vectors belong to the caller
    return true;
}

void AggPrecond::aggregate(const float* score, int cap_, int rounds_, hipStream_t s) {
    cap = std::max(2, std::min(cap_, kCapMax));
    rounds = std::max(0, rounds_);
    agg.ensure(G1); asize.ensure(G1); newroot.ensure(G1); best.ensure(G1);
    const int gf = grid_of(G1), gp = grid_of(U);
    hipLaunchKernelGGL(agg_init_kernel, dim3(gf), dim3(kB), 0, s, agg.p, asize.p, G1);
    for (int r = 0; r < rounds; ++r) {
        XFE_CHECK(hipMemsetAsync(best.p, 0, sizeof(unsigned long long) * G1, s));
        XFE_CHECK(hipMemsetAsync(newroot.p, 0xFF, sizeof(int) * G1, s));
        hipLaunchKernelGGL(agg_propose_kernel, dim3(gp), dim3(kB), 0, s, pf.p, pg.p, score, U, agg.p, asize.p, cap,
                           best.p);
        hipLaunchKernelGGL(agg_match_kernel, dim3(gf), dim3(kB), 0, s, pf.p, pg.p, agg.p, asize.p, best.p, G1,
                           newroot.p);
        hipLaunchKernelGGL(agg_relabel_kernel, dim3(gf), dim3(kB), 0, s, agg.p, newroot.p, G1);
    }
    XFE_CHECK(hipGetLastError());
}

void AggPrecond::build_blocks(hipStream_t s) {
    cnt.ensure(G1); bsz.ensure(G1); bsq.ensure(G1); bstart.ensure(G1); boff2.ensure(G1);
    skey.ensure(G1); sval.ensure(G1); iota.ensure(G1); loc.ensure(G1);
    const int gf = grid_of(G1);
    XFE_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int) * G1, s));
    hipLaunchKernelGGL(agg_count_kernel, dim3(gf), dim3(kB), 0, s, agg.p, G1, cnt.p);
    hipLaunchKernelGGL(block_sizes_kernel, dim3(gf), dim3(kB), 0, s, cnt.p, G1, bsz.p, bsq.p);
    exclusive_scan<int>(tmp, bsz.p, bstart.p, static_cast<size_t>(G1), s);
    exclusive_scan<int>(tmp, bsq.p, boff2.p, static_cast<size_t>(G1), s);
    int h[4];
    XFE_CHECK(hipMemcpyAsync(host_int, bstart.p + (G1 - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipMemcpyAsync(host_int + 1, bsz.p + (G1 - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipMemcpyAsync(host_int + 2, boff2.p + (G1 - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipMemcpyAsync(host_int + 3, bsq.p + (G1 - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < 4; ++k) h[k] = host_int[k];
    nb = h[0] + h[1];
    const int tot2 = h[2] + h[3];
    nblocks = 0;
    if (nb <= 0) return;
    hipLaunchKernelGGL(block_keys_kernel, dim3(gf), dim3(kB), 0, s, agg.p, bsz.p, G1, skey.p, iota.p);
    sort_pairs<int, int>(tmp, skey.p, cnt.p, iota.p, sval.p, static_cast<size_t>(G1),
                         bits_for(static_cast<unsigned long long>(G1)), s);  // sorted keys unused
    b_firm.ensure(nb); b_start.ensure(nb); b_size.ensure(nb); b_row.ensure(nb);
    hipLaunchKernelGGL(block_meta_kernel, dim3(grid_of(nb)), dim3(kB), 0, s, sval.p, nb, agg.p, bsz.p, bstart.p,
                       boff2.p, b_firm.p, b_start.p, b_size.p, b_row.p, loc.p);
    bmat.ensure(tot2);
    binv.ensure(tot2);
    XFE_CHECK(hipMemsetAsync(bmat.p, 0, sizeof(double) * tot2, s));
    hipLaunchKernelGGL(block_diag_kernel, dim3(grid_of(nb)), dim3(kB), 0, s, b_firm.p, b_start.p, b_row.p, nb, dex.p,
                       bmat.p);
    hipLaunchKernelGGL(block_off_kernel, dim3(grid_of(U)), dim3(kB), 0, s, pf.p, pg.p, W.p, U, agg.p, bsz.p, boff2.p,
                       loc.p, bmat.p);
    // block heads: the first blocked firm of each block
    b_hflag.ensure(nb); b_hidx.ensure(nb);
    hipLaunchKernelGGL(head_flag_kernel, dim3(grid_of(nb)), dim3(kB), 0, s, b_start.p, nb, b_hflag.p);
    exclusive_scan<int>(tmp, b_hflag.p, b_hidx.p, static_cast<size_t>(nb), s);
    XFE_CHECK(hipMemcpyAsync(host_int, b_hidx.p + (nb - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipMemcpyAsync(host_int + 1, b_hflag.p + (nb - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    nblocks = host_int[0] + host_int[1];
    b_head.ensure(nblocks);
    hipLaunchKernelGGL(head_compact_kernel, dim3(grid_of(nb)), dim3(kB), 0, s, b_hflag.p, b_hidx.p, nb, b_head.p);
    hipLaunchKernelGGL(block_invert_kernel, dim3(std::max(1, std::min(nblocks, 16384))), dim3(64), 0, s, b_head.p,
                       nblocks, b_size.p, b_row.p, bmat.p, binv.p);
    XFE_CHECK(hipGetLastError());
}

void AggPrecond::release() {
    for (auto* b : {&wcnt, &woff}) b->release();
    key_a.release(); key_b.release(); val_a.release(); val_b.release();
    for (auto* b : {&head, &run, &pos, &pf, &pg, &agg, &asize, &newroot, &cnt, &bsz, &bsq, &bstart, &boff2, &skey,
                    &sval, &iota, &loc, &b_firm, &b_start, &b_size, &b_row, &b_head, &b_hflag, &b_hidx})
        b->release();
    W.release(); dex.release(); st.release(); best.release(); bmat.release(); binv.release(); tmp.release();
    if (host_int) (void)hipHostFree(host_int);
    host_int = nullptr;
}

}  // namespace lp
}  // namespace xfe

using namespace xfe;

extern "C" {

int xfe_lp_blocks_apply(int F, int U, const int* pf, const int* pg, const double* W, const double* dex,
                        const float* score, int cap, int rounds, int C, const float* r, int* agg_out,
                        float* z_out) {
    clear_error();
    if (F <= 0 || U < 0 || C <= 0 || !dex || !r || !agg_out || !z_out || (U > 0 && (!pf || !pg || !W || !score))) {
        set_error("xfe_lp_blocks_apply: invalid arguments");
        return 0;
    }
    for (int p = 0; p < U; ++p)
        if (pf[p] < 0 || pf[p] >= F || pg[p] < 0 || pg[p] >= F || pf[p] >= pg[p] ||
            (p > 0 && (pf[p] < pf[p - 1] || (pf[p] == pf[p - 1] && pg[p] <= pg[p - 1])))) {
            set_error("xfe_lp_blocks_apply: pairs must be f < g, sorted and unique, ids in [0, F)");
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
        lp::AggPrecond ag;
        ag.G1 = F;
        ag.U = U;
        if (!ag.host_int) XFE_CHECK(hipHostMalloc(reinterpret_cast<void**>(&ag.host_int), 64 * sizeof(long long)));
        ag.pf.ensure(U); ag.pg.ensure(U); ag.W.ensure(U); ag.dex.ensure(F);
        DevBuf<float> d_score, d_r, d_z;
        d_score.ensure(U); d_r.ensure(static_cast<size_t>(F) * C); d_z.ensure(static_cast<size_t>(F) * C);
        if (U > 0) {
            XFE_CHECK(hipMemcpyAsync(ag.pf.p, pf, sizeof(int) * U, hipMemcpyHostToDevice, s));
            XFE_CHECK(hipMemcpyAsync(ag.pg.p, pg, sizeof(int) * U, hipMemcpyHostToDevice, s));
            XFE_CHECK(hipMemcpyAsync(ag.W.p, W, sizeof(double) * U, hipMemcpyHostToDevice, s));
            XFE_CHECK(hipMemcpyAsync(d_score.p, score, sizeof(float) * U, hipMemcpyHostToDevice, s));
        }
        XFE_CHECK(hipMemcpyAsync(ag.dex.p, dex, sizeof(double) * F, hipMemcpyHostToDevice, s));
        XFE_CHECK(hipMemcpyAsync(d_r.p, r, sizeof(float) * F * C, hipMemcpyHostToDevice, s));
        XFE_CHECK(hipMemsetAsync(d_z.p, 0, sizeof(float) * F * C, s));
        ag.aggregate(d_score.p, cap, rounds, s);
        ag.build_blocks(s);
        XFE_CHECK(lp::launch_block_apply(ag.view(), d_r.p, d_z.p, C, s));
        XFE_CHECK(hipMemcpyAsync(agg_out, ag.agg.p, sizeof(int) * F, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipMemcpyAsync(z_out, d_z.p, sizeof(float) * F * C, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        ag.release();
        d_score.release(); d_r.release(); d_z.release();
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

}  // extern "C"
