// Wide batched solver of the grounded firm Laplacian S = Df - B' Dw^{-1} B
// (see akm_wide.h): up to 256 right-hand sides per pass, FP32 Jacobi-PCG with
// FP64 residual replacement, S in its firm-pair form.
//
// Layout: vectors are level-major J x L (a firm's L lanes contiguous, L = 64 V
// with V = 1, 2 or 4 lanes per thread).  One wavefront takes one piece of 64
// firm-major pair entries (firm, neighbour, weight) and each thread holds V
// consecutive lanes, so every gather of a neighbour's lanes is one contiguous
// 256 V-byte (FP32) read by the wave; the entries are loaded one per thread
// and broadcast by shuffles.  Runs of one firm inside a piece are stored, the
// first and last run of each piece go to a boundary list reduced level by
// level (deterministic, no atomics).
//
// The per-lane CG scalars (rz, pq, alpha, beta, the FP32 and FP64 residual
// norms, active / converged flags, iteration counts) live on the device; the
// partial dot products are per block and lane and are summed in block order,
// so a solve is deterministic.  The host polls every kPoll iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "absorb_kernels.h"
#include "akm_wide.h"

namespace xfe {
namespace {

constexpr int kWPiece = 64;     // matches per wave piece
constexpr int kWBatch = 8;      // gathers issued together
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kPoll = 8;        // iterations between host polls
constexpr double kDelta2 = 1e-4;  // FP32 residual drop (in r.r) that triggers an FP64 replacement
// the restarted mode lets each FP32 run reduce |r| 1000-fold: far enough that
// the restart costs little, short of where FP32's rounding of the operator and
// of the block inverses makes the recurrence drift from the true residual
constexpr double kDelta2Restart = 1e-6;
// a recurrence that has dropped further than this (in r.r) since the last
// replacement has left FP32's accuracy (its residual no longer follows the
// true one), so the replacement restarts the direction as well: p = z
constexpr double kRestart2 = 1e-10;
// a lane is also due for a replacement after this many iterations without one
// (a recurrence that stagnates above the kDelta2 drop would otherwise drift)
constexpr int kMaxRun = 128;
// in the restarted mode a run that has not reached its kDelta2Restart drop
// after this many iterations is stagnating in FP32: it is replaced and
// restarted (and stops under the stall rule when that does not help)
constexpr int kMaxRunRestart = 512;
constexpr int kMaxDotGrid = 1024;
constexpr int kWideBlockCap = 16;    // firms per aggregate block
constexpr int kWideBlockRounds = 6;  // handshake rounds

// per-lane scalar slots (each kWideMaxLanes doubles), then two flags
enum {
    kSRz = 0, kSPq, kSAlpha, kSBeta, kSRr, kSRref, kSActive, kSIters, kSOk, kSBn, kSR64, kSRestart, kSStall, kSWant,
    kSSince, kSForced, kSSlots
};
constexpr int kSFlags = kSSlots * kWideMaxLanes;  // [0] replacement wanted, [1] active lanes
constexpr int kScTotal = kSFlags + 8;

template <typename T, int V>
struct alignas(sizeof(T) * V) LaneVec {
    T v[V];
};

template <typename T, int V>
__device__ __forceinline__ LaneVec<T, V> ldv(const T* p) {
    return *reinterpret_cast<const LaneVec<T, V>*>(p);
}

template <typename T, int V>
__device__ __forceinline__ void stv(T* p, const LaneVec<T, V>& a) {
    *reinterpret_cast<LaneVec<T, V>*>(p) = a;
}

template <typename T, int V>
__device__ __forceinline__ void zero(LaneVec<T, V>& a) {
#pragma unroll
    for (int k = 0; k < V; ++k) a.v[k] = T(0);
}

// Closes the run of `key` with total acc: interior runs are stored in out,
// the first run of a piece goes to boundary entry 2q (unless final).
template <typename T, int V>
__device__ __forceinline__ void close_run(bool& first, bool final, int key, int q, int lane, size_t lo, int L,
                                          const LaneVec<T, V>& acc, T* out, int* kout, T* vout) {
    if (first && !final) {
        stv(vout + static_cast<size_t>(2 * q) * L + lo, acc);
        if (lane == 0) kout[2 * q] = key;
    } else {
        stv(out + static_cast<size_t>(key) * L + lo, acc);
    }
    first = false;
}

// End of a piece: the last run (and a one-run piece's (key, 0) second entry).
template <typename T, int V>
__device__ __forceinline__ void close_piece(bool first, bool final, int key, int q, int lane, size_t lo, int L,
                                            const LaneVec<T, V>& acc, T* out, int* kout, T* vout) {
    if (final) {
        stv(out + static_cast<size_t>(key) * L + lo, acc);
        return;
    }
    T* e0 = vout + static_cast<size_t>(2 * q) * L + lo;
    T* e1 = vout + static_cast<size_t>(2 * q + 1) * L + lo;
    if (first) {
        stv(e0, acc);
        LaneVec<T, V> z;
        zero(z);
        stv(e1, z);
    } else {
        stv(e1, acc);
    }
    if (lane == 0) {
        if (first) kout[2 * q] = key;
        kout[2 * q + 1] = key;
    }
}

// Pair pass over fixed pieces of kWPiece firm-major entries (firm key,
// neighbour, weight): sum of W (p[f] - p[g]) per firm.
// With MULT, entry e belongs to worker fwk[e] and lane l weighs it by
// mult[fwk[e] * mstride + l / mcols] (the worker's multiplicity in the
// resample that lane l belongs to).
template <typename T, int V, bool MULT>
__global__ void __launch_bounds__(kBlock)
wide_pair_kernel(const int* __restrict__ fkey, const int* __restrict__ fw, const T* __restrict__ fc,
                 const T* __restrict__ p, int m, int L, T* __restrict__ out, int* __restrict__ kout,
                 T* __restrict__ vout, int final, const int* __restrict__ fwk, const int* __restrict__ mult,
                 int mstride, int mcols) {
    const int lane = threadIdx.x & (kWave - 1);
    const size_t lo = static_cast<size_t>(lane) * V;
    const int pieces = (m + kWPiece - 1) / kWPiece;
    for (int q = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; q < pieces; q += gridDim.x * kWavesPerBlock) {
        const int j0 = q * kWPiece;
        const int n = min(m - j0, kWPiece);
        const int kA = lane < n ? fkey[j0 + lane] : 0;
        const int wA = lane < n ? fw[j0 + lane] : 0;
        const T xA = lane < n ? fc[j0 + lane] : T(0);
        const int mA = MULT && lane < n ? fwk[j0 + lane] : 0;
        int key = __shfl(kA, 0, kWave);
        LaneVec<T, V> acc;
        zero(acc);
        bool first = true;
        for (int e0 = 0; e0 < n; e0 += kWBatch) {
            LaneVec<T, V> v[kWBatch];
#pragma unroll
            for (int u = 0; u < kWBatch; ++u) {
                const int k = __shfl(kA, (e0 + u) & (kWave - 1), kWave);
                const int w = __shfl(wA, (e0 + u) & (kWave - 1), kWave);
                const LaneVec<T, V> a = ldv<T, V>(p + static_cast<size_t>(k) * L + lo);
                const LaneVec<T, V> b = ldv<T, V>(p + static_cast<size_t>(w) * L + lo);
#pragma unroll
                for (int k2 = 0; k2 < V; ++k2) v[u].v[k2] = a.v[k2] - b.v[k2];
            }
#pragma unroll
            for (int u = 0; u < kWBatch; ++u) {
                const int k = __shfl(kA, (e0 + u) & (kWave - 1), kWave);
                const T x = __shfl(xA, (e0 + u) & (kWave - 1), kWave);
                if (e0 + u < n) {
                    if (k != key) {
                        close_run(first, final != 0, key, q, lane, lo, L, acc, out, kout, vout);
                        key = k;
                        zero(acc);
                    }
                    if constexpr (MULT) {
                        const int wk = __shfl(mA, (e0 + u) & (kWave - 1), kWave);
                        const int* mrow = mult + static_cast<size_t>(wk) * mstride;
#pragma unroll
                        for (int k2 = 0; k2 < V; ++k2)
                            acc.v[k2] += x * static_cast<T>(mrow[(static_cast<int>(lo) + k2) / mcols]) * v[u].v[k2];
                    } else {
#pragma unroll
                        for (int k2 = 0; k2 < V; ++k2) acc.v[k2] += x * v[u].v[k2];
                    }
                }
            }
        }
        if (n > 0) close_piece(first, final != 0, key, q, lane, lo, L, acc, out, kout, vout);
    }
}

// One level of the boundary-list reduction (entries sorted by key).
template <typename T, int V>
__global__ void __launch_bounds__(kBlock)
wide_list_kernel(const int* __restrict__ key_in, const T* __restrict__ vin, int m, int L, T* __restrict__ out,
                 int* __restrict__ kout, T* __restrict__ vout, int final) {
    const int lane = threadIdx.x & (kWave - 1);
    const size_t lo = static_cast<size_t>(lane) * V;
    const int pieces = (m + kWPiece - 1) / kWPiece;
    for (int q = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; q < pieces; q += gridDim.x * kWavesPerBlock) {
        const int j0 = q * kWPiece;
        const int n = min(m - j0, kWPiece);
        const int kA = lane < n ? key_in[j0 + lane] : 0;
        int key = __shfl(kA, 0, kWave);
        LaneVec<T, V> acc;
        zero(acc);
        bool first = true;
        for (int e0 = 0; e0 < n; e0 += kWBatch) {
            LaneVec<T, V> v[kWBatch];
#pragma unroll
            for (int u = 0; u < kWBatch; ++u) {
                const int j = e0 + u < n ? j0 + e0 + u : j0;
                v[u] = ldv<T, V>(vin + static_cast<size_t>(j) * L + lo);
            }
#pragma unroll
            for (int u = 0; u < kWBatch; ++u) {
                const int k = __shfl(kA, (e0 + u) & (kWave - 1), kWave);
                if (e0 + u < n) {
                    if (k != key) {
                        close_run(first, final != 0, key, q, lane, lo, L, acc, out, kout, vout);
                        key = k;
                        zero(acc);
                    }
#pragma unroll
                    for (int k2 = 0; k2 < V; ++k2) acc.v[k2] += v[u].v[k2];
                }
            }
        }
        if (n > 0) close_piece(first, final != 0, key, q, lane, lo, L, acc, out, kout, vout);
    }
}

// out = (pair pass, zero for firms without pairs) + corr p; ground pinned
template <typename T>
__global__ void __launch_bounds__(kBlock)
wide_finish_kernel(T* out, const T* p, const T* corr, const int* has, int J, int L, int ground) {
    const long long total = static_cast<long long>(J) * L;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int f = static_cast<int>(i / L);
        T v = has[f] ? out[i] : T(0);
        v += corr[f] * p[i];
        out[i] = f == ground ? T(0) : v;
    }
}

// Per block and lane partial sums of a[i] b[i] (rows strided over the grid;
// 256 / L rows per block step, combined in LDS in a fixed order).
template <typename T>
__global__ void __launch_bounds__(kBlock)
wide_dot_kernel(const T* __restrict__ a, const T* __restrict__ b, int J, int L, double* __restrict__ part) {
    __shared__ double red[kBlock];
    const int l = threadIdx.x % L, r0 = threadIdx.x / L, R = kBlock / L;
    double acc = 0.0;
    for (long long j = static_cast<long long>(blockIdx.x) * R + r0; j < J; j += static_cast<long long>(gridDim.x) * R) {
        const size_t i = static_cast<size_t>(j) * L + l;
        acc += static_cast<double>(a[i]) * static_cast<double>(b[i]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (r0 == 0) {
        double s = 0.0;
        for (int k = 0; k < R; ++k) s += red[k * L + l];
        part[static_cast<size_t>(blockIdx.x) * L + l] = s;
    }
}

// x += alpha p, r -= alpha q, z = Minv r; partials of r.z and r.r
__global__ void __launch_bounds__(kBlock)
wide_step_kernel(float* __restrict__ x, float* __restrict__ r, float* __restrict__ z, const float* __restrict__ p,
                 const float* __restrict__ q, const float* __restrict__ minv, const double* __restrict__ sc, int J,
                 int L, double* __restrict__ part, const float* __restrict__ minv_lane) {
    __shared__ double red[2][kBlock];
    const int l = threadIdx.x % L, r0 = threadIdx.x / L, R = kBlock / L;
    const float a = static_cast<float>(sc[kSAlpha * kWideMaxLanes + l]);
    double rz = 0.0, rr = 0.0;
    for (long long j = static_cast<long long>(blockIdx.x) * R + r0; j < J; j += static_cast<long long>(gridDim.x) * R) {
        const size_t i = static_cast<size_t>(j) * L + l;
        const float ri = r[i] - a * q[i];
        const float zi = (minv_lane ? minv_lane[i] : minv[j]) * ri;
        x[i] += a * p[i];
        r[i] = ri;
        z[i] = zi;
        rz += static_cast<double>(ri) * zi;
        rr += static_cast<double>(ri) * ri;
    }
    red[0][threadIdx.x] = rz;
    red[1][threadIdx.x] = rr;
    __syncthreads();
    if (r0 == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int k = 0; k < R; ++k) {
            s0 += red[0][k * L + l];
            s1 += red[1][k * L + l];
        }
        part[static_cast<size_t>(blockIdx.x) * L + l] = s0;
        part[static_cast<size_t>(gridDim.x + blockIdx.x) * L + l] = s1;
    }
}

// p = z + beta p
__global__ void __launch_bounds__(kBlock)
wide_dir_kernel(float* __restrict__ p, const float* __restrict__ z, const double* __restrict__ sc, int J, int L) {
    const long long total = static_cast<long long>(J) * L;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int l = static_cast<int>(i % L);
        p[i] = z[i] + static_cast<float>(sc[kSBeta * kWideMaxLanes + l]) * p[i];
    }
}

constexpr int kScalarThreads = 1024;  // threads of the per-lane scalar kernels

// Total over the nblk partials of the thread's lane (valid where threadIdx.x <
// L): kScalarThreads / L threads per lane sum strided slices with four
// accumulators, combined in LDS in slice order (deterministic).  Every thread
// of the block must call it.
__device__ __forceinline__ double lane_total(const double* __restrict__ part, int nblk, int L, int which,
                                             double* red) {
    const int l = threadIdx.x % L, k = threadIdx.x / L, K = kScalarThreads / L;
    const double* base = part + static_cast<size_t>(which) * nblk * L + l;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int b = k;
    for (; b + 3 * K < nblk; b += 4 * K) {
        a0 += base[static_cast<size_t>(b) * L];
        a1 += base[static_cast<size_t>(b + K) * L];
        a2 += base[static_cast<size_t>(b + 2 * K) * L];
        a3 += base[static_cast<size_t>(b + 3 * K) * L];
    }
    for (; b < nblk; b += K) a0 += base[static_cast<size_t>(b) * L];
    red[threadIdx.x] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    double t = 0.0;
    if (k == 0)
        for (int j = 0; j < K; ++j) t += red[j * L + l];
    __syncthreads();
    return t;
}

__device__ __forceinline__ double& S(double* sc, int slot, int l) { return sc[slot * kWideMaxLanes + l]; }

// Block-wide OR of want and count of active lanes into the flags (one block of kBlock threads).
__device__ __forceinline__ void store_flags(double* sc, bool want, bool active) {
    __shared__ int f_want, f_act;
    if (threadIdx.x == 0) {
        f_want = 0;
        f_act = 0;
    }
    __syncthreads();
    if (want) atomicOr(&f_want, 1);
    if (active) atomicAdd(&f_act, 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        sc[kSFlags] = f_want;
        sc[kSFlags + 1] = f_act;
    }
}

__global__ void __launch_bounds__(kBlock) wide_init_kernel(double* sc, int nb) {
    const int l = threadIdx.x;
    for (int s = 0; s < kSSlots; ++s) S(sc, s, l) = 0.0;
    S(sc, kSActive, l) = l < nb ? 1.0 : 0.0;
    S(sc, kSWant, l) = 1.0;  // the first FP64 check covers every lane
    if (l < 8) sc[kSFlags + l] = 0.0;
}

// bn = |b| per lane; b scaled to unit norm (zero lanes stay zero)
__global__ void __launch_bounds__(kScalarThreads) wide_norm_kernel(const double* part, int nblk, int L, double* sc) {
    __shared__ double red[kScalarThreads];
    const double n2 = lane_total(part, nblk, L, 0, red);
    const int l = threadIdx.x;
    if (l < L) S(sc, kSBn, l) = std::sqrt(n2);
}

__global__ void __launch_bounds__(kBlock) wide_scale_kernel(double* b, const double* sc, int J, int L) {
    const long long total = static_cast<long long>(J) * L;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const double bn = sc[kSBn * kWideMaxLanes + static_cast<int>(i % L)];
        b[i] = bn > 0.0 ? b[i] / bn : 0.0;
    }
}

// alpha = rz / pq per active lane; a lane whose pq is not in (0, 1e300) stops, unconverged
__global__ void __launch_bounds__(kScalarThreads) wide_alpha_kernel(const double* part, int nblk, int L, double* sc) {
    __shared__ double red[kScalarThreads];
    const double pq = lane_total(part, nblk, L, 0, red);
    const int l = threadIdx.x;
    if (l >= L) return;
    double a = 0.0;
    if (S(sc, kSActive, l) != 0.0) {
        if (pq > 0.0 && pq < 1e300) {
            a = S(sc, kSRz, l) / pq;
        } else {
            S(sc, kSActive, l) = 0.0;
            S(sc, kSOk, l) = 0.0;
        }
    }
    S(sc, kSPq, l) = pq;
    S(sc, kSAlpha, l) = a;
}

// beta = rz_new / rz; an FP64 replacement is wanted once r.r has dropped by
// kDelta2 since the last one, fallen below tol^2, or at max_iter
__global__ void __launch_bounds__(kScalarThreads)
wide_beta_kernel(const double* part, int nblk, int L, double* sc, double tol2, int max_iter, int max_run,
                 double delta2) {
    __shared__ double red[kScalarThreads];
    const double rzn = lane_total(part, nblk, L, 0, red);
    const double rr = lane_total(part, nblk, L, 1, red);
    const int l = threadIdx.x;
    bool want = false, act = false;
    if (l < L) {
        double beta = 0.0;
        if (S(sc, kSActive, l) != 0.0) {
            act = true;
            const double rz = S(sc, kSRz, l);
            beta = rz > 0.0 ? rzn / rz : 0.0;
            S(sc, kSRz, l) = rzn;
            S(sc, kSRr, l) = rr;
            S(sc, kSIters, l) += 1.0;
            S(sc, kSSince, l) += 1.0;
            want = rr <= delta2 * S(sc, kSRref, l) || rr <= tol2 || S(sc, kSIters, l) >= max_iter ||
                   S(sc, kSSince, l) >= max_run;
        }
        S(sc, kSBeta, l) = beta;
        S(sc, kSWant, l) = want ? 1.0 : 0.0;
        S(sc, kSForced, l) = 0.0;
    }
    store_flags(sc, want, act);
}

// FP64 residual check after a replacement: converged lanes stop (ok), lanes
// at max_iter stop unconverged, the others restart their reference norm
// A lane whose FP64 residual fails to halve over kStallUpdates replacements in
// a row has reached what the FP64 residual can resolve: it stops, converged
// when r.r <= accept2 (the caller's acceptance; accept2 = tol2 keeps the
// strict rule).
constexpr int kStallUpdates = 4;

__global__ void __launch_bounds__(kScalarThreads)
wide_check_kernel(const double* part, int nblk, int L, double* sc, double tol2, int max_iter, double accept2,
                  int restart_all) {
    __shared__ double red[kScalarThreads];
    const double r2 = lane_total(part, nblk, L, 0, red);
    const int l = threadIdx.x;
    bool act = false;
    if (l < L) {
        S(sc, kSRestart, l) = 0.0;
        if (S(sc, kSActive, l) != 0.0 && S(sc, kSWant, l) == 0.0) {
            act = true;  // a lane not due for a replacement keeps its recurrence
        } else if (S(sc, kSActive, l) != 0.0) {
            S(sc, kSR64, l) = r2;
            S(sc, kSSince, l) = 0.0;
            // the direction is kept only while the recurrence follows the true
            // residual: restart when it outran FP32 (kRestart2) or when the FP64
            // residual is more than 10x (in norm) what the recurrence claimed
            // (a lane replaced with the others before its own drop neither
            // restarts on the gap nor counts toward a stall)
            const double rr = S(sc, kSRr, l);
            const bool own = S(sc, kSForced, l) == 0.0;
            if (restart_all || rr < kRestart2 * S(sc, kSRref, l) || (own && r2 > 100.0 * rr))
                S(sc, kSRestart, l) = 1.0;
            const bool first = S(sc, kSRref, l) == 0.0;
            if (own) S(sc, kSStall, l) = (!first && r2 > 0.25 * S(sc, kSRref, l)) ? S(sc, kSStall, l) + 1.0 : 0.0;
            if (r2 <= tol2) {
                S(sc, kSActive, l) = 0.0;
                S(sc, kSOk, l) = 1.0;
            } else if (S(sc, kSStall, l) >= kStallUpdates) {
                S(sc, kSActive, l) = 0.0;
                S(sc, kSOk, l) = r2 <= accept2 ? 1.0 : 0.0;
            } else if (S(sc, kSIters, l) >= max_iter || !(r2 < 1e300)) {
                S(sc, kSActive, l) = 0.0;
                S(sc, kSOk, l) = 0.0;
            } else {
                S(sc, kSRref, l) = r2;
                act = true;
            }
        }
    }
    store_flags(sc, false, act);
}

__global__ void __launch_bounds__(kScalarThreads) wide_rz_kernel(const double* part, int nblk, int L, double* sc) {
    __shared__ double red[kScalarThreads];
    const double rz = lane_total(part, nblk, L, 0, red);
    const int l = threadIdx.x;
    if (l >= L) return;
    S(sc, kSRz, l) = S(sc, kSActive, l) != 0.0 ? rz : 0.0;
    S(sc, kSBeta, l) = 0.0;
}

// p = z on the lanes that restart their direction
__global__ void __launch_bounds__(kBlock) wide_restart_kernel(float* p, const float* z, const double* sc, int J, int L) {
    const long long total = static_cast<long long>(J) * L;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock)
        if (sc[kSRestart * kWideMaxLanes + static_cast<int>(i % L)] != 0.0) p[i] = z[i];
}

// x64 += x32, x32 = 0 on the lanes due for a replacement
__global__ void __launch_bounds__(kBlock)
wide_fold_kernel(double* x64, float* x32, long long total, int L, const double* sc) {
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        if (sc[kSWant * kWideMaxLanes + static_cast<int>(i % L)] == 0.0) continue;
        x64[i] += static_cast<double>(x32[i]);
        x32[i] = 0.0f;
    }
}

// every active lane is due (the last iteration)
__global__ void __launch_bounds__(kBlock) wide_want_all_kernel(double* sc) {
    const int l = threadIdx.x;
    if (S(sc, kSActive, l) != 0.0 && S(sc, kSWant, l) == 0.0) {
        S(sc, kSWant, l) = 1.0;
        S(sc, kSForced, l) = 1.0;  // replaced with the others, not on its own progress
    }
}

// r64 = b - q64 (ground rows zero)
__global__ void __launch_bounds__(kBlock)
wide_res_kernel(double* r64, const double* b, const double* q64, int J, int L, int ground) {
    const long long total = static_cast<long long>(J) * L;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int f = static_cast<int>(i / L);
        r64[i] = f == ground ? 0.0 : b[i] - q64[i];
    }
}

// the FP32 recurrence restarts from the FP64 residual (active lanes; the
// others are zeroed, direction included); partials of r.z
__global__ void __launch_bounds__(kBlock)
wide_reset_kernel(float* __restrict__ r, float* __restrict__ z, float* __restrict__ p, const double* __restrict__ r64,
                  const float* __restrict__ minv, const double* __restrict__ sc, int J, int L,
                  double* __restrict__ part, const float* __restrict__ minv_lane) {
    __shared__ double red[kBlock];
    const int l = threadIdx.x % L, r0 = threadIdx.x / L, R = kBlock / L;
    const bool act = sc[kSActive * kWideMaxLanes + l] != 0.0;
    const bool due = sc[kSWant * kWideMaxLanes + l] != 0.0;
    double rz = 0.0;
    for (long long j = static_cast<long long>(blockIdx.x) * R + r0; j < J; j += static_cast<long long>(gridDim.x) * R) {
        const size_t i = static_cast<size_t>(j) * L + l;
        if (act && !due) {
            rz += static_cast<double>(r[i]) * z[i];
        } else if (act) {
            const float ri = static_cast<float>(r64[i]);
            const float zi = (minv_lane ? minv_lane[i] : minv[j]) * ri;
            r[i] = ri;
            z[i] = zi;
            rz += static_cast<double>(ri) * zi;
        } else {
            r[i] = 0.0f;
            z[i] = 0.0f;
            p[i] = 0.0f;
        }
    }
    red[threadIdx.x] = rz;
    __syncthreads();
    if (r0 == 0) {
        double s = 0.0;
        for (int k = 0; k < R; ++k) s += red[k * L + l];
        part[static_cast<size_t>(blockIdx.x) * L + l] = s;
    }
}

// level-major b (J x L) from lane-major rhs (nb x J); ground rows and lanes past nb zero
__global__ void __launch_bounds__(kBlock)
wide_to_levels_kernel(const double* rhs, int J, int nb, int L, int ground, double* b) {
    const long long total = static_cast<long long>(J) * L;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int f = static_cast<int>(i / L);
        const int l = static_cast<int>(i % L);
        b[i] = (f == ground || l >= nb) ? 0.0 : rhs[static_cast<size_t>(l) * J + f];
    }
}

// lane-major z = bn x (unconverged lanes and the ground zero)
__global__ void __launch_bounds__(kBlock)
wide_to_lanes_kernel(const double* x, const double* sc, int J, int nb, int L, int ground, double* z) {
    const long long total = static_cast<long long>(J) * nb;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int l = static_cast<int>(i / J);
        const int f = static_cast<int>(i % J);
        const bool ok = sc[kSOk * kWideMaxLanes + l] != 0.0;
        z[i] = (f == ground || !ok) ? 0.0 : x[static_cast<size_t>(f) * L + l] * sc[kSBn * kWideMaxLanes + l];
    }
}

// x = bn x in place (level-major; lanes that did not converge and the ground zero)
__global__ void __launch_bounds__(kBlock) wide_unscale_kernel(double* x, const double* sc, int J, int L, int ground) {
    const long long total = static_cast<long long>(J) * L;
    for (long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; i < total;
         i += static_cast<long long>(gridDim.x) * kBlock) {
        const int f = static_cast<int>(i / L);
        const int l = static_cast<int>(i % L);
        const bool ok = sc[kSOk * kWideMaxLanes + l] != 0.0;
        x[i] = (f == ground || !ok) ? 0.0 : x[i] * sc[kSBn * kWideMaxLanes + l];
    }
}

int ew_grid(long long total) {
    const long long g = (total + kBlock - 1) / kBlock;
    return static_cast<int>(std::max(1LL, std::min(g, 8192LL)));
}

int wave_grid(long long waves) {
    const long long g = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
    return static_cast<int>(std::max(1LL, std::min(g, 16384LL)));
}

int dot_grid(int J, int L) {
    const long long rows_per_block = kBlock / L;
    const long long g = (static_cast<long long>(J) + rows_per_block * 8 - 1) / (rows_per_block * 8);
    return static_cast<int>(std::max(1LL, std::min<long long>(g, kMaxDotGrid)));
}

}  // namespace

void AkmWide::build(const AkmWideInput& in, hipStream_t stream, bool use_blocks) {
    s = stream;
    J = in.J;
    ground = in.ground;
    // directed pair entries (f, g, c_f c_g / Dw) of each worker's firms, and the
    // diagonal terms c_a (1 - C_w / Dw) of workers whose Dw is not their count total
    std::vector<std::pair<uint64_t, double>> ent;
    std::vector<double> corr(in.corr, in.corr + J);
    for (int k = 0; k < in.N; ++k) {
        const int a0 = in.wofs[k], a1 = in.wofs[k + 1];
        const double dw = in.dw[k];
        if (!(dw > 0.0)) continue;
        double cw = 0.0;
        for (int a = a0; a < a1; ++a) cw += in.rc[a];
        if (cw != dw)
            for (int a = a0; a < a1; ++a) corr[in.rf[a]] += in.rc[a] * (1.0 - cw / dw);
        for (int a = a0; a < a1; ++a)
            for (int b = a0; b < a1; ++b)
                if (a != b)
                    ent.emplace_back(static_cast<uint64_t>(in.rf[a]) * static_cast<uint64_t>(J) +
                                         static_cast<uint64_t>(in.rf[b]),
                                     in.rc[a] * in.rc[b] / dw);
    }
    // firm-major, neighbours ascending; a pair shared by several workers is one
    // entry (weights summed in sorted order: deterministic)
    std::sort(ent.begin(), ent.end());
    std::vector<int> h_key, h_nb, h_has(J, 0);
    std::vector<double> h_w64;
    for (size_t i = 0; i < ent.size(); ++i) {
        if (!h_key.empty() && i > 0 && ent[i].first == ent[i - 1].first) {
            h_w64.back() += ent[i].second;
            continue;
        }
        const int f = static_cast<int>(ent[i].first / static_cast<uint64_t>(J));
        h_key.push_back(f);
        h_nb.push_back(static_cast<int>(ent[i].first % static_cast<uint64_t>(J)));
        h_w64.push_back(ent[i].second);
        h_has[f] = 1;
    }
    E = static_cast<int>(h_key.size());
    std::vector<float> h_w32(h_w64.begin(), h_w64.end()), h_corr32(J), h_minv32(J);
    for (int f = 0; f < J; ++f) {
        h_corr32[f] = static_cast<float>(corr[f]);
        h_minv32[f] = f == ground ? 0.0f : static_cast<float>(in.inv_diag[f]);
    }
    auto up = [&](auto& d, const auto& h) {
        d.ensure(h.size());
        if (!h.empty())
            XFE_CHECK(hipMemcpyAsync(d.p, h.data(), sizeof(h[0]) * h.size(), hipMemcpyHostToDevice, s));
    };
    up(fkey, h_key); up(fnb, h_nb); up(has, h_has); up(w32, h_w32); up(w64, h_w64);
    up(corr32, h_corr32); up(corr64, corr); up(minv32, h_minv32);
    sc.ensure(kScTotal);
    if (!h_sc) XFE_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_sc), sizeof(double) * kScTotal, hipHostMallocDefault));
    // the blocks: undirected pairs f < g away from the ground firm, the exact
    // diagonal sum_g W_fg + corr_f of the grounded S
    blocks = false;
    if (use_blocks && E > 0) {
        std::vector<double> dex(corr);
        std::vector<int> pf, pg;
        std::vector<double> pw;
        for (int e = 0; e < E; ++e) {
            dex[h_key[e]] += h_w64[e];
            if (h_key[e] < h_nb[e] && h_key[e] != ground && h_nb[e] != ground) {
                pf.push_back(h_key[e]);
                pg.push_back(h_nb[e]);
                pw.push_back(h_w64[e]);
            }
        }
        if (ground >= 0) dex[ground] = 0.0;
        if (agg.set_pairs(J, pf, pg, pw, dex, s)) {
            agg.aggregate(agg.st.p, kWideBlockCap, kWideBlockRounds, s);
            agg.build_blocks(s);
            blocks = agg.nb > 0 && agg.nblocks > 0;
        }
    }
    XFE_CHECK(hipStreamSynchronize(s));  // the host vectors go out of scope
}

void AkmWide::build_worker_pairs(int J_, int ground_, const std::vector<int>& f, const std::vector<int>& g,
                                 const std::vector<int>& w, const std::vector<double>& base, hipStream_t stream) {
    s = stream;
    J = J_;
    ground = ground_;
    blocks = false;
    // firm-major order (stable in the given order within a firm)
    std::vector<int> order(f.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = static_cast<int>(i);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return f[a] < f[b]; });
    std::vector<int> h_key(order.size()), h_nb(order.size()), h_wk(order.size()), h_has(J, 0);
    std::vector<double> h_w64(order.size());
    for (size_t i = 0; i < order.size(); ++i) {
        const int e = order[i];
        h_key[i] = f[e];
        h_nb[i] = g[e];
        h_wk[i] = w[e];
        h_w64[i] = base[e];
        h_has[f[e]] = 1;
    }
    E = static_cast<int>(h_key.size());
    std::vector<float> h_w32(h_w64.begin(), h_w64.end()), zero32(J, 0.0f);
    std::vector<double> zero64(J, 0.0);
    auto up = [&](auto& d, const auto& h) {
        d.ensure(h.size());
        if (!h.empty())
            XFE_CHECK(hipMemcpyAsync(d.p, h.data(), sizeof(h[0]) * h.size(), hipMemcpyHostToDevice, s));
    };
    up(fkey, h_key); up(fnb, h_nb); up(fwk, h_wk); up(has, h_has); up(w32, h_w32); up(w64, h_w64);
    up(corr32, zero32); up(corr64, zero64); up(minv32, zero32);
    sc.ensure(kScTotal);
    if (!h_sc) XFE_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_sc), sizeof(double) * kScTotal, hipHostMallocDefault));
    XFE_CHECK(hipStreamSynchronize(s));
}

void AkmWide::unscale_x64() {
    hipLaunchKernelGGL(wide_unscale_kernel, dim3(ew_grid(static_cast<long long>(J) * L)), dim3(kBlock), 0, s, x64.p,
                       sc.p, J, L, ground);
    XFE_CHECK(hipGetLastError());
}

std::string AkmWide::lane_report(int l) const {
    char buf[256];
    std::snprintf(buf, sizeof(buf), "ok %g active %g iters %g bn %.3e r64 %.3e rref %.3e rr %.3e rz %.3e pq %.3e",
                  h_sc[kSOk * kWideMaxLanes + l], h_sc[kSActive * kWideMaxLanes + l],
                  h_sc[kSIters * kWideMaxLanes + l], h_sc[kSBn * kWideMaxLanes + l], h_sc[kSR64 * kWideMaxLanes + l],
                  h_sc[kSRref * kWideMaxLanes + l], h_sc[kSRr * kWideMaxLanes + l], h_sc[kSRz * kWideMaxLanes + l],
                  h_sc[kSPq * kWideMaxLanes + l]);
    return buf;
}

int AkmWide::lane_iterations(int l) const {
    return h_sc[kSOk * kWideMaxLanes + l] != 0.0 ? static_cast<int>(h_sc[kSIters * kWideMaxLanes + l]) : -1;
}

void AkmWide::mark_converged(int nb, const std::vector<char>& ok, const std::vector<int>& extra) {
    for (int l = 0; l < nb && l < kWideMaxLanes; ++l) {
        if (!ok[l]) continue;
        h_sc[kSOk * kWideMaxLanes + l] = 1.0;
        h_sc[kSIters * kWideMaxLanes + l] += extra[l];
    }
    XFE_CHECK(hipMemcpyAsync(sc.p + kSOk * kWideMaxLanes, h_sc + kSOk * kWideMaxLanes, sizeof(double) * kWideMaxLanes,
                             hipMemcpyHostToDevice, s));
    XFE_CHECK(hipStreamSynchronize(s));
}

void AkmWide::release() {
    agg.release();
    for (DevBuf<int>* d : {&fkey, &fnb, &fwk, &has, &bk[0], &bk[1]}) d->release();
    for (DevBuf<float>* d : {&w32, &corr32, &minv32, &bv32[0], &bv32[1], &x, &r, &z, &p, &q}) d->release();
    for (DevBuf<double>* d : {&w64, &corr64, &bv64[0], &bv64[1], &x64, &b64, &r64, &q64, &part, &sc}) d->release();
    if (h_sc) (void)hipHostFree(h_sc);
    h_sc = nullptr;
}

void AkmWide::set_operator(WideOp* o, int n_rows, hipStream_t stream, const float* minv) {
    op = o;
    s = stream;
    J = n_rows;
    ground = -1;
    E = 0;
    blocks = false;
    mult = nullptr;
    minv_lane = nullptr;
    minv32.ensure(n_rows);
    XFE_CHECK(hipMemcpyAsync(minv32.p, minv, sizeof(float) * n_rows, hipMemcpyDeviceToDevice, s));
    sc.ensure(kScTotal);
    if (!h_sc) XFE_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_sc), sizeof(double) * kScTotal, hipHostMallocDefault));
}

void AkmWide::ensure(int nb) {
    const int V = nb <= 64 ? 1 : nb <= 128 ? 2 : 4;
    L = (op && nb <= 32) ? 32 : 64 * V;
    const size_t JL = static_cast<size_t>(J) * L;
    const size_t BL = 2 * ((static_cast<size_t>(std::max(E, 1)) + kWPiece - 1) / kWPiece);
    for (DevBuf<float>* d : {&x, &r, &z, &p, &q}) d->ensure(JL);
    for (DevBuf<double>* d : {&x64, &b64, &r64, &q64}) d->ensure(JL);
    for (int k = 0; k < 2; ++k) {
        bk[k].ensure(BL);
        bv32[k].ensure(BL * L);
        bv64[k].ensure(BL * L);
    }
    part.ensure(2 * static_cast<size_t>(kMaxDotGrid) * kWideMaxLanes);
}

template <typename T, int V>
static void apply_v(AkmWide& w, const T* pv, T* out, const T* w_, const T* corr_, DevBuf<T>* bv) {
    const int L = w.L;
    const hipStream_t s = w.s;
    if (w.E > 0) {
        int m = w.E;
        int pieces = (m + kWPiece - 1) / kWPiece;
        const int slot = sizeof(T) == 4 ? profile_begin(s) : -1;
        if (w.mult)
            hipLaunchKernelGGL((wide_pair_kernel<T, V, true>), dim3(wave_grid(pieces)), dim3(kBlock), 0, s, w.fkey.p,
                               w.fnb.p, w_, pv, m, L, out, w.bk[0].p, bv[0].p, pieces == 1 ? 1 : 0, w.fwk.p, w.mult,
                               w.mstride, w.mcols);
        else
            hipLaunchKernelGGL((wide_pair_kernel<T, V, false>), dim3(wave_grid(pieces)), dim3(kBlock), 0, s, w.fkey.p,
                               w.fnb.p, w_, pv, m, L, out, w.bk[0].p, bv[0].p, pieces == 1 ? 1 : 0, nullptr, nullptr,
                               0, 1);
        XFE_CHECK(hipGetLastError());
        // the FP32 pair pass: entries (key, neighbour, weight), the neighbour's L
        // lanes per entry, the firm's lanes read and written once
        if (slot >= 0)
            profile_end(slot, s, static_cast<double>(w.E) * (12.0 + sizeof(T) * L) + 2.0 * sizeof(T) * L * w.J);
        int cur = 0;
        while (pieces > 1) {
            m = 2 * pieces;
            pieces = (m + kWPiece - 1) / kWPiece;
            hipLaunchKernelGGL((wide_list_kernel<T, V>), dim3(wave_grid(pieces)), dim3(kBlock), 0, s, w.bk[cur].p,
                               bv[cur].p, m, L, out, w.bk[1 - cur].p, bv[1 - cur].p, pieces == 1 ? 1 : 0);
            XFE_CHECK(hipGetLastError());
            cur = 1 - cur;
        }
    }
    hipLaunchKernelGGL((wide_finish_kernel<T>), dim3(ew_grid(static_cast<long long>(w.J) * L)), dim3(kBlock), 0, s,
                       out, pv, corr_, w.has.p, w.J, L, w.ground);
    XFE_CHECK(hipGetLastError());
}

template <typename T>
void AkmWide::apply(const T* pv, T* out, const T* w_, const T* corr_, DevBuf<T>* bv) {
    if (op) {
        op->apply(pv, out);
        return;
    }
    if (L == 64) apply_v<T, 1>(*this, pv, out, w_, corr_, bv);
    else if (L == 128) apply_v<T, 2>(*this, pv, out, w_, corr_, bv);
    else apply_v<T, 4>(*this, pv, out, w_, corr_, bv);
}

template void AkmWide::apply<float>(const float*, float*, const float*, const float*, DevBuf<float>*);
template void AkmWide::apply<double>(const double*, double*, const double*, const double*, DevBuf<double>*);

void AkmWide::solve_device(int nb, double tol, int max_iter, double accept) {
    const long long JL = static_cast<long long>(J) * L;
    const int eg = ew_grid(JL), dg = dot_grid(J, L);
    const double tol2 = tol * tol;
    const double accept2 = accept > tol ? accept * accept : tol2;
    // z = inv(block) r on the blocked firms (Jacobi elsewhere, from the step or
    // reset kernel) and the partials of r.z in place of the Jacobi ones; an
    // inactive lane has r = 0, so its z stays 0
    auto block_precondition = [&]() {
        if (op && op->has_blocks()) {
            op->precondition(r.p, z.p);
            hipLaunchKernelGGL((wide_dot_kernel<float>), dim3(dg), dim3(kBlock), 0, s, r.p, z.p, J, L, part.p);
            XFE_CHECK(hipGetLastError());
            return;
        }
        if (!blocks) return;
        XFE_CHECK(lp::launch_block_apply(agg.view(), r.p, z.p, L, s));
        hipLaunchKernelGGL((wide_dot_kernel<float>), dim3(dg), dim3(kBlock), 0, s, r.p, z.p, J, L, part.p);
        XFE_CHECK(hipGetLastError());
    };
    last_updates = 0;
    last_passes = 0;
    hipLaunchKernelGGL(wide_init_kernel, dim3(1), dim3(kWideMaxLanes), 0, s, sc.p, nb);
    XFE_CHECK(hipGetLastError());
    // unit-norm right-hand sides
    hipLaunchKernelGGL((wide_dot_kernel<double>), dim3(dg), dim3(kBlock), 0, s, b64.p, b64.p, J, L, part.p);
    hipLaunchKernelGGL(wide_norm_kernel, dim3(1), dim3(kScalarThreads), 0, s, part.p, dg, L, sc.p);
    hipLaunchKernelGGL(wide_scale_kernel, dim3(eg), dim3(kBlock), 0, s, b64.p, sc.p, J, L);
    XFE_CHECK(hipGetLastError());
    XFE_CHECK(hipMemsetAsync(x64.p, 0, sizeof(double) * JL, s));
    XFE_CHECK(hipMemsetAsync(x.p, 0, sizeof(float) * JL, s));
    XFE_CHECK(hipMemsetAsync(p.p, 0, sizeof(float) * JL, s));
    XFE_CHECK(hipMemcpyAsync(r64.p, b64.p, sizeof(double) * JL, hipMemcpyDeviceToDevice, s));
    // FP64 residual check, FP32 recurrence from it; p = z
    auto replace = [&]() {
        hipLaunchKernelGGL((wide_dot_kernel<double>), dim3(dg), dim3(kBlock), 0, s, r64.p, r64.p, J, L, part.p);
        hipLaunchKernelGGL(wide_check_kernel, dim3(1), dim3(kScalarThreads), 0, s, part.p, dg, L, sc.p, tol2,
                           max_iter, accept2, restart_each_replacement ? 1 : 0);
        hipLaunchKernelGGL(wide_reset_kernel, dim3(dg), dim3(kBlock), 0, s, r.p, z.p, p.p, r64.p, minv32.p, sc.p, J,
                           L, part.p, minv_lane);
        block_precondition();
        hipLaunchKernelGGL(wide_restart_kernel, dim3(eg), dim3(kBlock), 0, s, p.p, z.p, sc.p, J, L);
        hipLaunchKernelGGL(wide_rz_kernel, dim3(1), dim3(kScalarThreads), 0, s, part.p, dg, L, sc.p);
        XFE_CHECK(hipGetLastError());
    };
    replace();
    hipLaunchKernelGGL(wide_dir_kernel, dim3(eg), dim3(kBlock), 0, s, p.p, z.p, sc.p, J, L);
    XFE_CHECK(hipGetLastError());
    XFE_CHECK(hipMemcpyAsync(h_sc + kSFlags, sc.p + kSFlags, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    bool live = h_sc[kSFlags + 1] > 0.0;
    for (int it = 1; live && it <= max_iter; ++it) {
        apply<float>(p.p, q.p, w32.p, corr32.p, bv32);
        ++last_passes;
        hipLaunchKernelGGL((wide_dot_kernel<float>), dim3(dg), dim3(kBlock), 0, s, p.p, q.p, J, L, part.p);
        hipLaunchKernelGGL(wide_alpha_kernel, dim3(1), dim3(kScalarThreads), 0, s, part.p, dg, L, sc.p);
        hipLaunchKernelGGL(wide_step_kernel, dim3(dg), dim3(kBlock), 0, s, x.p, r.p, z.p, p.p, q.p, minv32.p, sc.p,
                           J, L, part.p, minv_lane);
        block_precondition();
        hipLaunchKernelGGL(wide_beta_kernel, dim3(1), dim3(kScalarThreads), 0, s, part.p, dg, L, sc.p, tol2,
                           max_iter, restart_each_replacement ? kMaxRunRestart : kMaxRun,
                           restart_each_replacement ? kDelta2Restart : kDelta2);
        hipLaunchKernelGGL(wide_dir_kernel, dim3(eg), dim3(kBlock), 0, s, p.p, z.p, sc.p, J, L);
        XFE_CHECK(hipGetLastError());
        if (it % kPoll != 0 && it != max_iter) continue;
        XFE_CHECK(hipMemcpyAsync(h_sc + kSFlags, sc.p + kSFlags, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        if (h_sc[kSFlags + 1] <= 0.0) break;
        if (trace && it % 256 == 0) {
            XFE_CHECK(hipMemcpyAsync(h_sc, sc.p, sizeof(double) * kScTotal, hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
            double least = 1.0;
            for (int l = 0; l < nb; ++l)
                if (h_sc[kSActive * kWideMaxLanes + l] != 0.0 && h_sc[kSRref * kWideMaxLanes + l] > 0.0)
                    least = std::min(least, h_sc[kSRr * kWideMaxLanes + l] / h_sc[kSRref * kWideMaxLanes + l]);
            std::fprintf(stderr, "[%s] iteration %d: %g lanes active, least r.r drop of the run %.3e\n", trace, it,
                         h_sc[kSFlags + 1], least);
        }
        if (h_sc[kSFlags] == 0.0 && it != max_iter) continue;
        // FP64 residual replacement: x64 += x32, r64 = b - S x64
        // the restarted mode replaces the lanes that are due; the kept-direction
        // mode replaces every active lane together (a lane whose recurrence has
        // not yet dropped keeps its direction, and a stale one is restarted)
        if (it == max_iter || !restart_each_replacement)
            hipLaunchKernelGGL(wide_want_all_kernel, dim3(1), dim3(kWideMaxLanes), 0, s, sc.p);
        hipLaunchKernelGGL(wide_fold_kernel, dim3(eg), dim3(kBlock), 0, s, x64.p, x.p, JL, L, sc.p);
        apply<double>(x64.p, q64.p, w64.p, corr64.p, bv64);
        hipLaunchKernelGGL(wide_res_kernel, dim3(eg), dim3(kBlock), 0, s, r64.p, b64.p, q64.p, J, L, ground);
        XFE_CHECK(hipGetLastError());
        replace();
        ++last_updates;
        XFE_CHECK(hipMemcpyAsync(h_sc + kSFlags, sc.p + kSFlags, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        if (trace)
            XFE_CHECK(hipMemcpyAsync(h_sc + kSR64 * kWideMaxLanes, sc.p + kSR64 * kWideMaxLanes,
                                     sizeof(double) * kWideMaxLanes, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        live = h_sc[kSFlags + 1] > 0.0;
        if (trace) {
            double worst = 0.0;
            for (int l = 0; l < nb; ++l) worst = std::max(worst, h_sc[kSR64 * kWideMaxLanes + l]);
            std::fprintf(stderr, "[%s] iteration %d replacement %d: %g lanes active, worst |r|/|b| %.3e\n", trace, it,
                         last_updates, h_sc[kSFlags + 1], std::sqrt(worst));
        }
    }
    // lanes still running at the end stop unconverged (their x32 is not folded)
    XFE_CHECK(hipMemcpyAsync(h_sc, sc.p, sizeof(double) * kScTotal, hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    for (int l = 0; l < nb; ++l)
        if (h_sc[kSActive * kWideMaxLanes + l] != 0.0) h_sc[kSOk * kWideMaxLanes + l] = 0.0;
    XFE_CHECK(hipMemcpyAsync(sc.p + kSOk * kWideMaxLanes, h_sc + kSOk * kWideMaxLanes, sizeof(double) * kWideMaxLanes,
                             hipMemcpyHostToDevice, s));
}

void AkmWide::solve(const double* rhs_pack, double* z_pack, int nb, double tol, int max_iter, int* iters_out,
                    bool device_ptrs) {
    ensure(nb);
    const long long JL = static_cast<long long>(J) * L;
    const size_t Jn = static_cast<size_t>(J) * nb;
    // lane-major input (staged in q64, J x L >= nb x J, when it is on the host),
    // then level-major b64
    const double* src = rhs_pack;
    if (!device_ptrs) {
        XFE_CHECK(hipMemcpyAsync(q64.p, rhs_pack, sizeof(double) * Jn, hipMemcpyHostToDevice, s));
        src = q64.p;
    }
    hipLaunchKernelGGL(wide_to_levels_kernel, dim3(ew_grid(JL)), dim3(kBlock), 0, s, src, J, nb, L, ground, b64.p);
    XFE_CHECK(hipGetLastError());
    solve_device(nb, tol, max_iter);
    double* dst = device_ptrs ? z_pack : q64.p;
    hipLaunchKernelGGL(wide_to_lanes_kernel, dim3(ew_grid(static_cast<long long>(Jn))), dim3(kBlock), 0, s, x64.p,
                       sc.p, J, nb, L, ground, dst);
    XFE_CHECK(hipGetLastError());
    if (!device_ptrs)
        XFE_CHECK(hipMemcpyAsync(z_pack, q64.p, sizeof(double) * Jn, hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    for (int l = 0; l < nb; ++l) {
        const bool ok = h_sc[kSOk * kWideMaxLanes + l] != 0.0;
        iters_out[l] = ok ? static_cast<int>(h_sc[kSIters * kWideMaxLanes + l]) : -1;
    }
}

}  // namespace xfe
