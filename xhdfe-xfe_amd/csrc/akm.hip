// AKM/KSS leverage solver: the grounded worker-partialled firm Laplacian
//     S = Df - B' Dw^{-1} B
// (B: worker x firm match counts m_c) and its Jacobi-preconditioned CG for
// up to kMaxCgCols right-hand sides at once.  The drop-in for the reference's
// akm_cuda_* functions (include/hdfe/akm_kss_cuda.hpp:19-74, impl
// src/akm_kss_cuda.cu:164-722).
//
// S p = Df p - B'(Dw^{-1} B p) is applied to up to 32 lanes at once in FP64
// (level-major J x nb vectors): one pass over the worker-major matches forms
// the worker means t_w = Dw^{-1} sum_a c_a p[f_a], one pass over the
// firm-major matches sums c_a (p[f] - t_w(a)) per firm with a deterministic
// hierarchical reduction of the runs that cross piece boundaries, and a last
// kernel adds (Df - deg) p and pins the ground firm (J - 1) to zero as in the
// reference.  Lanes are spread over threads (32 per worker or piece), so every
// gather reads a worker's or a firm's lanes as one contiguous block.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <utility>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <string>
#include <vector>

#include "absorb_kernels.h"
#include "akm_wide.h"
#include "common.h"
#include "xfe_hip.h"

namespace xfe {
namespace {

// level-major (f * nb + l) <- lane-major (l * J + f), ground rows zeroed
__global__ void __launch_bounds__(kBlock)
akm_to_levels(const double* lanes, int J, int nb, int ground, double* levels) {
    const long long total = static_cast<long long>(J) * nb;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int f = static_cast<int>(j / nb);
        const int l = static_cast<int>(j - static_cast<long long>(f) * nb);
        levels[j] = f == ground ? 0.0 : lanes[static_cast<size_t>(l) * J + f];
    }
}

// lane-major z = unscale_l * x, ground zero, lanes that did not converge zero
__global__ void __launch_bounds__(kBlock)
akm_to_lanes(const double* x, int J, int nb, int ground, const double* unscale, const double* cg,
             double* lanes) {
    const long long total = static_cast<long long>(J) * nb;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int l = static_cast<int>(j / J);
        const int f = static_cast<int>(j - static_cast<long long>(l) * J);
        const bool ok = cg[kCgOk + l] != 0.0;
        lanes[j] = (f == ground || !ok) ? 0.0 : x[static_cast<size_t>(f) * nb + l] * unscale[l];
    }
}

// out = Df p - B' t: the firm pass gave sum_a c_a (p[f] - t_w(a)) for firms
// with matches; add (Df - deg) p, take Df p for firms without matches, pin the ground
__global__ void __launch_bounds__(kBlock)
akm_finish(double* out, const double* p, const double* corr, const int* has, int J, int nb, int ground) {
    const long long total = static_cast<long long>(J) * nb;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int f = static_cast<int>(j / nb);
        double v = has[f] ? out[j] : 0.0;
        v += corr[f] * p[j];
        out[j] = f == ground ? 0.0 : v;
    }
}

__global__ void __launch_bounds__(kBlock) akm_iota64(int64_t* v, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * kBlock)
        v[i] = static_cast<int64_t>(i);
}

// S p = Df p - B'(Dw^{-1} B p) for nb <= 32 lanes at once, lane-parallel: a
// group of 32 threads takes one worker (or one piece of firm-major matches)
// and thread c holds lane c, so every gather of a worker's or firm's lanes is
// one contiguous 8 nb-byte read.
constexpr int kLaneGroup = 32;
constexpr int kGroups = kBlock / kLaneGroup;  // lane groups per block
constexpr int kPiece = 64;                    // firm-major matches (or list entries) per group

constexpr int kBatch = 16;  // gathers issued together by a lane group

// A group's records of a piece of at most kPiece consecutive matches: thread c
// loads entries j0 + c and j0 + 32 + c (coalesced); entry e is then read by
// every thread of the group with a shuffle.
template <typename T>
__device__ __forceinline__ T group_entry(T a, T b, int e) {
    return e < kLaneGroup ? __shfl(a, e, kLaneGroup) : __shfl(b, e - kLaneGroup, kLaneGroup);
}

// Worker means over pieces of whole workers (piece boundaries pstart fall on
// worker boundaries; a piece holds at most kPiece matches unless one worker has
// more): t[w * nb + c] = sum over the worker's matches a of
// (c_a / Dw[w]) p[f_a * nb + c].  Records come in kPiece at a time with two
// coalesced loads, the p gathers of kBatch matches are issued together, then
// the runs of equal worker are summed.
__global__ void __launch_bounds__(kBlock)
akm_wmean_kernel(const int* __restrict__ pstart, int npieces, const int* __restrict__ rw,
                 const int* __restrict__ rf, const double* __restrict__ rwt, const double* __restrict__ p, int nb,
                 double* __restrict__ t) {
    const int c = threadIdx.x % kLaneGroup;
    const int stride = gridDim.x * kGroups;
    // both groups of a wave run the same number of rounds and chunks (shuffles
    // need every lane of the wave)
    const int q0 = blockIdx.x * kGroups + threadIdx.x / kLaneGroup;
    const int rounds = (npieces - (q0 - q0 % 2) + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {
        const int q = q0 + r * stride;
        const bool live = q < npieces;
        const int j0 = live ? pstart[q] : 0;
        const int len = live ? pstart[q + 1] - j0 : 0;
        const int cl = c < nb ? c : 0;
        const int chunks = (len + kPiece - 1) / kPiece;
        const int cmax = max(chunks, __shfl_xor(chunks, kLaneGroup));
        int cur = -1;
        double acc = 0.0;
        for (int ch = 0; ch < cmax; ++ch) {
            const int base = j0 + ch * kPiece;
            const int n = min(max(len - ch * kPiece, 0), kPiece);
            const int wA = c < n ? rw[base + c] : -1, wB = c + kLaneGroup < n ? rw[base + kLaneGroup + c] : -1;
            const int fA = c < n ? rf[base + c] : 0, fB = c + kLaneGroup < n ? rf[base + kLaneGroup + c] : 0;
            const double xA = c < n ? rwt[base + c] : 0.0;
            const double xB = c + kLaneGroup < n ? rwt[base + kLaneGroup + c] : 0.0;
            const int nmax = max(n, __shfl_xor(n, kLaneGroup));
            if (ch == 0) cur = group_entry(wA, wB, 0);
            for (int e0 = 0; e0 < nmax; e0 += kBatch) {
                double v[kBatch];
                // unconditional loads (an entry past n reads firm 0, a lane past nb
                // lane 0; both are masked below), so the kBatch gathers are in flight
                // together
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    const int f = group_entry(fA, fB, e0 + u);
                    v[u] = p[static_cast<size_t>(f) * nb + cl];
                }
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    const int w = group_entry(wA, wB, e0 + u);
                    const double x = group_entry(xA, xB, e0 + u);
                    if (e0 + u < n) {
                        if (w != cur) {
                            if (c < nb) t[static_cast<size_t>(cur) * nb + c] = acc;
                            cur = w;
                            acc = 0.0;
                        }
                        acc += x * v[u];  // x = 0 past n
                    }
                }
            }
        }
        if (len > 0 && c < nb) t[static_cast<size_t>(cur) * nb + c] = acc;
    }
}

// Firm pass over fixed pieces of kPiece firm-major matches (records: firm key,
// worker, count): value c_a (p[f] - t[w_a]) per match with the t and p gathers
// of kBatch matches issued together, runs of equal firm summed; a run strictly
// inside its piece is stored in out, the first and the last run of the piece go
// to the boundary list (entries 2q and 2q + 1, a one-run piece adds (key, 0) as
// its second entry), unless final (the matches are one piece: every run stored).
__global__ void __launch_bounds__(kBlock)
akm_firm_kernel(const int* __restrict__ fkey, const int* __restrict__ fw, const double* __restrict__ fc,
                const double* __restrict__ p, const double* __restrict__ t, int m, int nb, double* __restrict__ out,
                int* __restrict__ kout, double* __restrict__ vout, int final) {
    const int c = threadIdx.x % kLaneGroup;
    const int stride = gridDim.x * kGroups;
    const int pieces = (m + kPiece - 1) / kPiece;
    const int q0 = blockIdx.x * kGroups + threadIdx.x / kLaneGroup;
    const int rounds = (pieces - (q0 - q0 % 2) + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {
        const int q = q0 + r * stride;
        const bool live = q < pieces;
        const int j0 = q * kPiece;
        const int n = live ? min(m - j0, kPiece) : 0;
        const int kA = c < n ? fkey[j0 + c] : 0, kB = c + kLaneGroup < n ? fkey[j0 + kLaneGroup + c] : 0;
        const int wA = c < n ? fw[j0 + c] : 0, wB = c + kLaneGroup < n ? fw[j0 + kLaneGroup + c] : 0;
        const double xA = c < n ? fc[j0 + c] : 0.0, xB = c + kLaneGroup < n ? fc[j0 + kLaneGroup + c] : 0.0;
        const int nmax = max(n, __shfl_xor(n, kLaneGroup));
        const int cl = c < nb ? c : 0;
        int key = group_entry(kA, kB, 0);
        double acc = 0.0;
        bool first = true;
        for (int e0 = 0; e0 < nmax; e0 += kBatch) {
            double v[kBatch];
            // unconditional loads (an entry past n reads firm 0 and worker 0, a lane
            // past nb lane 0; both are masked below), so the 2 kBatch gathers are in
            // flight together
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int k = group_entry(kA, kB, e0 + u);
                const int w = group_entry(wA, wB, e0 + u);
                v[u] = p[static_cast<size_t>(k) * nb + cl] - t[static_cast<size_t>(w) * nb + cl];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int k = group_entry(kA, kB, e0 + u);
                const double x = group_entry(xA, xB, e0 + u);
                if (e0 + u < n) {
                    if (k != key) {
                        if (c < nb) {
                            if (first && !final) vout[static_cast<size_t>(2 * q) * nb + c] = acc;
                            else out[static_cast<size_t>(key) * nb + c] = acc;
                        }
                        if (first && !final && c == 0) kout[2 * q] = key;
                        first = false;
                        key = k;
                        acc = 0.0;
                    }
                    acc += x * v[u];
                }
            }
        }
        if (n > 0) {
            if (final) {
                if (c < nb) out[static_cast<size_t>(key) * nb + c] = acc;
            } else {
                const size_t e0 = static_cast<size_t>(2 * q) * nb, e1 = static_cast<size_t>(2 * q + 1) * nb;
                if (c < nb) {
                    if (first) {
                        vout[e0 + c] = acc;
                        vout[e1 + c] = 0.0;
                    } else {
                        vout[e1 + c] = acc;
                    }
                }
                if (c == 0) {
                    if (first) kout[2 * q] = key;
                    kout[2 * q + 1] = key;
                }
            }
        }
    }
}

// One level of the boundary list reduction: entries (key, nb values) sorted by
// key, in pieces of kPiece entries per lane group, with the same run rule as
// akm_firm_kernel (a key's total is stored once or split over the next
// level's entries only); keys come in with two coalesced loads and the values
// of kBatch entries are loaded together.
__global__ void __launch_bounds__(kBlock)
akm_list_kernel(const int* __restrict__ key_in, const double* __restrict__ vin, int m, int nb,
                double* __restrict__ out, int* __restrict__ kout, double* __restrict__ vout, int final) {
    const int c = threadIdx.x % kLaneGroup;
    const int stride = gridDim.x * kGroups;
    const int pieces = (m + kPiece - 1) / kPiece;
    const int q0 = blockIdx.x * kGroups + threadIdx.x / kLaneGroup;
    const int rounds = (pieces - (q0 - q0 % 2) + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {
        const int q = q0 + r * stride;
        const bool live = q < pieces;
        const int j0 = q * kPiece;
        const int n = live ? min(m - j0, kPiece) : 0;
        const int kA = c < n ? key_in[j0 + c] : 0, kB = c + kLaneGroup < n ? key_in[j0 + kLaneGroup + c] : 0;
        const int nmax = max(n, __shfl_xor(n, kLaneGroup));
        const int cl = c < nb ? c : 0;
        int key = group_entry(kA, kB, 0);
        double acc = 0.0;
        bool first = true;
        for (int e0 = 0; e0 < nmax; e0 += kBatch) {
            // unconditional loads (an entry past n reads entry 0, a lane past nb
            // lane 0; both are masked below)
            double v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int j = e0 + u < n ? j0 + e0 + u : 0;
                v[u] = vin[static_cast<size_t>(j) * nb + cl];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int k = group_entry(kA, kB, e0 + u);
                if (e0 + u < n) {
                    if (k != key) {
                        if (c < nb) {
                            if (first && !final) vout[static_cast<size_t>(2 * q) * nb + c] = acc;
                            else out[static_cast<size_t>(key) * nb + c] = acc;
                        }
                        if (first && !final && c == 0) kout[2 * q] = key;
                        first = false;
                        key = k;
                        acc = 0.0;
                    }
                    acc += v[u];
                }
            }
        }
        if (n > 0) {
            if (final) {
                if (c < nb) out[static_cast<size_t>(key) * nb + c] = acc;
            } else {
                const size_t e0 = static_cast<size_t>(2 * q) * nb, e1 = static_cast<size_t>(2 * q + 1) * nb;
                if (c < nb) {
                    if (first) {
                        vout[e0 + c] = acc;
                        vout[e1 + c] = 0.0;
                    } else {
                        vout[e1 + c] = acc;
                    }
                }
                if (c == 0) {
                    if (first) kout[2 * q] = key;
                    kout[2 * q + 1] = key;
                }
            }
        }
    }
}

// ---- JLA draws on the device (src/akm_kss.cpp:3153-3285, the P_ii part)

// SplitMix64 (Steele, Lea and Flood 2014) and the reference's per-(draw,
// unit) stream state (src/akm_kss.cpp:212-222): the sign of working row m in
// draw d is bit m mod 64 of stream_seed(seed, 3 d, m / 64).
__host__ __device__ __forceinline__ uint64_t jla_mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

__device__ __forceinline__ double jla_sign(uint64_t seed, int d, int m) {
    const uint64_t s0 = jla_mix(seed ^ (0xA24BAED4963EE407ULL * (3ULL * static_cast<uint64_t>(d) + 1ULL)));
    const uint64_t st = jla_mix(s0 ^ (0x9FB21C651E98DF25ULL * (static_cast<uint64_t>(m >> 6) + 1ULL)));
    return ((st >> (m & 63)) & 1ULL) ? 1.0 : -1.0;
}

// tw[w nb + l] = sum over the worker's matches a of sqrt(c_a) r_a(d0 + l)
__global__ void __launch_bounds__(kBlock)
jla_tw_kernel(const int* wofs, int N, const int* orig, const double* sq, uint64_t seed, int d0, int nb,
              double* tw) {
    const long long total = static_cast<long long>(N) * nb;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int w = static_cast<int>(j / nb);
        const int l = static_cast<int>(j - static_cast<long long>(w) * nb);
        double acc = 0.0;
        for (int a = wofs[w]; a < wofs[w + 1]; ++a) acc += sq[a] * jla_sign(seed, d0 + l, orig[a]);
        tw[j] = acc;
    }
}

// Firm pass of the draws' Schur right-hand sides: per firm, the sum over its
// matches of c (r / sqrt(c) - tw[w] / Dw[w]) = sqrt(c) r - c tw[w] / Dw[w]
// (with su: s - c tw[w] / Dw[w], s = su of the row on lanes below h, else 0),
// with the run rule of akm_firm_kernel (pieces of kPiece firm-major matches,
// runs crossing a piece boundary into the boundary list).
__global__ void __launch_bounds__(kBlock)
jla_firm_kernel(const int* __restrict__ fkey, const int* __restrict__ fw, const double* __restrict__ fc,
                const int* __restrict__ forig, const double* __restrict__ tw, const double* __restrict__ idw,
                uint64_t seed, int d0, int m, int nb, double* __restrict__ out, int* __restrict__ kout,
                double* __restrict__ vout, int final, const double* __restrict__ su, int h) {
    const int c = threadIdx.x % kLaneGroup;
    const int stride = gridDim.x * kGroups;
    const int pieces = (m + kPiece - 1) / kPiece;
    const int q0 = blockIdx.x * kGroups + threadIdx.x / kLaneGroup;
    const int rounds = (pieces - (q0 - q0 % 2) + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {
        const int q = q0 + r * stride;
        const bool live = q < pieces;
        const int j0 = q * kPiece;
        const int n = live ? min(m - j0, kPiece) : 0;
        const int kA = c < n ? fkey[j0 + c] : 0, kB = c + kLaneGroup < n ? fkey[j0 + kLaneGroup + c] : 0;
        const int wA = c < n ? fw[j0 + c] : 0, wB = c + kLaneGroup < n ? fw[j0 + kLaneGroup + c] : 0;
        const int oA = c < n ? forig[j0 + c] : 0, oB = c + kLaneGroup < n ? forig[j0 + kLaneGroup + c] : 0;
        const double xA = c < n ? fc[j0 + c] : 1.0, xB = c + kLaneGroup < n ? fc[j0 + kLaneGroup + c] : 1.0;
        const int nmax = max(n, __shfl_xor(n, kLaneGroup));
        const int cl = c < nb ? c : 0;
        int key = group_entry(kA, kB, 0);
        double acc = 0.0;
        bool first = true;
        for (int e0 = 0; e0 < nmax; e0 += kBatch) {
            double v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int w = group_entry(wA, wB, e0 + u);
                v[u] = tw[static_cast<size_t>(w) * nb + cl] * idw[w];
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int k = group_entry(kA, kB, e0 + u);
                const int o = group_entry(oA, oB, e0 + u);
                const double x = group_entry(xA, xB, e0 + u);
                if (e0 + u < n) {
                    if (k != key) {
                        if (c < nb) {
                            if (first && !final) vout[static_cast<size_t>(2 * q) * nb + c] = acc;
                            else out[static_cast<size_t>(key) * nb + c] = acc;
                        }
                        if (first && !final && c == 0) kout[2 * q] = key;
                        first = false;
                        key = k;
                        acc = 0.0;
                    }
                    if (su) acc += (cl < h ? su[static_cast<size_t>(o) * h + cl] : 0.0) - x * v[u];
                    else acc += x * (jla_sign(seed, d0 + cl, o) / sqrt(x) - v[u]);
                }
            }
        }
        if (n > 0) {
            if (final) {
                if (c < nb) out[static_cast<size_t>(key) * nb + c] = acc;
            } else {
                const size_t e0 = static_cast<size_t>(2 * q) * nb, e1 = static_cast<size_t>(2 * q + 1) * nb;
                if (c < nb) {
                    if (first) {
                        vout[e0 + c] = acc;
                        vout[e1 + c] = 0.0;
                    } else {
                        vout[e1 + c] = acc;
                    }
                }
                if (c == 0) {
                    if (first) kout[2 * q] = key;
                    kout[2 * q + 1] = key;
                }
            }
        }
    }
}

// zf = unscale_l x (level-major), zero on the ground firm and on lanes that did not converge
__global__ void __launch_bounds__(kBlock)
jla_unscale_kernel(const double* x, int J, int nb, int ground, const double* unscale, const double* cg, double* zf) {
    const long long total = static_cast<long long>(J) * nb;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int f = static_cast<int>(j / nb);
        const int l = static_cast<int>(j - static_cast<long long>(f) * nb);
        zf[j] = (f == ground || cg[kCgOk + l] == 0.0) ? 0.0 : x[j] * unscale[l];
    }
}

// zw[w nb + l] = (tw[w nb + l] - sum over the worker's matches of c_a zf[f_a nb + l]) / Dw[w]
__global__ void __launch_bounds__(kBlock)
jla_zw_kernel(const int* wofs, int N, const int* mf, const double* mc, const double* idw, const double* tw,
              const double* zf, int nb, double* zw) {
    const long long total = static_cast<long long>(N) * nb;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int w = static_cast<int>(j / nb);
        const int l = static_cast<int>(j - static_cast<long long>(w) * nb);
        double acc = tw[j];
        for (int a = wofs[w]; a < wofs[w + 1]; ++a) acc -= mc[a] * zf[static_cast<size_t>(mf[a]) * nb + l];
        zw[j] = acc * idw[w];
    }
}

// Per working row (worker-major record a, row orig[a]): Z = sqrt(c)(zw + zf)
// and r for the nb draws; P += Z^2 / p, M += (r - Z)^2 / p; with mom (3 M:
// the means of Z^4, (r - Z)^4 and Z^2 (r - Z)^2, for the ratio estimator's
// non-linearity correction, src/akm_kss.cpp:3194-3206) those too
__global__ void __launch_bounds__(kBlock)
jla_accumulate_kernel(int M, const int* mw, const int* mf, const int* orig, const double* sq, const double* zw,
                      const double* zf, uint64_t seed, int d0, int nb, double inv_p, double* pacc, double* macc,
                      double* mom) {
    for (int a = blockIdx.x * kBlock + threadIdx.x; a < M; a += gridDim.x * kBlock) {
        const int o = orig[a];
        const double* zwa = zw + static_cast<size_t>(mw[a]) * nb;
        const double* zfa = zf + static_cast<size_t>(mf[a]) * nb;
        double pp = 0.0, mm = 0.0, p4 = 0.0, m4 = 0.0, pm = 0.0;
        for (int l = 0; l < nb; ++l) {
            const double Z = sq[a] * (zwa[l] + zfa[l]);
            const double e = jla_sign(seed, d0 + l, o) - Z;
            const double z2 = Z * Z, e2 = e * e;
            pp += z2;
            mm += e2;
            p4 += z2 * z2;
            m4 += e2 * e2;
            pm += z2 * e2;
        }
        pacc[o] += pp * inv_p;
        macc[o] += mm * inv_p;
        if (mom) {
            mom[o] += p4 * inv_p;
            mom[static_cast<size_t>(M) + o] += m4 * inv_p;
            mom[2 * static_cast<size_t>(M) + o] += pm * inv_p;
        }
    }
}

// pii = P / (P + M)
// ---- exact leverages (src/akm_kss.cpp:2667-2741): per match m, one lane:
// K [zw; zf] = [e_w; e_f] through the firm Schur complement

// the firm right-hand side e_f - B' Dw^-1 e_w of lane l (match m0 + l), and
// the lane's worker and firm
__global__ void __launch_bounds__(kBlock)
exact_rhs_kernel(int nb, int m0, const int* mlist, const int* rinv, const int* rw, const int* wofs, const int* r1,
                 const double* rcv, const double* idw, int L, double* b, int* lane_w, int* lane_f) {
    for (int l = blockIdx.x * kBlock + threadIdx.x; l < nb; l += gridDim.x * kBlock) {
        const int a = rinv[mlist ? mlist[l] : m0 + l];
        const int w = rw[a], f = r1[a];
        for (int a2 = wofs[w]; a2 < wofs[w + 1]; ++a2) b[static_cast<size_t>(r1[a2]) * L + l] -= rcv[a2] * idw[w];
        b[static_cast<size_t>(f) * L + l] += 1.0;
        lane_w[l] = w;
        lane_f[l] = f;
    }
}

// per block and lane: sum Df zf^2, sum Df zf over the firms (kBlock % L == 0)
__global__ void __launch_bounds__(kBlock)
exact_firm_kernel(const double* zf, const double* df, int J, int L, double* part) {
    __shared__ double r1s[kBlock], r2s[kBlock];
    const int l = threadIdx.x % L, rows = kBlock / L;
    double a1 = 0.0, a2 = 0.0;
    for (long long f = static_cast<long long>(blockIdx.x) * rows + threadIdx.x / L; f < J;
         f += static_cast<long long>(gridDim.x) * rows) {
        const double z = zf[static_cast<size_t>(f) * L + l];
        a1 += df[f] * z * z;
        a2 += df[f] * z;
    }
    r1s[threadIdx.x] = a1;
    r2s[threadIdx.x] = a2;
    __syncthreads();
    if (threadIdx.x < L) {
        double t1 = 0.0, t2 = 0.0;
        for (int k = 0; k < rows; ++k) {
            t1 += r1s[k * L + threadIdx.x];
            t2 += r2s[k * L + threadIdx.x];
        }
        part[(static_cast<size_t>(blockIdx.x) * 5 + 0) * L + threadIdx.x] = t1;
        part[(static_cast<size_t>(blockIdx.x) * 5 + 1) * L + threadIdx.x] = t2;
    }
}

// per block and lane over the workers: zw = (e_w - B zf) / Dw, then sum Dw zw^2,
// sum Dw zw and sum_m c_m zw[w_m] zf[f_m]; the lane's own P base zw[w] + zf[f]
__global__ void __launch_bounds__(kBlock)
exact_worker_kernel(const double* zf, const int* wofs, const int* r1, const double* rcv, const double* idw,
                    const double* dwv, const int* lane_w, const int* lane_f, int N, int L, int nb, double* part,
                    double* base_p) {
    __shared__ double s3[kBlock], s4[kBlock], sc[kBlock];
    const int l = threadIdx.x % L, rows = kBlock / L;
    double a3 = 0.0, a4 = 0.0, ac = 0.0;
    if (l < nb) {
        const int lw = lane_w[l];
        for (long long w = static_cast<long long>(blockIdx.x) * rows + threadIdx.x / L; w < N;
             w += static_cast<long long>(gridDim.x) * rows) {
            double bz = 0.0;
            for (int a = wofs[w]; a < wofs[w + 1]; ++a) bz += rcv[a] * zf[static_cast<size_t>(r1[a]) * L + l];
            const double zw = ((w == lw ? 1.0 : 0.0) - bz) * idw[w];
            a3 += dwv[w] * zw * zw;
            a4 += dwv[w] * zw;
            ac += zw * bz;
            if (w == lw) base_p[l] = zw + zf[static_cast<size_t>(lane_f[l]) * L + l];
        }
    }
    s3[threadIdx.x] = a3;
    s4[threadIdx.x] = a4;
    sc[threadIdx.x] = ac;
    __syncthreads();
    if (threadIdx.x < L) {
        double t3 = 0.0, t4 = 0.0, tc = 0.0;
        for (int k = 0; k < rows; ++k) {
            t3 += s3[k * L + threadIdx.x];
            t4 += s4[k * L + threadIdx.x];
            tc += sc[k * L + threadIdx.x];
        }
        part[(static_cast<size_t>(blockIdx.x) * 5 + 2) * L + threadIdx.x] = t3;
        part[(static_cast<size_t>(blockIdx.x) * 5 + 3) * L + threadIdx.x] = t4;
        part[(static_cast<size_t>(blockIdx.x) * 5 + 4) * L + threadIdx.x] = tc;
    }
}

__global__ void __launch_bounds__(kBlock) jla_ratio_kernel(int M, const double* pacc, const double* macc, double* pii) {
    for (int m = blockIdx.x * kBlock + threadIdx.x; m < M; m += gridDim.x * kBlock) {
        const double den = pacc[m] + macc[m];
        pii[m] = den > 0.0 ? pacc[m] / den : 0.0;
    }
}

// The non-linearity correction of the ratio estimator (src/akm_kss.cpp:
// 3273-3288): Ph = pii, Mh = 1 - Ph, V = (Mh^2 Psq + Ph^2 Msq - 2 Mh Ph PM) / p,
// B = (Mh Psq - Ph Msq + 2 (Mh - Ph) PM) / p, corr = 1 - V / Mh^2 + B / Mh
// (1 where Mh <= 0)
__global__ void __launch_bounds__(kBlock)
jla_corr_kernel(int M, const double* pii, const double* mom, double inv_p, double* corr) {
    for (int m = blockIdx.x * kBlock + threadIdx.x; m < M; m += gridDim.x * kBlock) {
        const double Ph = pii[m], Mh = 1.0 - Ph;
        double cr = 1.0;
        if (Mh > 0.0) {
            const double psq = mom[m], msq = mom[static_cast<size_t>(M) + m], pm = mom[2 * static_cast<size_t>(M) + m];
            const double V = (Mh * Mh * psq + Ph * Ph * msq - 2.0 * Mh * Ph * pm) * inv_p;
            const double B = (Mh * psq - Ph * msq + 2.0 * (Mh - Ph) * pm) * inv_p;
            cr = 1.0 - V / (Mh * Mh) + B / Mh;
        }
        corr[m] = cr;
    }
}

// ---- B_ii draws (src/akm_kss.cpp:2985-3047, :3128-3146): per draw, u_m is
// the sum of count[m] Rademacher draws of the stream (seed, 3 d + 1, m),
// scaled by 1 / sqrt(p) and demeaned over person years; the fe solve takes
// [0; F'u], the pe solve [W'u; 0].  A batch holds h draws: lanes [0, h) fe,
// [h, 2 h) pe.

constexpr int kBiiDraws = kMaxCgCols / 2;

// su[o h + l] = sum of the count[o] draws of row o in draw d0 + l; tot[l] +=
// their integer total (exact, order-free).  Thread slot l = threadIdx % 16.
__global__ void __launch_bounds__(kBlock)
jla_msum_kernel(int M, const long long* __restrict__ cnt, uint64_t seed, int d0, int h, double* __restrict__ su,
                unsigned long long* __restrict__ tot) {
    __shared__ long long red[kBlock];
    const int l = threadIdx.x % kBiiDraws;
    const int per = kBlock / kBiiDraws;
    long long part = 0;
    if (l < h) {
        const uint64_t s0 = jla_mix(seed ^ (0xA24BAED4963EE407ULL * (3ULL * static_cast<uint64_t>(d0 + l) + 2ULL)));
        for (int o = blockIdx.x * per + threadIdx.x / kBiiDraws; o < M; o += gridDim.x * per) {
            uint64_t st = jla_mix(s0 ^ (0x9FB21C651E98DF25ULL * (static_cast<uint64_t>(o) + 1ULL)));
            long long rem = cnt[o], acc = 0;
            while (rem > 0) {
                st = jla_mix(st);
                const int k = rem >= 64 ? 64 : static_cast<int>(rem);
                const uint64_t bits = k == 64 ? st : (st & ((1ULL << k) - 1ULL));
                acc += 2LL * __popcll(bits) - k;
                rem -= k;
            }
            su[static_cast<size_t>(o) * h + l] = static_cast<double>(acc);
            part += acc;
        }
    }
    red[threadIdx.x] = part;
    __syncthreads();
    if (threadIdx.x < h) {
        long long t = 0;
        for (int i = threadIdx.x; i < kBlock; i += kBiiDraws) t += red[i];
        atomicAdd(tot + threadIdx.x, static_cast<unsigned long long>(t));
    }
}

// u = msum / sqrt(p) - vbar count, vbar = (total / n_py) / sqrt(p)
__global__ void __launch_bounds__(kBlock)
jla_center_kernel(int M, const long long* cnt, int h, double scale, double n_py, const unsigned long long* tot,
                  double* su) {
    const long long total = static_cast<long long>(M) * h;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int o = static_cast<int>(j / h);
        const int l = static_cast<int>(j - static_cast<long long>(o) * h);
        const double vbar = (static_cast<double>(static_cast<long long>(tot[l])) / n_py) * scale;
        su[j] = su[j] * scale - vbar * static_cast<double>(cnt[o]);
    }
}

// Worker part of the B batch: tw = 0 on the fe lanes, W'u on the pe lanes
__global__ void __launch_bounds__(kBlock)
jla_tw_bii_kernel(const int* wofs, int N, const int* orig, const double* su, int h, double* tw) {
    const int nb = 2 * h;
    const long long total = static_cast<long long>(N) * nb;
    for (long long j = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kBlock) {
        const int w = static_cast<int>(j / nb);
        const int l = static_cast<int>(j - static_cast<long long>(w) * nb);
        double acc = 0.0;
        if (l >= h)
            for (int a = wofs[w]; a < wofs[w + 1]; ++a) acc += su[static_cast<size_t>(orig[a]) * h + (l - h)];
        tw[j] = acc;
    }
}

// Per working row: W2 = sqrt(c)(zw + zf) of the fe lane, W1 of the pe lane;
// Bfe += W2^2, Bpe += W1^2, Bcov += W1 W2 in ascending draw order
__global__ void __launch_bounds__(kBlock)
jla_bii_accumulate_kernel(int M, const int* mw, const int* mf, const int* orig, const double* sq, const double* zw,
                          const double* zf, int h, double* bfe, double* bpe, double* bcov) {
    const int nb = 2 * h;
    for (int a = blockIdx.x * kBlock + threadIdx.x; a < M; a += gridDim.x * kBlock) {
        const int o = orig[a];
        const double* zwa = zw + static_cast<size_t>(mw[a]) * nb;
        const double* zfa = zf + static_cast<size_t>(mf[a]) * nb;
        double fe = bfe[o], pe = bpe[o], co = bcov[o];
        for (int l = 0; l < h; ++l) {
            const double W2 = sq[a] * (zwa[l] + zfa[l]);
            const double W1 = sq[a] * (zwa[h + l] + zfa[h + l]);
            fe += W2 * W2;
            pe += W1 * W1;
            co += W1 * W2;
        }
        bfe[o] = fe;
        bpe[o] = pe;
        bcov[o] = co;
    }
}

constexpr int kAkmMaxGrid = 2048;

int akm_grid(long long work) {
    long long g = (work + kBlock - 1) / kBlock;
    return static_cast<int>(std::max(1LL, std::min<long long>(g, 2048)));
}

}  // namespace
}  // namespace xfe

using namespace xfe;

struct xfe_akm_context {
    int N = 0, J = 0, ground = 0, M = 0;
    hipStream_t s = nullptr;
    // JLA draws: worker offsets and per-record row index, sqrt(c), c in the
    // worker-major record order; row index per firm-major record; 1 / Dw
    DevBuf<int> wofs, rorig, forig;
    DevBuf<double> rsq, rcv, idw, jtw, pacc, macc, pii, jmom;
    // B_ii draws: person-year count per row, u per (row, draw), integer totals, sums
    DevBuf<long long> jcnt;
    DevBuf<double> jsu, bfe, bpe, bcov;
    DevBuf<unsigned long long> jtot;
    // worker-major matches
    int PW = 0;  // worker pieces
    DevBuf<int> wpiece, rw, r1;
    DevBuf<double> rwt, fc;
    // firm-major matches
    DevBuf<int> fkey, fw, has;
    DevBuf<double> inv_diag, corr;
    // workspace
    DevBuf<double> wmean, lanes, b, x, r, z, p, q, cg, part, scale;
    DevBuf<int> bk[2];
    DevBuf<double> bv[2];
    double* h_pack = nullptr;   // pinned J x lanes staging
    int h_pack_lanes = 0;
    double* h_small = nullptr;  // pinned scalars
    AkmWide wide;               // the 256-lane FP32 / FP64-replacement solver
    // exact leverages: the record of each match, Dw (new worker order) and Df,
    // per-lane worker / firm, partial sums
    DevBuf<int> rinv, lane_w, lane_f, lane_m;
    DevBuf<double> dwv, dfv, ex_part, ex_bp;
    double n_py = 0.0;

    ~xfe_akm_context() {
        wide.release();
        for (DevBuf<int>* d : {&wofs, &rorig, &forig, &wpiece, &rw, &r1, &fkey, &fw, &has, &bk[0], &bk[1], &rinv,
                               &lane_w, &lane_f, &lane_m})
            d->release();
        for (DevBuf<double>* d : {&dwv, &dfv, &ex_part, &ex_bp}) d->release();
        for (DevBuf<double>* d : {&rsq, &rcv, &idw, &jtw, &pacc, &macc, &pii, &jmom, &jsu, &bfe, &bpe, &bcov, &rwt, &fc,
                                  &inv_diag, &corr, &wmean, &lanes, &b, &x, &r, &z, &p, &q, &cg, &part, &scale,
                                  &bv[0], &bv[1]})
            d->release();
        jcnt.release();
        jtot.release();
        if (h_pack) (void)hipHostFree(h_pack);
        if (h_small) (void)hipHostFree(h_small);
        if (s) (void)hipStreamDestroy(s);
    }

    // out = S p for level-major J x nb vectors (ground pinned): the worker
    // means, the firm-major runs with the boundary list reduced level by level,
    // then the degree correction and the ground pin.
    void apply(const double* pv, double* out, int nb) {
        hipLaunchKernelGGL(akm_wmean_kernel, dim3(akm_grid(static_cast<long long>(PW) * kLaneGroup)), dim3(kBlock),
                           0, s, wpiece.p, PW, rw.p, r1.p, rwt.p, pv, nb, wmean.p);
        XFE_CHECK(hipGetLastError());
        int m = M;
        int pieces = (m + kPiece - 1) / kPiece;
        const int pslot = profile_begin(s);
        hipLaunchKernelGGL(akm_firm_kernel, dim3(akm_grid(static_cast<long long>(pieces) * kLaneGroup)),
                           dim3(kBlock), 0, s, fkey.p, fw.p, fc.p, pv, wmean.p, m, nb, out, bk[0].p, bv[0].p,
                           pieces == 1 ? 1 : 0);
        XFE_CHECK(hipGetLastError());
        // the firm pass alone: match records (firm key, worker, count), the worker's nb
        // lanes per match, the firm's lanes read and written once
        profile_end(pslot, s, static_cast<double>(M) * (12.0 + 8.0 * nb) + 16.0 * nb * J);
        int cur = 0;
        while (pieces > 1) {
            m = 2 * pieces;
            pieces = (m + kPiece - 1) / kPiece;
            hipLaunchKernelGGL(akm_list_kernel, dim3(akm_grid(static_cast<long long>(pieces) * kLaneGroup)),
                               dim3(kBlock), 0, s, bk[cur].p, bv[cur].p, m, nb, out, bk[1 - cur].p, bv[1 - cur].p,
                               pieces == 1 ? 1 : 0);
            XFE_CHECK(hipGetLastError());
            cur = 1 - cur;
        }
        hipLaunchKernelGGL(akm_finish, dim3(akm_grid(static_cast<long long>(J) * nb)), dim3(kBlock), 0, s, out, pv,
                           corr.p, has.p, J, nb, ground);
        XFE_CHECK(hipGetLastError());
    }

    void ensure_lanes(int nb) {
        const size_t JL = static_cast<size_t>(J) * nb;
        lanes.ensure(JL); b.ensure(JL); x.ensure(JL); r.ensure(JL); z.ensure(JL); p.ensure(JL); q.ensure(JL);
    }

    // One batch of draw solves from the worker part jtw (N x nb): the firm
    // part of the Schur right-hand sides (P_ii signs, or su on lanes below h),
    // reduced over the boundary list; the grounded solve; zf into lanes and
    // zw into wmean.  Returns true when some lane did not converge.
    bool jla_solve(uint64_t seed, int d0, int nb, const double* su, int h, double tol, int max_iter, int* its) {
        const long long NL = static_cast<long long>(N) * nb, JL = static_cast<long long>(J) * nb;
        XFE_CHECK(hipMemsetAsync(b.p, 0, sizeof(double) * JL, s));
        int m = M;
        int pieces = (m + kPiece - 1) / kPiece;
        hipLaunchKernelGGL(jla_firm_kernel, dim3(akm_grid(static_cast<long long>(pieces) * kLaneGroup)), dim3(kBlock),
                           0, s, fkey.p, fw.p, fc.p, forig.p, jtw.p, idw.p, seed, d0, m, nb, b.p, bk[0].p, bv[0].p,
                           pieces == 1 ? 1 : 0, su, h);
        XFE_CHECK(hipGetLastError());
        int cur = 0;
        while (pieces > 1) {
            m = 2 * pieces;
            pieces = (m + kPiece - 1) / kPiece;
            hipLaunchKernelGGL(akm_list_kernel, dim3(akm_grid(static_cast<long long>(pieces) * kLaneGroup)),
                               dim3(kBlock), 0, s, bk[cur].p, bv[cur].p, m, nb, b.p, bk[1 - cur].p, bv[1 - cur].p,
                               pieces == 1 ? 1 : 0);
            XFE_CHECK(hipGetLastError());
            cur = 1 - cur;
        }
        XFE_CHECK(hipMemsetAsync(b.p + static_cast<size_t>(ground) * nb, 0, sizeof(double) * nb, s));
        solve_levels(nb, tol, max_iter);
        lane_iterations(nb, its);
        bool failed = false;
        for (int l = 0; l < nb; ++l)
            if (its[l] < 0) failed = true;
        hipLaunchKernelGGL(jla_unscale_kernel, dim3(akm_grid(JL)), dim3(kBlock), 0, s, x.p, J, nb, ground,
                           scale.p + kMaxCgCols, cg.p, lanes.p);
        XFE_CHECK(hipGetLastError());
        hipLaunchKernelGGL(jla_zw_kernel, dim3(akm_grid(NL)), dim3(kBlock), 0, s, wofs.p, N, r1.p, rcv.p, idw.p,
                           jtw.p, lanes.p, nb, wmean.p);
        XFE_CHECK(hipGetLastError());
        return failed;
    }

    // Batched Jacobi-PCG (src/akm_kss_cuda.cu:532-695): per lane, the
    // right-hand side with its ground entry pinned is scaled to unit norm, so
    // the reference's relative stop r.r <= tol^2 |rhs|^2 becomes r.r <= tol^2;
    // a lane stops on convergence, on pAp outside (0, 1e300) or after max_iter
    // iterations.  Returns per-lane iteration counts (-1: not converged, the
    // lane's solution is zero).
    void solve(const double* rhs_pack, double* z_pack, int nb, double tol, int max_iter, int* iters_out) {
        ensure_lanes(nb);
        const size_t JL = static_cast<size_t>(J) * nb;
        XFE_CHECK(hipMemcpyAsync(lanes.p, rhs_pack, sizeof(double) * JL, hipMemcpyHostToDevice, s));
        const int gj = akm_grid(static_cast<long long>(JL));
        hipLaunchKernelGGL(akm_to_levels, dim3(gj), dim3(kBlock), 0, s, lanes.p, J, nb, ground, b.p);
        XFE_CHECK(hipGetLastError());
        solve_levels(nb, tol, max_iter);
        hipLaunchKernelGGL(akm_to_lanes, dim3(gj), dim3(kBlock), 0, s, x.p, J, nb, ground, scale.p + kMaxCgCols,
                           cg.p, lanes.p);
        XFE_CHECK(hipGetLastError());
        XFE_CHECK(hipMemcpyAsync(z_pack, lanes.p, sizeof(double) * JL, hipMemcpyDeviceToHost, s));
        lane_iterations(nb, iters_out);
    }

    // The batched PCG on the level-major right-hand sides already in b (ground
    // rows zero); leaves the scaled solutions in x, the scales in scale.
    void solve_levels(int nb, double tol, int max_iter) {
        const size_t JL = static_cast<size_t>(J) * nb;
        // per-lane norms of the pinned right-hand sides
        const int vg = akm_grid(J);
        part.ensure(std::max<size_t>(static_cast<size_t>(vg) * nb, 2 * static_cast<size_t>(kAkmMaxGrid) * kMaxCgCols));
        XFE_CHECK(launch_vdot<double>(b.p, b.p, J, nb, part.p, vg, s));
        XFE_CHECK(launch_reduce_partials(part.p, vg, nb, 0, scale.p + 2 * kMaxCgCols, s));
        XFE_CHECK(hipMemcpyAsync(h_small, scale.p + 2 * kMaxCgCols, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        double sc[2 * kMaxCgCols] = {};
        for (int l = 0; l < nb; ++l) {
            const double nrm = std::sqrt(h_small[l]);
            sc[l] = nrm > 0.0 ? 1.0 / nrm : 0.0;
            sc[kMaxCgCols + l] = nrm;
        }
        std::memcpy(h_small, sc, sizeof(sc));
        XFE_CHECK(hipMemcpyAsync(scale.p, h_small, sizeof(sc), hipMemcpyHostToDevice, s));
        const long long total = static_cast<long long>(JL);
        const int g = cg_grid(total, nb, kAkmMaxGrid);
        XFE_CHECK(launch_cg_init<double>(b.p, scale.p, inv_diag.p, x.p, r.p, z.p, p.p, total, nb, cg.p, part.p, g, s));
        constexpr int kPoll = 8;
        const double tol2 = tol * tol;
        for (int it = 1; it <= max_iter; ++it) {
            apply(p.p, q.p, nb);
            XFE_CHECK(launch_cg_dot<double>(p.p, q.p, total, nb, cg.p, part.p, g, s));
            XFE_CHECK(launch_cg_step<double>(x.p, r.p, z.p, p.p, q.p, inv_diag.p, total, nb, cg.p, part.p, tol2,
                                             max_iter, g, s));
            if (it % kPoll == 0 || it == max_iter) {
                XFE_CHECK(hipMemcpyAsync(h_small, cg.p + kCgCount, sizeof(double), hipMemcpyDeviceToHost, s));
                XFE_CHECK(hipStreamSynchronize(s));
                if (h_small[0] == 0.0) break;
            }
        }
    }

    // per-lane iteration counts (-1: not converged) of the last solve
    void lane_iterations(int nb, int* iters_out) {
        XFE_CHECK(hipMemcpyAsync(h_small, cg.p + kCgIters, sizeof(double) * 2 * kMaxCgCols, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        for (int l = 0; l < nb; ++l) {
            const bool ok = h_small[kMaxCgCols + l] != 0.0;
            iters_out[l] = ok ? static_cast<int>(h_small[l]) : -1;
        }
    }
};

extern "C" {

int xfe_akm_available(void) { return xfe_backend_available(); }

int xfe_akm_max_lanes(void) { return kMaxCgCols; }

int xfe_akm_stable_sort_u64(const uint64_t* keys, int64_t* order_out, size_t n) {
    clear_error();
    if (n == 0) return 1;
    if (!keys || !order_out || n > static_cast<size_t>(INT32_MAX)) {
        set_error("xfe_akm_stable_sort_u64: invalid arguments");
        return 0;
    }
    try {
        (void)hipGetLastError();
        hipStream_t s = nullptr;
        XFE_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        DevBuf<uint64_t> kin, kout;
        DevBuf<int64_t> vin, vout;
        DevBuf<unsigned char> tmp;
        kin.ensure(n); kout.ensure(n); vin.ensure(n); vout.ensure(n);
        XFE_CHECK(hipMemcpyAsync(kin.p, keys, sizeof(uint64_t) * n, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(akm_iota64, dim3(akm_grid(static_cast<long long>(n))), dim3(kBlock), 0, s, vin.p, n);
        XFE_CHECK(hipGetLastError());
        size_t bytes = 0;
        XFE_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin.p, kout.p, vin.p, vout.p, n, 0, 64, s));
        tmp.ensure(bytes);
        XFE_CHECK(rocprim::radix_sort_pairs(tmp.p, bytes, kin.p, kout.p, vin.p, vout.p, n, 0, 64, s));
        XFE_CHECK(hipMemcpyAsync(order_out, vout.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        kin.release(); kout.release(); vin.release(); vout.release(); tmp.release();
        (void)hipStreamDestroy(s);
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

xfe_akm_context* xfe_akm_create(int n_workers, int n_firms, const int* m_w, const int* m_f, const double* m_c,
                                size_t n_matches, const int64_t* w_ptr, const int64_t* f_ptr,
                                const int64_t* f_matches, const double* Dw, const double* Df,
                                const double* inv_diagS) {
    clear_error();
    if (n_workers <= 0 || n_firms <= 0 || n_matches == 0 || n_matches > static_cast<size_t>(INT32_MAX) || !m_w ||
        !m_f || !m_c || !w_ptr || !f_ptr || !f_matches || !Dw || !Df || !inv_diagS) {
        set_error("xfe_akm_create: invalid arguments");
        return nullptr;
    }
    const int N = n_workers, J = n_firms, M = static_cast<int>(n_matches);
    if (w_ptr[0] != 0 || w_ptr[N] != M || f_ptr[0] != 0 || f_ptr[J] != M) {
        set_error("xfe_akm_create: inconsistent match structure");
        return nullptr;
    }
    // host restatement of the structure into the device records
    std::vector<int> rw(M), fkey(M), fw(M), has(J, 0), wpiece;
    std::vector<double> rwt(M), fc(M), corr(J);
    for (int w = 0; w < N; ++w) {
        if (w_ptr[w + 1] < w_ptr[w]) {
            set_error("xfe_akm_create: w_ptr is not non-decreasing");
            return nullptr;
        }
        for (int64_t a = w_ptr[w]; a < w_ptr[w + 1]; ++a) {
            if (m_w[a] != w) {
                set_error("xfe_akm_create: matches are not grouped by worker as w_ptr says");
                return nullptr;
            }
        }
    }
    for (int a = 0; a < M; ++a) {
        const double c = m_c[a];
        if (!(c >= 1.0) || c != std::floor(c) || c > 2.0e9 || m_f[a] < 0 || m_f[a] >= J || m_w[a] < 0 ||
            m_w[a] >= N) {
            set_error("xfe_akm_create: match counts must be positive integers and ids in range");
            return nullptr;
        }
    }
    // Workers are renumbered by their primary firm (most person-years, then the
    // lowest firm id), so that the firm pass reads the worker means of a firm's
    // stayers as one contiguous block.  S is a firm-by-firm operator: the
    // numbering of workers does not change it.
    std::vector<int> primary(N, -1), order(N), newid(N);
    for (int w = 0; w < N; ++w) {
        double best = -1.0;
        for (int64_t a = w_ptr[w]; a < w_ptr[w + 1]; ++a)
            if (m_c[a] > best || (m_c[a] == best && m_f[a] < primary[w])) {
                best = m_c[a];
                primary[w] = m_f[a];
            }
    }
    for (int w = 0; w < N; ++w) order[w] = w;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return primary[a] < primary[b]; });
    for (int k = 0; k < N; ++k) newid[order[k]] = k;
    // worker-major records in the new order, and pieces of whole workers of at
    // most kPiece matches
    std::vector<int> rf(M), rorig(M), wofs(N + 1);
    std::vector<double> rsq(M), rcv(M), idw(N);
    wpiece.push_back(0);
    int pos = 0;
    for (int k = 0; k < N; ++k) {
        const int w = order[k];
        const int a0 = pos;
        const double iw = Dw[w] != 0.0 ? 1.0 / Dw[w] : 0.0;
        wofs[k] = pos;
        idw[k] = iw;
        for (int64_t a = w_ptr[w]; a < w_ptr[w + 1]; ++a, ++pos) {
            rw[pos] = k;
            rf[pos] = m_f[a];
            rwt[pos] = m_c[a] * iw;
            rorig[pos] = static_cast<int>(a);
            rsq[pos] = std::sqrt(m_c[a]);
            rcv[pos] = m_c[a];
        }
        if (pos > a0 && pos - wpiece.back() > kPiece && a0 > wpiece.back()) wpiece.push_back(a0);
    }
    wofs[N] = M;
    wpiece.push_back(M);
    // firm-major records, a firm's matches in the new worker order
    std::vector<double> deg(J, 0.0);
    std::vector<int> forig(M);
    struct FirmRec {
        int w;
        int m;
        double c;
        bool operator<(const FirmRec& o) const { return w < o.w || (w == o.w && m < o.m); }
    };
    std::vector<FirmRec> run;
    for (int f = 0; f < J; ++f) {
        run.clear();
        for (int64_t j = f_ptr[f]; j < f_ptr[f + 1]; ++j) {
            const int64_t m = f_matches[j];
            if (m < 0 || m >= M || m_f[m] != f) {
                set_error("xfe_akm_create: f_matches does not list the firm's matches");
                return nullptr;
            }
            run.push_back({newid[m_w[m]], static_cast<int>(m), m_c[m]});
            deg[f] += m_c[m];
        }
        std::sort(run.begin(), run.end());
        for (size_t i = 0; i < run.size(); ++i) {
            const int64_t j = f_ptr[f] + static_cast<int64_t>(i);
            fkey[j] = f;
            fw[j] = run[i].w;
            fc[j] = run[i].c;
            forig[j] = run[i].m;
        }
        has[f] = f_ptr[f + 1] > f_ptr[f] ? 1 : 0;
        corr[f] = Df[f] - deg[f];
    }
    xfe_akm_context* c = nullptr;
    try {
        (void)hipGetLastError();
        c = new xfe_akm_context();
        c->N = N; c->J = J; c->M = M; c->ground = J - 1;
        XFE_CHECK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
        auto up_i = [&](DevBuf<int>& d, const std::vector<int>& h) {
            d.ensure(h.size());
            XFE_CHECK(hipMemcpyAsync(d.p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, c->s));
        };
        auto up_d = [&](DevBuf<double>& d, const double* h, size_t n) {
            d.ensure(n);
            XFE_CHECK(hipMemcpyAsync(d.p, h, sizeof(double) * n, hipMemcpyHostToDevice, c->s));
        };
        c->PW = static_cast<int>(wpiece.size()) - 1;
        up_i(c->wpiece, wpiece);
        up_i(c->rw, rw);
        up_i(c->wofs, wofs);
        up_i(c->rorig, rorig);
        up_i(c->forig, forig);
        up_d(c->rsq, rsq.data(), M);
        up_d(c->rcv, rcv.data(), M);
        up_d(c->idw, idw.data(), N);
        up_d(c->rwt, rwt.data(), M);
        up_d(c->fc, fc.data(), M);
        up_i(c->r1, rf);
        up_i(c->fkey, fkey);
        up_i(c->fw, fw);
        up_i(c->has, has);
        up_d(c->inv_diag, inv_diagS, J);
        up_d(c->corr, corr.data(), J);
        {
            std::vector<int> rinv(M);
            std::vector<double> dwv(N);
            for (int a = 0; a < M; ++a) rinv[rorig[a]] = a;
            for (int k = 0; k < N; ++k) dwv[k] = Dw[order[k]];
            up_i(c->rinv, rinv);
            up_d(c->dwv, dwv.data(), N);
            up_d(c->dfv, Df, J);
            double npy = 0.0;
            for (int a = 0; a < M; ++a) npy += m_c[a];
            c->n_py = npy;
            XFE_CHECK(hipStreamSynchronize(c->s));
        }
        c->wmean.ensure(static_cast<size_t>(N) * kMaxCgCols);
        const size_t nb2 = 2 * ((static_cast<size_t>(M) + kPiece - 1) / kPiece);
        for (int k = 0; k < 2; ++k) {
            c->bk[k].ensure(nb2);
            c->bv[k].ensure(nb2 * kMaxCgCols);
        }
        c->cg.ensure(kCgSlots);
        XFE_CHECK(hipMemsetAsync(c->cg.p, 0, sizeof(double) * kCgSlots, c->s));
        c->part.ensure(2 * static_cast<size_t>(kAkmMaxGrid) * kMaxCgCols);
        c->scale.ensure(3 * kMaxCgCols);
        XFE_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c->h_small), 4 * kMaxCgCols * sizeof(double),
                                hipHostMallocDefault));
        std::vector<double> dw_new(N);
        for (int k = 0; k < N; ++k) dw_new[k] = Dw[order[k]];
        AkmWideInput wi;
        wi.N = N; wi.J = J; wi.M = M; wi.ground = J - 1;
        wi.wofs = wofs.data(); wi.rf = rf.data(); wi.rc = rcv.data(); wi.dw = dw_new.data();
        wi.corr = corr.data(); wi.inv_diag = inv_diagS;
        c->wide.build(wi, c->s, std::getenv("XFE_AKM_JACOBI") == nullptr);
        XFE_CHECK(hipStreamSynchronize(c->s));  // the host vectors above go out of scope
        return c;
    } catch (const std::exception& e) {
        set_error(e.what());
        delete c;
        return nullptr;
    }
}

void xfe_akm_destroy(xfe_akm_context* ctx) { delete ctx; }

double* xfe_akm_pack_buffer(xfe_akm_context* c, int nb) {
    clear_error();
    if (!c || nb < 1 || nb > kMaxCgCols) return nullptr;
    if (c->h_pack_lanes >= nb) return c->h_pack;
    if (c->h_pack) (void)hipHostFree(c->h_pack);
    c->h_pack = nullptr;
    c->h_pack_lanes = 0;
    if (hipHostMalloc(reinterpret_cast<void**>(&c->h_pack), sizeof(double) * static_cast<size_t>(c->J) * nb,
                      hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c->h_pack = nullptr;  // the caller uses its own buffer
        return nullptr;
    }
    c->h_pack_lanes = nb;
    return c->h_pack;
}

int xfe_akm_solve_S_multi(xfe_akm_context* c, const double* rhs_pack, double* z_pack, int nb, double tol,
                          int max_iter, int* iters_out) {
    clear_error();
    if (!c || !rhs_pack || !z_pack || !iters_out || nb < 1 || nb > kMaxCgCols || max_iter < 1) {
        set_error("xfe_akm_solve_S_multi: invalid arguments");
        return -1;
    }
    try {
        (void)hipGetLastError();
        c->solve(rhs_pack, z_pack, nb, tol, max_iter, iters_out);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

namespace {

// xfe_akm_jla_pii, and with corr the correction (xfe_akm_jla_leverages)
int jla_pii_impl(xfe_akm_context* c, int draws, uint64_t seed, double tol, int max_iter, double* pii,
                 double* corr, const char* name) {
    clear_error();
    if (!c || !pii || draws < 1 || max_iter < 1) {
        set_error(std::string(name) + ": invalid arguments");
        return -1;
    }
    try {
        (void)hipGetLastError();
        const int M = c->M, N = c->N, J = c->J;
        hipStream_t s = c->s;
        c->ensure_lanes(kMaxCgCols);
        c->jtw.ensure(static_cast<size_t>(N) * kMaxCgCols);
        c->pacc.ensure(M);
        c->macc.ensure(M);
        c->pii.ensure(M);
        XFE_CHECK(hipMemsetAsync(c->pacc.p, 0, sizeof(double) * M, s));
        XFE_CHECK(hipMemsetAsync(c->macc.p, 0, sizeof(double) * M, s));
        if (corr) {
            c->jmom.ensure(3 * static_cast<size_t>(M));
            XFE_CHECK(hipMemsetAsync(c->jmom.p, 0, sizeof(double) * 3 * M, s));
        }
        std::vector<int> its(kMaxCgCols);
        bool failed = false;
        for (int d0 = 0; d0 < draws; d0 += kMaxCgCols) {
            const int nb = std::min(kMaxCgCols, draws - d0);
            const long long NL = static_cast<long long>(N) * nb;
            // t = A' r: the worker part, then the firm part of the Schur right-hand side
            hipLaunchKernelGGL(jla_tw_kernel, dim3(akm_grid(NL)), dim3(kBlock), 0, s, c->wofs.p, N, c->rorig.p,
                               c->rsq.p, seed, d0, nb, c->jtw.p);
            XFE_CHECK(hipGetLastError());
            if (c->jla_solve(seed, d0, nb, nullptr, 0, tol, max_iter, its.data())) failed = true;
            hipLaunchKernelGGL(jla_accumulate_kernel, dim3(akm_grid(M)), dim3(kBlock), 0, s, M, c->rw.p, c->r1.p,
                               c->rorig.p, c->rsq.p, c->wmean.p, c->lanes.p, seed, d0, nb, 1.0 / draws, c->pacc.p,
                               c->macc.p, corr ? c->jmom.p : nullptr);
            XFE_CHECK(hipGetLastError());
        }
        hipLaunchKernelGGL(jla_ratio_kernel, dim3(akm_grid(M)), dim3(kBlock), 0, s, M, c->pacc.p, c->macc.p,
                           c->pii.p);
        XFE_CHECK(hipGetLastError());
        XFE_CHECK(hipMemcpyAsync(pii, c->pii.p, sizeof(double) * M, hipMemcpyDeviceToHost, s));
        if (corr) {
            // the correction over the moments, into pacc (no longer needed)
            hipLaunchKernelGGL(jla_corr_kernel, dim3(akm_grid(M)), dim3(kBlock), 0, s, M, c->pii.p, c->jmom.p,
                               1.0 / draws, c->pacc.p);
            XFE_CHECK(hipGetLastError());
            XFE_CHECK(hipMemcpyAsync(corr, c->pacc.p, sizeof(double) * M, hipMemcpyDeviceToHost, s));
        }
        XFE_CHECK(hipStreamSynchronize(s));
        if (failed) set_error(std::string(name) + ": some draw solves did not converge");
        return failed ? 1 : 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

}  // namespace

int xfe_akm_jla_pii(xfe_akm_context* c, int draws, uint64_t seed, double tol, int max_iter, double* pii) {
    return jla_pii_impl(c, draws, seed, tol, max_iter, pii, nullptr, "xfe_akm_jla_pii");
}

int xfe_akm_jla_leverages(xfe_akm_context* c, int draws, uint64_t seed, double tol, int max_iter, double* pii,
                          double* corr) {
    if (!corr) {
        clear_error();
        set_error("xfe_akm_jla_leverages: invalid arguments");
        return -1;
    }
    return jla_pii_impl(c, draws, seed, tol, max_iter, pii, corr, "xfe_akm_jla_leverages");
}

int xfe_akm_jla_bii(xfe_akm_context* c, int draws, uint64_t seed, double tol, int max_iter, const int64_t* py_count,
                    double n_py, double* bfe, double* bpe, double* bcov) {
    clear_error();
    if (!c || !py_count || !bfe || !bpe || !bcov || draws < 1 || max_iter < 1) {
        set_error("xfe_akm_jla_bii: invalid arguments");
        return -1;
    }
    try {
        (void)hipGetLastError();
        const int M = c->M, N = c->N;
        hipStream_t s = c->s;
        long long total = 0;
        for (int m = 0; m < M; ++m) {
            if (py_count[m] < 0) {
                set_error("xfe_akm_jla_bii: negative person-year count");
                return -1;
            }
            total += py_count[m];
        }
        if (!(n_py > 0.0)) n_py = static_cast<double>(total);
        c->ensure_lanes(kMaxCgCols);
        c->jtw.ensure(static_cast<size_t>(N) * kMaxCgCols);
        c->jcnt.ensure(M);
        c->jsu.ensure(static_cast<size_t>(M) * kBiiDraws);
        c->jtot.ensure(kBiiDraws);
        c->bfe.ensure(M);
        c->bpe.ensure(M);
        c->bcov.ensure(M);
        XFE_CHECK(hipMemcpyAsync(c->jcnt.p, py_count, sizeof(long long) * M, hipMemcpyHostToDevice, s));
        XFE_CHECK(hipMemsetAsync(c->bfe.p, 0, sizeof(double) * M, s));
        XFE_CHECK(hipMemsetAsync(c->bpe.p, 0, sizeof(double) * M, s));
        XFE_CHECK(hipMemsetAsync(c->bcov.p, 0, sizeof(double) * M, s));
        const double scale = 1.0 / std::sqrt(static_cast<double>(draws));
        std::vector<int> its(kMaxCgCols);
        bool failed = false;
        for (int d0 = 0; d0 < draws; d0 += kBiiDraws) {
            const int h = std::min(kBiiDraws, draws - d0), nb = 2 * h;
            XFE_CHECK(hipMemsetAsync(c->jtot.p, 0, sizeof(unsigned long long) * kBiiDraws, s));
            hipLaunchKernelGGL(jla_msum_kernel, dim3(akm_grid(static_cast<long long>(M) * kBiiDraws)), dim3(kBlock),
                               0, s, M, c->jcnt.p, seed, d0, h, c->jsu.p, c->jtot.p);
            XFE_CHECK(hipGetLastError());
            hipLaunchKernelGGL(jla_center_kernel, dim3(akm_grid(static_cast<long long>(M) * h)), dim3(kBlock), 0, s,
                               M, c->jcnt.p, h, scale, n_py, c->jtot.p, c->jsu.p);
            XFE_CHECK(hipGetLastError());
            hipLaunchKernelGGL(jla_tw_bii_kernel, dim3(akm_grid(static_cast<long long>(N) * nb)), dim3(kBlock), 0, s,
                               c->wofs.p, N, c->rorig.p, c->jsu.p, h, c->jtw.p);
            XFE_CHECK(hipGetLastError());
            if (c->jla_solve(seed, d0, nb, c->jsu.p, h, tol, max_iter, its.data())) failed = true;
            hipLaunchKernelGGL(jla_bii_accumulate_kernel, dim3(akm_grid(M)), dim3(kBlock), 0, s, M, c->rw.p, c->r1.p,
                               c->rorig.p, c->rsq.p, c->wmean.p, c->lanes.p, h, c->bfe.p, c->bpe.p, c->bcov.p);
            XFE_CHECK(hipGetLastError());
        }
        XFE_CHECK(hipMemcpyAsync(bfe, c->bfe.p, sizeof(double) * M, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipMemcpyAsync(bpe, c->bpe.p, sizeof(double) * M, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipMemcpyAsync(bcov, c->bcov.p, sizeof(double) * M, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        if (failed) set_error("xfe_akm_jla_bii: some draw solves did not converge");
        return failed ? 1 : 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int xfe_akm_wide_max_lanes(void) { return kWideMaxLanes; }

int xfe_akm_solve_S_wide(xfe_akm_context* c, const double* rhs_pack, double* z_pack, int nb, double tol,
                         int max_iter, int* iters_out) {
    clear_error();
    if (!c || !rhs_pack || !z_pack || !iters_out || nb < 1 || nb > kWideMaxLanes || max_iter < 1 || !(tol >= 0.0)) {
        set_error("xfe_akm_solve_S_wide: invalid arguments");
        return -1;
    }
    try {
        (void)hipGetLastError();
        c->wide.solve(rhs_pack, z_pack, nb, tol, max_iter, iters_out, false);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int xfe_akm_solve_S_wide_device(xfe_akm_context* c, const double* d_rhs_pack, double* d_z_pack, int nb, double tol,
                                int max_iter, int* iters_out) {
    clear_error();
    if (!c || !d_rhs_pack || !d_z_pack || !iters_out || nb < 1 || nb > kWideMaxLanes || max_iter < 1 ||
        !(tol >= 0.0)) {
        set_error("xfe_akm_solve_S_wide_device: invalid arguments");
        return -1;
    }
    try {
        (void)hipGetLastError();
        // the caller may have filled d_rhs_pack on another stream
        XFE_CHECK(hipDeviceSynchronize());
        c->wide.solve(d_rhs_pack, d_z_pack, nb, tol, max_iter, iters_out, true);
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

static thread_local int last_exact_retries = 0;

int xfe_akm_exact_leverages_last_retries(void) { return last_exact_retries; }

int xfe_akm_exact_leverages(xfe_akm_context* c, double tol, double accept, int max_iter, double* pii, double* bfe,
                            double* bpe, double* bcov) {
    clear_error();
    if (!c || !pii || !bfe || !bpe || !bcov || max_iter < 1 || !(tol >= 0.0) || !(accept >= 0.0)) {
        set_error("xfe_akm_exact_leverages: invalid arguments");
        return -1;
    }
    try {
        (void)hipGetLastError();
        const int M = c->M, N = c->N, J = c->J;
        hipStream_t s = c->s;
        AkmWide& wd = c->wide;
        std::vector<double> h, bp;
        std::vector<double> mc(M);
        // the match weights in the original order (for the match-level scaling)
        {
            std::vector<int> ri(M);
            std::vector<double> rc(M);
            XFE_CHECK(hipMemcpyAsync(ri.data(), c->rinv.p, sizeof(int) * M, hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipMemcpyAsync(rc.data(), c->rcv.p, sizeof(double) * M, hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
            for (int m = 0; m < M; ++m) mc[m] = rc[ri[m]];
        }
        // the sums over firms and workers of the solutions zf (J x L) of nb
        // lanes (lane l: match ms[l]) into the outputs
        auto finish = [&](const double* zf, int nb, int L, const std::vector<int>& ms) {
            const int rows = kBlock / L;
            const int gf = std::max(1, std::min(1024, (J + rows - 1) / rows));
            const int gw = std::max(1, std::min(1024, (N + rows - 1) / rows));
            const int g = std::max(gf, gw);
            c->ex_part.ensure(static_cast<size_t>(g) * 5 * L);
            XFE_CHECK(hipMemsetAsync(c->ex_part.p, 0, sizeof(double) * static_cast<size_t>(g) * 5 * L, s));
            hipLaunchKernelGGL(exact_firm_kernel, dim3(gf), dim3(kBlock), 0, s, zf, c->dfv.p, J, L, c->ex_part.p);
            hipLaunchKernelGGL(exact_worker_kernel, dim3(gw), dim3(kBlock), 0, s, zf, c->wofs.p, c->r1.p, c->rcv.p,
                               c->idw.p, c->dwv.p, c->lane_w.p, c->lane_f.p, N, L, nb, c->ex_part.p, c->ex_bp.p);
            XFE_CHECK(hipGetLastError());
            h.resize(static_cast<size_t>(g) * 5 * L);
            bp.resize(L);
            XFE_CHECK(hipMemcpyAsync(h.data(), c->ex_part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipMemcpyAsync(bp.data(), c->ex_bp.p, sizeof(double) * L, hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
            for (int l = 0; l < nb; ++l) {
                double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                for (int b = 0; b < g; ++b)
                    for (int k = 0; k < 5; ++k) t[k] += h[(static_cast<size_t>(b) * 5 + k) * L + l];
                const double npy = c->n_py;
                const int m = ms[l];
                const double cm = mc[m];
                pii[m] = cm * bp[l];
                bfe[m] = cm * (t[0] - t[1] * t[1] / npy);
                bpe[m] = cm * (t[2] - t[3] * t[3] / npy);
                bcov[m] = cm * (t[4] - t[3] * t[1] / npy);
            }
        };
        // 256 matches per wide solve (FP32 passes, FP64 residual replacement)
        std::vector<int> fails, ms;
        for (int m0 = 0; m0 < M; m0 += kWideMaxLanes) {
            const int nb = std::min(kWideMaxLanes, M - m0);
            wd.ensure(nb);
            const int L = wd.L;
            c->lane_w.ensure(kWideMaxLanes);
            c->lane_f.ensure(kWideMaxLanes);
            c->ex_bp.ensure(kWideMaxLanes);
            XFE_CHECK(hipMemsetAsync(wd.b64.p, 0, sizeof(double) * static_cast<size_t>(J) * L, s));
            hipLaunchKernelGGL(exact_rhs_kernel, dim3((nb + kBlock - 1) / kBlock), dim3(kBlock), 0, s, nb, m0,
                               static_cast<const int*>(nullptr), c->rinv.p, c->rw.p, c->wofs.p, c->r1.p, c->rcv.p,
                               c->idw.p, L, wd.b64.p, c->lane_w.p, c->lane_f.p);
            XFE_CHECK(hipGetLastError());
            XFE_CHECK(hipMemsetAsync(wd.b64.p + static_cast<size_t>(c->ground) * L, 0, sizeof(double) * L, s));
            wd.solve_device(nb, tol, max_iter, accept);
            wd.unscale_x64();
            ms.resize(nb);
            for (int l = 0; l < nb; ++l) {
                ms[l] = m0 + l;
                if (wd.lane_iterations(l) < 0) fails.push_back(m0 + l);
            }
            finish(wd.x64.p, nb, L, ms);
        }
        // the matches whose wide solve did not converge again, 32 at a time,
        // with the FP64 batched Jacobi-PCG
        int failed = 0;
        constexpr int kL = kMaxCgCols;
        c->lane_m.ensure(kL);
        for (size_t k0 = 0; k0 < fails.size(); k0 += kL) {
            const int nb = static_cast<int>(std::min<size_t>(kL, fails.size() - k0));
            ms.assign(fails.begin() + static_cast<long>(k0), fails.begin() + static_cast<long>(k0) + nb);
            XFE_CHECK(hipMemcpyAsync(c->lane_m.p, ms.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
            c->ensure_lanes(kL);
            XFE_CHECK(hipMemsetAsync(c->b.p, 0, sizeof(double) * static_cast<size_t>(J) * kL, s));
            hipLaunchKernelGGL(exact_rhs_kernel, dim3(1), dim3(kBlock), 0, s, nb, 0,
                               static_cast<const int*>(c->lane_m.p), c->rinv.p, c->rw.p, c->wofs.p, c->r1.p,
                               c->rcv.p, c->idw.p, kL, c->b.p, c->lane_w.p, c->lane_f.p);
            XFE_CHECK(hipGetLastError());
            XFE_CHECK(hipMemsetAsync(c->b.p + static_cast<size_t>(c->ground) * kL, 0, sizeof(double) * kL, s));
            c->solve_levels(kL, tol, max_iter);
            int its[kL];
            c->lane_iterations(kL, its);
            for (int l = 0; l < nb; ++l)
                if (its[l] < 0) ++failed;
            const long long JL = static_cast<long long>(J) * kL;
            hipLaunchKernelGGL(jla_unscale_kernel, dim3(akm_grid(JL)), dim3(kBlock), 0, s, c->x.p, J, kL, c->ground,
                               c->scale.p + kMaxCgCols, c->cg.p, c->lanes.p);
            XFE_CHECK(hipGetLastError());
            finish(c->lanes.p, nb, kL, ms);
        }
        last_exact_retries = static_cast<int>(fails.size());
        return failed;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -1;
    }
}

int xfe_akm_wide_last(xfe_akm_context* c, int* fp32_passes, int* fp64_replacements) {
    clear_error();
    if (!c || !fp32_passes || !fp64_replacements) {
        set_error("xfe_akm_wide_last: invalid arguments");
        return -1;
    }
    *fp32_passes = c->wide.last_passes;
    *fp64_replacements = c->wide.last_updates;
    return 0;
}

int xfe_akm_solve_S(xfe_akm_context* c, const double* rhs, double* z, double tol, int max_iter) {
    int it = -1;
    if (xfe_akm_solve_S_multi(c, rhs, z, 1, tol, max_iter, &it) != 0) return -1;
    return it;
}

}  // extern "C"
