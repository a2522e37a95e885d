// The learned preconditioner: the graph the network sees (firms, movers and
// their edges, input features), the pair scores that drive the aggregation,
// the probe that records the absorber's own iterates for training, and the
// training loop (forward, backward, Adam) on the device.
//
// The network (lp_gnn.hip) is this build's design; the reference has no learned
// component (its preconditioner for the FE normal equations is the randomized
// approximate Cholesky of src/schwarz_demean.cpp:100-133).  Its scores only
// choose which firms share a block of the aggregate-block preconditioner
// (lp_precond.hip); the solve's stopping rule is the FP64 residual, so the
// answer never depends on the network.
//
// Features are computed from integer counts and values summed in a fixed
// order (stable sorts, one wavefront per firm), so the same panel gives the
// same scores and the same aggregation on every run.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <vector>

#include "absorb_kernels.h"
#include "common.h"
#include "lp_gnn.h"
#include "lp_precond.h"
#include "xfe_hip.h"

namespace xfe {
namespace lp {

namespace {

constexpr int kB = 256;
constexpr int kMaxSeg = 64;  // segments of one worker the mover kernels look at (longer lists: the first 64)

int grid_of(long long work, int per_block = kB, int cap = 8192) {
    const long long g = (work + per_block - 1) / per_block;
    return static_cast<int>(std::max(1LL, std::min(g, static_cast<long long>(cap))));
}

int bits_for(unsigned long long v) {
    int b = 1;
    while (b < 64 && (1ULL << b) <= v) ++b;
    return b;
}

template <typename T>
void exclusive_scan(DevBuf<unsigned char>& tmp, const T* in, T* out, size_t m, hipStream_t s) {
    size_t bytes = 0;
    XFE_CHECK(rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), m, rocprim::plus<T>(), s));
    tmp.ensure(bytes);
    XFE_CHECK(rocprim::exclusive_scan(tmp.p, bytes, in, out, T(0), m, rocprim::plus<T>(), s));
}

// distinct firms among a worker's segments (at most kMaxSeg segments looked at)
__device__ __forceinline__ int distinct_firms(const int* r1, int b, int e) {
    int d = 0;
    for (int i = b; i < e; ++i) {
        bool seen = false;
        for (int j = b; j < i; ++j) seen |= r1[j] == r1[i];
        d += seen ? 0 : 1;
    }
    return d;
}

__global__ void __launch_bounds__(kB) mover_count_kernel(const int* wseg, const int* r1, int Nw, int* mcnt,
                                                         int* mflag) {
    for (int w = blockIdx.x * kB + threadIdx.x; w < Nw; w += gridDim.x * kB) {
        const int b = wseg[w];
        const int e = min(wseg[w + 1], b + kMaxSeg);
        const int d = e - b >= 2 ? distinct_firms(r1, b, e) : 1;
        mcnt[w] = d >= 2 ? d : 0;
        mflag[w] = d >= 2 ? 1 : 0;
    }
}

// A mover's firms in ascending order (its edges), the edge's mover index, and
// its 8 features: log n_w, log segments, log distinct firms, max and min share
// of n_w at one firm, the sum of squared shares, log(1 + returns), 1.
__global__ void __launch_bounds__(kB) mover_fill_kernel(const int* wseg, const int* r1, const int* len,
                                                        const double* inv_w, int Nw, const int* mflag,
                                                        const int* moff, const int* midx, int* w_ptr, int* w_idx,
                                                        int* e_mov, float* xw) {
    for (int w = blockIdx.x * kB + threadIdx.x; w < Nw; w += gridDim.x * kB) {
        if (!mflag[w]) continue;
        const int b = wseg[w];
        const int e = min(wseg[w + 1], b + kMaxSeg);
        const int m = midx[w];
        const int o = moff[w];
        w_ptr[m] = o;
        const double iw = inv_w[w];
        int last = -1, d = 0;
        double smax = 0.0, smin = 2.0, ssq = 0.0;
        for (;;) {
            int nxt = INT_MAX;
            for (int i = b; i < e; ++i)
                if (r1[i] > last && r1[i] < nxt) nxt = r1[i];
            if (nxt == INT_MAX) break;
            double L = 0.0;
            for (int i = b; i < e; ++i)
                if (r1[i] == nxt) L += len[i];
            const double sh = L * iw;
            smax = fmax(smax, sh);
            smin = fmin(smin, sh);
            ssq += sh * sh;
            w_idx[o + d] = nxt;
            e_mov[o + d] = m;
            ++d;
            last = nxt;
        }
        float* x = xw + static_cast<size_t>(m) * kInW;
        x[0] = static_cast<float>(-log(iw));
        x[1] = static_cast<float>(log(static_cast<double>(wseg[w + 1] - wseg[w])));
        x[2] = static_cast<float>(log(static_cast<double>(d)));
        x[3] = static_cast<float>(smax);
        x[4] = static_cast<float>(smin);
        x[5] = static_cast<float>(ssq);
        x[6] = static_cast<float>(log1p(static_cast<double>(wseg[w + 1] - wseg[w] - d)));
        x[7] = 1.0f;
    }
}

__global__ void __launch_bounds__(kB) iota_kernel(int* a, int m) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < m; i += gridDim.x * kB) a[i] = i;
}

// firm-major position i of the edge e = sidx[i]: its mover, and its rank in the firm's list
__global__ void __launch_bounds__(kB) transpose_kernel(const int* skey, const int* sidx, const int* e_mov,
                                                       const int* f_ptr, int E, int* f_idx, int* w_rank) {
    for (int i = blockIdx.x * kB + threadIdx.x; i < E; i += gridDim.x * kB) {
        const int e = sidx[i];
        f_idx[i] = e_mov[e];
        w_rank[e] = i - f_ptr[skey[i]];
    }
}

// the two half-edges of every pair: (f, strength) and (g, strength)
__global__ void __launch_bounds__(kB) half_edge_kernel(const int* pf, const int* pg, const float* st, int U, int* key,
                                                       float* val) {
    for (int p = blockIdx.x * kB + threadIdx.x; p < U; p += gridDim.x * kB) {
        const float v = fmaxf(st[p], 0.0f);
        key[2 * p] = pf[p];
        key[2 * p + 1] = pg[p];
        val[2 * p] = v;
        val[2 * p + 1] = v;
    }
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// firm features, one wavefront per firm (its pair half-edges strided over the
// lanes, then wave reductions in a fixed order; lanes 0-7 then form the eight
// features and store them together): log(1 + rows), log(1 +
// movers), log(1 + S_ff), S_ff / rows, log(1 + pairs), max strength, sum of
// strengths, 1
__global__ void __launch_bounds__(kB) firm_feat_kernel(const double* rows, const int* f_ptr, const double* dex,
                                                       const int* hoff, const float* hval, int F, float* xf) {
    const int lane = threadIdx.x & 63;
    for (int f = blockIdx.x * 4 + threadIdx.x / 64; f < F; f += gridDim.x * 4) {
        const int b = hoff[f], e = hoff[f + 1];
        float mx = 0.0f, sm = 0.0f;
        for (int q = b + lane; q < e; q += 64) {
            const float v = hval[q];
            mx = fmaxf(mx, v);
            sm += v;
        }
        mx = wave_max(mx);
        sm = wave_sum(sm);
        mx = __shfl(mx, 0, 64);
        sm = __shfl(sm, 0, 64);
        // lane k < 8 forms feature k: the four logarithms run side by side
        if (lane < kInF) {
            const double r = rows[f];
            const double d = fmax(dex[f], 0.0);
            const double arg = lane == 0 ? r
                               : lane == 1 ? static_cast<double>(f_ptr[f + 1] - f_ptr[f])
                               : lane == 2 ? d
                                           : static_cast<double>(e - b);
            float v = static_cast<float>(log1p(arg));
            if (lane == 3) v = static_cast<float>(d / fmax(r, 1.0));
            else if (lane == 5) v = mx;
            else if (lane == 6) v = sm;
            else if (lane == 7) v = 1.0f;
            xf[static_cast<size_t>(f) * kInF + lane] = v;
        }
    }
}

__global__ void __launch_bounds__(kB) probe_iterate_kernel(const double* beta, const float* x, const double* t,
                                                           long long total, int C, double* out) {
    for (long long j = blockIdx.x * static_cast<long long>(kB) + threadIdx.x; j < total;
         j += static_cast<long long>(gridDim.x) * kB)
        out[j] = beta[j] + static_cast<double>(x[j]) * t[j % C];
}

struct LearnedState {
    LearnedGraph graph;
    Gnn net;
    DevBuf<float> params, score;
    void release() {
        graph.release();
        net.release();
        params.release();
        score.release();
    }
};

thread_local LearnedState tls_learned;
thread_local ProbeState tls_probe;

}  // namespace

ProbeState& probe_state() { return tls_probe; }

hipError_t launch_probe_iterate(const double* beta, const float* x, const double* t, int N, int C, double* out,
                                hipStream_t s) {
    const long long total = static_cast<long long>(N) * C;
    hipLaunchKernelGGL(probe_iterate_kernel, dim3(grid_of(total)), dim3(kB), 0, s, beta, x, t, total, C, out);
    return hipGetLastError();
}

void LearnedGraph::build(const AggPrecond& ag, const SegView& sv, hipStream_t s) {
    F = sv.G1;
    const int Nw = sv.Nw;
    if (!host_int) XFE_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_int), 16 * sizeof(int)));
    mcnt.ensure(Nw); mflag.ensure(Nw); moff.ensure(Nw); midx.ensure(Nw);
    const int gw = grid_of(Nw);
    hipLaunchKernelGGL(mover_count_kernel, dim3(gw), dim3(kB), 0, s, sv.wseg, sv.r1, Nw, mcnt.p, mflag.p);
    exclusive_scan<int>(tmp, mcnt.p, moff.p, static_cast<size_t>(Nw), s);
    exclusive_scan<int>(tmp, mflag.p, midx.p, static_cast<size_t>(Nw), s);
    XFE_CHECK(hipMemcpyAsync(host_int, moff.p + (Nw - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipMemcpyAsync(host_int + 1, mcnt.p + (Nw - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipMemcpyAsync(host_int + 2, midx.p + (Nw - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipMemcpyAsync(host_int + 3, mflag.p + (Nw - 1), sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    E = host_int[0] + host_int[1];
    Wm = host_int[2] + host_int[3];
    w_ptr.ensure(static_cast<size_t>(Wm) + 1);
    w_idx.ensure(E); e_mov.ensure(E); f_idx.ensure(E); skey.ensure(E); eidx.ensure(E); sidx.ensure(E);
    w_rank.ensure(E);
    xw.ensure(static_cast<size_t>(std::max(Wm, 1)) * kInW);
    xf.ensure(static_cast<size_t>(F) * kInF);
    f_ptr.ensure(static_cast<size_t>(F) + 1);
    hipLaunchKernelGGL(mover_fill_kernel, dim3(gw), dim3(kB), 0, s, sv.wseg, sv.r1, sv.len, sv.inv_w, Nw, mflag.p,
                       moff.p, midx.p, w_ptr.p, w_idx.p, e_mov.p, xw.p);
    XFE_CHECK(hipMemcpyAsync(w_ptr.p + Wm, &E, sizeof(int), hipMemcpyHostToDevice, s));
    // firm -> mover lists: a stable sort of the edges by firm (movers ascending
    // inside a firm), the firm offsets from the sorted keys, and each edge's
    // rank in its firm's list
    const int fb = bits_for(static_cast<unsigned long long>(F));
    if (E > 0) {
        hipLaunchKernelGGL(iota_kernel, dim3(grid_of(E)), dim3(kB), 0, s, eidx.p, E);
        size_t bytes = 0;
        XFE_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, w_idx.p, skey.p, eidx.p, sidx.p,
                                            static_cast<size_t>(E), 0, fb, s));
        tmp.ensure(bytes);
        XFE_CHECK(rocprim::radix_sort_pairs(tmp.p, bytes, w_idx.p, skey.p, eidx.p, sidx.p,
                                            static_cast<size_t>(E), 0, fb, s));
        XFE_CHECK(launch_seg_offsets(skey.p, E, F, f_ptr.p, grid_of(E), s));
        hipLaunchKernelGGL(transpose_kernel, dim3(grid_of(E)), dim3(kB), 0, s, skey.p, sidx.p, e_mov.p, f_ptr.p, E,
                           f_idx.p, w_rank.p);
    } else {
        XFE_CHECK(hipMemsetAsync(f_ptr.p, 0, sizeof(int) * (static_cast<size_t>(F) + 1), s));
    }
    // the pairs' half-edges by firm (stable), for the per-firm pair features
    const int H = 2 * ag.U;
    hkey.ensure(H); hkey2.ensure(H); hval.ensure(H); hval2.ensure(H); hoff.ensure(static_cast<size_t>(F) + 1);
    if (H > 0) {
        hipLaunchKernelGGL(half_edge_kernel, dim3(grid_of(ag.U)), dim3(kB), 0, s, ag.pf.p, ag.pg.p, ag.st.p, ag.U,
                           hkey.p, hval.p);
        size_t bytes = 0;
        XFE_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, hkey.p, hkey2.p, hval.p, hval2.p, static_cast<size_t>(H),
                                            0, fb, s));
        tmp.ensure(bytes);
        XFE_CHECK(rocprim::radix_sort_pairs(tmp.p, bytes, hkey.p, hkey2.p, hval.p, hval2.p, static_cast<size_t>(H),
                                            0, fb, s));
        XFE_CHECK(launch_seg_offsets(hkey2.p, H, F, hoff.p, grid_of(H), s));
    } else {
        XFE_CHECK(hipMemsetAsync(hoff.p, 0, sizeof(int) * (static_cast<size_t>(F) + 1), s));
    }
    hipLaunchKernelGGL(firm_feat_kernel, dim3(grid_of(F, 4)), dim3(kB), 0, s, sv.rows, f_ptr.p, ag.dex.p, hoff.p,
                       hval2.p, F, xf.p);
    XFE_CHECK(hipGetLastError());
}

void LearnedGraph::release() {
    for (auto* b : {&mcnt, &mflag, &moff, &midx, &w_ptr, &w_idx, &e_mov, &f_ptr, &f_idx, &skey, &eidx, &sidx, &w_rank,
                    &hkey, &hkey2, &hoff})
        b->release();
    hval.release(); hval2.release(); xf.release(); xw.release(); tmp.release();
    if (host_int) (void)hipHostFree(host_int);
    host_int = nullptr;
}

const float* learned_pair_scores(AggPrecond& ag, const SegView& sv, const float* params, hipStream_t s) {
    if (!params) throw Error("learned preconditioner: no network parameters (xfe_options.lp_params)");
    LearnedState& L = tls_learned;
    L.graph.build(ag, sv, s);
    L.params.ensure(kNumParams);
    XFE_CHECK(hipMemcpyAsync(L.params.p, params, sizeof(float) * kNumParams, hipMemcpyHostToDevice, s));
    const LearnedGraph& g = L.graph;
    Graph gv{g.xf.p, g.xw.p, g.f_ptr.p, g.f_idx.p, g.w_ptr.p, g.w_idx.p, ag.pf.p, ag.pg.p, ag.st.p, g.w_rank.p};
    L.net.ensure(g.F, g.Wm, ag.U, false);
    L.net.forward(gv, L.params.p, s);
    return L.net.t.p;
}

}  // namespace lp
}  // namespace xfe

using namespace xfe;

extern "C" {

int xfe_lp_probe_fetch(float* xf, float* xw, int* f_ptr, int* f_idx, int* w_ptr, int* w_idx, int* pf, int* pg,
                       float* pq, double* W, double* dex, double* err) {
    clear_error();
    lp::ProbeState& P = lp::probe_state();
    if (!P.done) {
        set_error("xfe_lp_probe_fetch: no recorded panel (call xfe_lp_probe first)");
        return 0;
    }
    auto put = [](auto* dst, const auto& v) {
        if (dst && !v.empty()) std::copy(v.begin(), v.end(), dst);
    };
    put(xf, P.xf); put(xw, P.xw); put(f_ptr, P.f_ptr); put(f_idx, P.f_idx); put(w_ptr, P.w_ptr);
    put(w_idx, P.w_idx); put(pf, P.pf); put(pg, P.pg); put(pq, P.pq); put(W, P.W); put(dex, P.dex); put(err, P.err);
    return 1;
}

int xfe_lp_train(int npanels, const xfe_lp_graph* panels, int steps, float lr, float* theta, double* loss_hist) {
    clear_error();
    if (npanels <= 0 || !panels || steps < 0 || !theta || !(lr > 0.0f)) {
        set_error("xfe_lp_train: invalid arguments");
        return 0;
    }
    for (int k = 0; k < npanels; ++k) {
        const xfe_lp_graph& g = panels[k];
        if (g.F <= 0 || g.Wm < 0 || g.U < 0 || !g.xf || !g.f_ptr || !g.u_tgt || !g.u_mask ||
            (g.Wm > 0 && (!g.xw || !g.w_ptr)) || (g.U > 0 && (!g.pf || !g.pg || !g.y_tgt))) {
            set_error("xfe_lp_train: invalid panel");
            return 0;
        }
        const int E = g.f_ptr[g.F];
        if (E != (g.Wm > 0 ? g.w_ptr[g.Wm] : 0) || (E > 0 && (!g.f_idx || !g.w_idx))) {
            set_error("xfe_lp_train: the two edge lists differ in length");
            return 0;
        }
        for (int e = 0; e < E; ++e)
            if (g.f_idx[e] < 0 || g.f_idx[e] >= g.Wm || g.w_idx[e] < 0 || g.w_idx[e] >= g.F) {
                set_error("xfe_lp_train: edge endpoint out of range");
                return 0;
            }
        for (int p = 0; p < g.U; ++p)
            if (g.pf[p] < 0 || g.pf[p] >= g.F || g.pg[p] < 0 || g.pg[p] >= g.F) {
                set_error("xfe_lp_train: pair endpoint out of range");
                return 0;
            }
    }
    try {
        (void)hipGetLastError();
        hipStream_t s = nullptr;
        XFE_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{s};
        struct Dev {
            DevBuf<float> xf, xw, pq, ut, um, y;
            DevBuf<int> fp, fi, wp, wi, pf, pg, rk;
            double nm = 0.0;
        };
        std::vector<Dev> dev(npanels);
        auto up = [&](auto& buf, const auto* src, size_t n) {
            buf.ensure(n);
            if (n && src) XFE_CHECK(hipMemcpyAsync(buf.p, src, sizeof(*src) * n, hipMemcpyHostToDevice, s));
        };
        int maxF = 1, maxW = 1, maxU = 1;
        for (int k = 0; k < npanels; ++k) {
            const xfe_lp_graph& g = panels[k];
            Dev& d = dev[k];
            const int E = g.f_ptr[g.F];
            up(d.xf, g.xf, static_cast<size_t>(g.F) * lp::kInF);
            up(d.xw, g.xw, static_cast<size_t>(g.Wm) * lp::kInW);
            up(d.fp, g.f_ptr, static_cast<size_t>(g.F) + 1);
            up(d.fi, g.f_idx, static_cast<size_t>(E));
            up(d.wp, g.w_ptr, static_cast<size_t>(g.Wm) + 1);
            up(d.wi, g.w_idx, static_cast<size_t>(E));
            const std::vector<int> rk = lp::host_w_rank(g.F, g.Wm, g.f_ptr, g.f_idx, g.w_ptr, g.w_idx);
            up(d.rk, rk.data(), static_cast<size_t>(E));
            up(d.pf, g.pf, static_cast<size_t>(g.U));
            up(d.pg, g.pg, static_cast<size_t>(g.U));
            if (g.pq) up(d.pq, g.pq, static_cast<size_t>(g.U));
            up(d.ut, g.u_tgt, static_cast<size_t>(g.F));
            up(d.um, g.u_mask, static_cast<size_t>(g.F));
            up(d.y, g.y_tgt, static_cast<size_t>(g.U));
            for (int f = 0; f < g.F; ++f) d.nm += g.u_mask[f];
            maxF = std::max(maxF, g.F);
            maxW = std::max(maxW, g.Wm);
            maxU = std::max(maxU, g.U);
        }
        DevBuf<float> th, m, v, grad;
        th.ensure(lp::kNumParams); m.ensure(lp::kNumParams); v.ensure(lp::kNumParams); grad.ensure(lp::kNumParams);
        XFE_CHECK(hipMemcpyAsync(th.p, theta, sizeof(float) * lp::kNumParams, hipMemcpyHostToDevice, s));
        XFE_CHECK(hipMemsetAsync(m.p, 0, sizeof(float) * lp::kNumParams, s));
        XFE_CHECK(hipMemsetAsync(v.p, 0, sizeof(float) * lp::kNumParams, s));
        lp::Gnn net;
        net.ensure(maxF, maxW, maxU, true);
        for (int step = 1; step <= steps; ++step) {
            const int k = (step - 1) % npanels;
            const xfe_lp_graph& g = panels[k];
            Dev& d = dev[k];
            net.ensure(g.F, g.Wm, g.U, true);
            lp::Graph gv{d.xf.p, d.xw.p, d.fp.p, d.fi.p, d.wp.p, d.wi.p, d.pf.p, d.pg.p, g.pq ? d.pq.p : nullptr,
                         d.rk.p};
            lp::Targets tg{d.ut.p, d.um.p, d.y.p, d.nm};
            net.forward(gv, th.p, s);
            const double loss = net.loss_backward(gv, th.p, tg, grad.p, s);
            if (loss_hist) loss_hist[step - 1] = loss;
            lp::adam_step(th.p, m.p, v.p, grad.p, step, lr, s);
        }
        XFE_CHECK(hipMemcpyAsync(theta, th.p, sizeof(float) * lp::kNumParams, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        net.release();
        for (auto& d : dev) {
            for (auto* b : {&d.xf, &d.xw, &d.pq, &d.ut, &d.um, &d.y}) b->release();
            for (auto* b : {&d.fp, &d.fi, &d.wp, &d.wi, &d.pf, &d.pg, &d.rk}) b->release();
        }
        th.release(); m.release(); v.release(); grad.release();
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

}  // extern "C"
