// Host driver of the MI355X fixed-effects absorber: the C ABI of
// include/xfe_hip.h for absorption (xfe_absorb, xfe_absorb_device).
//
// The solver follows the reference CPU/GPU absorber's algorithm
// (src/fe_absorption.cpp:5993-7088 packed Irons-Tuck path, mirrored by
// src/fe_absorption_cuda.cu:2294-4613): Gauss-Seidel (or symmetric) sweeps of
// FE projections, Irons-Tuck extrapolation with grand acceleration every 4
// iterations, the divergence safeguard with CG hand-off, the honest
// (reghdfe-comparable) stopping rule ||G(x)-x|| <= tol max(1,||x||), the
// two-stage warm-up for 3+ FEs, and the tightening polish.  The arithmetic is
// organised MI355X-first: every projection is one fused streaming pass over
// the column block (absorb_kernels.hip), rows are permuted once so the largest
// dimension is level-sorted, and each iteration does a single blocking copy
// of 3 scalars.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "absorb_kernels.h"
#include "common.h"
#include "lp_precond.h"
#include "slope_mixed.h"
#include "xfe_hip.h"

namespace xfe {

Profile g_profile;

namespace {
std::vector<hipEvent_t> g_prof_events;  // 2 per slot, reused
std::vector<double> g_prof_bytes;
size_t g_prof_used = 0;
}  // namespace

int profile_begin(hipStream_t s) {
    if (!g_profile.enabled) return -1;
    if (2 * (g_prof_used + 1) > g_prof_events.size()) {
        for (int k = 0; k < 2; ++k) {
            hipEvent_t e = nullptr;
            XFE_CHECK(hipEventCreate(&e));
            g_prof_events.push_back(e);
        }
        g_prof_bytes.push_back(0.0);
    }
    XFE_CHECK(hipEventRecord(g_prof_events[2 * g_prof_used], s));
    return static_cast<int>(g_prof_used++);
}

void profile_end(int slot, hipStream_t s, double bytes) {
    if (slot < 0) return;
    XFE_CHECK(hipEventRecord(g_prof_events[2 * static_cast<size_t>(slot) + 1], s));
    g_prof_bytes[static_cast<size_t>(slot)] = bytes;
}

void profile_resolve() {
    for (size_t i = 0; i < g_prof_used; ++i) {
        XFE_CHECK(hipEventSynchronize(g_prof_events[2 * i + 1]));
        float ms = 0.f;
        XFE_CHECK(hipEventElapsedTime(&ms, g_prof_events[2 * i], g_prof_events[2 * i + 1]));
        g_profile.ms += ms;
        g_profile.launches += 1;
        g_profile.bytes += g_prof_bytes[i];
    }
    g_prof_used = 0;
}

namespace {

constexpr int kMaxGrid = 2048;
constexpr size_t kLdsBudget = 32 * 1024;  // bytes of LDS a privatised reduce may use

int grid_for(long long work) {
    long long g = (work + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > kMaxGrid) g = kMaxGrid;
    return static_cast<int>(g);
}

// Per-thread device workspace (mirrors the reference's thread-local
// persistent CUDA workspace, src/fe_absorption_cuda.cu:1536).
struct Workspace {
    int device = -1;
    hipStream_t stream = nullptr;
    DevBuf<double> vec[7];  // x, gx, ggx, acc_a, acc_b, acc_c, spare
    DevBuf<int> gid[XFE_MAX_DIMS];
    DevBuf<double> sums[XFE_MAX_DIMS];
    DevBuf<double> means[XFE_MAX_DIMS];
    DevBuf<double> wsum[XFE_MAX_DIMS];
    DevBuf<double> inv[XFE_MAX_DIMS];
    DevBuf<int> perm, keys_tmp, vals_tmp, flag;
    DevBuf<unsigned char> sort_tmp;
    DevBuf<double> w;
    DevBuf<double> partials;
    DevBuf<double> prof_out;  // level-size profiles, 4 per dimension
    DevBuf<double> scalars;
    DevBuf<double> lanes;  // per-column scalars (CG alpha/beta)
    DevBuf<double> stage;  // host-API staging of [y | X]
    DevBuf<int> stage_ids;
    DevBuf<double> stage_w;
    DevBuf<double> stage_z[XFE_MAX_DIMS];  // host-API staging of slope variables
    DevBuf<double> slope_z[XFE_MAX_DIMS];  // slope variables in the slope path's row order
    SlopeMixed* slope_mixed = nullptr;     // the mixed-precision slope path's buffers
    // level-aligned tiles of the sorted dimension and spell structures
    DevBuf<int> tile_start, level_count, lcount, tmp_flags, tmp_inc, tmp_iota, tmp_key;
    DevBuf<int> sp_key[XFE_MAX_DIMS], sp_order[XFE_MAX_DIMS], sp_base[XFE_MAX_DIMS];
    DevBuf<double> sp_sums[XFE_MAX_DIMS];
    // mixed-precision solver (coefficient space of the non-sorted dimensions)
    DevBuf<double> f_beta, f_b, f_diag, f_counts, f_scale, cg, cg_part;
    DevBuf<float> f_x, f_r, f_z, f_p, f_q, f_minv, f_invp;
    DevBuf<float> f_sp[XFE_MAX_DIMS];
    // segment form of the Schur operator (worker-major and firm-major records)
    DevBuf<int> seg_start, seg_w, seg_r1, seg_len, seg_a, wseg, f_key, f_order, f_w, f_len, f_a;
    DevBuf<int> w_pf, w_pi, w_rank, seg_rk, p_order, p_rk, p_r1, p_len, p_a, p_wseg;
    DevBuf<float> f_wmean, f_hpart, f_invw, f_invwm;
    DevBuf<double> d_wmean, d_hpart, seg_tot, f_b0, f_tmp, d_invw, d_invwm;
    DevBuf<double> s64_z, s64_p, s64_q, s64_coef;  // the FP64 Schur PCG (precision 2)
    DevBuf<int> m_flag, m_idx, m_list;  // movers among the worker ranks
    DevBuf<int> fm_key, fm_w, fm_len, fm_a, stay_h;  // movers' firm-major records; stayer histogram
    DevBuf<double> stay_K;              // the stayers' year block (G2 x G2)
    DevBuf<int> seg_bk[2];
    DevBuf<float> f_bv[2];
    DevBuf<double> d_bv[2];
    // group()/individual() structure: incidence lists, scales, colour classes
    DevBuf<int> gi_iptr, gi_igroup, gi_members;
    DevBuf<double> gi_scale, gi_denom;
    // FE recovery: the untransformed column and the level coefficients
    DevBuf<double> rec_p, rec_contrib;
    DevBuf<double> rec_alpha[3];
    // aggregate-block preconditioner of the mixed-precision solver
    xfe::lp::AggPrecond agg;
    // the last mixed-precision solve of this workspace: primary dimension and
    // where each other dimension's coefficients sit in f_beta (level-major, C wide)
    struct {
        bool used = false;
        int p = -1, NR = 0, C = 0;
        int rdim[2] = {-1, -1};
        int off[2] = {0, 0};
        // the certificate's per-column sums of squared level sums of the final
        // transform, taken from the transform pass itself (valid for cert_V)
        bool cert_ready = false;
        const double* cert_V = nullptr;
        std::vector<double> cert_sq;
    } fast_rec;
    DevBuf<double> cert_psum;  // per level of p: sums of the final transform's rows
    double* host_scalars = nullptr;  // pinned

    // Frees every device resource of this thread's workspace.
    void release() {
        slope_mixed_delete(slope_mixed);
        slope_mixed = nullptr;
        for (auto& b : vec) b.release();
        for (int d = 0; d < XFE_MAX_DIMS; ++d) {
            gid[d].release(); sums[d].release(); means[d].release();
            wsum[d].release(); inv[d].release();
        }
        perm.release(); keys_tmp.release(); vals_tmp.release(); flag.release();
        sort_tmp.release(); w.release(); partials.release(); scalars.release(); prof_out.release();
        lanes.release(); stage.release(); stage_ids.release(); stage_w.release();
        tile_start.release(); level_count.release(); lcount.release(); tmp_flags.release(); tmp_inc.release();
        tmp_iota.release(); tmp_key.release();
        for (int d = 0; d < XFE_MAX_DIMS; ++d) {
            sp_key[d].release(); sp_order[d].release(); sp_base[d].release(); sp_sums[d].release();
            f_sp[d].release(); stage_z[d].release(); slope_z[d].release();
        }
        cg.release(); cg_part.release(); f_beta.release(); f_b.release(); f_diag.release(); f_counts.release(); f_scale.release();
        f_x.release(); f_r.release(); f_z.release(); f_p.release(); f_q.release(); f_minv.release();
        f_invp.release();
        seg_start.release(); seg_w.release(); seg_r1.release(); seg_len.release(); seg_a.release();
        wseg.release(); f_key.release(); f_order.release(); f_w.release(); f_len.release(); f_a.release();
        f_wmean.release(); f_hpart.release(); d_wmean.release(); d_hpart.release(); seg_tot.release();
        f_b0.release(); f_tmp.release();
        s64_z.release(); s64_p.release(); s64_q.release(); s64_coef.release();
        w_pf.release(); w_pi.release(); w_rank.release(); seg_rk.release(); p_order.release();
        p_rk.release(); p_r1.release(); p_len.release(); p_a.release(); p_wseg.release();
        f_invw.release(); d_invw.release(); f_invwm.release(); d_invwm.release();
        m_flag.release(); m_idx.release(); m_list.release();
        fm_key.release(); fm_w.release(); fm_len.release(); fm_a.release(); stay_h.release(); stay_K.release();
        for (int k = 0; k < 2; ++k) { seg_bk[k].release(); f_bv[k].release(); d_bv[k].release(); }
        rec_p.release(); rec_contrib.release();
        for (auto& b : rec_alpha) b.release();
        agg.release();
        gi_iptr.release(); gi_igroup.release(); gi_members.release(); gi_scale.release(); gi_denom.release();
        fast_rec.used = false;
        fast_rec.cert_ready = false;
        cert_psum.release();
        if (host_scalars) (void)hipHostFree(host_scalars);
        if (stream) (void)hipStreamDestroy(stream);
        host_scalars = nullptr; stream = nullptr; device = -1;
    }

    void init() {
        int dev = 0;
        XFE_CHECK(hipGetDevice(&dev));
        if (stream && dev == device) return;
        device = dev;
        XFE_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        XFE_CHECK(hipHostMalloc(reinterpret_cast<void**>(&host_scalars), 4096 * sizeof(double),
                                hipHostMallocDefault));
    }
};

thread_local Workspace tls_ws;

// Phase timings of one absorption on stderr when XFE_PHASES is set (each mark
// waits for the stream, so the marks themselves cost time; off otherwise).
struct PhaseClock {
    bool on = false;
    hipStream_t s = nullptr;
    std::chrono::steady_clock::time_point t;
    void start(hipStream_t st) {
        on = std::getenv("XFE_PHASES") != nullptr;
        s = st;
        if (on) {
            XFE_CHECK(hipStreamSynchronize(s));
            t = std::chrono::steady_clock::now();
        }
    }
    void mark(const char* name) {
        if (!on) return;
        XFE_CHECK(hipStreamSynchronize(s));
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[xfe phase] %-28s %8.3f ms\n", name,
                     std::chrono::duration<double, std::milli>(now - t).count());
        t = now;
    }
};
thread_local PhaseClock tls_phase;

// Source of a sweep: a stored block, or the Irons-Tuck extrapolation
// P - coef (P - Q) computed on the fly and written to xout.
struct SweepSrc {
    const double* src = nullptr;
    const double* P = nullptr;
    const double* Q = nullptr;
    double coef = 0.0;
    double* xout = nullptr;
    SweepSrc() = default;
    explicit SweepSrc(const double* s_) : src(s_) {}
    SweepSrc(const double* p_, const double* q_, double c_, double* x_) : P(p_), Q(q_), coef(c_), xout(x_) {}
};

struct Solver {
    Workspace& ws;
    hipStream_t s;
    int n = 0, C = 0, ld = 0, D = 0;
    int G[XFE_MAX_DIMS] = {};
    const int* gid[XFE_MAX_DIMS] = {};
    const double* w = nullptr;
    int grid = 1;
    int nchunks = 1;
    int sorted_dim = -1;          // the dimension whose levels are contiguous runs
    int xcd_copies = kXcdCopies;  // private accumulators for unsorted dimensions
    // tiled mode: level-aligned tiles of sorted_dim, spells for large dimensions
    bool tiled = false;
    const int* tile_start = nullptr;
    int num_tiles = 0;
    int tile_grid = 1;
    struct Spells {
        int M = 0;
        double* S = nullptr;
        const int* order = nullptr;
        const int* key = nullptr;
        const int* base = nullptr;
    } spell[XFE_MAX_DIMS];
    double pull_ms = 0.0;

    Solver(Workspace& w_, int n_, int C_, int ld_) : ws(w_), s(w_.stream), n(n_), C(C_), ld(ld_) {
        grid = pass_grid(n);
        nchunks = (C + 7) / 8;
        if (const char* e = std::getenv("XFE_XCD_COPIES")) xcd_copies = std::max(1, std::min(8, std::atoi(e)));
    }

    static int pass_grid(int rows) {
        static const int max_grid = [] {
            const char* e = std::getenv("XFE_MAX_GRID");
            const int v = e ? std::atoi(e) : 0;
            return v > 0 ? std::min(v, 65536) : kMaxGrid;
        }();
        const long long tiles = (static_cast<long long>(rows) + kBlock - 1) / kBlock;
        return static_cast<int>(std::max(1LL, std::min<long long>(tiles, max_grid)));
    }

    // accumulation mode of dimension d for a pass over `cols` columns per launch
    int mode_of(int d, int cols) const {
        if (static_cast<size_t>(G[d]) * std::min(8, cols) * sizeof(double) <= kLdsBudget) return kRedLds;
        if (d == sorted_dim) return kRedSorted;
        return kRedGlobal;
    }
    int copies_of(int d, int cols) const { return mode_of(d, cols) == kRedGlobal ? xcd_copies : 1; }

    // size and clear the accumulators of dimension d (call once per solve)
    void prepare_dim(int d) {
        const size_t Gd = static_cast<size_t>(std::max(G[d], 1));
        const size_t cs = static_cast<size_t>(std::max(copies_of(d, C), copies_of(d, 1)));
        ws.sums[d].ensure(Gd * C * cs);
        ws.means[d].ensure(Gd * C);
        ws.wsum[d].ensure(Gd);
        ws.inv[d].ensure(Gd);
        XFE_CHECK(hipMemsetAsync(ws.sums[d].p, 0, sizeof(double) * Gd * C * cs, s));
    }

    // sum of weights (counts when w == nullptr) per level of d into ws.wsum[d]
    // and its inverse into ws.inv[d]; runs the fused pass on a column of ones
    void weight_sums(int d) {
        if (!w && (d == sorted_dim || G[d] <= kLevelCountLds)) {
            // counts: integer histograms (exact, as the FP64 sums of ones)
            ws.lcount.ensure(static_cast<size_t>(std::max(G[d], 1)));
            XFE_CHECK(hipMemsetAsync(ws.lcount.p, 0, sizeof(int) * std::max(G[d], 1), s));
            XFE_CHECK(launch_level_weights(gid[d], n, G[d], d == sorted_dim, ws.lcount.p, ws.wsum[d].p, grid, s));
            XFE_CHECK(launch_inverse(ws.wsum[d].p, ws.inv[d].p, G[d], grid_for(G[d]), s));
            return;
        }
        PassArgs a;
        a.n = n; a.ld = ld; a.C = 1; a.c0 = 0; a.one = 1;
        a.gid_r = gid[d]; a.sums_r = ws.sums[d].p; a.G_r = G[d]; a.w = w;
        const int mode = mode_of(d, 1);
        a.xcd_copies = copies_of(d, 1);
        a.copy_stride = static_cast<size_t>(G[d]);
        XFE_CHECK(launch_pass(a, 1, mode, 0, grid, s));
        XFE_CHECK(launch_finalize(ws.sums[d].p, ws.wsum[d].p, nullptr, G[d], 1, a.xcd_copies,
                                  nullptr, grid_for(G[d]), s));
        XFE_CHECK(launch_inverse(ws.wsum[d].p, ws.inv[d].p, G[d], grid_for(G[d]), s));
    }

    // One fused pass over the column block (all column chunks).
    void pass(const double* src, double* dst, const double* P, const double* Q, double coef,
              double* xout, int a_dim, int r_dim, int nq, const double* it_x, const double* it_gx,
              double* partials) {
        for (int ch = 0; ch < nchunks; ++ch) {
            const int c0 = ch * 8;
            const int cc = std::min(8, C - c0);
            PassArgs a;
            a.n = n; a.ld = ld; a.C = C; a.c0 = c0;
            a.src = src; a.dst = dst; a.P = P; a.Q = Q; a.coef = coef; a.xout = xout;
            if (a_dim >= 0) { a.gid_a = gid[a_dim]; a.means_a = ws.means[a_dim].p; }
            if (r_dim >= 0) {
                a.gid_r = gid[r_dim]; a.sums_r = ws.sums[r_dim].p; a.G_r = G[r_dim];
                a.w = w;
            }
            a.it_x = it_x; a.it_gx = it_gx;
            a.partials = partials ? partials + static_cast<size_t>(ch) * grid * nq : nullptr;
            const int mode = r_dim >= 0 ? mode_of(r_dim, C) : kRedNone;
            if (r_dim >= 0) {
                a.xcd_copies = copies_of(r_dim, C);
                a.copy_stride = static_cast<size_t>(G[r_dim]) * C;
            }
            const int pslot = profile_begin(s);
            XFE_CHECK(launch_pass(a, cc, mode, nq, grid, s));
            if (pslot >= 0) {
                // algorithmic bytes of this launch
                double vec_rw = (P ? 2.0 : 1.0) + (xout ? 1.0 : 0.0) + (dst ? 1.0 : 0.0) +
                                (it_x ? 2.0 : 0.0);
                double bytes = vec_rw * 8.0 * static_cast<double>(n) * cc;
                if (a_dim >= 0) bytes += 4.0 * n + 8.0 * static_cast<double>(G[a_dim]) * cc;
                if (r_dim >= 0) bytes += 4.0 * n + 16.0 * static_cast<double>(G[r_dim]) * cc;
                if (r_dim >= 0 && w) bytes += 8.0 * n;
                profile_end(pslot, s, bytes);
            }
        }
    }

    void finalize(int d, double* partials) {
        const long long total = static_cast<long long>(G[d]) * C;
        XFE_CHECK(launch_finalize(ws.sums[d].p, ws.means[d].p, ws.inv[d].p, G[d], C,
                                  copies_of(d, C), partials, grid_for(total), s));
    }
    int finalize_grid(int d) const { return grid_for(static_cast<long long>(G[d]) * C); }

    void clear_sums(int d) {
        XFE_CHECK(hipMemsetAsync(ws.sums[d].p, 0,
                                 sizeof(double) * static_cast<size_t>(G[d]) * C * copies_of(d, C), s));
    }

    // Reduce `count` partial records of nq doubles on device and copy the nq
    // results to host memory (blocking).
    void fetch(const double* partials, int count, int nq, int mode, double* out) {
        XFE_CHECK(launch_reduce_partials(partials, count, nq, mode, ws.scalars.p, s));
        XFE_CHECK(hipMemcpyAsync(ws.host_scalars, ws.scalars.p, sizeof(double) * nq,
                                 hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        for (int j = 0; j < nq; ++j) out[j] = ws.host_scalars[j];
    }

    // dst <- G(src) over the projection sequence `seq`.  If pre_reduced, the
    // sums of seq[0] already hold src's reduce.  tail_reduce >= 0 starts the
    // reduce of that dim on the output (fusion with the next sweep).
    // nq_last: 1 -> sum of squares of the output, 3 -> IT stats vs (it_x, it_gx).
    void sweep(const double* src, double* dst, const std::vector<int>& seq, bool pre_reduced,
               int tail_reduce, int nq_last, const double* it_x, const double* it_gx,
               double* out_q) {
        const int m = static_cast<int>(seq.size());
        if (m == 0) {
            if (src != dst) XFE_CHECK(hipMemcpyAsync(dst, src, sizeof(double) * ld * C,
                                                     hipMemcpyDeviceToDevice, s));
            return;
        }
        if (!pre_reduced) pass(src, nullptr, nullptr, nullptr, 0.0, nullptr, -1, seq[0], 0,
                               nullptr, nullptr, nullptr);
        for (int j = 0; j < m; ++j) {
            finalize(seq[j], nullptr);
            const bool last = j + 1 == m;
            const int r = last ? tail_reduce : seq[j + 1];
            const int nq = last ? nq_last : 0;
            pass(j == 0 ? src : dst, dst, nullptr, nullptr, 0.0, nullptr, seq[j], r, nq,
                 last ? it_x : nullptr, last ? it_gx : nullptr, nq ? ws.partials.p : nullptr);
        }
        if (nq_last && out_q) fetch(ws.partials.p, grid * nchunks, nq_last, 0, out_q);
    }

    double sumsq(const double* v, int reduce_dim) {
        pass(v, nullptr, nullptr, nullptr, 0.0, nullptr, -1, reduce_dim, 1, nullptr, nullptr,
             ws.partials.p);
        double q = 0.0;
        fetch(ws.partials.p, grid * nchunks, 1, 0, &q);
        return q;
    }

    void copy(double* dst, const double* src) {
        XFE_CHECK(hipMemcpyAsync(dst, src, sizeof(double) * static_cast<size_t>(ld) * C,
                                 hipMemcpyDeviceToDevice, s));
    }

    // ------------------------------------------------------------ tiled mode

    int tile_mode_of(int d) const {
        if (static_cast<size_t>(G[d]) * std::min(8, C) * sizeof(double) <= kLdsBudget) return kRedLds;
        return kRedSpell;
    }

    // One tiled pass over all column chunks: [source] -> [apply a] -> [project
    // sorted_dim inside the tile if pp] -> [start reduce r] -> [dst, NQ].
    void tile_pass(const SweepSrc& in, double* dst, int a_dim, bool pp, int r_dim, int nq,
                   const double* it_x, const double* it_gx, double* partials, double* maxp) {
        const int rr = r_dim >= 0 ? tile_mode_of(r_dim) : 0;
        for (int ch = 0; ch < nchunks; ++ch) {
            const int c0 = ch * 8;
            const int cc = std::min(8, C - c0);
            TileArgs a;
            a.n = n; a.ld = ld; a.C = C; a.c0 = c0;
            a.src = in.src; a.P = in.P; a.Q = in.Q; a.coef = in.coef; a.xout = in.xout;
            a.dst = dst;
            if (a_dim >= 0) { a.gid_a = gid[a_dim]; a.means_a = ws.means[a_dim].p; }
            a.tile_start = tile_start; a.num_tiles = num_tiles;
            a.gid_p = gid[sorted_dim]; a.inv_p = ws.inv[sorted_dim].p;
            if (r_dim >= 0) {
                a.gid_r = gid[r_dim]; a.G_r = G[r_dim]; a.sums_r = ws.sums[r_dim].p;
                a.spell_sums = spell[r_dim].S; a.tile_spell_base = spell[r_dim].base;
            }
            a.w = w;
            a.it_x = it_x; a.it_gx = it_gx;
            a.partials = partials ? partials + static_cast<size_t>(ch) * tile_grid * nq : nullptr;
            a.maxp_partials = maxp ? maxp + static_cast<size_t>(ch) * tile_grid : nullptr;
            const int pslot = profile_begin(s);
            XFE_CHECK(launch_tile_pass(a, cc, rr, pp, nq, tile_grid, s));
            if (pslot >= 0) {
                // algorithmic bytes: vectors read/written, id arrays, level tables
                const double vec_rw = (in.P ? 2.0 : 1.0) + (in.xout ? 1.0 : 0.0) +
                                      (dst ? 1.0 : 0.0) + (it_x ? 2.0 : 0.0);
                double bytes = vec_rw * 8.0 * static_cast<double>(n) * cc;
                if (a_dim >= 0) bytes += 4.0 * n + 8.0 * static_cast<double>(G[a_dim]) * cc;
                if (pp || rr == kRedSpell) bytes += 4.0 * n;
                if (pp) bytes += 8.0 * static_cast<double>(G[sorted_dim]);
                if (rr != 0) bytes += 4.0 * n;
                if (rr == kRedSpell) bytes += 8.0 * static_cast<double>(spell[r_dim].M) * cc;
                if (w && (pp || rr != 0)) bytes += 8.0 * n;
                profile_end(pslot, s, bytes);
            }
        }
    }

    // Level sums of r from its spell sums, then means.
    void pull_and_finalize(int r, double* partials) {
        if (tile_mode_of(r) == kRedSpell) {
            for (int ch = 0; ch < nchunks; ++ch) {
                const int c0 = ch * 8;
                const int cc = std::min(8, C - c0);
                XFE_CHECK(launch_spell_pull(spell[r].S, C, c0, cc, spell[r].order, spell[r].key,
                                            spell[r].M, ws.sums[r].p, grid_for(spell[r].M), s));
            }
        }
        XFE_CHECK(launch_finalize(ws.sums[r].p, ws.means[r].p, ws.inv[r].p, G[r], C, 1, partials,
                                  grid_for(static_cast<long long>(G[r]) * C), s));
    }

    // dst <- G(in) over the projection sequence seq (see sweep()).  In tiled
    // mode each pass is [apply pending] [project sorted_dim in the tile] [start
    // the next reduce]; otherwise the chain of reduce / apply passes.  When
    // maxp_out is given it receives the largest |mean| of all projections of
    // the sweep (polish rule).
    void sweep2(SweepSrc in, double* dst, const std::vector<int>& seq, int nq_last,
                const double* it_x, const double* it_gx, double* out_q, double* maxp_out = nullptr) {
        const int m = static_cast<int>(seq.size());
        if (m == 0) {
            if (in.P) {
                pass(nullptr, nullptr, in.P, in.Q, in.coef, in.xout, -1, -1, 0, nullptr, nullptr, nullptr);
                in = SweepSrc{in.xout};
            }
            if (in.src != dst) copy(dst, in.src);
            return;
        }
        double* part_max = maxp_out ? ws.partials.p + maxp_offset() : nullptr;
        size_t max_off = 0;
        if (!tiled) {
            if (in.P) {
                pass(nullptr, nullptr, in.P, in.Q, in.coef, in.xout, -1, seq[0], 0, nullptr, nullptr,
                     nullptr);
                in = SweepSrc{in.xout};
            } else {
                pass(in.src, nullptr, nullptr, nullptr, 0.0, nullptr, -1, seq[0], 0, nullptr,
                     nullptr, nullptr);
            }
            for (int j = 0; j < m; ++j) {
                finalize(seq[j], part_max ? part_max + max_off : nullptr);
                if (part_max) max_off += finalize_grid(seq[j]);
                const bool last = j + 1 == m;
                const int r = last ? -1 : seq[j + 1];
                const int nq = last ? nq_last : 0;
                pass(j == 0 ? in.src : dst, dst, nullptr, nullptr, 0.0, nullptr, seq[j], r, nq,
                     last ? it_x : nullptr, last ? it_gx : nullptr, nq ? ws.partials.p : nullptr);
            }
            if (nq_last && out_q) fetch(ws.partials.p, grid * nchunks, nq_last, 0, out_q);
        } else {
            const int p = sorted_dim;
            int i = 0, pending = -1;
            bool wrote = false;
            while (true) {
                const int a = pending;
                bool pp = false;
                if (i < m && seq[i] == p) { pp = true; ++i; }
                int r = -1;
                if (i < m) { r = seq[i]; ++i; }
                const bool last = r < 0;
                const int nq = last ? nq_last : 0;
                const bool writes = a >= 0 || pp;
                SweepSrc src_now = wrote ? SweepSrc{dst} : in;
                double* mp = nullptr;
                if (part_max && pp) {
                    mp = part_max + max_off;
                    max_off += static_cast<size_t>(tile_grid) * nchunks;
                }
                tile_pass(src_now, writes ? dst : nullptr, a, pp, r, nq, last ? it_x : nullptr,
                          last ? it_gx : nullptr, nq ? ws.partials.p : nullptr, mp);
                if (!wrote && in.P) in = SweepSrc{in.xout};  // the extrapolated values now live in xout
                wrote = wrote || writes;
                if (last) break;
                double* fp = nullptr;
                if (part_max) {
                    fp = part_max + max_off;
                    max_off += finalize_grid(r);
                }
                pull_and_finalize(r, fp);
                pending = r;
            }
            if (!wrote) {  // a sweep that never applied a projection (not reachable with m > 0)
                if (in.src != dst) copy(dst, in.src);
            }
            if (nq_last && out_q) fetch(ws.partials.p, tile_grid * nchunks, nq_last, 0, out_q);
        }
        if (maxp_out) {
            double mx = 0.0;
            fetch(part_max, static_cast<int>(max_off), 1, 1, &mx);
            *maxp_out = mx;
        }
    }

    // partial-buffer offset reserved for polish maxima (after the pass partials)
    size_t maxp_offset() const {
        return static_cast<size_t>(std::max(grid, tile_grid)) * nchunks * 3 + 64;
    }
};

// Build the projection sequence of one sweep: forward over `order`, backward
// when symmetric, with immediately repeated projections dropped (P_d is
// idempotent, so d,d acts as d).
std::vector<int> sweep_sequence(const std::vector<int>& order, bool symmetric) {
    std::vector<int> seq;
    for (int d : order) {
        if (seq.empty() || seq.back() != d) seq.push_back(d);
    }
    if (symmetric) {
        for (int p = static_cast<int>(order.size()); p-- > 0;) {
            if (seq.empty() || seq.back() != order[p]) seq.push_back(order[p]);
        }
    }
    return seq;
}

struct ProfileStats {
    int num_groups = 0;
    double mean = 0.0, cv = 0.0, singleton_ratio = 0.0, max_val = 0.0;
};

// profile_fe, src/fe_absorption.cpp:4945-4982.
ProfileStats profile_from_sums(double sum, double sq, double mx, double single, int G, bool unit) {
    ProfileStats p;
    p.num_groups = G;
    if (G <= 0) return p;
    p.mean = sum / G;
    const double var = std::max(0.0, sq / G - p.mean * p.mean);
    p.cv = p.mean > 0.0 ? std::sqrt(var) / p.mean : 0.0;
    p.max_val = mx;
    p.singleton_ratio = unit ? single / G : 0.0;
    return p;
}

// order_by_profile, src/fe_absorption.cpp:5472-5491.
std::vector<int> order_by_profile(const std::vector<ProfileStats>& p) {
    std::vector<int> order(p.size());
    for (size_t i = 0; i < p.size(); ++i) order[i] = static_cast<int>(i);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (p[a].num_groups != p[b].num_groups) return p[a].num_groups > p[b].num_groups;
        if (p[a].singleton_ratio != p[b].singleton_ratio) return p[a].singleton_ratio < p[b].singleton_ratio;
        if (p[a].cv != p[b].cv) return p[a].cv < p[b].cv;
        return a < b;
    });
    return order;
}

struct AccelState {
    double *cur, *gx, *ggx, *acc_a, *acc_b, *acc_c;
    bool diverged = false;
};

// packed_accel_phase, src/fe_absorption.cpp:6599-6782 (device form).
// Leaves the iterate in st.cur.  Returns converged.
bool accel_phase(Solver& sv, AccelState& st, const std::vector<int>& seq, int max_iter,
                 bool honest, double honest_tol, double conv_tol, int& iter_used, int& sweeps) {
    iter_used = 0;
    if (max_iter <= 0 || seq.empty()) return false;
    const double kDiverge = 4.0, kBail = 1.5;
    const int cg_iter_cap = 1000, grand_every = 4;
    int last_check = -1;
    double prev_norm = std::sqrt(sv.sumsq(st.cur, -1));
    int grand = 0;
    double best = prev_norm;
    bool suspended = false;
    sv.sweep2(SweepSrc(st.cur), st.gx, seq, 0, nullptr, nullptr, nullptr);  // gx = G(cur)
    ++sweeps;
    for (int iter = 0; iter < max_iter; ++iter) {
        double q[3];
        sv.sweep2(SweepSrc(st.gx), st.ggx, seq, 3, st.cur, st.gx, q);  // ggx = G(gx), IT stats
        ++sweeps;
        const double vprod = q[0], ssq = q[1], d1sq = q[2];
        if (honest && std::sqrt(d1sq) <= honest_tol * std::max(1.0, prev_norm)) {
            std::swap(st.cur, st.ggx);
            iter_used = iter + 1;
            return true;
        }
        if (ssq == 0.0) {
            sv.copy(st.cur, st.gx);
            iter_used = iter + 1;
            return true;
        }
        double coef = vprod / ssq;
        if (suspended) coef = 0.0;
        // cur = ggx - coef (ggx - gx); gx = G(cur) with its norm
        double sq = 0.0;
        sv.sweep2(SweepSrc(st.ggx, st.gx, coef, st.cur), st.gx, seq, 1, nullptr, nullptr, &sq);
        ++sweeps;
        {
            const double curr = std::sqrt(sq);
            const double rel_change = std::fabs(curr - prev_norm) / std::max(1.0, prev_norm);
            const int step = iter - last_check;
            last_check = iter;
            prev_norm = curr;
            if (curr <= best) {
                best = curr;
                suspended = false;
            } else {
                if (curr > kBail * best) {
                    st.diverged = true;
                    std::swap(st.cur, st.gx);
                    iter_used = iter + 1;
                    return false;
                }
                if (curr > kDiverge * best) {
                    if (!suspended) grand = 0;
                    suspended = true;
                }
            }
            if (!honest && step > 0 && rel_change / static_cast<double>(step) < conv_tol) {
                std::swap(st.cur, st.gx);
                iter_used = iter + 1;
                return true;
            }
        }
        if (honest && cg_iter_cap > 0 && !st.diverged && iter + 1 >= cg_iter_cap) {
            st.diverged = true;
            std::swap(st.cur, st.gx);
            iter_used = iter + 1;
            return false;
        }
        if (!suspended && grand_every > 0 && (iter + 1) % grand_every == 0) {
            ++grand;
            if (grand == 1) {
                sv.copy(st.acc_a, st.gx);
            } else if (grand == 2) {
                sv.copy(st.acc_b, st.gx);
            } else {
                sv.copy(st.acc_c, st.gx);
                XFE_CHECK(launch_it_stats(st.acc_a, st.acc_b, st.acc_c, sv.n, sv.C, sv.ld,
                                          sv.ws.partials.p, sv.grid, sv.s));
                double g3[3];
                sv.fetch(sv.ws.partials.p, sv.grid, 3, 0, g3);
                if (g3[1] != 0.0) {
                    // cur = acc_c - ac (acc_c - acc_b); gx = G(cur)
                    const double ac = g3[0] / g3[1];
                    sv.sweep2(SweepSrc(st.acc_c, st.acc_b, ac, st.cur), st.gx, seq, 0, nullptr,
                              nullptr, nullptr);
                } else {
                    sv.copy(st.cur, st.acc_a);
                    sv.sweep2(SweepSrc(st.cur), st.gx, seq, 0, nullptr, nullptr, nullptr);
                }
                ++sweeps;
                grand = 0;
            }
        }
    }
    std::swap(st.cur, st.gx);
    iter_used = max_iter;
    return false;
}

// run_packed_cg, src/fe_absorption.cpp:6790-6973 (device form): per-column
// CG on the symmetric operator (I - G_sym).  Works in place on `cur`.
bool run_cg(Solver& sv, double* cur, double* r, double* u, double* v, const std::vector<int>& order,
            int max_iter, double tol, int& iters_out, int& sweeps) {
    const int C = sv.C;
    const std::vector<int> seq = sweep_sequence(order, true);
    std::vector<double> improvement(C), ssr(C), ssr_old(C), alpha(C), beta(C), recent(C, 0.0),
        denom(C), next_imp(C);
    std::vector<unsigned char> solved(C, 0);
    auto coldot = [&](const double* a, const double* b, std::vector<double>& out) {
        XFE_CHECK(launch_col_dot(a, b, sv.n, C, sv.ld, sv.ws.partials.p, sv.grid, sv.s));
        sv.fetch(sv.ws.partials.p, sv.grid, C, 0, out.data());
    };
    auto proj = [&](const double* src, double* dst) {
        sv.sweep2(SweepSrc(src), dst, seq, 0, nullptr, nullptr, nullptr);
        ++sweeps;
        XFE_CHECK(launch_sub(dst, src, dst, sv.n, C, sv.ld, sv.grid, sv.s));
    };
    auto upload_lanes = [&](const std::vector<double>& x) {
        XFE_CHECK(hipMemcpyAsync(sv.ws.lanes.p, x.data(), sizeof(double) * C,
                                 hipMemcpyHostToDevice, sv.s));
    };
    coldot(cur, cur, improvement);
    proj(cur, r);
    coldot(r, r, ssr);
    sv.copy(u, r);
    bool converged = false;
    int iters = 0;
    for (int iter = 1; iter <= max_iter; ++iter) {
        proj(u, v);
        coldot(u, v, denom);
        bool bd = false;
        for (int k = 0; k < C; ++k) { alpha[k] = 0.0; next_imp[k] = improvement[k]; }
        for (int k = 0; k < C && !bd; ++k) {
            if (!std::isfinite(improvement[k]) || !std::isfinite(ssr[k]) || ssr[k] < 0.0) { bd = true; break; }
            if (solved[k]) { if (ssr[k] != 0.0) bd = true; recent[k] = 0.0; continue; }
            if (ssr[k] == 0.0) { solved[k] = 1; recent[k] = 0.0; continue; }
            if (!std::isfinite(denom[k]) || denom[k] <= 0.0) { bd = true; break; }
            alpha[k] = ssr[k] / denom[k];
            recent[k] = alpha[k] * ssr[k];
            next_imp[k] = improvement[k] - recent[k];
            if (!std::isfinite(alpha[k]) || alpha[k] <= 0.0 || !std::isfinite(recent[k]) ||
                recent[k] <= 0.0 || !std::isfinite(next_imp[k])) { bd = true; break; }
        }
        iters = iter;
        if (bd) break;
        improvement = next_imp;
        upload_lanes(alpha);
        XFE_CHECK(launch_col_axpy(cur, cur, u, sv.ws.lanes.p, sv.n, C, sv.ld, -1.0, sv.grid, sv.s));
        XFE_CHECK(launch_col_axpy(r, r, v, sv.ws.lanes.p, sv.n, C, sv.ld, -1.0, sv.grid, sv.s));
        ssr_old = ssr;
        coldot(r, r, ssr);
        for (int k = 0; k < C && !bd; ++k) {
            beta[k] = 0.0;
            if (!std::isfinite(ssr[k]) || ssr[k] < 0.0 || !std::isfinite(ssr_old[k]) || ssr_old[k] < 0.0) { bd = true; break; }
            if (solved[k]) { if (ssr[k] != 0.0) bd = true; continue; }
            if (ssr[k] == 0.0) { solved[k] = 1; continue; }
            if (ssr_old[k] <= 0.0) { bd = true; break; }
            beta[k] = ssr[k] / ssr_old[k];
            if (!std::isfinite(beta[k]) || beta[k] <= 0.0) { bd = true; break; }
        }
        if (bd) break;
        upload_lanes(beta);
        XFE_CHECK(launch_col_axpy(u, r, u, sv.ws.lanes.p, sv.n, C, sv.ld, 1.0, sv.grid, sv.s));
        double update_error = 0.0;
        for (int k = 0; k < C; ++k) {
            if (solved[k]) continue;
            const double den = std::fabs(improvement[k]);
            if (!(recent[k] > 0.0) || !(den > 0.0)) { bd = true; break; }
            const double e = std::sqrt(recent[k] / den);
            if (!std::isfinite(e)) { bd = true; break; }
            update_error = std::max(update_error, e);
        }
        if (bd) break;
        if (update_error <= tol) { converged = true; break; }
    }
    iters_out = iters > 0 ? iters : max_iter;
    return converged;
}

}  // namespace

// ---------------------------------------------------------------------------
// Mixed-precision solver (xfe_options.precision = 1).
//
// Eliminates the level-sorted dimension p exactly (per-level means inside the
// tiles) and solves the Schur system S x = b over the coefficients of the other
// dimensions R by Jacobi-preconditioned CG in FP32 (FP64 dot products, step
// lengths and CG scalars), inside FP64 iterative refinement: the right-hand side
// b = D_R' M_p (v - D_R beta) is recomputed from the FP64 data after every inner
// solve, beta accumulates in FP64, and the final within transform
// M_p (v - D_R beta) is formed in FP64.  Convergence is judged only on the FP64
// reduced residual.  Applicable with 2 or 3 dimensions, unit weights, tiles and
// at most kMaxCgCols columns; returns false otherwise (the caller then runs the
// plain FP64 method).
struct FastSolver {
    Solver& sv;
    Workspace& ws;
    hipStream_t s;
    int p = -1, NR = 0, N = 0, C = 0;
    int rdim[2] = {-1, -1};
    int off[2] = {0, 0};
    int mode[2] = {0, 0};
    size_t lds_extra32 = 0, lds_extra64 = 0;
    int lds_off[2] = {0, 0};
    int vgrid = 1;
    int inner_total = 0;
    double pull_apply_ms = 0.0;
    // segment form: k1 indexes the spell dimension r1 in rdim, k2 the range
    // dimension r2 (-1 without one)
    bool seg = false;
    int k1 = 0, k2 = -1, M = 0, Nw = 0, G2 = 0, seg_grid_w = 1, seg_grid_f = 1;
    int Nm = 0;  // movers (workers with two or more segments)
    int Mf = 0;  // their segments: the firm-major records of the firm pass
    bool stay_years = false;  // stayers add the year block stay_K
    bool empty_firms = true;  // some level of r1 has no segment
    // aggregate-block preconditioner on the firm levels (segment form only):
    // precond 0 Jacobi, 1 aggregates from the coupling strengths, 2 from the
    // learned network's scores (params: its parameter vector, device)
    int precond = 0;
    const float* lp_params = nullptr;
    bool blocks = false;
    bool block_fallback = false;
    lp::BlockView bview;
    double precond_ms = 0.0;
    lp::ProbeState* probe = nullptr;  // training-data probe (xfe_lp_probe)

    explicit FastSolver(Solver& sv_) : sv(sv_), ws(sv_.ws), s(sv_.s) {}

    // Stable radix sort of m (key, value) pairs with keys in [0, max_key].
    void sort_pairs(const int* kin, int* kout, const int* vin, int* vout, int m, int max_key) {
        int bits = 1;
        while ((1LL << bits) <= static_cast<long long>(max_key)) ++bits;
        size_t bytes = 0;
        XFE_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, static_cast<size_t>(m), 0,
                                            bits, s));
        ws.sort_tmp.ensure(bytes);
        XFE_CHECK(rocprim::radix_sort_pairs(ws.sort_tmp.p, bytes, kin, kout, vin, vout, static_cast<size_t>(m),
                                            0, bits, s));
    }

    // Segment records of the sorted rows (see seg_kernels.hip).  Returns false
    // when the range dimension has too many levels for the LDS tables.
    bool setup_segments() {
        if (std::getenv("XFE_NO_SEGMENTS")) return false;
        if (NR == 2) {
            k1 = sv.G[rdim[0]] >= sv.G[rdim[1]] ? 0 : 1;
            k2 = 1 - k1;
            if (sv.G[rdim[k2]] > kSegMaxRange) return false;
            G2 = sv.G[rdim[k2]];
        } else {
            k1 = 0;
            k2 = -1;
            G2 = 0;
        }
        const int n = sv.n;
        Nw = sv.G[p];
        const int* gp = sv.gid[p];
        const int* g1 = sv.gid[rdim[k1]];
        const int* g2 = k2 >= 0 ? sv.gid[rdim[k2]] : nullptr;
        const int grid = grid_for(n);
        ws.tmp_flags.ensure(n);
        ws.tmp_inc.ensure(n);
        XFE_CHECK(launch_seg_heads(gp, g1, g2, n, ws.tmp_flags.p, grid, s));
        size_t bytes = 0;
        XFE_CHECK(rocprim::inclusive_scan(nullptr, bytes, ws.tmp_flags.p, ws.tmp_inc.p,
                                          static_cast<size_t>(n), rocprim::plus<int>(), s));
        ws.sort_tmp.ensure(bytes);
        XFE_CHECK(rocprim::inclusive_scan(ws.sort_tmp.p, bytes, ws.tmp_flags.p, ws.tmp_inc.p,
                                          static_cast<size_t>(n), rocprim::plus<int>(), s));
        XFE_CHECK(hipMemcpyAsync(&M, ws.tmp_inc.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        if (M <= 0) return false;
        ws.seg_start.ensure(M); ws.seg_w.ensure(M); ws.seg_r1.ensure(M); ws.seg_len.ensure(M);
        ws.seg_a.ensure(M); ws.wseg.ensure(static_cast<size_t>(Nw) + 1);
        ws.f_key.ensure(M); ws.f_order.ensure(M); ws.f_w.ensure(M); ws.f_len.ensure(M); ws.f_a.ensure(M);
        ws.tmp_iota.ensure(std::max(M, Nw));
        ws.w_pf.ensure(Nw); ws.w_pi.ensure(Nw); ws.w_rank.ensure(Nw); ws.seg_rk.ensure(M);
        ws.p_order.ensure(M); ws.p_rk.ensure(M); ws.p_r1.ensure(M); ws.p_len.ensure(M); ws.p_a.ensure(M);
        ws.p_wseg.ensure(static_cast<size_t>(Nw) + 1);
        ws.f_invw.ensure(Nw); ws.d_invw.ensure(Nw);
        const int mgrid = grid_for(M);
        const int wgrid = grid_for(Nw);
        const int G1 = sv.G[rdim[k1]];
        XFE_CHECK(launch_seg_fill(ws.tmp_flags.p, ws.tmp_inc.p, gp, g1, g2, n, ws.seg_start.p, ws.seg_w.p,
                                  ws.seg_r1.p, ws.seg_a.p, grid, s));
        XFE_CHECK(launch_seg_len(ws.seg_start.p, M, n, ws.seg_len.p, mgrid, s));
        XFE_CHECK(launch_seg_offsets(ws.seg_w.p, M, Nw, ws.wseg.p, mgrid, s));
        // Renumber workers by their primary firm (stable), so that the workers
        // of one firm are adjacent: the worker pass then gathers firm
        // coefficients of neighbouring firms and the firm pass reads the means
        // of neighbouring workers.
        XFE_CHECK(launch_seg_primary(ws.wseg.p, ws.seg_r1.p, Nw, G1, ws.w_pf.p, wgrid, s));
        XFE_CHECK(launch_iota(ws.tmp_iota.p, Nw, wgrid, s));
        sort_pairs(ws.w_pf.p, ws.w_rank.p, ws.tmp_iota.p, ws.w_pi.p, Nw, G1);  // sorted keys unused
        XFE_CHECK(launch_seg_rank(ws.w_pi.p, Nw, ws.w_rank.p, wgrid, s));
        XFE_CHECK(launch_seg_gather<int>(ws.w_rank.p, ws.seg_w.p, ws.seg_rk.p, M, mgrid, s));
        XFE_CHECK(launch_iota(ws.tmp_iota.p, M, mgrid, s));
        sort_pairs(ws.seg_rk.p, ws.p_rk.p, ws.tmp_iota.p, ws.p_order.p, M, Nw);
        XFE_CHECK(launch_seg_gather<int>(ws.seg_r1.p, ws.p_order.p, ws.p_r1.p, M, mgrid, s));
        XFE_CHECK(launch_seg_gather<int>(ws.seg_len.p, ws.p_order.p, ws.p_len.p, M, mgrid, s));
        if (k2 >= 0) XFE_CHECK(launch_seg_gather<int>(ws.seg_a.p, ws.p_order.p, ws.p_a.p, M, mgrid, s));
        XFE_CHECK(launch_seg_offsets(ws.p_rk.p, M, Nw, ws.p_wseg.p, mgrid, s));
        XFE_CHECK(launch_seg_gather<float>(ws.f_invp.p, ws.w_pi.p, ws.f_invw.p, Nw, wgrid, s));
        XFE_CHECK(launch_seg_gather<double>(ws.inv[p].p, ws.w_pi.p, ws.d_invw.p, Nw, wgrid, s));
        // firm-major order of the renumbered segments (stable: by worker rank within a firm)
        XFE_CHECK(launch_iota(ws.tmp_iota.p, M, mgrid, s));
        sort_pairs(ws.p_r1.p, ws.f_key.p, ws.tmp_iota.p, ws.f_order.p, M, G1);
        XFE_CHECK(launch_seg_gather<int>(ws.p_rk.p, ws.f_order.p, ws.f_w.p, M, mgrid, s));
        XFE_CHECK(launch_seg_gather<int>(ws.p_len.p, ws.f_order.p, ws.f_len.p, M, mgrid, s));
        if (k2 >= 0) XFE_CHECK(launch_seg_gather<int>(ws.p_a.p, ws.f_order.p, ws.f_a.p, M, mgrid, s));
        // movers: only their worker means are formed and read (a stayer's one
        // segment is handled by the firm pass alone)
        ws.m_flag.ensure(Nw); ws.m_idx.ensure(Nw); ws.m_list.ensure(Nw);
        XFE_CHECK(launch_seg_movers(ws.p_wseg.p, Nw, ws.m_flag.p, wgrid, s));
        {
            size_t sb = 0;
            XFE_CHECK(rocprim::exclusive_scan(nullptr, sb, ws.m_flag.p, ws.m_idx.p, 0, static_cast<size_t>(Nw),
                                              rocprim::plus<int>(), s));
            ws.sort_tmp.ensure(sb);
            XFE_CHECK(rocprim::exclusive_scan(ws.sort_tmp.p, sb, ws.m_flag.p, ws.m_idx.p, 0, static_cast<size_t>(Nw),
                                              rocprim::plus<int>(), s));
        }
        int mcount[2] = {0, 0};
        XFE_CHECK(hipMemcpyAsync(&mcount[0], ws.m_idx.p + (Nw - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipMemcpyAsync(&mcount[1], ws.m_flag.p + (Nw - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        Nm = mcount[0] + mcount[1];
        XFE_CHECK(launch_seg_mover_relabel(ws.m_flag.p, ws.m_idx.p, Nw, M, ws.m_list.p, ws.f_w.p, ws.f_len.p,
                                           std::max(wgrid, mgrid), s));
        ws.f_invwm.ensure(std::max(Nm, 1)); ws.d_invwm.ensure(std::max(Nm, 1));
        if (Nm > 0) {
            XFE_CHECK(launch_seg_gather<float>(ws.f_invw.p, ws.m_list.p, ws.f_invwm.p, Nm, grid_for(Nm), s));
            XFE_CHECK(launch_seg_gather<double>(ws.d_invw.p, ws.m_list.p, ws.d_invwm.p, Nm, grid_for(Nm), s));
        }
        // the firm pass takes the movers' segments only (a stayer's segment adds
        // nothing to its firm); the stayers' year block is the fixed matrix stay_K
        XFE_CHECK(launch_seg_mover_segments(ws.f_len.p, M, ws.tmp_flags.p, mgrid, s));
        {
            size_t sb = 0;
            XFE_CHECK(rocprim::exclusive_scan(nullptr, sb, ws.tmp_flags.p, ws.tmp_inc.p, 0, static_cast<size_t>(M),
                                              rocprim::plus<int>(), s));
            ws.sort_tmp.ensure(sb);
            XFE_CHECK(rocprim::exclusive_scan(ws.sort_tmp.p, sb, ws.tmp_flags.p, ws.tmp_inc.p, 0,
                                              static_cast<size_t>(M), rocprim::plus<int>(), s));
        }
        XFE_CHECK(hipMemcpyAsync(&mcount[0], ws.tmp_inc.p + (M - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipMemcpyAsync(&mcount[1], ws.tmp_flags.p + (M - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        Mf = mcount[0] + mcount[1];
        if (Mf <= 0) return false;
        ws.fm_key.ensure(Mf); ws.fm_w.ensure(Mf); ws.fm_len.ensure(Mf); ws.fm_a.ensure(Mf);
        XFE_CHECK(launch_seg_compact(ws.tmp_flags.p, ws.tmp_inc.p, M, ws.f_key.p, ws.f_w.p, ws.f_len.p,
                                     k2 >= 0 ? ws.f_a.p : nullptr, ws.fm_key.p, ws.fm_w.p, ws.fm_len.p, ws.fm_a.p,
                                     mgrid, s));
        stay_years = false;
        if (k2 >= 0 && Nw > Nm) {
            const int E = G2 * (G2 + 1);
            ws.stay_h.ensure(E);
            XFE_CHECK(hipMemsetAsync(ws.stay_h.p, 0, sizeof(int) * E, s));
            XFE_CHECK(launch_seg_stayer_hist(ws.f_len.p, ws.f_a.p, M, G2, ws.stay_h.p, mgrid, s));
            std::vector<int> H(E);
            XFE_CHECK(hipMemcpyAsync(H.data(), ws.stay_h.p, sizeof(int) * E, hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
            // K[y][y'] = -sum over stayer segments covering y and y' of 1 / L
            std::vector<double> K(static_cast<size_t>(G2) * G2, 0.0);
            for (int a0 = 0; a0 < G2; ++a0)
                for (int L = 1; a0 + L <= G2; ++L) {
                    const int cnt = H[a0 * (G2 + 1) + L];
                    if (!cnt) continue;
                    const double v = static_cast<double>(cnt) / L;
                    for (int y = a0; y < a0 + L; ++y)
                        for (int t = a0; t < a0 + L; ++t) K[static_cast<size_t>(y) * G2 + t] -= v;
                }
            ws.stay_K.ensure(K.size());
            XFE_CHECK(hipMemcpyAsync(ws.stay_K.p, K.data(), sizeof(double) * K.size(), hipMemcpyHostToDevice, s));
            XFE_CHECK(hipStreamSynchronize(s));
            stay_years = true;
        }
        // firms without a mover segment (their sums must be cleared before each apply)
        ws.w_pf.ensure(std::max(Nw, G1 + 1));
        XFE_CHECK(launch_seg_offsets(ws.fm_key.p, Mf, G1, ws.w_pf.p, grid_for(Mf), s));
        ws.flag.ensure(1);
        XFE_CHECK(hipMemsetAsync(ws.flag.p, 0, sizeof(int), s));
        XFE_CHECK(launch_seg_empty(ws.w_pf.p, G1, ws.flag.p, grid_for(G1), s));
        int nempty = 0;
        XFE_CHECK(hipMemcpyAsync(&nempty, ws.flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        empty_firms = nempty > 0;
        seg_grid_w = grid_for(std::max(Nm, 1));
        seg_grid_f = grid_for(Mf);
        const size_t tabs = static_cast<size_t>(seg_grid_f) * (G2 + 1) * 8;
        ws.f_wmean.ensure(static_cast<size_t>(std::max(Nm, 1)) * 8);
        ws.d_wmean.ensure(static_cast<size_t>(std::max(Nm, 1)) * 8);
        ws.f_hpart.ensure(tabs);
        ws.d_hpart.ensure(tabs);
        ws.seg_tot.ensure(static_cast<size_t>(G2 + 1) * 8);
        const size_t nb = 2 * ((static_cast<size_t>(M) + kWave - 1) / kWave);
        for (int k = 0; k < 2; ++k) {
            ws.seg_bk[k].ensure(nb);
            ws.f_bv[k].ensure(nb * 8);
            ws.d_bv[k].ensure(nb * 8);
        }
        const size_t NC = static_cast<size_t>(N) * C;
        ws.f_b0.ensure(NC);
        ws.f_tmp.ensure(NC);
        return true;
    }

    template <typename T>
    T* bv_buf(int k) {
        if constexpr (std::is_same<T, float>::value) return ws.f_bv[k].p;
        else return ws.d_bv[k].p;
    }

    // Finish the firm reduction: reduce the boundary lists level by level
    // (stride: the row stride of out1); large levels with the full grid, the
    // rest in one block.
    template <typename T>
    void seg_boundaries(T* out1, int stride, int c0, int cc) {
        int m = static_cast<int>(2 * ((static_cast<long long>(Mf) + kWave - 1) / kWave));
        int cur = 0;
        while (m > kBndTail) {
            XFE_CHECK(launch_seg_bnd<T>(ws.seg_bk[cur].p, bv_buf<T>(cur), m, out1, stride, c0, cc,
                                        ws.seg_bk[1 - cur].p, bv_buf<T>(1 - cur), 0, grid_for(m), s));
            m = static_cast<int>(2 * ((static_cast<long long>(m) + kWave - 1) / kWave));
            cur = 1 - cur;
        }
        XFE_CHECK(launch_seg_bnd_tail<T>(ws.seg_bk[cur].p, bv_buf<T>(cur), ws.seg_bk[1 - cur].p,
                                         bv_buf<T>(1 - cur), m, out1, stride, c0, cc, s));
    }

    // out = S x through the segment records (T = float for the inner CG,
    // double for the refinement residual).
    template <typename T>
    const T* inv_movers() const {
        if constexpr (std::is_same<T, float>::value) return ws.f_invwm.p;
        else return ws.d_invwm.p;
    }

    template <typename T>
    void seg_apply(const T* x, T* out, const T* inv_w, T* wmean, double* hpart) {
        // every firm with a segment and every year is written; clear only when some firm has none
        if (empty_firms) XFE_CHECK(hipMemsetAsync(out, 0, sizeof(T) * static_cast<size_t>(N) * C, s));
        for (int ch = 0; ch < sv.nchunks; ++ch) {
            const int c0 = ch * 8;
            const int cc = std::min(8, C - c0);
            SegArgs<T> a;
            a.C = C; a.c0 = c0; a.Nw = Nw; a.M = Mf; a.G2 = G2;
            a.Nm = Nm; a.mlist = ws.m_list.p; a.inv_wm = inv_movers<T>();
            a.x1 = x + static_cast<size_t>(off[k1]) * C;
            a.x2 = k2 >= 0 ? x + static_cast<size_t>(off[k2]) * C : nullptr;
            a.inv_w = inv_w;
            a.wseg = ws.p_wseg.p; a.seg_r1 = ws.p_r1.p; a.seg_len = ws.p_len.p; a.seg_a = ws.p_a.p;
            a.wmean = wmean;
            a.f_key = ws.fm_key.p; a.f_w = ws.fm_w.p; a.f_len = ws.fm_len.p; a.f_a = ws.fm_a.p;
            a.out1 = out + static_cast<size_t>(off[k1]) * C;
            a.hpart = hpart;
            a.bkey = ws.seg_bk[0].p;
            a.bval = bv_buf<T>(0);
            if (Nm > 0) XFE_CHECK(launch_seg_worker<T>(a, cc, k2 >= 0, seg_grid_w, s));
            const int pslot = profile_begin(s);
            XFE_CHECK(launch_seg_firm<T>(a, cc, k2 >= 0, false, seg_grid_f, s));
            if (pslot >= 0) {
                // the firm pass alone: the movers' segment records (key, mover, length,
                // start), one worker mean per segment, the firm coefficients read and
                // written once
                const double e = sizeof(T);
                const double G1 = sv.G[rdim[k1]];
                profile_end(pslot, s,
                            static_cast<double>(Mf) * (12.0 + (k2 >= 0 ? 4.0 : 0.0) + e * cc) + 2.0 * e * cc * G1);
            }
            seg_boundaries<T>(a.out1, C, c0, cc);
            if (k2 >= 0)
                XFE_CHECK(launch_seg_year<T>(hpart, seg_grid_f, G2, cc, ws.seg_tot.p, ws.wsum[rdim[k2]].p,
                                             a.x2, C, c0, stay_years ? ws.stay_K.p : nullptr,
                                             out + static_cast<size_t>(off[k2]) * C, s));
        }
    }

    // Jacobi diagonal from the segments: firms L - L^2 / n_w per mover segment
    // (0 for a stayer's), years n_y - sum of 1 / n_w over the segments covering
    // the year (the stayers' part: the diagonal of stay_K).
    void seg_diag() {
        XFE_CHECK(hipMemsetAsync(ws.f_diag.p, 0, sizeof(double) * N, s));
        SegArgs<double> a;
        a.C = 1; a.c0 = 0; a.Nw = Nw; a.M = Mf; a.G2 = G2;
        a.inv_w = ws.d_invw.p;
        a.Nm = Nm; a.mlist = ws.m_list.p; a.inv_wm = ws.d_invwm.p;
        a.f_key = ws.fm_key.p; a.f_w = ws.fm_w.p; a.f_len = ws.fm_len.p; a.f_a = ws.fm_a.p;
        a.out1 = ws.f_diag.p + off[k1];
        a.hpart = ws.d_hpart.p;
        a.bkey = ws.seg_bk[0].p;
        a.bval = bv_buf<double>(0);
        XFE_CHECK(launch_seg_firm<double>(a, 1, k2 >= 0, true, seg_grid_f, s));
        seg_boundaries<double>(a.out1, 1, 0, 1);
        if (k2 >= 0)
            XFE_CHECK(launch_seg_year<double>(ws.d_hpart.p, seg_grid_f, G2, 1, ws.seg_tot.p,
                                              ws.wsum[rdim[k2]].p, nullptr, 1, 0, stay_years ? ws.stay_K.p : nullptr,
                                              ws.f_diag.p + off[k2], s));
    }

    // Refinement residual b = b0 - S beta in FP64 through the segments.
    void residual_seg() {
        seg_apply<double>(ws.f_beta.p, ws.f_tmp.p, ws.d_invw.p, ws.d_wmean.p, ws.d_hpart.p);
        const int NC = N * C;
        XFE_CHECK(launch_sub(ws.f_b.p, ws.f_b0.p, ws.f_tmp.p, NC, 1, NC, grid_for(NC), s));
    }

    bool setup() {
        if (!sv.tiled || sv.w != nullptr || sv.D < 2 || sv.D > 3 || sv.C > kMaxCgCols) return false;
        p = sv.sorted_dim;
        C = sv.C;
        NR = 0;
        N = 0;
        int lo32 = 0, lo64 = 0;
        for (int d = 0; d < sv.D; ++d) {
            if (d == p) continue;
            rdim[NR] = d;
            off[NR] = N;
            N += sv.G[d];
            const int cc = std::min(8, C);
            const bool lds = static_cast<size_t>(sv.G[d]) * cc * sizeof(double) <= kLdsBudget;
            mode[NR] = lds ? kRedLds : kRedSpell;
            if (!lds && sv.spell[d].S == nullptr) return false;
            if (lds) {
                lds_off[NR] = lo64;  // same element offset for both precisions
                lo32 += sv.G[d] * cc;
                lo64 += sv.G[d] * cc;
            }
            ++NR;
        }
        lds_extra32 = static_cast<size_t>(lo32) * sizeof(float);
        lds_extra64 = static_cast<size_t>(lo64) * sizeof(double);
        const size_t NC = static_cast<size_t>(N) * C;
        ws.f_beta.ensure(NC); ws.f_b.ensure(NC);
        ws.f_diag.ensure(N); ws.f_counts.ensure(N); ws.f_scale.ensure(6 * kMaxCgCols);
        ws.f_x.ensure(NC); ws.f_r.ensure(NC); ws.f_z.ensure(NC); ws.f_p.ensure(NC); ws.f_q.ensure(NC);
        ws.f_minv.ensure(N); ws.f_invp.ensure(sv.G[p]);
        for (int k = 0; k < NR; ++k)
            if (mode[k] == kRedSpell) ws.f_sp[rdim[k]].ensure(static_cast<size_t>(sv.spell[rdim[k]].M) * C);
        vgrid = std::max(1, std::min(2048, grid_for(N)));
        ws.partials.ensure(static_cast<size_t>(std::max(vgrid, sv.tile_grid)) * 2 * kMaxCgCols + 64);
        ws.scalars.ensure(2 * kMaxCgCols);
        ws.cg.ensure(kCgSlots);
        ws.cg_part.ensure(2 * static_cast<size_t>(kMaxGrid) * kMaxCgCols);  // 2 x grid x C partials
        XFE_CHECK(hipMemsetAsync(ws.cg.p, 0, sizeof(double) * kCgSlots, s));
        XFE_CHECK(launch_to_float(ws.inv[p].p, ws.f_invp.p, sv.G[p], grid_for(sv.G[p]), s));
        seg = setup_segments();
        tls_phase.mark("segments");
        // Jacobi preconditioner: diag(S) from the segments, or from runs of each
        // R level inside levels of p
        if (seg) {
            seg_diag();
        } else {
            XFE_CHECK(hipMemsetAsync(ws.f_diag.p, 0, sizeof(double) * N, s));
            for (int k = 0; k < NR; ++k)
                XFE_CHECK(launch_schur_diag(sv.tile_start, sv.num_tiles, sv.gid[p], sv.gid[rdim[k]],
                                            ws.inv[p].p, ws.f_diag.p + off[k], sv.tile_grid, s));
        }
        for (int k = 0; k < NR; ++k) {
            XFE_CHECK(hipMemcpyAsync(ws.f_counts.p + off[k], ws.wsum[rdim[k]].p,
                                     sizeof(double) * sv.G[rdim[k]], hipMemcpyDeviceToDevice, s));
        }
        tls_phase.mark("jacobi diagonal");
        if (seg && precond > 0) setup_blocks();
        tls_phase.mark("block preconditioner");
        XFE_CHECK(launch_diag_invert(ws.f_diag.p, ws.f_counts.p, ws.f_minv.p, N, grid_for(N), s));
        XFE_CHECK(hipMemsetAsync(ws.f_beta.p, 0, sizeof(double) * NC, s));
        return true;
    }

    // The aggregate blocks: firm pairs and their strengths from the worker-major
    // segments, the scores (the strengths, or the learned network's), the
    // handshake aggregation and the block inverses; the firm part of the
    // Jacobi diagonal becomes the exact diagonal (the per-segment formula
    // misses the cross terms of a worker's returns to a firm).
    void setup_blocks() {
        auto t0 = std::chrono::steady_clock::now();
        lp::AggPrecond& ag = ws.agg;
        lp::SegView v;
        v.Nw = Nw;
        v.G1 = sv.G[rdim[k1]];
        v.wseg = ws.p_wseg.p;
        v.r1 = ws.p_r1.p;
        v.len = ws.p_len.p;
        v.inv_w = ws.d_invw.p;
        v.diag = ws.f_diag.p + off[k1];
        v.rows = ws.wsum[rdim[k1]].p;
        if (!ag.build_pairs(v, s)) return;
        tls_phase.mark("  pairs");
        const float* score = ag.st.p;
        if (precond == 2) score = lp::learned_pair_scores(ag, v, lp_params, s);
        tls_phase.mark("  network scores");
        const int cap = std::getenv("XFE_LP_CAP") ? std::atoi(std::getenv("XFE_LP_CAP")) : 16;
        const int rounds = std::getenv("XFE_LP_ROUNDS") ? std::atoi(std::getenv("XFE_LP_ROUNDS")) : 6;
        ag.aggregate(score, cap, rounds, s);
        tls_phase.mark("  aggregation");
        ag.build_blocks(s);
        XFE_CHECK(hipMemcpyAsync(ws.f_diag.p + off[k1], ag.dex.p, sizeof(double) * v.G1, hipMemcpyDeviceToDevice, s));
        XFE_CHECK(hipStreamSynchronize(s));
        blocks = ag.nb > 0;
        bview = ag.view();
        precond_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    // Training data of the learned preconditioner: the error of the iterate
    // recorded at probe_iter against the final solution on the firm levels,
    // and the pairs, graph and features of this panel (host copies).
    void record_probe() {
        lp::ProbeState& P = *probe;
        const int G1 = sv.G[rdim[k1]];
        std::vector<double> fin(static_cast<size_t>(N) * C), itv(static_cast<size_t>(N) * C);
        XFE_CHECK(hipMemcpyAsync(fin.data(), ws.f_beta.p, sizeof(double) * fin.size(), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipMemcpyAsync(itv.data(), P.iter.p, sizeof(double) * itv.size(), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        P.C = C;
        P.F = G1;
        P.inner = inner_total;
        P.err.assign(static_cast<size_t>(G1) * C, 0.0);
        for (int f = 0; f < G1; ++f)
            for (int c = 0; c < C; ++c) {
                const size_t j = static_cast<size_t>(off[k1] + f) * C + c;
                P.err[static_cast<size_t>(f) * C + c] = fin[j] - itv[j];
            }
        lp::AggPrecond& ag = ws.agg;
        lp::SegView v;
        v.Nw = Nw;
        v.G1 = G1;
        v.wseg = ws.p_wseg.p;
        v.r1 = ws.p_r1.p;
        v.len = ws.p_len.p;
        v.inv_w = ws.d_invw.p;
        v.diag = ws.f_diag.p + off[k1];
        v.rows = ws.wsum[rdim[k1]].p;
        if (!ag.build_pairs(v, s)) throw Error("xfe_lp_probe: the panel has no firm pair");
        lp::LearnedGraph gr;
        gr.build(ag, v, s);
        auto down = [&](auto& vec, const auto* src, size_t m) {
            vec.resize(m);
            if (m) XFE_CHECK(hipMemcpyAsync(vec.data(), src, sizeof(vec[0]) * m, hipMemcpyDeviceToHost, s));
        };
        P.Wm = gr.Wm;
        P.E = gr.E;
        P.U = ag.U;
        down(P.xf, gr.xf.p, static_cast<size_t>(G1) * 8);
        down(P.xw, gr.xw.p, static_cast<size_t>(gr.Wm) * 8);
        down(P.f_ptr, gr.f_ptr.p, static_cast<size_t>(G1) + 1);
        down(P.f_idx, gr.f_idx.p, static_cast<size_t>(gr.E));
        down(P.w_ptr, gr.w_ptr.p, static_cast<size_t>(gr.Wm) + 1);
        down(P.w_idx, gr.w_idx.p, static_cast<size_t>(gr.E));
        down(P.pf, ag.pf.p, static_cast<size_t>(ag.U));
        down(P.pg, ag.pg.p, static_cast<size_t>(ag.U));
        down(P.pq, ag.st.p, static_cast<size_t>(ag.U));
        down(P.W, ag.W.p, static_cast<size_t>(ag.U));
        down(P.dex, ag.dex.p, static_cast<size_t>(G1));
        XFE_CHECK(hipStreamSynchronize(s));
        gr.release();
        P.done = true;
    }

    template <typename T>
    void schur_pass(const T* coef, T* out, const T* V, T* dst, T sign, const T* inv_p,
                    T* const* spell_sums, size_t lds_extra, double* partials, T* p_sums = nullptr) {
        XFE_CHECK(hipMemsetAsync(out, 0, sizeof(T) * static_cast<size_t>(N) * C, s));
        for (int ch = 0; ch < sv.nchunks; ++ch) {
            const int c0 = ch * 8;
            const int cc = std::min(8, C - c0);
            SchurArgs<T> a;
            a.n = sv.n; a.ld = sv.ld; a.C = C; a.c0 = c0;
            a.V = V; a.dst = dst;
            a.tile_start = sv.tile_start; a.num_tiles = sv.num_tiles;
            a.gid_p = sv.gid[p]; a.inv_p = inv_p; a.sign = sign;
            for (int k = 0; k < NR; ++k) {
                const int d = rdim[k];
                a.gid_r[k] = sv.gid[d];
                a.coef_r[k] = coef + static_cast<size_t>(off[k]) * C;
                a.G_r[k] = sv.G[d];
                a.mode_r[k] = mode[k];
                a.spell_sums[k] = spell_sums[k];
                a.spell_base[k] = sv.spell[d].base;
                a.lds_target[k] = out + static_cast<size_t>(off[k]) * C;
                a.lds_off[k] = lds_off[k] / 8 * cc;  // element offset for this chunk width
            }
            a.partials = partials ? partials + static_cast<size_t>(ch) * sv.tile_grid : nullptr;
            a.p_sums = p_sums;
            // LDS offsets were laid out for full 8-wide chunks; recompute for cc
            int lo = 0;
            for (int k = 0; k < NR; ++k) {
                if (mode[k] != kRedLds) continue;
                // one copy per wave when the copies stay within 16 KB
                const size_t one = static_cast<size_t>(sv.G[rdim[k]]) * cc * sizeof(T);
                a.lds_copies[k] = one * (kBlock / kWave) <= 16 * 1024 ? kBlock / kWave : 1;
                a.lds_off[k] = lo;
                lo += sv.G[rdim[k]] * cc * a.lds_copies[k];
            }
            const size_t extra = static_cast<size_t>(lo) * sizeof(T);
            (void)lds_extra;
            const int pslot = seg ? -1 : profile_begin(s);
            XFE_CHECK(launch_schur<T>(a, cc, NR, V != nullptr, dst != nullptr, extra, sv.tile_grid, s));
            if (pslot >= 0) {
                const double e = sizeof(T);
                double bytes = 4.0 * sv.n * (1 + NR);                     // ids
                if (V) bytes += e * static_cast<double>(sv.n) * cc;       // data read
                if (dst) bytes += e * static_cast<double>(sv.n) * cc;     // data write
                for (int k = 0; k < NR; ++k) {
                    bytes += e * static_cast<double>(sv.G[rdim[k]]) * cc;  // coefficient gather
                    if (mode[k] == kRedSpell) bytes += e * static_cast<double>(sv.spell[rdim[k]].M) * cc;
                }
                bytes += e * static_cast<double>(sv.G[p]);                // 1 / level sizes of p
                profile_end(pslot, s, bytes);
            }
        }
        for (int k = 0; k < NR; ++k) {
            if (mode[k] != kRedSpell) continue;
            const auto& sp = sv.spell[rdim[k]];
            for (int ch = 0; ch < sv.nchunks; ++ch) {
                const int c0 = ch * 8;
                const int cc = std::min(8, C - c0);
                XFE_CHECK(launch_pull<T>(spell_sums[k], C, c0, cc, sp.order, sp.key, sp.M,
                                         out + static_cast<size_t>(off[k]) * C, grid_for(sp.M), s));
            }
        }
    }

    void apply32(const float* x, float* q) {
        if (seg) {
            seg_apply<float>(x, q, ws.f_invw.p, ws.f_wmean.p, ws.d_hpart.p);
            return;
        }
        float* sp[2] = {nullptr, nullptr};
        for (int k = 0; k < NR; ++k) if (mode[k] == kRedSpell) sp[k] = ws.f_sp[rdim[k]].p;
        schur_pass<float>(x, q, nullptr, nullptr, 1.0f, ws.f_invp.p, sp, lds_extra32, nullptr);
    }

    // b = D_R' M_p (V - D_R beta) in FP64; with dst, also writes M_p (V - D_R beta)
    // and, with p_sums, that block's sums over the levels of p.
    double residual64(double* V, double* dst, double* p_sums = nullptr) {
        double* sp[2] = {nullptr, nullptr};
        for (int k = 0; k < NR; ++k) if (mode[k] == kRedSpell) sp[k] = ws.sp_sums[rdim[k]].p;
        schur_pass<double>(ws.f_beta.p, ws.f_b.p, V, dst, -1.0, ws.inv[p].p, sp, lds_extra64,
                           dst ? ws.partials.p : nullptr, p_sums);
        double sq = 0.0;
        if (dst) sv.fetch(ws.partials.p, sv.tile_grid * sv.nchunks, 1, 0, &sq);
        return sq;
    }

    template <typename T>
    void coldot(const T* a, const T* b, double* out) {
        XFE_CHECK(launch_vdot<T>(a, b, N, C, ws.partials.p, vgrid, s));
        sv.fetch(ws.partials.p, vgrid, C, 0, out);
    }

    void upload(const double* h, double* d, int cnt) {
        for (int j = 0; j < cnt; ++j) ws.host_scalars[j] = h[j];
        XFE_CHECK(hipMemcpyAsync(d, ws.host_scalars, sizeof(double) * cnt, hipMemcpyHostToDevice, s));
        XFE_CHECK(hipStreamSynchronize(s));
    }

    // One inner FP32 PCG solve of S x = b / |b| per column; beta64 += |b| x.
    // The CG scalars live on the device; the host polls the count of active
    // columns every kPoll iterations (iterations after the last column stops
    // are no-ops: alpha = 0).
    int inner_solve(double tol_in, int max_inner) {
        constexpr int kPoll = 8;
        double bb[kMaxCgCols], scale[2 * kMaxCgCols];
        coldot<double>(ws.f_b.p, ws.f_b.p, bb);
        for (int c = 0; c < C; ++c) {
            const double nb = std::sqrt(bb[c]);
            scale[c] = nb > 0.0 ? 1.0 / nb : 0.0;        // rhs scaling into FP32
            scale[kMaxCgCols + c] = nb;                  // and back
        }
        for (int c = C; c < kMaxCgCols; ++c) scale[c] = scale[kMaxCgCols + c] = 0.0;
        upload(scale, ws.f_scale.p, 2 * kMaxCgCols);
        const long long total = static_cast<long long>(N) * C;
        const int g = cg_grid(total, C, kMaxGrid);
        double* cg = ws.cg.p;
        double* part = ws.cg_part.p;
        XFE_CHECK(launch_cg_init<float>(ws.f_b.p, ws.f_scale.p, ws.f_minv.p, ws.f_x.p, ws.f_r.p, ws.f_z.p,
                                        ws.f_p.p, total, C, cg, part, g, s));
        int it = 0;
        for (it = 1; it <= max_inner; ++it) {
            apply32(ws.f_p.p, ws.f_q.p);
            XFE_CHECK(launch_cg_dot<float>(ws.f_p.p, ws.f_q.p, total, C, cg, part, g, s));
            XFE_CHECK(launch_cg_step<float>(ws.f_x.p, ws.f_r.p, ws.f_z.p, ws.f_p.p, ws.f_q.p, ws.f_minv.p, total, C,
                                            cg, part, tol_in * tol_in, max_inner + 1, g, s));
            if (it % kPoll == 0 || it == max_inner) {
                XFE_CHECK(hipMemcpyAsync(ws.host_scalars, cg + kCgCount, sizeof(double), hipMemcpyDeviceToHost, s));
                XFE_CHECK(hipStreamSynchronize(s));
                if (ws.host_scalars[0] == 0.0) break;
            }
        }
        XFE_CHECK(launch_refine_update(ws.f_beta.p, ws.f_x.p, ws.f_scale.p + kMaxCgCols, N, C,
                                       grid_for(total), s));
        inner_total += std::min(it, max_inner);
        return it;
    }

    // The FP64 residual b0 - S beta into f_b (through the segments, or from the
    // rows of V).
    void true_residual(double* V) {
        if (seg) residual_seg();
        else residual64(V, nullptr);
    }

    // Mixed-precision CG with reliable updates (residual replacement; Sleijpen &
    // van der Vorst, Computing 56 (1996); the mixed-precision form of Clark et
    // al., Comput. Phys. Commun. 181 (2010)).  One FP32 Jacobi-PCG recurrence on
    // the Schur system per column, in units of the current residual norm t_c.
    // When the iterated FP32 residual has fallen to delta (or below the target),
    // the FP32 iterate is added to the FP64 beta, the residual b0 - S beta is
    // recomputed in FP64 and replaces the FP32 one, and the recurrence continues
    // with its old search direction (rescaled), so the Krylov space is kept
    // across the FP64 updates instead of restarting the solve (the direction
    // keeps its beta p_prev part and takes the new residual's z).  A column stops
    // when its FP64 residual is at most target |b0_c|, or when four updates in a
    // row each fail to halve it.  Returns the FP64 relative residual (max over
    // columns); iterations and updates through the out arguments.
    // S x in FP64 for any level-major x (the segment operator, or the row pass)
    void apply64(const double* x, double* q) {
        if (seg) {
            seg_apply<double>(x, q, ws.d_invw.p, ws.d_wmean.p, ws.d_hpart.p);
            return;
        }
        double* sp[2] = {nullptr, nullptr};
        for (int k = 0; k < NR; ++k) if (mode[k] == kRedSpell) sp[k] = ws.sp_sums[rdim[k]].p;
        schur_pass<double>(x, q, nullptr, nullptr, 1.0, ws.inv[p].p, sp, lds_extra64, nullptr);
    }

    // precision 2: Jacobi-PCG on S beta = b0 entirely in FP64 (the same
    // elimination and preconditioner as the mixed path's Jacobi variant, so
    // the two differ only in precision); r in f_b, beta in f_beta; a column
    // stops at |r| <= target |b0|.  Returns the worst |r| / |b0|.
    double fp64_solve(const double* b0n, double target, int max_iter, int& iters) {
        const long long total = static_cast<long long>(N) * C;
        const int g = grid_for(total);
        for (DevBuf<double>* b : {&ws.s64_z, &ws.s64_p, &ws.s64_q}) b->ensure(static_cast<size_t>(total));
        ws.s64_coef.ensure(2 * static_cast<size_t>(kMaxCgCols));
        double* r = ws.f_b.p;
        double *z = ws.s64_z.p, *pv = ws.s64_p.p, *q = ws.s64_q.p;
        double* d_alpha = ws.s64_coef.p;
        double* d_beta = ws.s64_coef.p + kMaxCgCols;
        double rz[kMaxCgCols], rz_new[kMaxCgCols], pq[kMaxCgCols], rr[kMaxCgCols], coef[kMaxCgCols];
        bool act[kMaxCgCols];
        XFE_CHECK(launch_s64_jacobi(r, ws.f_minv.p, z, total, C, g, s));
        XFE_CHECK(hipMemcpyAsync(pv, z, sizeof(double) * total, hipMemcpyDeviceToDevice, s));
        coldot<double>(r, z, rz);
        coldot<double>(r, r, rr);
        double worst = 0.0;
        int nact = 0;
        for (int c = 0; c < C; ++c) {
            act[c] = b0n[c] > 0.0 && std::sqrt(std::max(0.0, rr[c])) > target * b0n[c];
            nact += act[c] ? 1 : 0;
        }
        iters = 0;
        for (int it = 1; it <= max_iter && nact > 0; ++it) {
            apply64(pv, q);
            coldot<double>(pv, q, pq);
            for (int c = 0; c < C; ++c) coef[c] = (act[c] && pq[c] > 0.0 && rz[c] > 0.0) ? rz[c] / pq[c] : 0.0;
            upload(coef, d_alpha, C);
            XFE_CHECK(launch_s64_axpy(ws.f_beta.p, pv, d_alpha, total, C, 1.0, g, s));
            XFE_CHECK(launch_s64_axpy(r, q, d_alpha, total, C, -1.0, g, s));
            coldot<double>(r, r, rr);
            XFE_CHECK(launch_s64_jacobi(r, ws.f_minv.p, z, total, C, g, s));
            coldot<double>(r, z, rz_new);
            for (int c = 0; c < C; ++c) {
                if (act[c] && (std::sqrt(std::max(0.0, rr[c])) <= target * b0n[c] || !(pq[c] > 0.0))) {
                    act[c] = false;
                    --nact;
                }
                coef[c] = (act[c] && rz[c] > 0.0) ? rz_new[c] / rz[c] : 0.0;
                rz[c] = rz_new[c];
            }
            upload(coef, d_beta, C);
            XFE_CHECK(launch_s64_dir(pv, z, d_beta, total, C, g, s));
            iters = it;
        }
        for (int c = 0; c < C; ++c)
            if (b0n[c] > 0.0) worst = std::max(worst, std::sqrt(std::max(0.0, rr[c])) / b0n[c]);
        inner_total += iters;
        return worst;
    }

    double reliable_solve(double* V, const double* b0n, double target, double delta, int max_inner, int& iters,
                          int& updates) {
        constexpr int kPoll = 8;
        constexpr int K = kMaxCgCols;
        const long long total = static_cast<long long>(N) * C;
        const int g = cg_grid(total, C, kMaxGrid);
        double* cg = ws.cg.p;
        double* part = ws.cg_part.p;
        double* d_t = ws.f_scale.p + K;        // the units of the FP32 iterate, per column
        double* d_rel = ws.f_scale.p + 2 * K;  // inv_t | f | tol2 | act
        double hb[2 * K], t[K], prev[K];
        int stall[K];
        bool act[K];
        XFE_CHECK(hipMemsetAsync(cg, 0, sizeof(double) * kCgSlots, s));
        coldot<double>(ws.f_b.p, ws.f_b.p, hb);
        double worst = 0.0;
        auto plan = [&](bool first) {
            double rel[4 * K + K];
            std::fill(rel, rel + 5 * K, 0.0);
            int nact = 0;
            for (int c = 0; c < C; ++c) {
                if (act[c]) {
                    rel[c] = 1.0 / t[c];
                    rel[K + c] = first ? 0.0 : prev[c] / t[c];  // old units -> new units (0: fresh start)
                    const double lim = 0.5 * target * b0n[c] / t[c];  // half the target: FP32 drift margin
                    rel[2 * K + c] = std::max(delta * delta, lim * lim);
                    rel[3 * K + c] = 1.0;
                    ++nact;
                }
                rel[4 * K + c] = t[c];
            }
            // d_t and the reset parameters are contiguous: one upload
            upload(rel + 4 * K, d_t, K);
            upload(rel, d_rel, 4 * K);
            return nact;
        };
        for (int c = 0; c < C; ++c) {
            t[c] = std::sqrt(hb[c]);
            prev[c] = t[c];
            stall[c] = 0;
            act[c] = b0n[c] > 0.0 && t[c] > target * b0n[c];
            if (b0n[c] > 0.0) worst = std::max(worst, t[c] / b0n[c]);
        }
        iters = 0;
        updates = 0;
        if (plan(true) == 0) return worst;
        const int foff = seg ? off[k1] : 0;
        XFE_CHECK(launch_cg_reliable_reset(ws.f_b.p, d_rel, ws.f_minv.p, ws.f_x.p, ws.f_r.p, ws.f_z.p, ws.f_p.p,
                                           total, C, cg, part, g, blocks ? &bview : nullptr, foff, s));
        auto one_iteration = [&]() {
            apply32(ws.f_p.p, ws.f_q.p);
            XFE_CHECK(launch_cg_dot<float>(ws.f_p.p, ws.f_q.p, total, C, cg, part, g, s));
            XFE_CHECK(launch_cg_step_rel(ws.f_x.p, ws.f_r.p, ws.f_z.p, ws.f_p.p, ws.f_q.p, ws.f_minv.p, total, C, cg,
                                         part, max_inner + 1, g, blocks ? &bview : nullptr, foff, s));
        };
        // the kPoll iterations between two polls as one HIP graph (about a
        // dozen launches each, all on this stream, no host work and no memset
        // nodes): captured at first use, replayed while the preconditioner is
        // the same; only with the segment operator when every firm has a
        // segment (the row pass and the other case clear with a memset), and
        // not with the probe or the profile
        struct GraphGuard {
            hipGraphExec_t e = nullptr;
            ~GraphGuard() { if (e) (void)hipGraphExecDestroy(e); }
        } gexec;
        bool gexec_blocks = false;
        const bool graphs = seg && !empty_firms && !probe && !g_profile.enabled && !std::getenv("XFE_NO_GRAPH");
        for (int it = 1; it <= max_inner; ++it) {
            if (graphs && (it - 1) % kPoll == 0 && it + kPoll - 1 <= max_inner) {
                if (gexec.e && gexec_blocks != blocks) {
                    XFE_CHECK(hipGraphExecDestroy(gexec.e));
                    gexec.e = nullptr;
                }
                if (!gexec.e) {
                    hipGraph_t graph = nullptr;
                    XFE_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    for (int k = 0; k < kPoll; ++k) one_iteration();
                    XFE_CHECK(hipStreamEndCapture(s, &graph));
                    XFE_CHECK(hipGraphInstantiate(&gexec.e, graph, nullptr, nullptr, 0));
                    XFE_CHECK(hipGraphDestroy(graph));
                    gexec_blocks = blocks;
                }
                XFE_CHECK(hipGraphLaunch(gexec.e, s));
                it += kPoll - 1;
            } else {
                one_iteration();
                if (probe && inner_total + it == probe->probe_iter)
                    XFE_CHECK(lp::launch_probe_iterate(ws.f_beta.p, ws.f_x.p, d_t, N, C, probe->iter.p, s));
            }
            iters = it;
            if (it % kPoll != 0 && it != max_inner) continue;
            XFE_CHECK(hipMemcpyAsync(ws.host_scalars, cg + kCgCount, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
            const bool want = ws.host_scalars[1] != 0.0 || ws.host_scalars[0] == 0.0 || it == max_inner;
            if (!want) continue;
            // reliable update: beta += t_c x; FP64 residual; replace
            XFE_CHECK(launch_refine_update(ws.f_beta.p, ws.f_x.p, d_t, N, C, grid_for(total), s));
            true_residual(V);
            ++updates;
            coldot<double>(ws.f_b.p, ws.f_b.p, hb);
            worst = 0.0;
            bool stalled = false;
            for (int c = 0; c < C; ++c) {
                const double tn = std::sqrt(hb[c]);
                if (b0n[c] > 0.0) worst = std::max(worst, tn / b0n[c]);
                if (!act[c]) continue;
                if (tn <= target * b0n[c] || !(tn > 0.0)) {
                    act[c] = false;
                } else {
                    stall[c] = tn > 0.5 * t[c] ? stall[c] + 1 : 0;
                    if (stall[c] >= 4) {
                        act[c] = false;
                        stalled = true;
                    }
                }
                prev[c] = t[c];
                t[c] = tn > 0.0 ? tn : t[c];
            }
            bool fresh = false;
            // XFE_LP_STALL_TEST: the host declares the block preconditioner stalled at
            // its first residual replacement (tests of the fallback; never set otherwise)
            if (blocks && std::getenv("XFE_LP_STALL_TEST")) stalled = true;
            if (stalled && blocks) {
                // the block preconditioner stalled: Jacobi-PCG continues from beta (fresh directions)
                blocks = false;
                block_fallback = true;
                fresh = true;
                for (int c = 0; c < C; ++c) {
                    stall[c] = 0;
                    act[c] = b0n[c] > 0.0 && t[c] > target * b0n[c];
                }
            }
            if (it == max_inner || plan(fresh) == 0) break;
            XFE_CHECK(launch_cg_reliable_reset(ws.f_b.p, d_rel, ws.f_minv.p, ws.f_x.p, ws.f_r.p, ws.f_z.p, ws.f_p.p,
                                               total, C, cg, part, g, blocks ? &bview : nullptr, foff, s));
        }
        inner_total += iters;
        return worst;
    }
};

static int C_of(const Solver& sv) { return sv.C; }

static bool solve_fast(Solver& sv, double* V, const xfe_options& opt, xfe_absorb_result& res) {
    FastSolver fs(sv);
    fs.precond = opt.preconditioner;
    fs.lp_params = opt.lp_params;
    lp::ProbeState& probe = lp::probe_state();
    if (probe.active) fs.precond = 0;
    if (!fs.setup()) return false;
    if (probe.active && fs.seg) {
        fs.probe = &probe;
        probe.iter.ensure(static_cast<size_t>(fs.N) * C_of(sv));
        XFE_CHECK(hipMemsetAsync(probe.iter.p, 0, sizeof(double) * static_cast<size_t>(fs.N) * C_of(sv), sv.s));
    }
    const int C = sv.C;
    const double target = opt.refine_tol > 0.0 ? opt.refine_tol
                                               : std::min(3e-13, (opt.tol > 0.0 ? opt.tol : 1e-8) * 3e-5);
    res.refine_target = target;
    const double tol_in = std::getenv("XFE_FAST_INNER") ? std::atof(std::getenv("XFE_FAST_INNER")) : 1e-5;
    const int max_inner = std::max(10, std::min(opt.max_iter, 5000));
    double b0[kMaxCgCols];
    fs.residual64(V, nullptr);
    if (fs.seg) {
        XFE_CHECK(hipMemcpyAsync(sv.ws.f_b0.p, sv.ws.f_b.p, sizeof(double) * static_cast<size_t>(fs.N) * C,
                                 hipMemcpyDeviceToDevice, sv.s));
    }
    fs.coldot<double>(sv.ws.f_b.p, sv.ws.f_b.p, b0);
    tls_phase.mark("b0 = D_R' M_p V");
    double b0n[kMaxCgCols];
    for (int c = 0; c < C; ++c) b0n[c] = std::sqrt(b0[c]);
    const double delta = std::getenv("XFE_RELIABLE_DELTA") ? std::atof(std::getenv("XFE_RELIABLE_DELTA")) : 0.01;
    int iters = 0, updates = 0;
    const double rel =
        opt.precision == 2
            ? fs.fp64_solve(b0n, target, std::max(10, std::min(opt.max_iter, 20000)), iters)
            : fs.reliable_solve(V, b0n, target, delta, std::max(10, std::min(opt.max_iter, 20000)), iters, updates);
    const int refinements = updates;
    const bool converged = rel <= std::max(target, 1e-9);
    (void)tol_in;
    (void)max_inner;
    tls_phase.mark("PCG");
    if (fs.probe) fs.record_probe();
    // final within transform, in place; its pass also leaves D_R' v~ in f_b and,
    // for the certificate, the sums of v~ over the levels of p
    double* psum = nullptr;
    if (opt.certify) {
        sv.ws.cert_psum.ensure(static_cast<size_t>(sv.G[fs.p]) * C);
        XFE_CHECK(hipMemsetAsync(sv.ws.cert_psum.p, 0, sizeof(double) * static_cast<size_t>(sv.G[fs.p]) * C, sv.s));
        psum = sv.ws.cert_psum.p;
    }
    fs.residual64(V, V, psum);
    tls_phase.mark("final transform");
    auto& rec = sv.ws.fast_rec;
    rec.cert_ready = false;
    if (psum) {
        // the certificate's squared level sums (unweighted: the mixed path takes
        // no weights): every level of the R dimensions from f_b, of p from psum
        rec.cert_sq.assign(C, 0.0);
        std::vector<double> tmp(C, 0.0);
        auto add = [&](const double* t, long long total) {
            const int g = cg_grid(total, C, kMaxGrid);
            sv.ws.partials.ensure(static_cast<size_t>(g) * C);
            XFE_CHECK(launch_table_sumsq(t, total, C, sv.ws.partials.p, g, sv.s));
            sv.fetch(sv.ws.partials.p, g, C, 0, tmp.data());
            for (int c = 0; c < C; ++c) rec.cert_sq[c] += tmp[c];
        };
        add(sv.ws.f_b.p, static_cast<long long>(fs.N) * C);
        add(psum, static_cast<long long>(sv.G[fs.p]) * C);
        rec.cert_V = V;
        rec.cert_ready = true;
    }
    rec.used = true;
    rec.p = fs.p;
    rec.NR = fs.NR;
    rec.C = C;
    for (int k = 0; k < 2; ++k) { rec.rdim[k] = fs.rdim[k]; rec.off[k] = fs.off[k]; }
    res.iterations = fs.inner_total;
    res.refinements = refinements;
    res.preconditioner_used = fs.blocks || fs.block_fallback ? fs.precond : 0;
    res.precond_fallback = fs.block_fallback ? 1 : 0;
    res.precond_blocks = fs.blocks || fs.block_fallback ? sv.ws.agg.nb : 0;
    res.precond_ms = fs.precond_ms;
    res.converged = converged ? 1 : 0;
    res.used_accel = 0;
    res.used_cg = 1;
    res.sweeps = 2 * refinements + 2;
    return true;
}

// Core solve on device buffers already in the workspace layout.  V: n x C,
// ld; gids[d] already permuted; ws.wsum/inv filled.  Returns 0 or throws.
static void solve_core(Solver& sv, double* V, const std::vector<int>& order_in,
                       const std::vector<ProfileStats>& profs, const xfe_options& opt,
                       xfe_absorb_result& res) {
    Workspace& ws = sv.ws;
    const int n = sv.n, C = sv.C, ld = sv.ld, D = sv.D;
    const double* w = sv.w;
    std::vector<int> G(sv.G, sv.G + D);
    for (int d = 0; d < D; ++d) res.sweep_order_used[d] = order_in[d];
    if (opt.precision == 1 || opt.precision == 2) {
        if (solve_fast(sv, V, opt, res)) {
            res.precision_used = opt.precision;
            return;
        }
        res.precision_fallback = 1;
    }
    res.precision_used = 0;
    const bool honest = opt.tolerance_mode == 1;
    const double tol_eff = (opt.tol > 0.0 && opt.tolerance_mode == 1) ? std::min(opt.tol, 1e-9) : opt.tol;
    std::vector<int> order = order_in;
    for (int d = 0; d < D; ++d) res.sweep_order_used[d] = order[d];
    const bool symmetric = opt.method == 2 && D > 1;
    const std::vector<int> seq = sweep_sequence(order, symmetric);

    // gates: src/fe_absorption_cuda.cu:2346-2390 (mirror of the CPU gates)
    double min_occ = -1.0;
    for (int d = 0; d < D; ++d)
        if (profs[d].num_groups > 0 && (min_occ < 0.0 || profs[d].mean < min_occ)) min_occ = profs[d].mean;
    const bool ill = D >= 2 && min_occ >= 0.0 && min_occ <= 12.0;
    int total_levels = 0;
    for (int d = 0; d < D; ++d) total_levels += G[d];
    bool use_cg = false;
    if (w == nullptr) {
        if (opt.force_cg && D >= 2) use_cg = true;
        if (opt.tolerance_mode == 1 && D >= 3) {
            double share = 0.0;
            for (int d = 0; d < D; ++d) share = std::max(share, profs[d].max_val / static_cast<double>(n));
            if (share >= 0.37) use_cg = true;
        }
        if (D == 2 && n <= 20000 && total_levels >= n - 2 && total_levels <= n + 2) use_cg = true;
    }
    const bool use_accel = use_cg || (D >= 2 && (n >= 200000 || ill));
    const int budget = (use_accel && n < 200000) ? std::max(1, (opt.max_iter * 2) / 5) : opt.max_iter;

    // two-stage warm-up (src/fe_absorption.cpp:6293-6331)
    int iter_warmup = 0, iter_two = 0;
    std::vector<int> two_order;
    if (D >= 3 && n >= 200000) {
        int first = -1, second = -1, third = -1;
        for (int d = 0; d < D; ++d) {
            const int g = profs[d].num_groups;
            if (g > first) { third = second; second = first; first = g; }
            else if (g > second) { third = second; second = g; }
            else if (g > third) { third = g; }
        }
        if (first > 0 && third > 0 && static_cast<double>(third) <= 0.05 * first) {
            double msing = 0.0, mcv = 0.0;
            for (int d = 0; d < D; ++d) { msing = std::max(msing, profs[d].singleton_ratio); mcv = std::max(mcv, profs[d].cv); }
            int diff;
            if (D >= 4 && (msing >= 0.30 || mcv >= 2.0)) diff = 3;
            else if (D >= 4 || msing >= 0.20 || mcv >= 1.5) diff = 2;
            else if (D >= 3 || msing >= 0.10 || mcv >= 1.0) diff = 1;
            else diff = 0;
            if (diff == 3) { iter_warmup = 5; iter_two = 5; }
            else if (diff >= 2) { iter_warmup = 8; iter_two = 8; }
            else { iter_warmup = 5; iter_two = 8; }
            int best = -1, sec = -1, bg = -1, sg = -1;
            for (int d = 0; d < D; ++d) {
                const int g = profs[d].num_groups;
                if (g > bg) { sec = best; sg = bg; best = d; bg = g; }
                else if (g > sg) { sec = d; sg = g; }
            }
            for (int d : order) if (d == best || d == sec) two_order.push_back(d);
            if (two_order.size() != 2) { iter_warmup = 0; iter_two = 0; two_order.clear(); }
        }
    }

    const size_t vec_bytes = sizeof(double) * static_cast<size_t>(ld) * C;
    for (int j = 1; j < 7; ++j) ws.vec[j].ensure(static_cast<size_t>(ld) * C);
    bool converged = false;
    int sweeps = 0;
    double* cur = V;
    if (use_accel) {
        res.used_accel = 1;
        AccelState st{V, ws.vec[1].p, ws.vec[2].p, ws.vec[3].p, ws.vec[4].p, ws.vec[5].p};
        if (use_cg) {
            res.used_cg = 1;
            converged = run_cg(sv, V, ws.vec[1].p, ws.vec[2].p, ws.vec[3].p, order, opt.max_iter,
                               opt.tol, res.iterations, sweeps);
        } else {
            int remaining = budget, total = 0, used = 0;
            const std::vector<int> seq2 = sweep_sequence(two_order, symmetric);
            if (iter_warmup > 0 && remaining > 0) {
                converged = accel_phase(sv, st, seq, std::min(iter_warmup, remaining), honest,
                                        opt.tol, tol_eff, used, sweeps);
                total += used; remaining -= used;
            }
            if (!converged && !st.diverged && !two_order.empty() && iter_two > 0 && remaining > 0) {
                accel_phase(sv, st, seq2, std::min(iter_two, remaining), honest, opt.tol, tol_eff,
                            used, sweeps);
                total += used; remaining -= used;
            }
            if (!converged && !st.diverged && remaining > 0) {
                converged = accel_phase(sv, st, seq, remaining, honest, opt.tol, tol_eff, used,
                                        sweeps);
                total += used; remaining -= used;
            }
            cur = st.cur;
            if (!converged && st.diverged) {
                res.used_cg = 1;
                // CG needs three work vectors distinct from cur
                std::vector<double*> spare;
                for (int j = 0; j < 7 && spare.size() < 3; ++j) {
                    double* p = j == 0 ? V : ws.vec[j].p;
                    if (p != cur) spare.push_back(p);
                }
                converged = run_cg(sv, cur, spare[0], spare[1], spare[2], order, opt.max_iter,
                                   opt.tol, res.iterations, sweeps);
            } else if (converged) {
                res.iterations = total;
            }
        }
    } else {
        // plain MAP, src/fe_absorption.cpp:7368-7412
        double prev = std::sqrt(sv.sumsq(V, -1));
        int last_check = -1;
        for (int iter = 0; iter < opt.max_iter; ++iter) {
            double sq = 0.0;
            sv.sweep2(SweepSrc(V), V, seq, 1, nullptr, nullptr, &sq);
            ++sweeps;
            const double curr = std::sqrt(sq);
            const double rel = std::fabs(curr - prev) / std::max(1.0, prev);
            const int step = iter - last_check;
            last_check = iter;
            prev = curr;
            if (step > 0 && rel / static_cast<double>(step) < tol_eff) {
                res.iterations = iter + 1;
                converged = true;
                break;
            }
        }
    }
    if (!converged) {
        res.iterations = opt.max_iter;
        res.converged = 0;
    } else {
        res.converged = 1;
        if (res.iterations == 0) res.iterations = 1;
        // tightening polish, src/fe_absorption.cpp:7016-7068
        const bool strict = opt.tolerance_mode == 2;
        const int max_polish = strict ? std::max(0, opt.max_iter - res.iterations) : 6;
        const double polish_tol = strict ? opt.tol : (tol_eff > 0.0 ? tol_eff * 0.01 : tol_eff);
        double final_max = 0.0;
        int done = 0;
        for (int p = 0; p < max_polish; ++p) {
            // one sweep in place; the largest |mean| over its projections
            double mx = 0.0;
            sv.sweep2(SweepSrc(cur), cur, seq, 0, nullptr, nullptr, nullptr, &mx);
            ++sweeps;
            final_max = mx;
            ++done;
            if (mx <= polish_tol) break;
        }
        res.polish_sweeps = done;
        res.polish_final_max = final_max;
    }
    if (cur != V) XFE_CHECK(hipMemcpyAsync(V, cur, vec_bytes, hipMemcpyDeviceToDevice, ws.stream));
    res.sweeps = sweeps;
}

// select_sweep_order (src/fe_absorption.cpp:5396-5470), on the device: for
// 2 <= D <= 4 dimensions try every permutation for two Gauss-Seidel passes over
// the first min(n, 200000) rows of y (caller's row order, weight sums over those
// rows) and keep the first permutation with the smallest remaining norm.
static std::vector<int> select_sweep_order_device(Workspace& ws, const double* d_y, int n, int C,
                                                  const std::vector<const int*>& gids,
                                                  const std::vector<int>& G, const double* w) {
    const int D = static_cast<int>(G.size());
    std::vector<int> order(D);
    for (int d = 0; d < D; ++d) order[d] = d;
    if (D <= 1 || D > 4) return order;
    const int np = std::min(n, 200000);
    if (np <= 0) return order;
    const int ldp = ((np + 63) / 64) * 64;
    hipStream_t s = ws.stream;
    Solver sv(ws, np, 1, ldp);
    sv.D = D;
    sv.w = w;
    (void)C;
    for (int d = 0; d < D; ++d) {
        sv.G[d] = G[d];
        sv.gid[d] = gids[d];
    }
    for (int d = 0; d < D; ++d) {
        sv.prepare_dim(d);
        sv.weight_sums(d);
    }
    ws.partials.ensure(static_cast<size_t>(sv.grid) * 3 + 1024);
    ws.scalars.ensure(8);
    ws.vec[6].ensure(ldp);
    double* tmp = ws.vec[6].p;
    std::vector<int> perm = order;
    double best = INFINITY;
    do {
        XFE_CHECK(hipMemcpyAsync(tmp, d_y, sizeof(double) * np, hipMemcpyDeviceToDevice, s));
        std::vector<int> seq(perm);
        seq.insert(seq.end(), perm.begin(), perm.end());  // two passes
        double sq = 0.0;
        sv.sweep(tmp, tmp, seq, false, -1, 1, nullptr, nullptr, &sq);
        const double norm = std::sqrt(sq);
        if (norm < best) {
            best = norm;
            order = perm;
        }
    } while (std::next_permutation(perm.begin(), perm.end()));
    return order;
}

// Level-aligned tiles of the sorted dimension p (tile_kernels.hip) and, for
// every other dimension too large for LDS, its spells: maximal runs of equal
// level inside a level of p, with their order by level for the pull.
// counts_known: ws.wsum[p] holds the level sizes (unit weights, sums computed here)
static void setup_tiles(Solver& sv, Workspace& ws, int n, int D, const std::vector<int>& G, bool counts_known) {
    sv.tiled = false;
    const int p = sv.sorted_dim;
    if (p < 0 || n < kTileRows || std::getenv("XFE_NO_TILES")) return;
    hipStream_t s = ws.stream;
    const int grid = grid_for(n);
    ws.level_count.ensure(static_cast<size_t>(G[p]) + 1);
    if (counts_known) {
        // unit weights: the level sizes are the weight sums already on the device
        XFE_CHECK(hipMemsetAsync(ws.level_count.p + G[p], 0, sizeof(int), s));
        XFE_CHECK(launch_max_count(ws.wsum[p].p, G[p], ws.level_count.p + G[p], grid_for(G[p]), s));
    } else {
        XFE_CHECK(hipMemsetAsync(ws.level_count.p, 0, sizeof(int) * (static_cast<size_t>(G[p]) + 1), s));
        XFE_CHECK(launch_level_count(sv.gid[p], n, ws.level_count.p, grid, s));
        XFE_CHECK(launch_max_int(ws.level_count.p, G[p], ws.level_count.p + G[p], grid_for(G[p]), s));
    }
    int lmax = 0;
    XFE_CHECK(hipMemcpyAsync(&lmax, ws.level_count.p + G[p], sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    if (lmax <= 0 || lmax > kTileRows / 2) return;  // a level too large for one tile
    const int step = kTileRows - lmax;
    const int tiles = static_cast<int>((static_cast<long long>(n) + step - 1) / step);
    ws.tile_start.ensure(static_cast<size_t>(tiles) + 1);
    XFE_CHECK(launch_tile_start(sv.gid[p], n, step, tiles, ws.tile_start.p, grid_for(tiles + 1), s));
    sv.tile_start = ws.tile_start.p;
    sv.num_tiles = tiles;
    sv.tile_grid = std::max(1, std::min(tiles, 2048));
    for (int d = 0; d < D; ++d) {
        if (d == p || sv.tile_mode_of(d) != kRedSpell) continue;
        // spell heads -> ids -> keys; sort spells by level of d
        ws.tmp_flags.ensure(n);
        ws.tmp_inc.ensure(n);
        XFE_CHECK(launch_spell_heads(sv.gid[p], sv.gid[d], n, ws.tmp_flags.p, grid, s));
        size_t bytes = 0;
        XFE_CHECK(rocprim::inclusive_scan(nullptr, bytes, ws.tmp_flags.p, ws.tmp_inc.p,
                                          static_cast<size_t>(n), rocprim::plus<int>(), s));
        ws.sort_tmp.ensure(bytes);
        XFE_CHECK(rocprim::inclusive_scan(ws.sort_tmp.p, bytes, ws.tmp_flags.p, ws.tmp_inc.p,
                                          static_cast<size_t>(n), rocprim::plus<int>(), s));
        int M = 0;
        XFE_CHECK(hipMemcpyAsync(&M, ws.tmp_inc.p + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        ws.tmp_key.ensure(M);
        ws.tmp_iota.ensure(M);
        ws.sp_key[d].ensure(M);
        ws.sp_order[d].ensure(M);
        ws.sp_base[d].ensure(tiles);
        ws.sp_sums[d].ensure(static_cast<size_t>(M) * sv.C);
        XFE_CHECK(launch_spell_keys(ws.tmp_flags.p, ws.tmp_inc.p, sv.gid[d], n, ws.tmp_key.p, grid, s));
        XFE_CHECK(launch_tile_spell_base(ws.tile_start.p, tiles, ws.tmp_inc.p, n, ws.sp_base[d].p,
                                         grid_for(tiles), s));
        XFE_CHECK(launch_iota(ws.tmp_iota.p, M, grid_for(M), s));
        int bits = 1;
        while ((1LL << bits) < static_cast<long long>(G[d])) ++bits;
        bytes = 0;
        XFE_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, ws.tmp_key.p, ws.sp_key[d].p,
                                            ws.tmp_iota.p, ws.sp_order[d].p, static_cast<size_t>(M),
                                            0, bits, s));
        ws.sort_tmp.ensure(bytes);
        XFE_CHECK(rocprim::radix_sort_pairs(ws.sort_tmp.p, bytes, ws.tmp_key.p, ws.sp_key[d].p,
                                            ws.tmp_iota.p, ws.sp_order[d].p, static_cast<size_t>(M),
                                            0, bits, s));
        sv.spell[d].M = M;
        sv.spell[d].S = ws.sp_sums[d].p;
        sv.spell[d].order = ws.sp_order[d].p;
        sv.spell[d].key = ws.sp_key[d].p;
        sv.spell[d].base = ws.sp_base[d].p;
    }
    XFE_CHECK(hipStreamSynchronize(s));
    sv.tiled = true;
}

// Per column, ||D'W v_c||^2 over all dimensions: the squared level sums of
// w * v (the fused reduce pass into level sums, then per-column sums of
// squares of the level-major G x C table), summed over the dimensions.  With
// keep_dim >= 0, level 0 of every other dimension is left out.
static std::vector<double> normal_sq(Solver& sv, const double* V, int keep_dim = -1) {
    Workspace& ws = sv.ws;
    const int C = sv.C;
    std::vector<double> out(C, 0.0), tmp(C, 0.0);
    for (int d = 0; d < sv.D; ++d) {
        sv.clear_sums(d);
        sv.pass(V, nullptr, nullptr, nullptr, 0.0, nullptr, -1, d, 0, nullptr, nullptr, nullptr);
        const long long total = static_cast<long long>(sv.G[d]) * C;
        XFE_CHECK(launch_finalize(ws.sums[d].p, ws.means[d].p, nullptr, sv.G[d], C, sv.copies_of(d, C),
                                  nullptr, grid_for(total), sv.s));
        const int g = cg_grid(total, C, kMaxGrid);
        ws.partials.ensure(static_cast<size_t>(g) * C);
        XFE_CHECK(launch_table_sumsq(ws.means[d].p, total, C, ws.partials.p, g, sv.s));
        sv.fetch(ws.partials.p, g, C, 0, tmp.data());
        for (int c = 0; c < C; ++c) out[c] += tmp[c];
        if (keep_dim >= 0 && d != keep_dim) {
            XFE_CHECK(hipMemcpyAsync(tmp.data(), ws.means[d].p, sizeof(double) * C, hipMemcpyDeviceToHost, sv.s));
            XFE_CHECK(hipStreamSynchronize(sv.s));
            for (int c = 0; c < C; ++c) out[c] -= tmp[c] * tmp[c];
        }
    }
    for (int c = 0; c < C; ++c) out[c] = std::max(0.0, out[c]);
    return out;
}

// Normal-equation certificate of the transformed block Vt (n x C, ld), the
// device form of certify_absorption_result (src/fe_absorption.cpp:14942-15190)
// for standard FEs: per column, the squared level sums of w * vt summed over
// all dimensions (the fused reduce pass into level sums, then per-column sums
// of squares of the G x C table); ||D'W||_F^2 = D sum_i w_i^2; orig_sq holds
// the per-column sums of squares of the untransformed columns.
static void certificate(Solver& sv, const double* Vt, const std::vector<double>& orig_sq,
                        const xfe_options& opt, int converged, double& abs_out, double& rel_out,
                        int& certified) {
    Workspace& ws = sv.ws;
    const int C = sv.C;
    const bool prof = g_profile.enabled;
    g_profile.enabled = false;  // the profile covers the solver's passes only
    // the mixed path's transform pass already summed its output over every
    // level of every dimension; otherwise one reduce pass per dimension
    const auto& rec = ws.fast_rec;
    const std::vector<double> resid =
        (rec.used && rec.cert_ready && rec.cert_V == Vt && !sv.w) ? rec.cert_sq : normal_sq(sv, Vt);
    g_profile.enabled = prof;
    double wsq = static_cast<double>(sv.n);
    if (sv.w) {
        XFE_CHECK(launch_col_dot(sv.w, sv.w, sv.n, 1, sv.n, ws.partials.p, sv.grid, sv.s));
        sv.fetch(ws.partials.p, sv.grid, 1, 0, &wsq);
    }
    const double op = std::sqrt(std::max(0.0, sv.D * wsq));
    double amax = 0.0, rmax = 0.0;
    for (int c = 0; c < C; ++c) {
        const double a = std::sqrt(std::max(0.0, resid[c]));
        const double scale = op * std::sqrt(std::max(0.0, orig_sq[c]));
        const double rel = scale > 0.0 ? a / scale : (a > 0.0 ? DBL_MAX : 0.0);
        amax = std::max(amax, a);
        rmax = std::max(rmax, rel);
    }
    const double tol_eff = (opt.tol > 0.0 && opt.tolerance_mode == 1) ? std::min(opt.tol, 1e-9) : opt.tol;
    const double limit = std::max(8.0 * std::max(0.0, tol_eff), 64.0 * DBL_EPSILON);
    abs_out = amax;
    rel_out = rmax;
    certified = (converged && rmax <= limit) ? 1 : 0;
}

// per-column sums of squares of an n x C block (ld)
static std::vector<double> column_sumsq(Solver& sv, const double* V, int ld) {
    std::vector<double> out(sv.C, 0.0);
    sv.ws.partials.ensure(static_cast<size_t>(sv.grid) * sv.C);
    XFE_CHECK(launch_col_dot(V, V, sv.n, sv.C, ld, sv.ws.partials.p, sv.grid, sv.s));
    sv.fetch(sv.ws.partials.p, sv.grid, sv.C, 0, out.data());
    return out;
}

// The stopping rule of schwarz_demean_raw (src/schwarz_demean.cpp:283-378),
// checked on the result: every column with ||D'v_c|| > 1e-300 must reach
// ||D'(v_c - D u_c)|| / ||D'v_c|| < tol, the norms leaving out level 0 of
// every dimension but the one with the most levels (the reference pins those
// coefficients, :224-225, :283-284).
struct SchwarzRule {
    double tol = 0.0;
    bool converged = false;
    double achieved = INFINITY;  // max over active columns of the rule's value
    int polish = 0;              // FP64 sweeps taken after the solve
};

// Solve under the schwarz rule: the mixed solver runs to a Schur refinement
// target below tol and the rule is evaluated with the level-sum pass.  If it
// does not hold -- the Schur residual b - S beta is evaluated to about
// eps ||S|| ||beta||, which on ill-conditioned panels leaves the result near
// 1e-12 of ||D'v|| -- the plain FP64 solver, with its strict-residual stopping
// rule, continues from the transformed block (its within transform is the
// same) on the rows themselves, followed by up to kSchwarzPolish projection
// sweeps while the rule still fails.  Columns with ||D'v_c|| <= 1e-300 keep their input
// values, as in the reference.
constexpr int kSchwarzPolish = 8;

static void solve_schwarz(Solver& sv, double* V, const double* d_Vin, int ld_in, const int* perm,
                          const std::vector<int>& order, const std::vector<ProfileStats>& profs,
                          const xfe_options& opt, xfe_absorb_result& res, SchwarzRule& rule) {
    const int C = sv.C, n = sv.n, ld = sv.ld;
    const int grid = grid_for(n);
    int keep = 0;  // the reference's free dimension: the first with the most levels (:65-69)
    for (int d = 1; d < sv.D; ++d)
        if (sv.G[d] > sv.G[keep]) keep = d;
    const std::vector<double> bn = normal_sq(sv, V, keep);
    std::vector<char> active(C, 0);
    int nactive = 0;
    for (int c = 0; c < C; ++c) {
        if (!std::isfinite(bn[c]) || bn[c] < 0.0) return;  // not converged
        active[c] = std::sqrt(bn[c]) > 1e-300;
        nactive += active[c];
    }
    if (nactive == 0) {
        rule.converged = true;
        return;
    }
    if (opt.max_iter <= 0) return;
    // the rule's value on the current block (INFINITY if not finite)
    auto evaluate = [&]() {
        const std::vector<double> rn = normal_sq(sv, V, keep);
        double worst = 0.0;
        for (int c = 0; c < C; ++c) {
            if (!active[c]) continue;
            const double rel = std::sqrt(rn[c]) / std::sqrt(bn[c]);
            if (!std::isfinite(rel)) return static_cast<double>(INFINITY);
            worst = std::max(worst, rel);
        }
        return worst;
    };
    xfe_options o = opt;
    o.refine_tol = std::min(1e-11, 0.25 * rule.tol);
    solve_core(sv, V, order, profs, o, res);
    rule.achieved = evaluate();
    if (std::isfinite(rule.achieved) && rule.achieved >= rule.tol) {
        xfe_options p = opt;
        p.precision = 0;
        p.tolerance_mode = 2;  // strict-residual stopping
        p.tol = 0.1 * rule.tol;
        p.max_iter = std::min(opt.max_iter, 2000);
        p.certify = 0;
        xfe_absorb_result r2{};
        solve_core(sv, V, order, profs, p, r2);
        res.iterations += r2.iterations;
        rule.achieved = evaluate();
        const std::vector<int> seq = sweep_sequence(order, false);
        for (int sweep = 0; sweep < kSchwarzPolish && std::isfinite(rule.achieved) && rule.achieved >= rule.tol;
             ++sweep) {
            sv.sweep2(SweepSrc(V), V, seq, 0, nullptr, nullptr, nullptr);
            ++rule.polish;
            rule.achieved = evaluate();
        }
    }
    rule.converged = std::isfinite(rule.achieved) && rule.achieved < rule.tol;
    if (rule.converged && nactive < C) {
        for (int c = 0; c < C; ++c)
            if (!active[c])
                XFE_CHECK(launch_gather_cols(d_Vin + static_cast<size_t>(c) * ld_in, ld_in, perm,
                                             V + static_cast<size_t>(c) * ld, ld, n, 1, grid, sv.s));
    }
}

// Shared driver for the entry points.  d_Vin: caller's n x C (ld_in) block,
// d_gid_in[d]: caller's ids; the result is written back to d_Vin.  With a
// schwarz rule the solve follows solve_schwarz.
static void absorb_device_impl(double* d_Vin, int n, int C, int ld_in,
                               const std::vector<const int*>& d_gid_in, const std::vector<int>& G,
                               const std::vector<const double*>& h_wsum, const double* d_w_in,
                               const int* sweep_order, const xfe_options& opt,
                               xfe_absorb_result& res, SchwarzRule* rule = nullptr) {
    Workspace& ws = tls_ws;
    ws.init();
    ws.fast_rec.used = false;
    ws.fast_rec.cert_ready = false;
    hipStream_t s = ws.stream;
    const int D = static_cast<int>(G.size());
    const auto t0 = std::chrono::steady_clock::now();
    tls_phase.start(s);
    for (int d = 0; d < D; ++d) res.fe_levels[d] = G[d];
    const int ld = ((n + 63) / 64) * 64;
    const int grid = grid_for(n);
    std::vector<int> probe_order(D);
    for (int d = 0; d < D; ++d) probe_order[d] = d;
    if (!sweep_order && D == 2) {
        probe_order = select_sweep_order_device(ws, d_Vin, n, C, d_gid_in, G, d_w_in);
    }
    // primary (sorted) dimension: the one with the most levels
    int p = 0;
    for (int d = 1; d < D; ++d) if (G[d] > G[p]) p = d;
    ws.flag.ensure(1);
    XFE_CHECK(hipMemsetAsync(ws.flag.p, 0, sizeof(int), s));
    XFE_CHECK(launch_sorted_check(d_gid_in[p], n, ws.flag.p, grid, s));
    int unsorted = 0;
    XFE_CHECK(hipMemcpyAsync(&unsorted, ws.flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    const int* perm = nullptr;
    if (unsorted) {
        ws.perm.ensure(n);
        ws.keys_tmp.ensure(n);
        ws.vals_tmp.ensure(n);
        XFE_CHECK(launch_iota(ws.vals_tmp.p, n, grid, s));
        int bits = 1;
        while ((1LL << bits) < static_cast<long long>(G[p])) ++bits;
        size_t tmp_bytes = 0;
        XFE_CHECK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, d_gid_in[p], ws.keys_tmp.p,
                                            ws.vals_tmp.p, ws.perm.p, static_cast<size_t>(n), 0,
                                            bits, s));
        ws.sort_tmp.ensure(tmp_bytes);
        XFE_CHECK(rocprim::radix_sort_pairs(ws.sort_tmp.p, tmp_bytes, d_gid_in[p], ws.keys_tmp.p,
                                            ws.vals_tmp.p, ws.perm.p, static_cast<size_t>(n), 0,
                                            bits, s));
        perm = ws.perm.p;
    }
    // permuted working copies; rows already in level order with the padded
    // stride and an aligned block are worked on in place (no copy in or out)
    const bool in_place = !perm && !rule && ld_in == ld && (reinterpret_cast<uintptr_t>(d_Vin) & 255) == 0;
    double* V = d_Vin;
    if (!in_place) {
        ws.vec[0].ensure(static_cast<size_t>(ld) * C);
        V = ws.vec[0].p;
        XFE_CHECK(launch_gather_cols(d_Vin, ld_in, perm, V, ld, n, C, grid, s));
    }
    const double* w = nullptr;
    if (d_w_in) {
        ws.w.ensure(n);
        XFE_CHECK(launch_gather_vec(d_w_in, perm, ws.w.p, n, grid, s));
        w = ws.w.p;
    }
    tls_phase.mark("level sort + gather");
    Solver sv(ws, n, C, ld);
    sv.D = D;
    sv.w = w;
    sv.sorted_dim = p;  // level-sorted after the permutation
    std::vector<ProfileStats> profs(D);
    for (int d = 0; d < D; ++d) {
        if (perm) {
            ws.gid[d].ensure(n);
            XFE_CHECK(launch_gather_ids(d_gid_in[d], perm, ws.gid[d].p, n, grid, s));
            sv.gid[d] = ws.gid[d].p;
        } else {
            sv.gid[d] = d_gid_in[d];  // already in level order: read in place
        }
        sv.G[d] = G[d];
        sv.prepare_dim(d);
        if (h_wsum[d]) {
            XFE_CHECK(hipMemcpyAsync(ws.wsum[d].p, h_wsum[d], sizeof(double) * G[d],
                                     hipMemcpyHostToDevice, s));
            XFE_CHECK(launch_inverse(ws.wsum[d].p, ws.inv[d].p, G[d], grid_for(G[d]), s));
        } else {
            sv.weight_sums(d);
        }
        // the level-size profile (order_by_profile) from a device reduction: no
        // download of the level table
        const int pgrid = std::max(1, std::min(1024, grid_for(G[d])));
        ws.partials.ensure(static_cast<size_t>(4) * pgrid);
        ws.prof_out.ensure(4 * XFE_MAX_DIMS);
        XFE_CHECK(launch_level_profile(ws.wsum[d].p, G[d], ws.partials.p, pgrid, ws.prof_out.p + 4 * d, s));
    }
    XFE_CHECK(hipMemcpyAsync(ws.host_scalars, ws.prof_out.p, sizeof(double) * 4 * D, hipMemcpyDeviceToHost, s));
    XFE_CHECK(hipStreamSynchronize(s));
    for (int d = 0; d < D; ++d) {
        const double* h = ws.host_scalars + 4 * d;
        profs[d] = profile_from_sums(h[0], h[1], h[2], h[3], G[d], w == nullptr);
    }
    tls_phase.mark("ids + weight sums");
    setup_tiles(sv, ws, n, D, G, w == nullptr && h_wsum[p] == nullptr);
    tls_phase.mark("tiles");
    // partial buffer: pass partials (3 per block and chunk), per-column CG dots,
    // then the polish maxima of one sweep (<= 2D projections)
    const size_t pg = static_cast<size_t>(std::max(std::max(sv.grid, grid), sv.tile_grid));
    size_t fgmax = 0;
    for (int d = 0; d < D; ++d)
        fgmax = std::max(fgmax, static_cast<size_t>(grid_for(static_cast<long long>(G[d]) * C)));
    const size_t part = sv.maxp_offset() + pg * C + 2 * static_cast<size_t>(D) * (pg * sv.nchunks + fgmax);
    ws.partials.ensure(part + 1024);
    ws.scalars.ensure(std::max(C, 8));
    ws.lanes.ensure(std::max(C, 8));

    std::vector<int> order(D);
    if (sweep_order) {
        for (int d = 0; d < D; ++d) order[d] = sweep_order[d];
    } else if (D > 2) {
        order = order_by_profile(profs);
    } else {
        order = probe_order;
    }
    if (rule) {
        solve_schwarz(sv, V, d_Vin, ld_in, perm, order, profs, opt, res, *rule);
        res.converged = rule->converged ? 1 : 0;
        XFE_CHECK(launch_scatter_cols(V, ld, perm, d_Vin, ld_in, n, C, grid, s));
        XFE_CHECK(hipStreamSynchronize(s));
        res.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return;
    }
    std::vector<double> orig_sq;
    if (opt.certify) orig_sq = column_sumsq(sv, V, ld);
    tls_phase.mark("input norms");
    solve_core(sv, V, order, profs, opt, res);
    if (opt.certify) {
        certificate(sv, V, orig_sq, opt, res.converged, res.abs_residual, res.abs_residual_rel,
                    res.precision_certified);
        res.certificate_computed = 1;
        tls_phase.mark("certificate");
    }
    if (!in_place) XFE_CHECK(launch_scatter_cols(V, ld, perm, d_Vin, ld_in, n, C, grid, s));
    XFE_CHECK(hipStreamSynchronize(s));
    tls_phase.mark("scatter");
    const auto t1 = std::chrono::steady_clock::now();
    res.solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
}

// Absorption with heterogeneous slopes (absorb_fixed_effects_v6_mixed,
// src/fe_absorption.cpp:8162, with slope_project_inplace :4096-4300 for a slope
// dimension), in FP64.  Every dimension's map v -> v - P_d v is a W-orthogonal
// projection (a slope dimension projects each level onto span{1, z}, or span{z}
// without the intercept), so the symmetric sweep T = (I - P_1)...(I - P_D)...
// (I - P_1) is self-adjoint and the within transform is v - y with y the
// solution of (I - T) y = (I - T) v in the span of the fixed effects: CG on
// that system per column (the reference's accelerator for slow alternating
// projections, src/fe_absorption.cpp:6790-6973), from y = 0, in the W inner
// product.  The residual r = (I - T)(v - y) is the change one symmetric sweep
// would make to the current transform; the solve stops when every column has
// ||r||_W <= tol_eff max(1, ||v||_W).  V is n x C (ld), in the caller's row
// order on entry and exit.
static void absorb_slopes_impl(double* V_in, int n, int C, int ld, const std::vector<const int*>& gid_in,
                               const std::vector<int>& G, const std::vector<const double*>& z_in,
                               const std::vector<int>& mode, const double* w_in, const std::vector<int>& order,
                               const xfe_options& opt, xfe_absorb_result& res) {
    Workspace& ws = tls_ws;
    hipStream_t s = ws.stream;
    const int D = static_cast<int>(G.size());
    const int grid = grid_for(n);
    // Rows are put in the order of the (first) slope dimension's levels (a
    // stable radix sort), so its level sums reduce runs of rows inside the
    // waves; the result is scattered back at the end.
    int sd = -1;
    for (int d = 0; d < D && sd < 0; ++d)
        if (z_in[d]) sd = d;
    double* V = V_in;
    std::vector<const int*> gid = gid_in;
    std::vector<const double*> z = z_in;
    const double* w = w_in;
    const int* perm = nullptr;
    if (sd >= 0) {
        ws.perm.ensure(n);
        ws.keys_tmp.ensure(n);
        ws.vals_tmp.ensure(n);
        XFE_CHECK(launch_iota(ws.vals_tmp.p, n, grid, s));
        int bits = 1;
        while ((1LL << bits) < static_cast<long long>(G[sd])) ++bits;
        size_t tmp_bytes = 0;
        XFE_CHECK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, gid_in[sd], ws.keys_tmp.p, ws.vals_tmp.p, ws.perm.p,
                                            static_cast<size_t>(n), 0, bits, s));
        ws.sort_tmp.ensure(tmp_bytes);
        XFE_CHECK(rocprim::radix_sort_pairs(ws.sort_tmp.p, tmp_bytes, gid_in[sd], ws.keys_tmp.p, ws.vals_tmp.p,
                                            ws.perm.p, static_cast<size_t>(n), 0, bits, s));
        perm = ws.perm.p;
        ws.vec[0].ensure(static_cast<size_t>(ld) * C);
        V = ws.vec[0].p;
        XFE_CHECK(launch_gather_cols(V_in, ld, perm, V, ld, n, C, grid, s));
        for (int d = 0; d < D; ++d) {
            ws.gid[d].ensure(n);
            XFE_CHECK(launch_gather_ids(gid_in[d], perm, ws.gid[d].p, n, grid, s));
            gid[d] = ws.gid[d].p;
            if (z_in[d]) {
                ws.slope_z[d].ensure(n);
                XFE_CHECK(launch_gather_vec(z_in[d], perm, ws.slope_z[d].p, n, grid, s));
                z[d] = ws.slope_z[d].p;
            }
        }
        if (w_in) {
            ws.w.ensure(n);
            XFE_CHECK(launch_gather_vec(w_in, perm, ws.w.p, n, grid, s));
            w = ws.w.p;
        }
    }
    // copies of each level-sum table: a level's rows spread their atomics over
    // up to 8 copies (at most 64 MB per table)
    std::vector<int> copies(D);
    for (int d = 0; d < D; ++d) {
        const size_t tab = 2 * static_cast<size_t>(G[d]) * C;
        copies[d] = static_cast<int>(std::max<size_t>(1, std::min<size_t>(8, (size_t(8) << 20) / tab)));
        ws.wsum[d].ensure(3 * static_cast<size_t>(G[d]));
        ws.sums[d].ensure(tab * copies[d]);
        ws.means[d].ensure(tab);
        XFE_CHECK(hipMemsetAsync(ws.wsum[d].p, 0, sizeof(double) * 3 * G[d], s));
        XFE_CHECK(launch_slope_moments(gid[d], z[d], w, n, ws.wsum[d].p, grid, s));
        res.sweep_order_used[d] = order[d];
    }
    const size_t block = static_cast<size_t>(ld) * C;
    for (int k = 1; k <= 5; ++k) ws.vec[k].ensure(block);
    double* y = ws.vec[1].p;
    double* r = ws.vec[2].p;
    double* pv = ws.vec[3].p;
    double* q = ws.vec[4].p;
    double* t = ws.vec[5].p;
    ws.partials.ensure(static_cast<size_t>(grid) * C);
    ws.scalars.ensure(std::max(C, 8));
    ws.lanes.ensure(std::max(C, 8));
    auto wdot = [&](const double* a, const double* b, double* out) {
        XFE_CHECK(launch_slope_wdot(a, b, w, n, C, ld, ws.partials.p, grid, s));
        XFE_CHECK(launch_reduce_partials(ws.partials.p, grid, C, 0, ws.scalars.p, s));
        XFE_CHECK(hipMemcpyAsync(out, ws.scalars.p, sizeof(double) * C, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
    };
    auto upload = [&](const std::vector<double>& h) {
        XFE_CHECK(hipMemcpyAsync(ws.lanes.p, h.data(), sizeof(double) * C, hipMemcpyHostToDevice, s));
    };
    const std::vector<int> seq = sweep_sequence(order, true);
    int sweeps = 0;
    // out = (I - T) x, through t
    auto apply_A = [&](const double* x, double* out) {
        XFE_CHECK(hipMemcpyAsync(t, x, sizeof(double) * block, hipMemcpyDeviceToDevice, s));
        for (int d : seq) {
            XFE_CHECK(hipMemsetAsync(ws.sums[d].p, 0,
                                     sizeof(double) * 2 * static_cast<size_t>(G[d]) * C * copies[d], s));
            const int pslot = z[d] ? profile_begin(s) : -1;
            XFE_CHECK(launch_slope_reduce(t, ld, C, n, gid[d], z[d], w, ws.sums[d].p, G[d], copies[d], grid, s));
            // the slope dimension's reduce: ids, z, (w,) C values per row, and the 2 C
            // level sums of each row's level
            profile_end(pslot, s,
                        static_cast<double>(n) * (12.0 + (w ? 8.0 : 0.0) + 8.0 * C) + 16.0 * static_cast<double>(G[d]) * C);
            XFE_CHECK(launch_slope_coef(ws.sums[d].p, ws.wsum[d].p, G[d], C, mode[d], copies[d], ws.means[d].p,
                                        grid_for(static_cast<long long>(G[d]) * C), s));
            XFE_CHECK(launch_slope_apply(t, ld, C, n, gid[d], z[d], ws.means[d].p, grid, s));
        }
        ++sweeps;
        XFE_CHECK(launch_sub(out, x, t, n, C, ld, grid, s));
    };
    const double tol_eff = (opt.tol > 0.0 && opt.tolerance_mode == 1) ? std::min(opt.tol, 1e-9) : opt.tol;
    std::vector<double> vv(C), rr(C), rr_new(C), pq(C), alpha(C), beta(C), limit(C);
    wdot(V, V, vv.data());
    for (int c = 0; c < C; ++c) limit[c] = tol_eff * std::max(1.0, std::sqrt(std::max(0.0, vv[c])));
    // y = 0, r = p = (I - T) v
    XFE_CHECK(hipMemsetAsync(y, 0, sizeof(double) * block, s));
    apply_A(V, r);
    XFE_CHECK(hipMemcpyAsync(pv, r, sizeof(double) * block, hipMemcpyDeviceToDevice, s));
    wdot(r, r, rr.data());
    bool converged = false;
    int it = 0;
    const bool trace = std::getenv("XFE_SLOPE_TRACE") != nullptr;  // progress lines on stderr
    while (true) {
        bool done = true;
        for (int c = 0; c < C; ++c)
            if (std::sqrt(std::max(0.0, rr[c])) > limit[c]) done = false;
        if (done) {
            converged = true;
            break;
        }
        if (it >= opt.max_iter) break;
        ++it;
        apply_A(pv, q);
        wdot(pv, q, pq.data());
        for (int c = 0; c < C; ++c) alpha[c] = (pq[c] > 0.0 && rr[c] > 0.0) ? rr[c] / pq[c] : 0.0;
        upload(alpha);
        XFE_CHECK(launch_col_axpy(y, y, pv, ws.lanes.p, n, C, ld, 1.0, grid, s));
        XFE_CHECK(launch_col_axpy(r, r, q, ws.lanes.p, n, C, ld, -1.0, grid, s));
        wdot(r, r, rr_new.data());
        for (int c = 0; c < C; ++c) beta[c] = rr[c] > 0.0 ? rr_new[c] / rr[c] : 0.0;
        upload(beta);
        XFE_CHECK(launch_col_axpy(pv, r, pv, ws.lanes.p, n, C, ld, 1.0, grid, s));
        rr = rr_new;
        if (trace && it % 100 == 0) {
            double worst = 0.0;
            for (int c = 0; c < C; ++c)
                if (limit[c] > 0.0) worst = std::max(worst, std::sqrt(std::max(0.0, rr[c])) / limit[c]);
            std::fprintf(stderr, "[slope fp64] iteration %d: worst |r| / limit %.3e\n", it, worst);
        }
    }
    XFE_CHECK(launch_sub(V, V, y, n, C, ld, grid, s));
    if (perm) XFE_CHECK(launch_scatter_cols(V, ld, perm, V_in, ld, n, C, grid, s));
    XFE_CHECK(hipStreamSynchronize(s));
    res.iterations = it;
    res.converged = converged ? 1 : 0;
    res.used_cg = 1;
    res.sweeps = sweeps;
}

}  // namespace xfe

using namespace xfe;

// One Gauss-Seidel pass over the residual in ws.vec[6] (ld = n rounded up to
// 64; modified) with the ids staged in ws.stage_ids (dimension-major, n):
// max_d max_g |weighted mean_g| (fe_recovery_max_delta_cuda,
// src/fe_absorption_cuda.cu:5227).  w: device weights or nullptr.
static double max_delta_staged(Workspace& ws, int n, const xfe_fe_input* fes, int num_fes, const double* w,
                               const int* sweep_order) {
    hipStream_t s = ws.stream;
    const int ld = ((n + 63) / 64) * 64;
    Solver sv(ws, n, 1, ld);
    sv.D = num_fes;
    sv.w = w;
    size_t part = 0;
    for (int d = 0; d < num_fes; ++d) {
        if (!fes[d].group_ids || fes[d].num_groups <= 0) throw Error("invalid FE input");
        const int G = fes[d].num_groups;
        int* dst = ws.stage_ids.p + static_cast<size_t>(d) * n;
        sv.gid[d] = dst;
        sv.G[d] = G;
        sv.prepare_dim(d);
        if (fes[d].weight_sums) {
            XFE_CHECK(hipMemcpyAsync(ws.wsum[d].p, fes[d].weight_sums, sizeof(double) * G,
                                     hipMemcpyHostToDevice, s));
            XFE_CHECK(launch_inverse(ws.wsum[d].p, ws.inv[d].p, G, grid_for(G), s));
        } else {
            sv.weight_sums(d);
        }
        part += grid_for(G);
    }
    ws.partials.ensure(part + static_cast<size_t>(std::max(sv.grid, grid_for(n))) * 3 + 1024);
    ws.scalars.ensure(8);
    std::vector<int> order(num_fes);
    for (int d = 0; d < num_fes; ++d) order[d] = sweep_order ? sweep_order[d] : d;
    // one Gauss-Seidel pass: reduce d, means (+ max |mean|), subtract
    size_t off = 0;
    sv.pass(ws.vec[6].p, nullptr, nullptr, nullptr, 0.0, nullptr, -1, order[0], 0, nullptr,
            nullptr, nullptr);
    for (int j = 0; j < num_fes; ++j) {
        const int d = order[j];
        sv.finalize(d, ws.partials.p + off);
        off += sv.finalize_grid(d);
        const int r = j + 1 < num_fes ? order[j + 1] : -1;
        sv.pass(ws.vec[6].p, ws.vec[6].p, nullptr, nullptr, 0.0, nullptr, d, r, 0, nullptr,
                nullptr, nullptr);
    }
    double mx = 0.0;
    sv.fetch(ws.partials.p, static_cast<int>(off), 1, 1, &mx);
    return mx;
}

// Certificate of a group()/individual() candidate on the host
// (certify_group_individual_candidate, src/fe_absorption.cpp:15192-15300,
// through certify_absorption_result :14942-15190 with the individual
// dimension): abs_c^2 = sum_d ||D_d' W v~_c||^2 + ||A' W v~_c||^2 with A the
// individual design (row g, individual j: s_g), ||D'W||_F^2 = sum_d sum_i w_i^2
// + sum over incidences of (w_g s_g)^2, rel_c = abs_c / (||D'W||_F ||v_c||);
// certified iff converged and max_c rel_c <= max(8 tol_eff, 64 eps).
static void gi_certificate(const double* V0, const double* Vt, int n, int C, const xfe_fe_input* fes, int D,
                           const xfe_group_individual& gi, const double* w, double tol_eff,
                           xfe_absorb_result& res) {
    long double op = 0.0L;
    std::vector<long double> abs2(C, 0.0L), orig2(C, 0.0L);
    for (int c = 0; c < C; ++c)
        for (int i = 0; i < n; ++i) {
            const long double v = V0[static_cast<size_t>(c) * n + i];
            orig2[c] += v * v;
        }
    for (int d = 0; d < D; ++d) {
        const int G = fes[d].num_groups;
        std::vector<long double> sums(static_cast<size_t>(G) * C, 0.0L);
        for (int i = 0; i < n; ++i) {
            const long double wi = w ? w[i] : 1.0;
            op += wi * wi;
            long double* sg = &sums[static_cast<size_t>(fes[d].group_ids[i]) * C];
            for (int c = 0; c < C; ++c) sg[c] += wi * Vt[static_cast<size_t>(c) * n + i];
        }
        for (size_t q = 0; q < sums.size(); ++q) abs2[q % C] += sums[q] * sums[q];
    }
    const int m = gi.num_individuals;
    std::vector<long double> isum(static_cast<size_t>(m) * C, 0.0L);
    for (int g = 0; g < n; ++g) {
        const long double mult = static_cast<long double>(w ? w[g] : 1.0) * gi.group_scale[g];
        for (int q = gi.group_ptr[g]; q < gi.group_ptr[g + 1]; ++q) {
            op += mult * mult;
            long double* sj = &isum[static_cast<size_t>(gi.group_individual[q]) * C];
            for (int c = 0; c < C; ++c) sj[c] += mult * Vt[static_cast<size_t>(c) * n + g];
        }
    }
    for (size_t q = 0; q < isum.size(); ++q) abs2[q % C] += isum[q] * isum[q];
    const double opn = std::sqrt(static_cast<double>(op));
    double max_abs = 0.0, max_rel = 0.0;
    for (int c = 0; c < C; ++c) {
        const double a = std::sqrt(static_cast<double>(abs2[c]));
        const double scale = opn * std::sqrt(static_cast<double>(orig2[c]));
        const double rel = scale > 0.0 ? a / scale : (a > 0.0 ? DBL_MAX : 0.0);
        max_abs = std::max(max_abs, a);
        max_rel = std::max(max_rel, rel);
    }
    res.abs_residual = max_abs;
    res.abs_residual_rel = max_rel;
    res.precision_certified = max_rel <= std::max(8.0 * std::max(0.0, tol_eff), 64.0 * DBL_EPSILON) ? 1 : 0;
    res.certificate_computed = 1;
}

// Greedy colouring of the individuals in index order (the reference's host
// step, src/fe_absorption_cuda.cu:4913-4947): each individual takes the
// smallest colour not used by an earlier individual sharing a row with it.
// Returns the members of each colour, colour by colour, and the offsets.
static void gi_colour(const xfe_group_individual& gi, std::vector<int>& members, std::vector<int>& offsets) {
    const int m = gi.num_individuals;
    std::vector<int> colour(m, 0), used;
    int ncol = 1;
    for (int j = 0; j < m; ++j) {
        for (int q = gi.individual_ptr[j]; q < gi.individual_ptr[j + 1]; ++q) {
            const int g = gi.individual_group[q];
            for (int r = gi.group_ptr[g]; r < gi.group_ptr[g + 1]; ++r) {
                const int o = gi.group_individual[r];
                if (o >= j) continue;
                const int co = colour[o];
                if (co >= static_cast<int>(used.size())) used.resize(co + 1, -1);
                used[co] = j;
            }
        }
        int c = 0;
        while (c < static_cast<int>(used.size()) && used[c] == j) ++c;
        colour[j] = c;
        ncol = std::max(ncol, c + 1);
    }
    offsets.assign(ncol + 1, 0);
    for (int j = 0; j < m; ++j) ++offsets[colour[j] + 1];
    for (int c = 0; c < ncol; ++c) offsets[c + 1] += offsets[c];
    members.assign(m, 0);
    std::vector<int> cur(offsets.begin(), offsets.end() - 1);
    for (int j = 0; j < m; ++j) members[cur[colour[j]]++] = j;
}

// AbsorptionMethod values (include/hdfe/hdfe_regressor.hpp:18) the device
// absorber accepts: Auto, GaussSeidel, SymmetricGaussSeidel, and Schwarz /
// AutoMlsmr, which the reference's device absorber sweeps as Gauss-Seidel
// (every method but Jacobi, src/fe_absorption_cuda.cu:3383).  Jacobi is never
// sent to the device by the reference's dispatcher (src/fe_absorption.cpp:9141)
// and is refused with slopes and by the group()/individual() absorber
// (src/fe_absorption_cuda.cu:2324, :4771); Lsmr and Mlsmr are CPU-only there
// (src/fe_absorption.cpp:9127).  A refused method fails closed (return 0), so
// the caller keeps its CPU path.
// Mixed-precision absorption with heterogeneous slopes (precision 1): the
// Schur PCG of slope_mixed.hip, then the normal-equation certificate with the
// slope moments (certify_absorption_result with slopes,
// src/fe_absorption.cpp:14942-15100: per dimension the level sums of w v~ and,
// for a slope, of w z v~; ||D'W||_F^2 = sum over dimensions and moments of
// sum_i (w_i m_i)^2).  V: n x C (ld), device, transformed in place.
// FP32 iterations the mixed slope path gets; the columns they leave
// unconverged continue with up to kSlopeFp64MaxIter FP64 block-PCG
// iterations, and only what that leaves goes to the FP64 slope path
constexpr int kSlopeMixedMaxIter = 2000;
constexpr int kSlopeFp64MaxIter = 10000;

// false when a column did not converge: V is then untouched and the caller
// runs the FP64 slope path on it
static bool absorb_slopes_mixed(double* V, int n, int C, int ld, const std::vector<const int*>& gid,
                                const std::vector<int>& G, const std::vector<const double*>& z,
                                const std::vector<int>& mode, const xfe_options& opt, xfe_absorb_result& res) {
    Workspace& ws = tls_ws;
    hipStream_t s = ws.stream;
    const int D = static_cast<int>(G.size());
    const int grid = grid_for(n);
    // both buffers of the column sums below (this entry allocates them: no
    // solver set-up has run on this path)
    ws.partials.ensure(static_cast<size_t>(grid) * std::max(C, 1) + 64);
    ws.scalars.ensure(std::max(C, 64));
    if (!ws.host_scalars) throw Error("absorb_slopes_mixed: workspace not initialised");
    auto col_sumsq = [&](const double* a, int cols, int lda, double* out) {
        XFE_CHECK(launch_col_dot(a, a, n, cols, lda, ws.partials.p, grid, s));
        XFE_CHECK(launch_reduce_partials(ws.partials.p, grid, cols, 0, ws.scalars.p, s));
        XFE_CHECK(hipMemcpyAsync(ws.host_scalars, ws.scalars.p, sizeof(double) * cols, hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        for (int c = 0; c < cols; ++c) out[c] = ws.host_scalars[c];
    };
    std::vector<double> orig(C, 0.0);
    if (opt.certify) col_sumsq(V, C, ld, orig.data());
    if (!ws.slope_mixed) ws.slope_mixed = slope_mixed_new();
    const double target = opt.refine_tol > 0.0 ? opt.refine_tol
                                               : std::min(3e-13, (opt.tol > 0.0 ? opt.tol : 1e-8) * 3e-5);
    SlopeMixedResult out;
    const auto t0 = std::chrono::steady_clock::now();
    slope_mixed_run(*ws.slope_mixed, V, n, C, ld, gid, G, z, mode, target,
                    std::max(10, std::min(opt.max_iter, kSlopeMixedMaxIter)),
                    std::max(10, std::min(opt.max_iter, kSlopeFp64MaxIter)), s, out);
    if (!out.converged) return false;
    res.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    res.iterations = out.iterations;  // FP32 and FP64 continuation iterations of the slowest column
    res.refinements = out.replacements;
    res.sweeps = out.passes + out.fp64_iterations;
    res.converged = out.converged ? 1 : 0;
    res.used_cg = 1;
    res.used_accel = 0;
    res.precision_used = 1;
    res.refine_target = target;
    for (int d = 0; d < D; ++d) {
        res.fe_levels[d] = G[d];
        res.sweep_order_used[d] = d;
    }
    if (!opt.certify) return true;
    // the dropped intercepts' level sums, and ||D||_F^2
    std::vector<double> resid = out.resid_sq;
    double op = static_cast<double>(n);  // the eliminated dimension
    for (int d = 0; d < D; ++d) {
        if (d == out.primary) continue;
        if (mode[d] == 0 || mode[d] == 1) op += static_cast<double>(n);
        if (mode[d] == 1 || mode[d] == 2) {
            double zz = 0.0;
            col_sumsq(z[d], 1, n, &zz);
            op += zz;
        }
    }
    {
        // every slope intercept that slope_mixed_run left out belongs to a
        // dimension with mode 1 whose intercept a finer dimension spans; the
        // run reports their count, and their level sums are added here
        int left = out.intercepts_dropped;
        for (int d = 0; d < D && left > 0; ++d) {
            if (mode[d] != 1 || d == out.primary) continue;
            if (!slope_mixed_intercept_dropped(*ws.slope_mixed, d)) continue;
            slope_mixed_level_sumsq(*ws.slope_mixed, V, n, C, ld, gid[d], G[d], s, resid);
            --left;
        }
    }
    double amax = 0.0, rmax = 0.0;
    for (int c = 0; c < C; ++c) {
        const double a = std::sqrt(std::max(0.0, resid[c]));
        const double scale = std::sqrt(std::max(0.0, op)) * std::sqrt(std::max(0.0, orig[c]));
        const double rel = scale > 0.0 ? a / scale : (a > 0.0 ? DBL_MAX : 0.0);
        amax = std::max(amax, a);
        rmax = std::max(rmax, rel);
    }
    const double tol_eff = (opt.tol > 0.0 && opt.tolerance_mode == 1) ? std::min(opt.tol, 1e-9) : opt.tol;
    const double limit = std::max(8.0 * std::max(0.0, tol_eff), 64.0 * DBL_EPSILON);
    res.abs_residual = amax;
    res.abs_residual_rel = rmax;
    res.precision_certified = (res.converged && rmax <= limit) ? 1 : 0;
    res.certificate_computed = 1;
    return true;
}

static bool method_ok(int method, const char* fn) {
    if (method == 0 || method == 1 || method == 2 || method == 4 || method == 7) return true;
    const char* what = method == 3 ? "Jacobi is not a device method"
                     : (method == 5 || method == 6) ? "LSMR/MLSMR are CPU-only methods"
                                                    : "unknown absorption method";
    set_error(std::string(fn) + ": " + what + " (method " + std::to_string(method) + ")");
    return false;
}

static bool gi_valid(const xfe_group_individual* gi, int n) {
    if (!gi || gi->num_groups != n || gi->num_individuals <= 0 || !gi->group_ptr || !gi->group_individual ||
        !gi->individual_ptr || !gi->individual_group || !gi->group_scale)
        return false;
    const int m = gi->num_individuals;
    if (gi->group_ptr[0] != 0 || gi->individual_ptr[0] != 0) return false;
    for (int g = 0; g < n; ++g)
        if (gi->group_ptr[g + 1] < gi->group_ptr[g]) return false;
    for (int j = 0; j < m; ++j)
        if (gi->individual_ptr[j + 1] < gi->individual_ptr[j]) return false;
    const int E = gi->group_ptr[n];
    if (gi->individual_ptr[m] != E) return false;
    for (int q = 0; q < E; ++q)
        if (gi->group_individual[q] < 0 || gi->group_individual[q] >= m || gi->individual_group[q] < 0 ||
            gi->individual_group[q] >= n)
            return false;
    return true;
}

extern "C" {

int xfe_backend_available(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return count > 0 ? 1 : 0;
}

int xfe_absorb_device(double* d_V, int n, int C, int ld, const int* d_group_ids,
                      const int* num_groups, int num_fes, const double* d_weights,
                      const int* sweep_order, const xfe_options* opt, xfe_absorb_result* res) {
    clear_error();
    if (!res || !opt || !d_V || n < 0 || C <= 0 || C > XFE_MAX_COLS || ld < n || num_fes < 0 ||
        num_fes > XFE_MAX_DIMS) {
        set_error("xfe_absorb_device: invalid arguments");
        return 0;
    }
    if (!method_ok(opt->method, "xfe_absorb_device")) return 0;
    std::memset(res, 0, sizeof(*res));
    if (num_fes == 0 || n == 0) {
        res->converged = 1;
        return 1;
    }
    try {
        (void)hipGetLastError();  // clear sticky errors (src/fe_absorption_cuda.cu:2484-2491)
        // The library works on its own non-blocking stream: wait for the
        // caller's pending work on the device (e.g. the copy that filled d_V)
        // before reading its buffers.
        XFE_CHECK(hipDeviceSynchronize());
        std::vector<const int*> gids(num_fes);
        std::vector<int> G(num_fes);
        std::vector<const double*> wsum(num_fes, nullptr);
        for (int d = 0; d < num_fes; ++d) {
            gids[d] = d_group_ids + static_cast<size_t>(d) * n;
            G[d] = num_groups[d];
            if (G[d] <= 0) throw Error("num_groups must be positive");
        }
        absorb_device_impl(d_V, n, C, ld, gids, G, wsum, d_weights, sweep_order, *opt, *res);
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

int xfe_absorb_device_slopes(double* d_V, int n, int C, int ld, const int* d_group_ids, const int* num_groups,
                             int num_fes, const double* const* d_slope_values, const int* slope_mode,
                             const double* d_weights, const int* sweep_order, const xfe_options* opt,
                             xfe_absorb_result* res) {
    clear_error();
    if (!res || !opt || !d_V || !d_group_ids || !num_groups || !slope_mode || n < 0 || C <= 0 ||
        C > XFE_MAX_COLS || ld < n || num_fes < 1 || num_fes > XFE_MAX_DIMS) {
        set_error("xfe_absorb_device_slopes: invalid arguments");
        return 0;
    }
    if (!method_ok(opt->method, "xfe_absorb_device_slopes")) return 0;
    for (int d = 0; d < num_fes; ++d) {
        if (slope_mode[d] < 0 || slope_mode[d] > 2 || num_groups[d] <= 0 ||
            (slope_mode[d] != 0 && (!d_slope_values || !d_slope_values[d]))) {
            set_error("xfe_absorb_device_slopes: each dimension needs levels, a mode in 0..2 and, for a slope, "
                      "its values");
            return 0;
        }
        if (sweep_order && (sweep_order[d] < 0 || sweep_order[d] >= num_fes)) {
            set_error("xfe_absorb_device_slopes: invalid sweep_order");
            return 0;
        }
    }
    std::memset(res, 0, sizeof(*res));
    if (n == 0) {
        res->converged = 1;
        return 1;
    }
    try {
        (void)hipGetLastError();
        XFE_CHECK(hipDeviceSynchronize());
        Workspace& ws = tls_ws;
        ws.init();
        std::vector<const int*> gids(num_fes);
        std::vector<int> G(num_fes), mode(num_fes), order(num_fes);
        std::vector<const double*> z(num_fes, nullptr);
        for (int d = 0; d < num_fes; ++d) {
            gids[d] = d_group_ids + static_cast<size_t>(d) * n;
            G[d] = num_groups[d];
            mode[d] = slope_mode[d];
            z[d] = mode[d] != 0 ? d_slope_values[d] : nullptr;
            order[d] = sweep_order ? sweep_order[d] : d;
            res->fe_levels[d] = G[d];
        }
        if (opt->precision == 1 && slope_mixed_eligible(C, num_fes, mode.data(), d_weights) &&
            absorb_slopes_mixed(d_V, n, C, ld, gids, G, z, mode, *opt, *res)) {
        } else {
            absorb_slopes_impl(d_V, n, C, ld, gids, G, z, mode, d_weights, order, *opt, *res);
            res->precision_fallback = opt->precision == 1 ? 1 : 0;
        }
        XFE_CHECK(hipStreamSynchronize(ws.stream));
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

int xfe_absorb(const double* y, const double* X, int n, int k, const xfe_fe_input* fes,
               int num_fes, const double* weights, const int* sweep_order, const xfe_options* opt,
               xfe_absorb_result* res, double* y_out, double* X_out) {
    clear_error();
    if (!res || !opt || !y || (k > 0 && !X) || n < 0 || k < 0 || k + 1 > XFE_MAX_COLS ||
        num_fes < 0 || num_fes > XFE_MAX_DIMS || !y_out || (k > 0 && !X_out)) {
        set_error("xfe_absorb: invalid arguments");
        return 0;
    }
    if (!method_ok(opt->method, "xfe_absorb")) return 0;
    std::memset(res, 0, sizeof(*res));
    bool slopes = false;
    for (int d = 0; d < num_fes; ++d) {
        if (fes[d].is_slope) {
            if (!fes[d].slope_values) {
                set_error("xfe_absorb: a slope dimension needs slope_values");
                return 0;
            }
            slopes = true;
        }
        if (!fes[d].group_ids || fes[d].num_groups <= 0) {
            set_error("xfe_absorb: invalid FE input");
            return 0;
        }
    }
    if (num_fes == 0 || n == 0) {
        std::memcpy(y_out, y, sizeof(double) * n);
        if (k > 0) std::memcpy(X_out, X, sizeof(double) * static_cast<size_t>(n) * k);
        res->converged = 1;
        return 1;
    }
    try {
        (void)hipGetLastError();
        Workspace& ws = tls_ws;
        ws.init();
        hipStream_t s = ws.stream;
        const int C = k + 1;
        ws.stage.ensure(static_cast<size_t>(n) * C);
        XFE_CHECK(hipMemcpyAsync(ws.stage.p, y, sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (k > 0)
            XFE_CHECK(hipMemcpyAsync(ws.stage.p + n, X, sizeof(double) * static_cast<size_t>(n) * k,
                                     hipMemcpyHostToDevice, s));
        ws.stage_ids.ensure(static_cast<size_t>(n) * num_fes);
        std::vector<const int*> gids(num_fes);
        std::vector<int> G(num_fes);
        std::vector<const double*> wsum(num_fes);
        for (int d = 0; d < num_fes; ++d) {
            int* dst = ws.stage_ids.p + static_cast<size_t>(d) * n;
            XFE_CHECK(hipMemcpyAsync(dst, fes[d].group_ids, sizeof(int) * n, hipMemcpyHostToDevice, s));
            gids[d] = dst;
            G[d] = fes[d].num_groups;
            wsum[d] = fes[d].weight_sums;
        }
        if (weights) {
            ws.stage_w.ensure(n);
            XFE_CHECK(hipMemcpyAsync(ws.stage_w.p, weights, sizeof(double) * n,
                                     hipMemcpyHostToDevice, s));
        }
        if (slopes) {
            std::vector<const double*> z(num_fes, nullptr);
            std::vector<int> mode(num_fes, 0), order(num_fes);
            for (int d = 0; d < num_fes; ++d) {
                res->fe_levels[d] = G[d];
                order[d] = sweep_order ? sweep_order[d] : d;
                if (order[d] < 0 || order[d] >= num_fes) throw Error("invalid sweep_order");
                if (!fes[d].is_slope) continue;
                ws.stage_z[d].ensure(n);
                XFE_CHECK(hipMemcpyAsync(ws.stage_z[d].p, fes[d].slope_values, sizeof(double) * n,
                                         hipMemcpyHostToDevice, s));
                z[d] = ws.stage_z[d].p;
                mode[d] = fes[d].slope_has_intercept ? 1 : 2;
            }
            if (opt->precision == 1 && slope_mixed_eligible(C, num_fes, mode.data(), weights) &&
                absorb_slopes_mixed(ws.stage.p, n, C, n, gids, G, z, mode, *opt, *res)) {
            } else {
                absorb_slopes_impl(ws.stage.p, n, C, n, gids, G, z, mode, weights ? ws.stage_w.p : nullptr,
                                   order, *opt, *res);
                res->precision_fallback = opt->precision == 1 ? 1 : 0;
            }
        } else {
            absorb_device_impl(ws.stage.p, n, C, n, gids, G, wsum, weights ? ws.stage_w.p : nullptr,
                               sweep_order, *opt, *res);
        }
        XFE_CHECK(hipMemcpyAsync(y_out, ws.stage.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        if (k > 0)
            XFE_CHECK(hipMemcpyAsync(X_out, ws.stage.p + n, sizeof(double) * static_cast<size_t>(n) * k,
                                     hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

int xfe_lp_probe(const double* y, const double* X, int n, int k, const xfe_fe_input* fes, int num_fes,
                 int probe_iter, xfe_lp_panel_sizes* sizes) {
    clear_error();
    if (!sizes || probe_iter < 1 || num_fes < 2 || num_fes > 3 || k + 1 > 32) {
        set_error("xfe_lp_probe: invalid arguments (2 or 3 dimensions, at most 32 columns, probe_iter >= 1)");
        return 0;
    }
    for (int d = 0; d < num_fes; ++d)
        if (fes[d].is_slope || fes[d].weight_sums) {
            set_error("xfe_lp_probe: standard unweighted dimensions only");
            return 0;
        }
    lp::ProbeState& P = lp::probe_state();
    P.active = true;
    P.done = false;
    P.probe_iter = probe_iter;
    xfe_options o;
    std::memset(&o, 0, sizeof(o));
    o.tol = 1e-8;
    o.tolerance_mode = 1;
    o.max_iter = 100000;
    o.method = 1;
    o.precision = 1;
    xfe_absorb_result r;
    std::vector<double> yo(static_cast<size_t>(n)), Xo(static_cast<size_t>(n) * k + 1);
    const int ok = xfe_absorb(y, X, n, k, fes, num_fes, nullptr, nullptr, &o, &r, yo.data(), Xo.data());
    P.active = false;
    P.iter.release();
    if (!ok) return 0;
    if (!P.done) {
        set_error("xfe_lp_probe: the panel did not take the mixed-precision segment path");
        return 0;
    }
    sizes->F = P.F;
    sizes->Wm = P.Wm;
    sizes->E = P.E;
    sizes->U = P.U;
    sizes->C = P.C;
    sizes->inner_iterations = P.inner;
    sizes->converged = r.converged;
    return 1;
}

int xfe_schwarz_demean(double* Y, double* X, int64_t n64, int k, const int32_t* const* fe,
                       const int* nlev, int ndim, int threads, double tol, int max_iter,
                       int* iters_out, int collapse) {
    (void)threads;
    (void)collapse;
    clear_error();
    // the reference's shape gates (src/schwarz_demean.cpp:422, :436)
    if (ndim < 2 || n64 <= 0 || k < 0 || k + 1 > 16) {
        set_error("xfe_schwarz_demean: needs ndim >= 2, n > 0 and 0 <= k <= 15");
        return 0;
    }
    if (ndim > XFE_MAX_DIMS || n64 > INT32_MAX || !Y || (k > 0 && !X) || !fe || !nlev) {
        set_error("xfe_schwarz_demean: invalid arguments");
        return 0;
    }
    const int n = static_cast<int>(n64);
    for (int d = 0; d < ndim; ++d) {
        if (!fe[d] || nlev[d] <= 0) {
            set_error("xfe_schwarz_demean: invalid FE input");
            return 0;
        }
        const int32_t* f = fe[d];
        const int L = nlev[d];
        for (int i = 0; i < n; ++i)
            if (f[i] < 0 || f[i] >= L) {
                set_error("xfe_schwarz_demean: id out of range");
                return 0;
            }
    }
    try {
        (void)hipGetLastError();
        Workspace& ws = tls_ws;
        ws.init();
        hipStream_t s = ws.stream;
        const int C = k + 1;
        ws.stage.ensure(static_cast<size_t>(n) * C);
        XFE_CHECK(hipMemcpyAsync(ws.stage.p, Y, sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (k > 0)
            XFE_CHECK(hipMemcpyAsync(ws.stage.p + n, X, sizeof(double) * static_cast<size_t>(n) * k,
                                     hipMemcpyHostToDevice, s));
        ws.stage_ids.ensure(static_cast<size_t>(n) * ndim);
        std::vector<const int*> gids(ndim);
        std::vector<int> G(ndim);
        std::vector<const double*> wsum(ndim, nullptr);
        for (int d = 0; d < ndim; ++d) {
            int* dst = ws.stage_ids.p + static_cast<size_t>(d) * n;
            XFE_CHECK(hipMemcpyAsync(dst, fe[d], sizeof(int) * n, hipMemcpyHostToDevice, s));
            gids[d] = dst;
            G[d] = nlev[d];
        }
        xfe_options opt{};
        opt.tol = tol;
        opt.tolerance_mode = 0;
        opt.max_iter = max_iter;
        opt.precision = 1;
        xfe_absorb_result res{};
        SchwarzRule rule;
        rule.tol = tol;
        absorb_device_impl(ws.stage.p, n, C, n, gids, G, wsum, nullptr, nullptr, opt, res, &rule);
        if (iters_out) *iters_out = res.iterations;
        if (!rule.converged) {
            char msg[160];
            std::snprintf(msg, sizeof(msg),
                          "xfe_schwarz_demean: tolerance not reached (rule %.3e after the FP64 solve and %d sweeps)",
                          rule.achieved, rule.polish);
            set_error(msg);
            return 0;
        }
        XFE_CHECK(hipMemcpyAsync(Y, ws.stage.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        if (k > 0)
            XFE_CHECK(hipMemcpyAsync(X, ws.stage.p + n, sizeof(double) * static_cast<size_t>(n) * k,
                                     hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

int xfe_certify(const double* y, const double* X, const double* y_t, const double* X_t, int n, int k,
                const xfe_fe_input* fes, int num_fes, const double* weights, const xfe_options* opt,
                int converged, xfe_certificate* out) {
    clear_error();
    if (!out || !opt || !y || !y_t || (k > 0 && (!X || !X_t)) || n < 0 || k < 0 || k + 1 > XFE_MAX_COLS ||
        num_fes < 0 || num_fes > XFE_MAX_DIMS || (num_fes > 0 && !fes)) {
        set_error("xfe_certify: invalid arguments");
        return 0;
    }
    std::memset(out, 0, sizeof(*out));
    out->precision_certified = converged ? 1 : 0;
    if (num_fes == 0 || n == 0) return 1;  // nothing to certify (src/fe_absorption.cpp:14971-14973)
    for (int d = 0; d < num_fes; ++d) {
        if (fes[d].is_slope || !fes[d].group_ids || fes[d].num_groups <= 0) {
            set_error("xfe_certify: invalid FE input");
            return 0;
        }
    }
    try {
        (void)hipGetLastError();
        Workspace& ws = tls_ws;
        ws.init();
        hipStream_t s = ws.stream;
        const int C = k + 1;
        const size_t nc = static_cast<size_t>(n) * C;
        ws.stage.ensure(nc);
        ws.vec[1].ensure(nc);
        XFE_CHECK(hipMemcpyAsync(ws.stage.p, y, sizeof(double) * n, hipMemcpyHostToDevice, s));
        XFE_CHECK(hipMemcpyAsync(ws.vec[1].p, y_t, sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (k > 0) {
            XFE_CHECK(hipMemcpyAsync(ws.stage.p + n, X, sizeof(double) * n * k, hipMemcpyHostToDevice, s));
            XFE_CHECK(hipMemcpyAsync(ws.vec[1].p + n, X_t, sizeof(double) * n * k, hipMemcpyHostToDevice, s));
        }
        ws.stage_ids.ensure(static_cast<size_t>(n) * num_fes);
        Solver sv(ws, n, C, n);
        sv.D = num_fes;
        for (int d = 0; d < num_fes; ++d) {
            int* dst = ws.stage_ids.p + static_cast<size_t>(d) * n;
            XFE_CHECK(hipMemcpyAsync(dst, fes[d].group_ids, sizeof(int) * n, hipMemcpyHostToDevice, s));
            sv.G[d] = fes[d].num_groups;
            sv.gid[d] = dst;
            sv.prepare_dim(d);
        }
        if (weights) {
            ws.stage_w.ensure(n);
            XFE_CHECK(hipMemcpyAsync(ws.stage_w.p, weights, sizeof(double) * n, hipMemcpyHostToDevice, s));
            sv.w = ws.stage_w.p;
        }
        ws.scalars.ensure(std::max(C, 8));
        ws.partials.ensure(static_cast<size_t>(sv.grid) * C + 1024);
        const std::vector<double> orig_sq = column_sumsq(sv, ws.stage.p, n);
        certificate(sv, ws.vec[1].p, orig_sq, *opt, converged, out->abs_residual, out->abs_residual_rel,
                    out->precision_certified);
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

int xfe_recovery_max_delta(const double* residual, int n, const xfe_fe_input* fes, int num_fes,
                           const double* weights, const int* sweep_order, double* out_max_delta) {
    clear_error();
    if (!residual || !out_max_delta || n < 0 || num_fes < 0 || num_fes > XFE_MAX_DIMS) {
        set_error("xfe_recovery_max_delta: invalid arguments");
        return 0;
    }
    *out_max_delta = 0.0;
    if (n == 0 || num_fes == 0) return 1;
    try {
        (void)hipGetLastError();
        Workspace& ws = tls_ws;
        ws.init();
        hipStream_t s = ws.stream;
        const int ld = ((n + 63) / 64) * 64;
        ws.vec[6].ensure(ld);
        XFE_CHECK(hipMemcpyAsync(ws.vec[6].p, residual, sizeof(double) * n, hipMemcpyHostToDevice, s));
        const double* w = nullptr;
        if (weights) {
            ws.stage_w.ensure(n);
            XFE_CHECK(hipMemcpyAsync(ws.stage_w.p, weights, sizeof(double) * n, hipMemcpyHostToDevice, s));
            w = ws.stage_w.p;
        }
        ws.stage_ids.ensure(static_cast<size_t>(n) * num_fes);
        for (int d = 0; d < num_fes; ++d) {
            if (!fes[d].group_ids || fes[d].num_groups <= 0) throw Error("invalid FE input");
            XFE_CHECK(hipMemcpyAsync(ws.stage_ids.p + static_cast<size_t>(d) * n, fes[d].group_ids,
                                     sizeof(int) * n, hipMemcpyHostToDevice, s));
        }
        const double mx = max_delta_staged(ws, n, fes, num_fes, w, sweep_order);
        *out_max_delta = mx;
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

int xfe_recover_fixed_effects(const double* partial, int n, const xfe_fe_input* fes, int num_fes,
                              const double* weights, const int* sweep_order, const xfe_options* opt,
                              double fe_tol, double* const* alpha_out, double* contributions_out,
                              xfe_recovery_result* out) {
    clear_error();
    if (!partial || !opt || !out || !alpha_out || n < 0 || num_fes < 0 || num_fes > 3 || !fes) {
        set_error("xfe_recover_fixed_effects: invalid arguments");
        return 0;
    }
    std::memset(out, 0, sizeof(*out));
    for (int d = 0; d < num_fes; ++d) {
        if (!fes[d].group_ids || fes[d].num_groups <= 0 || fes[d].is_slope || !alpha_out[d]) {
            set_error("xfe_recover_fixed_effects: invalid FE input (ids, levels, no slopes, alpha_out)");
            return 0;
        }
    }
    if (weights && num_fes >= 2) {
        set_error("xfe_recover_fixed_effects: weighted recovery over two or more dimensions is not supported");
        return 0;
    }
    if (fe_tol <= 0.0) fe_tol = 1e-6;  // HdfeOptions::fe_tolerance
    if (n == 0 || num_fes == 0) {
        for (int d = 0; d < num_fes; ++d) std::fill(alpha_out[d], alpha_out[d] + fes[d].num_groups, 0.0);
        out->converged = 1;
        out->absorb_converged = 1;
        return 1;
    }
    try {
        (void)hipGetLastError();
        Workspace& ws = tls_ws;
        ws.init();
        hipStream_t s = ws.stream;
        const int grid = grid_for(n);
        ws.stage.ensure(n);
        ws.rec_p.ensure(n);
        XFE_CHECK(hipMemcpyAsync(ws.rec_p.p, partial, sizeof(double) * n, hipMemcpyHostToDevice, s));
        XFE_CHECK(hipMemcpyAsync(ws.stage.p, ws.rec_p.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        ws.stage_ids.ensure(static_cast<size_t>(n) * num_fes);
        std::vector<const int*> gids(num_fes);
        std::vector<int> G(num_fes);
        std::vector<const double*> wsum(num_fes);
        for (int d = 0; d < num_fes; ++d) {
            int* dst = ws.stage_ids.p + static_cast<size_t>(d) * n;
            XFE_CHECK(hipMemcpyAsync(dst, fes[d].group_ids, sizeof(int) * n, hipMemcpyHostToDevice, s));
            gids[d] = dst;
            G[d] = fes[d].num_groups;
            wsum[d] = fes[d].weight_sums;
        }
        const double* w = nullptr;
        if (weights) {
            ws.stage_w.ensure(n);
            XFE_CHECK(hipMemcpyAsync(ws.stage_w.p, weights, sizeof(double) * n, hipMemcpyHostToDevice, s));
            w = ws.stage_w.p;
        }
        // within transform by the Schur path: its FP64 coefficients are those of
        // the non-primary dimensions
        xfe_options o = *opt;
        o.precision = 1;
        o.certify = 0;
        xfe_absorb_result ares;
        std::memset(&ares, 0, sizeof(ares));
        absorb_device_impl(ws.stage.p, n, 1, n, gids, G, wsum, w, sweep_order, o, ares);
        const auto& rec = ws.fast_rec;
        if (num_fes >= 2 && !rec.used)
            throw Error("xfe_recover_fixed_effects: the panel does not take the Schur-complement path");
        RecoverArgs a;
        a.partial = ws.rec_p.p;
        a.vt = ws.stage.p;
        a.w = w;
        a.D = num_fes;
        a.p = num_fes >= 2 ? rec.p : 0;
        for (int d = 0; d < num_fes; ++d) {
            ws.rec_alpha[d].ensure(G[d]);
            a.gid[d] = gids[d];
            a.alpha[d] = ws.rec_alpha[d].p;
        }
        for (int k = 0; num_fes >= 2 && k < rec.NR; ++k) {
            const int d = rec.rdim[k];
            XFE_CHECK(hipMemcpyAsync(ws.rec_alpha[d].p, ws.f_beta.p + rec.off[k], sizeof(double) * G[d],
                                     hipMemcpyDeviceToDevice, s));
        }
        const int p = a.p;
        XFE_CHECK(hipMemsetAsync(ws.rec_alpha[p].p, 0, sizeof(double) * G[p], s));
        XFE_CHECK(launch_recover_primary(a, ws.rec_alpha[p].p, n, grid, s));
        // ws.wsum[p]: the primary dimension's weight sums, filled by the absorber
        XFE_CHECK(launch_recover_divide(ws.rec_alpha[p].p, ws.wsum[p].p, G[p], grid_for(G[p]), s));
        // the reference's check: one Gauss-Seidel pass over partial - sum of contributions
        const int ld = ((n + 63) / 64) * 64;
        ws.vec[6].ensure(ld);
        XFE_CHECK(launch_recover_residual(a, ws.vec[6].p, n, grid, s));
        for (int d = 0; d < num_fes; ++d)
            XFE_CHECK(hipMemcpyAsync(alpha_out[d], ws.rec_alpha[d].p, sizeof(double) * G[d],
                                     hipMemcpyDeviceToHost, s));
        XFE_CHECK(hipStreamSynchronize(s));
        const double mx = max_delta_staged(ws, n, fes, num_fes, w, sweep_order);
        if (contributions_out) {
            // expanded on the device, downloaded once
            ws.rec_contrib.ensure(static_cast<size_t>(n) * num_fes);
            XFE_CHECK(launch_recover_expand(a, ws.rec_contrib.p, n, grid, s));
            XFE_CHECK(hipMemcpyAsync(contributions_out, ws.rec_contrib.p, sizeof(double) * n * num_fes,
                                     hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
        }
        out->iterations = ares.iterations;
        out->absorb_converged = ares.converged;
        out->max_delta = mx;
        out->converged = (ares.converged && mx <= fe_tol) ? 1 : 0;
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

int xfe_absorb_group_individual(const double* y, const double* X, int n, int k, const xfe_fe_input* fes,
                                int num_fes, const xfe_group_individual* gi, const double* weights,
                                const int* sweep_order, const xfe_options* opt, xfe_absorb_result* res,
                                double* y_out, double* X_out) {
    clear_error();
    if (!res || !opt || !y || (k > 0 && !X) || n <= 0 || k < 0 || k + 1 > XFE_MAX_COLS || num_fes < 0 ||
        num_fes > XFE_MAX_DIMS || (num_fes > 0 && !fes) || !y_out || (k > 0 && !X_out)) {
        set_error("xfe_absorb_group_individual: invalid arguments");
        return 0;
    }
    if (!method_ok(opt->method, "xfe_absorb_group_individual")) return 0;
    std::memset(res, 0, sizeof(*res));
    if (!gi_valid(gi, n)) {
        set_error("xfe_absorb_group_individual: invalid group/individual structure");
        return 0;
    }
    for (int d = 0; d < num_fes; ++d) {
        if (!fes[d].group_ids || fes[d].num_groups <= 0 || fes[d].is_slope) {
            set_error("xfe_absorb_group_individual: invalid FE input");
            return 0;
        }
        for (int i = 0; i < n; ++i)
            if (fes[d].group_ids[i] < 0 || fes[d].group_ids[i] >= fes[d].num_groups) {
                set_error("xfe_absorb_group_individual: FE id out of range");
                return 0;
            }
    }
    if (opt->tolerance_mode == 2) {
        set_error("xfe_absorb_group_individual: strict-residual tolerance mode is not supported");
        return 0;
    }
    try {
        (void)hipGetLastError();
        Workspace& ws = tls_ws;
        ws.init();
        hipStream_t s = ws.stream;
        const int C = k + 1;
        const int m = gi->num_individuals;
        const int E = gi->group_ptr[n];
        const size_t NC = static_cast<size_t>(n) * C;
        ws.stage.ensure(NC);
        XFE_CHECK(hipMemcpyAsync(ws.stage.p, y, sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (k > 0) XFE_CHECK(hipMemcpyAsync(ws.stage.p + n, X, sizeof(double) * n * k, hipMemcpyHostToDevice, s));
        double* V = ws.stage.p;
        const double* w = nullptr;
        if (weights) {
            ws.stage_w.ensure(n);
            XFE_CHECK(hipMemcpyAsync(ws.stage_w.p, weights, sizeof(double) * n, hipMemcpyHostToDevice, s));
            w = ws.stage_w.p;
        }
        const int grid = grid_for(n);
        ws.stage_ids.ensure(static_cast<size_t>(n) * std::max(num_fes, 1));
        std::vector<const int*> gid(num_fes);
        for (int d = 0; d < num_fes; ++d) {
            int* dst = ws.stage_ids.p + static_cast<size_t>(d) * n;
            XFE_CHECK(hipMemcpyAsync(dst, fes[d].group_ids, sizeof(int) * n, hipMemcpyHostToDevice, s));
            gid[d] = dst;
            const int G = fes[d].num_groups;
            ws.wsum[d].ensure(G);
            ws.sums[d].ensure(static_cast<size_t>(G) * C);
            if (fes[d].weight_sums) {
                XFE_CHECK(hipMemcpyAsync(ws.wsum[d].p, fes[d].weight_sums, sizeof(double) * G, hipMemcpyHostToDevice, s));
            } else {
                XFE_CHECK(hipMemsetAsync(ws.wsum[d].p, 0, sizeof(double) * G, s));
                XFE_CHECK(launch_gi_weight_sums(n, dst, w, ws.wsum[d].p, grid, s));
            }
            res->fe_levels[d] = G;
        }
        ws.gi_iptr.ensure(m + 1);
        ws.gi_igroup.ensure(std::max(E, 1));
        ws.gi_scale.ensure(n);
        ws.gi_denom.ensure(m);
        ws.gi_members.ensure(m);
        XFE_CHECK(hipMemcpyAsync(ws.gi_iptr.p, gi->individual_ptr, sizeof(int) * (m + 1), hipMemcpyHostToDevice, s));
        if (E > 0)
            XFE_CHECK(hipMemcpyAsync(ws.gi_igroup.p, gi->individual_group, sizeof(int) * E, hipMemcpyHostToDevice, s));
        XFE_CHECK(hipMemcpyAsync(ws.gi_scale.p, gi->group_scale, sizeof(double) * n, hipMemcpyHostToDevice, s));
        std::vector<int> members, offsets;
        gi_colour(*gi, members, offsets);
        XFE_CHECK(hipMemcpyAsync(ws.gi_members.p, members.data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
        XFE_CHECK(launch_gi_denom(ws.gi_iptr.p, ws.gi_igroup.p, ws.gi_scale.p, w, m, ws.gi_denom.p, grid_for(m), s));
        // sweep order: the caller's permutation, else 0..D-1 (src/fe_absorption_cuda.cu:4771-4800)
        std::vector<int> order(num_fes);
        for (int d = 0; d < num_fes; ++d) order[d] = d;
        if (sweep_order) {
            std::vector<char> seen(num_fes, 0);
            bool ok = true;
            for (int d = 0; d < num_fes && ok; ++d) {
                const int v = sweep_order[d];
                if (v < 0 || v >= num_fes || seen[v]) ok = false;
                else seen[v] = 1;
            }
            if (ok) for (int d = 0; d < num_fes; ++d) order[d] = sweep_order[d];
        }
        for (int d = 0; d < num_fes; ++d) res->sweep_order_used[d] = order[d];
        const bool symmetric = opt->method == 2;
        const double relax = 1.0;  // HdfeOptions::jacobi_relaxation's default
        auto std_demean = [&](int d) {
            const int G = fes[d].num_groups;
            XFE_CHECK(hipMemsetAsync(ws.sums[d].p, 0, sizeof(double) * G * C, s));
            XFE_CHECK(launch_gi_std_accumulate(V, n, C, gid[d], w, ws.sums[d].p, G, grid, s));
            XFE_CHECK(launch_gi_std_apply(V, n, C, gid[d], ws.sums[d].p, ws.wsum[d].p, grid, s));
        };
        auto individual_sweep = [&]() {
            for (size_t c = 0; c + 1 < offsets.size(); ++c) {
                const int cnt = offsets[c + 1] - offsets[c];
                if (cnt <= 0) continue;
                XFE_CHECK(launch_gi_color_sweep(V, n, C, ws.gi_iptr.p, ws.gi_igroup.p, ws.gi_scale.p, w,
                                                ws.gi_denom.p, ws.gi_members.p + offsets[c], cnt, relax,
                                                grid_for(static_cast<long long>(cnt) * kWave), s));
            }
        };
        auto full_sweep = [&]() {
            for (int d : order) std_demean(d);
            individual_sweep();
            if (symmetric) {
                individual_sweep();
                for (int j = num_fes - 1; j >= 0; --j) std_demean(order[j]);
            }
        };
        const int ngrid = grid_for(static_cast<long long>(NC));
        ws.partials.ensure(ngrid);
        std::vector<double> hpart(ngrid);
        auto norm = [&]() {
            XFE_CHECK(launch_gi_sumsq(V, static_cast<long long>(NC), ws.partials.p, ngrid, s));
            XFE_CHECK(hipMemcpyAsync(hpart.data(), ws.partials.p, sizeof(double) * ngrid, hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
            double t = 0.0;
            for (double v : hpart) t += v;
            return std::sqrt(t);
        };
        // change-of-norm rule (src/fe_absorption_cuda.cu:5119-5150)
        const double tol_eff = (opt->tol > 0.0 && opt->tolerance_mode == 1) ? std::min(opt->tol, 1e-9) : opt->tol;
        const double gi_tol = (tol_eff > 0.0 && opt->tolerance_mode == 1) ? std::min(tol_eff, 1e-12) : tol_eff;
        const int max_iter = opt->max_iter;
        double prev = norm();
        int iterations = 0;
        bool converged = false;
        for (int it = 0; it < max_iter; ++it) {
            full_sweep();
            const double cur = norm();
            const double rel = std::abs(cur - prev) / std::max(1.0, prev);
            prev = cur;
            if (rel < gi_tol) {
                iterations = it + 1;
                converged = true;
                break;
            }
        }
        iterations = converged ? std::max(1, iterations) : max_iter;
        // 16 polish sweeps, then certify; 64 more sweeps per failed certificate
        if (converged && iterations < max_iter) {
            const int extra = std::min(16, max_iter - iterations);
            for (int j = 0; j < extra; ++j) full_sweep();
            iterations += extra;
        }
        std::vector<double> Vt(NC), V0(NC);
        std::memcpy(V0.data(), y, sizeof(double) * n);
        if (k > 0) std::memcpy(V0.data() + n, X, sizeof(double) * n * k);
        auto fetch = [&]() {
            XFE_CHECK(hipMemcpyAsync(Vt.data(), V, sizeof(double) * NC, hipMemcpyDeviceToHost, s));
            XFE_CHECK(hipStreamSynchronize(s));
        };
        fetch();
        gi_certificate(V0.data(), Vt.data(), n, C, fes, num_fes, *gi, weights, tol_eff, *res);
        while (!res->precision_certified && iterations < max_iter) {
            const int batch = std::min(64, max_iter - iterations);
            for (int j = 0; j < batch; ++j) full_sweep();
            iterations += batch;
            fetch();
            gi_certificate(V0.data(), Vt.data(), n, C, fes, num_fes, *gi, weights, tol_eff, *res);
        }
        std::memcpy(y_out, Vt.data(), sizeof(double) * n);
        if (k > 0) std::memcpy(X_out, Vt.data() + n, sizeof(double) * n * k);
        res->iterations = iterations;
        res->sweeps = iterations;
        res->converged = res->precision_certified;
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 0;
    }
}

void xfe_release_workspace(void) { tls_ws.release(); }

void xfe_profile_enable(int on) { g_profile.enabled = on != 0; }
void xfe_profile_reset(void) {
    try {
        profile_resolve();
    } catch (const std::exception& e) {
        set_error(e.what());
    }
    g_profile.ms = 0.0;
    g_profile.launches = 0;
    g_profile.bytes = 0.0;
}
void xfe_profile_get(double* pass_ms, long long* pass_launches, double* pass_bytes) {
    try {
        profile_resolve();
    } catch (const std::exception& e) {
        set_error(e.what());
    }
    if (pass_ms) *pass_ms = g_profile.ms;
    if (pass_launches) *pass_launches = g_profile.launches;
    if (pass_bytes) *pass_bytes = g_profile.bytes;
}

}  // extern "C"
