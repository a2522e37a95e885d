// The aggregate-block preconditioner of the mixed-precision Schur PCG
// (lp_precond.hip): firm pairs coupled by movers, their coupling weights and
// strengths, a parallel handshake aggregation over pair scores, and the exact
// inverse of each aggregate's block of the Schur complement.
//
// On the firm levels the Schur complement S = D_R' M_p D_R of the worker
// dimension is a weighted graph Laplacian (plus the year coupling):
//   S_ff = sum_w (L_wf - L_wf^2 / n_w),  S_fg = -W_fg,  W_fg = sum_w L_wf L_wg / n_w
// (L_wf: rows of worker w at firm f, n_w: rows of w).  Its slowest modes under
// Jacobi scaling live on small groups of firms coupled to each other far more
// than to the rest; solving each such group's block exactly removes them.
#ifndef XFE_LP_PRECOND_H
#define XFE_LP_PRECOND_H

#include <hip/hip_runtime.h>

#include <vector>

#include "common.h"

namespace xfe {
namespace lp {

// The segment records the preconditioner is built from (worker-major order).
struct SegView {
    int Nw = 0;                    // workers (segment groups)
    int G1 = 0;                    // firms
    const int* wseg = nullptr;     // Nw + 1 offsets
    const int* r1 = nullptr;       // firm of each segment
    const int* len = nullptr;      // rows of each segment
    const double* inv_w = nullptr; // 1 / n_w
    const double* diag = nullptr;  // sum over segments of L - L^2 / n_w per firm (G1)
    const double* rows = nullptr;  // rows per firm (G1)
};

// Device view of the blocks for the apply kernel.
struct BlockView {
    int nb = 0;                    // firms inside blocks
    int nblocks = 0;               // blocks
    const int* heads = nullptr;    // first blocked-firm position of each block
    const int* firm = nullptr;     // blocked firms, grouped by block, ascending inside a block
    const int* start = nullptr;    // per blocked firm: first position of its block
    const int* size = nullptr;     // per blocked firm: its block's size
    const int* row = nullptr;      // per blocked firm: offset of its row of the block inverse
    const float* inv = nullptr;    // block inverses, row-major, block after block
};

// z[f] = sum_g inv[f][g] r[g] for the blocked firms (level-major r, z with C
// columns; r and z point at the firm levels).
hipError_t launch_block_apply(const BlockView& b, const float* r, float* z, int C, hipStream_t s);

struct AggPrecond {
    int G1 = 0;
    int U = 0;             // unique firm pairs (f < g) with a mover between them
    long long emitted = 0; // segment pairs emitted
    int nblocks = 0, nb = 0, cap = 0, rounds = 0;
    DevBuf<long long> wcnt, woff;
    DevBuf<unsigned long long> key_a, key_b;
    DevBuf<double> val_a, val_b;
    DevBuf<int> head, run, pos;
    DevBuf<int> pf, pg;           // unique pairs f < g
    DevBuf<double> W;             // coupling weights
    DevBuf<double> dex;           // exact firm diagonal
    DevBuf<float> st;             // strengths W / sqrt(d_f d_g)
    DevBuf<int> agg, asize, newroot, cnt, bsz, bsq, bstart, boff2, skey, sval, iota, loc;
    DevBuf<unsigned long long> best;
    DevBuf<int> b_firm, b_start, b_size, b_row, b_head, b_hflag, b_hidx;
    DevBuf<double> bmat;
    DevBuf<float> binv;
    DevBuf<unsigned char> tmp;
    int* host_int = nullptr;      // pinned scratch for counts

    // Pairs, weights, the exact diagonal and the strengths.  Returns false
    // when there is no pair or too many to index with int.
    bool build_pairs(const SegView& sv, hipStream_t s);
    // The same from host arrays: U pairs f < g with weights w, the exact
    // diagonal d (G1); the strengths on the device.  Returns false when empty.
    bool set_pairs(int G1, const std::vector<int>& f, const std::vector<int>& g, const std::vector<double>& w,
                   const std::vector<double>& d, hipStream_t s);
    // Handshake aggregation over the U pairs: per round every aggregate picks
    // its best-scoring pair (ties: lower pair index) to an aggregate it can
    // join without exceeding cap firms, and mutual picks merge.
    void aggregate(const float* score, int cap, int rounds, hipStream_t s);
    // Blocks of the aggregates with two or more firms and their inverses.
    void build_blocks(hipStream_t s);
    BlockView view() const {
        return BlockView{nb, nblocks, b_head.p, b_firm.p, b_start.p, b_size.p, b_row.p, binv.p};
    }
    void release();
};

// The learned preconditioner's graph (lp_learned.hip): firms and movers
// (workers with segments at two or more firms), one edge per (mover, firm)
// with the firm listed once, both CSR directions, the input features of firms
// and movers; the pairs of an AggPrecond are the scored firm pairs.
struct LearnedGraph {
    int F = 0, Wm = 0, E = 0;
    DevBuf<int> mcnt, mflag, moff, midx, w_ptr, w_idx, e_mov, f_ptr, f_idx, skey, eidx, sidx, w_rank;
    DevBuf<int> hkey, hkey2, hoff;  // half-edges of the pairs (both ends), sorted by firm
    DevBuf<float> hval, hval2;
    DevBuf<float> xf, xw;
    DevBuf<unsigned char> tmp;
    int* host_int = nullptr;
    void build(const AggPrecond& ag, const SegView& sv, hipStream_t s);
    void release();
};

// Aggregation scores of the learned preconditioner: the graph network with
// the parameters params (host, float32) scores each pair of ag (after
// build_pairs): score = t_fg + (u_f + u_g) / 2.  Returns a device array of
// ag.U scores.
const float* learned_pair_scores(AggPrecond& ag, const SegView& sv, const float* params, hipStream_t s);

// out = beta + t_c x (the current iterate of the mixed-precision CG, FP64)
hipError_t launch_probe_iterate(const double* beta, const float* x, const double* t, int N, int C, double* out,
                                hipStream_t s);

// Training data of one panel, recorded by a mixed-precision solve in probe mode
// (xfe_lp_probe): the iterate at inner iteration probe_iter and the final
// solution give the error vectors on the firm levels; the pairs, graph and
// features are the ones the learned preconditioner would see.
struct ProbeState {
    bool active = false;
    bool done = false;
    int probe_iter = 0;
    int C = 0, F = 0, Wm = 0, E = 0, U = 0, inner = 0;
    DevBuf<double> iter;            // N x C iterate at probe_iter
    std::vector<double> err;        // F x C (firm-major): final - iterate
    std::vector<float> xf, xw, pq;
    std::vector<int> f_ptr, f_idx, w_ptr, w_idx, pf, pg;
    std::vector<double> W, dex;
};
ProbeState& probe_state();

}  // namespace lp
}  // namespace xfe

#endif
