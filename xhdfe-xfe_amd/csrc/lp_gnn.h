// The learned preconditioner's graph network (lp_gnn.hip): parameter layout,
// graph and target views, and the device workspace of one network evaluation.
#ifndef XFE_LP_GNN_H
#define XFE_LP_GNN_H

#include <hip/hip_runtime.h>

#include <vector>

#include "common.h"

namespace xfe {
namespace lp {

constexpr int kH = 64;      // width
constexpr int kInF = 8;     // firm input features
constexpr int kInW = 8;     // mover input features
constexpr int kLayers = 3;  // message-passing layers
constexpr int kFirmCap = 256;  // movers a firm's mean looks at: its first kFirmCap (ascending ids)

// Parameter vector (float32), in this order: Wf0 (8 x 64), bf0, Ww0 (8 x 64), bw0,
// then per layer l = 1..3: Wwl (128 x 64), bwl, Wfl (128 x 64), bfl; then wa
// (64), ba, we (64), be, ws (the weight of the pair feature in the pair head).
// Matrices are row-major (input x output).
constexpr int kOffWf0 = 0;
constexpr int kOffBf0 = kOffWf0 + kInF * kH;
constexpr int kOffWw0 = kOffBf0 + kH;
constexpr int kOffBw0 = kOffWw0 + kInW * kH;
constexpr int kOffL1 = kOffBw0 + kH;
constexpr int kLayerSize = 2 * (2 * kH * kH + kH);
constexpr int kOffWa = kOffL1 + kLayers * kLayerSize;
constexpr int kOffBa = kOffWa + kH;
constexpr int kOffWe = kOffBa + 1;
constexpr int kOffBe = kOffWe + kH;
constexpr int kOffWs = kOffBe + 1;
constexpr int kNumParams = kOffWs + 1;

constexpr int off_Ww(int l) { return kOffL1 + (l - 1) * kLayerSize; }
constexpr int off_bw(int l) { return off_Ww(l) + 2 * kH * kH; }
constexpr int off_Wf(int l) { return off_bw(l) + kH; }
constexpr int off_bf(int l) { return off_Wf(l) + 2 * kH * kH; }

// Device views of one mobility graph: firm features xf (F x 8), mover features
// xw (Wm x 8), firm -> mover edge lists (f_ptr, f_idx), mover -> firm edge lists
// (w_ptr, w_idx) with w_rank = the position of each mover -> firm edge in
// the firm's list (a firm's mean over movers takes the first kFirmCap of its
// list, like a fixed-size neighbour sample), and the firm pairs (pf, pg)
// scored by the pair head, with one feature pq per pair (null: 0).
struct Graph {
    const float* xf;
    const float* xw;
    const int* f_ptr;
    const int* f_idx;
    const int* w_ptr;
    const int* w_idx;
    const int* pf;
    const int* pg;
    const float* pq;
    const int* w_rank;
};

// Training targets: u_tgt / u_mask per firm, y_tgt per pair; nmask = sum of u_mask.
struct Targets {
    const float* u_tgt;
    const float* u_mask;
    const float* y_tgt;
    double nmask;
};

// Activations and gradients of one evaluation.  keep = store every layer (the
// backward pass needs them); without it two alternating buffers suffice.
struct Gnn {
    int F = 0, Wm = 0, np = 0;
    bool keep = false;
    DevBuf<float> hf, hw, af, aw, u, t;
    DevBuf<float> dhf, dhw, dtf, dtw, dagg, part;
    DevBuf<double> dpart;
    std::vector<double> hdpart;
    double last_loss_u = 0.0, last_loss_e = 0.0;

    void ensure(int F, int Wm, int np, bool keep);
    float* hf_at(int l);
    float* hw_at(int l);
    float* af_at(int l);
    float* aw_at(int l);
    // u (F) and t (np) from the parameters th (device)
    void forward(const Graph& g, const float* th, hipStream_t s);
    // after forward (keep = true): loss, and its gradient into grad (device, kNumParams)
    double loss_backward(const Graph& g, const float* th, const Targets& tg, float* grad, hipStream_t s);
    void release() {
        for (DevBuf<float>* b : {&hf, &hw, &af, &aw, &u, &t, &dhf, &dhw, &dtf, &dtw, &dagg, &part}) b->release();
        dpart.release();
    }
};

// w_rank of host edge lists (the position of each mover -> firm edge in the firm's list).
std::vector<int> host_w_rank(int F, int Wm, const int* f_ptr, const int* f_idx, const int* w_ptr, const int* w_idx);

// One Adam step (step counts from 1).
void adam_step(float* th, float* m, float* v, const float* g, int step, float lr, hipStream_t s);

}  // namespace lp
}  // namespace xfe

#endif
