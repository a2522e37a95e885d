// Wide batched solver of the AKM/KSS grounded firm Laplacian
//     S = Df - B' Dw^{-1} B
// for up to kWideMaxLanes right-hand sides per operator pass: FP32 Jacobi-PCG
// whose residual is replaced by the FP64 residual b - S x (x accumulated in
// FP64) each time the FP32 r.r has dropped 1e4-fold, so the
// answer is certified in FP64 while the bulk of the passes move FP32 lanes.
//
// S is applied in its firm-pair form: per worker, diag(c) - c c' / Dw is the
// Laplacian of the complete graph on the worker's firms with edge weights
// c_a c_b / Dw (plus c_a (1 - C_w / Dw) on the diagonal when Dw differs from
// the worker's count total C_w), so
//     (S p)_f = sum_g W_fg (p_f - p_g) + corr_f p_f,
// W_fg summed over the workers employed by both firms and corr = Df - deg plus
// those diagonal terms.  Differences of neighbouring firms' values are formed
// before they are weighted, so FP32 keeps the accuracy of the slow (smooth)
// modes that a worker mean t_w = sum (c / Dw) p_f, rounded to FP32, loses.
#ifndef XFE_AKM_WIDE_H
#define XFE_AKM_WIDE_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "common.h"
#include "lp_precond.h"

namespace xfe {

constexpr int kWideMaxLanes = 256;

// Host-side records handed over by the context (workers renumbered by their
// primary firm): worker-major matches (rf firm, rc count) with worker offsets
// wofs (N + 1) and Dw per worker; per firm corr = Df - deg and inv_diag.
struct AkmWideInput {
    int N = 0, J = 0, M = 0, ground = 0;
    const int* wofs = nullptr;
    const int* rf = nullptr;
    const double* rc = nullptr;
    const double* dw = nullptr;
    const double* corr = nullptr;
    const double* inv_diag = nullptr;
};

// An operator other than the firm-pair form (the absorber's slope path): q =
// S p on level-major n_rows x L vectors, FP32 for the PCG passes and FP64 for
// the residual replacements.
struct WideOp {
    virtual void apply(const float* p, float* q) = 0;
    virtual void apply(const double* p, double* q) = 0;
    // a block preconditioner of its own: z = B^-1 r on its blocks (z keeps
    // the Jacobi values elsewhere)
    virtual bool has_blocks() const { return false; }
    virtual void precondition(const float* r, float* z) { (void)r; (void)z; }
    virtual ~WideOp() = default;
};

struct AkmWide {
    int J = 0, ground = 0;
    int E = 0;  // directed firm-pair entries (firm-major)
    int L = 0;  // lane stride of the allocated vectors (64, 128 or 256)
    hipStream_t s = nullptr;
    DevBuf<int> fkey, fnb, fwk, has;
    DevBuf<float> w32, corr32, minv32;
    DevBuf<double> w64, corr64;
    DevBuf<int> bk[2];
    DevBuf<float> bv32[2];
    DevBuf<double> bv64[2];
    DevBuf<float> x, r, z, p, q;
    DevBuf<double> x64, b64, r64, q64, part, sc;
    // aggregate blocks of strongly coupled firms solved exactly inside the
    // preconditioner (Jacobi on the other firms)
    lp::AggPrecond agg;
    bool blocks = false;
    // per-lane entry weights (resampling): entry e of worker fwk[e] weighs
    // base_e mult[fwk[e] * mstride + l / mcols] in lane l; per-lane Jacobi
    // inverse diagonal minv_lane (J x L, FP32) in place of minv32
    const int* mult = nullptr;
    int mstride = 0, mcols = 1;
    const float* minv_lane = nullptr;
    // a custom operator (then J is its row count, ground -1 for none, minv32
    // its Jacobi inverse diagonal, and 32 lanes are allowed)
    WideOp* op = nullptr;
    double* h_sc = nullptr;  // pinned per-lane scalars
    int last_updates = 0;    // FP64 residual replacements of the last solve
    int last_passes = 0;     // FP32 operator passes of the last solve
    const char* trace = nullptr;  // when set: one line per FP64 replacement on stderr, with this tag
    // restart the direction (p = z) at every FP64 replacement: restarted
    // iterative refinement, each FP32 run a fresh PCG on the current residual
    // (robust where the FP32 operator loses the slow modes' accuracy); by
    // default the direction is kept unless the recurrence outran FP32
    bool restart_each_replacement = false;

    // use_blocks: build the aggregate blocks (cap firms per block, rounds of
    // handshake aggregation over the coupling strengths)
    void build(const AkmWideInput& in, hipStream_t stream, bool use_blocks);
    void release();
    // rhs_pack / z_pack: lane-major arrays (nb x J), on the host or, with
    // device_ptrs, on the device; iters_out[l]: FP32 iterations of lane l, -1
    // when it did not converge (its solution zero)
    void solve(const double* rhs_pack, double* z_pack, int nb, double tol, int max_iter, int* iters_out,
               bool device_ptrs);
    // the same on right-hand sides already on the device, level-major J x L
    // FP64 (ground rows ignored); leaves the unscaled solutions level-major in
    // x64 and the per-lane results in h_sc
    // accept (> tol): a lane whose FP64 residual stalls (fails to halve over
    // four replacements in a row) counts as converged at |r| <= accept; the
    // default keeps the strict rule |r| <= tol
    void solve_device(int nb, double tol, int max_iter, double accept = 0.0);

    // entries given per worker, not merged (f, g, worker, base weight), no
    // diagonal correction, no blocks: the operator sum_w mult_w L_w of a
    // resample once mult is set
    void build_worker_pairs(int J, int ground, const std::vector<int>& f, const std::vector<int>& g,
                            const std::vector<int>& w, const std::vector<double>& base, hipStream_t stream);
    // vectors for nb lanes (sets L)
    void ensure(int nb);
    // a custom operator over n_rows levels with the Jacobi inverse diagonal
    // minv (device, FP32, n_rows); no pair entries
    void set_operator(WideOp* o, int n_rows, hipStream_t stream, const float* minv);
    // x64 = bn x64 (the unscaled solutions; lanes that did not converge zero)
    void unscale_x64();
    // iterations of lane l in the last solve, -1 when it did not converge
    int lane_iterations(int l) const;
    // after a caller's own continuation of the last solve on x64 (scaled
    // units): lane l (< nb) converged there when ok[l], with extra[l] more
    // iterations (unscale_x64 then keeps its solution)
    void mark_converged(int nb, const std::vector<char>& ok, const std::vector<int>& extra);
    // lane l's final scalars as text (diagnostics)
    std::string lane_report(int l) const;

  private:
    template <typename T>
    void apply(const T* pv, T* out, const T* w_, const T* corr_, DevBuf<T>* bv);
};

}  // namespace xfe

#endif
