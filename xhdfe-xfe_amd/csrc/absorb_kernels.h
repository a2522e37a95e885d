// Launch interface of the absorber kernels (absorb_kernels.hip).
#ifndef XFE_ABSORB_KERNELS_H
#define XFE_ABSORB_KERNELS_H

#include <hip/hip_runtime.h>

namespace xfe {

namespace lp {
struct BlockView;
}

constexpr int kBlock = 256;
constexpr int kWave = 64;

// how a pass accumulates level sums of the dimension it starts (0: no reduce)
enum { kRedNone = 0, kRedLds = 1, kRedSorted = 2, kRedGlobal = 3, kRedSpell = 4 };
constexpr int kTileRows = 512;  // rows per level-aligned tile (tile_kernels.hip)
constexpr int kXcdCopies = 8;  // XCD-private accumulator copies (MI355X: 8 XCDs)

// Arguments of the fused streaming pass (see pass_kernel).
struct PassArgs {
    int n = 0;
    int ld = 0;
    int C = 0;   // total columns (stride of the level-major tables)
    int c0 = 0;  // first column handled by this launch
    const double* src = nullptr;
    double* dst = nullptr;
    const double* P = nullptr;  // optional IT extrapolation v = P - coef (P - Q)
    const double* Q = nullptr;
    double coef = 0.0;
    double* xout = nullptr;
    const int* gid_a = nullptr;  // projection being finished
    const double* means_a = nullptr;
    const int* gid_r = nullptr;  // projection being started
    double* sums_r = nullptr;
    int G_r = 0;
    const double* w = nullptr;
    const double* it_x = nullptr;
    const double* it_gx = nullptr;
    double* partials = nullptr;
    int one = 0;           // values are 1 (weight sums)
    int xcd_copies = 1;    // kRedGlobal: private accumulator copies
    size_t copy_stride = 0;
    int lds_copies = 1;    // kRedLds: LDS tables per block (set by the launcher)
};

// Arguments of the tiled pass (tile_kernels.hip).
struct TileArgs {
    int n = 0;
    int ld = 0;
    int C = 0;
    int c0 = 0;
    const double* src = nullptr;
    double* dst = nullptr;
    const double* P = nullptr;
    const double* Q = nullptr;
    double coef = 0.0;
    double* xout = nullptr;
    const int* gid_a = nullptr;
    const double* means_a = nullptr;
    const int* tile_start = nullptr;  // num_tiles + 1 row offsets
    int num_tiles = 0;
    const int* gid_p = nullptr;       // the level-sorted dimension
    const double* inv_p = nullptr;
    const int* gid_r = nullptr;
    int G_r = 0;
    double* sums_r = nullptr;         // kRedLds flush target
    double* spell_sums = nullptr;     // kRedSpell output (spells x C)
    const int* tile_spell_base = nullptr;
    const double* w = nullptr;
    const double* it_x = nullptr;
    const double* it_gx = nullptr;
    double* partials = nullptr;
    double* maxp_partials = nullptr;  // per-block max |mean| of p (polish)
};

hipError_t launch_tile_pass(const TileArgs& a, int cc, int rr, bool pp, int nq, int grid,
                            hipStream_t s);
hipError_t launch_spell_pull(const double* S, int C, int c0, int cc, const int* order,
                             const int* skey, int M, double* sums, int grid, hipStream_t s);
hipError_t launch_level_count(const int* gid, int n, int* cnt, int grid, hipStream_t s);
// rows per level into cnt (zeroed by the caller, G ints) and their FP64 copy
// wsum, for ids that are non-decreasing (sorted: each level's rows
// contiguous) or that have at most kLevelCountLds levels (an LDS histogram);
// other dimensions count through the FP64 pass, whose XCD-private copies of
// the level table beat global integer atomics there
constexpr int kLevelCountLds = 8192;
hipError_t launch_level_weights(const int* gid, int n, int G, bool sorted, int* cnt, double* wsum, int grid,
                                hipStream_t s);
hipError_t launch_max_int(const int* v, int m, int* out, int grid, hipStream_t s);
// out = max over m of (int) v (level sizes from unit-weight sums)
hipError_t launch_max_count(const double* v, int m, int* out, int grid, hipStream_t s);
hipError_t launch_tile_start(const int* gid, int n, int step, int num_tiles, int* tile_start,
                             int grid, hipStream_t s);
hipError_t launch_spell_heads(const int* gp, const int* gr, int n, int* flags, int grid,
                              hipStream_t s);
hipError_t launch_spell_keys(const int* flags, const int* inc, const int* gr, int n, int* skey,
                             int grid, hipStream_t s);
hipError_t launch_tile_spell_base(const int* tile_start, int num_tiles, const int* inc, int n,
                                  int* base, int grid, hipStream_t s);

// Arguments of the Schur-complement pass of the mixed-precision solver
// (fast_kernels.hip).  T = float for the inner CG, double for the refinement.
constexpr int kMaxCgCols = 32;  // columns solved together by the mixed-precision path
template <typename T>
struct SchurArgs {
    int n = 0;
    int ld = 0;
    int C = 0;
    int c0 = 0;
    const T* V = nullptr;   // data block (refinement residual / final transform)
    T* dst = nullptr;       // final transform output
    const int* tile_start = nullptr;
    int num_tiles = 0;
    const int* gid_p = nullptr;
    const T* inv_p = nullptr;
    T sign = T(1);
    const int* gid_r[2] = {nullptr, nullptr};
    const T* coef_r[2] = {nullptr, nullptr};  // level-major G_r x C coefficients
    int G_r[2] = {0, 0};
    int mode_r[2] = {0, 0};                   // kRedSpell or kRedLds
    T* spell_sums[2] = {nullptr, nullptr};
    const int* spell_base[2] = {nullptr, nullptr};
    T* lds_target[2] = {nullptr, nullptr};
    int lds_off[2] = {0, 0};
    int lds_copies[2] = {1, 1};               // LDS-mode sums: copies (one per wave when they fit)
    double* partials = nullptr;               // STORE: per-block sum of squares
    T* p_sums = nullptr;                      // STORE: per level of p, the sums of the stored rows (G_p x C)
};

// Arguments of the segment form of the Schur operator (seg_kernels.hip).
constexpr int kSegMaxRange = 200;  // most levels of the range dimension r2 (LDS tables <= 64 KB)
template <typename T>
struct SegArgs {
    int C = 0;                    // stride of the level-major coefficient tables
    int c0 = 0;                   // first column of this chunk
    int Nw = 0;                   // levels of the eliminated dimension p
    int M = 0;                    // segments
    int G2 = 0;                   // levels of r2 (0: no r2)
    const T* x1 = nullptr;        // coefficients of r1 (G1 x C)
    const T* x2 = nullptr;        // coefficients of r2 (G2 x C)
    const T* inv_w = nullptr;     // 1 / n_w (per worker rank)
    int Nm = 0;                   // movers: workers with two or more segments
    const int* mlist = nullptr;   // worker rank of each mover
    const T* inv_wm = nullptr;    // 1 / n_w per mover
    const int* wseg = nullptr;    // Nw + 1 segment offsets per level of p
    const int* seg_r1 = nullptr;  // worker-major segment records
    const int* seg_len = nullptr;
    const int* seg_a = nullptr;
    T* wmean = nullptr;           // Nm x cc means of u over the movers
    const int* f_key = nullptr;   // firm-major segment records (sorted by r1)
    const int* f_w = nullptr;     // mover index (segments of movers only)
    const int* f_len = nullptr;   // length; negative for the one segment of a stayer
    const int* f_a = nullptr;
    T* out1 = nullptr;            // (S x)_1, zeroed before the pass
    double* hpart = nullptr;      // per-block (G2 + 1) x cc year differences
    int* bkey = nullptr;          // boundary list of the firm reduction:
    T* bval = nullptr;            // 2 entries (key, cc values) per wave of segments
};

hipError_t launch_seg_heads(const int* gp, const int* g1, const int* g2, int n, int* flags, int grid,
                            hipStream_t s);
hipError_t launch_seg_fill(const int* flags, const int* inc, const int* gp, const int* g1, const int* g2,
                           int n, int* seg_start, int* seg_w, int* seg_r1, int* seg_a, int grid,
                           hipStream_t s);
hipError_t launch_seg_len(const int* seg_start, int M, int n, int* seg_len, int grid, hipStream_t s);
hipError_t launch_seg_offsets(const int* key, int M, int Nw, int* off, int grid, hipStream_t s);
// flag[w] = 1 for workers with two or more segments
hipError_t launch_seg_movers(const int* wseg, int Nw, int* flag, int grid, hipStream_t s);
// mlist[idx[w]] = w for movers; firm-major records: f_w = mover index for a
// mover's segments, f_len negated for a stayer's
hipError_t launch_seg_mover_relabel(const int* flag, const int* idx, int Nw, int M, int* mlist, int* f_w, int* f_len,
                                    int grid, hipStream_t s);
hipError_t launch_seg_primary(const int* wseg, const int* seg_r1, int Nw, int G1, int* pf, int grid,
                              hipStream_t s);
hipError_t launch_seg_rank(const int* perm, int m, int* rank, int grid, hipStream_t s);
hipError_t launch_seg_empty(const int* off, int G, int* flag, int grid, hipStream_t s);
template <typename T>
hipError_t launch_seg_gather(const T* src, const int* perm, T* dst, int m, int grid, hipStream_t s);
// pass 1 (worker means) and pass 2 (firm sums, year differences; diag: the Jacobi diagonal)
template <typename T>
hipError_t launch_seg_worker(const SegArgs<T>& a, int cc, bool has_y, int grid, hipStream_t s);
template <typename T>
hipError_t launch_seg_firm(const SegArgs<T>& a, int cc, bool has_y, bool diag, int grid, hipStream_t s);
template <typename T>
hipError_t launch_seg_bnd(const int* kin, const T* vin, int m, T* out, int C, int c0, int cc, int* kout, T* vout,
                          int final, int grid, hipStream_t s);
// the remaining levels (list of m entries) in one block of kBndTail threads
constexpr int kBndTail = 1024;
template <typename T>
hipError_t launch_seg_bnd_tail(int* k0, T* v0, int* k1, T* v1, int m, T* out, int C, int c0, int cc,
                               hipStream_t s);
template <typename T>
hipError_t launch_seg_year(const double* hpart, int nblk, int G2, int cc, double* tot, const double* counts,
                           const T* x2, int C, int c0, const double* K, T* out2, hipStream_t s);
// stayer segments (negative length) counted by (first year, length) into H (G2 x (G2 + 1), zeroed)
hipError_t launch_seg_stayer_hist(const int* f_len, const int* f_a, int M, int G2, int* H, int grid, hipStream_t s);
// flag[j] = 1 for the records with a positive length (movers' segments)
hipError_t launch_seg_mover_segments(const int* f_len, int M, int* flag, int grid, hipStream_t s);
// the flagged records, in order, at positions idx (exclusive scan of flag); a_in may be null
hipError_t launch_seg_compact(const int* flag, const int* idx, int M, const int* k_in, const int* w_in,
                              const int* l_in, const int* a_in, int* k_out, int* w_out, int* l_out, int* a_out,
                              int grid, hipStream_t s);

template <typename T>
hipError_t launch_schur(const SchurArgs<T>& a, int cc, int nr, bool has_v, bool store,
                        size_t lds_extra, int grid, hipStream_t s);
template <typename T>
hipError_t launch_pull(const T* S, int C, int c0, int cc, const int* order, const int* skey, int M,
                       T* sums, int grid, hipStream_t s);
template <typename T>
hipError_t launch_vdot(const T* a, const T* b, int N, int C, double* partials, int grid, hipStream_t s);
// the FP64 Schur PCG (precision 2): y += sign coef[c] x; z = minv[level] r;
// p = z + beta[c] p over level-major N x C vectors (total = N C)
hipError_t launch_s64_axpy(double* y, const double* x, const double* coef, long long total, int C, double sign,
                           int grid, hipStream_t s);
hipError_t launch_s64_jacobi(const double* r, const float* minv, double* z, long long total, int C, int grid,
                             hipStream_t s);
hipError_t launch_s64_dir(double* p, const double* z, const double* beta, long long total, int C, int grid,
                          hipStream_t s);
// Scalar block of the device-side CG (doubles, kMaxCgCols per field).
enum {
    kCgAlpha = 0 * kMaxCgCols, kCgBeta = 1 * kMaxCgCols, kCgRz = 2 * kMaxCgCols,
    kCgActive = 3 * kMaxCgCols, kCgPq = 4 * kMaxCgCols, kCgRr = 5 * kMaxCgCols,
    kCgRzNew = 6 * kMaxCgCols, kCgIters = 7 * kMaxCgCols, kCgOk = 8 * kMaxCgCols,
    kCgTol2 = 9 * kMaxCgCols,  // reliable-update CG: per-column r.r that calls for an FP64 residual
    kCgCount = 10 * kMaxCgCols, kCgUpd = 10 * kMaxCgCols + 1, kCgSlots = 10 * kMaxCgCols + 8
};
// grid for the CG vector kernels: a multiple of C (each thread keeps one column)
inline int cg_grid(long long total, int C, int max_grid) {
    long long per = (total + static_cast<long long>(kBlock) * C - 1) / (static_cast<long long>(kBlock) * C);
    const long long cap = max_grid / C > 0 ? max_grid / C : 1;
    if (per > cap) per = cap;
    if (per < 1) per = 1;
    return static_cast<int>(per * C);
}
// pq -> alpha (cg_dot + cg_alpha); part: grid x C doubles (2 x for cg_step).
// T = float (mixed-precision inner solves) or double (AKM/KSS solves).
template <typename T>
hipError_t launch_cg_dot(const T* a, const T* b, long long total, int C, double* cg, double* part, int grid,
                         hipStream_t s);
// x, r, z update, r.r and r.z -> beta, per-column iteration counts, the
// max_iter cap and the active count, then p = z + beta p
template <typename T>
hipError_t launch_cg_step(T* x, T* r, T* z, T* p, const T* q, const T* minv, long long total, int C, double* cg,
                          double* part, double tol2, int max_iter, int grid, hipStream_t s);
template <typename T>
hipError_t launch_cg_init(const double* b, const double* scale, const T* minv, T* x, T* r, T* z, T* p,
                          long long total, int C, double* cg, double* part, int grid, hipStream_t s);
hipError_t launch_refine_update(double* beta, const float* x, const double* inv_scale, int N, int C,
                                int grid, hipStream_t s);
// Reliable-update CG (mixed-precision inner solve): an iteration's step with the
// update test (r.r <= cg[kCgTol2 + c] sets cg[kCgUpd]) instead of the stopping
// test, then p = z + beta p.
// blk (may be null): the aggregate blocks of the firm levels, which start at
// level firm_off: z = inv r on the blocked firms, minv r elsewhere.
hipError_t launch_cg_step_rel(float* x, float* r, float* z, float* p, const float* q, const float* minv,
                              long long total, int C, double* cg, double* part, int max_iter, int grid,
                              const lp::BlockView* blk, int firm_off, hipStream_t s);
// Residual replacement: r = r64 / t_c, z = minv r, p = f_c (p - z_old) + z
// (p = z for a restarted column, f_c = 0), x = 0; then per column rz, the
// activity act_c and the update threshold tol2_c.
// rel: [inv_t | f | tol2 | act], kMaxCgCols each.
hipError_t launch_cg_reliable_reset(const double* r64, const double* rel, const float* minv, float* x, float* r,
                                    float* z, float* p, long long total, int C, double* cg, double* part, int grid,
                                    const lp::BlockView* blk, int firm_off, hipStream_t s);
hipError_t launch_schur_diag(const int* tile_start, int num_tiles, const int* gid_p,
                             const int* gid_r, const double* inv_p, double* diag, int grid,
                             hipStream_t s);
hipError_t launch_diag_invert(const double* diag, const double* counts, float* minv, int N,
                              int grid, hipStream_t s);
hipError_t launch_to_float(const double* a, float* b, long long m, int grid, hipStream_t s);

hipError_t launch_pass(const PassArgs& a, int cc, int rmode, int nq, int grid, hipStream_t s);
hipError_t launch_finalize(double* sums, double* means, const double* inv, int G, int C,
                           int copies, double* partials, int grid, hipStream_t s);
hipError_t launch_reduce_partials(const double* partials, int count, int nq, int mode, double* out,
                                  hipStream_t s);
hipError_t launch_it_stats(const double* x, const double* gx, const double* ggx, int n, int C,
                           int ld, double* partials, int grid, hipStream_t s);
// sum, sum of squares, max and count of entries <= 1 + 1e-12 of a level table
// (G values) into out[0..3]; part: 4 x grid scratch
hipError_t launch_level_profile(const double* ws, int G, double* part, int grid, double* out, hipStream_t s);
hipError_t launch_col_dot(const double* a, const double* b, int n, int C, int ld, double* partials,
                          int grid, hipStream_t s);
// per-column sums of squares of a level-major table (certificate), grid from cg_grid
constexpr int kMaxTableCols = 256;
hipError_t launch_table_sumsq(const double* a, long long total, int C, double* part, int grid, hipStream_t s);
hipError_t launch_col_axpy(double* y, const double* a, const double* b, const double* alpha, int n,
                           int C, int ld, double sign, int grid, hipStream_t s);
hipError_t launch_sub(double* y, const double* a, const double* b, int n, int C, int ld, int grid,
                      hipStream_t s);
// heterogeneous slopes: level moments, level sums of w v and w z v, the
// per-level (alpha, gamma) with the reference's rank rule, and the apply
hipError_t launch_slope_wdot(const double* a, const double* b, const double* w, int n, int C, int ld,
                             double* partials, int grid, hipStream_t s);
hipError_t launch_slope_moments(const int* gid, const double* z, const double* w, int n, double* mom, int grid,
                                hipStream_t s);
hipError_t launch_slope_reduce(const double* V, int ld, int C, int n, const int* gid, const double* z,
                               const double* w, double* sums, int G, int copies, int grid, hipStream_t s);
hipError_t launch_slope_coef(const double* sums, const double* mom, int G, int C, int mode, int copies,
                             double* coef, int grid, hipStream_t s);
hipError_t launch_slope_apply(double* V, int ld, int C, int n, const int* gid, const double* z, const double* coef,
                              int grid, hipStream_t s);
hipError_t launch_inverse(const double* ws, double* inv, int G, int grid, hipStream_t s);
hipError_t launch_iota(int* v, int n, int grid, hipStream_t s);
hipError_t launch_gather_ids(const int* src, const int* perm, int* dst, int n, int grid,
                             hipStream_t s);
hipError_t launch_gather_cols(const double* src, int ld_src, const int* perm, double* dst,
                              int ld_dst, int n, int C, int grid, hipStream_t s);
hipError_t launch_scatter_cols(const double* src, int ld_src, const int* perm, double* dst,
                               int ld_dst, int n, int C, int grid, hipStream_t s);
hipError_t launch_gather_vec(const double* src, const int* perm, double* dst, int n, int grid,
                             hipStream_t s);
hipError_t launch_sorted_check(const int* gid, int n, int* flag, int grid, hipStream_t s);

// Device time and algorithmic bytes of the dominant kernel of the last solves
// (xfe_profile_*): a pair of HIP events around each profiled launch, recorded
// without waiting; their times are read when the numbers are asked for.
struct Profile {
    bool enabled = false;
    double ms = 0.0;
    long long launches = 0;
    double bytes = 0.0;
};
extern Profile g_profile;
// Records the start of a profiled launch on s; returns its slot (-1 when
// profiling is off).
int profile_begin(hipStream_t s);
// Records the end of the launch in `slot` with its algorithmic bytes.
void profile_end(int slot, hipStream_t s, double bytes);
// Waits for the recorded launches and adds them to g_profile.
void profile_resolve();

// FE recovery: rows in the caller's order; gid[d] n ids, alpha[d] level
// coefficients of dimension d (alpha[p] is being formed by recover_primary).
struct RecoverArgs {
    const double* partial = nullptr;
    const double* vt = nullptr;  // within transform of partial
    const double* w = nullptr;
    const int* gid[3] = {nullptr, nullptr, nullptr};
    const double* alpha[3] = {nullptr, nullptr, nullptr};
    int D = 0;
    int p = 0;
};
hipError_t launch_recover_primary(const RecoverArgs& a, double* sums, int n, int grid, hipStream_t s);
hipError_t launch_recover_divide(double* alpha, const double* wsum, int G, int grid, hipStream_t s);
hipError_t launch_recover_residual(const RecoverArgs& a, double* res, int n, int grid, hipStream_t s);
hipError_t launch_recover_expand(const RecoverArgs& a, double* contrib, int n, int grid, hipStream_t s);

// group()/individual() absorber (gi_kernels.hip); V is n x C, ld n.
hipError_t launch_gi_std_accumulate(const double* V, int n, int C, const int* gid, const double* w, double* sums,
                                    int G, int grid, hipStream_t s);
hipError_t launch_gi_std_apply(double* V, int n, int C, const int* gid, const double* sums, const double* wsum,
                               int grid, hipStream_t s);
hipError_t launch_gi_weight_sums(int n, const int* gid, const double* w, double* wsum, int grid, hipStream_t s);
hipError_t launch_gi_denom(const int* ind_ptr, const int* ind_group, const double* scale, const double* w, int m,
                           double* denom, int grid, hipStream_t s);
hipError_t launch_gi_color_sweep(double* V, int n, int C, const int* ind_ptr, const int* ind_group,
                                 const double* scale, const double* w, const double* denom, const int* members,
                                 int count, double relax, int grid, hipStream_t s);
hipError_t launch_gi_sumsq(const double* V, long long total, double* partials, int grid, hipStream_t s);

}  // namespace xfe



#endif
