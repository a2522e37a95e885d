/* xfe_hip.h -- C ABI of the MI355X (gfx950) fixed-effects absorber.
 *
 * Plain C: POD structs, pointers and sizes, int status codes, no exceptions,
 * no Eigen/torch types.  Each entry point names the reference interface it
 * replaces (xhdfe-xfe, file:line).  A C++ caller that used
 * hdfe::detail::absorb_fixed_effects_cuda binds these through the thin adapter
 * shown in INTEGRATION.md.
 *
 * Threading: one host thread per call; the library keeps one device
 * workspace per calling thread (thread_local), like the reference's
 * persistent CUDA workspace (src/fe_absorption_cuda.cu:1536).
 */
#ifndef XFE_HIP_H
#define XFE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XFE_MAX_DIMS 8
#define XFE_MAX_COLS 256

/* Per-dimension view, mirrors hdfe::detail::GpuFeInput
 * (src/fe_absorption_cuda.hpp:12-21).  group_ids are compact 0-based ids
 * (length n); weight_sums has num_groups entries (sum of weights, or counts). */
typedef struct {
    const int* group_ids;
    int num_groups;
    int num_levels_present;
    const double* weight_sums;
    int is_slope;              /* heterogeneous slope term on this dimension (xfe_absorb only): each
                                  level is projected onto span{1, z} (with slope_has_intercept) or
                                  span{z}, z = slope_values (n, host) */
    int slope_has_intercept;
    const double* slope_values;
    const double* slope_sum_z;
    const double* slope_sum_zz;
} xfe_fe_input;

/* Options: the subset of hdfe::HdfeOptions (include/hdfe/hdfe_regressor.hpp:44-97)
 * that absorb_fixed_effects_cuda reads, plus the AbsorptionMethod argument. */
typedef struct {
    double tol;                 /* HdfeOptions::tol (default 1e-8) */
    int tolerance_mode;         /* ToleranceMode: 0 xhdfe-fast, 1 reghdfe-comparable, 2 strict-residual */
    int max_iter;               /* HdfeOptions::max_iter */
    int method;                 /* AbsorptionMethod (include/hdfe/hdfe_regressor.hpp:18): 0 Auto, 1 GaussSeidel,
                                   2 SymmetricGaussSeidel; 4 Schwarz and 7 AutoMlsmr sweep as GaussSeidel, as the
                                   reference's device absorber does for every method but Jacobi
                                   (src/fe_absorption_cuda.cu:3383); 3 Jacobi (never sent to the device by the
                                   reference's dispatcher, src/fe_absorption.cpp:9141), 5 Lsmr and 6 Mlsmr
                                   (CPU-only there, :9127) and any other value are rejected: the call returns 0 */
    int force_cg;               /* explicit reghdfe convergence criterion -> CG accelerator */
    int precision;              /* 0: FP64 throughout (the plain method); 1: FP32 sweeps under FP64 refinement;
                                   2: the Schur-complement Jacobi-PCG of precision 1 entirely in FP64 (its
                                   precision ablation; no counterpart in the reference) */
    int certify;                /* 1: compute the normal-equation certificate of the result on the device
                                   (certify_absorption_result, src/fe_absorption.cpp:14942-15190) */
    double refine_tol;          /* precision 1: FP64 refinement target, ||b - S beta|| / ||b|| of the
                                   Schur system over the non-primary levels; <= 0 selects
                                   min(3e-13, 3e-5 tol) */
    int preconditioner;         /* precision 1: 0 Jacobi; 1 aggregate blocks from the firm coupling
                                   strengths; 2 the learned preconditioner (aggregates from the graph
                                   network's scores, lp_params) -- the exact inverse of each
                                   aggregate's block of the Schur system on the firm levels, Jacobi on
                                   the others.  Convergence is always decided on the FP64 residual;
                                   a stall hands the solve to Jacobi-PCG (precond_fallback). */
    const float* lp_params;     /* preconditioner 2: the network's parameters (host, float32,
                                   xfe_lp_param_count() values) */
} xfe_options;

/* Outputs of one absorption (the fields of hdfe::detail::AbsorptionResult,
 * include/fe_absorption.hpp:24-47, that the GPU backend fills). */
typedef struct {
    int iterations;
    int converged;
    int used_accel;
    int used_cg;
    int polish_sweeps;
    double polish_final_max;
    int sweep_order_used[XFE_MAX_DIMS];
    int fe_levels[XFE_MAX_DIMS];
    int sweeps;                 /* projection sweeps executed (diagnostic) */
    int refinements;            /* FP64 refinement rounds (precision 1) */
    double solve_ms;            /* device wall time of the solve */
    /* certificate (options.certify): AbsorptionResult::abs_residual,
     * abs_residual_rel and precision_certified (include/fe_absorption.hpp:38-40) */
    double abs_residual;
    double abs_residual_rel;
    int precision_certified;
    int certificate_computed;
    /* which method ran: 0 the plain FP64 method, 1 the mixed-precision Schur PCG, 2 the FP64
       Schur PCG.  A request
       for precision 1 that the mixed path cannot take (weights, slopes, 1 or more than 3
       dimensions, more than 32 columns, no level-aligned tiles) runs the plain FP64 method and
       reports 0 here with precision_fallback = 1. */
    int precision_used;
    int precision_fallback;
    double refine_target;       /* precision 1: the FP64 refinement target that was applied */
    int preconditioner_used;    /* precision 1: the options.preconditioner that ran (0 when no firm pair
                                   exists or the path has no segment form) */
    int precond_fallback;       /* the block preconditioner stalled and Jacobi-PCG finished the solve */
    int precond_blocks;         /* firms inside aggregate blocks */
    double precond_ms;          /* set-up time of the preconditioner (pairs, network, aggregation, blocks) */
} xfe_absorb_result;

/* The certificate on its own (certify_cuda_absorption_result,
 * src/fe_absorption_cuda_certificate.cpp:318; certify_absorption_result,
 * src/fe_absorption.cpp:14942-15190, standard FEs): for every column c,
 * abs_c = ||D'W v~_c||_2 over all FE dimensions, ||D'W||_F^2 = sum_d sum_i w_i^2,
 * rel_c = abs_c / (||D'W||_F ||v_c||_2) with v the untransformed column;
 * certified = converged && max_c rel_c <= max(8 tol_eff, 64 eps), tol_eff =
 * min(tol, 1e-9) in reghdfe-comparable mode.  Host buffers: y, X (original),
 * y_t, X_t (transformed), all n rows, X column-major (ld n); weights may be
 * NULL.  Returns 1 on success, 0 on failure. */
typedef struct {
    double abs_residual;        /* max_c abs_c */
    double abs_residual_rel;    /* max_c rel_c */
    int precision_certified;
} xfe_certificate;

int xfe_certify(const double* y, const double* X, const double* y_t, const double* X_t, int n, int k,
                const xfe_fe_input* fes, int num_fes, const double* weights,
                const xfe_options* opt, int converged, xfe_certificate* out);

/* Shadow-mode device certificate (cuda_cert_shadow_begin / _finish / _abort and
 * CudaCertShadowResult, src/fe_absorption_cuda.hpp:55-103, impl
 * src/fe_absorption_cuda_certificate.cu): brackets Res_sq(c) = sum_d sum_g
 * (sum_{i in g} w_i r_ic)^2, Op_sq = sum_d sum_i w_i^2 and Orig_sq(c) =
 * sum_i v_ic^2 with rigorous a-priori rounding bounds and decides per column,
 * without sqrt or division: PASS iff R_upper <= L^2 A_lower B_lower, FAIL iff
 * R_lower > L^2 A_upper B_upper, else INDETERMINATE; ERROR for non-finite
 * input or any failure (fail-closed).  begin takes the original columns
 * (device, column-major, ld), finish the residual columns after the solve;
 * only scalars come back.  Standard dimensions only (a non-NULL slope value
 * pointer is rejected).  finish and abort free the handle. */
typedef struct {
    int verdict;              /* 0 PASS, 1 FAIL, 2 INDETERMINATE, 3 ERROR */
    int host_like_pass;       /* to-nearest rel_max <= L */
    double rel_max;           /* to-nearest sqrt(Res_sq / (Op_sq Orig_sq)), max over columns */
    double rel_max_perdim;    /* the same per dimension, max over dimensions and columns */
    double r_lower;           /* Res_sq bracket of the deciding column */
    double r_upper;
    double wsum_maxdiff;      /* max_g |recomputed - given weight sum| */
    double wsum_tol;          /* the gamma-scaled tolerance at the worst level */
    int wsum_ok;              /* 1 iff every level's difference is within its tolerance */
    long long empty_groups;   /* levels without rows, counted */
    double shadow_begin_ms;
    double shadow_finish_ms;
    int dims;
    int rhs;
} xfe_cert_shadow_result;

typedef struct xfe_cert_shadow xfe_cert_shadow;

xfe_cert_shadow* xfe_cert_shadow_begin(const double* d_y, const double* d_x, int n, int cols, int ld,
                                       const double* d_w, const int* const* d_group_ids,
                                       const double* const* d_weight_sums, const int* num_groups, int num_dims,
                                       const double* const* d_slope_values,
                                       const unsigned char* slope_has_intercept);
int xfe_cert_shadow_finish(xfe_cert_shadow* handle, const double* d_y_res, const double* d_x_res,
                           double certificate_limit, xfe_cert_shadow_result* out);
void xfe_cert_shadow_abort(xfe_cert_shadow* handle);

/* cuda_backend_available (src/fe_absorption_cuda.hpp:125,
 * impl src/fe_absorption_cuda.cu:2281).  1 if a gfx950 device is usable. */
int xfe_backend_available(void);

/* absorb_fixed_effects_cuda (src/fe_absorption_cuda.hpp:127-134,
 * impl src/fe_absorption_cuda.cu:2294-4738; slope dimensions as
 * slope_project_inplace, src/fe_absorption.cpp:4096-4300, by FP64 alternating
 * projections without the certificate).  Host buffers: y (n), X (n x k,
 * column-major, ld = n), optional weights (n).  sweep_order: num_fes entries
 * (0-based dims) or NULL for the default.  Writes y_out (n) and X_out (n x k).
 * Returns 1 on success (check res->converged, as the reference does), 0 on a
 * device/allocation/launch failure (the caller maps it to gpu status 4). */
int xfe_absorb(const double* y, const double* X, int n, int k,
               const xfe_fe_input* fes, int num_fes, const double* weights,
               const int* sweep_order, const xfe_options* opt,
               xfe_absorb_result* res, double* y_out, double* X_out);

/* group()/individual() structure, mirrors hdfe::detail::GroupIndividualStructure
 * (include/fe_absorption.hpp:13-21): num_groups = n rows (one observation per
 * group); group g's individuals are group_individual[group_ptr[g] ..
 * group_ptr[g+1]), individual j's groups individual_group[individual_ptr[j] ..
 * individual_ptr[j+1]) (the same incidences); group_scale (n) multiplies the
 * individual effects of a row (1 for sum, 1/size for mean aggregation). */
typedef struct {
    int num_groups;
    int num_individuals;
    const int* group_ptr;
    const int* group_individual;
    const int* individual_ptr;
    const int* individual_group;
    const double* group_scale;
} xfe_group_individual;

/* absorb_fixed_effects_group_individual_cuda (src/fe_absorption_cuda.hpp:136-145,
 * impl src/fe_absorption_cuda.cu:4740-5219): within transform of [y | X] over
 * the standard dimensions fes and the individual dimension, by Gauss-Seidel
 * sweeps (standard dimensions in sweep order, then the individuals colour
 * class by colour class; mirrored for method 2), the change-of-norm stop at
 * min(tol_eff, 1e-12), 16 polish sweeps, then the certificate with the
 * individual dimension (batches of 64 more sweeps until certified or
 * max_iter).  res->converged = res->precision_certified, as the reference sets
 * result.converged.  Host buffers as xfe_absorb.  Tolerance modes 0 and 1.
 * Returns 1 on success, 0 on invalid input or a device failure. */
int xfe_absorb_group_individual(const double* y, const double* X, int n, int k, const xfe_fe_input* fes,
                                int num_fes, const xfe_group_individual* gi, const double* weights,
                                const int* sweep_order, const xfe_options* opt, xfe_absorb_result* res,
                                double* y_out, double* X_out);

/* Same computation with device-resident buffers (inputs already in HBM):
 * d_V holds [y | X] column-major with leading dimension ld >= n and is
 * overwritten with the within transform; d_group_ids is num_fes x n
 * (dimension-major, ld n) compact ids; d_weight_sums may be NULL (computed on
 * the device).  No counterpart in the reference (its backend always stages
 * host memory); it serves callers that keep their panel on the device.
 * Ordering: the call first waits for all work already queued on the device
 * (any stream), so buffers filled asynchronously by the caller are complete,
 * and it returns after its own work has finished.  When the ids of the
 * dimension with the most levels are already sorted, ld is n rounded up to a
 * multiple of 64 and d_V is 256-byte aligned, the solve works in d_V itself
 * (no working copy): rows n..ld-1 may then be used as scratch, and after a
 * device failure (return 0) d_V's content is unspecified. */
int xfe_absorb_device(double* d_V, int n, int C, int ld,
                      const int* d_group_ids, const int* num_groups, int num_fes,
                      const double* d_weights, const int* sweep_order,
                      const xfe_options* opt, xfe_absorb_result* res);

/* xfe_absorb_device with heterogeneous slope dimensions (the slope terms of
 * absorb_fixed_effects_v6_mixed, src/fe_absorption.cpp:8162, whose device call
 * at :8336 stages host memory; no device-resident counterpart in the
 * reference).  slope_mode[d]: 0 standard, 1 slope with intercept, 2 slope
 * only; d_slope_values[d] (device, n) for slope dimensions.  With
 * opt->precision 1 an unweighted problem with at most 32 columns and a
 * standard dimension runs the mixed-precision Schur PCG (res->precision_used
 * 1), otherwise the FP64 slope path (precision_fallback 1 when precision 1
 * was asked); columns the mixed run leaves unconverged after 2000 FP32
 * iterations continue with up to 10000 FP64 block-PCG iterations from their
 * iterate, and only if that fails are the (untouched) columns absorbed by the
 * FP64 slope path (precision_fallback 1).  The certificate (opt->certify) includes the slope
 * moments.
 * Ordering as xfe_absorb_device.  Returns 1 on success, 0 on failure. */
int xfe_absorb_device_slopes(double* d_V, int n, int C, int ld, const int* d_group_ids, const int* num_groups,
                             int num_fes, const double* const* d_slope_values, const int* slope_mode,
                             const double* d_weights, const int* sweep_order, const xfe_options* opt,
                             xfe_absorb_result* res);

/* xhdfe::schwarz_demean_raw (include/schwarz_demean.hpp:30, impl
 * src/schwarz_demean.cpp:419-438; called by absorb_fixed_effects,
 * src/fe_absorption.cpp:5185-5215): unweighted within transform of Y (n) and
 * X (n x k, column-major, ld n) over ndim >= 2 dimensions, in place.  fe[d]
 * holds n ids in [0, nlev[d]).  Converged means, for every column c with
 * ||D'v_c|| > 1e-300, ||D'(v_c - D u_c)|| / ||D'v_c|| < tol over all
 * dimensions (the reference's stopping rule, evaluated on the result).
 * Returns 1 when converged (Y, X overwritten with the residuals), 0 otherwise
 * (ndim < 2, n <= 0, k < 0, k + 1 > 16, invalid ids, tolerance not met or a
 * device failure; Y and X are left unchanged).  iters_out (optional)
 * receives the inner Schur PCG iterations.  threads and collapse are
 * accepted for the signature and do not change the result. */
int xfe_schwarz_demean(double* Y, double* X, int64_t n, int k, const int32_t* const* fe,
                       const int* nlev, int ndim, int threads, double tol, int max_iter,
                       int* iters_out, int collapse);

/* build_indexer (src/fe_absorption.cpp:845-940) with finalize_contiguous_groups
 * (:547-590), on the device: raw int32 ids (n) become compact 0-based ids in
 * ascending raw-id order.  When max - min + 1 <= min(2n, 50M) the id is
 * raw - min (unused slots kept, num_levels_present < num_groups) unless fewer
 * than 90 % of the slots are used, then the used slots are ranked; otherwise
 * the distinct ids are ranked.  contiguous = 1 iff the compact ids are
 * non-decreasing over the rows (the reference's groups_contiguous).  Host
 * buffers; returns 1 on success, 0 on failure. */
int xfe_build_indexer(const int* raw, int n, int* group_ids, int* num_groups, int* num_levels_present,
                      int* contiguous);

/* The same on device buffers (d_raw, d_group_ids in HBM; the counts come back
 * to host pointers).  No counterpart in the reference, whose indexer runs on
 * the host; it serves callers that keep their panel on the device.  Waits for
 * all work already queued on the device before reading d_raw. */
int xfe_build_indexer_device(const int* d_raw, int n, int* d_group_ids, int* num_groups,
                             int* num_levels_present, int* contiguous);

/* Allocation-free form of xfe_build_indexer_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same kernels take their scratch (min/max, flags, rank tables and the
 * radix-sort and scan storage) from a 40 MiB file-scope device array that each
 * device holds until the process ends; the call makes no hipMalloc/hipFree and
 * creates no stream (it runs on the null stream).  n may be at most 1,048,576
 * rows (2^20); larger inputs are rejected with an error.  Every call on one
 * device shares that device's array: calls that overlap on one device race
 * (serialize them); each call clears what it needs.  Same outputs, ordering
 * and error convention as xfe_build_indexer_device. */
int xfe_build_indexer_device_noalloc(const int* d_raw, int n, int* d_group_ids, int* num_groups,
                                     int* num_levels_present, int* contiguous);

/* compute_weight_sums (src/fe_absorption.cpp:942-1037): out[g] = sum of
 * weights (or the count, weights NULL) over the rows with group_ids[i] = g,
 * g < num_groups.  Ids outside [0, num_groups) are rejected.  Host buffers;
 * returns 1 on success, 0 on failure. */
int xfe_weight_sums(const int* group_ids, int n, int num_groups, const double* weights, double* out);

/* compute_keep_mask_drop_singletons (src/hdfe_regressor_v11.cpp:4056-4302):
 * iterative singleton drop on the device, bit-exact.  fes: num_fes x n raw
 * int32 ids (host).  keep (n bytes) out.  Returns dropped rows, -1 on error. */
int xfe_drop_singletons(const int* fes, int n, int num_fes, int max_rounds,
                        unsigned char* keep, int* rounds_used);

/* count_bipartite_components (src/hdfe_regressor_v11.cpp:3579-3685) with the
 * node numbering of build_lookup_compact (:1354-1417): mobility groups of two
 * FE dimensions.  a, b: raw ids (host, n).  expected_a/expected_b: level
 * counts as the reference's caller passes them (-1 if unknown).  labels (n,
 * optional) receives 1-based component ids ranked by the smallest node.
 * Returns the component count, -1 on error. */
int xfe_bipartite_components(const int* a, const int* b, int n, int expected_a,
                             int expected_b, int* labels);

/* compute_fe_dof_reghdfe (src/hdfe_regressor_v11.cpp:3743-3986), pairwise
 * ("all") method: redundant levels per dimension from mobility groups, with the
 * nested-FE coarsening.  fes: num_fes x n raw ids; levels: levels per dim.
 * Writes redundant / num_coefs (num_fes each).  Returns df_a, -1 on error. */
int xfe_fe_dof(const int* fes, int n, int num_fes, const int* levels,
               int* redundant, int* num_coefs);

/* X~'X~ Gram and cluster meat for the OLS step after absorption
 * (src/ols.cpp:1338 Gram, :889-1046 one-way cluster meat).  V: n x C
 * column-major [y~ | X~] (host).  Outputs: gram (C x C, column-major, over all
 * columns incl. y), and when cluster != NULL the k x k meat
 * sum_g (X_g' e_g)(X_g' e_g)' for residuals e = y~ - X~ beta (beta: k).
 * Returns 0 on success. */
int xfe_gram(const double* V, int n, int C, double* gram);
int xfe_cluster_meat(const double* V, int n, int C, const double* beta,
                     const int* cluster, int num_clusters, double* meat);

/* fe_recovery_max_delta_cuda (src/fe_absorption_cuda.hpp:152-156, impl
 * src/fe_absorption_cuda.cu:5227): one Gauss-Seidel pass of `residual` over the
 * dims in sweep_order, returns max_d max_g |weighted mean_g| in *out_max_delta.
 * residual (host, n) is not modified.  Returns 1 on success, 0 on failure. */
int xfe_recovery_max_delta(const double* residual, int n, const xfe_fe_input* fes,
                           int num_fes, const double* weights, const int* sweep_order,
                           double* out_max_delta);

/* The GPU branch of recover_fixed_effects_group_ids (src/fe_absorption.cpp:15470-15555,
 * called for savefe by fit, src/hdfe_regressor_v11.cpp:9335-9345): absorbs
 * `partial` with the fixed effects retained (absorb_fixed_effects_cuda with
 * retain_fixed_effects, src/fe_absorption_cuda.cu:4663-4690) and writes each
 * dimension's level coefficients alpha_out[d] (num_groups doubles, host) and,
 * when contributions_out is not NULL, contributions_out[d*n + i] =
 * alpha_out[d][ids_d[i]] (FeRecoveryResult::contributions).  max_delta is
 * fe_recovery_max_delta of partial - sum_d contributions_d (one Gauss-Seidel
 * pass, as xfe_recovery_max_delta); converged = absorber converged &&
 * max_delta <= fe_tol (fe_tol <= 0 selects HdfeOptions::fe_tolerance's 1e-6).
 * The coefficients are one solution of D alpha = partial - M_D partial; like
 * the reference's, they are identified up to a constant per dimension and
 * mobility group.  1 to 3 dimensions; weights only with one dimension.
 * Returns 1 on success, 0 on failure (then see xfe_last_error). */
typedef struct {
    int iterations;        /* FeRecoveryResult::iterations (inner iterations of the solve) */
    int converged;         /* FeRecoveryResult::converged */
    double max_delta;      /* FeRecoveryResult::max_delta */
    int absorb_converged;  /* the absorber's own flag */
} xfe_recovery_result;

int xfe_recover_fixed_effects(const double* partial, int n, const xfe_fe_input* fes, int num_fes,
                              const double* weights, const int* sweep_order, const xfe_options* opt,
                              double fe_tol, double* const* alpha_out, double* contributions_out,
                              xfe_recovery_result* out);

/* articulation_points (src/akm_kss.cpp:444-493; called by build_leave_out_sample's
 * pruning loop, :736-870): artic[v] = 1 iff removing vertex v disconnects its
 * connected component, for the undirected graph with neighbours
 * adj[adj_ptr[v] .. adj_ptr[v+1]) of vertex v (each edge listed from both ends,
 * no self loops or repeated edges).  Computed on the device by the
 * Tarjan-Vishkin method on a breadth-first spanning forest; the graph is
 * validated on the host.  Returns 1 on success, 0 on invalid input or a device
 * failure. */
int xfe_articulation_points(int n_vertices, const int* adj_ptr, const int* adj, unsigned char* artic);

/* leave_out_connected_set (include/hdfe/akm_kss.hpp:55-90, impl
 * src/akm_kss.cpp:2233-2290 over build_leave_out_sample :494-900): the rows
 * of the largest worker-firm connected set (most firms, then most person-year
 * rows, then the smallest original firm id) that stays connected when any
 * one worker is left out: movers that are cut vertices of the mover-firm
 * graph are dropped, with the largest set re-taken, until none is; then
 * workers observed once leave.  keep (n) marks the kept input rows; fweights
 * (optional, n) are frequency weights.  The match-key sort
 * (xfe_akm_stable_sort_u64) and the cut-vertex search
 * (xfe_articulation_points) run on the device, the bookkeeping on the host.
 * Returns 1 on success, 0 on invalid input or a device failure. */
typedef struct {
    long long n_obs_input;
    long long n_obs_connected;  /* rows of the largest connected set before pruning */
    long long n_obs;            /* rows of the leave-out set */
    int n_workers;
    int n_firms;
    long long n_matches;
    int n_movers;
    int n_stayers;
    int prune_iterations;
    int gpu_used;               /* the match-key sort ran on the device */
} xfe_leave_out_result;

int xfe_akm_leave_out_set(const int* worker_ids, const int* firm_ids, int n, const long long* fweights,
                          unsigned char* keep, xfe_leave_out_result* out);

/* ---- AKM/KSS leverage solver (include/hdfe/akm_kss_cuda.hpp:19-74, impl
 * src/akm_kss_cuda.cu:164-722): the grounded worker-partialled firm Laplacian
 * S = Df - B' Dw^{-1} B over n_firms levels (ground: the last firm), B the
 * worker x firm match counts.  Same arguments and error convention as the
 * reference's akm_cuda_* functions. */
typedef struct xfe_akm_context xfe_akm_context;

/* akm_cuda_available (:21). */
int xfe_akm_available(void);

/* akm_cuda_stable_sort_u64 (:27-29): order_out receives the 0-based row
 * indices of the host keys in ascending key order, equal keys in input order.
 * Returns 1 on success, 0 on failure. */
int xfe_akm_stable_sort_u64(const uint64_t* keys, int64_t* order_out, size_t n);

/* akm_cuda_create (:33-40): uploads the match structure (matches grouped by
 * worker as w_ptr says; f_ptr / f_matches list each firm's matches).  Match
 * counts m_c must be positive integers.  Returns NULL on failure (the caller
 * falls back to its CPU solver). */
xfe_akm_context* xfe_akm_create(int n_workers, int n_firms, const int* m_w, const int* m_f,
                                const double* m_c, size_t n_matches, const int64_t* w_ptr,
                                const int64_t* f_ptr, const int64_t* f_matches,
                                const double* Dw, const double* Df, const double* inv_diagS);

/* akm_cuda_destroy (:42). */
void xfe_akm_destroy(xfe_akm_context* ctx);

/* akm_cuda_solve_S (:47-48): Jacobi-PCG of S z = rhs (length n_firms, ground
 * entry pinned to zero), stop at |r|^2 <= tol^2 |rhs|^2.  Returns the
 * iteration count, or -1 when not converged within max_iter. */
int xfe_akm_solve_S(xfe_akm_context* ctx, const double* rhs, double* z, double tol, int max_iter);

/* akm_cuda_solve_S_multi (:59-61): nb <= xfe_akm_max_lanes() independent
 * lanes at once; rhs_pack / z_pack hold nb contiguous length-n_firms columns
 * (may alias).  iters_out[l]: iterations, or -1 (that lane's z is zero).
 * Returns 0, or -1 on a device error. */
int xfe_akm_solve_S_multi(xfe_akm_context* ctx, const double* rhs_pack, double* z_pack, int nb,
                          double tol, int max_iter, int* iters_out);

/* JLA leverages of the working rows (the P_ii part of the JLA draws,
 * src/akm_kss.cpp:3153-3285, match level), on the device: the rows are the
 * context's matches in the order given to xfe_akm_create, row m standing for
 * sqrt(m_c[m]) (e_worker + e_firm).  Per draw d, signs r from the reference's
 * SplitMix64 streams (seed, 3 d, block of 64 rows), the two-way normal
 * equations solved for A'r (firm part by the grounded S solver, batches of
 * xfe_akm_max_lanes() draws, stop at tol), Z = A [zw; zf]; then
 * pii = mean Z^2 / (mean Z^2 + mean (r - Z)^2) (host array of M).  Returns 0, 1 when some
 * draw's solve did not converge (its z is zero, as in the reference), -1 on
 * failure. */
int xfe_akm_jla_pii(xfe_akm_context* ctx, int draws, uint64_t seed, double tol, int max_iter,
                    double* pii);

/* xfe_akm_jla_pii with the reference's non-linearity correction of the JLA
 * ratio estimator per working row (src/akm_kss.cpp:3194-3206 and :3273-3288,
 * the corr factor of its sigma_i): with the draw means P, M of Z^2 and
 * (r - Z)^2 and Psq, Msq, PM of Z^4, (r - Z)^4 and Z^2 (r - Z)^2, Ph = pii,
 * Mh = 1 - Ph, V = (Mh^2 Psq + Ph^2 Msq - 2 Mh Ph PM) / p,
 * B = (Mh Psq - Ph Msq + 2 (Mh - Ph) PM) / p, corr = 1 - V / Mh^2 + B / Mh
 * (1 where Mh <= 0).  pii and corr: host arrays of M.  Returns as
 * xfe_akm_jla_pii. */
int xfe_akm_jla_leverages(xfe_akm_context* ctx, int draws, uint64_t seed, double tol, int max_iter,
                          double* pii, double* corr);

/* JLA estimates of the KSS quadratic-form weights B_ii of the variance
 * components (src/akm_kss.cpp:2985-3047 and :3128-3146, match level), on the
 * device, rows as in xfe_akm_jla_pii.  Per draw d, u_m = sum of py_count[m]
 * Rademacher draws of the reference's stream (seed, 3 d + 1, m), scaled by
 * 1 / sqrt(draws) and demeaned over person years (u_m - vbar py_count[m],
 * vbar = sum u / n_py); the fe solve of [0; F'u] and the pe solve of
 * [W'u; 0] (draws in batches of xfe_akm_max_lanes() / 2, stop at tol) give
 * W2 = A z_fe and W1 = A z_pe; bfe += W2^2, bpe += W1^2, bcov += W1 W2 (host
 * arrays of M).  py_count: person-year rows per match; n_py <= 0 selects
 * sum(py_count) (the reference's n_py is the frequency-weight total).
 * Returns 0, 1 when some draw's solve did not converge, -1 on failure. */
int xfe_akm_jla_bii(xfe_akm_context* ctx, int draws, uint64_t seed, double tol, int max_iter,
                    const int64_t* py_count, double n_py, double* bfe, double* bpe, double* bcov);

/* akm_cuda_max_lanes (:65). */
int xfe_akm_max_lanes(void);

/* Wide batched solve: no counterpart in the reference, whose caller
 * (src/akm_kss.cpp:1659-1707) loops over akm_cuda_solve_S_multi in batches
 * of akm_cuda_max_lanes() = 32.  Up to xfe_akm_wide_max_lanes() (256) lanes
 * per operator pass, FP32 PCG (aggregate blocks of strongly coupled firms
 * solved exactly, Jacobi elsewhere) whose residual is replaced by the FP64
 * residual b - S x (x accumulated in FP64) whenever r.r has dropped 1e4-fold;
 * a lane stops when its FP64 residual meets |r|^2 <= tol^2 |rhs|^2.  Same
 * pack layout, ground pin and per-lane results as xfe_akm_solve_S_multi
 * (iters_out[l]: FP32 iterations, or -1 with that lane's z zero).  Only the
 * movers' matches enter the passes (a single-match worker with c = Dw adds
 * nothing to S beyond its diagonal).  Returns 0, or -1 on a device error. */
int xfe_akm_solve_S_wide(xfe_akm_context* ctx, const double* rhs_pack, double* z_pack, int nb,
                         double tol, int max_iter, int* iters_out);

/* xfe_akm_solve_S_wide on device arrays (nb x n_firms lane-major, FP64;
 * may alias): waits for the work already queued on the device, returns after
 * its own work has ended.  No counterpart in the reference (its solves take
 * host packs). */
int xfe_akm_solve_S_wide_device(xfe_akm_context* ctx, const double* d_rhs_pack, double* d_z_pack, int nb,
                                double tol, int max_iter, int* iters_out);

/* 256: the lane limit of xfe_akm_solve_S_wide. */
int xfe_akm_wide_max_lanes(void);

/* FP32 operator passes and FP64 residual replacements of the context's last
 * xfe_akm_solve_S_wide call.  Returns 0, or -1 on invalid arguments. */
int xfe_akm_wide_last(xfe_akm_context* ctx, int* fp32_passes, int* fp64_replacements);

/* Exact leverages and quadratic-form weights, the reference's exact path of
 * akm_kss_decompose (LeverageMethod::Exact, src/akm_kss.cpp:2667-2741, match
 * level): per match m one solve of K [zw; zf] = [e_w; e_f] (here through the
 * firm Schur complement, 256 matches per wide solve), then
 *   pii[m]  = c_m (zw[w] + zf[f]),
 *   bfe[m]  = c_m (sum_f Df zf^2 - (sum_f Df zf)^2 / n_py),
 *   bpe[m]  = c_m (sum_w Dw zw^2 - (sum_w Dw zw)^2 / n_py),
 *   bcov[m] = c_m (sum_k c_k zw[w_k] zf[f_k] - (sum_w Dw zw)(sum_f Df zf) / n_py),
 * n_py = sum_m c_m, in the context's match order (outputs of M doubles each).
 * tol / max_iter: the wide solver's (FP32 passes under FP64 residual
 * replacement, stop on the FP64 residual |r| <= tol |rhs|); accept (>= tol, or
 * 0 for the strict rule): a solve whose FP64 residual stops improving counts
 * as converged at |r| <= accept |rhs|.  Matches whose wide solve does not
 * converge are solved again by the FP64 batched Jacobi-PCG.  Returns 0, the
 * number of matches whose solve did not converge there either (their values
 * are those of a zero solution), or -1 on failure. */
int xfe_akm_exact_leverages(xfe_akm_context* ctx, double tol, double accept, int max_iter, double* pii, double* bfe,
                            double* bpe, double* bcov);

/* Matches of this thread's last xfe_akm_exact_leverages call whose wide solve
 * did not converge and that were solved again by the FP64 batched Jacobi-PCG
 * (32 lanes per solve).  No counterpart in the reference. */
int xfe_akm_exact_leverages_last_retries(void);

/* ---- Cluster bootstrap refits of the two-way FE regression on the wide
 * solver's launch path.  No counterpart in the reference (its bootstrap is the
 * Gelbach decomposition's, src/akm_kss.cpp:4965-7305, outside this path); the
 * refits reuse the absorber's weighted semantics (rows weighted by their
 * worker's multiplicity, src/fe_absorption.cpp:942-1037). */
typedef struct xfe_boot_context xfe_boot_context;

/* Panel rows (worker, firm: compact 0-based ids; the firm graph should be
 * connected, the last firm grounds the firm effects) and the columns
 * V = [y | X] (n x C column-major with leading dimension ldv, C <= 4).
 * Returns NULL on failure. */
xfe_boot_context* xfe_boot_create(int n, int n_workers, int n_firms, const int* worker, const int* firm,
                                  const double* V, int ldv, int C);

void xfe_boot_destroy(xfe_boot_context* ctx);

/* R refits: mult (R x n_workers, resample-major) holds each worker's
 * multiplicity in each resample (its rows enter that many times; 0 leaves it
 * out; resample sizes may differ).  gram_out (R x C x C) receives the Gram
 * X~' W X~ of the within-transformed columns of each refit (W: the
 * multiplicities), from which beta = G[1:,1:]^-1 G[1:,0]; iters_out[r]: the
 * FP32 iterations of refit r's firm-effect solves, -1 when one did not
 * converge to tol.  64 refits share each operator pass.  Returns 0, 1 when
 * some solve did not converge, -1 on failure. */
int xfe_boot_refits(xfe_boot_context* ctx, int R, const int* mult, double tol, int max_iter, double* gram_out,
                    int* iters_out);

/* FP32 operator passes and FP64 residual replacements summed over the
 * batches of the context's last xfe_boot_refits call. */
int xfe_boot_last(xfe_boot_context* ctx, int* fp32_passes, int* fp64_replacements);

/* akm_cuda_pack_buffer (:74): pinned host buffer of nb x n_firms doubles owned
 * by the context, or NULL (then the caller uses its own buffer). */
double* xfe_akm_pack_buffer(xfe_akm_context* ctx, int nb);

/* ---- Learned preconditioner (no counterpart in the reference, whose
 * preconditioner for the FE normal equations is the randomized approximate
 * Cholesky of src/schwarz_demean.cpp:100-133).  On the firm levels the Schur
 * system of the mixed-precision absorber is a weighted graph Laplacian whose
 * slowest modes (under Jacobi scaling) live on small groups of firms coupled
 * to each other far more than to the rest.  A graph network over the
 * worker-firm mobility graph (firms; movers = workers with segments at two or
 * more firms; one edge per (mover, firm)), width 64, three message-passing
 * layers, BF16 MFMA with FP32 accumulation, emits a per-level score u_f and a
 * score t_fg per coupled firm pair; the aggregates are formed by a parallel
 * handshake over t_fg + (u_f + u_g) / 2 (at most 16 firms each), and the
 * preconditioner solves each aggregate's block exactly (xfe_options
 * .preconditioner = 2).  The parameter vector is float32, xfe_lp_param_count()
 * long (the package ships trained values). */
int xfe_lp_param_count(void);

/* One evaluation of the network, its loss and gradient (host buffers): firm
 * features xf (F x 8), mover features xw (Wm x 8), firm -> mover edges
 * (f_ptr: F + 1, f_idx: mover ids), mover -> firm edges (w_ptr: Wm + 1, w_idx:
 * firm ids; the same edge set), firm pairs (pf, pg) with one feature each (pq,
 * may be NULL: 0), targets u_tgt / u_mask (F) and y_tgt (npairs), parameters
 * theta.  loss = mean over masked firms of (u - u_tgt)^2 + mean binary
 * cross-entropy of sigmoid(t) against y_tgt; grad (kNumParams), u_out (F),
 * t_out (npairs) optional.  Returns 1 on success. */
int xfe_lp_gnn_eval(int F, int Wm, const float* xf, const float* xw, const int* f_ptr, const int* f_idx,
                    const int* w_ptr, const int* w_idx, int npairs, const int* pf, const int* pg, const float* pq,
                    const float* u_tgt, const float* u_mask, const float* y_tgt, const float* theta,
                    double* loss, float* grad, float* u_out, float* t_out);

/* Training data from the absorber itself: xfe_lp_probe runs the
 * mixed-precision absorber (Jacobi-PCG, tol 1e-8) on [y | X] (host, n rows, X
 * column-major) with 2 or 3 unweighted dimensions (the first is the
 * eliminated one), records the FP64 iterate at inner iteration probe_iter, and
 * keeps on the host: per firm and column the error of that iterate against the
 * final solution, the firm pairs (pf < pg, coupling W, strength pq), the exact
 * firm diagonal, and the network's graph and features as the preconditioner
 * builds them.  xfe_lp_probe_fetch copies them into caller buffers of the
 * sizes reported (xf F x 8, xw Wm x 8, f_ptr F + 1, f_idx E, w_ptr Wm + 1,
 * w_idx E, pf / pg / pq / W U, dex F, err F x C firm-major); NULL skips one. */
typedef struct {
    int F, Wm, E, U, C;
    int inner_iterations;
    int converged;
} xfe_lp_panel_sizes;

int xfe_lp_probe(const double* y, const double* X, int n, int k, const xfe_fe_input* fes, int num_fes,
                 int probe_iter, xfe_lp_panel_sizes* sizes);
int xfe_lp_probe_fetch(float* xf, float* xw, int* f_ptr, int* f_idx, int* w_ptr, int* w_idx, int* pf, int* pg,
                       float* pq, double* W, double* dex, double* err);

/* The aggregation and block solves of the preconditioner on given pairs (a
 * unit of the preconditioner, for tests and for callers with their own
 * scores): firm pairs pf < pg (sorted, unique) with couplings W, the firm
 * diagonal dex, pair scores; the handshake aggregation (cap firms per
 * aggregate, `rounds` rounds) into agg_out (the root firm of each firm), and
 * z_out = inverse of each aggregate's block times r on the firms of
 * aggregates of two or more firms, 0 elsewhere (r, z: F x C level-major,
 * float32).  Returns 1 on success. */
int xfe_lp_blocks_apply(int F, int U, const int* pf, const int* pg, const double* W, const double* dex,
                        const float* score, int cap, int rounds, int C, const float* r, int* agg_out,
                        float* z_out);

/* Training on the device: Adam (lr, beta 0.9 / 0.999) over `steps` steps, one
 * panel per step in turn, each step one forward and backward pass of the
 * network (the loss of xfe_lp_gnn_eval).  theta: parameters in/out (host);
 * loss_hist: the loss of every step (may be NULL).  Returns 1 on success. */
typedef struct {
    int F, Wm, U;
    const float* xf;
    const float* xw;
    const int* f_ptr;
    const int* f_idx;
    const int* w_ptr;
    const int* w_idx;
    const int* pf;
    const int* pg;
    const float* pq;
    const float* u_tgt;
    const float* u_mask;
    const float* y_tgt;
} xfe_lp_graph;

int xfe_lp_train(int npanels, const xfe_lp_graph* panels, int steps, float lr, float* theta, double* loss_hist);

/* Frees the calling thread's device workspace (it otherwise lives until the
 * process ends).  No counterpart in the reference. */
void xfe_release_workspace(void);

/* Last error message of the calling thread ("" if none). */
const char* xfe_last_error(void);

/* Kernel timing for the roofline figure: when enabled, the library brackets
 * every launch of the fused sweep pass with HIP events on its own stream and
 * accumulates the total milliseconds, launch count and algorithmic bytes. */
void xfe_profile_enable(int on);
void xfe_profile_reset(void);
void xfe_profile_get(double* pass_ms, long long* pass_launches, double* pass_bytes);

#ifdef __cplusplus
}
#endif
#endif /* XFE_HIP_H */
