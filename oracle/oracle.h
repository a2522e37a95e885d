/* CPU oracle for the fixed-effects absorber hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference CPU algorithms (xhdfe-xfe).  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker; the product never links it.
 *
 * Parity status per function (what pins each one) is listed in oracle.c's
 * header comment.
 */
#ifndef XFE_ORACLE_H
#define XFE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* build_indexer (src/fe_absorption.cpp:845-940): raw ids -> compact ids in
 * ascending raw-id order.  Returns 0 on success. */
int or_build_indexer(int n, const int* raw, int* gid, int* num_groups, int* levels_present,
                     int* contiguous);

/* compute_weight_sums (src/fe_absorption.cpp:942-1037). w may be NULL. */
void or_weight_sums(int n, const int* gid, int G, const double* w, double* out);

/* Absorber options (subset of HdfeOptions, include/hdfe/hdfe_regressor.hpp:44-97). */
typedef struct {
    double tol;            /* HdfeOptions::tol */
    int tolerance_mode;    /* 0 xhdfe-fast, 1 reghdfe-comparable, 2 strict-residual */
    int max_iter;          /* HdfeOptions::max_iter */
    int symmetric;         /* symmetric sweep (SGS) */
    int force_cg;          /* explicit reghdfe criterion -> CG accelerator */
    int threads;           /* OpenMP threads for the projections (>=1) */
    const int* sweep_order;/* optional explicit order (D entries) or NULL */
} or_options;

typedef struct {
    int iterations;
    int converged;
    int used_accel;
    int used_cg;
    int polish_sweeps;
    double polish_final_max;
    int sweep_order[8];
} or_result;

/* absorb_fixed_effects (src/fe_absorption.cpp:5993-8100), packed path.
 * V: n x C column-major [y | X], overwritten with the within transform.
 * gids: D x n compact ids (dimension-major); G[d] level counts; w: weights or NULL. */
int or_absorb(int n, int C, double* V, int D, const int* gids, const int* G,
              const double* w, const or_options* opt, or_result* res);

/* One FE projection v <- v - P_d v over C columns (demean_inplace_contiguous,
 * src/fe_absorption.cpp:2380). */
void or_project(int n, int C, double* V, const int* gid, int G, const double* w);

/* compute_keep_mask_drop_singletons (src/hdfe_regressor_v11.cpp:4056-4302).
 * fes: D x n raw ids.  keep[n] out.  Returns the number dropped. */
int or_drop_singletons(int n, int D, const int* fes, int max_rounds, unsigned char* keep,
                       int* rounds_used);

/* count_bipartite_components (src/hdfe_regressor_v11.cpp:3579-3685) with the
 * node numbering of build_lookup_compact (:1354-1417).  labels[n] (1-based,
 * may be NULL).  expected_a/expected_b: the caller's level counts (-1 = none). */
int or_components(int n, const int* a, const int* b, int expected_a, int expected_b,
                  int* labels);

/* compute_fe_dof_reghdfe (src/hdfe_regressor_v11.cpp:3743-3986), method "all". */
int or_dof_reghdfe(int n, int D, const int* fes, const int* levels, int* redundant,
                   int* num_coefs);

/* fe_is_function_of_refined_fe (src/hdfe_regressor_v11.cpp:3686-3741). */
int or_fe_is_function_of(int n, const int* coarse, const int* refined);

#ifdef __cplusplus
}
#endif
#endif
