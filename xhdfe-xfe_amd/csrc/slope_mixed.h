// Mixed-precision absorption with heterogeneous slopes (slope_mixed.hip): the
// standard dimension with the most levels eliminated exactly, FP32 PCG with
// FP64 residual replacement on the Schur system of the remaining levels and
// slopes, the final transform in FP64.
#ifndef XFE_SLOPE_MIXED_H
#define XFE_SLOPE_MIXED_H

#include <hip/hip_runtime.h>

#include <vector>

namespace xfe {

constexpr int kSmMaxR = 6;  // remaining (level or slope) blocks of unknowns

struct SlopeMixed;

struct SlopeMixedResult {
    int iterations = 0;          // FP32 PCG iterations (max over columns)
    int passes = 0;              // FP32 operator passes
    int replacements = 0;        // FP64 residual replacements
    int fp64_iterations = 0;     // FP64 block-PCG iterations after the FP32 runs
    int unknowns = 0;            // levels and slopes of the Schur system
    int primary = -1;            // the eliminated dimension
    int intercepts_dropped = 0;  // slope intercepts spanned by a finer dimension
    int pinned = 0;              // levels pinned to zero (one per connected component and standard block)
    bool converged = false;
    // per column: the sum over every dimension's levels (intercept and slope
    // moments) of the squared level sums of the transform, except the dropped
    // intercepts' (added by the caller)
    std::vector<double> resid_sq;
};

SlopeMixed* slope_mixed_new();
void slope_mixed_delete(SlopeMixed* sm);

// mode per dimension: 0 standard, 1 slope with intercept, 2 slope only.  The
// path takes unweighted problems with at most 32 columns and at least one
// standard dimension.
bool slope_mixed_eligible(int C, int D, const int* mode, const double* w);

// V: n x C column-major (ld), device, the caller's row order on entry and on
// exit (transformed in place when every column converged, untouched
// otherwise); gid / z: device arrays of n.  tol: the target of the Schur
// residual |b - S x| / |b| per column; a column whose FP64 residual stalls
// above it is accepted at 1e-9 (the absorber's mixed rule).  max_iter bounds
// the FP32 iterations; the columns they leave unconverged continue with up to
// fp64_max_iter FP64 block-PCG iterations from their iterate.
void slope_mixed_run(SlopeMixed& sm, double* V, int n, int C, int ld, const std::vector<const int*>& gid,
                     const std::vector<int>& G, const std::vector<const double*>& z, const std::vector<int>& mode,
                     double tol, int max_iter, int fp64_max_iter, hipStream_t s, SlopeMixedResult& out);

// whether the last run left the intercept of dimension d out of the solve
bool slope_mixed_intercept_dropped(const SlopeMixed& sm, int d);

// the dropped intercepts' contribution to resid_sq: squared level sums of
// the transform V (rows in the caller's order) over dimension gid's levels
void slope_mixed_level_sumsq(SlopeMixed& sm, const double* V, int n, int C, int ld, const int* gid, int G,
                             hipStream_t s, std::vector<double>& acc);

}  // namespace xfe

#endif
