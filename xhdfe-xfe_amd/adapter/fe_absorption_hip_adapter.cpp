// Drop-in for the reference's CUDA absorber translation units
// (src/fe_absorption_cuda.cu and src/fe_absorption_cuda_certificate.{cpp,cu}):
// the hdfe::detail functions declared in src/fe_absorption_cuda.hpp, defined
// over the C ABI of include/xfe_hip.h.  A maintainer compiles this file in
// place of those units, with -DHDFE_USE_CUDA so the callers' #ifdef
// HDFE_USE_CUDA blocks stay active, and links libxfe_hip.so; every caller
// (src/fe_absorption.cpp:8336, :9203, :9871, :15484; src/hdfe_regressor_v11.cpp)
// is unchanged.  No HIP header is needed on this side.
//
// Status semantics stay with the reference's dispatcher
// (src/fe_absorption.cpp:9209-9237): false -> gpu status 4, true with
// converged = false -> status 3, true and converged -> status 1 and the host
// certificate.  No exception crosses the boundary: the library reports
// failures by return codes and xfe_last_error().
#include <cstddef>
#include <stdexcept>
#include <vector>

#include "fe_absorption_cuda.hpp"
#include "xfe_hip.h"

namespace hdfe {
namespace detail {

namespace {

std::vector<xfe_fe_input> to_inputs(const std::vector<GpuFeInput>& in) {
    std::vector<xfe_fe_input> out(in.size());
    for (std::size_t d = 0; d < in.size(); ++d) {
        const GpuFeInput& f = in[d];
        out[d].group_ids = f.group_ids;
        out[d].num_groups = f.num_groups;
        out[d].num_levels_present = f.num_levels_present;
        out[d].weight_sums = f.weight_sums;
        out[d].is_slope = f.is_slope ? 1 : 0;
        out[d].slope_has_intercept = f.slope_has_intercept ? 1 : 0;
        out[d].slope_values = f.slope_values;
        out[d].slope_sum_z = f.slope_sum_z;
        out[d].slope_sum_zz = f.slope_sum_zz;
    }
    return out;
}

std::vector<int> to_order(const std::vector<std::size_t>& order) {
    return std::vector<int>(order.begin(), order.end());
}

xfe_options to_options(const HdfeOptions& options, AbsorptionMethod method, int precision) {
    xfe_options o{};
    o.tol = options.tol;
    o.tolerance_mode = static_cast<int>(options.tolerance_mode);
    o.max_iter = options.max_iter;
    o.method = static_cast<int>(method);  // Jacobi / Lsmr / Mlsmr are refused by the library
    o.force_cg = (options.convergence_criterion == ConvergenceCriterion::Reghdfe ||
                  options.convergence_criterion == ConvergenceCriterion::Both)
                     ? 1
                     : 0;
    o.precision = precision;  // 0: the plain FP64 method; 1: mixed-precision Schur PCG
    o.certify = 1;            // the normal-equation certificate, computed on the device
    o.refine_tol = 0.0;       // min(3e-13, 3e-5 tol)
    return o;
}

// savefe (recover_fixed_effects_group_ids, src/fe_absorption.cpp:15470-15555)
// calls the absorber with retain_fixed_effects and an empty X, then reads
// fe_alpha_y and checks it with fe_recovery_max_delta_cuda itself.  The
// library solves for the level values in one call.
bool recover_into_result(const Eigen::Ref<const Eigen::VectorXd>& partial, const std::vector<GpuFeInput>& fe_inputs,
                         const Eigen::VectorXd* weights, const std::vector<std::size_t>& sweep_order,
                         const HdfeOptions& options, AbsorptionResult& result) {
    const int n = static_cast<int>(partial.size());
    const int D = static_cast<int>(fe_inputs.size());
    std::vector<xfe_fe_input> fes = to_inputs(fe_inputs);
    std::vector<int> order = to_order(sweep_order);
    result.fe_alpha_y.assign(D, Eigen::VectorXd());
    result.fe_alpha_X.assign(D, Eigen::MatrixXd());
    std::vector<double*> alpha(D);
    for (int d = 0; d < D; ++d) {
        result.fe_alpha_y[d].resize(fe_inputs[d].num_groups);
        result.fe_alpha_X[d].resize(fe_inputs[d].num_groups, 0);
        alpha[d] = result.fe_alpha_y[d].data();
    }
    xfe_options o = to_options(options, AbsorptionMethod::Auto, 1);
    o.certify = 0;
    xfe_recovery_result r{};
    if (!xfe_recover_fixed_effects(partial.data(), n, fes.data(), D, weights ? weights->data() : nullptr,
                                   order.empty() ? nullptr : order.data(), &o, options.fe_tolerance, alpha.data(),
                                   /*contributions_out=*/nullptr, &r))
        return false;  // e.g. weights with two or more dimensions: the caller's CPU recovery
    result.y_tilde = partial;  // not read by the savefe caller
    result.X_tilde.resize(n, 0);
    result.iterations = r.iterations;
    result.converged = r.absorb_converged != 0;
    result.fe_group_ids.resize(D);
    result.fe_weight_sums.resize(D);
    for (int d = 0; d < D; ++d) {
        result.fe_group_ids[d].assign(fe_inputs[d].group_ids, fe_inputs[d].group_ids + n);
        result.fe_weight_sums[d] = Eigen::Map<const Eigen::VectorXd>(fe_inputs[d].weight_sums, fe_inputs[d].num_groups);
    }
    return true;
}

}  // namespace

bool cuda_backend_available() { return xfe_backend_available() != 0; }

bool absorb_fixed_effects_cuda(const Eigen::Ref<const Eigen::VectorXd>& y, const Eigen::Ref<const Eigen::MatrixXd>& X,
                               const std::vector<GpuFeInput>& fe_inputs, const Eigen::VectorXd* weights,
                               const std::vector<std::size_t>& sweep_order, const HdfeOptions& options,
                               AbsorptionMethod method, AbsorptionResult& result) {
    const int n = static_cast<int>(y.size());
    const int k = static_cast<int>(X.cols());
    if (X.rows() != n || (k > 0 && X.outerStride() != n)) return false;  // the ABI takes ld = n
    if (options.retain_fixed_effects && k == 0)
        return recover_into_result(y, fe_inputs, weights, sweep_order, options, result);
    std::vector<xfe_fe_input> fes = to_inputs(fe_inputs);
    std::vector<int> order = to_order(sweep_order);
    const xfe_options o = to_options(options, method, 0);
    xfe_absorb_result r{};
    result.y_tilde.resize(n);
    result.X_tilde.resize(n, k);
    if (!xfe_absorb(y.data(), k > 0 ? X.data() : nullptr, n, k, fes.data(), static_cast<int>(fes.size()),
                    weights ? weights->data() : nullptr, order.empty() ? nullptr : order.data(), &o, &r,
                    result.y_tilde.data(), k > 0 ? result.X_tilde.data() : nullptr))
        return false;
    result.iterations = r.iterations;
    result.converged = r.converged != 0;
    result.sweep_order_used.assign(r.sweep_order_used, r.sweep_order_used + fes.size());
    result.fe_levels.assign(r.fe_levels, r.fe_levels + fes.size());
    if (r.certificate_computed) {
        result.abs_residual = r.abs_residual;
        result.abs_residual_rel = r.abs_residual_rel;
        result.precision_certified = r.precision_certified != 0;
    }
    if (options.retain_fixed_effects) {
        // savefe with regressors: the ids and weight sums are returned, the per-level
        // alphas are not, so the estimator skips its hybrid recovery
        // (src/hdfe_regressor_v11.cpp:9169-9190, "missing_alpha_buffers") and recovers the
        // fixed effects through recover_fixed_effects_group_ids, which lands in
        // recover_into_result above.
        result.fe_group_ids.resize(fe_inputs.size());
        result.fe_weight_sums.resize(fe_inputs.size());
        for (std::size_t d = 0; d < fe_inputs.size(); ++d) {
            result.fe_group_ids[d].assign(fe_inputs[d].group_ids, fe_inputs[d].group_ids + n);
            if (fe_inputs[d].weight_sums)
                result.fe_weight_sums[d] =
                    Eigen::Map<const Eigen::VectorXd>(fe_inputs[d].weight_sums, fe_inputs[d].num_groups);
        }
        result.fe_alpha_y.clear();
        result.fe_alpha_X.clear();
        result.fe_means.clear();
    }
    return true;
}

bool absorb_fixed_effects_group_individual_cuda(const Eigen::Ref<const Eigen::VectorXd>& y,
                                                const Eigen::Ref<const Eigen::MatrixXd>& X,
                                                const std::vector<GpuFeInput>& standard_fe_inputs,
                                                const GroupIndividualStructure& gi, const Eigen::VectorXd* weights,
                                                const std::vector<std::size_t>& sweep_order,
                                                const HdfeOptions& options, AbsorptionMethod method,
                                                AbsorptionResult& result) {
    const int n = static_cast<int>(y.size());
    const int k = static_cast<int>(X.cols());
    if (X.rows() != n || (k > 0 && X.outerStride() != n)) return false;
    std::vector<xfe_fe_input> fes = to_inputs(standard_fe_inputs);
    const xfe_group_individual g{gi.num_groups,           gi.num_individuals,       gi.group_ptr.data(),
                                 gi.group_individual.data(), gi.individual_ptr.data(), gi.individual_group.data(),
                                 gi.group_scale.data()};
    std::vector<int> order = to_order(sweep_order);
    const xfe_options o = to_options(options, method, 0);
    xfe_absorb_result r{};
    result.y_tilde.resize(n);
    result.X_tilde.resize(n, k);
    if (!xfe_absorb_group_individual(y.data(), k > 0 ? X.data() : nullptr, n, k, fes.data(),
                                     static_cast<int>(fes.size()), &g, weights ? weights->data() : nullptr,
                                     order.empty() ? nullptr : order.data(), &o, &r, result.y_tilde.data(),
                                     k > 0 ? result.X_tilde.data() : nullptr))
        return false;
    result.iterations = r.iterations;
    result.converged = r.converged != 0;  // = certified, as the reference sets it
    result.precision_certified = r.precision_certified != 0;
    result.abs_residual = r.abs_residual;
    result.abs_residual_rel = r.abs_residual_rel;
    result.sweep_order_used.assign(r.sweep_order_used, r.sweep_order_used + fes.size());
    return true;
}

bool fe_recovery_max_delta_cuda(const Eigen::Ref<const Eigen::VectorXd>& residual,
                                const std::vector<GpuFeInput>& fe_inputs, const Eigen::VectorXd* weights,
                                const std::vector<std::size_t>& sweep_order, double& out_max_delta) {
    std::vector<xfe_fe_input> fes = to_inputs(fe_inputs);
    std::vector<int> order = to_order(sweep_order);
    return xfe_recovery_max_delta(residual.data(), static_cast<int>(residual.size()), fes.data(),
                                  static_cast<int>(fes.size()), weights ? weights->data() : nullptr,
                                  order.empty() ? nullptr : order.data(), &out_max_delta) != 0;
}

// The reference caches the uploaded ids between savefe checks; the library keeps its
// device workspace between calls, so the cached form is the same call.
bool fe_recovery_max_delta_cuda_cached(const Eigen::Ref<const Eigen::VectorXd>& residual,
                                       const std::vector<GpuFeInput>& fe_inputs, const Eigen::VectorXd* weights,
                                       const std::vector<std::size_t>& sweep_order, double& out_max_delta) {
    return fe_recovery_max_delta_cuda(residual, fe_inputs, weights, sweep_order, out_max_delta);
}

// certify_cuda_absorption_result (src/fe_absorption_cuda_certificate.cpp:318), standard FEs.
void certify_cuda_absorption_result(const Eigen::Ref<const Eigen::VectorXd>& y,
                                    const Eigen::Ref<const Eigen::MatrixXd>& X, const std::vector<Eigen::VectorXi>& fes,
                                    const Eigen::VectorXd* weights, const HdfeOptions& options,
                                    const std::vector<HeterogeneousSlopeTerm>& slopes,
                                    const std::vector<GpuFeCertificateView>& views, AbsorptionResult& result) {
    if (!slopes.empty() || views.size() != fes.size()) {
        // slope terms: the host certificate (certify_absorption_result) is the one to use
        certify_absorption_result(y, X, fes, weights, options, slopes, result);
        return;
    }
    std::vector<xfe_fe_input> in(views.size());
    for (std::size_t d = 0; d < views.size(); ++d) {
        in[d] = xfe_fe_input{};
        in[d].group_ids = views[d].group_ids;
        in[d].num_groups = views[d].num_groups;
        in[d].num_levels_present = views[d].num_groups;
    }
    xfe_options o{};
    o.tol = options.tol;
    o.tolerance_mode = static_cast<int>(options.tolerance_mode);
    xfe_certificate c{};
    const int k = static_cast<int>(X.cols());
    if (!xfe_certify(y.data(), k > 0 ? X.data() : nullptr, result.y_tilde.data(),
                     k > 0 ? result.X_tilde.data() : nullptr, static_cast<int>(y.size()), k, in.data(),
                     static_cast<int>(in.size()), weights ? weights->data() : nullptr, &o, result.converged ? 1 : 0,
                     &c)) {
        certify_absorption_result(y, X, fes, weights, options, slopes, result);  // device failure: host path
        return;
    }
    result.abs_residual = c.abs_residual;
    result.abs_residual_rel = c.abs_residual_rel;
    result.precision_certified = c.precision_certified != 0;
}

// Shadow-mode device certificate (src/fe_absorption_cuda.hpp:55-103).
struct CudaCertShadowHandle {
    xfe_cert_shadow* h;
};

CudaCertShadowHandle* cuda_cert_shadow_begin(const double* d_y_original, const double* d_x_original, int n, int cols,
                                             int ld, const double* d_weights_or_null, const int* const* d_group_ids,
                                             const double* const* d_weight_sums, const int* num_groups, int num_dims,
                                             const double* const* d_slope_values,
                                             const unsigned char* slope_has_intercept) {
    xfe_cert_shadow* h = xfe_cert_shadow_begin(d_y_original, d_x_original, n, cols, ld, d_weights_or_null,
                                               d_group_ids, d_weight_sums, num_groups, num_dims, d_slope_values,
                                               slope_has_intercept);
    return h ? new CudaCertShadowHandle{h} : nullptr;
}

bool cuda_cert_shadow_finish(CudaCertShadowHandle* handle, const double* d_y_residual, const double* d_x_residual,
                             double certificate_limit, CudaCertShadowResult* out) {
    xfe_cert_shadow_result r{};
    const int ok = handle ? xfe_cert_shadow_finish(handle->h, d_y_residual, d_x_residual, certificate_limit, &r) : 0;
    delete handle;
    out->verdict = ok ? r.verdict : 3;
    out->host_like_pass = r.host_like_pass;
    out->rel_max = r.rel_max;
    out->rel_max_perdim = r.rel_max_perdim;
    out->r_lower = r.r_lower;
    out->r_upper = r.r_upper;
    out->wsum_maxdiff = r.wsum_maxdiff;
    out->wsum_tol = r.wsum_tol;
    out->wsum_ok = r.wsum_ok;
    out->empty_groups = r.empty_groups;
    out->shadow_begin_ms = r.shadow_begin_ms;
    out->shadow_finish_ms = r.shadow_finish_ms;
    out->dims = r.dims;
    out->rhs = r.rhs;
    return ok != 0;
}

void cuda_cert_shadow_abort(CudaCertShadowHandle* handle) {
    if (handle) xfe_cert_shadow_abort(handle->h);
    delete handle;
}

}  // namespace detail
}  // namespace hdfe
