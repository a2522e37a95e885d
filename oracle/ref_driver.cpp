// Test infrastructure only: a C-ABI shim over the reference's own CPU core
// (xhdfe-xfe src/*.cpp compiled from /root/reference by oracle/Makefile into
// oracle/_ref/libxhdfe_ref.so).  It is the checker used to pin the oracle's
// restatement in this container; it never ships and never runs on a GPU.
//
// Entry points follow the reference's public estimator interface:
//   HdfeRegressorV11::partial_out  (src/hdfe_regressor_v11.cpp:10275) -- xfe
//   HdfeRegressorV11::fit          (src/hdfe_regressor_v11.cpp:7522)  -- xhdfe
//   detail::absorb_fixed_effects   (src/fe_absorption.cpp:5993)       -- raw CPU absorber
//   detail::certify_absorption_result (src/fe_absorption.cpp:14942)   -- normal-equation certificate
#include <Eigen/Dense>

#include <cstring>
#include <exception>
#include <cstdio>
#include <vector>

#include "fe_absorption.hpp"
#include "schwarz_demean.hpp"
#include "hdfe/hdfe_regressor.hpp"
#include "hdfe/hdfe_regressor_v11.hpp"
#include "hdfe/akm_kss.hpp"

namespace {

std::vector<Eigen::VectorXi> make_fes(const int* fes, int n, int d) {
    std::vector<Eigen::VectorXi> out;
    for (int k = 0; k < d; ++k) {
        Eigen::VectorXi v(n);
        std::memcpy(v.data(), fes + static_cast<std::size_t>(k) * n, sizeof(int) * n);
        out.push_back(std::move(v));
    }
    return out;
}

hdfe::HdfeOptions make_options(double tol, int tolerance_mode, int threads) {
    hdfe::HdfeOptions o;
    o.tol = tol;
    o.tolerance_mode = static_cast<hdfe::ToleranceMode>(tolerance_mode);
    o.num_threads = threads;
    return o;
}

}  // namespace

extern "C" {

// Raw CPU absorber on already-filtered rows (no singleton drop).
// y: n, X: n*k column-major, fes: d*n (dimension-major). Outputs overwrite
// y_out/X_out.  Returns 0 on success, -1 on exception.
// w: optional analytic weights (n), nullptr for unit weights.
int ref_absorb(int n, int k, int d, const double* y, const double* X, const int* fes, const double* w,
               double tol, int tolerance_mode, int threads, int symmetric,
               double* y_out, double* X_out, int* iterations, int* converged) {
    try {
        Eigen::VectorXd yv = Eigen::Map<const Eigen::VectorXd>(y, n);
        Eigen::MatrixXd Xm = Eigen::Map<const Eigen::MatrixXd>(X, n, k);
        auto fv = make_fes(fes, n, d);
        hdfe::HdfeOptions o = make_options(tol, tolerance_mode, threads);
        o.symmetric_sweep = symmetric != 0;
        Eigen::VectorXd wv;
        if (w) wv = Eigen::Map<const Eigen::VectorXd>(w, n);
        auto r = hdfe::detail::absorb_fixed_effects(yv, Xm, fv, w ? &wv : nullptr, o);
        std::memcpy(y_out, r.y_tilde.data(), sizeof(double) * n);
        if (k > 0) std::memcpy(X_out, r.X_tilde.data(), sizeof(double) * n * k);
        *iterations = r.iterations;
        *converged = r.converged ? 1 : 0;
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_absorb: %s\n", e.what());
        return -1;
    }
}

// The reference's FE indexer and weight sums (build_indexer, src/fe_absorption.cpp:845-940;
// compute_weight_sums, :942-1037), read back through absorb_fixed_effects_v6 with
// retain_fixed_effects, which returns indexers[d].group_ids, the weight sums and
// num_levels_present (src/fe_absorption.cpp:9444-9452; fe_levels :9043).  fes: d x n raw ids;
// w optional.  Outputs: gids (d x n), num_groups (d), levels_present (d), and the weight sums
// of dimension j at wsums + j * n (capacity n per dimension).  Returns 0, -1 on failure.
int ref_indexer(int n, int d, const int* fes, const double* w, int* gids, int* num_groups,
                int* levels_present, double* wsums) {
    try {
        Eigen::VectorXd y = Eigen::VectorXd::LinSpaced(n, 1.0, 2.0);
        Eigen::MatrixXd X(n, 0);
        auto fv = make_fes(fes, n, d);
        hdfe::HdfeOptions o = make_options(1e-8, 1, 1);
        o.retain_fixed_effects = true;
        Eigen::VectorXd wv;
        if (w) wv = Eigen::Map<const Eigen::VectorXd>(w, n);
        auto r = hdfe::detail::absorb_fixed_effects_v6(y, X, fv, w ? &wv : nullptr, o,
                                                       hdfe::AbsorptionMethod::Auto);
        if (static_cast<int>(r.fe_group_ids.size()) != d || static_cast<int>(r.fe_weight_sums.size()) != d ||
            static_cast<int>(r.fe_levels.size()) != d)
            return -1;
        for (int j = 0; j < d; ++j) {
            if (static_cast<int>(r.fe_group_ids[j].size()) != n) return -1;
            std::memcpy(gids + static_cast<std::size_t>(j) * n, r.fe_group_ids[j].data(), sizeof(int) * n);
            const int G = static_cast<int>(r.fe_weight_sums[j].size());
            num_groups[j] = G;
            levels_present[j] = r.fe_levels[j];
            std::memcpy(wsums + static_cast<std::size_t>(j) * n, r.fe_weight_sums[j].data(),
                        sizeof(double) * G);
        }
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_indexer: %s\n", e.what());
        return -1;
    }
}

// xfe path: singleton drop + absorption.  Returns the number of kept rows
// (>= 0) and fills sample_index (n_full capacity), y_out/X_out (kept rows).
int ref_partial_out(int n, int k, int d, const double* y, const double* X, const int* fes,
                    double tol, int tolerance_mode, int threads,
                    int* sample_index, double* y_out, double* X_out, int* iterations,
                    int* converged) {
    try {
        Eigen::VectorXd yv = Eigen::Map<const Eigen::VectorXd>(y, n);
        Eigen::MatrixXd Xm = Eigen::Map<const Eigen::MatrixXd>(X, n, k);
        auto fv = make_fes(fes, n, d);
        hdfe::v11::HdfeRegressorV11 reg(make_options(tol, tolerance_mode, threads));
        auto r = reg.partial_out(yv, Xm, fv, nullptr, nullptr, nullptr);
        const auto& res = reg.results();
        const int kept = static_cast<int>(r.y_tilde.size());
        for (int i = 0; i < kept; ++i) {
            sample_index[i] = res.sample_index.size() == kept ? res.sample_index[i] : i;
        }
        std::memcpy(y_out, r.y_tilde.data(), sizeof(double) * kept);
        if (k > 0) std::memcpy(X_out, r.X_tilde.data(), sizeof(double) * kept * k);
        *iterations = r.iterations;
        *converged = r.converged ? 1 : 0;
        return kept;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_partial_out: %s\n", e.what());
        return -1;
    }
}

// xhdfe path: singletons, absorption, OLS with SEs clustered on `cluster`
// (or robust when cluster == nullptr), reghdfe DoF, first-pair mobility
// groups.  Output arrays are sized by the caller: coef/se k, sample_index
// and groupvar n, fe_redundant/fe_levels d.  Returns kept rows or -1.
int ref_fit(int n, int k, int d, const double* y, const double* X, const int* fes,
            const int* cluster, double tol, int tolerance_mode, int threads,
            int want_groupvar, double* coef, double* se, int* sample_index, int* groupvar,
            double* df_a, double* df_resid, int* fe_redundant, int* fe_levels,
            int* num_singletons, int* iterations, double* rss) {
    try {
        Eigen::VectorXd yv = Eigen::Map<const Eigen::VectorXd>(y, n);
        Eigen::MatrixXd Xm = Eigen::Map<const Eigen::MatrixXd>(X, n, k);
        auto fv = make_fes(fes, n, d);
        hdfe::HdfeOptions o = make_options(tol, tolerance_mode, threads);
        o.save_groupvar = want_groupvar != 0;
        std::vector<Eigen::VectorXi> cl;
        if (cluster) {
            o.se_type = hdfe::StandardErrorType::Cluster;
            Eigen::VectorXi c(n);
            std::memcpy(c.data(), cluster, sizeof(int) * n);
            cl.push_back(std::move(c));
        }
        hdfe::v11::HdfeRegressorV11 reg(o);
        reg.fit(yv, Xm, fv, nullptr, cluster ? &cl : nullptr);
        const auto& r = reg.results();
        const int kc = static_cast<int>(r.coefficients.size());
        for (int j = 0; j < k && j < kc; ++j) {
            coef[j] = r.coefficients[j];
            se[j] = r.std_errors[j];
        }
        for (int i = 0; i < r.nobs; ++i) {
            sample_index[i] = r.sample_index.size() == r.nobs ? r.sample_index[i] : i;
            if (groupvar) groupvar[i] = r.groupvar.size() == r.nobs ? r.groupvar[i] : 0;
        }
        *df_a = r.df_a;
        *df_resid = r.df_resid;
        for (int j = 0; j < d; ++j) {
            fe_redundant[j] = j < static_cast<int>(r.fe_redundant.size()) ? r.fe_redundant[j] : -1;
            fe_levels[j] = j < static_cast<int>(r.fe_num_levels.size()) ? r.fe_num_levels[j] : -1;
        }
        *num_singletons = r.num_singletons;
        *iterations = r.num_iterations;
        *rss = r.rss;
        return r.nobs;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_fit: %s\n", e.what());
        return -1;
    }
}

// Certificate of a given within transform (y_t, X_t) of (y, X): weights may be
// NULL.  Outputs abs_residual, abs_residual_rel and precision_certified.
int ref_certify(int n, int k, int d, const double* y, const double* X, const int* fes, const double* w,
                double tol, int tolerance_mode, int converged, const double* y_t, const double* X_t,
                double* abs_residual, double* abs_residual_rel, int* certified) {
    try {
        Eigen::VectorXd yv = Eigen::Map<const Eigen::VectorXd>(y, n);
        Eigen::MatrixXd Xm = Eigen::Map<const Eigen::MatrixXd>(X, n, k);
        auto fv = make_fes(fes, n, d);
        hdfe::HdfeOptions o = make_options(tol, tolerance_mode, 1);
        hdfe::detail::AbsorptionResult r;
        r.y_tilde = Eigen::Map<const Eigen::VectorXd>(y_t, n);
        r.X_tilde = Eigen::Map<const Eigen::MatrixXd>(X_t, n, k);
        r.converged = converged != 0;
        Eigen::VectorXd wv;
        if (w) wv = Eigen::Map<const Eigen::VectorXd>(w, n);
        hdfe::detail::certify_absorption_result(yv, Xm, fv, w ? &wv : nullptr, o, {}, r);
        *abs_residual = r.abs_residual;
        *abs_residual_rel = r.abs_residual_rel;
        *certified = r.precision_certified ? 1 : 0;
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_certify: %s\n", e.what());
        return -1;
    }
}

// xhdfe::schwarz_demean_raw (src/schwarz_demean.cpp:419): PCG within transform.
// Returns 1 (converged; y_out/X_out hold the residuals), 0 (not converged or
// unsupported), -1 on exception.
int ref_schwarz(int n, int k, int d, const double* y, const double* X, const int* fes, const int* nlev,
                double tol, int max_iter, double* y_out, double* X_out, int* iters) {
    try {
        std::memcpy(y_out, y, sizeof(double) * n);
        if (k > 0) std::memcpy(X_out, X, sizeof(double) * static_cast<std::size_t>(n) * k);
        std::vector<const int32_t*> fp;
        std::vector<int> lev;
        for (int j = 0; j < d; ++j) {
            fp.push_back(reinterpret_cast<const int32_t*>(fes + static_cast<std::size_t>(j) * n));
            lev.push_back(nlev[j]);
        }
        const bool ok = xhdfe::schwarz_demean_raw(y_out, X_out, n, k, fp, lev, 1, tol, max_iter, iters, -1);
        return ok ? 1 : 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_schwarz: %s\n", e.what());
        return -1;
    }
}

// absorb_fixed_effects_v6 (src/fe_absorption.cpp:8750) with one heterogeneous
// slope term on dimension slope_dim (values z, with or without the level
// intercept), CPU backend.  Returns 0 on success, -1 on exception.
int ref_absorb_slope(int n, int k, int d, const double* y, const double* X, const int* fes, int slope_dim,
                     const double* z, int include_intercept, double tol, int tolerance_mode, int threads,
                     double* y_out, double* X_out, int* iterations, int* converged) {
    try {
        Eigen::VectorXd yv = Eigen::Map<const Eigen::VectorXd>(y, n);
        Eigen::MatrixXd Xm = Eigen::Map<const Eigen::MatrixXd>(X, n, k);
        auto fv = make_fes(fes, n, d);
        hdfe::HdfeOptions o = make_options(tol, tolerance_mode, threads);
        std::vector<hdfe::detail::HeterogeneousSlopeTerm> slopes(1);
        slopes[0].fe_index = slope_dim;
        slopes[0].values = Eigen::Map<const Eigen::VectorXd>(z, n);
        slopes[0].include_intercept = include_intercept != 0;
        auto r = hdfe::detail::absorb_fixed_effects_v6(yv, Xm, fv, nullptr, o, hdfe::AbsorptionMethod::Auto,
                                                       slopes);
        std::memcpy(y_out, r.y_tilde.data(), sizeof(double) * n);
        if (k > 0) std::memcpy(X_out, r.X_tilde.data(), sizeof(double) * n * k);
        *iterations = r.iterations;
        *converged = r.converged ? 1 : 0;
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_absorb_slope: %s\n", e.what());
        return -1;
    }
}

// akm_kss_decompose (src/akm_kss.cpp:2286) with JLA leverages at match level,
// no pruning (the caller passes a leave-out connected sample): the working
// rows (original worker / firm ids, person-years) and their P_ii.  Outputs
// have capacity n; returns the number of working rows, -1 on exception.
int ref_akm_jla_pii(int n, const double* y, const int* worker, const int* firm, int draws,
                    unsigned long long seed, int* row_worker, int* row_firm, double* row_weight,
                    double* pii) {
    try {
        Eigen::VectorXd yv = Eigen::Map<const Eigen::VectorXd>(y, n);
        Eigen::VectorXi wv = Eigen::Map<const Eigen::VectorXi>(worker, n);
        Eigen::VectorXi fv = Eigen::Map<const Eigen::VectorXi>(firm, n);
        hdfe::akm::AkmOptions o;
        o.leverage_method = hdfe::akm::LeverageMethod::Jla;
        o.jla_draws = draws;
        o.seed = seed;
        o.prune = false;
        o.num_threads = 1;
        auto r = hdfe::akm::akm_kss_decompose(yv, wv, fv, nullptr, o);
        const int R = static_cast<int>(r.pii.size());
        for (int k = 0; k < R && k < n; ++k) {
            row_worker[k] = r.row_worker[k];
            row_firm[k] = r.row_firm[k];
            row_weight[k] = r.row_weight[k];
            pii[k] = r.pii[k];
        }
        return R;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_akm_jla_pii: %s\n", e.what());
        return -1;
    }
}

// akm_kss_decompose with exact leverages (LeverageMethod::Exact,
// src/akm_kss.cpp:2667-2741), match level, no pruning: the working rows, their
// P_ii, the components (as ref_akm_jla_components) and sigma_i.  Row outputs
// have capacity n; returns the number of working rows, -1 on exception.
int ref_akm_exact(int n, const double* y, const int* worker, const int* firm, int* row_worker, int* row_firm,
                  double* row_weight, double* pii, double* comp, double* sigma) {
    try {
        Eigen::VectorXd yv = Eigen::Map<const Eigen::VectorXd>(y, n);
        Eigen::VectorXi wv = Eigen::Map<const Eigen::VectorXi>(worker, n);
        Eigen::VectorXi fv = Eigen::Map<const Eigen::VectorXi>(firm, n);
        hdfe::akm::AkmOptions o;
        o.leverage_method = hdfe::akm::LeverageMethod::Exact;
        o.prune = false;
        o.num_threads = 1;
        auto r = hdfe::akm::akm_kss_decompose(yv, wv, fv, nullptr, o);
        const int R = static_cast<int>(r.pii.size());
        for (int k = 0; k < R && k < n; ++k) {
            row_worker[k] = r.row_worker[k];
            row_firm[k] = r.row_firm[k];
            row_weight[k] = r.row_weight[k];
            pii[k] = r.pii[k];
            sigma[k] = r.sigma_i[k];
        }
        const hdfe::akm::AkmComponents* cs[3] = {&r.plugin, &r.agsu, &r.kss};
        for (int i = 0; i < 3; ++i) {
            comp[3 * i] = cs[i]->var_alpha;
            comp[3 * i + 1] = cs[i]->var_psi;
            comp[3 * i + 2] = cs[i]->cov_alpha_psi;
        }
        comp[9] = r.sigma2_ho;
        comp[10] = r.row_weight.sum();
        return R;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_akm_exact: %s\n", e.what());
        return -1;
    }
}

// The same decomposition's variance components and KSS sigma_i
// (src/akm_kss.cpp:3354-3412): comp = plug-in, AGSU and KSS
// (var_alpha, var_psi, cov_alpha_psi) each, then sigma2_ho and n_py; sigma
// has capacity n.  Returns the number of working rows, -1 on exception.
int ref_akm_jla_components(int n, const double* y, const int* worker, const int* firm, int draws,
                           unsigned long long seed, double* comp, double* sigma) {
    try {
        Eigen::VectorXd yv = Eigen::Map<const Eigen::VectorXd>(y, n);
        Eigen::VectorXi wv = Eigen::Map<const Eigen::VectorXi>(worker, n);
        Eigen::VectorXi fv = Eigen::Map<const Eigen::VectorXi>(firm, n);
        hdfe::akm::AkmOptions o;
        o.leverage_method = hdfe::akm::LeverageMethod::Jla;
        o.jla_draws = draws;
        o.seed = seed;
        o.prune = false;
        o.num_threads = 1;
        auto r = hdfe::akm::akm_kss_decompose(yv, wv, fv, nullptr, o);
        const hdfe::akm::AkmComponents* cs[3] = {&r.plugin, &r.agsu, &r.kss};
        for (int i = 0; i < 3; ++i) {
            comp[3 * i] = cs[i]->var_alpha;
            comp[3 * i + 1] = cs[i]->var_psi;
            comp[3 * i + 2] = cs[i]->cov_alpha_psi;
        }
        comp[9] = r.sigma2_ho;
        comp[10] = r.row_weight.sum();
        const int R = static_cast<int>(r.sigma_i.size());
        for (int k = 0; k < R && k < n; ++k) sigma[k] = r.sigma_i[k];
        return R;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_akm_jla_components: %s\n", e.what());
        return -1;
    }
}

// detail::recover_fixed_effects (src/fe_absorption.cpp:15316-15354, impl
// :14447-14935): FE values of `partial` by the reference's CPU recovery.
// contrib has capacity d*n (dimension-major): contributions[dim][i] =
// alpha_dim[fes_dim[i]]; tol is HdfeOptions::fe_tolerance.  Returns 0 on success, -1 on exception.
int ref_recover_fe(int n, int d, const double* partial, const int* fes, double tol, int threads,
                   double* contrib, int* iterations, double* max_delta, int* converged) {
    try {
        Eigen::VectorXd pv = Eigen::Map<const Eigen::VectorXd>(partial, n);
        auto fv = make_fes(fes, n, d);
        hdfe::HdfeOptions o = make_options(1e-8, 1, threads);
        o.fe_tolerance = tol;
        o.max_iter = 100000;
        auto r = hdfe::detail::recover_fixed_effects(pv, fv, nullptr, o);
        for (int k = 0; k < d; ++k)
            std::memcpy(contrib + static_cast<std::size_t>(k) * n, r.contributions[k].data(), sizeof(double) * n);
        *iterations = r.iterations;
        *max_delta = r.max_delta;
        *converged = r.converged ? 1 : 0;
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_recover_fe: %s\n", e.what());
        return -1;
    }
}

// detail::absorb_fixed_effects_group_individual (src/fe_absorption.cpp:9741):
// the reference's CPU group()/individual() absorber.  gi arrays as in
// GroupIndividualStructure (include/fe_absorption.hpp:13-21); w may be NULL;
// method 1 GaussSeidel, 2 SymmetricGaussSeidel.  Returns 0 on success.
int ref_absorb_gi(int n, int k, int d, const double* y, const double* X, const int* fes, int m,
                  const int* group_ptr, const int* group_individual, const int* individual_ptr,
                  const int* individual_group, const double* group_scale, const double* w, double tol,
                  int method, double* y_out, double* X_out, int* iterations, int* converged) {
    try {
        Eigen::VectorXd yv = Eigen::Map<const Eigen::VectorXd>(y, n);
        Eigen::MatrixXd Xm = Eigen::Map<const Eigen::MatrixXd>(X, n, k);
        auto fv = make_fes(fes, n, d);
        hdfe::detail::GroupIndividualStructure gi;
        gi.num_groups = n;
        gi.num_individuals = m;
        const int E = group_ptr[n];
        gi.group_ptr.assign(group_ptr, group_ptr + n + 1);
        gi.group_individual.assign(group_individual, group_individual + E);
        gi.individual_ptr.assign(individual_ptr, individual_ptr + m + 1);
        gi.individual_group.assign(individual_group, individual_group + E);
        gi.group_scale.assign(group_scale, group_scale + n);
        hdfe::HdfeOptions o = make_options(tol, 1, 1);
        Eigen::VectorXd wv;
        if (w) wv = Eigen::Map<const Eigen::VectorXd>(w, n);
        auto r = hdfe::detail::absorb_fixed_effects_group_individual(
            yv, Xm, fv, gi, w ? &wv : nullptr, o,
            method == 2 ? hdfe::AbsorptionMethod::SymmetricGaussSeidel : hdfe::AbsorptionMethod::GaussSeidel);
        std::memcpy(y_out, r.y_tilde.data(), sizeof(double) * n);
        if (k > 0) std::memcpy(X_out, r.X_tilde.data(), sizeof(double) * n * k);
        *iterations = r.iterations;
        *converged = r.converged ? 1 : 0;
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_absorb_gi: %s\n", e.what());
        return -1;
    }
}

// hdfe::akm::leave_out_connected_set (src/akm_kss.cpp:2233): keep (n) and
// info = n_obs_input, n_obs_connected, n_obs, n_workers, n_firms, n_matches,
// n_movers, n_stayers, prune_iterations.  Returns 0 on success.
int ref_leave_out_set(int n, const int* worker, const int* firm, unsigned char* keep, long long* info) {
    try {
        Eigen::VectorXi wv = Eigen::Map<const Eigen::VectorXi>(worker, n);
        Eigen::VectorXi fv = Eigen::Map<const Eigen::VectorXi>(firm, n);
        hdfe::akm::LeaveOutSetOptions o;
        o.num_threads = 1;
        auto r = hdfe::akm::leave_out_connected_set(wv, fv, nullptr, o);
        for (int i = 0; i < n; ++i) keep[i] = r.keep[i];
        const long long v[9] = {r.n_obs_input, r.n_obs_connected, r.n_obs, r.n_workers, r.n_firms,
                                r.n_matches, r.n_movers, r.n_stayers, r.prune_iterations};
        for (int k = 0; k < 9; ++k) info[k] = v[k];
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ref_leave_out_set: %s\n", e.what());
        return -1;
    }
}

}  // extern "C"
