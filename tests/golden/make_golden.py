"""Writes tests/golden/ref_cases.npz from the reference's own CPU core.

Run in the build container only (it needs oracle/_ref/libxhdfe_ref.so, which
oracle/Makefile compiles from the reference sources).  Inputs are regenerated
from the seeded panel generators, so the fixture holds only the reference's
outputs plus checksums of the inputs they were computed from.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "xhdfe-xfe_amd"))

import oracle  # noqa: E402
from golden_cases import (CASES, JLA_DRAWS, JLA_SEED, SLOPE_CASES, certificate_weights, indexer_cases,  # noqa: E402
                          gi_panel, gi_weights, input_checksum, jla_panel, leave_out_panels, schwarz_edge_cases, slope_values)


def main():
    if oracle.ref() is None:
        raise SystemExit("oracle/_ref/libxhdfe_ref.so is not built (make -C oracle ref)")
    out = {}
    for name, make in CASES.items():
        p = make()
        y, X, fes = p["y"], p["X"], p["fes"]
        out[f"{name}/input_checksum"] = input_checksum(p)
        # xfe path (singletons + the estimator's own absorber choice; small
        # panels may take its sparse direct solve)
        idx, yo, Xo, it, conv = oracle.ref_partial_out(y, X, fes, tol=1e-12)
        out[f"{name}/xfe/sample_index"] = idx
        out[f"{name}/xfe/y_tilde"] = yo
        out[f"{name}/xfe/X_tilde"] = Xo
        # the MAP absorber itself (absorb_fixed_effects) on the kept rows, with
        # the method the estimator picks (symmetric GS for two FEs)
        yk, Xk = y[idx], X[idx]
        fk = [f[idx] for f in fes]
        for tag, tol in (("tight", 1e-12), ("default", 1e-8)):
            yo, Xo, it, conv = oracle.ref_absorb(yk, Xk, fk, tol=tol, symmetric=len(fes) == 2)
            out[f"{name}/{tag}/y_tilde"] = yo
            out[f"{name}/{tag}/X_tilde"] = Xo
            out[f"{name}/{tag}/iterations"] = np.int64(it)
            out[f"{name}/{tag}/converged"] = np.int64(conv)
            # the reference's certificate of its own output (certify_absorption_result)
            c = oracle.ref_certify(yk, Xk, yo, Xo, fk, tol=tol, converged=conv)
            out[f"{name}/{tag}/cert"] = np.array([c["abs_residual"], c["abs_residual_rel"],
                                                  float(c["certified"])])
            # and with row weights (the certificate's operator norm and sums use w)
            w = certificate_weights(yk.size)
            c = oracle.ref_certify(yk, Xk, yo, Xo, fk, weights=w, tol=tol, converged=conv)
            out[f"{name}/{tag}/cert_w"] = np.array([c["abs_residual"], c["abs_residual_rel"],
                                                    float(c["certified"])])
        # the weighted within transform (absorb_fixed_effects with analytic weights)
        w = certificate_weights(yk.size)
        for tag, tol in (("tight", 1e-12), ("default", 1e-8)):
            yo, Xo, it, conv = oracle.ref_absorb(yk, Xk, fk, tol=tol, symmetric=len(fes) == 2, weights=w)
            out[f"{name}/w_{tag}/y_tilde"] = yo
            out[f"{name}/w_{tag}/X_tilde"] = Xo
            out[f"{name}/w_{tag}/iterations"] = np.int64(it)
            out[f"{name}/w_{tag}/converged"] = np.int64(conv)
        # savefe recovery of y's FE values (recover_fixed_effects), at the
        # default fe_tolerance and tightly
        for tag, ftol in (("default", 1e-6), ("tight", 1e-12)):
            contrib, it, md, conv = oracle.ref_recover_fe(yk, fk, tol=ftol)
            out[f"{name}/recover_{tag}/contributions"] = contrib
            out[f"{name}/recover_{tag}/info"] = np.array([it, md, conv], dtype=np.float64)
        # schwarz_demean_raw (the PCG demeaner) on the kept rows with compact ids
        ck = [oracle.build_indexer(f) for f in fk]
        for tag, tol in (("tight", 1e-12), ("default", 1e-8)):
            ok, yo, Xo, it = oracle.ref_schwarz(yk, Xk, [c[0] for c in ck], [c[1] for c in ck], tol=tol)
            out[f"{name}/schwarz_{tag}/y_tilde"] = yo
            out[f"{name}/schwarz_{tag}/X_tilde"] = Xo
            out[f"{name}/schwarz_{tag}/iterations"] = np.int64(it)
            out[f"{name}/schwarz_{tag}/ok"] = np.int64(ok)
        # one heterogeneous slope term (absorb_fixed_effects_v6 with slopes)
        for sname, sdim, icpt in SLOPE_CASES:
            if sname != name:
                continue
            z = slope_values(yk.size)
            yo, Xo, it, conv = oracle.ref_absorb_slope(yk, Xk, fk, sdim, z, intercept=icpt, tol=1e-12)
            tag = f"slope{sdim}_{'icpt' if icpt else 'noicpt'}"
            out[f"{name}/{tag}/y_tilde"] = yo
            out[f"{name}/{tag}/X_tilde"] = Xo
            out[f"{name}/{tag}/iterations"] = np.int64(it)
            out[f"{name}/{tag}/converged"] = np.int64(conv)
        # certificate of the untransformed columns (large residual, not certified)
        c = oracle.ref_certify(yk, Xk, yk, Xk, fk, tol=1e-8, converged=True)
        out[f"{name}/raw/cert"] = np.array([c["abs_residual"], c["abs_residual_rel"], float(c["certified"])])
        f = oracle.ref_fit(y, X, fes, cluster=p["cluster"], tol=1e-12)
        for key in ("coef", "se", "sample_index", "groupvar", "fe_redundant", "fe_levels"):
            out[f"{name}/fit/{key}"] = np.asarray(f[key])
        for key in ("df_a", "df_resid", "num_singletons", "nobs", "rss"):
            out[f"{name}/fit/{key}"] = np.asarray(f[key])
    # the reference's indexer and weight sums (build_indexer, compute_weight_sums)
    for name, (raw, w) in indexer_cases().items():
        (g, G, lp, ws), = oracle.ref_indexer([raw])
        (gw, Gw, lpw, wsw), = oracle.ref_indexer([raw], weights=w)
        assert np.array_equal(g, gw) and G == Gw and lp == lpw
        out[f"indexer/{name}/gids"] = g
        out[f"indexer/{name}/info"] = np.array([G, lp], dtype=np.int64)
        out[f"indexer/{name}/counts"] = ws
        out[f"indexer/{name}/wsums"] = wsw
    # schwarz_demean_raw's gates and edge cases on the two-way panel: the
    # returned flag and iteration count per case (order of SCHWARZ_EDGE_CASES)
    out["schwarz_edges"] = np.array([list(oracle.ref_schwarz(*args, tol=1e-10, max_iter=mi)[0::3])
                                     for args, mi in schwarz_edge_cases()], dtype=np.int64)
    # JLA leverages of the reference's KSS decomposition (match level, no pruning)
    jp = jla_panel()
    rw, rf, rc, pii = oracle.ref_akm_jla_pii(jp["y"], jp["worker"], jp["firm"], JLA_DRAWS, JLA_SEED)
    out["jla/row_worker"], out["jla/row_firm"], out["jla/row_weight"], out["jla/pii"] = rw, rf, rc, pii
    # its variance components (plug-in, AGSU, KSS; sigma2_ho, n_py) and KSS
    # sigma_i: linear functionals of its B_ii draws
    out["jla/components"], out["jla/sigma_i"] = oracle.ref_akm_jla_components(
        jp["y"], jp["worker"], jp["firm"], JLA_DRAWS, JLA_SEED)
    # group()/individual() absorption (absorb_fixed_effects_group_individual)
    gp = gi_panel()
    for tag, w, method in (("gs", None, 1), ("sgs_w", gi_weights(gp["y"].size), 2)):
        yo, Xo, it, conv = oracle.ref_absorb_gi(gp["y"], gp["X"], gp["fes"], gp["gi"], weights=w, tol=1e-8,
                                                method=method)
        out[f"gi/{tag}/y_tilde"], out[f"gi/{tag}/X_tilde"] = yo, Xo
        out[f"gi/{tag}/info"] = np.array([it, conv], dtype=np.int64)
    # the leave-out connected set (leave_out_connected_set)
    for name, (w, f) in leave_out_panels().items():
        keep, info = oracle.ref_leave_out_set(w, f)
        out[f"los/{name}/keep"] = keep
        out[f"los/{name}/info"] = info
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_cases.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
