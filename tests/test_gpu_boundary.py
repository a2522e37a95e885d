"""The boundary rows on an MI355X, through the C ABI: the device indexer and
weight sums against the reference's build_indexer / compute_weight_sums
(fixtures), the weighted within transform against the reference's weighted
absorber (fixtures) with the weight sums computed on the device, the
not-converged and refused-method paths, the precision report, and the OLS
coefficients and worker-clustered SEs formed from the device's transform,
Gram and cluster meat against the reference's fit.

Tolerances: integer results exact; weight sums 1e-13 relative; transforms
within 1e-10 (tol 1e-12) and 1e-8 (tol 1e-8) of the column scale; coefficients
1e-9 and SEs 1e-7 relative (the stated bar for coefficients and clustered SEs).
"""
import sys

import numpy as np
import pytest

import oracle
import xfe_amd
from golden_cases import certificate_weights, indexer_cases

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _kept(p):
    keep, _, _ = xfe_amd.drop_singletons(p["fes"])
    return keep


def test_build_indexer_matches_reference(golden):
    _log("build_indexer on the device vs the reference's indexer")
    for name, (raw, w) in indexer_cases().items():
        gid, G, lp, cont = xfe_amd.build_indexer(raw)
        np.testing.assert_array_equal(gid, golden[f"indexer/{name}/gids"], err_msg=name)
        assert [G, lp] == list(golden[f"indexer/{name}/info"]), name
        assert cont == bool(np.all(np.diff(gid) >= 0)), name
        np.testing.assert_array_equal(xfe_amd.weight_sums(gid, G), golden[f"indexer/{name}/counts"])
        np.testing.assert_allclose(xfe_amd.weight_sums(gid, G, w), golden[f"indexer/{name}/wsums"],
                                   rtol=1e-13, atol=1e-13)


def test_build_indexer_large_and_device_form():
    _log("build_indexer at 2M rows: dense, sparse and device-tensor forms vs the oracle")
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    n = 2_000_000
    for raw in (rng.integers(0, 600_000, n).astype(np.int32),                 # dense, gaps kept
                rng.integers(-2 ** 31, 2 ** 31 - 1, n, dtype=np.int64).astype(np.int32),  # sparse
                (rng.integers(0, 300_000, n) * 7).astype(np.int32)):          # dense range, re-ranked
        g, G, lp, c = xfe_amd.build_indexer(raw)
        og, oG, olp, oc = oracle.build_indexer(raw)
        np.testing.assert_array_equal(g, og)
        assert (G, lp, c) == (oG, olp, oc)
        t = torch.from_numpy(raw).to("cuda")
        tg, tG, tlp, tc = xfe_amd.build_indexer_device(t)
        np.testing.assert_array_equal(tg.cpu().numpy(), og)
        assert (tG, tlp, tc) == (oG, olp, oc)


def test_weight_sums_rejects_bad_ids():
    with pytest.raises(xfe_amd.XfeError, match="outside"):
        xfe_amd.weight_sums(np.array([0, 1, 5], dtype=np.int32), 3)


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
@pytest.mark.parametrize("tag", ["tight", "default"])
def test_absorb_weighted_matches_reference(golden, cases, name, tag):
    """Analytic weights, weight sums computed on the device (weight_sums NULL),
    against the reference's weighted absorber (fixtures)."""
    _log(f"weighted absorb vs reference: {name} {tag}")
    p = cases[name]
    keep = _kept(p)
    fes = [f[keep] for f in p["fes"]]
    idx = [xfe_amd.build_indexer(f) for f in fes]
    gids, G = [i[0] for i in idx], [i[1] for i in idx]
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    w = certificate_weights(V.shape[0])
    tol = 1e-12 if tag == "tight" else 1e-8
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, weights=w, tol=tol)
    ref = np.column_stack([golden[f"{name}/w_{tag}/y_tilde"], golden[f"{name}/w_{tag}/X_tilde"]])
    err = np.abs(np.column_stack([yg, Xg]) - ref).max() / np.abs(V).max()
    _log(f"  iterations {info['iterations']} (reference {int(golden[f'{name}/w_{tag}/iterations'])}) err {err:.2e}")
    assert info["converged"]
    assert err <= (1e-10 if tag == "tight" else 1e-8), err


def test_absorb_not_converged_path(cases):
    """max_iter too small: the call succeeds (return 1) with converged = 0,
    the reference's 'true with converged=false' (caller status 3)."""
    p = cases["akm_three_way"]
    keep = _kept(p)
    fes = [f[keep] for f in p["fes"]]
    idx = [xfe_amd.build_indexer(f) for f in fes]
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], [i[0] for i in idx], [i[1] for i in idx], tol=1e-12,
                                  max_iter=1)
    assert not info["converged"]
    assert np.all(np.isfinite(yg)) and np.all(np.isfinite(Xg))


def test_refused_method_on_device(cases):
    p = cases["two_way"]
    keep = _kept(p)
    fes = [f[keep] for f in p["fes"]]
    idx = [xfe_amd.build_indexer(f) for f in fes]
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    with pytest.raises(xfe_amd.XfeError, match="Jacobi"):
        xfe_amd.absorb(V[:, 0], V[:, 1:], [i[0] for i in idx], [i[1] for i in idx], method=3)
    # Schwarz sweeps as Gauss-Seidel, as the reference's device absorber does
    y4, X4, info4 = xfe_amd.absorb(V[:, 0], V[:, 1:], [i[0] for i in idx], [i[1] for i in idx], method=4,
                                   tol=1e-12)
    y1, X1, info1 = xfe_amd.absorb(V[:, 0], V[:, 1:], [i[0] for i in idx], [i[1] for i in idx], method=1,
                                   tol=1e-12)
    assert abs(info4["iterations"] - info1["iterations"]) <= 2
    assert np.abs(y4 - y1).max() <= 1e-10 * np.abs(V).max()


def test_precision_report(cases):
    """precision 1 reports the mixed path; with weights it falls back to FP64
    and says so; the refinement target follows tol."""
    p = cases["akm_three_way"]
    keep = _kept(p)
    fes = [f[keep] for f in p["fes"]]
    idx = [xfe_amd.build_indexer(f) for f in fes]
    gids, G = [i[0] for i in idx], [i[1] for i in idx]
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    _, _, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, precision=1, tol=1e-8)
    assert info["precision_used"] == 1 and not info["precision_fallback"]
    assert info["refine_target"] == pytest.approx(3e-13)
    _, _, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, precision=1, tol=1e-10)
    assert info["refine_target"] == pytest.approx(3e-15)
    _, _, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, precision=1, tol=1e-8,
                                weights=certificate_weights(V.shape[0]))
    assert info["precision_used"] == 0 and info["precision_fallback"]


@pytest.mark.parametrize("name", ["two_way", "akm_three_way"])
def test_ols_cluster_from_device_matches_reference_fit(golden, cases, name):
    """Singletons, indexer, transform (tol 1e-12), Gram and cluster meat on the
    device; beta = Gram solve, V = bread meat bread * G/(G-1) * (N-1)/df_r with
    the estimator's FE degrees of freedom; against the reference's fit
    (coefficients 1e-9, worker-clustered SEs 1e-7)."""
    _log(f"OLS + clustered SE from the device transform: {name}")
    p = cases[name]
    keep = _kept(p)
    np.testing.assert_array_equal(np.flatnonzero(keep), golden[f"{name}/fit/sample_index"])
    fes = [f[keep] for f in p["fes"]]
    idx = [xfe_amd.build_indexer(f) for f in fes]
    gids, G = [i[0] for i in idx], [i[1] for i in idx]
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-12)
    assert info["converged"]
    Vt = np.column_stack([yg, Xg])
    levels = [lp for _, _, lp, _ in idx]
    df_a, _, _ = oracle.fit_dof(fes, levels, p["cluster"][keep])
    omit = oracle.fe_collinear(V[:, 1:], Xg)
    cols = np.concatenate([[0], 1 + np.flatnonzero(~omit)])
    Vk = np.asfortranarray(Vt[:, cols])
    gram = xfe_amd.gram(Vk)
    k = Vk.shape[1] - 1
    xtx, xty = gram[1:, 1:], gram[1:, 0]
    beta = np.linalg.solve(xtx, xty)
    _, cl = np.unique(p["cluster"][keep], return_inverse=True)
    ncl = int(cl.max()) + 1
    meat = xfe_amd.cluster_meat(Vk, beta, cl.astype(np.int32), ncl)
    bread = np.linalg.inv(xtx)
    n = Vk.shape[0]
    Vc = bread @ meat @ bread * (ncl / (ncl - 1.0)) * ((n - 1.0) / (n - k - 1 - df_a))
    coef = np.zeros(V.shape[1] - 1)
    se = np.full(V.shape[1] - 1, np.nan)
    coef[~omit] = beta
    se[~omit] = np.sqrt(np.diag(Vc))
    np.testing.assert_allclose(coef, golden[f"{name}/fit/coef"], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(se, golden[f"{name}/fit/se"], rtol=1e-7)
