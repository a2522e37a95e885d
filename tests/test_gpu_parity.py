"""Parity of the HIP path with the oracle, through the C ABI, on an MI355X.

Tolerances: integer results (singleton masks, component labels, DoF) are
compared exactly; demeaned columns within 1e-8 relative to the column scale
(max |v~_gpu - v~_oracle| / max |v|) at the default tolerance 1e-8 and within
1e-10 at tol 1e-12; Gram / meat within 1e-12 relative.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
import xfe_amd
from xfe_amd import panels

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _prep(p):
    keep, _, _ = oracle.drop_singletons(p["fes"])
    fes = [f[keep] for f in p["fes"]]
    idx = [oracle.build_indexer(f) for f in fes]
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    return keep, fes, [i[0] for i in idx], [i[1] for i in idx], V


def _rel(a, b, scale):
    return np.abs(a - b).max() / scale


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
@pytest.mark.parametrize("tol", [1e-8, 1e-12])
def test_absorb_matches_oracle(cases, name, tol):
    _log(f"absorb {name} tol={tol}")
    p = cases[name]
    _, fes, gids, G, V = _prep(p)
    Vo, info_o = oracle.absorb(V, gids, G, tol=tol)
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=tol)
    assert info["converged"]
    scale = np.abs(V).max()
    err = max(_rel(yg, Vo[:, 0], scale), _rel(Xg, Vo[:, 1:], scale))
    assert err <= (1e-8 if tol == 1e-8 else 1e-10), (err, info, info_o)


def test_absorb_config0_full(golden):
    _log("absorb config0 200k")
    p = panels.config0()
    _, fes, gids, G, V = _prep(p)
    Vo, info_o = oracle.absorb(V, gids, G, tol=1e-8)
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8)
    scale = np.abs(V).max()
    err = max(_rel(yg, Vo[:, 0], scale), _rel(Xg, Vo[:, 1:], scale))
    _log(f"  gpu iters {info['iterations']} oracle iters {info_o['iterations']} err {err:.3e}")
    assert info["converged"] and err <= 1e-8


def test_absorb_akm_scaled_three_way():
    _log("absorb akm 1% scale")
    p = panels.akm_small(scale=0.01)
    _, fes, gids, G, V = _prep(p)
    Vo, info_o = oracle.absorb(V, gids, G, tol=1e-8)
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8)
    scale = np.abs(V).max()
    err = max(_rel(yg, Vo[:, 0], scale), _rel(Xg, Vo[:, 1:], scale))
    _log(f"  gpu iters {info['iterations']} oracle iters {info_o['iterations']} err {err:.3e}")
    assert info["converged"] and err <= 1e-8


def test_absorb_weighted():
    _log("absorb weighted")
    p = panels.config0(seed=11, workers=800, years=6, firms=90)
    _, fes, gids, G, V = _prep(p)
    w = np.random.default_rng(2).uniform(0.5, 2.0, V.shape[0])
    Vo, _ = oracle.absorb(V, gids, G, weights=w, tol=1e-12)
    ws = [np.bincount(g, weights=w, minlength=Gd) for g, Gd in zip(gids, G)]
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, weight_sums=ws, weights=w, tol=1e-12)
    scale = np.abs(V).max()
    assert max(_rel(yg, Vo[:, 0], scale), _rel(Xg, Vo[:, 1:], scale)) <= 1e-10


def test_absorb_cg_path():
    _log("absorb forced CG")
    p = panels.config0(seed=12, workers=700, years=5, firms=60)
    _, fes, gids, G, V = _prep(p)
    Vo, info_o = oracle.absorb(V, gids, G, tol=1e-12, force_cg=True)
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-12, force_cg=True)
    assert info["used_cg"] and info_o["used_cg"]
    scale = np.abs(V).max()
    assert max(_rel(yg, Vo[:, 0], scale), _rel(Xg, Vo[:, 1:], scale)) <= 1e-9


def test_absorb_plain_map_path():
    _log("absorb plain MAP (no acceleration)")
    rng = np.random.default_rng(13)
    n = 5000
    a = rng.integers(0, 50, n).astype(np.int32)   # 100 obs per level -> no acceleration
    b = rng.integers(0, 40, n).astype(np.int32)
    V = rng.standard_normal((n, 3))
    Vo, info_o = oracle.absorb(V, [a, b], [50, 40], tol=1e-10)
    assert not info_o["used_accel"]
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], [a, b], [50, 40], tol=1e-10)
    assert not info["used_accel"] and info["converged"]
    assert info["iterations"] == info_o["iterations"]
    assert max(_rel(yg, Vo[:, 0], 1.0), _rel(Xg, Vo[:, 1:], 1.0)) <= 1e-10


def test_absorb_many_columns():
    _log("absorb 20 columns (chunked passes)")
    p = panels.config0(seed=14, workers=500, years=8, firms=70)
    _, fes, gids, G, V = _prep(p)
    rng = np.random.default_rng(3)
    V = np.column_stack([V, rng.standard_normal((V.shape[0], 16))])
    Vo, _ = oracle.absorb(V, gids, G, tol=1e-12)
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-12)
    scale = np.abs(V).max()
    assert max(_rel(yg, Vo[:, 0], scale), _rel(Xg, Vo[:, 1:], scale)) <= 1e-10


def test_singletons_exact(cases):
    _log("singletons")
    rng = np.random.default_rng(21)
    for trial in range(20):
        n = int(rng.integers(1, 400))
        D = int(rng.integers(1, 4))
        fes = [rng.integers(-3, int(rng.integers(2, 200)), n).astype(np.int32) for _ in range(D)]
        k1, d1, r1 = oracle.drop_singletons(fes)
        k2, d2, r2 = xfe_amd.drop_singletons(fes)
        np.testing.assert_array_equal(k1, k2)
        assert d1 == d2 and r1 == r2
    for p in list(cases.values()) + [panels.config0()]:
        k1, d1, r1 = oracle.drop_singletons(p["fes"])
        k2, d2, r2 = xfe_amd.drop_singletons(p["fes"])
        np.testing.assert_array_equal(k1, k2)
        assert (d1, r1) == (d2, r2)
    # round cap: a chain that needs many rounds stops at the same place
    n = 300
    a = np.arange(n, dtype=np.int32) // 2
    b = (np.arange(n, dtype=np.int32) + 1) // 2
    for cap in (1, 5, 100):
        k1, d1, r1 = oracle.drop_singletons([a, b], max_rounds=cap)
        k2, d2, r2 = xfe_amd.drop_singletons([a, b], max_rounds=cap)
        np.testing.assert_array_equal(k1, k2)
        assert (d1, r1) == (d2, r2)


def test_components_exact(cases):
    _log("components")
    rng = np.random.default_rng(22)
    for trial in range(20):
        n = int(rng.integers(1, 500))
        a = (rng.integers(0, int(rng.integers(1, 80)), n) * int(rng.integers(1, 4))).astype(np.int32)
        b = rng.integers(-30, int(rng.integers(-28, 60)), n).astype(np.int32)
        la = np.unique(a).size
        for exp in (-1, la):
            c1, l1 = oracle.components(a, b, exp, -1)
            c2, l2 = xfe_amd.components(a, b, exp, -1)
            assert c1 == c2
            np.testing.assert_array_equal(l1, l2)
    for p in list(cases.values()) + [panels.config0()]:
        keep, _, _ = oracle.drop_singletons(p["fes"])
        a, b = p["fes"][0][keep], p["fes"][1][keep]
        la, lb = np.unique(a).size, np.unique(b).size
        c1, l1 = oracle.components(a, b, la, lb)
        c2, l2 = xfe_amd.components(a, b, la, lb)
        assert c1 == c2
        np.testing.assert_array_equal(l1, l2)


def test_dof_exact(cases):
    _log("dof")
    for p in cases.values():
        keep, _, _ = oracle.drop_singletons(p["fes"])
        fes = [f[keep] for f in p["fes"]]
        levels = [np.unique(f).size for f in fes]
        df1, r1, c1 = oracle.dof_reghdfe(fes, levels)
        df2, r2, c2 = xfe_amd.fe_dof(fes, levels)
        assert df1 == df2
        np.testing.assert_array_equal(r1, r2)
        np.testing.assert_array_equal(c1, c2)
    # coarsening: an FE that is a function of a finer one
    rng = np.random.default_rng(23)
    fine = rng.integers(0, 60, 2000).astype(np.int32)
    coarse = (fine // 7).astype(np.int32)
    other = rng.integers(0, 30, 2000).astype(np.int32)
    fes = [other, coarse, fine]
    levels = [np.unique(f).size for f in fes]
    assert oracle.dof_reghdfe(fes, levels)[0] == xfe_amd.fe_dof(fes, levels)[0]


def test_gram_and_cluster_meat():
    _log("gram / meat")
    rng = np.random.default_rng(24)
    n, C = 10000, 11
    V = rng.standard_normal((n, C))
    G = xfe_amd.gram(V)
    ref = V.T @ V
    assert np.abs(G - ref).max() <= 1e-12 * np.abs(ref).max()
    cl = rng.integers(0, 300, n).astype(np.int32)
    beta = rng.standard_normal(C - 1)
    M = xfe_amd.cluster_meat(V, beta, cl, 300)
    e = V[:, 0] - V[:, 1:] @ beta
    S = np.zeros((300, C - 1))
    np.add.at(S, cl, V[:, 1:] * e[:, None])
    ref = S.T @ S
    assert np.abs(M - ref).max() <= 1e-11 * np.abs(ref).max()


def test_recovery_max_delta():
    _log("recovery max delta")
    rng = np.random.default_rng(25)
    n = 3000
    a = rng.integers(0, 40, n).astype(np.int32)
    b = rng.integers(0, 25, n).astype(np.int32)
    r = rng.standard_normal(n)
    # one GS pass: max over dims of max |mean| at the time the dim is projected
    v = r.copy()
    mx = 0.0
    for g, Gd in ((a, 40), (b, 25)):
        m = np.bincount(g, weights=v, minlength=Gd) / np.bincount(g, minlength=Gd)
        mx = max(mx, np.abs(m).max())
        v = v - m[g]
    fe = (xfe_amd.FeInput * 2)()
    for d, (g, Gd) in enumerate(((a, 40), (b, 25))):
        fe[d].group_ids = g.ctypes.data
        fe[d].num_groups = Gd
    import ctypes
    out = ctypes.c_double()
    ok = xfe_amd.lib().xfe_recovery_max_delta(r.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(n),
                                             fe, ctypes.c_int(2), None, None, ctypes.byref(out))
    assert ok == 1
    assert abs(out.value - mx) <= 1e-13 * max(1.0, mx)


def test_device_entry_point_matches_host_entry_point():
    _log("device entry point")
    torch = pytest.importorskip("torch")
    p = panels.config0(seed=15, workers=1000, years=7, firms=100)
    _, fes, gids, G, V = _prep(p)
    yg, Xg, _ = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-10)
    n, C = V.shape
    ld = ((n + 63) // 64) * 64
    dV = torch.zeros((C, ld), dtype=torch.float64, device="cuda")
    dV[:, :n] = torch.from_numpy(np.ascontiguousarray(V.T)).cuda()
    ids = torch.from_numpy(np.ascontiguousarray(np.stack(gids))).cuda()
    torch.cuda.synchronize()
    info = xfe_amd.absorb_device(dV.data_ptr(), n, C, ld, ids.data_ptr(), G, tol=1e-10)
    out = dV[:, :n].cpu().numpy().T
    scale = np.abs(V).max()
    assert info["converged"]
    assert max(_rel(out[:, 0], yg, scale), _rel(out[:, 1:], Xg, scale)) <= 1e-12


@pytest.mark.parametrize("precision", [0, 1])
def test_device_entry_point_waits_for_queued_work(precision):
    """The device entry point runs on the library's own stream; a caller that
    fills d_V asynchronously (torch's stream) and calls at once must still get
    the transform of the filled data."""
    _log(f"device entry point ordering, precision {precision}")
    torch = pytest.importorskip("torch")
    p = panels.akm_small(scale=0.02)
    _, fes, gids, G, V = _prep(p)
    yg, Xg, _ = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=precision)
    n, C = V.shape
    ld = ((n + 63) // 64) * 64
    V0 = torch.zeros((C, ld), dtype=torch.float64, device="cuda")
    V0[:, :n] = torch.from_numpy(np.ascontiguousarray(V.T)).cuda()
    ids = torch.from_numpy(np.ascontiguousarray(np.stack(gids))).cuda()
    dV = torch.full_like(V0, 1.0e6)
    torch.cuda.synchronize()
    big = torch.empty(64 << 20, dtype=torch.float64, device="cuda")
    for _ in range(8):            # a backlog of ~4 GB of copies on torch's stream
        big.copy_(big.flip(0))
    dV.copy_(V0)                  # no synchronize before the call
    info = xfe_amd.absorb_device(dV.data_ptr(), n, C, ld, ids.data_ptr(), G, tol=1e-8, precision=precision)
    out = dV[:, :n].cpu().numpy().T
    scale = np.abs(V).max()
    assert info["converged"]
    # the two entry points run the same method on the same data (FP32 inner
    # solves may round differently between runs)
    assert max(_rel(out[:, 0], yg, scale), _rel(out[:, 1:], Xg, scale)) <= (1e-12 if precision == 0 else 1e-9)


@pytest.mark.parametrize("precision", [0, 1])
def test_device_entry_point_in_place_and_copied(precision):
    """Rows already sorted by the dimension with the most levels and the
    padded stride: the solve works in the caller's block itself; with a wider
    stride it works in a copy, and with the rows shuffled in a sorted
    permutation.  All three equal the host entry point on the same rows."""
    _log(f"device entry point in place / copied, precision {precision}")
    torch = pytest.importorskip("torch")
    p = panels.akm_small(scale=0.02)
    _, fes, gids, G, V = _prep(p)
    pdim = int(np.argmax(G))
    srt = np.argsort(gids[pdim], kind="stable")
    rng = np.random.default_rng(3)
    n, C = V.shape
    for label, rows, extra in (("in place", srt, 0), ("copied", srt, 64), ("permuted", rng.permutation(n), 0)):
        Vr = V[rows]
        gr = [g[rows] for g in gids]
        yg, Xg, _ = xfe_amd.absorb(Vr[:, 0], Vr[:, 1:], gr, G, tol=1e-8, precision=precision)
        ld = ((n + 63) // 64) * 64 + extra
        dV = torch.zeros((C, ld), dtype=torch.float64, device="cuda")
        dV[:, :n] = torch.from_numpy(np.ascontiguousarray(Vr.T)).cuda()
        ids = torch.from_numpy(np.ascontiguousarray(np.stack(gr))).cuda()
        torch.cuda.synchronize()
        info = xfe_amd.absorb_device(dV.data_ptr(), n, C, ld, ids.data_ptr(), G, tol=1e-8, precision=precision)
        out = dV[:, :n].cpu().numpy().T
        scale = np.abs(V).max()
        err = max(_rel(out[:, 0], yg, scale), _rel(out[:, 1:], Xg, scale))
        _log(f"  {label}: {err:.2e}")
        assert info["converged"]
        assert err <= (1e-12 if precision == 0 else 1e-9), label


# ---------------------------------------------------------------- mixed precision
# The FP32 Schur-complement CG under FP64 refinement converges to the
# projection itself; it is compared with the oracle run to tol 1e-12 (error
# bound: 1e-9 of the column scale) -- a tighter fixed point than either method
# reaches at the default tolerance.

@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_mixed_precision_matches_tight_oracle(cases, name):
    _log(f"mixed precision {name}")
    p = cases[name]
    _, fes, gids, G, V = _prep(p)
    Vo, _ = oracle.absorb(V, gids, G, tol=1e-12)
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1)
    scale = np.abs(V).max()
    err = max(_rel(yg, Vo[:, 0], scale), _rel(Xg, Vo[:, 1:], scale))
    _log(f"  inner iterations {info['iterations']} refinements {info['refinements']} err {err:.2e}")
    assert info["converged"] and info["refinements"] >= 1
    assert err <= 1e-9


@pytest.mark.parametrize("name", ["two_way", "akm_three_way"])
def test_fp64_schur_matches_tight_oracle_and_mixed(cases, name):
    """precision 2 (the Schur-complement Jacobi-PCG entirely in FP64, the
    mixed path's FP64 ablation): certified, within 1e-9 of the tight oracle,
    and within 1e-9 of the mixed path run with the same (Jacobi)
    preconditioner."""
    _log(f"fp64 schur {name}")
    p = cases[name]
    _, fes, gids, G, V = _prep(p)
    Vo, _ = oracle.absorb(V, gids, G, tol=1e-12)
    y2, X2, i2 = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=2, preconditioner=0)
    y1, X1, i1 = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1, preconditioner=0)
    scale = np.abs(V).max()
    err = max(_rel(y2, Vo[:, 0], scale), _rel(X2, Vo[:, 1:], scale))
    err12 = max(_rel(y2, y1, scale), _rel(X2, X1, scale))
    _log(f"  FP64 iterations {i2['iterations']} (mixed {i1['iterations']}) err {err:.2e} vs mixed {err12:.2e}")
    assert i2["converged"] and i2["precision_used"] == 2 and i2["precision_certified"], i2
    assert err <= 1e-9 and err12 <= 1e-9


def test_mixed_precision_graph_replays_match_eager(monkeypatch):
    """The mixed path's inner iterations replayed as HIP graphs (eight per
    graph, captured per solve, the FP64 replacements eager between the
    replays) give the iteration and replacement counts of the same launches
    issued one by one (XFE_NO_GRAPH) and the same result to 1e-12 of the
    column scale (two eager solves already differ in the last bits, about
    4e-14: the per-tile FP64 LDS sums of the row passes add in arrival
    order), on three solves in a row with a plain FP64 absorption between."""
    _log("mixed precision graph replays")
    p = panels.akm_small(scale=0.01)
    _, fes, gids, G, V = _prep(p)
    monkeypatch.setenv("XFE_NO_GRAPH", "1")
    y0, X0, i0 = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1)
    monkeypatch.delenv("XFE_NO_GRAPH")
    assert i0["converged"] and i0["precision_used"] == 1
    for k in range(3):
        y, X, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1)
        _log(f"  solve {k}: iterations {info['iterations']} (eager {i0['iterations']})")
        assert info["iterations"] == i0["iterations"] and info["refinements"] == i0["refinements"]
        scale = np.abs(V).max()
        assert max(_rel(y, y0, scale), _rel(X, X0, scale)) <= 1e-12
        xfe_amd.absorb(V[:, 0], V[:, 1:2], gids, G, tol=1e-6, precision=0)


@pytest.mark.parametrize("which", ["config0", "akm_1pct"])
def test_mixed_precision_spell_path(which):
    _log(f"mixed precision {which}")
    p = panels.config0() if which == "config0" else panels.akm_small(scale=0.01)
    _, fes, gids, G, V = _prep(p)
    Vo, _ = oracle.absorb(V, gids, G, tol=1e-12)
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1)
    scale = np.abs(V).max()
    err = max(_rel(yg, Vo[:, 0], scale), _rel(Xg, Vo[:, 1:], scale))
    _log(f"  inner iterations {info['iterations']} refinements {info['refinements']} err {err:.2e}")
    assert info["converged"]
    assert err <= 1e-9


# ---------------------------------------------------------------- certificate

def _rel_close(got, want, rtol):
    # values at the rounding floor (about 1e-17 relative) carry no digits
    return abs(got - want) <= rtol * want + 64 * np.finfo(np.float64).eps


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_certify_matches_oracle(cases, name):
    from golden_cases import certificate_weights
    _log(f"certify {name}")
    p = cases[name]
    _, fes, gids, G, V = _prep(p)
    Vo, info = oracle.absorb(V, gids, G, tol=1e-8)
    w = certificate_weights(V.shape[0])
    for Vt, weights, rtol in ((V, None, 1e-11), (Vo, None, 1e-5), (Vo, w, 1e-11)):
        g = xfe_amd.certify(V[:, 0], V[:, 1:], Vt[:, 0], Vt[:, 1:], gids, G, weights=weights, tol=1e-8)
        o = oracle.certify(V, Vt, gids, G, weights=weights, tol=1e-8)
        assert _rel_close(g["abs_residual"], o["abs_residual"], rtol), (g, o)
        assert _rel_close(g["abs_residual_rel"], o["abs_residual_rel"], rtol), (g, o)
        assert g["certified"] == o["certified"]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", ["two_way", "akm_three_way"])
def test_absorb_reports_certificate(cases, name, precision):
    _log(f"absorb certificate {name} precision {precision}")
    p = cases[name]
    _, fes, gids, G, V = _prep(p)
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=precision)
    assert info["certificate_computed"] and info["precision_certified"]
    o = oracle.certify(V, np.column_stack([yg, Xg]), gids, G, tol=1e-8)
    assert o["certified"]
    assert _rel_close(info["abs_residual_rel"], o["abs_residual_rel"], 1e-5), (info, o)


# ---------------------------------------------------------------- AKM/KSS solver

def _akm_graph(seed, N, J, movers):
    rng = np.random.default_rng(seed)
    mw, mf, mc = [], [], []
    for w in range(N):
        k = 1 + int(rng.random() < movers) + int(rng.random() < movers / 3)
        firms = np.unique(np.minimum(J - 1, (rng.pareto(1.2, size=k) * 3).astype(int)))
        for f in firms:
            mw.append(w); mf.append(int(f)); mc.append(int(rng.integers(1, 6)))
    # a chain of movers between consecutive firms keeps the firm graph connected
    # (the reference solves on the leave-out connected set)
    for f in range(1, J):
        w = N + f - 1
        mw += [w, w]; mf += [f - 1, f]; mc += [1, 1]
    N = N + J - 1
    return oracle.akm_build(mw, mf, mc, N, J)


@pytest.mark.parametrize("nb", [1, 9, 32])
def test_akm_solve_matches_oracle(nb):
    _log(f"akm solve, {nb} lanes")
    st = _akm_graph(7, N=20000, J=1500, movers=0.3)
    J = st["J"]
    rng = np.random.default_rng(nb)
    rhs = rng.normal(size=(nb, J))
    solver = xfe_amd.AkmSolver(st)
    z, its = solver.solve_multi(rhs, tol=1e-10, max_iter=5000)
    for l in range(nb):
        zo, ito = oracle.akm_solve_S(st, rhs[l], 1e-10, 5000)
        assert ito > 0 and its[l] > 0
        assert abs(int(its[l]) - ito) <= 2, (l, its[l], ito)
        scale = np.abs(zo).max()
        assert np.abs(z[l] - zo).max() <= 1e-8 * scale, (l, np.abs(z[l] - zo).max() / scale)
        assert z[l, J - 1] == 0.0
    if nb == 1:
        z1, it1 = solver.solve(rhs[0], tol=1e-10, max_iter=5000)
        assert it1 == its[0] and np.array_equal(z1, z[0])
    solver.close()


def test_akm_lane_rules():
    _log("akm lane rules")
    st = _akm_graph(8, N=3000, J=300, movers=0.3)
    J = st["J"]
    rng = np.random.default_rng(3)
    rhs = rng.normal(size=(4, J))
    rhs[1] = 0.0                       # zero right-hand side: converged at once, z = 0
    rhs[2] = 0.0; rhs[2, J - 1] = 7.0  # only the ground entry: pinned away
    solver = xfe_amd.AkmSolver(st)
    z, its = solver.solve_multi(rhs, tol=1e-10, max_iter=5000)
    assert its[0] > 0 and its[1] == 0 and its[2] == 0 and its[3] > 0
    assert not z[1].any() and not z[2].any()
    z, its = solver.solve_multi(rhs[[0, 3]], tol=1e-14, max_iter=2)   # too few iterations
    assert list(its) == [-1, -1] and not z.any()
    assert solver.solve(rhs[0], tol=1e-14, max_iter=2)[1] == -1
    solver.close()


def test_akm_stable_sort_u64():
    _log("akm stable sort")
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 1 << 62, size=100000, dtype=np.uint64)
    keys[::7] = keys[3]                # many equal keys: input order must be kept
    keys[-5:] = np.uint64(2**64 - 1)
    np.testing.assert_array_equal(xfe_amd.akm_stable_sort_u64(keys), np.argsort(keys, kind="stable"))


def test_akm_create_rejects_bad_input():
    st = _akm_graph(10, N=200, J=30, movers=0.3)
    bad = dict(st)
    bad["m_c"] = st["m_c"].copy()
    bad["m_c"][0] = 1.5                # counts must be integers
    with pytest.raises(xfe_amd.XfeError):
        xfe_amd.AkmSolver(bad)


# ------------------------------------------------------------ schwarz_demean_raw
# xfe_schwarz_demean returns the projection (mixed Schur PCG under FP64
# refinement) once the reference's stopping rule, evaluated on the result,
# holds for every column; it is compared with the oracle at tol 1e-14 and with
# the reference's own output (golden fixture).

@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
@pytest.mark.parametrize("tag", ["tight", "default"])
def test_schwarz_matches_reference(golden, cases, name, tag):
    _log(f"schwarz {name} {tag}")
    _, fes, gids, G, V = _prep(cases[name])
    tol = 1e-12 if tag == "tight" else 1e-8
    y = V[:, 0].copy()
    X = np.asfortranarray(V[:, 1:])
    ok, it = xfe_amd.schwarz_demean(y, X, gids, G, tol=tol)
    assert ok and it > 0, xfe_amd.last_error()
    Vg = np.column_stack([y, X])
    assert oracle.schwarz_rule(V, Vg, gids, G).max() < tol
    scale = np.abs(V).max()
    Vo, _ = oracle.absorb(V, gids, G, tol=1e-14)
    Vr = np.column_stack([golden[f"{name}/schwarz_{tag}/y_tilde"], golden[f"{name}/schwarz_{tag}/X_tilde"]])
    err_o = np.abs(Vg - Vo).max() / scale
    err_r = np.abs(Vg - Vr).max() / scale
    _log(f"  iterations {it} vs oracle {err_o:.2e} vs reference {err_r:.2e}")
    # the rule bounds the normal-equation residual; the error itself may be
    # larger by the conditioning of the panel (the module's 1e-10 at tol 1e-12)
    assert err_o <= (1e-10 if tag == "tight" else 1e-8)
    assert err_r <= (1e-10 if tag == "tight" else 1e-7)


def test_schwarz_tight_tolerance(cases):
    """A tolerance below the default refinement target takes the FP64 polish sweeps."""
    _log("schwarz tol 1e-13")
    _, fes, gids, G, V = _prep(cases["akm_three_way"])
    y = V[:, 0].copy()
    X = np.asfortranarray(V[:, 1:])
    ok, _ = xfe_amd.schwarz_demean(y, X, gids, G, tol=1e-13)
    assert ok, xfe_amd.last_error()
    assert oracle.schwarz_rule(V, np.column_stack([y, X]), gids, G).max() < 1e-13


def test_schwarz_edge_cases(golden):
    """The reference's gates and edge cases (golden flags, same inputs).  Where
    the reference fails on an empty level of a third dimension (its
    preconditioner divides by the level's zero count) this path returns the
    projection; iteration counts are this solver's own."""
    from golden_cases import SCHWARZ_EDGE_NAMES, schwarz_edge_cases
    ref_flags = golden["schwarz_edges"]
    for j, ((y0, X0, ids, lev), mi) in enumerate(schwarz_edge_cases()):
        name = SCHWARZ_EDGE_NAMES[j]
        _log(f"schwarz edge {name}")
        y = y0.copy()
        X = np.asfortranarray(X0.copy())
        ok, it = xfe_amd.schwarz_demean(y, X, ids, lev, tol=1e-10, max_iter=mi)
        want = True if name == "empty_level_extra" else bool(ref_flags[j][0])
        assert ok == want, (name, xfe_amd.last_error())
        if not ok:
            assert np.array_equal(y, y0) and np.array_equal(X, X0), name
            continue
        V = np.column_stack([y0, X0])
        Vg = np.column_stack([y, X])
        if name == "zero_columns":
            assert it == 0 and np.array_equal(Vg, V)
            continue
        assert oracle.schwarz_rule(V, Vg, ids, lev).max() < 1e-10, name
        Vo, _ = oracle.absorb(V, ids, lev, tol=1e-14)
        assert np.abs(Vg - Vo).max() / np.abs(V).max() <= 1e-9, name
        if name == "one_zero_column":
            assert np.array_equal(X[:, 1], X0[:, 1])


# ------------------------------------------------------------ heterogeneous slopes
# The slope path runs FP64 CG on the symmetric projection sweep until a sweep
# would change the transform by at most tol_eff max(1, ||v||); it is compared
# with the exact projection (oracle) and with the reference's absorber
# (fixtures at tol 1e-12, themselves 9e-9 of the column scale from the
# projection on these slowly converging panels).

@pytest.mark.parametrize("name,sdim,icpt", [("two_way", 1, True), ("two_way", 1, False),
                                             ("akm_three_way", 1, True)])
def test_absorb_slopes_matches_oracle(golden, cases, name, sdim, icpt):
    from golden_cases import slope_values
    _log(f"slopes {name} dim {sdim} intercept {icpt}")
    _, fes, gids, G, V = _prep(cases[name])
    z = slope_values(V.shape[0])
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-12, slopes={sdim: (z, icpt)})
    assert info["converged"], info
    Vg = np.column_stack([yg, Xg])
    Vo = oracle.slope_projection(V, gids, G, sdim, z, intercept=icpt)
    tag = f"slope{sdim}_{'icpt' if icpt else 'noicpt'}"
    Vr = np.column_stack([golden[f"{name}/{tag}/y_tilde"], golden[f"{name}/{tag}/X_tilde"]])
    scale = np.abs(V).max()
    err_o = np.abs(Vg - Vo).max() / scale
    err_r = np.abs(Vg - Vr).max() / scale
    _log(f"  sweeps {info['iterations']} vs oracle {err_o:.2e} vs reference {err_r:.2e}")
    assert err_o <= 1e-8 and err_r <= 2e-8


def test_absorb_slope_single_dimension_weighted():
    """One weighted slope dimension is exact after one CG step: each level's
    weighted least-squares line of v on z."""
    _log("slopes single dimension weighted")
    rng = np.random.default_rng(11)
    n = 900
    g = rng.integers(0, 30, n).astype(np.int32)
    z = rng.normal(size=n)
    w = rng.uniform(0.5, 2.0, n)
    V = rng.normal(size=(n, 3))
    for icpt in (True, False):
        yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], [g], [30], weights=w, tol=1e-12,
                                      slopes={0: (z, icpt)})
        Vo = oracle.slope_projection(V, [g], [30], 0, z, intercept=icpt, weights=w)
        assert info["converged"] and info["iterations"] <= 2
        assert np.abs(np.column_stack([yg, Xg]) - Vo).max() <= 1e-12


def test_absorb_slope_requires_values():
    rng = np.random.default_rng(1)
    g = rng.integers(0, 5, 50).astype(np.int32)
    fes = (xfe_amd.FeInput * 1)()
    fes[0].group_ids = g.ctypes.data
    fes[0].num_groups = 5
    fes[0].num_levels_present = 5
    fes[0].is_slope = 1
    y = rng.normal(size=50)
    yo = np.empty(50)
    o = xfe_amd._opts(1e-8, 1, 1000, 1, False, 0)
    r = xfe_amd.AbsorbResult()
    ok = xfe_amd.lib().xfe_absorb(xfe_amd._ptr(y), None, 50, 0, fes, 1, None, None,
                                  ctypes.byref(o), ctypes.byref(r), xfe_amd._ptr(yo), None)
    assert ok == 0 and "slope_values" in xfe_amd.last_error()


# ------------------------------------------------------------ JLA leverages

def test_akm_jla_pii_matches_reference(golden):
    """xfe_akm_jla_pii with the reference's streams against the reference's
    P_ii (50 draws, seed 42; the solves stop at 1e-12 here, the reference's
    direct solve is exact) and against the oracle at 200 draws."""
    from golden_cases import JLA_DRAWS, JLA_SEED
    _log("akm jla pii")
    rw, rf, rc = golden["jla/row_worker"], golden["jla/row_firm"], golden["jla/row_weight"]
    wu, wi = np.unique(rw, return_inverse=True)
    fu, fi = np.unique(rf, return_inverse=True)
    st = oracle.akm_build(wi, fi, rc, wu.size, fu.size)
    solver = xfe_amd.AkmSolver(st)
    pii, ok = solver.jla_pii(JLA_DRAWS, JLA_SEED, tol=1e-12, max_iter=5000)
    assert ok, xfe_amd.last_error()
    err = np.abs(pii - golden["jla/pii"]).max()
    _log(f"  vs reference {err:.2e}")
    assert err <= 1e-8
    pii2, ok = solver.jla_pii(200, 9, tol=1e-12, max_iter=5000)
    assert ok and np.abs(pii2 - oracle.akm_jla_pii(st, 200, 9)).max() <= 1e-8
    solver.close()


def test_akm_exact_leverages_match_reference():
    """xfe_akm_exact_leverages (one solve per match, 256 per wide solve)
    against the reference's exact path (LeverageMethod::Exact, fixture): P_ii,
    and the B_ii weights through the components they enter (AGSU: sum B;
    KSS: B . sigma_i); against the oracle's exact leverages per match; sum
    P_ii = N + J - 1."""
    import os
    _log("akm exact leverages")
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "exact_leverages.npz"))
    wu, wi = np.unique(fx["row_worker"], return_inverse=True)
    fu, fi = np.unique(fx["row_firm"], return_inverse=True)
    st = oracle.akm_build(wi, fi, fx["row_weight"], wu.size, fu.size)
    solver = xfe_amd.AkmSolver(st)
    P, Bfe, Bpe, Bcov, failed = solver.exact_leverages(tol=1e-11, max_iter=5000)
    solver.close()
    assert failed == 0, xfe_amd.last_error()
    err = np.abs(P - fx["pii"]).max()
    _log(f"  P_ii vs reference {err:.2e}; sum {P.sum():.6f} (N + J - 1 = {st['N'] + st['J'] - 1})")
    assert err <= 1e-8
    assert abs(P.sum() - (st["N"] + st["J"] - 1)) <= 1e-6
    comp, sig = fx["components"], fx["sigma_i"]
    dof = comp[10] - 1.0
    for i, B in enumerate((Bpe, Bfe, Bcov)):
        want_sum = (comp[i] - comp[3 + i]) * dof / comp[9]
        want_sig = (comp[i] - comp[6 + i]) * dof
        assert abs(B.sum() - want_sum) <= 1e-8 * abs(want_sum)
        assert abs(B @ sig - want_sig) <= 1e-8 * abs(want_sig)
    want = oracle.akm_exact_leverages(st)
    for got, w in zip((P, Bfe, Bpe, Bcov), want):
        assert np.abs(got - w).max() <= 1e-8 * max(1.0, np.abs(w).max())


def test_kss_components_exact_match_reference():
    """The plug-in, AGSU and KSS variance components with exact leverages
    (the device's two-way fit and exact leverages, the O(rows) sums on the
    host) against the reference's akm_kss_decompose with LeverageMethod::Exact
    on the same panel (fixture): every component and sigma2_ho to 1e-8
    relative (the corrected ones relative to their plug-in), sigma_i per match
    to 1e-8 of its scale."""
    import os
    from golden_cases import jla_panel
    from xfe_amd import akm
    _log("kss components (exact)")
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "exact_leverages.npz"))
    jp = jla_panel()
    r = akm.kss_components(jp["y"], jp["worker"], jp["firm"], method="exact")
    comp = fx["components"]
    got = list(r["plugin"]) + list(r["agsu"]) + list(r["kss"]) + [r["sigma2_ho"], r["n_py"]]
    # a corrected component is the plug-in less a correction of nearly its
    # size: its error is measured against the plug-in's magnitude
    scale = [abs(w) for w in comp]
    for i in range(3, 9):
        scale[i] = max(scale[i], abs(comp[i % 3]))
    err = max(abs(g - w) / s for g, w, s in zip(got, comp, scale))
    _log(f"  components vs reference: max rel {err:.2e}")
    assert err <= 1e-8, (got, comp)
    assert r["unit_leverage"] == 0
    # sigma_i in the reference's working-row order (its rows are matches)
    key = {(int(w), int(f)): k for k, (w, f) in enumerate(zip(r["match_worker"], r["match_firm"]))}
    order = np.array([key[(int(w), int(f))] for w, f in zip(fx["row_worker"], fx["row_firm"])])
    sig = r["sigma_i"][order]
    assert np.abs(sig - fx["sigma_i"]).max() <= 1e-8 * np.abs(fx["sigma_i"]).max()
    assert np.abs(r["pii"][order] - fx["pii"]).max() <= 1e-8


def test_kss_components_jla_match_reference(golden):
    """The JLA decomposition (50 draws, seed 42: the reference's streams on
    its working-row order) with the device's fit, leverages, non-linearity
    correction and B_ii draws against the reference's akm_kss_decompose with
    LeverageMethod::Jla on the same panel (fixture): the 11 numbers to 1e-8
    (corrected ones relative to their plug-in), P_ii and sigma_i to 1e-8; the
    correction against the oracle's."""
    from golden_cases import JLA_DRAWS, JLA_SEED, jla_panel
    from xfe_amd import akm
    _log("kss components (jla)")
    jp = jla_panel()
    r = akm.kss_components(jp["y"], jp["worker"], jp["firm"], method="jla", draws=JLA_DRAWS, seed=JLA_SEED,
                           tol=1e-12, max_iter=5000)
    np.testing.assert_array_equal(r["match_worker"], golden["jla/row_worker"])
    np.testing.assert_array_equal(r["match_firm"], golden["jla/row_firm"])
    comp = golden["jla/components"]
    got = list(r["plugin"]) + list(r["agsu"]) + list(r["kss"]) + [r["sigma2_ho"], r["n_py"]]
    scale = [abs(w) for w in comp]
    for i in range(3, 9):
        scale[i] = max(scale[i], abs(comp[i % 3]))
    err = max(abs(g - w) / s for g, w, s in zip(got, comp, scale))
    _log(f"  components vs reference: max rel {err:.2e}")
    assert err <= 1e-8, (got, comp)
    assert np.abs(r["pii"] - golden["jla/pii"]).max() <= 1e-8
    sig = golden["jla/sigma_i"]
    assert np.abs(r["sigma_i"] - sig).max() <= 1e-8 * np.abs(sig).max()
    wu, fu, wi, fi, ridx, mw, mf, c = akm.match_panel(jp["worker"], jp["firm"])
    st = oracle.akm_build(mw, mf, c, wu.size, fu.size)
    _, corr = oracle.akm_jla_moments(st, JLA_DRAWS, JLA_SEED)
    assert np.abs(r["corr"] - corr).max() <= 1e-8


def test_akm_jla_bii_matches_reference(golden):
    """xfe_akm_jla_bii (50 draws, seed 42: batches of 16 + 16 + 16 + 2)
    against the oracle's weights per row, and against the reference: its AGSU
    components are plug-in - sigma2_ho sum(B) / dof and its KSS components
    plug-in - sum(B sigma_i) / dof.  Solves stop at 1e-12 (the reference's
    and the oracle's are direct), so the bound is 1e-8 relative."""
    from golden_cases import JLA_DRAWS, JLA_SEED
    _log("akm jla bii")
    rw, rf, rc = golden["jla/row_worker"], golden["jla/row_firm"], golden["jla/row_weight"]
    wu, wi = np.unique(rw, return_inverse=True)
    fu, fi = np.unique(rf, return_inverse=True)
    st = oracle.akm_build(wi, fi, rc, wu.size, fu.size)
    cnt = rc.astype(np.int64)
    solver = xfe_amd.AkmSolver(st)
    bfe, bpe, bcov, ok = solver.jla_bii(cnt, JLA_DRAWS, JLA_SEED, tol=1e-12, max_iter=5000)
    assert ok, xfe_amd.last_error()
    want = oracle.akm_jla_bii(st, JLA_DRAWS, JLA_SEED, cnt)
    for got, w in zip((bfe, bpe, bcov), want):
        err = np.abs(got - w).max() / np.abs(w).max()
        _log(f"  vs oracle {err:.2e}")
        assert err <= 1e-8
    comp, sig = golden["jla/components"], golden["jla/sigma_i"]
    dof = comp[10] - 1.0
    for i, B in enumerate((bpe, bfe, bcov)):
        want_sum = (comp[i] - comp[3 + i]) * dof / comp[9]
        want_sig = (comp[i] - comp[6 + i]) * dof
        assert abs(B.sum() - want_sum) <= 1e-8 * abs(want_sum)
        assert abs(B @ sig - want_sig) <= 1e-8 * abs(want_sig)
    # default counts (the integral match weights), a ragged last batch
    b2 = solver.jla_bii(None, 21, 5, tol=1e-12, max_iter=5000)
    assert b2[3]
    for got, w in zip(b2[:3], oracle.akm_jla_bii(st, 21, 5, cnt)):
        assert np.abs(got - w).max() <= 1e-8 * np.abs(w).max()
    solver.close()


# ---- savefe FE recovery (xfe_recover_fixed_effects)

@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_recover_fixed_effects_matches_reference(golden, cases, name):
    """The device recovery against the reference's own recovery and the
    oracle: the per-row sum of the FE values (unique, unlike the values) within
    1e-8 of the column scale of the reference's tight recovery and of the
    exact fitted component; contributions expand the level values; max_delta
    is the oracle's check of the same residual."""
    _log(f"recover {name}")
    _, fes, gids, G, V = _prep(cases[name])
    y = V[:, 0]
    alphas, contrib, info = xfe_amd.recover_fixed_effects(y, gids, G, fe_tol=1e-6)
    assert info["converged"] and info["absorb_converged"], (info, xfe_amd.last_error())
    for d in range(len(gids)):
        np.testing.assert_array_equal(contrib[d], alphas[d][gids[d]])
    tot = contrib.sum(0)
    scale = np.abs(y).max()
    ref = golden[f"{name}/recover_tight/contributions"]
    assert _rel(tot, ref.sum(0), scale) < 1e-8
    Vo, _ = oracle.absorb(V, gids, G, tol=1e-14)
    assert _rel(tot, y - Vo[:, 0], scale) < 1e-8
    md = oracle.fe_recovery_max_delta(y - tot, gids, G)
    assert info["max_delta"] < 1e-8
    assert abs(info["max_delta"] - md) <= 1e-12 + 1e-6 * md, (info["max_delta"], md)


def test_recover_fixed_effects_akm_scaled():
    """A 1 % AKM-shape panel (three dimensions): the sum of the recovered FE
    values is the exact fitted component, and the check pass is tiny."""
    _log("recover akm 1%")
    p = panels.akm_small(scale=0.01)
    _, fes, gids, G, V = _prep(p)
    y = V[:, 0]
    alphas, contrib, info = xfe_amd.recover_fixed_effects(y, gids, G)
    assert info["converged"], info
    Vo, _ = oracle.absorb(V[:, :1], gids, G, tol=1e-13)
    scale = np.abs(y).max()
    assert _rel(contrib.sum(0), y - Vo[:, 0], scale) < 1e-8
    assert info["max_delta"] < 1e-8


def test_recover_fixed_effects_one_dimension_weighted():
    """One weighted dimension: the values are the weighted group means of
    partial (closed form)."""
    _log("recover one dim weighted")
    rng = np.random.default_rng(11)
    n, G = 20000, 300
    g = rng.integers(0, G, n).astype(np.int32)
    w = 0.5 + rng.random(n)
    y = rng.standard_normal(n)
    alphas, contrib, info = xfe_amd.recover_fixed_effects(y, [g], [G], weights=w)
    assert info["converged"]
    exact = np.bincount(g, w * y, G) / np.bincount(g, w, G)
    np.testing.assert_allclose(alphas[0], exact, rtol=0, atol=1e-13)
    assert info["max_delta"] < 1e-13


def test_recover_fixed_effects_rejects_weighted_multi_dim():
    rng = np.random.default_rng(12)
    n = 1000
    a = rng.integers(0, 50, n).astype(np.int32)
    b = rng.integers(0, 20, n).astype(np.int32)
    with pytest.raises(xfe_amd.XfeError):
        xfe_amd.recover_fixed_effects(rng.standard_normal(n), [a, b], [50, 20], weights=np.ones(n))


# ---- group()/individual() absorption (xfe_absorb_group_individual)

@pytest.mark.parametrize("tag", ["gs", "sgs_w"])
def test_group_individual_matches_reference(golden, tag):
    """The device group()/individual() absorber: certified, within 1e-7 of the
    column scale of the reference's CPU absorber (both stop certified at
    8 tol_eff), within 1e-8 of the exact weighted least squares projection,
    and its certificate equals the oracle's on its output."""
    from golden_cases import gi_panel, gi_weights
    _log(f"group_individual {tag}")
    p = gi_panel()
    V = np.column_stack([p["y"], p["X"]])
    w = gi_weights(V.shape[0]) if tag == "sgs_w" else None
    method = 2 if tag == "sgs_w" else 1
    yg, Xg, info = xfe_amd.absorb_group_individual(p["y"], p["X"], p["fes"], p["G"], p["gi"], weights=w,
                                                   method=method)
    assert info["converged"] and info["precision_certified"], info
    Vg = np.column_stack([yg, Xg])
    scale = np.abs(V).max()
    ref = np.column_stack([golden[f"gi/{tag}/y_tilde"], golden[f"gi/{tag}/X_tilde"]])
    assert _rel(Vg, ref, scale) < 1e-7
    ex = oracle.gi_exact(V, p["fes"], p["G"], p["gi"], weights=w)
    assert _rel(Vg, ex, scale) < 1e-8
    a, rel, cert = oracle.certify_group_individual(V, Vg, p["fes"], p["G"], p["gi"], weights=w)
    # both at the rounding floor here (~1e-18, the decision limit is 8e-9):
    # long double on the device host side, double in the oracle
    assert cert and abs(rel - info["abs_residual_rel"]) <= 1e-6 * rel + 1e-15
    # same colouring, same sweeps: the oracle's restatement needs the same count
    Vo, it, _ = oracle.absorb_group_individual(V, p["fes"], p["G"], p["gi"], weights=w, method=method)
    assert abs(it - info["iterations"]) <= 1, (it, info["iterations"])
    assert _rel(Vg, Vo, scale) < 1e-10


def test_group_individual_only_individuals():
    """No standard dimension: the individual projection alone (exact lstsq)."""
    from golden_cases import gi_panel
    _log("group_individual no standard fes")
    p = gi_panel(seed=77, n=2000, individuals=500)
    V = np.column_stack([p["y"], p["X"]])
    yg, Xg, info = xfe_amd.absorb_group_individual(p["y"], p["X"], [], [], p["gi"])
    assert info["converged"]
    ex = oracle.gi_exact(V, [], [], p["gi"])
    assert _rel(np.column_stack([yg, Xg]), ex, np.abs(V).max()) < 1e-8


def test_group_individual_rejects_bad_structure():
    from golden_cases import gi_panel
    p = gi_panel(seed=78, n=500, individuals=100)
    gi = dict(p["gi"])
    gi["group_individual"] = gi["group_individual"].copy()
    gi["group_individual"][3] = gi["num_individuals"]  # out of range
    with pytest.raises(xfe_amd.XfeError):
        xfe_amd.absorb_group_individual(p["y"], p["X"], p["fes"], p["G"], gi)


# ---- shadow device certificate (xfe_cert_shadow_*)

def _dev(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to("cuda")


def _shadow(V0, Vr, gids, G, L, w=None):
    import torch
    y0, X0 = _dev(V0[:, 0]), _dev(V0[:, 1:].T)
    yr, Xr = _dev(Vr[:, 0]), _dev(Vr[:, 1:].T)
    g = [_dev(x.astype(np.int32)) for x in gids]
    wd = None if w is None else _dev(w)
    out = xfe_amd.cert_shadow(y0, X0, yr, Xr, g, G, L, weights=wd)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_cert_shadow_matches_oracle(golden, cases, name):
    """The device shadow certificate against the oracle's decision on the
    reference's converged output (PASS at the tight limit) and on the raw
    columns (FAIL; to-nearest rel within 1e-11 of the reference's own
    certificate value); counts and weight sums consistent."""
    _log(f"cert shadow {name}")
    _, fes, gids, G, V0 = _prep(cases[name])
    Vt = np.column_stack([golden[f"{name}/tight/y_tilde"], golden[f"{name}/tight/X_tilde"]])
    conv = _shadow(V0, Vt, gids, G, 8e-12)
    assert conv["verdict"] == "PASS" and conv["host_like_pass"] == 1, conv
    # the reference's dense indexer keeps id gaps: empty levels are counted, not dropped
    empty = sum(int((np.bincount(g, minlength=Gd) == 0).sum()) for g, Gd in zip(gids, G))
    assert conv["wsum_ok"] == 1 and conv["empty_groups"] == empty and conv["rhs"] == V0.shape[1]
    raw = _shadow(V0, V0, gids, G, 8e-9)
    assert raw["verdict"] == "FAIL" and raw["host_like_pass"] == 0
    ref_rel = golden[f"{name}/raw/cert"][1]
    assert abs(raw["rel_max"] - ref_rel) <= 1e-11 * ref_rel
    o = oracle.cert_shadow(V0, V0, gids, G, 8e-9)
    assert abs(raw["rel_max"] - max(o["rel"])) <= 1e-12 * raw["rel_max"]


def test_cert_shadow_bracket_and_straddle(golden, cases):
    """One column: the device bracket contains the exact rational Res_sq;
    at L = rel the verdict is INDETERMINATE, 1e-6 either side decides."""
    _log("cert shadow straddle")
    _, fes, gids, G, V0 = _prep(cases["sparse_two_way"])
    y = V0[:, :1]
    R, A, B = oracle.cert_shadow_exact(y, y, gids, G)
    r = _shadow(y, y, gids, G, 1.0)
    assert r["r_lower"] <= R[0] <= r["r_upper"]
    rel = r["rel_max"]
    assert _shadow(y, y, gids, G, rel)["verdict"] == "INDETERMINATE"
    assert _shadow(y, y, gids, G, rel * (1 + 1e-6))["verdict"] == "PASS"
    assert _shadow(y, y, gids, G, rel * (1 - 1e-6))["verdict"] == "FAIL"


def test_cert_shadow_weighted_matches_oracle(golden, cases):
    from golden_cases import certificate_weights
    _log("cert shadow weighted")
    _, fes, gids, G, V0 = _prep(cases["two_way"])
    w = certificate_weights(V0.shape[0])
    Vr = V0 * 1e-3
    d = _shadow(V0, Vr, gids, G, 8e-9, w=w)
    o = oracle.cert_shadow(V0, Vr, gids, G, 8e-9, weights=w)
    assert d["verdict"] == ("FAIL" if "FAIL" in o["verdict"] else o["verdict"][0])
    ref = max(o["rel"])
    # the oracle's rel uses sum w^2 per dimension as the device's
    assert abs(d["rel_max"] - ref) <= 1e-11 * ref


def test_cert_shadow_rejects_bad_ids():
    _log("cert shadow bad ids")
    n = 1000
    rng = np.random.default_rng(3)
    V = rng.standard_normal((n, 2))
    g = rng.integers(0, 10, n).astype(np.int32)
    g[5] = 10  # outside [0, 10): found on the device, the handle is not created
    with pytest.raises(xfe_amd.XfeError):
        _shadow(V, V, [g], [10], 1e-8)


def test_cert_shadow_akm_scaled():
    """A 1 % AKM panel (worker and firm tables beyond the LDS copies, the year
    table inside them): raw columns FAIL with the oracle's to-nearest value
    (1e-12), the device absorber's converged output PASSES."""
    _log("cert shadow akm 1%")
    p = panels.akm_small(scale=0.01)
    _, fes, gids, G, V0 = _prep(p)
    raw = _shadow(V0, V0, gids, G, 8e-9)
    o = oracle.cert_shadow(V0, V0, gids, G, 8e-9)
    assert raw["verdict"] == "FAIL"
    assert abs(raw["rel_max"] - max(o["rel"])) <= 1e-12 * raw["rel_max"]
    yg, Xg, info = xfe_amd.absorb(V0[:, 0], V0[:, 1:], gids, G, tol=1e-8)
    conv = _shadow(V0, np.column_stack([yg, Xg]), gids, G, 8e-9)
    assert conv["verdict"] == "PASS" and conv["host_like_pass"] == 1, conv
    assert o["verdict"] == ["FAIL"] * V0.shape[1]


# ---- articulation points (xfe_articulation_points)

def _random_graph_list(seed, count, nmax):
    rng = np.random.default_rng(seed)
    for t in range(count):
        n = int(rng.integers(1, nmax))
        kind = t % 4
        if kind == 0:
            raw = rng.integers(0, n, (int(rng.integers(0, 2 * n + 1)), 2))
            edges = {tuple(sorted(map(int, e))) for e in raw if e[0] != e[1]}
        elif kind == 1:
            edges = {(i, i + 1) for i in range(n - 1)}
            edges |= {tuple(sorted(map(int, e))) for e in rng.integers(0, n, (int(rng.integers(0, 3)), 2)) if e[0] != e[1]}
        elif kind == 2:
            edges = {(int(rng.integers(0, i)), i) for i in range(1, n)}
        else:
            h = max(n // 2, 1)
            edges = {tuple(sorted(map(int, e))) for e in rng.integers(0, h, (n, 2)) if e[0] != e[1]}
            edges |= {tuple(sorted(map(int, e))) for e in rng.integers(h, n, (n, 2)) if e[0] != e[1]} if n > h else set()
            if n > 2:
                edges.add((0, n - 1))
        yield n, sorted(edges)


def test_articulation_points_random_graphs():
    """The device cut-vertex search equals the oracle's depth-first search
    (itself pinned by brute force) on 400 small graphs, bit for bit."""
    _log("articulation random graphs")
    for n, edges in _random_graph_list(11, 400, 40):
        ptr, adj = oracle.graph_csr(n, edges)
        got = xfe_amd.articulation_points(ptr, adj)
        want = oracle.articulation_points(ptr, adj)
        assert (got == want).all(), (n, edges)


def test_articulation_points_path_and_isolated():
    """A 3,000-vertex path (every inner vertex cut, 2,999 BFS levels from the
    root), isolated vertices and a cycle (no cut vertex)."""
    _log("articulation path")
    n = 3000
    ptr, adj = oracle.graph_csr(n + 5, [(i, i + 1) for i in range(n - 1)] + [(n + 2, n + 3), (n + 3, n + 4),
                                                                             (n + 4, n + 2)])
    got = xfe_amd.articulation_points(ptr, adj)
    want = np.zeros(n + 5, dtype=np.uint8)
    want[1:n - 1] = 1
    np.testing.assert_array_equal(got, want)


def test_articulation_points_mover_graph():
    """The mover-firm graph of an AKM panel (workers with two or more firms
    and their firms, the graph build_leave_out_sample prunes): the device
    equals the oracle."""
    _log("articulation mover graph")
    p = panels.akm_panel(workers=40000, firms=6000, years=20, p_move=0.04)
    w, f = p["fes"][0], p["fes"][1]
    key = np.unique(w.astype(np.int64) * 6000 + f)
    mw, mf = key // 6000, key % 6000
    deg = np.bincount(mw, minlength=40000)
    keep = deg[mw] >= 2
    mw, mf = mw[keep], mf[keep]
    _, wi = np.unique(mw, return_inverse=True)
    fu, fi = np.unique(mf, return_inverse=True)
    nm = wi.max() + 1
    ptr, adj = oracle.graph_csr(nm + fu.size, np.column_stack([wi, nm + fi]))
    got = xfe_amd.articulation_points(ptr, adj)
    want = oracle.articulation_points(ptr, adj)
    _log(f"  vertices {ptr.size - 1}, edges {adj.size // 2}, cut vertices {int(want.sum())}")
    assert want.sum() > 0
    np.testing.assert_array_equal(got, want)


def test_articulation_points_rejects_bad_graph():
    ptr = np.array([0, 1, 2], dtype=np.int32)
    adj = np.array([1, 2], dtype=np.int32)  # vertex 2 does not exist
    with pytest.raises(xfe_amd.XfeError):
        xfe_amd.articulation_points(ptr, adj)


_LOS_KEYS = ("n_obs_input", "n_obs_connected", "n_obs", "n_workers", "n_firms", "n_matches", "n_movers",
             "n_stayers", "prune_iterations")


@pytest.mark.parametrize("name", ["low_mobility", "moderate", "gapped_ids"])
def test_leave_out_set_matches_reference(golden, name):
    """xfe_akm_leave_out_set equals the reference's leave_out_connected_set
    (fixtures): the same rows, counts and pruning rounds."""
    from golden_cases import leave_out_panels
    _log(f"leave-out set {name}")
    w, f = leave_out_panels()[name]
    keep, info = xfe_amd.leave_out_set(w, f)
    np.testing.assert_array_equal(keep, golden[f"los/{name}/keep"].astype(bool))
    assert [info[k] for k in _LOS_KEYS] == golden[f"los/{name}/info"].tolist()
    assert info["gpu_used"] == 1


def test_leave_out_set_kss_scaled():
    """A 5 % KSS-shape panel (60k workers, 7.5k firms, 0.5M rows, three
    pruning rounds): the device result equals the oracle's, also with
    frequency weights."""
    _log("leave-out set kss 5%")
    p = panels.kss_panel(workers=60000, firms=7500)
    keep, info = xfe_amd.leave_out_set(p["worker"], p["firm"])
    k2, i2 = oracle.leave_out_set(p["worker"], p["firm"])
    np.testing.assert_array_equal(keep, k2)
    assert [info[k] for k in _LOS_KEYS] == [i2[k] for k in _LOS_KEYS]
    _log(f"  rows {info['n_obs']} of {info['n_obs_input']}, rounds {info['prune_iterations']}")
    # frequency weights: a worker's weighted count decides the once-observed
    # filter, as in the reference
    fw = 1 + (np.arange(p["worker"].size) % 3).astype(np.int64)
    k3, i3 = xfe_amd.leave_out_set(p["worker"], p["firm"], fweights=fw)
    k4, i4 = oracle.leave_out_set(p["worker"], p["firm"], fweights=fw)
    np.testing.assert_array_equal(k3, k4)
    assert [i3[k] for k in _LOS_KEYS] == [i4[k] for k in _LOS_KEYS]
