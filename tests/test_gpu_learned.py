"""The learned preconditioner on an MI355X, through the C ABI.

Network parity: the device's forward pass, loss and gradient against the
oracle's restatement with the same bfloat16 rounding of every matrix-product
operand (oracle/learned.py, whose gradient is pinned by finite differences);
the device accumulates in FP32, the oracle exactly, so the tolerances are
relative: outputs 2e-3 of their scale, loss 1e-4, gradient 5e-3 of each
parameter block's largest entry.
"""
import sys

import numpy as np
import pytest

import xfe_amd
from oracle import learned as L

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("size", [(30, 40), (3000, 5000), (20, 3000)])
def test_gnn_matches_oracle(size):
    F, Wm = size
    _log(f"gnn eval F={F} Wm={Wm}")
    g = L.random_graph(11 + F, F=F, Wm=Wm)
    th = L.init_params(5)
    loss, grad, u, t = xfe_amd.lp_gnn_eval(g, th)
    lo, go, out = L.forward_backward(th.astype(np.float64), g, bf16_products=True)
    _log(f"  loss device {loss:.6g} oracle {lo:.6g}")
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert np.abs(u - out["u"]).max() <= 2e-3 * max(1.0, np.abs(out["u"]).max())
    assert np.abs(t - out["t"]).max() <= 2e-3 * max(1.0, np.abs(out["t"]).max())
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in L.LAYOUT])
    for k, (name, _) in enumerate(L.LAYOUT):
        a, b = grad[offs[k]:offs[k + 1]], go[offs[k]:offs[k + 1]]
        scale = max(np.abs(b).max(), 1e-12)
        err = np.abs(a - b).max() / scale
        assert err <= 5e-3, (name, err)


# ---- the aggregate-block preconditioner and the learned aggregation ----

import oracle  # noqa: E402
from oracle import precond as P  # noqa: E402
from xfe_amd import panels  # noqa: E402


def _akm(scale, p_move, seed):
    p = panels.akm_panel(seed=panels.SEED0 + seed, workers=int(5_950_000 * scale),
                         firms=int(800_000 * scale), p_move=p_move)
    keep, _, _ = xfe_amd.drop_singletons(p["fes"])
    fes = [f[keep] for f in p["fes"]]
    idx = [oracle.build_indexer(f) for f in fes]
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    return V, [i[0] for i in idx], [i[1] for i in idx]


def test_blocks_apply_matches_oracle():
    """Handshake aggregation bit for bit and the block solves (FP32 inverses)
    within 1e-5 of the oracle's, with the coupling strengths and with random
    scores as the pair scores."""
    V, gids, G = _akm(0.003, 0.05, 7)
    pf, pg, W, dex = P.pairs_dense(gids[0], gids[1], Nw=G[0], F=G[1])
    F = G[1]
    rng = np.random.default_rng(1)
    r = rng.random((F, 3)).astype(np.float32)
    for name, score in (("strength", P.strength(pf, pg, W, dex)),
                        ("random", rng.standard_normal(pf.size).astype(np.float32))):
        _log(f"blocks apply: F={F} pairs={pf.size} scores={name}")
        agg, z = xfe_amd.lp_blocks_apply(pf, pg, W, dex, score, r)
        agg_o = P.handshake(pf, pg, score, F)
        assert (agg == agg_o).all()
        z_o = P.blocks_apply(pf, pg, W, dex, agg_o, r)
        blocked = np.bincount(agg_o, minlength=F)[agg_o] >= 2
        assert blocked.sum() > F // 10
        assert np.abs(z - z_o).max() <= 1e-5 * np.abs(z_o).max()


def test_probe_pairs_and_features_match_oracle():
    """The device's firm pairs, couplings, exact diagonal and strengths (from the
    segment records) equal the oracle's sparse Schur formula; the firm
    features and the movers' first features are their closed forms."""
    V, gids, G = _akm(0.003, 0.05, 8)
    _log(f"probe: n={V.shape[0]} G={G}")
    pr = xfe_amd.lp_probe(V[:, 0], V[:, 1:], gids, G, probe_iter=4)
    pf, pg, W, dex = P.pairs_dense(gids[0], gids[1], Nw=G[0], F=G[1])
    assert pr["F"] == G[1]
    assert (pr["pf"] == pf).all() and (pr["pg"] == pg).all()
    assert np.abs(pr["W"] - W).max() <= 1e-12 * W.max()
    assert np.abs(pr["dex"] - dex).max() <= 1e-12 * dex.max()
    st = P.strength(pf, pg, W, dex)
    assert np.abs(pr["pq"] - st).max() <= 1e-6
    rows = np.bincount(gids[1], minlength=G[1]).astype(np.float64)
    xf = pr["xf"]
    assert np.allclose(xf[:, 0], np.log1p(rows), rtol=1e-6)
    assert np.allclose(xf[:, 2], np.log1p(np.maximum(dex, 0)), rtol=1e-6, atol=1e-6)
    pc = np.bincount(pf, minlength=G[1]) + np.bincount(pg, minlength=G[1])
    assert np.allclose(xf[:, 4], np.log1p(pc), rtol=1e-6)
    smax = np.zeros(G[1], np.float32)
    np.maximum.at(smax, pf, st)
    np.maximum.at(smax, pg, st)
    assert np.abs(xf[:, 5] - smax).max() <= 1e-6
    ssum = np.bincount(pf, st, G[1]) + np.bincount(pg, st, G[1])
    assert np.abs(xf[:, 6] - ssum).max() <= 1e-5 * max(1.0, ssum.max())
    # movers: workers at two or more firms; their row counts and firm counts
    B = {}
    for w_, f_ in zip(gids[0], gids[1]):
        B.setdefault(int(w_), set()).add(int(f_))
    nw = np.bincount(gids[0], minlength=G[0])
    movers = sorted((float(np.log(nw[w_])), float(np.log(len(fs)))) for w_, fs in B.items() if len(fs) >= 2)
    got = sorted(zip(pr["xw"][:, 0].astype(float), pr["xw"][:, 2].astype(float)))
    assert len(got) == len(movers) == pr["Wm"]
    assert np.abs(np.array(got) - np.array(movers)).max() <= 1e-5
    assert pr["err"].shape == (G[1], V.shape[1]) and np.isfinite(pr["err"]).all()


@pytest.fixture(scope="module")
def held_out():
    """A held-out panel of the AKM shape at 5 % scale (2.7M rows after
    singletons) with mover rate 0.02, not among the training panels; its
    mixed-precision Jacobi-PCG solution (certified by the FP64 normal-equation
    residual) and the plain FP64 absorber's solution to 1e-14."""
    V, gids, G = _akm(0.05, 0.02, 900)
    y0, X0, i0 = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1, preconditioner=0)
    assert i0["converged"] and i0["precision_certified"]
    yp, Xp, _ = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-14, precision=0, certify=False)
    return V, gids, G, np.column_stack([y0, X0]), np.column_stack([yp, Xp])


def test_learned_preconditioner_gain(held_out):
    """The trained network's aggregation needs fewer inner iterations than the
    mixed-precision method without it (Jacobi-PCG) on the held-out panel, about
    as few as the coupling-strength aggregation.  Each answer is certified by
    the FP64 residual, and the three preconditioners reach the same fixed point
    within 1e-9 of the column scale.  Against the plain FP64 absorber run to
    1e-14 the bound is 1e-6: on this ill-conditioned panel (mover rate 0.02)
    that method stops on the change of a sweep, and at 1e-14 it sits at the
    FP64 noise floor, where the order of its FP64 atomics decides when it stops
    (3e-10 from the certified answer in most runs, 1.9e-7 in one)."""
    V, gids, G, ref, plain = held_out
    scale = np.abs(V).max()
    its = {}
    for pc in (0, 1, 2):
        _log(f"held-out n={V.shape[0]} preconditioner={pc}")
        y, X, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1, preconditioner=pc)
        got = np.column_stack([y, X])
        err = np.abs(got - ref).max() / scale
        err_plain = np.abs(got - plain).max() / scale
        _log(f"  inner {info['iterations']} blocks {info['precond_blocks']} err {err:.2e} "
             f"(plain FP64 {err_plain:.2e}) cert {info['abs_residual_rel']:.2e}")
        assert info["precision_used"] == 1 and info["converged"] and info["precision_certified"]
        assert info["preconditioner_used"] == pc and not info["precond_fallback"]
        assert err <= 1e-9
        assert err_plain <= 1e-6
        its[pc] = info["iterations"]
    assert its[2] * 2 < its[0]
    assert its[2] <= 1.25 * its[1] + 8


def test_learned_preconditioner_fallback(held_out, monkeypatch):
    """When the block preconditioner stalls (declared by the host here), the
    solve continues with Jacobi-PCG and reaches the same answer."""
    V, gids, G, ref, _ = held_out
    monkeypatch.setenv("XFE_LP_STALL_TEST", "1")
    _log("held-out fallback")
    y, X, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1, preconditioner=2)
    assert info["precond_fallback"] and info["converged"] and info["precision_certified"]
    assert np.abs(np.column_stack([y, X]) - ref).max() / np.abs(V).max() <= 1e-9


def test_learned_preconditioner_rejects_missing_params():
    V, gids, G = _akm(0.001, 0.05, 9)
    with pytest.raises(xfe_amd.XfeError):
        xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, precision=1, preconditioner=2,
                       lp_params=np.zeros(3, np.float32))
