"""The wide leverage solver (xfe_akm_solve_S_wide): up to 256 right-hand
sides per operator pass, FP32 PCG under FP64 residual replacement, against
the oracle's FP64 Jacobi-PCG (itself pinned by a dense solve in the CPU suite)
and against the 32-lane FP64 solver, with the lane rules of
akm_cuda_solve_S_multi (src/akm_kss_cuda.cu:532-695)."""
import sys

import numpy as np
import pytest

import oracle
import xfe_amd

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _graph(seed, N, J, movers, weights=True):
    rng = np.random.default_rng(seed)
    mw, mf, mc = [], [], []
    for w in range(N):
        k = 1 + int(rng.random() < movers) + int(rng.random() < movers / 3)
        firms = np.unique(np.minimum(J - 1, (rng.pareto(1.2, size=k) * 3).astype(int)))
        for f in firms:
            mw.append(w); mf.append(int(f)); mc.append(int(rng.integers(1, 6)) if weights else 1)
    for f in range(1, J):  # a chain of movers keeps the firm graph connected
        w = N + f - 1
        mw += [w, w]; mf += [f - 1, f]; mc += [1, 1]
    return oracle.akm_build(mw, mf, mc, N + J - 1, J)


def _residual_ld(st, r, z):
    """|r - S z| with S z = Df z - B' Dw^-1 B z in long double."""
    ld = np.longdouble
    c = st["m_c"].astype(ld)
    zl = z.astype(ld)
    t = np.zeros(st["N"], dtype=ld)
    np.add.at(t, st["m_w"], c * zl[st["m_f"]])
    t /= st["Dw"].astype(ld)
    Sz = st["Df"].astype(ld) * zl
    np.add.at(Sz, st["m_f"], -c * t[st["m_w"]])
    Sz[st["J"] - 1] = 0
    d = r.astype(ld) - Sz
    return float(np.sqrt(np.sum(d * d)))


@pytest.mark.parametrize("nb", [3, 64, 130, 256])
def test_wide_solve_matches_oracle_and_fp64_lanes(nb):
    _log(f"wide solve, {nb} lanes")
    st = _graph(7, N=20000, J=1500, movers=0.3)
    J = st["J"]
    rng = np.random.default_rng(100 + nb)
    rhs = rng.normal(size=(nb, J))
    solver = xfe_amd.AkmSolver(st)
    z, its, (passes, upd) = solver.solve_wide(rhs, tol=1e-10, max_iter=5000)
    assert (its > 0).all() and upd >= 2 and passes >= its.max()
    assert (z[:, J - 1] == 0.0).all()
    # the FP64 32-lane solver on the same lanes (tol 1e-10 each)
    for b0 in range(0, nb, 32):
        zm, itm = solver.solve_multi(rhs[b0:b0 + 32], tol=1e-10, max_iter=5000)
        assert (itm > 0).all()
        for l in range(zm.shape[0]):
            scale = np.abs(zm[l]).max()
            err = np.abs(z[b0 + l] - zm[l]).max() / scale
            assert err <= 1e-8, (b0 + l, err)
            # FP32 recurrences lose orthogonality sooner on this ill-conditioned
            # graph (about 770 FP64 iterations), and each replacement point adds
            # at most a poll interval
            assert its[b0 + l] <= 1.25 * itm[l] + 8 * (upd + 1), (b0 + l, its[b0 + l], itm[l])
    # the oracle's PCG on a few lanes
    for l in sorted({0, nb // 2, nb - 1}):
        zo, ito = oracle.akm_solve_S(st, rhs[l], 1e-10, 5000)
        scale = np.abs(zo).max()
        assert np.abs(z[l] - zo).max() <= 1e-8 * scale
        # the residual of the returned solution meets the stop: evaluated in
        # extended precision (an FP64 application of S through worker means
        # rounds to several 1e-9 here), with 5 % for the device's FP64 rounding
        r = rhs[l].copy(); r[J - 1] = 0.0
        res = _residual_ld(st, r, z[l])
        assert res <= 1.05e-10 * np.linalg.norm(r), (l, res, np.linalg.norm(r))
    solver.close()


def test_wide_lane_rules_and_unit_weights():
    _log("wide lane rules")
    st = _graph(8, N=3000, J=300, movers=0.3, weights=False)
    J = st["J"]
    rng = np.random.default_rng(3)
    rhs = rng.normal(size=(70, J))
    rhs[1] = 0.0                        # zero right-hand side: converged at once, z = 0
    rhs[2] = 0.0; rhs[2, J - 1] = 7.0   # only the ground entry: pinned away
    rhs[5] *= 1e-30                     # tiny and huge scales: unit-norm lanes
    rhs[6] *= 1e+30
    solver = xfe_amd.AkmSolver(st)
    z, its, _ = solver.solve_wide(rhs, tol=1e-10, max_iter=5000)
    assert its[1] == 0 and its[2] == 0 and not z[1].any() and not z[2].any()
    assert (np.delete(its, [1, 2]) > 0).all()
    zm, itm = solver.solve_multi(rhs[[0, 5, 6, 69]], tol=1e-10, max_iter=5000)
    for k, l in enumerate([0, 5, 6, 69]):
        assert np.abs(z[l] - zm[k]).max() <= 1e-8 * np.abs(zm[k]).max()
    z2, its2, _ = solver.solve_wide(rhs[[0, 3]], tol=1e-14, max_iter=2)   # too few iterations
    assert list(its2) == [-1, -1] and not z2.any()
    # determinism: the same lanes give the same bits
    z3, its3, _ = solver.solve_wide(rhs, tol=1e-10, max_iter=5000)
    assert np.array_equal(z3, z) and np.array_equal(its3, its)
    with pytest.raises(xfe_amd.XfeError):
        solver.solve_wide(np.zeros((257, J)))
    # the device entry point: the same bits on HBM-resident lanes
    import torch
    rd = torch.from_numpy(rhs).to("cuda:0")
    zd = torch.empty_like(rd)
    itd, _ = solver.solve_wide_device(rd.data_ptr(), zd.data_ptr(), rhs.shape[0], tol=1e-10, max_iter=5000)
    assert np.array_equal(zd.cpu().numpy(), z) and np.array_equal(itd, its)
    solver.close()


def test_wide_blocks_match_jacobi_with_fewer_iterations(monkeypatch):
    """The aggregate blocks (exactly solved groups of strongly coupled firms)
    change the preconditioner only: the same solutions to the tolerance, in
    fewer iterations than Jacobi on a mover graph with weakly linked clusters."""
    _log("wide blocks vs Jacobi")
    rng = np.random.default_rng(11)
    J, N = 1200, 30000
    mw, mf, mc = [], [], []
    for w in range(N):
        # clusters of 8 firms with most moves inside a cluster
        base = 8 * int(rng.integers(0, J // 8))
        k = 1 + int(rng.random() < 0.4)
        firms = {base + int(rng.integers(0, 8))}
        if k == 2:
            firms.add(base + int(rng.integers(0, 8)) if rng.random() < 0.95 else int(rng.integers(0, J)))
        for f in sorted(firms):
            mw.append(w); mf.append(f); mc.append(int(rng.integers(1, 4)))
    for f in range(1, J):
        w = N + f - 1
        mw += [w, w]; mf += [f - 1, f]; mc += [1, 1]
    st = oracle.akm_build(mw, mf, mc, N + J - 1, J)
    rhs = rng.choice([-1.0, 1.0], size=(40, J))
    blk = xfe_amd.AkmSolver(st)
    zb, itb, _ = blk.solve_wide(rhs, tol=1e-10, max_iter=20000)
    blk.close()
    monkeypatch.setenv("XFE_AKM_JACOBI", "1")
    jac = xfe_amd.AkmSolver(st)
    zj, itj, _ = jac.solve_wide(rhs, tol=1e-10, max_iter=20000)
    jac.close()
    assert (itb > 0).all() and (itj > 0).all()
    scale = np.abs(zj).max(axis=1)
    assert (np.abs(zb - zj).max(axis=1) <= 1e-8 * scale).all()
    _log(f"iterations: blocks {itb.max()}, Jacobi {itj.max()}")
    assert itb.max() < itj.max()


def _boot_panel(seed, workers, years, firms):
    from xfe_amd import panels
    p = panels.config0(seed=seed, workers=workers, years=years, firms=firms)
    worker = p["fes"][0]
    firm = np.unique(p["fes"][1], return_inverse=True)[1].astype(np.int32)
    V = np.column_stack([p["y"], p["X"]])
    return worker, firm, V


def test_boot_refits_match_exact_weighted_projection():
    """Cluster-bootstrap refits (64 per operator pass, two batches, uneven
    resample sizes) against the dense weighted two-way projection: the Gram of
    each refit and its beta."""
    _log("bootstrap refits")
    worker, firm, V = _boot_panel(20261203 + 900, 1500, 6, 60)
    Nw = int(worker.max()) + 1
    rng = np.random.default_rng(4)
    R = 70
    mult = np.stack([rng.multinomial(int(rng.uniform(0.7, 1.3) * Nw), np.ones(Nw) / Nw) for _ in range(R)])
    mult[5] = 1                                        # the original sample
    boot = xfe_amd.BootRefits(worker, firm, V)
    gram, its, (passes, upd), ok = boot.refits(mult, tol=1e-10, max_iter=5000)
    assert ok and (its > 0).all() and upd >= 1
    betas = xfe_amd.BootRefits.betas(gram)
    for r in (0, 5, 31, 63, 64, 69):
        Go = oracle.boot_gram_exact(worker, firm, V, mult[r])
        assert np.abs(gram[r] - Go).max() <= 1e-8 * np.abs(Go).max(), (r, np.abs(gram[r] - Go).max())
        bo = np.linalg.solve(Go[1:, 1:], Go[1:, 0])
        assert np.abs(betas[r] - bo).max() <= 1e-7 * np.abs(bo).max(), (r, betas[r], bo)
    # the original sample: the plain FP64 absorber and OLS give the same beta
    yt, Xt, info = xfe_amd.absorb(V[:, 0], V[:, 1:], [worker.astype(np.int32), firm],
                                  [Nw, int(firm.max()) + 1], tol=1e-12)
    b_abs = np.linalg.lstsq(Xt, yt, rcond=None)[0]
    assert np.abs(betas[5] - b_abs).max() <= 1e-7 * np.abs(b_abs).max()
    # deterministic, and a multiplicity below zero is refused
    gram2, its2, _, _ = boot.refits(mult, tol=1e-10, max_iter=5000)
    assert np.array_equal(gram2, gram) and np.array_equal(its2, its)
    bad = mult[:2].copy()
    bad[0, 3] = -1
    with pytest.raises(xfe_amd.XfeError):
        boot.refits(bad)
    boot.close()
