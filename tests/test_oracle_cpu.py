"""The oracle pinned against the reference's outputs, closed forms and brute force.

CPU only.  ref_cases.npz holds the reference CPU core's outputs on seeded
panels (written by tests/golden/make_golden.py).
"""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from golden_cases import input_checksum


def _compact(fes):
    out = [oracle.build_indexer(f) for f in fes]
    return [o[0] for o in out], [o[1] for o in out]


def _kept(case):
    keep, _, _ = oracle.drop_singletons(case["fes"])
    return keep


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_inputs_regenerate(golden, cases, name):
    assert input_checksum(cases[name]) == golden[f"{name}/input_checksum"]


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_singletons_match_reference(golden, cases, name):
    keep = _kept(cases[name])
    idx = np.flatnonzero(keep)
    np.testing.assert_array_equal(idx, golden[f"{name}/xfe/sample_index"])
    np.testing.assert_array_equal(idx, golden[f"{name}/fit/sample_index"])
    assert (~keep).sum() == golden[f"{name}/fit/num_singletons"]


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
@pytest.mark.parametrize("tag", ["tight", "default"])
def test_absorb_matches_reference(golden, cases, name, tag):
    p = cases[name]
    keep = _kept(p)
    fes = [f[keep] for f in p["fes"]]
    gids, G = _compact(fes)
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    tol = 1e-12 if tag == "tight" else 1e-8
    Vo, info = oracle.absorb(V, gids, G, tol=tol)
    ref_y = golden[f"{name}/{tag}/y_tilde"]
    ref_X = golden[f"{name}/{tag}/X_tilde"]
    scale = np.abs(V).max()
    err = max(np.abs(Vo[:, 0] - ref_y).max(), np.abs(Vo[:, 1:] - ref_X).max()) / scale
    assert err <= (1e-11 if tag == "tight" else 1e-9), err
    assert info["converged"]
    # same algorithm, same trajectory: iteration counts agree with the reference
    assert info["iterations"] == int(golden[f"{name}/{tag}/iterations"])


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_absorb_matches_reference_xfe(golden, cases, name):
    """Against the estimator's own partial_out (which may take its exact sparse
    solver on small panels): the within transform agrees to the solver tolerance."""
    p = cases[name]
    keep = _kept(p)
    fes = [f[keep] for f in p["fes"]]
    gids, G = _compact(fes)
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    Vo, _ = oracle.absorb(V, gids, G, tol=1e-12)
    ref = np.column_stack([golden[f"{name}/xfe/y_tilde"], golden[f"{name}/xfe/X_tilde"]])
    assert np.abs(Vo - ref).max() / np.abs(V).max() < 1e-9


def _direct_residual(V, fes):
    """Least-squares residual of V on the FE dummies by a sparse direct solve
    of the grounded normal equations (closed form of the within transform)."""
    from scipy.sparse.csgraph import connected_components
    n = V.shape[0]
    _, ia = np.unique(fes[0], return_inverse=True)
    _, ib = np.unique(fes[1], return_inverse=True)
    Ga, Gb = ia.max() + 1, ib.max() + 1
    Da = sp.csr_matrix((np.ones(n), (np.arange(n), ia)), shape=(n, Ga))
    Db = sp.csr_matrix((np.ones(n), (np.arange(n), ib)), shape=(n, Gb))
    adj = sp.bmat([[None, Da.T @ Db], [Db.T @ Da, None]])
    _, comp = connected_components(adj, directed=False)
    # ground one b-level per mobility group (the dummies are rank deficient by
    # exactly one per group)
    first_b = {}
    for j in range(Gb):
        first_b.setdefault(comp[Ga + j], j)
    keep_b = np.setdiff1d(np.arange(Gb), np.array(list(first_b.values())))
    Dm = sp.hstack([Da, Db[:, keep_b]]).tocsc()
    A = (Dm.T @ Dm).tocsc()
    out = np.empty_like(V)
    lu = spla.splu(A)
    for c in range(V.shape[1]):
        coef = lu.solve(Dm.T @ V[:, c])
        out[:, c] = V[:, c] - Dm @ coef
    return out


@pytest.mark.parametrize("name", ["two_way", "sparse_two_way"])
def test_absorb_matches_direct_solve(cases, name):
    p = cases[name]
    keep = _kept(p)
    fes = [f[keep] for f in p["fes"]]
    gids, G = _compact(fes)
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    Vo, _ = oracle.absorb(V, gids, G, tol=1e-13, tolerance_mode=2)
    ref = _direct_residual(V, fes)
    assert np.abs(Vo - ref).max() / np.abs(V).max() < 1e-9


def test_one_way_is_group_demeaning():
    rng = np.random.default_rng(1)
    g = rng.integers(0, 50, 3000).astype(np.int32)
    V = rng.standard_normal((3000, 3))
    out = oracle.project(V, g, 50)
    means = np.zeros((50, 3))
    np.add.at(means, g, V)
    means /= np.bincount(g, minlength=50)[:, None]
    np.testing.assert_allclose(out, V - means[g], atol=1e-14)


def test_absorb_invariants(cases):
    p = cases["akm_three_way"]
    keep = _kept(p)
    fes = [f[keep] for f in p["fes"]]
    gids, G = _compact(fes)
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    Vo, _ = oracle.absorb(V, gids, G, tol=1e-12)
    for g, Gd in zip(gids, G):   # D' v~ = 0 for every dimension
        s = np.zeros((Gd, V.shape[1]))
        np.add.at(s, g, Vo)
        assert np.abs(s).max() < 1e-8 * np.abs(V).max()
    V2, _ = oracle.absorb(Vo, gids, G, tol=1e-12)  # idempotence
    assert np.abs(V2 - Vo).max() < 1e-9 * np.abs(V).max()


def _brute_singletons(fes, max_rounds=100):
    keep = np.ones(fes[0].size, bool)
    for _ in range(max_rounds):
        drop = np.zeros_like(keep)
        for f in fes:
            vals, cnt = np.unique(f[keep], return_counts=True)
            single = set(vals[cnt == 1].tolist())
            drop |= keep & np.isin(f, list(single))
        if not drop.any():
            break
        keep &= ~drop
    return keep


def test_singletons_brute_force():
    rng = np.random.default_rng(3)
    for trial in range(30):
        n = int(rng.integers(1, 60))
        D = int(rng.integers(1, 4))
        fes = [rng.integers(-5, int(rng.integers(2, 30)), n).astype(np.int32) for _ in range(D)]
        keep, dropped, _ = oracle.drop_singletons(fes)
        np.testing.assert_array_equal(keep, _brute_singletons(fes))
        assert dropped == (~keep).sum()


def _bfs_labels(a, b):
    """Components of the bipartite graph; label = rank of the component's first
    a-node in first-appearance order of a's values (DenseMap/Sparse numbering)."""
    order_a = list(dict.fromkeys(a.tolist()))
    node_a = {v: i for i, v in enumerate(order_a)}
    adj = {}
    for x, y in zip(a.tolist(), b.tolist()):
        adj.setdefault(("a", x), set()).add(("b", y))
        adj.setdefault(("b", y), set()).add(("a", x))
    comp = {}
    c = 0
    for v in order_a:
        if ("a", v) in comp:
            continue
        c += 1
        stack = [("a", v)]
        comp[("a", v)] = c
        while stack:
            u = stack.pop()
            for w in adj[u]:
                if w not in comp:
                    comp[w] = c
                    stack.append(w)
    return c, np.array([comp[("a", x)] for x in a.tolist()], dtype=np.int32)


def test_components_bfs():
    rng = np.random.default_rng(4)
    for trial in range(30):
        n = int(rng.integers(1, 80))
        a = rng.integers(0, int(rng.integers(1, 40)), n).astype(np.int32) * 3 + 7  # sparse ids
        b = rng.integers(-20, int(rng.integers(-18, 30)), n).astype(np.int32)
        c, lab = oracle.components(a, b)
        c2, lab2 = _bfs_labels(a, b)
        assert c == c2
        np.testing.assert_array_equal(lab, lab2)


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_groupvar_and_dof_match_reference(golden, cases, name):
    p = cases[name]
    keep = _kept(p)
    fes = [f[keep] for f in p["fes"]]
    levels = [np.unique(f).size for f in fes]
    c, lab = oracle.components(fes[0], fes[1], levels[0], levels[1])
    np.testing.assert_array_equal(lab, golden[f"{name}/fit/groupvar"])
    df, red, nc = oracle.dof_reghdfe(fes, levels)
    assert red[1] == c or len(fes) > 2
    df_a, red_fit, _ = oracle.fit_dof(fes, levels, p["cluster"][keep])
    np.testing.assert_array_equal(red_fit, golden[f"{name}/fit/fe_redundant"])
    assert df_a == golden[f"{name}/fit/df_a"]


def _dense_rank_dof(fes):
    n = fes[0].size
    cols = []
    for f in fes:
        _, inv = np.unique(f, return_inverse=True)
        D = np.zeros((n, inv.max() + 1))
        D[np.arange(n), inv] = 1.0
        cols.append(D)
    return np.linalg.matrix_rank(np.hstack(cols))


def test_dof_two_way_is_exact_rank():
    rng = np.random.default_rng(5)
    for trial in range(20):
        n = int(rng.integers(5, 120))
        fes = [rng.integers(0, int(rng.integers(2, 25)), n).astype(np.int32) for _ in range(2)]
        levels = [np.unique(f).size for f in fes]
        df, red, nc = oracle.dof_reghdfe(fes, levels)
        # for two FEs the mobility-group count is exact: rank = levels - components
        assert df == _dense_rank_dof(fes) + (1 if red[1] == 0 else 0) or df == _dense_rank_dof(fes)


def test_ols_cluster_matches_reference(golden, cases):
    for name in ("two_way", "akm_three_way"):
        p = cases[name]
        keep = _kept(p)
        fes = [f[keep] for f in p["fes"]]
        gids, G = _compact(fes)
        V = np.column_stack([p["y"][keep], p["X"][keep]])
        Vo, _ = oracle.absorb(V, gids, G, tol=1e-12)
        levels = [np.unique(f).size for f in fes]
        df_a, _, _ = oracle.fit_dof(fes, levels, p["cluster"][keep])
        omit = oracle.fe_collinear(V[:, 1:], Vo[:, 1:])
        beta, se, _, _ = oracle.ols_cluster(Vo[:, 0], Vo[:, 1:], p["cluster"][keep], df_a, omit)
        np.testing.assert_allclose(beta, golden[f"{name}/fit/coef"], rtol=1e-9)
        np.testing.assert_allclose(se, golden[f"{name}/fit/se"], rtol=1e-7)


# ---------------------------------------------------------------- certificate
# certify_absorption_result (src/fe_absorption.cpp:14942-15190): the oracle's
# restatement against the reference's own certificate of its own outputs
# (fixtures), with and without row weights, and against a dense ||D'W v||.

@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
@pytest.mark.parametrize("tag", ["tight", "default", "raw"])
def test_certificate_matches_reference(golden, cases, name, tag):
    from golden_cases import certificate_weights
    p = cases[name]
    keep = _kept(p)
    fes = [f[keep] for f in p["fes"]]
    gids, G = _compact(fes)
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    if tag == "raw":
        Vt, tol, variants = V, 1e-8, (("cert", None),)
    else:
        Vt = np.column_stack([golden[f"{name}/{tag}/y_tilde"], golden[f"{name}/{tag}/X_tilde"]])
        tol = 1e-12 if tag == "tight" else 1e-8
        variants = (("cert", None), ("cert_w", certificate_weights(V.shape[0])))
    for key, w in variants:
        ref_abs, ref_rel, ref_cert = golden[f"{name}/{tag}/{key}"]
        o = oracle.certify(V, Vt, gids, G, weights=w, tol=tol, converged=True)
        # converged outputs: the certificate sums cancel to a small residue, so
        # summation order shows at ~1e-7; large residuals agree to rounding
        rtol = 1e-5 if (tag != "raw" and key == "cert") else 1e-11
        assert abs(o["abs_residual"] - ref_abs) <= rtol * ref_abs, (key, o, ref_abs)
        assert abs(o["abs_residual_rel"] - ref_rel) <= rtol * ref_rel, (key, o, ref_rel)
        assert o["certified"] == bool(ref_cert)


def test_certificate_dense_formula():
    """||D'W v~_c|| with a dense D (one indicator block per dimension),
    ||D'W||_F and the untransformed norm, on a tiny weighted case."""
    rng = np.random.default_rng(11)
    n, G = 60, [7, 5, 3]
    gids = [rng.integers(0, g, n) for g in G]
    V = rng.normal(size=(n, 3))
    Vt = rng.normal(size=(n, 3)) * 1e-3
    w = rng.uniform(0.5, 2.0, n)
    D = np.hstack([np.eye(g)[ids] for g, ids in zip(G, gids)])
    A = D.T * w                          # D'W
    resid = np.linalg.norm(A @ Vt, axis=0)
    rel = resid / (np.linalg.norm(A) * np.linalg.norm(V, axis=0))
    o = oracle.certify(V, Vt, gids, G, weights=w, tol=1e-8)
    assert abs(o["abs_residual"] - resid.max()) <= 1e-12 * resid.max()
    assert abs(o["abs_residual_rel"] - rel.max()) <= 1e-12 * rel.max()
    np.testing.assert_allclose(o["per_column_rel"], rel, rtol=1e-12)
    # the certification limit max(8 tol_eff, 64 eps), tol_eff = min(tol, 1e-9) (reghdfe-comparable)
    small = Vt * (5e-9 / rel.max())
    assert oracle.certify(V, small, gids, G, weights=w, tol=1e-8)["certified"]          # 5e-9 <= 8e-9
    assert not oracle.certify(V, small * 2, gids, G, weights=w, tol=1e-8)["certified"]  # 1e-8 > 8e-9
    assert not oracle.certify(V, small, gids, G, weights=w, tol=1e-8, converged=False)["certified"]
    assert oracle.certify(V, small * 2, gids, G, weights=w, tol=1e-8, tolerance_mode=0)["certified"]


# ---------------------------------------------------------------- AKM/KSS grounded solve
# TwoWaySolver::build (src/akm_kss.cpp:1042-1091) and the grounded Jacobi-PCG
# (src/akm_kss_cuda.cu:532-695), restated in the oracle, against a dense solve.

def akm_graph(seed, N=400, J=60, movers=0.4):
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


def _dense_S(st):
    B = np.zeros((st["N"], st["J"]))
    np.add.at(B, (st["m_w"], st["m_f"]), st["m_c"])
    return np.diag(st["Df"]) - B.T @ (B / st["Dw"][:, None])


def test_akm_oracle_matches_dense_solve():
    for seed in (1, 2, 3):
        st = akm_graph(seed)
        J = st["J"]
        S = _dense_S(st)
        rng = np.random.default_rng(seed + 10)
        p = rng.normal(size=J); p[J - 1] = 0.0
        Sp = S @ p; Sp[J - 1] = 0.0
        np.testing.assert_allclose(oracle.akm_apply_S(st, p), Sp, rtol=1e-12, atol=1e-12 * np.abs(Sp).max())
        rhs = rng.normal(size=J)
        z, it = oracle.akm_solve_S(st, rhs, 1e-12, 10000)
        # firms unreachable from the rest make S singular: solve on the reachable block by lstsq
        zd = np.zeros(J)
        zd[:J - 1] = np.linalg.lstsq(S[:J - 1, :J - 1], rhs[:J - 1], rcond=None)[0]
        assert it > 0
        assert np.abs(z - zd).max() <= 1e-8 * np.abs(zd).max(), (seed, np.abs(z - zd).max())
        assert z[J - 1] == 0.0


def test_akm_oracle_lane_rules():
    st = akm_graph(4)
    J = st["J"]
    z, it = oracle.akm_solve_S(st, np.zeros(J), 1e-10, 100)
    assert it == 0 and not z.any()
    rhs = np.zeros(J); rhs[J - 1] = 5.0              # only the ground entry: pinned away
    z, it = oracle.akm_solve_S(st, rhs, 1e-10, 100)
    assert it == 0 and not z.any()
    rhs = np.random.default_rng(5).normal(size=J)
    z, it = oracle.akm_solve_S(st, rhs, 1e-14, 2)    # cannot converge in 2 iterations
    assert it == -1 and not z.any()


# ------------------------------------------------------------ schwarz_demean_raw

def _schwarz_inputs(case):
    keep = _kept(case)
    gids, G = _compact([f[keep] for f in case["fes"]])
    return gids, G, np.column_stack([case["y"][keep], case["X"][keep]])


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
@pytest.mark.parametrize("tag", ["tight", "default"])
def test_schwarz_fixture_matches_oracle(golden, cases, name, tag):
    """The reference's PCG demeaner (schwarz_demean_raw at tol 1e-12 / 1e-8)
    lands on the projection the oracle computes, within its tolerance."""
    gids, G, V = _schwarz_inputs(cases[name])
    assert golden[f"{name}/schwarz_{tag}/ok"] == 1
    Vr = np.column_stack([golden[f"{name}/schwarz_{tag}/y_tilde"], golden[f"{name}/schwarz_{tag}/X_tilde"]])
    Vo, _ = oracle.absorb(V, gids, G, tol=1e-14)
    err = np.abs(Vr - Vo).max() / np.abs(V).max()
    assert err <= (1e-11 if tag == "tight" else 1e-7), err
    tol = 1e-12 if tag == "tight" else 1e-8
    assert oracle.schwarz_rule(V, Vr, gids, G).max() < tol


def test_schwarz_rule_closed_forms(cases):
    gids, G, V = _schwarz_inputs(cases["akm_three_way"])
    # the untransformed block leaves the whole right-hand side: rel = 1
    np.testing.assert_allclose(oracle.schwarz_rule(V, V, gids, G), 1.0, rtol=1e-14)
    # the projection itself: rel at the rounding floor
    Vo, _ = oracle.absorb(V, gids, G, tol=1e-14)
    assert oracle.schwarz_rule(V, Vo, gids, G).max() < 1e-12
    # a zero column is inactive (rel 0), the others unchanged
    Z = V.copy()
    Z[:, 1] = 0.0
    rel = oracle.schwarz_rule(Z, Z, gids, G)
    assert rel[1] == 0.0 and np.all(rel[[0, 2]] == 1.0)
    # homogeneity: the rule does not depend on the scale of a column
    # and halfway to the projection the residual halves: rel = 1/2
    H = 0.5 * (V + Vo)
    np.testing.assert_allclose(oracle.schwarz_rule(3.0 * V, 3.0 * H, gids, G), 0.5, rtol=1e-10)


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_schwarz_rule_matches_reference(cases, name):
    """The oracle's rule reads the reference's own convergence test: at a
    tolerance 1e-4 above the rule's value on the reference's output the
    reference stops at the same iteration, 1e-4 below it needs one more
    (leaving the pinned levels in moves the value by 1.6e-4 to 7.7e-3 here)."""
    if oracle.ref() is None:
        pytest.skip("oracle/_ref is not built (needs the reference sources)")
    gids, G, V = _schwarz_inputs(cases[name])
    ok, yo, Xo, it = oracle.ref_schwarz(V[:, 0], V[:, 1:], gids, G, tol=1e-8)
    assert ok
    r = oracle.schwarz_rule(V, np.column_stack([yo, Xo]), gids, G).max()
    assert oracle.ref_schwarz(V[:, 0], V[:, 1:], gids, G, tol=r * 1.0001)[3] == it
    assert oracle.ref_schwarz(V[:, 0], V[:, 1:], gids, G, tol=r * 0.9999)[3] == it + 1


def test_schwarz_edge_fixture_regenerates(golden):
    """The committed edge-case flags are what the reference returns today."""
    if oracle.ref() is None:
        pytest.skip("oracle/_ref is not built (needs the reference sources)")
    from golden_cases import schwarz_edge_cases
    got = [list(oracle.ref_schwarz(*args, tol=1e-10, max_iter=mi)[0::3]) for args, mi in schwarz_edge_cases()]
    np.testing.assert_array_equal(np.array(got, dtype=np.int64), golden["schwarz_edges"])


def test_akm_host_structure_matches_oracle():
    """The package's TwoWaySolver::build restatement (what the device solver is
    handed) equals the oracle's on a random match list."""
    from xfe_amd import akm
    rng = np.random.default_rng(5)
    worker = rng.integers(0, 300, 4000)
    firm = rng.integers(0, 40, 4000)
    mw, mf, mc, N, J = akm.panel_matches(worker, firm, connected=False)
    assert np.all(np.diff(mw) >= 0) and mc.sum() == 4000
    a = akm.build_structure(mw, mf, mc, N, J)
    b = oracle.akm_build(mw, mf, mc, N, J)
    for k in b:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


# ------------------------------------------------------------ heterogeneous slopes

def _slope_inputs(case):
    keep = _kept(case)
    gids, G = _compact([f[keep] for f in case["fes"]])
    return gids, G, np.column_stack([case["y"][keep], case["X"][keep]])


@pytest.mark.parametrize("name,sdim,icpt", [("two_way", 1, True), ("two_way", 1, False),
                                             ("akm_three_way", 1, True)])
def test_slope_projection_matches_reference(golden, cases, name, sdim, icpt):
    """The exact slope projection against the reference's absorber with the
    slope term at tol 1e-12 (its alternating projections converge slowly on
    these panels: 53 - 2,508 iterations; 9e-9 of the column scale left)."""
    from golden_cases import slope_values
    gids, G, V = _slope_inputs(cases[name])
    z = slope_values(V.shape[0])
    tag = f"slope{sdim}_{'icpt' if icpt else 'noicpt'}"
    assert golden[f"{name}/{tag}/converged"] == 1
    Vr = np.column_stack([golden[f"{name}/{tag}/y_tilde"], golden[f"{name}/{tag}/X_tilde"]])
    Vo = oracle.slope_projection(V, gids, G, sdim, z, intercept=icpt)
    assert np.abs(Vr - Vo).max() / np.abs(V).max() <= 2e-8


def test_slope_projection_matches_reference_tight():
    """Where the reference converges fast it lands on the oracle's projection
    to rounding (slope on either dimension, with and without the intercept)."""
    if oracle.ref() is None:
        pytest.skip("oracle/_ref is not built (needs the reference sources)")
    rng = np.random.default_rng(3)
    n = 3000
    a = rng.integers(0, 200, n)
    b = rng.integers(0, 40, n)
    z = rng.normal(size=n) + 0.5
    y = rng.normal(size=n)
    X = rng.normal(size=(n, 2))
    gids, G = _compact([a, b])
    V = np.column_stack([y, X])
    for sdim in (0, 1):
        for icpt in (True, False):
            yo, Xo, _, conv = oracle.ref_absorb_slope(y, X, [a, b], sdim, z, intercept=icpt, tol=1e-13)
            Vo = oracle.slope_projection(V, gids, G, sdim, z, intercept=icpt)
            assert conv and np.abs(np.column_stack([yo, Xo]) - Vo).max() <= 1e-12


def test_slope_projection_single_dimension():
    """One dimension with a slope: per level, the residual of the least-squares
    line of v on z (numpy polyfit); without the intercept, of v on z alone."""
    rng = np.random.default_rng(11)
    n = 900
    g = rng.integers(0, 30, n)
    z = rng.normal(size=n)
    V = rng.normal(size=(n, 2))
    for icpt in (True, False):
        Vo = oracle.slope_projection(V, [g], [30], 0, z, intercept=icpt)
        for lev in range(30):
            m = g == lev
            for c in range(2):
                if icpt:
                    slope, inter = np.polyfit(z[m], V[m, c], 1)
                    want = V[m, c] - (inter + slope * z[m])
                else:
                    want = V[m, c] - z[m] * (z[m] @ V[m, c]) / (z[m] @ z[m])
                np.testing.assert_allclose(Vo[m, c], want, atol=1e-12)


# ------------------------------------------------------------ JLA leverages

def test_splitmix64_known_answers():
    """SplitMix64's published outputs (Vigna's reference generator, state 0:
    the first two outputs)."""
    assert int(oracle.splitmix64(np.uint64(0))) == 0xE220A8397B1DCDAF
    assert int(oracle.splitmix64(np.uint64(0x9E3779B97F4A7C15))) == 0x6E789E6AA1B965F4


def _jla_structure(golden):
    rw, rf, rc = golden["jla/row_worker"], golden["jla/row_firm"], golden["jla/row_weight"]
    wu, wi = np.unique(rw, return_inverse=True)
    fu, fi = np.unique(rf, return_inverse=True)
    assert np.all(np.diff(wi) >= 0)  # the reference's working rows are worker-major
    return oracle.akm_build(wi, fi, rc, wu.size, fu.size)


def test_jla_pii_matches_reference(golden):
    """The oracle's JLA leverages with the reference's streams equal the
    reference's akm_kss_decompose P_ii (50 draws, seed 42) to rounding."""
    from golden_cases import JLA_DRAWS, JLA_SEED
    st = _jla_structure(golden)
    pii = oracle.akm_jla_pii(st, JLA_DRAWS, JLA_SEED)
    assert np.abs(pii - golden["jla/pii"]).max() <= 1e-12


def test_jla_pii_limit_is_exact_leverage(golden):
    """Many draws approach the exact leverages x_m' (A'A)^+ x_m of the
    match-level design (dense), within the draws' sampling error."""
    st = _jla_structure(golden)
    N, J = st["N"], st["J"]
    A = np.zeros((st["m_w"].size, N + J))
    sq = np.sqrt(st["m_c"])
    A[np.arange(A.shape[0]), st["m_w"]] = sq
    A[np.arange(A.shape[0]), N + st["m_f"]] = sq
    H = A @ np.linalg.pinv(A.T @ A) @ A.T
    exact = np.diag(H)
    est = oracle.akm_jla_pii(st, 400, 7)
    assert np.abs(est - exact).mean() < 0.02 and np.corrcoef(est, exact)[0, 1] > 0.95


def test_rademacher_sums_closed_forms():
    """A count of 0 sums to 0; a count of 1 is the low bit of the advanced
    state (1 -> +1); a count of 64 + k is the 64-bit popcount of the first
    advance plus the k-bit one of the second; parity equals count's."""
    cnt = np.array([0, 1, 64, 70, 129, 5])
    got = oracle.rademacher_sums(11, 4, cnt)
    s0 = oracle.stream_seed(11, 4, np.arange(cnt.size, dtype=np.uint64))
    s1 = oracle.splitmix64(s0)
    s2 = oracle.splitmix64(s1)
    s3 = oracle.splitmix64(s2)
    pop = lambda x, k: bin(int(x) & ((1 << k) - 1)).count("1")
    assert got[0] == 0.0
    assert got[1] == (1.0 if int(s1[1]) & 1 else -1.0)
    assert got[2] == 2 * pop(s1[2], 64) - 64
    assert got[3] == 2 * pop(s1[3], 64) - 64 + 2 * pop(s2[3], 6) - 6
    assert got[4] == 2 * (pop(s1[4], 64) + pop(s2[4], 64) + pop(s3[4], 1)) - 129
    assert np.all((got - cnt) % 2 == 0) and np.all(np.abs(got) <= cnt)


def _jla_bii(golden):
    from golden_cases import JLA_DRAWS, JLA_SEED
    st = _jla_structure(golden)
    rc = golden["jla/row_weight"]
    return st, rc, oracle.akm_jla_bii(st, JLA_DRAWS, JLA_SEED, rc.astype(np.int64))


def test_jla_bii_matches_reference_components(golden):
    """The reference's AGSU components are plug-in - sigma2_ho sum(B)/dof and
    its KSS components plug-in - sum(B sigma_i)/dof (50 draws, seed 42):
    the oracle's B_ii draws reproduce both linear functionals of each of the
    three weight vectors to rounding."""
    st, rc, (bfe, bpe, bcov) = _jla_bii(golden)
    comp, sig = golden["jla/components"], golden["jla/sigma_i"]
    assert comp[10] == rc.sum()
    dof = comp[10] - 1.0
    for i, B in enumerate((bpe, bfe, bcov)):  # var_alpha, var_psi, cov
        want_sum = (comp[i] - comp[3 + i]) * dof / comp[9]
        want_sig = (comp[i] - comp[6 + i]) * dof
        assert abs(B.sum() - want_sum) <= 1e-11 * abs(want_sum)
        assert abs(B @ sig - want_sig) <= 1e-11 * abs(want_sig)


def test_jla_bii_limit_is_exact_quadratic_form(golden):
    """Many draws approach the exact weights: with u's covariance over
    person years C = diag(count) - count count'/n_py and
    L = A K^+ [0; F'] (fe) or A K^+ [W'; 0] (pe), E Bfe = diag(L_fe C L_fe'),
    E Bpe = diag(L_pe C L_pe'), E Bcov = diag(L_pe C L_fe') (dense)."""
    st, rc, _ = _jla_bii(golden)
    N, J = st["N"], st["J"]
    M = st["m_w"].size
    sq = np.sqrt(st["m_c"])
    A = np.zeros((M, N + J))
    A[np.arange(M), st["m_w"]] = sq
    A[np.arange(M), N + st["m_f"]] = sq
    Kp = np.linalg.pinv(A.T @ A)
    Wm = np.zeros((N, M))
    Wm[st["m_w"], np.arange(M)] = 1.0
    Fm = np.zeros((J, M))
    Fm[st["m_f"], np.arange(M)] = 1.0
    Lfe = A @ Kp @ np.vstack([np.zeros((N, M)), Fm])
    Lpe = A @ Kp @ np.vstack([Wm, np.zeros((J, M))])
    C = np.diag(rc) - np.outer(rc, rc) / rc.sum()
    efe = np.einsum("ij,jk,ik->i", Lfe, C, Lfe)
    epe = np.einsum("ij,jk,ik->i", Lpe, C, Lpe)
    eco = np.einsum("ij,jk,ik->i", Lpe, C, Lfe)
    bfe, bpe, bcov = oracle.akm_jla_bii(st, 600, 3, rc.astype(np.int64))
    for est, ex in ((bfe, efe), (bpe, epe), (bcov, eco)):
        assert abs(est.sum() - ex.sum()) <= 0.05 * np.abs(ex).sum()
        assert np.corrcoef(est, ex)[0, 1] > 0.9


# ---- savefe FE recovery (recover_fixed_effects, src/fe_absorption.cpp:14447-15700)

def _recover_case(cases, name):
    p = cases[name]
    keep = _kept(p)
    fk = [f[keep] for f in p["fes"]]
    ck = [oracle.build_indexer(f) for f in fk]
    return p["y"][keep], p["X"][keep], fk, [c[0] for c in ck], [c[1] for c in ck]


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_recover_matches_reference(golden, cases, name):
    """The oracle's recovery against the reference's own (tight fe_tolerance):
    the FE values are identified only up to constants per dimension and
    mobility group, so their sum per row is what both must agree on; and it
    is the exact fitted component partial - M_D partial."""
    y, X, fk, gids, G = _recover_case(cases, name)
    alphas, it, md, conv = oracle.recover_fixed_effects(y, gids, G, fe_tol=1e-12)
    assert conv and md < 1e-12
    tot = sum(a[g] for a, g in zip(alphas, gids))
    ref = golden[f"{name}/recover_tight/contributions"]
    scale = np.abs(y).max()
    assert np.abs(tot - ref.sum(0)).max() / scale < 1e-8
    Vo, _ = oracle.absorb(np.column_stack([y, X]), gids, G, tol=1e-14)
    assert np.abs(tot - (y - Vo[:, 0])).max() / scale < 1e-9
    # the reference's default fe_tolerance stops early: its sum is within what
    # its max_delta allows
    refd = golden[f"{name}/recover_default/contributions"]
    it_d, md_d, conv_d = golden[f"{name}/recover_default/info"]
    assert conv_d == 1 and md_d < 1e-6
    assert np.abs(tot - refd.sum(0)).max() / scale < 1e-4


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_recovery_max_delta_of_reference_output(golden, cases, name):
    """fe_recovery_max_delta of the reference's converged recovery: one more
    Gauss-Seidel pass moves no level by more than the reference's stopping
    threshold allows."""
    y, _, _, gids, G = _recover_case(cases, name)
    for tag, ftol in (("tight", 1e-12), ("default", 1e-6)):
        C = golden[f"{name}/recover_{tag}/contributions"]
        md = oracle.fe_recovery_max_delta(y - C.sum(0), gids, G)
        _, md_ref, _ = golden[f"{name}/recover_{tag}/info"]
        assert md <= 2.0 * ftol, (tag, md, md_ref)


def test_recover_tiny_is_least_squares():
    """On a tiny weighted two-way design the recovered coefficients fit
    partial exactly as dense weighted least squares does (the fitted values
    are unique even where the coefficients are not)."""
    rng = np.random.default_rng(5)
    n = 60
    a = rng.integers(0, 7, n).astype(np.int32)
    b = rng.integers(0, 5, n).astype(np.int32)
    w = 0.5 + rng.random(n)
    y = rng.standard_normal(n)
    alphas, _, md, conv = oracle.recover_fixed_effects(y, [a, b], [7, 5], weights=w, fe_tol=1e-13)
    assert conv
    Dm = np.hstack([np.eye(7)[a], np.eye(5)[b]])
    sw = np.sqrt(w)
    coef = np.linalg.lstsq(Dm * sw[:, None], y * sw, rcond=None)[0]
    np.testing.assert_allclose(alphas[0][a] + alphas[1][b], Dm @ coef, atol=1e-11)
    # one dimension: the weighted group means, exactly
    al, it, _, _ = oracle.recover_fixed_effects(y, [a], [7], weights=w)
    assert it == 2 or it == 1
    np.testing.assert_allclose(al[0], np.bincount(a, w * y, 7) / np.bincount(a, w, 7), rtol=1e-14)


# ---- group()/individual() absorption (absorb_fixed_effects_group_individual)

def _gi_case():
    from golden_cases import gi_panel
    p = gi_panel()
    return p, np.column_stack([p["y"], p["X"]])


def test_gi_colouring_is_proper_and_greedy():
    """No two individuals of one colour share a row, and every individual's
    colour is the smallest not taken by an earlier neighbour."""
    p, _ = _gi_case()
    gi = p["gi"]
    colour, classes = oracle.gi_colour(gi)
    gp, gind = gi["group_ptr"], gi["group_individual"]
    for g in range(len(gp) - 1):
        team = gind[gp[g]:gp[g + 1]]
        assert len(set(colour[team].tolist())) == len(team)
    assert sum(len(c) for c in classes) == gi["num_individuals"]
    assert colour.max() + 1 >= max(np.diff(gp))


@pytest.mark.parametrize("tag", ["gs", "sgs_w"])
def test_gi_oracle_matches_reference_and_exact(golden, tag):
    """The oracle's group()/individual() transform against the reference's CPU
    absorber (1e-7 of the column scale: both stop certified at 8 tol_eff) and
    against the exact weighted least squares projection (1e-8)."""
    from golden_cases import gi_weights
    p, V = _gi_case()
    w = gi_weights(V.shape[0]) if tag == "sgs_w" else None
    Vt, it, cert = oracle.absorb_group_individual(V, p["fes"], p["G"], p["gi"], weights=w,
                                                  method=2 if tag == "sgs_w" else 1)
    assert cert
    ref = np.column_stack([golden[f"gi/{tag}/y_tilde"], golden[f"gi/{tag}/X_tilde"]])
    scale = np.abs(V).max()
    assert golden[f"gi/{tag}/info"][1] == 1
    assert np.abs(Vt - ref).max() / scale < 1e-7
    ex = oracle.gi_exact(V, p["fes"], p["G"], p["gi"], weights=w)
    assert np.abs(Vt - ex).max() / scale < 1e-8
    # the reference's own output passes the certificate restatement
    assert oracle.certify_group_individual(V, ref, p["fes"], p["G"], p["gi"], weights=w)[2]


def test_gi_certificate_rejects_untransformed():
    p, V = _gi_case()
    a, rel, cert = oracle.certify_group_individual(V, V, p["fes"], p["G"], p["gi"])
    assert not cert and rel > 1e-3


# ---- shadow device certificate (cuda_cert_shadow_*): the oracle's brackets

def _cert_case(golden, cases, name):
    p = cases[name]
    keep = _kept(p)
    fk = [f[keep] for f in p["fes"]]
    ck = [oracle.build_indexer(f) for f in fk]
    V0 = np.column_stack([p["y"][keep], p["X"][keep]])
    Vt = np.column_stack([golden[f"{name}/tight/y_tilde"], golden[f"{name}/tight/X_tilde"]])
    return V0, Vt, [c[0] for c in ck], [c[1] for c in ck]


def test_cert_shadow_brackets_contain_exact(golden, cases):
    """Exact rational Res_sq, Op_sq and Orig_sq of the reference's converged
    output and of the raw columns lie inside the oracle's brackets; converged
    passes at the tight limit, raw fails."""
    V0, Vt, gids, G = _cert_case(golden, cases, "sparse_two_way")
    for Vr, want in ((Vt, "PASS"), (V0, "FAIL")):
        sh = oracle.cert_shadow(V0, Vr, gids, G, 8e-12)
        R, A, B = oracle.cert_shadow_exact(V0, Vr, gids, G)
        assert sh["A"][0] <= A <= sh["A"][1]
        for c in range(V0.shape[1]):
            assert sh["R"][c][0] <= R[c] <= sh["R"][c][1]
            assert sh["B"][c][0] <= B[c] <= sh["B"][c][1]
        assert set(sh["verdict"]) == {want}


def test_cert_shadow_weighted_brackets_contain_exact(golden, cases):
    from golden_cases import certificate_weights
    V0, Vt, gids, G = _cert_case(golden, cases, "sparse_two_way")
    w = certificate_weights(V0.shape[0])
    sh = oracle.cert_shadow(V0[:, :1], V0[:, :1] * 0.5 + Vt[:, :1], gids, G, 8e-9, weights=w)
    R, A, B = oracle.cert_shadow_exact(V0[:, :1], V0[:, :1] * 0.5 + Vt[:, :1], gids, G, weights=w)
    assert sh["A"][0] <= A <= sh["A"][1]
    assert sh["R"][0][0] <= R[0] <= sh["R"][0][1]


@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
def test_cert_shadow_value_matches_reference(golden, cases, name):
    """The to-nearest value behind the decision is the reference's own
    certificate of the raw columns (fixture, 1e-11); its converged outputs
    pass, as the reference certified them."""
    V0, Vt, gids, G = _cert_case(golden, cases, name)
    raw = oracle.cert_shadow(V0, V0, gids, G, 8e-9)
    ref_rel = golden[f"{name}/raw/cert"][1]
    assert abs(max(raw["rel"]) - ref_rel) <= 1e-11 * ref_rel
    assert set(raw["verdict"]) == {"FAIL"}
    assert golden[f"{name}/tight/cert"][2] == 1
    assert set(oracle.cert_shadow(V0, Vt, gids, G, 8e-12)["verdict"]) == {"PASS"}


def test_cert_shadow_limit_straddle(golden, cases):
    """At L = the to-nearest rel the brackets straddle (INDETERMINATE); a
    relative 1e-6 either side decides."""
    V0, _, gids, G = _cert_case(golden, cases, "two_way")
    y = V0[:, :1]
    rel = oracle.cert_shadow(y, y, gids, G, 1.0)["rel"][0]
    assert oracle.cert_shadow(y, y, gids, G, rel)["verdict"] == ["INDETERMINATE"]
    assert oracle.cert_shadow(y, y, gids, G, rel * (1 + 1e-6))["verdict"] == ["PASS"]
    assert oracle.cert_shadow(y, y, gids, G, rel * (1 - 1e-6))["verdict"] == ["FAIL"]


# ---- articulation points (the leave-out set's cut-vertex search)

def _random_graphs(seed, count, nmax):
    rng = np.random.default_rng(seed)
    for t in range(count):
        n = int(rng.integers(1, nmax))
        kind = t % 4
        if kind == 0:
            raw = rng.integers(0, n, (int(rng.integers(0, 2 * n + 1)), 2))
            edges = {tuple(sorted(map(int, e))) for e in raw if e[0] != e[1]}
        elif kind == 1:  # chains with a few chords
            edges = {(i, i + 1) for i in range(n - 1)}
            edges |= {tuple(sorted(map(int, e))) for e in rng.integers(0, n, (int(rng.integers(0, 3)), 2)) if e[0] != e[1]}
        elif kind == 2:  # trees
            edges = {(int(rng.integers(0, i)), i) for i in range(1, n)}
        else:  # two dense halves joined by one edge
            h = max(n // 2, 1)
            edges = {tuple(sorted(map(int, e))) for e in rng.integers(0, h, (n, 2)) if e[0] != e[1]}
            edges |= {tuple(sorted(map(int, e))) for e in rng.integers(h, n, (n, 2)) if e[0] != e[1]} if n > h else set()
            if n > 2:
                edges.add((0, n - 1))
        yield n, sorted(edges)


def _brute_cut_vertices(n, ptr, adj):
    A = sp.csr_matrix((np.ones(len(adj)), adj, ptr), shape=(n, n))
    base = sp.csgraph.connected_components(A, directed=False)[0]
    out = np.zeros(n, dtype=np.uint8)
    for x in range(n):
        keep = np.ones(n, bool)
        keep[x] = False
        c = sp.csgraph.connected_components(A[keep][:, keep], directed=False)[0]
        out[x] = c > base - (1 if ptr[x + 1] == ptr[x] else 0)
    return out


def test_articulation_oracle_is_brute_force():
    """The restated depth-first search flags exactly the vertices whose
    removal adds a component (400 small graphs: random, chains with chords,
    trees, two halves joined by a bridge)."""
    import scipy.sparse.csgraph  # noqa: F401
    for n, edges in _random_graphs(7, 400, 22):
        ptr, adj = oracle.graph_csr(n, edges)
        np.testing.assert_array_equal(oracle.articulation_points(ptr, adj), _brute_cut_vertices(n, ptr, adj))


_LOS_KEYS = ("n_obs_input", "n_obs_connected", "n_obs", "n_workers", "n_firms", "n_matches", "n_movers",
             "n_stayers", "prune_iterations")


@pytest.mark.parametrize("name", ["low_mobility", "moderate", "gapped_ids"])
def test_leave_out_set_matches_reference(golden, name):
    """The oracle's leave-out connected set equals the reference's
    leave_out_connected_set row for row, with the same counts and pruning
    rounds (fixtures)."""
    from golden_cases import leave_out_panels
    w, f = leave_out_panels()[name]
    keep, info = oracle.leave_out_set(w, f)
    np.testing.assert_array_equal(keep, golden[f"los/{name}/keep"].astype(bool))
    assert [info[k] for k in _LOS_KEYS] == golden[f"los/{name}/info"].tolist()


# ---- the indexer and weight sums (build_indexer, compute_weight_sums) ----

def test_indexer_matches_reference(golden):
    """Every branch of the reference's build_indexer (fixtures written from the
    compiled reference through retain_fixed_effects): ids, level counts,
    levels present; compute_weight_sums' counts and weighted sums."""
    from golden_cases import indexer_cases
    for name, (raw, w) in indexer_cases().items():
        gid, G, lp, cont = oracle.build_indexer(raw)
        np.testing.assert_array_equal(gid, golden[f"indexer/{name}/gids"], err_msg=name)
        assert [G, lp] == list(golden[f"indexer/{name}/info"]), name
        assert cont == bool(np.all(np.diff(gid) >= 0)), name
        np.testing.assert_array_equal(oracle.weight_sums(gid, G), golden[f"indexer/{name}/counts"])
        ref = golden[f"indexer/{name}/wsums"]
        np.testing.assert_allclose(oracle.weight_sums(gid, G, w), ref, rtol=1e-13, atol=1e-13)


def test_indexer_rules_by_brute_force():
    """The documented rules restated independently on random vectors: ids
    follow ascending raw ids; dense ranges keep unused slots unless fewer than
    90 % are used; the sparse path ranks the distinct ids."""
    rng = np.random.default_rng(7)
    for t in range(400):
        n = int(rng.integers(1, 300))
        span = int(rng.choice([n // 2 + 1, n, 2 * n, 3 * n, 10 ** 7]))
        raw = (rng.integers(0, span, n) - rng.integers(0, 3) * span // 2).astype(np.int32)
        gid, G, lp, _ = oracle.build_indexer(raw)
        uniq = np.unique(raw)
        rng_ = int(raw.max()) - int(raw.min()) + 1
        if rng_ <= 2 * n and uniq.size * 10 >= rng_ * 9:
            assert G == rng_ and lp == uniq.size
            np.testing.assert_array_equal(gid, raw.astype(np.int64) - raw.min())
        else:
            assert G == lp == uniq.size
            np.testing.assert_array_equal(gid, np.searchsorted(uniq, raw))


def test_indexer_against_compiled_reference_random():
    if oracle.ref() is None:
        pytest.skip("compiled reference not built here")
    rng = np.random.default_rng(11)
    for t in range(200):
        n = int(rng.integers(1, 500))
        raw = rng.integers(-(10 ** int(rng.integers(1, 9))), 10 ** int(rng.integers(1, 9)), n).astype(np.int32)
        (g, G, lp, _), = oracle.ref_indexer([raw])
        og, oG, olp, _ = oracle.build_indexer(raw)
        np.testing.assert_array_equal(og, g)
        assert (oG, olp) == (G, lp)


# ---- the weighted within transform ----

@pytest.mark.parametrize("name", ["two_way", "akm_three_way", "sparse_two_way"])
@pytest.mark.parametrize("tag", ["tight", "default"])
def test_absorb_weighted_matches_reference(golden, cases, name, tag):
    """The oracle's weighted absorber against the reference's
    absorb_fixed_effects with the same analytic weights: same iterations
    (one more on the low-mobility panel at tol 1e-12) and the same transform."""
    from golden_cases import certificate_weights
    p = cases[name]
    keep = _kept(p)
    fes = [f[keep] for f in p["fes"]]
    gids, G = _compact(fes)
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    w = certificate_weights(V.shape[0])
    tol = 1e-12 if tag == "tight" else 1e-8
    Vo, info = oracle.absorb(V, gids, G, weights=w, tol=tol)
    ref = np.column_stack([golden[f"{name}/w_{tag}/y_tilde"], golden[f"{name}/w_{tag}/X_tilde"]])
    err = np.abs(Vo - ref).max() / np.abs(V).max()
    assert err <= (1e-10 if tag == "tight" else 1e-9), err
    assert info["converged"] and bool(golden[f"{name}/w_{tag}/converged"])
    assert abs(info["iterations"] - int(golden[f"{name}/w_{tag}/iterations"])) <= 1


def test_absorb_weighted_closed_form():
    """Weighted least squares by a dense solve on a tiny two-way case: the
    weighted within transform is the W-orthogonal residual, D'W v~ = 0, and
    it is idempotent."""
    rng = np.random.default_rng(5)
    n = 240
    a = rng.integers(0, 30, n).astype(np.int32)
    b = rng.integers(0, 12, n).astype(np.int32)
    keep, _, _ = oracle.drop_singletons([a, b])
    a, b = a[keep], b[keep]
    n = a.size
    gids, G = _compact([a, b])
    V = rng.normal(size=(n, 3))
    w = rng.uniform(0.2, 3.0, n)
    Vo, info = oracle.absorb(V, gids, G, weights=w, tol=1e-14, tolerance_mode=0)
    D = np.zeros((n, G[0] + G[1]))
    D[np.arange(n), gids[0]] = 1
    D[np.arange(n), G[0] + gids[1]] = 1
    sw = np.sqrt(w)[:, None]
    coef, *_ = np.linalg.lstsq(D * sw, V * sw, rcond=None)
    exact = V - D @ coef
    assert np.abs(Vo - exact).max() / np.abs(V).max() < 1e-10
    assert np.abs(D.T @ (w[:, None] * Vo)).max() < 1e-9
    Vo2, _ = oracle.absorb(Vo, gids, G, weights=w, tol=1e-14, tolerance_mode=0)
    assert np.abs(Vo2 - Vo).max() < 1e-10


# ---- the learned preconditioner's network (oracle/learned.py) ----

@pytest.mark.parametrize("F,Wm", [(30, 40), (3, 400)])
def test_lp_gnn_gradient_finite_differences(F, Wm):
    """The oracle's backward pass against central finite differences of its
    forward pass (exact products, float64), on a random mobility graph: every
    parameter block is probed.  With 3 firms and 400 movers every firm has
    more movers than the firm mean's cap of 256, so the capped mean and its
    transpose are checked too."""
    from oracle import learned as L
    g = L.random_graph(3, F=F, Wm=Wm, deg=(2, 3) if F < 4 else (2, 4))
    assert (F >= 4) or np.diff(g["f_ptr"]).min() > L.FIRM_CAP
    th = L.init_params(1).astype(np.float64)
    _, grad, _ = L.forward_backward(th, g, bf16_products=False)
    rng = np.random.default_rng(0)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in L.LAYOUT])
    probes = [int(rng.integers(offs[k], offs[k + 1])) for k in range(len(L.LAYOUT)) for _ in range(3)]
    worst = 0.0
    for i in probes:
        h = 1e-6
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (L.forward_backward(tp, g, False, False)[0] - L.forward_backward(tm, g, False, False)[0]) / (2 * h)
        worst = max(worst, abs(fd - grad[i]) / (abs(fd) + abs(grad[i]) + 1e-7))
    assert worst < 1e-5, worst


def test_lp_gnn_bf16_rounding():
    """bf16() rounds to nearest even on the 16 dropped bits (known answers)."""
    from oracle import learned as L
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3], dtype=np.float32)
    got = L.bf16(x)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, 197 * 2.0 ** -16], dtype=np.float64)  # 3e-3 -> 197 x 2^-16
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9)


# ---- the aggregate-block preconditioner (oracle/precond.py) ----

def _tiny_akm(seed, workers=60, firms=12, years=6, p_move=0.3):
    rng = np.random.default_rng(seed)
    T = rng.integers(2, years + 1, workers)
    w = np.repeat(np.arange(workers), T)
    y = np.concatenate([np.arange(t) for t in T])
    f = np.empty(w.size, np.int64)
    cur = rng.integers(0, firms, workers)
    for i in range(w.size):
        if i > 0 and w[i] == w[i - 1] and rng.random() < p_move:
            cur[w[i]] = rng.integers(0, firms)
        f[i] = cur[w[i]]
    return w, f, y


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_precond_pairs_equal_dense_schur(seed):
    """The firm block of the Schur complement of the worker dimension from the
    sparse formula (pairs_dense) equals the dense D_R' M_p D_R: the diagonal,
    every coupled pair, and zeros elsewhere."""
    from oracle import precond as P
    w, f, y = _tiny_akm(seed)
    F = f.max() + 1
    pf, pg, W, dex = P.pairs_dense(w, f, F=F)
    S = P.schur_dense(w, f, y)[:F, :F]
    assert np.abs(np.diag(S) - dex).max() < 1e-10
    off = np.zeros((F, F))
    off[pf, pg] = -W
    off = off + off.T
    mask = ~np.eye(F, dtype=bool)
    assert np.abs(S[mask] - off[mask]).max() < 1e-10
    assert (pf < pg).all() and (np.diff(pf.astype(np.int64) * F + pg) > 0).all()


def test_precond_handshake_rules():
    """One round merges exactly the mutual best pairs (brute force over the
    firms); every aggregate respects the cap and is connected by pairs."""
    from oracle import precond as P
    rng = np.random.default_rng(5)
    F = 40
    cand = {(int(a), int(b)) for a, b in rng.integers(0, F, (120, 2)) if a < b}
    pf, pg = map(np.array, zip(*sorted(cand)))
    score = rng.random(pf.size).astype(np.float32)
    agg1 = P.handshake(pf, pg, score, F, cap=16, rounds=1)
    best = {}
    for p in range(pf.size):
        for a in (pf[p], pg[p]):
            k = (score[p], -p)
            if a not in best or k > best[a][0]:
                best[a] = (k, p)
    expect = np.arange(F)
    for a, (_, p) in best.items():
        b = pg[p] if pf[p] == a else pf[p]
        if best.get(b, (None, -1))[1] == p and a < b:
            expect[b] = a
    assert (agg1 == expect).all()
    for cap in (2, 4, 16):
        agg = P.handshake(pf, pg, score, F, cap=cap, rounds=8)
        for r in np.unique(agg):
            mem = set(np.flatnonzero(agg == r))
            assert len(mem) <= cap
            seen, todo = {min(mem)}, [min(mem)]
            while todo:
                x = todo.pop()
                for a, b in zip(pf, pg):
                    for u, v in ((a, b), (b, a)):
                        if u == x and v in mem and v not in seen:
                            seen.add(v)
                            todo.append(v)
            assert seen == mem


def test_precond_block_inverse():
    """A non-singular block: the plain inverse.  The Laplacian of an isolated
    group (constant null vector): a symmetric generalized inverse, B X B = B."""
    from oracle import precond as P
    rng = np.random.default_rng(2)
    A = rng.standard_normal((6, 6))
    B = A @ A.T + 0.5 * np.eye(6)
    assert np.abs(P.block_inverse(B) - np.linalg.inv(B)).max() < 1e-10 * np.abs(np.linalg.inv(B)).max()
    Wm = rng.random((5, 5))
    Wm = np.triu(Wm, 1)
    Wm = Wm + Wm.T
    Lap = np.diag(Wm.sum(1)) - Wm
    X = P.block_inverse(Lap)
    assert np.abs(X - X.T).max() < 1e-10
    assert np.abs(Lap @ X @ Lap - Lap).max() < 1e-9 * np.abs(Lap).max()
    assert np.linalg.eigvalsh(X).min() > -1e-10


def test_lp_trained_parameters_shipped():
    """The package's trained network parameters load and match the network's size."""
    import xfe_amd
    th = xfe_amd.default_lp_params()
    assert th.size == xfe_amd.lp_param_count() and np.isfinite(th).all()


def test_boot_gram_exact_pins():
    """The bootstrap refit's Gram: a worker drawn k times equals k copies of
    its rows (the replicated panel, unweighted), and the weighted projection
    equals the C restatement's weighted absorber (pinned to the reference)."""
    rng = np.random.default_rng(5)
    Nw, T, J = 120, 5, 15
    worker = np.repeat(np.arange(Nw), T)
    firm = rng.integers(0, J, size=Nw * T)
    V = rng.normal(size=(Nw * T, 3))
    mult = rng.multinomial(int(0.9 * Nw), np.ones(Nw) / Nw)
    G = oracle.boot_gram_exact(worker, firm, V, mult)
    rep = np.repeat(np.arange(Nw * T), mult[worker])
    Gr = oracle.boot_gram_exact(worker[rep], firm[rep], V[rep], np.ones(Nw, dtype=int))
    assert np.abs(G - Gr).max() <= 1e-10 * np.abs(G).max()
    # weighted absorber (C restatement) on the rows that stay, tight tolerance
    keep = mult[worker] > 0
    wk, fk = worker[keep], firm[keep]
    gw = np.unique(wk, return_inverse=True)[1].astype(np.int32)
    gf = np.unique(fk, return_inverse=True)[1].astype(np.int32)
    wts = mult[wk].astype(np.float64)
    Vt, info = oracle.absorb(V[keep], [gw, gf], [int(gw.max()) + 1, int(gf.max()) + 1], weights=wts, tol=1e-14)
    Ga = (Vt * wts[:, None]).T @ Vt
    assert np.abs(G - Ga).max() <= 1e-9 * np.abs(G).max()


def _exact_fixture():
    import os
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "exact_leverages.npz"))


def test_exact_leverages_match_reference():
    """The oracle's exact leverages (one dense two-way solve per match) equal
    the reference's exact path (LeverageMethod::Exact, fixture): P_ii to
    rounding; the B_ii weights through the components they enter, AGSU =
    plug-in - sigma2_ho sum(B) / dof and KSS = plug-in - sum(B sigma_i) / dof;
    and sum P_ii = N + J - 1, the rank of the match-level design."""
    fx = _exact_fixture()
    wu, wi = np.unique(fx["row_worker"], return_inverse=True)
    fu, fi = np.unique(fx["row_firm"], return_inverse=True)
    st = oracle.akm_build(wi, fi, fx["row_weight"], wu.size, fu.size)
    P, Bfe, Bpe, Bcov = oracle.akm_exact_leverages(st)
    assert np.abs(P - fx["pii"]).max() <= 1e-12
    assert abs(P.sum() - (st["N"] + st["J"] - 1)) <= 1e-9
    comp, sig = fx["components"], fx["sigma_i"]
    dof = comp[10] - 1.0
    for i, B in enumerate((Bpe, Bfe, Bcov)):
        want_sum = (comp[i] - comp[3 + i]) * dof / comp[9]
        want_sig = (comp[i] - comp[6 + i]) * dof
        assert abs(B.sum() - want_sum) <= 1e-12 * abs(want_sum)
        assert abs(B @ sig - want_sig) <= 1e-12 * abs(want_sig)


def test_exact_leverages_brute_force():
    """P_ii is the diagonal of the hat matrix A (A'A)^+ A' of the match-level
    design (rows sqrt(c)(e_w + e_f)), and Bfe the c-scaled D_f-weighted
    variance of the firm part of (A'A)^+ x_m, both from a pseudo-inverse of
    the full normal matrix (no grounding: the weighted variance does not see
    the shift between the worker and firm parts)."""
    rng = np.random.default_rng(5)
    N, J = 30, 8
    # a connected panel: a chain through the firms, then random extra matches
    mw = list(range(N))
    mf = [w % J for w in range(N)]
    for f in range(J - 1):
        mw.append(f)
        mf.append(f + 1)
    for _ in range(40):
        mw.append(int(rng.integers(0, N)))
        mf.append(int(rng.integers(0, J)))
    key = np.unique(np.array(mw, dtype=np.int64) * J + np.array(mf))
    mw, mf = (key // J).astype(np.int32), (key % J).astype(np.int32)
    c = rng.integers(1, 5, mw.size).astype(np.float64)
    st = oracle.akm_build(mw, mf, c, N, J)
    A = np.zeros((mw.size, N + J))
    sq = np.sqrt(c)
    A[np.arange(mw.size), mw] = sq
    A[np.arange(mw.size), N + mf] = sq
    Kp = np.linalg.pinv(A.T @ A)
    P, Bfe, Bpe, Bcov = oracle.akm_exact_leverages(st)
    assert np.allclose(P, np.diag(A @ Kp @ A.T), atol=1e-10)
    n_py = c.sum()
    for m in range(mw.size):
        e = np.zeros(N + J)
        e[mw[m]] = 1.0
        e[N + mf[m]] = 1.0
        zf = (Kp @ e)[N:]
        want = c[m] * (st["Df"] @ zf ** 2 - (st["Df"] @ zf) ** 2 / n_py)
        assert abs(Bfe[m] - want) <= 1e-9 * max(1.0, abs(want))
