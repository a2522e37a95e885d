"""CPU oracle for the fixed-effects absorber -- TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg, always as the checker.  The product (the HIP library
under ``xhdfe-xfe_amd/``) never imports it and never falls back to it.

* ``liboracle.so`` (oracle.c): plain-C restatement of the reference CPU code
  (citations per function in oracle.c).
* ``_ref/libxhdfe_ref.so``: the reference's own CPU core compiled from
  /root/reference by ``oracle/Makefile`` (present only where the reference
  is; used to pin the restatement).
* ``boot_gram_exact``: the Gram of the weighted two-way within transform of a
  cluster-bootstrap refit (rows weighted by their worker's multiplicity, the
  reference's weighted absorber semantics, src/fe_absorption.cpp:942-1037)
  by a dense weighted projection; pinned against the explicitly replicated
  panel (a worker drawn k times is k copies of its rows) and against the C
  restatement's weighted absorber.
* ``ols_cluster``: numpy restatement of the OLS + one-way cluster VCV
  (src/ols.cpp:1237-1556, scale G/(G-1) (N-1)/df_resid at :682-687).
* ``akm_build`` / ``akm_solve_S``: numpy restatement of the AKM/KSS grounded
  firm system and its Jacobi-PCG (src/akm_kss.cpp:1042-1091,
  src/akm_kss_cuda.cu:532-695); pinned by a dense grounded solve.
* ``certify``: numpy restatement of the normal-equation certificate
  (src/fe_absorption.cpp:14942-15190); pinned against the reference's own
  certificate (``ref_certify``) and a dense ||D'W v|| on tiny cases.
* ``schwarz_rule``: the stopping rule of schwarz_demean_raw
  (src/schwarz_demean.cpp:283-378) evaluated on a result; pinned against the
  reference's own convergence decisions (``ref_schwarz``).
* ``slope_projection``: the exact within transform with a heterogeneous slope
  term (the fixed point of slope_project_inplace, src/fe_absorption.cpp:
  4096-4300) by dense least squares; pinned against the reference's slope
  absorber (``ref_absorb_slope``) and per-level polyfit.
* ``recover_fixed_effects`` / ``fe_recovery_max_delta``: numpy restatement of
  the savefe recovery's Gauss-Seidel phase with coefficient accumulation and
  its max-delta check (src/fe_absorption.cpp:14447-14935, :15563-15700);
  pinned against the reference's own recovery (``ref_recover_fe``) on the
  golden panels and against the exact fitted component partial - M_D partial.
* ``absorb_group_individual`` / ``gi_colour`` / ``certify_group_individual``:
  numpy restatement of the group()/individual() absorber
  (src/fe_absorption_cuda.cu:4740-5219, colouring :4913-4947; certificate
  src/fe_absorption.cpp:15192-15300 with :15092-15190); pinned against the
  reference's CPU absorber (``ref_absorb_gi``) and the exact weighted least
  squares projection ``gi_exact``.
* ``cert_shadow``: numpy restatement of the shadow certificate's brackets and
  decision (src/fe_absorption_cuda_certificate.cu:27-101 decision contract);
  pinned by exact rational arithmetic (``cert_shadow_exact``: the brackets
  contain the exact Res_sq, Op_sq, Orig_sq) and by the reference's own
  certificate values of its outputs (its to-nearest rel).
* ``articulation_points``: restatement of the reference's iterative Tarjan
  depth-first search (src/akm_kss.cpp:444-493); pinned by brute force (remove
  each vertex, count components) on small graphs.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.csgraph  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

_i = ctypes.c_int
_d = ctypes.c_double
_p = ctypes.c_void_p


class _Opts(ctypes.Structure):
    _fields_ = [("tol", _d), ("tolerance_mode", _i), ("max_iter", _i), ("symmetric", _i),
                ("force_cg", _i), ("threads", _i), ("sweep_order", _p)]


class _Res(ctypes.Structure):
    _fields_ = [("iterations", _i), ("converged", _i), ("used_accel", _i), ("used_cg", _i),
                ("polish_sweeps", _i), ("polish_final_max", _d), ("sweep_order", _i * 8)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-C", _HERE, "oracle"], check=True, capture_output=True)
        _lib = ctypes.CDLL(path)
    return _lib


def ref():
    """The compiled reference CPU core, or None where it is not built."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libxhdfe_ref.so")
        if not os.path.exists(path):
            return None
        _ref = ctypes.CDLL(path)
    return _ref


def _ptr(a):
    return a.ctypes.data_as(_p) if a is not None else None


def build_indexer(raw):
    raw = np.ascontiguousarray(raw, dtype=np.int32)
    gid = np.empty_like(raw)
    G, lp, cont = _i(), _i(), _i()
    lib().or_build_indexer(_i(raw.size), _ptr(raw), _ptr(gid), ctypes.byref(G), ctypes.byref(lp),
                           ctypes.byref(cont))
    return gid, G.value, lp.value, bool(cont.value)


def weight_sums(gid, G, weights=None):
    """compute_weight_sums (src/fe_absorption.cpp:942-1037): level sums of the
    weights, or counts."""
    g = np.ascontiguousarray(gid, dtype=np.int32)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    out = np.empty(int(G))
    lib().or_weight_sums(_i(g.size), _ptr(g), _i(int(G)), _ptr(w), _ptr(out))
    return out


def absorb(V, gids, G, weights=None, tol=1e-8, tolerance_mode=1, max_iter=100000,
           symmetric=None, force_cg=False, threads=None, sweep_order=None):
    """Within transform of the n x C block V (column-major [y | X]).

    gids: list of compact id arrays; G: level counts.  ``symmetric`` defaults to
    the reference's method choice (SGS for two FEs, GS otherwise;
    src/fe_absorption.cpp:5568-5589)."""
    V = np.array(V, dtype=np.float64, order="F", copy=True)
    n, C = V.shape
    D = len(gids)
    ids = np.ascontiguousarray(np.stack([np.asarray(g, dtype=np.int32) for g in gids])) if D else \
        np.zeros((0, n), dtype=np.int32)
    Garr = np.asarray(G, dtype=np.int32)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    if symmetric is None:
        symmetric = D == 2
    order = None if sweep_order is None else np.asarray(sweep_order, dtype=np.int32)
    o = _Opts(tol, tolerance_mode, max_iter, int(bool(symmetric)), int(bool(force_cg)),
              threads or min(8, os.cpu_count() or 1), _ptr(order) if order is not None else None)
    r = _Res()
    rc = lib().or_absorb(_i(n), _i(C), _ptr(V), _i(D), _ptr(ids), _ptr(Garr), _ptr(w),
                         ctypes.byref(o), ctypes.byref(r))
    if rc != 0:
        raise RuntimeError("or_absorb failed")
    info = dict(iterations=r.iterations, converged=bool(r.converged), used_accel=bool(r.used_accel),
                used_cg=bool(r.used_cg), polish_sweeps=r.polish_sweeps,
                polish_final_max=r.polish_final_max, sweep_order=list(r.sweep_order)[:D])
    return V, info


def project(V, gid, G, weights=None):
    V = np.array(V, dtype=np.float64, order="F", copy=True)
    n, C = V.shape
    gid = np.ascontiguousarray(gid, dtype=np.int32)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    lib().or_project(_i(n), _i(C), _ptr(V), _ptr(gid), _i(G), _ptr(w))
    return V


def drop_singletons(fes, max_rounds=100):
    fes = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.int32) for f in fes]))
    D, n = fes.shape
    keep = np.empty(n, dtype=np.uint8)
    rounds = _i()
    dropped = lib().or_drop_singletons(_i(n), _i(D), _ptr(fes), _i(max_rounds), _ptr(keep),
                                       ctypes.byref(rounds))
    return keep.astype(bool), dropped, rounds.value


def components(a, b, expected_a=-1, expected_b=-1, labels=True):
    a = np.ascontiguousarray(a, dtype=np.int32)
    b = np.ascontiguousarray(b, dtype=np.int32)
    lab = np.empty(a.size, dtype=np.int32) if labels else None
    c = lib().or_components(_i(a.size), _ptr(a), _ptr(b), _i(expected_a), _i(expected_b), _ptr(lab))
    return c, lab


def dof_reghdfe(fes, levels):
    fes = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.int32) for f in fes]))
    D, n = fes.shape
    lv = np.asarray(levels, dtype=np.int32)
    red = np.empty(D, dtype=np.int32)
    nc = np.empty(D, dtype=np.int32)
    df = lib().or_dof_reghdfe(_i(n), _i(D), _ptr(fes), _ptr(lv), _ptr(red), _ptr(nc))
    return df, red, nc


def fe_is_function_of(coarse, refined):
    coarse = np.ascontiguousarray(coarse, dtype=np.int32)
    refined = np.ascontiguousarray(refined, dtype=np.int32)
    return bool(lib().or_fe_is_function_of(_i(coarse.size), _ptr(coarse), _ptr(refined)))


def fit_dof(fes, levels, cluster=None):
    """The estimator's FE degrees of freedom (src/hdfe_regressor_v11.cpp:9540-9700,
    no slopes): FEs nested in the cluster variable (fe_nested_within_cluster,
    :6765 -- the cluster id is a function of the FE level) are fully redundant,
    pairwise mobility groups over the others (compute_fe_dof_reghdfe), then one
    redundant level for every intercept dimension except the first when none is
    nested.  Returns (df_a, redundant, num_coefs)."""
    D = len(fes)
    levels = [int(x) for x in levels]
    nested = [cluster is not None and fe_is_function_of(cluster, fes[d]) for d in range(D)]
    red = [0] * D
    nn = [d for d in range(D) if not nested[d]]
    if nn:
        _, r2, _ = dof_reghdfe([fes[d] for d in nn], [levels[d] for d in nn])
        for q, d in enumerate(nn):
            red[d] = int(r2[q])
    nested_total = sum(levels[d] for d in range(D) if nested[d])
    for d in range(D):
        if nested[d]:
            red[d] = levels[d]
    skip_next = nested_total == 0
    for d in nn:
        if red[d] >= levels[d]:
            continue
        if skip_next:
            skip_next = False
            continue
        red[d] = max(red[d], 1)
    nc = [max(0, levels[d] - red[d]) for d in range(D)]
    return sum(nc), np.array(red), np.array(nc)


def fe_collinear(X_raw, X_t):
    """Columns omitted as collinear with the FEs: ||M_D x||^2 <= 1e-9 ||x||^2
    (src/hdfe_regressor_v11.cpp:8668-8700)."""
    raw = (np.asarray(X_raw) ** 2).sum(axis=0)
    til = (np.asarray(X_t) ** 2).sum(axis=0)
    return np.where(raw <= 0.0, til <= 1e-9, til <= 1e-9 * raw)


def ols_cluster(y_t, X_t, cluster, df_a, omit=None):
    """OLS on the within-transformed data with a one-way cluster VCV.

    V = (X'X)^-1 (sum_g s_g s_g') (X'X)^-1 * G/(G-1) * (N-1)/df_r
    (src/ols.cpp:657-687) with the estimator's residual df
    df_r = N - k - 1 - df_a (intercept dof, src/hdfe_regressor_v11.cpp:9821-9824)."""
    X_all = np.asarray(X_t, dtype=np.float64)
    y = np.asarray(y_t, dtype=np.float64)
    keep_cols = np.ones(X_all.shape[1], bool) if omit is None else ~np.asarray(omit)
    X = X_all[:, keep_cols]
    n, k = X.shape
    xtx = X.T @ X
    beta = np.linalg.solve(xtx, X.T @ y)
    e = y - X @ beta
    _, inv_c = np.unique(cluster, return_inverse=True)
    G = int(inv_c.max()) + 1
    S = np.zeros((G, k))
    np.add.at(S, inv_c, X * e[:, None])
    meat = S.T @ S
    bread = np.linalg.inv(xtx)
    df_resid = n - k - 1 - df_a
    V = bread @ meat @ bread * (G / (G - 1.0)) * ((n - 1.0) / df_resid)
    beta_all = np.zeros(X_all.shape[1])
    se_all = np.full(X_all.shape[1], np.nan)  # omitted: missing SE, as the estimator reports
    beta_all[keep_cols] = beta
    se_all[keep_cols] = np.sqrt(np.diag(V))
    return beta_all, se_all, e, G


# ------------------------------------------------------------ reference checker

def effective_tolerance(tol, tolerance_mode):
    """effective_absorption_tolerance (src/fe_absorption.cpp:339-345)."""
    if tol > 0.0 and tolerance_mode == 1:
        return min(tol, 1.0e-9)
    return tol


def certify(V, Vt, gids, G, weights=None, tol=1e-8, tolerance_mode=1, converged=True):
    """certify_absorption_result (src/fe_absorption.cpp:14942-15190) for
    standard FEs: per column c, abs_c = ||D'W vt_c|| over all dimensions
    (level sums of w * vt, squared and summed), ||D'W||_F^2 = sum_d sum_i w_i^2,
    rel_c = abs_c / (||D'W||_F ||v_c||) with v the untransformed column
    (unweighted norm); certified iff converged and max_c rel_c <=
    max(8 tol_eff, 64 eps) (:15181-15189).  Accumulates squares in extended
    precision like the reference (long double)."""
    V = np.asarray(V, dtype=np.float64)
    Vt = np.asarray(Vt, dtype=np.float64)
    if V.ndim == 1:
        V, Vt = V[:, None], Vt[:, None]
    n, C = V.shape
    w = np.ones(n) if weights is None else np.asarray(weights, dtype=np.float64)
    orig = (V.astype(np.longdouble) ** 2).sum(axis=0)
    resid = np.zeros(C, dtype=np.longdouble)
    op = np.longdouble(0.0)
    for g, Gd in zip(gids, G):
        g = np.asarray(g)
        op += (w.astype(np.longdouble) ** 2).sum()
        for c in range(C):
            sums = np.bincount(g, weights=w * Vt[:, c], minlength=Gd)
            resid[c] += (sums.astype(np.longdouble) ** 2).sum()
    absolute = np.sqrt(np.maximum(resid, 0.0)).astype(np.float64)
    scale = float(np.sqrt(max(op, 0.0))) * np.sqrt(np.maximum(orig, 0.0)).astype(np.float64)
    rel = np.where(scale > 0.0, absolute / np.where(scale > 0.0, scale, 1.0),
                   np.where(absolute > 0.0, np.finfo(np.float64).max, 0.0))
    tol_eff = effective_tolerance(tol, tolerance_mode)
    limit = max(8.0 * max(0.0, tol_eff), 64.0 * np.finfo(np.float64).eps)
    return dict(abs_residual=float(absolute.max()), abs_residual_rel=float(rel.max()),
                certified=bool(converged and rel.max() <= limit), per_column_rel=rel)


def ref_certify(y, X, yt, Xt, fes, weights=None, tol=1e-8, tolerance_mode=1, converged=True):
    """The reference's certify_absorption_result through oracle/_ref (checker)."""
    r = ref()
    if r is None:
        return None
    y = np.ascontiguousarray(y, dtype=np.float64)
    X = np.asfortranarray(np.asarray(X, dtype=np.float64).reshape(y.size, -1))
    yt = np.ascontiguousarray(yt, dtype=np.float64)
    Xt = np.asfortranarray(np.asarray(Xt, dtype=np.float64).reshape(y.size, -1))
    n, k = X.shape
    ids = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.int32) for f in fes]))
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    a, rl, cert = _d(), _d(), _i()
    rc = r.ref_certify(_i(n), _i(k), _i(len(fes)), _ptr(y), _ptr(X), _ptr(ids), _ptr(w), _d(tol),
                       _i(tolerance_mode), _i(int(bool(converged))), _ptr(yt), _ptr(Xt),
                       ctypes.byref(a), ctypes.byref(rl), ctypes.byref(cert))
    if rc != 0:
        raise RuntimeError("ref_certify failed")
    return dict(abs_residual=a.value, abs_residual_rel=rl.value, certified=bool(cert.value))


def akm_build(mw, mf, mc, N, J):
    """TwoWaySolver::build (src/akm_kss.cpp:1042-1091): from matches (worker,
    firm, count) grouped by worker, the CSR offsets w_ptr, the firm offsets
    f_ptr and per-firm match lists f_matches (ascending match index), the
    degrees Dw, Df and the Jacobi inverse diagonal of the grounded S
    (diag S = Df - sum c^2 / Dw; 0 for the ground firm J-1 and non-positive
    entries)."""
    mw = np.asarray(mw, dtype=np.int32)
    mf = np.asarray(mf, dtype=np.int32)
    c = np.asarray(mc, dtype=np.float64)
    Dw = np.bincount(mw, weights=c, minlength=N)
    Df = np.bincount(mf, weights=c, minlength=J)
    w_ptr = np.concatenate([[0], np.cumsum(np.bincount(mw, minlength=N))]).astype(np.int64)
    f_ptr = np.concatenate([[0], np.cumsum(np.bincount(mf, minlength=J))]).astype(np.int64)
    f_matches = np.argsort(mf, kind="stable").astype(np.int64)
    diag = Df - np.bincount(mf, weights=c * c / Dw[mw], minlength=J)
    inv_diag = np.where((np.arange(J) < J - 1) & (diag > 0.0), 1.0 / np.where(diag > 0.0, diag, 1.0), 0.0)
    return dict(N=N, J=J, m_w=mw, m_f=mf, m_c=c, w_ptr=w_ptr, f_ptr=f_ptr, f_matches=f_matches,
                Dw=Dw, Df=Df, inv_diag=inv_diag)


def akm_apply_S(st, p):
    """Grounded S p = Df p - B'(Dw^-1 B p), S p[ground] = 0
    (src/akm_kss_cuda.cu:347-354 with the kernels at :33-57)."""
    t = np.bincount(st["m_w"], weights=st["m_c"] * p[st["m_f"]], minlength=st["N"]) / st["Dw"]
    out = st["Df"] * p - np.bincount(st["m_f"], weights=st["m_c"] * t[st["m_w"]], minlength=st["J"])
    out[st["J"] - 1] = 0.0
    return out


def akm_solve_S(st, rhs, tol, max_iter):
    """Jacobi-PCG on the grounded S for one lane (src/akm_kss_cuda.cu:532-695,
    the per-lane rules of akm_cuda_solve_S_multi): ground entry of rhs pinned
    to 0, x = 0, stop when r.r <= tol^2 |rhs|^2 (after the update of that
    iteration), give up on pAp outside (0, 1e300) or after max_iter iterations.
    Returns (z, iterations) with z = 0 and iterations = -1 when not converged."""
    J = st["J"]
    r = np.array(rhs, dtype=np.float64)
    r[J - 1] = 0.0
    x = np.zeros(J)
    rhs2 = float(r @ r)
    if rhs2 == 0.0:
        return x, 0
    tol2 = tol * tol * rhs2
    z = st["inv_diag"] * r
    p = z.copy()
    rz = float(r @ z)
    its = 0
    while True:
        Ap = akm_apply_S(st, p)
        pAp = float(p @ Ap)
        if not (pAp > 0.0) or not (pAp < 1e300):
            return np.zeros(J), -1
        a = rz / pAp
        x += a * p
        r -= a * Ap
        z = st["inv_diag"] * r
        rr = float(r @ r)
        rz_new = float(r @ z)
        its += 1
        if rr <= tol2:
            x[J - 1] = 0.0
            return x, its
        if its >= max_iter:
            return np.zeros(J), -1
        p = z + (rz_new / rz) * p
        rz = rz_new


_U64 = (1 << 64) - 1


def splitmix64(x):
    """SplitMix64 (Steele, Lea and Flood 2014), as src/akm_kss.cpp:212-217,
    on numpy uint64 arrays (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        x = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def stream_seed(seed, tag, unit):
    """The per-(draw, unit) stream state of src/akm_kss.cpp:219-222."""
    with np.errstate(over="ignore"):
        a = np.uint64((0xA24BAED4963EE407 * (int(tag) + 1)) & _U64)
        s = splitmix64(np.uint64(seed) ^ a)
        u = (np.uint64(0x9FB21C651E98DF25) * (np.asarray(unit, dtype=np.uint64) + np.uint64(1)))
        return splitmix64(s ^ u)


def jla_signs(seed, draw, R):
    """The +-1 vector of draw `draw` over R working rows (src/akm_kss.cpp:
    3157-3170): 64 rows per stream state stream_seed(seed, 3 draw, block),
    row r taking bit r mod 64 (1 -> +1)."""
    nb = (R + 63) // 64
    with np.errstate(over="ignore"):
        bits = stream_seed(seed, 3 * draw, np.arange(nb, dtype=np.uint64))
    r = np.arange(R)
    b = (bits[r // 64] >> (r % 64).astype(np.uint64)) & np.uint64(1)
    return np.where(b == 1, 1.0, -1.0)


def akm_jla_pii(st, draws, seed):
    """JLA leverages of the match-level working rows (src/akm_kss.cpp:3153-3285,
    the P_ii part): per draw, r = +-1 signs, t = A' r with A the rows
    sqrt(c_m) (e_w + e_f), the two-way normal equations K [zw; zf] = t solved
    exactly (firms grounded at J-1, a dense solve of the Schur complement S),
    Z = A [zw; zf]; with P = mean Z^2 and M = mean (r - Z)^2 over the draws,
    P_ii = P / (P + M).  st: akm_build's structure (rows = matches)."""
    return akm_jla_moments(st, draws, seed)[0]


def akm_jla_moments(st, draws, seed):
    """akm_jla_pii with the reference's non-linearity correction of the JLA
    ratio estimator (src/akm_kss.cpp:3194-3206 for the moments, :3273-3288 for
    the correction): besides P = mean Z^2 and M = mean (r - Z)^2, the means of
    Z^4, (r - Z)^4 and Z^2 (r - Z)^2 (Psq, Msq, PM); with Ph = P / (P + M)
    (0 when P + M = 0) and Mh = 1 - Ph,
    V = (Mh^2 Psq + Ph^2 Msq - 2 Mh Ph PM) / p,
    B = (Mh Psq - Ph Msq + 2 (Mh - Ph) PM) / p and
    corr = 1 - V / Mh^2 + B / Mh (1 where Mh <= 0).  Returns (P_ii, corr)."""
    mw, mf, c = st["m_w"], st["m_f"], st["m_c"]
    N, J = st["N"], st["J"]
    R = mw.size
    sq = np.sqrt(c)
    S = np.zeros((J, J))
    for f in range(J):
        e = np.zeros(J)
        e[f] = 1.0
        S[:, f] = akm_apply_S(st, e)
    Sg = S[:J - 1, :J - 1]
    acc = np.zeros((5, R))
    for d in range(draws):
        r = jla_signs(seed, d, R)
        v = sq * r
        tw = np.bincount(mw, weights=v, minlength=N)
        tf = np.bincount(mf, weights=v, minlength=J)
        rhs = tf - np.bincount(mf, weights=c * (tw / st["Dw"])[mw], minlength=J)
        zf = np.zeros(J)
        zf[:J - 1] = np.linalg.solve(Sg, rhs[:J - 1])
        zw = (tw - np.bincount(mw, weights=c * zf[mf], minlength=N)) / st["Dw"]
        Z = sq * (zw[mw] + zf[mf])
        z2, m2 = Z * Z, (r - Z) ** 2
        acc += np.array([z2, m2, z2 * z2, m2 * m2, z2 * m2]) / draws
    P, M, Psq, Msq, PM = acc
    den = P + M
    Ph = np.where(den > 0.0, P / np.where(den > 0.0, den, 1.0), 0.0)
    Mh = 1.0 - Ph
    corr = np.ones(R)
    ok = Mh > 0.0
    V = (Mh * Mh * Psq + Ph * Ph * Msq - 2.0 * Mh * Ph * PM) / draws
    B = (Mh * Psq - Ph * Msq + 2.0 * (Mh - Ph) * PM) / draws
    corr[ok] = 1.0 - V[ok] / (Mh[ok] * Mh[ok]) + B[ok] / Mh[ok]
    return Ph, corr


def rademacher_sums(seed, tag, count):
    """Per unit m, the sum of count[m] +-1 draws of the stream
    stream_seed(seed, tag, m) (src/akm_kss.cpp:224-238): the state advances
    by one SplitMix64 application per 64 draws, and each advance contributes
    2 popcount(low k bits) - k."""
    count = np.asarray(count, dtype=np.int64)
    with np.errstate(over="ignore"):
        s = stream_seed(seed, tag, np.arange(count.size, dtype=np.uint64))
        acc = np.zeros(count.size)
        rem = count.copy()
        while (rem > 0).any():
            live = rem > 0
            s = np.where(live, splitmix64(s), s)
            k = np.minimum(rem, 64)
            sh = np.minimum(k, 63).astype(np.uint64)
            mask = np.where(k >= 64, np.uint64(_U64), (np.uint64(1) << sh) - np.uint64(1))
            pc = np.bitwise_count(s & mask).astype(np.float64)
            acc += np.where(live, 2.0 * pc - k, 0.0)
            rem = np.where(live, rem - k, rem)
    return acc


def _two_way_solve(st, Sg, tw, tf):
    """(zw, zf) of K [zw; zf] = [tw; tf] for the match-level normal equations,
    firms grounded at J-1 (dense Schur solve)."""
    mw, mf, c, J = st["m_w"], st["m_f"], st["m_c"], st["J"]
    rhs = tf - np.bincount(mf, weights=c * (tw / st["Dw"])[mw], minlength=J)
    zf = np.zeros(J)
    zf[:J - 1] = np.linalg.solve(Sg, rhs[:J - 1])
    zw = (tw - np.bincount(mw, weights=c * zf[mf], minlength=st["N"])) / st["Dw"]
    return zw, zf


def akm_jla_bii(st, draws, seed, count):
    """JLA estimates of the KSS quadratic-form weights B_ii of the three
    components, per match-level working row (src/akm_kss.cpp:2985-3047 and
    :3128-3146): per draw d, u_m = sum of count[m] Rademacher draws of the
    stream (seed, 3 d + 1, m) scaled by 1/sqrt(p) and demeaned over person
    years (u_m - vbar count[m], vbar = sum u / n_py), then two solves,
    fe: K z = [0; F'u] and pe: K z = [W'u; 0], with W2 = A z_fe and
    W1 = A z_pe on the rows (A rows sqrt(c)(e_w + e_f));
    Bfe += W2^2, Bpe += W1^2, Bcov += W1 W2.  count: person-year rows per
    match (the match weights when unweighted); n_py = sum of the weights."""
    mw, mf, c = st["m_w"], st["m_f"], st["m_c"]
    N, J = st["N"], st["J"]
    R = mw.size
    count = np.asarray(count, dtype=np.int64)
    sq = np.sqrt(c)
    S = np.zeros((J, J))
    for f in range(J):
        e = np.zeros(J)
        e[f] = 1.0
        S[:, f] = akm_apply_S(st, e)
    Sg = S[:J - 1, :J - 1]
    scale = 1.0 / np.sqrt(float(draws))
    n_py = float(c.sum())
    bfe, bpe, bcov = np.zeros(R), np.zeros(R), np.zeros(R)
    for d in range(draws):
        ms = rademacher_sums(seed, 3 * d + 1, count)
        vbar = (float(ms.sum()) / n_py) * scale
        u = ms * scale - vbar * count
        tw = np.bincount(mw, weights=u, minlength=N)
        tf = np.bincount(mf, weights=u, minlength=J)
        zw, zf = _two_way_solve(st, Sg, np.zeros(N), tf)
        W2 = sq * (zw[mw] + zf[mf])
        zw, zf = _two_way_solve(st, Sg, tw, np.zeros(J))
        W1 = sq * (zw[mw] + zf[mf])
        bfe += W2 * W2
        bpe += W1 * W1
        bcov += W1 * W2
    return bfe, bpe, bcov


def ref_akm_exact(y, worker, firm):
    """The reference's akm_kss_decompose with exact leverages at match level
    (no pruning) through oracle/_ref (checker): (row_worker, row_firm,
    row_weight, pii, comp, sigma_i), comp as in ref_akm_jla_components."""
    r = ref()
    if r is None:
        return None
    n = len(y)
    y = np.ascontiguousarray(y, dtype=np.float64)
    w = np.ascontiguousarray(worker, dtype=np.int32)
    f = np.ascontiguousarray(firm, dtype=np.int32)
    rw = np.empty(n, dtype=np.int32)
    rf = np.empty(n, dtype=np.int32)
    rc = np.empty(n)
    pii = np.empty(n)
    sig = np.empty(n)
    comp = np.empty(11)
    r.ref_akm_exact.argtypes = [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p]
    R = r.ref_akm_exact(n, _ptr(y), _ptr(w), _ptr(f), _ptr(rw), _ptr(rf), _ptr(rc), _ptr(pii), _ptr(comp),
                        _ptr(sig))
    if R < 0:
        raise RuntimeError("ref_akm_exact failed")
    return rw[:R], rf[:R], rc[:R], pii[:R], comp, sig[:R]


def akm_exact_leverages(st):
    """The exact path of akm_kss_decompose (src/akm_kss.cpp:2667-2741, match
    level): per match m, K [zw; zf] = [e_w; e_f] solved exactly (firms
    grounded at J-1, dense Schur solve); P_ii = c (zw[w] + zf[f]); with
    s1 = sum Df zf^2, s2 = sum Df zf, s3 = sum Dw zw^2, s4 = sum Dw zw and
    cross = sum_k c_k zw[w_k] zf[f_k]: Bfe = c (s1 - s2^2 / n_py), Bpe =
    c (s3 - s4^2 / n_py), Bcov = c (cross - s4 s2 / n_py), n_py = sum c."""
    mw, mf, c = st["m_w"], st["m_f"], st["m_c"]
    N, J = st["N"], st["J"]
    S = np.zeros((J, J))
    for f in range(J):
        e = np.zeros(J)
        e[f] = 1.0
        S[:, f] = akm_apply_S(st, e)
    Sg = S[:J - 1, :J - 1]
    n_py = float(c.sum())
    Dw, Df = st["Dw"], st["Df"]
    out = np.zeros((4, mw.size))
    for m in range(mw.size):
        tw = np.zeros(N)
        tf = np.zeros(J)
        tw[mw[m]] = 1.0
        tf[mf[m]] = 1.0
        zw, zf = _two_way_solve(st, Sg, tw, tf)
        s1, s2 = float(Df @ (zf * zf)), float(Df @ zf)
        s3, s4 = float(Dw @ (zw * zw)), float(Dw @ zw)
        cross = float(np.sum(c * zw[mw] * zf[mf]))
        out[:, m] = c[m] * np.array([zw[mw[m]] + zf[mf[m]], s1 - s2 * s2 / n_py, s3 - s4 * s4 / n_py,
                                     cross - s4 * s2 / n_py])
    return out[0], out[1], out[2], out[3]


def ref_akm_jla_components(y, worker, firm, draws, seed):
    """The reference's akm_kss_decompose (JLA, match level, no pruning)
    through oracle/_ref (checker): (comp, sigma_i) with comp = plug-in, AGSU
    and KSS (var_alpha, var_psi, cov_alpha_psi) each, sigma2_ho, n_py."""
    r = ref()
    if r is None:
        return None
    n = len(y)
    y = np.ascontiguousarray(y, dtype=np.float64)
    w = np.ascontiguousarray(worker, dtype=np.int32)
    f = np.ascontiguousarray(firm, dtype=np.int32)
    comp = np.empty(11)
    sig = np.empty(n)
    r.ref_akm_jla_components.argtypes = [_i, _p, _p, _p, _i, ctypes.c_ulonglong, _p, _p]
    R = r.ref_akm_jla_components(n, _ptr(y), _ptr(w), _ptr(f), int(draws), int(seed), _ptr(comp), _ptr(sig))
    if R < 0:
        raise RuntimeError("ref_akm_jla_components failed")
    return comp, sig[:R]


def _level_means(r, gid, G, w):
    if w is None:
        return np.bincount(gid, weights=r, minlength=G) / np.maximum(np.bincount(gid, minlength=G), 1)
    ws = np.bincount(gid, weights=w, minlength=G)
    return np.where(ws > 0, np.bincount(gid, weights=w * r, minlength=G) / np.where(ws > 0, ws, 1.0), 0.0)


def fe_recovery_max_delta(residual, gids, G, weights=None, order=None):
    """fe_recovery_max_delta (src/fe_absorption.cpp:15563-15700): one
    Gauss-Seidel pass over the residual in sweep order; the largest absolute
    weighted level mean met on the way."""
    r = np.array(residual, dtype=np.float64)
    mx = 0.0
    for d in (order if order is not None else range(len(gids))):
        m = _level_means(r, gids[d], G[d], weights)
        mx = max(mx, float(np.abs(m).max()) if m.size else 0.0)
        r -= m[gids[d]]
    return mx


def recover_fixed_effects(partial, gids, G, weights=None, fe_tol=1e-6, max_iter=100000, order=None):
    """The savefe recovery (recover_fixed_effects_impl,
    src/fe_absorption.cpp:14447-14935, its plain Gauss-Seidel phase): from
    alpha = 0, sweep the dimensions in order, each taking the weighted level
    means m of the current residual partial - sum_d alpha_d[ids_d] into
    alpha_d; stop once the largest |m| of a sweep is below fe_tol.  Returns
    (alphas, iterations, max_delta, converged).  The reference's accelerated
    phases change the trajectory, not the fixed point D alpha = partial -
    M_D partial."""
    r = np.array(partial, dtype=np.float64)
    D = len(gids)
    alphas = [np.zeros(int(G[d])) for d in range(D)]
    seq = list(order if order is not None else range(D))
    mx = 0.0
    for it in range(1, max_iter + 1):
        mx = 0.0
        for d in seq:
            m = _level_means(r, gids[d], G[d], weights)
            alphas[d] += m
            r -= m[gids[d]]
            mx = max(mx, float(np.abs(m).max()) if m.size else 0.0)
        if mx < fe_tol:
            return alphas, it, mx, True
    return alphas, max_iter, mx, False


def gi_colour(gi):
    """Greedy colouring of the individuals in index order
    (src/fe_absorption_cuda.cu:4913-4947): (colour per individual, classes as
    lists of individuals in index order)."""
    m = int(gi["num_individuals"])
    gp, gind = gi["group_ptr"], gi["group_individual"]
    ip, ig = gi["individual_ptr"], gi["individual_group"]
    colour = np.zeros(m, dtype=np.int64)
    for j in range(m):
        taken = set()
        for g in ig[ip[j]:ip[j + 1]]:
            for o in gind[gp[g]:gp[g + 1]]:
                if o < j:
                    taken.add(int(colour[o]))
        c = 0
        while c in taken:
            c += 1
        colour[j] = c
    classes = [np.flatnonzero(colour == c) for c in range(int(colour.max()) + 1)]
    return colour, classes


def gi_design(n, gi):
    """The individual design A (n x individuals, A[g, j] = group_scale[g])."""
    gp = np.asarray(gi["group_ptr"])
    rows = np.repeat(np.arange(n), np.diff(gp))
    A = np.zeros((n, int(gi["num_individuals"])))
    A[rows, gi["group_individual"]] = np.asarray(gi["group_scale"])[rows]
    return A


def gi_exact(V, gids, G, gi, weights=None):
    """The within transform as dense weighted least squares: the residual of
    each column of V on [D_std, A] (small cases)."""
    V = np.asarray(V, dtype=np.float64).reshape(len(gids[0]) if gids else len(gi["group_scale"]), -1)
    n = V.shape[0]
    blocks = [np.eye(int(G[d]))[gids[d]] for d in range(len(gids))] + [gi_design(n, gi)]
    M = np.hstack(blocks)
    sw = np.ones(n) if weights is None else np.sqrt(np.asarray(weights, dtype=np.float64))
    coef = np.linalg.lstsq(M * sw[:, None], V * sw[:, None], rcond=None)[0]
    return V - M @ coef


def certify_group_individual(V, Vt, gids, G, gi, weights=None, tol=1e-8, tolerance_mode=1):
    """The certificate with the individual dimension
    (src/fe_absorption.cpp:15092-15190): (abs, rel, certified)."""
    V = np.asarray(V, dtype=np.float64)
    Vt = np.asarray(Vt, dtype=np.float64)
    n, C = Vt.shape
    w = np.ones(n) if weights is None else np.asarray(weights, dtype=np.float64)
    op = 0.0
    abs2 = np.zeros(C)
    for d in range(len(gids)):
        op += float((w * w).sum())
        for c in range(C):
            abs2[c] += float((np.bincount(gids[d], weights=w * Vt[:, c], minlength=int(G[d])) ** 2).sum())
    gp = np.asarray(gi["group_ptr"])
    rows = np.repeat(np.arange(n), np.diff(gp))
    mult = (w * np.asarray(gi["group_scale"]))[rows]
    op += float((mult * mult).sum())
    m = int(gi["num_individuals"])
    for c in range(C):
        abs2[c] += float((np.bincount(gi["group_individual"], weights=mult * Vt[rows, c], minlength=m) ** 2).sum())
    a = np.sqrt(abs2)
    rel = a / (np.sqrt(op) * np.sqrt((V * V).sum(0)))
    lim = max(8.0 * max(0.0, effective_tolerance(tol, tolerance_mode)), 64.0 * np.finfo(float).eps)
    return float(a.max()), float(rel.max()), bool(rel.max() <= lim)


def absorb_group_individual(V, gids, G, gi, weights=None, tol=1e-8, tolerance_mode=1, max_iter=100000,
                            method=1, sweep_order=None):
    """The group()/individual() absorber (absorb_fixed_effects_group_individual_cuda,
    src/fe_absorption_cuda.cu:4740-5219): sweeps of the standard dimensions
    (weighted level means) in order, then the individuals colour class by
    colour class (a_j = sum_g w_g s_g v_g / sum_g w_g s_g^2, v_g -= s_g a_j),
    mirrored for method 2; stop when |norm - previous norm| / max(1,
    previous) < min(tol_eff, 1e-12) (reghdfe-comparable), then 16 polish
    sweeps and the certificate, with 64 more sweeps per failed certificate.
    Returns (V~, iterations, certified)."""
    V0 = np.asarray(V, dtype=np.float64)
    V0 = V0.reshape(V0.shape[0], -1)
    Vt = V0.copy()
    n, C = Vt.shape
    w = np.ones(n) if weights is None else np.asarray(weights, dtype=np.float64)
    order = list(range(len(gids)) if sweep_order is None else sweep_order)
    wsum = [np.bincount(gids[d], weights=w, minlength=int(G[d])) for d in range(len(gids))]
    scale = np.asarray(gi["group_scale"], dtype=np.float64)
    ip, ig = gi["individual_ptr"], gi["individual_group"]
    m = int(gi["num_individuals"])
    owner_all = np.repeat(np.arange(m), np.diff(ip))
    denom = np.bincount(owner_all, weights=w[ig] * scale[ig] ** 2, minlength=m)
    denom = np.where(denom > 0, denom, 1.0)
    _, classes = gi_colour(gi)
    plans = []
    for cls in classes:
        pos = np.concatenate([np.arange(ip[j], ip[j + 1]) for j in cls]) if len(cls) else np.zeros(0, int)
        own = np.repeat(np.arange(len(cls)), [ip[j + 1] - ip[j] for j in cls])
        plans.append((cls, ig[pos], own))

    def std(d):
        for c in range(C):
            s = np.bincount(gids[d], weights=w * Vt[:, c], minlength=int(G[d]))
            Vt[:, c] -= (s / wsum[d])[gids[d]]

    def indiv():
        for cls, rows, own in plans:
            for c in range(C):
                num = np.bincount(own, weights=w[rows] * scale[rows] * Vt[rows, c], minlength=len(cls))
                a = num / denom[cls]
                Vt[rows, c] -= scale[rows] * a[own]

    def sweep():
        for d in order:
            std(d)
        indiv()
        if method == 2:
            indiv()
            for d in reversed(order):
                std(d)

    tol_eff = effective_tolerance(tol, tolerance_mode)
    gi_tol = min(tol_eff, 1e-12) if (tol_eff > 0 and tolerance_mode == 1) else tol_eff
    prev = float(np.sqrt((Vt * Vt).sum()))
    it = 0
    conv = False
    for k in range(max_iter):
        sweep()
        cur = float(np.sqrt((Vt * Vt).sum()))
        rel = abs(cur - prev) / max(1.0, prev)
        prev = cur
        if rel < gi_tol:
            it, conv = k + 1, True
            break
    it = max(1, it) if conv else max_iter
    if conv and it < max_iter:
        extra = min(16, max_iter - it)
        for _ in range(extra):
            sweep()
        it += extra
    cert = certify_group_individual(V0, Vt, gids, G, gi, weights, tol, tolerance_mode)[2]
    while not cert and it < max_iter:
        b = min(64, max_iter - it)
        for _ in range(b):
            sweep()
        it += b
        cert = certify_group_individual(V0, Vt, gids, G, gi, weights, tol, tolerance_mode)[2]
    return Vt, it, cert


def articulation_points(adj_ptr, adj):
    """articulation_points (src/akm_kss.cpp:444-493): iterative depth-first
    search from every undiscovered vertex in index order, disc/low times,
    a non-root u with a child c of low(c) >= disc(u) is a cut vertex, a root
    with two or more children is one."""
    adj_ptr = np.asarray(adj_ptr)
    adj = np.asarray(adj)
    n = adj_ptr.size - 1
    artic = np.zeros(n, dtype=np.uint8)
    disc = np.full(n, -1, dtype=np.int64)
    low = np.zeros(n, dtype=np.int64)
    parent = np.full(n, -1, dtype=np.int64)
    child = np.zeros(n, dtype=np.int64)
    it = np.zeros(n, dtype=np.int64)
    timer = 0
    for root in range(n):
        if disc[root] != -1:
            continue
        stack = [root]
        disc[root] = low[root] = timer
        timer += 1
        it[root] = adj_ptr[root]
        while stack:
            u = stack[-1]
            if it[u] < adj_ptr[u + 1]:
                v = int(adj[it[u]])
                it[u] += 1
                if disc[v] == -1:
                    parent[v] = u
                    child[u] += 1
                    disc[v] = low[v] = timer
                    timer += 1
                    it[v] = adj_ptr[v]
                    stack.append(v)
                elif v != parent[u]:
                    low[u] = min(low[u], disc[v])
            else:
                stack.pop()
                p = parent[u]
                if p != -1:
                    low[p] = min(low[p], low[u])
                    if p != root and low[u] >= disc[p]:
                        artic[p] = 1
        if child[root] >= 2:
            artic[root] = 1
    return artic


def ref_leave_out_set(worker, firm):
    """The reference's leave_out_connected_set through oracle/_ref (checker):
    (keep, info) with info = n_obs_input, n_obs_connected, n_obs, n_workers,
    n_firms, n_matches, n_movers, n_stayers, prune_iterations."""
    r = ref()
    if r is None:
        return None
    n = len(worker)
    w = np.ascontiguousarray(worker, dtype=np.int32)
    f = np.ascontiguousarray(firm, dtype=np.int32)
    keep = np.zeros(max(n, 1), dtype=np.uint8)
    info = np.zeros(9, dtype=np.int64)
    r.ref_leave_out_set.argtypes = [_i, _p, _p, _p, _p]
    if r.ref_leave_out_set(n, _ptr(w), _ptr(f), _ptr(keep), _ptr(info)) != 0:
        raise RuntimeError("ref_leave_out_set failed")
    return keep[:n].astype(bool), info


def leave_out_set(worker, firm, fweights=None):
    """build_leave_out_sample with pruning (src/akm_kss.cpp:494-900): dense
    labels in first-appearance order, the (worker, firm) match table, the
    largest connected set (most firms, then rows, then smallest original
    firm id), rounds that drop movers which are cut vertices of the
    mover-firm graph and re-take the largest set, then workers observed once
    leave.  Returns (keep per row, info dict)."""
    worker = np.asarray(worker)
    firm = np.asarray(firm)
    n = worker.size
    w0 = np.unique(worker, return_index=True)
    _, first_w = np.unique(worker, return_index=True)
    wl = np.empty(n, dtype=np.int64)
    ordw = np.argsort(first_w)
    uw = np.unique(worker)
    rank_w = np.empty(uw.size, dtype=np.int64)
    rank_w[ordw] = np.arange(uw.size)
    wl = rank_w[np.searchsorted(uw, worker)]
    uf_, first_f = np.unique(firm, return_index=True)
    rank_f = np.empty(uf_.size, dtype=np.int64)
    rank_f[np.argsort(first_f)] = np.arange(uf_.size)
    fl = rank_f[np.searchsorted(uf_, firm)]
    W0, F0 = uw.size, uf_.size
    cnt = np.ones(n, dtype=np.int64) if fweights is None else np.asarray(fweights, dtype=np.int64)
    key = wl * F0 + fl
    mkey, row_match = np.unique(key, return_inverse=True)
    mw, mf = mkey // F0, mkey % F0
    mc = np.bincount(row_match, weights=cnt, minlength=mkey.size).astype(np.int64)
    M = mkey.size
    active = np.ones(M, dtype=bool)
    firm_orig = np.empty(F0, dtype=np.int64)
    firm_orig[fl] = firm

    def largest_cc():
        a = np.flatnonzero(active)
        g = sp.coo_matrix((np.ones(a.size), (mw[a], W0 + mf[a])), shape=(W0 + F0, W0 + F0))
        _, lab = sp.csgraph.connected_components(g, directed=False)
        obs = np.bincount(lab[mw[a]], weights=mc[a], minlength=W0 + F0)
        fs = np.unique(mf[a])
        firms = np.bincount(lab[W0 + fs], minlength=W0 + F0)
        minf = np.full(W0 + F0, np.iinfo(np.int64).max)
        np.minimum.at(minf, lab[W0 + fs], firm_orig[fs])
        cand = np.flatnonzero((firms > 0) | (obs > 0))
        best = min(cand, key=lambda r: (-firms[r], -obs[r], minf[r]))
        active[active] = lab[mw[active]] == best

    largest_cc()
    n_conn = int(mc[active].sum())
    rounds = 0
    while True:
        deg = np.bincount(mw[active], minlength=W0)
        sel = active & (deg[mw] >= 2)
        if not sel.any():
            break
        movers = np.unique(mw[sel])
        firms_ = np.unique(mf[sel])
        mi = np.full(W0, -1)
        mi[movers] = np.arange(movers.size)
        fi = np.full(F0, -1)
        fi[firms_] = np.arange(firms_.size)
        ptr, adj = graph_csr(movers.size + firms_.size,
                             np.column_stack([mi[mw[sel]], movers.size + fi[mf[sel]]]))
        art = articulation_points(ptr, adj)
        bad = np.zeros(W0, dtype=bool)
        bad[movers[art[:movers.size] == 1]] = True
        if not bad.any():
            break
        active &= ~bad[mw]
        largest_cc()
        rounds += 1
    wobs = np.bincount(mw[active], weights=mc[active], minlength=W0)
    active &= ~(wobs[mw] == 1)
    keep = active[row_match]
    wdeg = np.bincount(mw[active], minlength=W0)
    info = dict(n_obs_input=n, n_obs_connected=n_conn, n_obs=int(cnt[keep].sum()), n_workers=int((wdeg > 0).sum()),
                n_firms=int(np.unique(mf[active]).size), n_matches=int(active.sum()), n_movers=int((wdeg >= 2).sum()),
                n_stayers=int((wdeg == 1).sum()), prune_iterations=rounds)
    return keep, info


def graph_csr(n, edges):
    """CSR (adj_ptr, adj) of an undirected edge list, neighbours sorted."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    u = np.concatenate([e[:, 0], e[:, 1]])
    v = np.concatenate([e[:, 1], e[:, 0]])
    o = np.lexsort((v, u))
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(ptr, u[o] + 1, 1)
    return np.cumsum(ptr).astype(np.int32), v[o].astype(np.int32)


_U = 2.0 ** -53
_ETA = 2.0 ** -1021


def _gamma(k):
    ku = k * _U
    return ku / (1.0 - ku) * (1.0 + 4.0 * _U)


def _up(x):
    return np.nextafter(x, np.inf)


def _down(x):
    return max(0.0, float(np.nextafter(x, -np.inf)))


def _bracket_sum(lo_sum, hi_sum, k):
    g = _gamma(k + 2.0)
    return (_down(_down(lo_sum * _down(1.0 - g)) - (k + 2.0) * _ETA),
            float(_up(_up(hi_sum * _up(1.0 + g)) + (k + 2.0) * _ETA)))


def cert_shadow(V0, Vr, gids, G, limit, weights=None):
    """The shadow certificate's brackets and decision per column (the
    decision contract of src/fe_absorption_cuda_certificate.cu:15-25):
    returns dict(verdict per column, R/A/B brackets, rel to nearest).  Level
    sums S_g are bracketed by |S_g| -+ gamma_{n_g+1} T_g (T_g = sum |w r|), the
    squares and totals by gamma_{K+2}; every bound is pushed one ulp outward;
    2^-1021 per operation covers underflow."""
    V0 = np.asarray(V0, dtype=np.float64).reshape(len(gids[0]), -1)
    Vr = np.asarray(Vr, dtype=np.float64).reshape(V0.shape)
    n, C = V0.shape
    w = np.ones(n) if weights is None else np.asarray(weights, dtype=np.float64)
    Q = float((w * w).sum())
    D = len(gids)
    if weights is None:
        A_lo = A_hi = D * Q
    else:
        qlo, qhi = _bracket_sum(Q, Q, n)
        A_lo, A_hi = _down(D * qlo), float(_up(D * qhi))
    out = dict(verdict=[], R=[], A=(A_lo, A_hi), B=[], rel=[])
    for c in range(C):
        b = float((V0[:, c] ** 2).sum())
        B = _bracket_sum(b, b, n)
        lo = hi = mid = 0.0
        K = 0.0
        for d in range(D):
            cnt = np.bincount(gids[d], minlength=int(G[d]))
            S = np.abs(np.bincount(gids[d], weights=w * Vr[:, c], minlength=int(G[d])))
            T = np.bincount(gids[d], weights=np.abs(w * Vr[:, c]), minlength=int(G[d]))
            m = cnt > 0
            e = np.array([_gamma(float(k) + 1.0) for k in cnt[m]]) * T[m] * (1.0 + 2.0 * _U) + (cnt[m] + 2.0) * _ETA
            l = np.maximum(0.0, (S[m] - e) * (1.0 - 2.0 * _U))
            h = (S[m] + e) * (1.0 + 2.0 * _U)
            lo += float((l * l).sum())
            hi += float((h * h).sum())
            mid += float((S[m] ** 2).sum())
            K += float(G[d])
        R = _bracket_sum(lo, hi, K)
        llo, lhi = _down(limit * limit), float(_up(limit * limit))
        t_lo = _down(_down(llo * A_lo) * B[0])
        t_hi = float(_up(_up(lhi * A_hi) * B[1]))
        v = "PASS" if R[1] <= t_lo else ("FAIL" if R[0] > t_hi else "INDETERMINATE")
        out["verdict"].append(v)
        out["R"].append(R)
        out["B"].append(B)
        out["rel"].append(np.sqrt(mid) / (np.sqrt(D * Q) * np.sqrt(b)))
    return out


def cert_shadow_exact(V0, Vr, gids, G, weights=None):
    """Exact Res_sq(c), Op_sq and Orig_sq(c) of the FP64 inputs in rational
    arithmetic (small cases): the values the shadow brackets must contain."""
    from fractions import Fraction
    V0 = np.asarray(V0, dtype=np.float64).reshape(len(gids[0]), -1)
    Vr = np.asarray(Vr, dtype=np.float64).reshape(V0.shape)
    n, C = V0.shape
    w = [Fraction(1)] * n if weights is None else [Fraction(float(x)) for x in weights]
    A = len(gids) * sum(x * x for x in w)
    R, B = [], []
    for c in range(C):
        B.append(sum(Fraction(float(x)) ** 2 for x in V0[:, c]))
        tot = Fraction(0)
        for d in range(len(gids)):
            sums = [Fraction(0)] * int(G[d])
            for i, g in enumerate(gids[d]):
                sums[g] += w[i] * Fraction(float(Vr[i, c]))
            tot += sum(s * s for s in sums)
        R.append(tot)
    return R, A, B


def ref_absorb_gi(y, X, fes, gi, weights=None, tol=1e-8, method=1):
    """The reference's CPU group()/individual() absorber through oracle/_ref
    (checker): (y_tilde, X_tilde, iterations, converged)."""
    r = ref()
    if r is None:
        return None
    n = len(y)
    y = np.ascontiguousarray(y, dtype=np.float64)
    X = np.asfortranarray(np.asarray(X, dtype=np.float64).reshape(n, -1))
    k = X.shape[1]
    F = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.int32) for f in fes])) if fes else np.zeros(1, np.int32)
    a = {key: np.ascontiguousarray(gi[key], dtype=np.int32)
         for key in ("group_ptr", "group_individual", "individual_ptr", "individual_group")}
    sc = np.ascontiguousarray(gi["group_scale"], dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    yo, Xo = np.empty(n), np.empty((n, k), order="F")
    it, conv = ctypes.c_int(), ctypes.c_int()
    r.ref_absorb_gi.argtypes = [_i, _i, _i, _p, _p, _p, _i, _p, _p, _p, _p, _p, _p, ctypes.c_double, _i, _p, _p,
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    rc = r.ref_absorb_gi(n, k, len(fes), _ptr(y), _ptr(X), _ptr(F), int(gi["num_individuals"]), _ptr(a["group_ptr"]),
                         _ptr(a["group_individual"]), _ptr(a["individual_ptr"]), _ptr(a["individual_group"]),
                         _ptr(sc), _ptr(w), float(tol), int(method), _ptr(yo), _ptr(Xo), ctypes.byref(it),
                         ctypes.byref(conv))
    if rc != 0:
        raise RuntimeError("ref_absorb_gi failed")
    return yo, Xo, it.value, conv.value


def ref_recover_fe(partial, fes, tol=1e-8):
    """The reference's recover_fixed_effects (src/fe_absorption.cpp:15316)
    through oracle/_ref (checker): (contributions d x n, iterations,
    max_delta, converged)."""
    r = ref()
    if r is None:
        return None
    n, d = len(partial), len(fes)
    pv = np.ascontiguousarray(partial, dtype=np.float64)
    F = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.int32) for f in fes]))
    out = np.empty((d, n))
    it, conv = ctypes.c_int(), ctypes.c_int()
    md = ctypes.c_double()
    r.ref_recover_fe.argtypes = [_i, _i, _p, _p, ctypes.c_double, _i, _p,
                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_int)]
    rc = r.ref_recover_fe(n, d, _ptr(pv), _ptr(F), float(tol), 1, _ptr(out), ctypes.byref(it),
                          ctypes.byref(md), ctypes.byref(conv))
    if rc != 0:
        raise RuntimeError("ref_recover_fe failed")
    return out, it.value, md.value, conv.value


def ref_akm_jla_pii(y, worker, firm, draws, seed):
    """The reference's akm_kss_decompose with JLA leverages at match level (no
    pruning) through oracle/_ref (checker): (row_worker, row_firm,
    row_weight, pii) of its working rows, in its order."""
    r = ref()
    if r is None:
        return None
    n = len(y)
    y = np.ascontiguousarray(y, dtype=np.float64)
    w = np.ascontiguousarray(worker, dtype=np.int32)
    f = np.ascontiguousarray(firm, dtype=np.int32)
    rw = np.empty(n, dtype=np.int32)
    rf = np.empty(n, dtype=np.int32)
    rc = np.empty(n)
    pii = np.empty(n)
    r.ref_akm_jla_pii.argtypes = [_i, _p, _p, _p, _i, ctypes.c_ulonglong, _p, _p, _p, _p]
    R = r.ref_akm_jla_pii(n, _ptr(y), _ptr(w), _ptr(f), int(draws), int(seed), _ptr(rw), _ptr(rf), _ptr(rc),
                          _ptr(pii))
    if R < 0:
        raise RuntimeError("ref_akm_jla_pii failed")
    return rw[:R], rf[:R], rc[:R], pii[:R]


def schwarz_rule(V, Vt, gids, G):
    """The stopping rule of schwarz_demean_raw evaluated on a result
    (src/schwarz_demean.cpp:283-305, :358-378): per column c,
    rel_c = ||D'vt_c|| / ||D'v_c|| over the normal equations of all
    dimensions, leaving out the pinned levels -- level 0 of every dimension
    except the one with the most levels (the first such, :65-69; pins at
    :224-225, :283-284).  Columns with ||D'v_c|| <= 1e-300 are inactive (rel 0).
    Returns the per-column rel array."""
    V = np.asarray(V, dtype=np.float64)
    Vt = np.asarray(Vt, dtype=np.float64)
    if V.ndim == 1:
        V, Vt = V[:, None], Vt[:, None]
    C = V.shape[1]
    pa = int(np.argmax(np.asarray(G)))
    bn = np.zeros(C)
    rn = np.zeros(C)
    for d, (g, Gd) in enumerate(zip(gids, G)):
        g = np.asarray(g)
        for c in range(C):
            b = np.bincount(g, weights=V[:, c], minlength=Gd)
            r = np.bincount(g, weights=Vt[:, c], minlength=Gd)
            if d != pa:
                b[0] = 0.0
                r[0] = 0.0
            bn[c] += b @ b
            rn[c] += r @ r
    bn = np.sqrt(bn)
    return np.where(bn > 1e-300, np.sqrt(rn) / np.where(bn > 1e-300, bn, 1.0), 0.0)


def slope_projection(V, gids, G, slope_dim, z, intercept=True, weights=None):
    """The within transform with a heterogeneous slope term on dimension
    slope_dim (slope_project_inplace, src/fe_absorption.cpp:4096-4300, as the
    fixed point of the alternating projections): v - A (A'WA)^+ A'W v, where A
    holds one indicator column per level of every other dimension and, for
    slope_dim, z times the level indicator (and the indicator itself with the
    intercept).  Dense least squares: small panels only."""
    V = np.asarray(V, dtype=np.float64)
    if V.ndim == 1:
        V = V[:, None]
    n = V.shape[0]
    cols = []
    for d, (g, Gd) in enumerate(zip(gids, G)):
        g = np.asarray(g)
        ind = np.zeros((n, Gd))
        ind[np.arange(n), g] = 1.0
        if d == slope_dim:
            if intercept:
                cols.append(ind)
            cols.append(ind * np.asarray(z, dtype=np.float64)[:, None])
        else:
            cols.append(ind)
    A = np.hstack(cols)
    sw = np.ones(n) if weights is None else np.sqrt(np.asarray(weights, dtype=np.float64))
    coef, *_ = np.linalg.lstsq(A * sw[:, None], V * sw[:, None], rcond=1e-12)
    return V - A @ coef


def ref_schwarz(y, X, gids, G, tol=1e-10, max_iter=10000):
    """The reference's schwarz_demean_raw through oracle/_ref (checker):
    (ok, y_t, X_t, iterations)."""
    r = ref()
    if r is None:
        return None
    y = np.ascontiguousarray(y, dtype=np.float64)
    X = np.asfortranarray(np.asarray(X, dtype=np.float64).reshape(y.size, -1))
    n, k = X.shape
    ids = np.ascontiguousarray(np.stack([np.asarray(g, dtype=np.int32) for g in gids]))
    lev = np.ascontiguousarray(G, dtype=np.int32)
    yo = np.empty(n)
    Xo = np.empty((n, k), order="F")
    it = _i()
    rc = r.ref_schwarz(_i(n), _i(k), _i(len(gids)), _ptr(y), _ptr(X), _ptr(ids), _ptr(lev), _d(tol),
                       _i(max_iter), _ptr(yo), _ptr(Xo), ctypes.byref(it))
    if rc < 0:
        raise RuntimeError("ref_schwarz failed")
    return bool(rc), yo, Xo, it.value


def ref_absorb_slope(y, X, fes, slope_dim, z, intercept=True, tol=1e-8, tolerance_mode=1, threads=1):
    """The reference's absorb_fixed_effects_v6 with one heterogeneous slope
    term through oracle/_ref (checker): (y_t, X_t, iterations, converged)."""
    r = ref()
    if r is None:
        return None
    y = np.ascontiguousarray(y, dtype=np.float64)
    X = np.asfortranarray(np.asarray(X, dtype=np.float64).reshape(y.size, -1))
    n, k = X.shape
    ids = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.int32) for f in fes]))
    zz = np.ascontiguousarray(z, dtype=np.float64)
    yo = np.empty(n)
    Xo = np.empty((n, k), order="F")
    it, cv = _i(), _i()
    rc = r.ref_absorb_slope(_i(n), _i(k), _i(len(fes)), _ptr(y), _ptr(X), _ptr(ids), _i(slope_dim), _ptr(zz),
                            _i(int(bool(intercept))), _d(tol), _i(tolerance_mode), _i(threads), _ptr(yo),
                            _ptr(Xo), ctypes.byref(it), ctypes.byref(cv))
    if rc != 0:
        raise RuntimeError("ref_absorb_slope failed")
    return yo, Xo, it.value, bool(cv.value)


def ref_indexer(fes, weights=None):
    """The reference's build_indexer and compute_weight_sums through oracle/_ref
    (checker): a list of (gids, num_groups, levels_present, weight_sums) per
    dimension."""
    r = ref()
    if r is None:
        return None
    ids = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.int32) for f in fes]))
    D, n = ids.shape
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    gids = np.empty((D, n), dtype=np.int32)
    G = np.zeros(D, dtype=np.int32)
    lp = np.zeros(D, dtype=np.int32)
    ws = np.zeros((D, max(n, 1)))
    rc = r.ref_indexer(_i(n), _i(D), _ptr(ids), _ptr(w), _ptr(gids), _ptr(G), _ptr(lp), _ptr(ws))
    if rc != 0:
        raise RuntimeError("ref_indexer failed")
    return [(gids[j].copy(), int(G[j]), int(lp[j]), ws[j, :G[j]].copy()) for j in range(D)]


def ref_absorb(y, X, fes, tol=1e-8, tolerance_mode=1, threads=1, symmetric=False, weights=None):
    r = ref()
    if r is None:
        return None
    y = np.ascontiguousarray(y, dtype=np.float64)
    X = np.asfortranarray(X, dtype=np.float64)
    n, k = X.shape
    ids = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.int32) for f in fes]))
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    yo = np.empty(n)
    Xo = np.empty((n, k), order="F")
    it, conv = _i(), _i()
    rc = r.ref_absorb(_i(n), _i(k), _i(len(fes)), _ptr(y), _ptr(X), _ptr(ids), _ptr(w), _d(tol),
                      _i(tolerance_mode), _i(threads), _i(int(symmetric)), _ptr(yo), _ptr(Xo),
                      ctypes.byref(it), ctypes.byref(conv))
    if rc != 0:
        raise RuntimeError("ref_absorb failed")
    return yo, Xo, it.value, bool(conv.value)


def ref_partial_out(y, X, fes, tol=1e-8, tolerance_mode=1, threads=1):
    r = ref()
    if r is None:
        return None
    y = np.ascontiguousarray(y, dtype=np.float64)
    X = np.asfortranarray(X, dtype=np.float64)
    n, k = X.shape
    ids = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.int32) for f in fes]))
    idx = np.empty(n, dtype=np.int32)
    yo = np.empty(n)
    Xo = np.empty((n, k), order="F")
    it, conv = _i(), _i()
    kept = r.ref_partial_out(_i(n), _i(k), _i(len(fes)), _ptr(y), _ptr(X), _ptr(ids), _d(tol),
                             _i(tolerance_mode), _i(threads), _ptr(idx), _ptr(yo), _ptr(Xo),
                             ctypes.byref(it), ctypes.byref(conv))
    if kept < 0:
        raise RuntimeError("ref_partial_out failed")
    return idx[:kept].copy(), yo[:kept].copy(), np.asfortranarray(Xo.reshape(-1, order="F")[:kept * k].reshape(kept, k, order="F")), it.value, bool(conv.value)


def ref_fit(y, X, fes, cluster=None, tol=1e-8, tolerance_mode=1, threads=1, groupvar=True):
    r = ref()
    if r is None:
        return None
    y = np.ascontiguousarray(y, dtype=np.float64)
    X = np.asfortranarray(X, dtype=np.float64)
    n, k = X.shape
    D = len(fes)
    ids = np.ascontiguousarray(np.stack([np.asarray(f, dtype=np.int32) for f in fes]))
    cl = None if cluster is None else np.ascontiguousarray(cluster, dtype=np.int32)
    coef = np.zeros(k)
    se = np.zeros(k)
    idx = np.empty(n, dtype=np.int32)
    gv = np.zeros(n, dtype=np.int32)
    df_a, df_r, rss = _d(), _d(), _d()
    red = np.zeros(D, dtype=np.int32)
    lev = np.zeros(D, dtype=np.int32)
    ns, it = _i(), _i()
    kept = r.ref_fit(_i(n), _i(k), _i(D), _ptr(y), _ptr(X), _ptr(ids), _ptr(cl), _d(tol),
                     _i(tolerance_mode), _i(threads), _i(int(groupvar)), _ptr(coef), _ptr(se),
                     _ptr(idx), _ptr(gv), ctypes.byref(df_a), ctypes.byref(df_r), _ptr(red),
                     _ptr(lev), ctypes.byref(ns), ctypes.byref(it), ctypes.byref(rss))
    if kept < 0:
        raise RuntimeError("ref_fit failed")
    return dict(coef=coef, se=se, sample_index=idx[:kept].copy(), groupvar=gv[:kept].copy(),
                df_a=df_a.value, df_resid=df_r.value, fe_redundant=red, fe_levels=lev,
                num_singletons=ns.value, iterations=it.value, rss=rss.value, nobs=kept)


def boot_gram_exact(worker, firm, V, mult):
    """Gram R' W R of the within transform R of V (n x C) on worker and firm
    dummies under the row weights W = mult[worker] (rows of workers with
    multiplicity 0 drop out), by a dense weighted least-squares projection."""
    worker = np.asarray(worker)
    firm = np.asarray(firm)
    V = np.asarray(V, dtype=np.float64)
    w = np.asarray(mult, dtype=np.float64)[worker]
    keep = w > 0
    wk, fk, Vk, sw = worker[keep], firm[keep], V[keep], np.sqrt(w[keep])
    _, wi = np.unique(wk, return_inverse=True)
    _, fi = np.unique(fk, return_inverse=True)
    n = wk.size
    D = np.zeros((n, wi.max() + 1 + fi.max() + 1))
    D[np.arange(n), wi] = 1.0
    D[np.arange(n), wi.max() + 1 + fi] = 1.0
    A = D * sw[:, None]
    B = Vk * sw[:, None]
    coef, *_ = np.linalg.lstsq(A, B, rcond=None)
    R = B - A @ coef
    return R.T @ R
