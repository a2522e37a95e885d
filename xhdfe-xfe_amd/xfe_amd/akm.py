"""Host side of the AKM/KSS leverage solver: the match structure that the
reference's TwoWaySolver::build hands to akm_cuda_create
(src/akm_kss.cpp:1042-1091, the call at :1154), built from a worker-firm
panel.  numpy marshalling around the device solver (AkmSolver)."""
from __future__ import annotations

import numpy as np


def build_structure(mw, mf, mc, N, J):
    """TwoWaySolver::build (src/akm_kss.cpp:1042-1091): matches (worker, firm,
    count) grouped by worker -> CSR worker offsets w_ptr, firm offsets f_ptr,
    per-firm match lists f_matches (ascending match index), degrees Dw and Df
    and the Jacobi inverse diagonal of the grounded S (diag S = Df -
    sum c^2 / Dw; 0 for the ground firm J-1 and for non-positive entries)."""
    mw = np.ascontiguousarray(mw, dtype=np.int32)
    mf = np.ascontiguousarray(mf, dtype=np.int32)
    c = np.ascontiguousarray(mc, dtype=np.float64)
    Dw = np.bincount(mw, weights=c, minlength=N)
    Df = np.bincount(mf, weights=c, minlength=J)
    w_ptr = np.zeros(N + 1, dtype=np.int64)
    np.cumsum(np.bincount(mw, minlength=N), out=w_ptr[1:])
    f_ptr = np.zeros(J + 1, dtype=np.int64)
    np.cumsum(np.bincount(mf, minlength=J), out=f_ptr[1:])
    f_matches = np.argsort(mf, kind="stable").astype(np.int64)
    diag = Df - np.bincount(mf, weights=c * c / Dw[mw], minlength=J)
    ok = (np.arange(J) < J - 1) & (diag > 0.0)
    inv_diag = np.where(ok, 1.0 / np.where(ok, diag, 1.0), 0.0)
    return dict(N=int(N), J=int(J), m_w=mw, m_f=mf, m_c=c, w_ptr=w_ptr, f_ptr=f_ptr, f_matches=f_matches,
                Dw=Dw, Df=Df, inv_diag=inv_diag)


def panel_matches(worker, firm, connected=True):
    """Collapse panel rows into matches (worker, firm, person-years), grouped
    by worker with firms ascending, on compact 0-based ids.  With connected,
    only the largest mobility group (connected component of the worker-firm
    graph, from the device's xfe_bipartite_components) is kept: the
    Laplacian is grounded by one firm, so it must be connected.  Returns
    (mw, mf, mc, N, J)."""
    from . import components  # device kernel; imported here to keep the module light
    worker = np.asarray(worker, dtype=np.int64)
    firm = np.asarray(firm, dtype=np.int64)
    if connected:
        _, labels = components(worker.astype(np.int32), firm.astype(np.int32))
        labels = np.asarray(labels)
        big = np.bincount(labels).argmax()
        keep = labels == big
        worker, firm = worker[keep], firm[keep]
    wu, wi = np.unique(worker, return_inverse=True)
    fu, fi = np.unique(firm, return_inverse=True)
    N, J = wu.size, fu.size
    key = wi.astype(np.int64) * J + fi
    uk, counts = np.unique(key, return_counts=True)
    mw = (uk // J).astype(np.int32)
    mf = (uk % J).astype(np.int32)
    return mw, mf, counts.astype(np.float64), int(N), int(J)



def _first_appearance(ids):
    """Dense labels in first-appearance order and the original id of each."""
    u, first, inv = np.unique(np.asarray(ids), return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty(u.size, dtype=np.int64)
    rank[order] = np.arange(u.size)
    return u[order], rank[inv.reshape(-1)]


def match_panel(worker, firm):
    """Panel rows -> matches without pruning, as the reference's working rows
    (src/akm_kss.cpp:873-960): workers and firms take dense labels in
    first-appearance order over the rows, matches are sorted by (worker
    label, firm label).  Returns (wu, fu, wi, fi, ridx, mw, mf, c) with wu /
    fu the original ids of the labels, wi / fi the rows' labels, ridx each
    row's match and mw / mf / c the matches and their person-year counts."""
    wu, wi = _first_appearance(worker)
    fu, fi = _first_appearance(firm)
    J = fu.size
    key = wi * J + fi
    uk, ridx, counts = np.unique(key, return_inverse=True, return_counts=True)
    return (wu, fu, wi, fi, ridx.reshape(-1), (uk // J).astype(np.int32), (uk % J).astype(np.int32),
            counts.astype(np.float64))


def two_way_rhs(y, ridx, mw, mf, c, N, J, Dw):
    """Right-hand side of the firm Schur system of the two-way fit
    K [alpha; psi] = [W'y; F'y]: F'y - B' Dw^-1 W'y, with the per-match sums
    of y and W'y (rhs, m_ysum, tw)."""
    m_ysum = np.bincount(ridx, weights=y, minlength=mw.size)
    tw = np.bincount(mw, weights=m_ysum, minlength=N)
    tf = np.bincount(mf, weights=m_ysum, minlength=J)
    return tf - np.bincount(mf, weights=c * (tw / Dw)[mw], minlength=J), m_ysum, tw


def kss_sums(y, wi, fi, ridx, mw, mf, c, alpha, psi, P, Bfe, Bpe, Bcov, corr=None):
    """The O(rows) part of the reference's variance decomposition at match
    level with unweighted rows (src/akm_kss.cpp:2632-2662 for yt / eta / the
    stayer flags, :3299-3349 for sigma_i, :3353-3410 for the components), from
    the two-way fit (alpha, psi) and the per-match leverages P and weights
    (Bfe, Bpe, Bcov):

      yt = sqrt(c) ybar_m, eta = yt - sqrt(c) (alpha_w + psi_f);
      sigma_i = (yt - mean yt) corr (eta / M_ii), M_ii = 1 - P_ii, on a mover's match
      (corr: the JLA estimator's non-linearity correction, 1 for exact leverages)
      (0 where M_ii <= 1e-12: such a match is counted in unit_leverage); on a
      stayer's match the mean over its person-years of
      (y - mean y) e / (1 - 1 / T_w), e = y - alpha - psi, T_w = Dw;
      plug-in: D-weighted variances and the covariance of alpha and psi over
      n_py - 1; KSS = plug-in - sum(B sigma) / (n_py - 1);
      AGSU = plug-in - sigma2_ho sum(B) / (n_py - 1), sigma2_ho =
      RSS / (n_py - (N + J - 1)) (0 when that is not positive)."""
    N, J = alpha.size, psi.size
    M = mw.size
    Dw = np.bincount(mw, weights=c, minlength=N)
    Df = np.bincount(mf, weights=c, minlength=J)
    sc = np.sqrt(c)
    m_ysum = np.bincount(ridx, weights=y, minlength=M)
    yt = sc * (m_ysum / c)
    eta = yt - sc * (alpha[mw] + psi[mf])
    stayer = np.bincount(mw, minlength=N)[mw] < 2
    Mii = 1.0 - P
    sigma = np.zeros(M)
    mov = ~stayer & (Mii > 1e-12)
    unit = int(np.count_nonzero(~stayer & ~(Mii > 1e-12)))
    cf = np.ones(M) if corr is None else np.asarray(corr, dtype=np.float64)
    sigma[mov] = (yt[mov] - yt.mean()) * cf[mov] * (eta[mov] / Mii[mov])
    T = Dw[wi]
    e = y - alpha[wi] - psi[fi]
    rows = stayer[ridx]
    acc = np.bincount(ridx[rows], weights=((y[rows] - y.mean()) * (e[rows] / (1.0 - 1.0 / T[rows]))),
                      minlength=M)
    sigma[stayer] = acc[stayer] / c[stayer]
    n_py = float(c.sum())
    dof = n_py - 1.0
    pb = float(Df @ psi) / n_py
    ab = float(Dw @ alpha) / n_py
    plugin = (float(Dw @ ((alpha - ab) ** 2)) / dof, float(Df @ ((psi - pb) ** 2)) / dof,
              float(np.sum(c * (alpha[mw] - ab) * (psi[mf] - pb))) / dof)
    B = (Bpe, Bfe, Bcov)  # the weights of var_alpha, var_psi, cov_alpha_psi
    kss = tuple(plugin[i] - float(B[i] @ sigma) / dof for i in range(3))
    dof_ho = n_py - (N + J - 1)
    s2 = float(e @ e) / dof_ho if dof_ho > 0 else 0.0
    agsu = tuple(plugin[i] - s2 * float(B[i].sum()) / dof for i in range(3))
    return dict(plugin=plugin, agsu=agsu, kss=kss, sigma2_ho=s2, n_py=n_py, pii=P, sigma_i=sigma,
                unit_leverage=unit)


def kss_components(y, worker, firm, method="exact", draws=200, seed=20260705, tol=1e-11, max_iter=10000,
                   accept=0.0):
    """The variance decomposition of the reference's akm_kss_decompose at
    match level with unweighted rows, no controls and no pruning (the panel
    must be leave-out connected; src/akm_kss.cpp:2551-2741, :2753-3288,
    :3299-3410), for method "exact" (LeverageMethod::Exact) or "jla"
    (LeverageMethod::Jla: `draws` Rademacher draws of the reference's streams
    from `seed`).  The two-way fit (the grounded firm Schur solve), the
    leverages, the JLA correction and the quadratic-form weights run on the
    device (xfe_akm_exact_leverages, or xfe_akm_jla_leverages and
    xfe_akm_jla_bii); kss_sums forms the O(rows) sums.  The rows are
    match_panel's, the reference's working-row order, so the JLA draws equal
    the reference's.  accept: as in AkmSolver.exact_leverages.  Returns
    kss_sums' dict (plugin, agsu, kss as (var_alpha, var_psi,
    cov_alpha_psi); sigma2_ho; n_py; per match P_ii and sigma_i) with the
    matches' original ids (match_worker, match_firm), counts, the weights
    bfe / bpe / bcov, the correction corr and the seconds of the fit, the
    leverages and the sums."""
    import time
    from . import AkmSolver, XfeError
    if method not in ("exact", "jla"):
        raise XfeError(f"kss_components: unknown method {method!r}")
    y = np.asarray(y, dtype=np.float64)
    wu, fu, wi, fi, ridx, mw, mf, c = match_panel(worker, firm)
    N, J = wu.size, fu.size
    st = build_structure(mw, mf, c, N, J)
    t0 = time.perf_counter()
    rhs, _, tw = two_way_rhs(y, ridx, mw, mf, c, N, J, st["Dw"])
    solver = AkmSolver(st)
    corr = None
    try:
        z, its = solver.solve_multi(rhs[None, :], tol=tol, max_iter=max_iter)
        if its[0] < 0:
            raise XfeError("kss_components: the two-way fit did not converge")
        t1 = time.perf_counter()
        if method == "exact":
            P, Bfe, Bpe, Bcov, failed = solver.exact_leverages(tol=tol, max_iter=max_iter, accept=accept)
            if failed:
                raise XfeError(f"kss_components: {failed} leverage solves did not converge")
        else:
            P, corr, ok = solver.jla_leverages(draws=draws, seed=seed, tol=tol, max_iter=max_iter)
            Bfe, Bpe, Bcov, ok2 = solver.jla_bii(None, draws=draws, seed=seed, tol=tol, max_iter=max_iter)
            if not (ok and ok2):
                raise XfeError("kss_components: some JLA draw solves did not converge")
        t2 = time.perf_counter()
    finally:
        solver.close()
    psi = np.asarray(z[0], dtype=np.float64)
    alpha = (tw - np.bincount(mw, weights=c * psi[mf], minlength=N)) / st["Dw"]
    r = kss_sums(y, wi, fi, ridx, mw, mf, c, alpha, psi, P, Bfe, Bpe, Bcov, corr=corr)
    t3 = time.perf_counter()
    r.update(match_worker=wu[mw], match_firm=fu[mf], match_count=c, bfe=Bfe, bpe=Bpe, bcov=Bcov, corr=corr,
             seconds=dict(fit=t1 - t0, leverages=t2 - t1, sums=t3 - t2))
    return r


def kss_components_exact(y, worker, firm, tol=1e-11, max_iter=10000, accept=0.0):
    """kss_components with exact leverages."""
    return kss_components(y, worker, firm, method="exact", tol=tol, max_iter=max_iter, accept=accept)
