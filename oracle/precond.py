"""Test infrastructure only: a numpy restatement of the aggregate-block
preconditioner of the mixed-precision absorber (xhdfe-xfe_amd/csrc/
lp_precond.hip) and of the learned preconditioner's input features
(lp_learned.hip), the checker of those kernels.  Nothing in the product
imports it.

The reference has no such preconditioner (its FE normal equations are
preconditioned by the randomized approximate Cholesky of
src/schwarz_demean.cpp:100-133); the object it acts on is the reference's own:
the Schur complement of the worker dimension, S = D_R' M_p D_R, whose firm
block is

    S_ff = sum_w (L_wf - L_wf^2 / n_w),   S_fg = -W_fg,   W_fg = sum_w L_wf L_wg / n_w

(L_wf: rows of worker w at firm f, n_w: rows of w).  pairs_dense() computes it
from the rows by sparse products; tests pin it against the dense
D_R' (I - D_p (D_p' D_p)^-1 D_p') D_R of small panels.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def pairs_dense(worker, firm, Nw=None, F=None):
    """(pf, pg, W, dex): the firm pairs f < g with W_fg > 0 in (f, g) order,
    their coupling weights, and the exact firm diagonal S_ff."""
    worker = np.asarray(worker, np.int64)
    firm = np.asarray(firm, np.int64)
    Nw = int(worker.max()) + 1 if Nw is None else Nw
    F = int(firm.max()) + 1 if F is None else F
    B = sp.csr_matrix((np.ones(worker.size), (worker, firm)), shape=(Nw, F))  # L_wf
    nw = np.asarray(B.sum(axis=1)).ravel()
    Bs = sp.diags(1.0 / np.maximum(nw, 1)) @ B
    G = (B.T @ Bs).tocoo()  # sum_w L_wf L_wg / n_w
    rows = np.asarray(B.sum(axis=0)).ravel()
    dex = rows - np.asarray(sp.csr_matrix(G).diagonal())
    m = G.row < G.col
    pf, pg, W = G.row[m], G.col[m], G.data[m]
    o = np.lexsort((pg, pf))
    return pf[o].astype(np.int32), pg[o].astype(np.int32), W[o], dex


def schur_dense(worker, firm, year):
    """Dense D_R' M_p D_R over (firms, years) of a small panel."""
    worker, firm, year = (np.asarray(a, np.int64) for a in (worker, firm, year))
    n = worker.size
    Nw, F, Y = worker.max() + 1, firm.max() + 1, year.max() + 1
    Dp = np.zeros((n, Nw))
    Dp[np.arange(n), worker] = 1.0
    DR = np.zeros((n, F + Y))
    DR[np.arange(n), firm] = 1.0
    DR[np.arange(n), F + year] = 1.0
    M = np.eye(n) - Dp @ np.linalg.solve(Dp.T @ Dp, Dp.T)
    return DR.T @ M @ DR


def strength(pf, pg, W, dex):
    d = np.sqrt(np.maximum(dex[pf], 0) * np.maximum(dex[pg], 0))
    return np.where(d > 0, W / np.where(d > 0, d, 1.0), 0.0).astype(np.float32)


def _ord(score):
    u = np.asarray(score, np.float32).view(np.uint32).astype(np.uint64)
    return np.where(u & 0x80000000, (~u) & 0xFFFFFFFF, u | 0x80000000)


def handshake(pf, pg, score, F, cap=16, rounds=6):
    """The parallel handshake aggregation: per round every aggregate (root)
    takes the largest key (score, then lower pair index) over its pairs to
    other aggregates that together hold at most cap firms; a root whose best
    pair is also the best pair of the other end merges the larger root into the
    smaller.  Returns the root of every firm."""
    pf = np.asarray(pf, np.int64)
    pg = np.asarray(pg, np.int64)
    U = pf.size
    key = (_ord(score) << np.uint64(32)) | (np.uint64(0xFFFFFFFF) - np.arange(U, dtype=np.uint64))
    agg = np.arange(F)
    size = np.ones(F, np.int64)
    for _ in range(rounds):
        a, b = agg[pf], agg[pg]
        ok = (a != b) & (size[a] + size[b] <= cap) & np.isfinite(np.asarray(score, np.float32))
        best = np.zeros(F, np.uint64)
        np.maximum.at(best, a[ok], key[ok])
        np.maximum.at(best, b[ok], key[ok])
        newroot = np.full(F, -1)
        roots = np.flatnonzero((best != 0) & (agg == np.arange(F)))
        for r in roots:
            p = int(np.uint64(0xFFFFFFFF) - (best[r] & np.uint64(0xFFFFFFFF)))
            x, y = agg[pf[p]], agg[pg[p]]
            other = y if x == r else x
            if other > r and best[other] == best[r]:
                newroot[other] = r
                size[r] += size[other]
        m = newroot[agg] >= 0
        agg = np.where(m, newroot[agg], agg)
    return agg


def block_inverse(Bm):
    """Generalized inverse of a symmetric positive semi-definite block: L D L'
    with a pivot dropped when it falls to 1e-10 of its diagonal entry, then
    L^-T D^+ L^-1.  B inv B = B; the plain inverse when B is non-singular."""
    s = Bm.shape[0]
    A = np.array(Bm, dtype=np.float64)
    L = np.eye(s)
    D = np.zeros(s)
    for k in range(s):
        a0 = A[k, k]
        d = a0 - np.sum(L[k, :k] ** 2 * D[:k])
        keep = d > 1e-10 * a0 and a0 > 0
        D[k] = d if keep else 0.0
        for m in range(k + 1, s):
            v = A[m, k] - np.sum(L[m, :k] * L[k, :k] * D[:k])
            L[m, k] = v / d if keep else 0.0
    Y = np.linalg.inv(L)
    Dp = np.where(D > 0, 1.0 / np.where(D > 0, D, 1.0), 0.0)
    return Y.T @ np.diag(Dp) @ Y


def blocks_apply(pf, pg, W, dex, agg, r):
    """z = inverse of each aggregate's block (size >= 2) times r on its firms;
    0 on firms outside blocks.  r: F x C."""
    F = dex.size
    r = np.asarray(r, np.float64).reshape(F, -1)
    z = np.zeros_like(r)
    pos = {}
    for f in range(F):
        pos.setdefault(int(agg[f]), []).append(f)
    Wd = {(int(a), int(b)): w for a, b, w in zip(pf, pg, W)}
    for root, mem in pos.items():
        if len(mem) < 2:
            continue
        mem = sorted(mem)
        s = len(mem)
        Bm = np.zeros((s, s))
        for i, f in enumerate(mem):
            Bm[i, i] = dex[f]
            for j, g in enumerate(mem):
                if f < g and (f, g) in Wd:
                    Bm[i, j] = Bm[j, i] = -Wd[(f, g)]
        z[mem] = block_inverse(Bm) @ r[mem]
    return z
