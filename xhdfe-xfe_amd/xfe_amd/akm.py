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
