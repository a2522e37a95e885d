"""Mixed-precision absorption with heterogeneous slopes (xfe_absorb with a
slope dimension and precision 1, and xfe_absorb_device_slopes): the worker
dimension eliminated, FP32 Schur PCG with FP64 residual replacement, against
the oracle's exact slope projection (dense least squares, pinned against the
reference's slope absorber in the CPU suite), the reference's slope fixtures
and the FP64 slope path."""
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


def _prep(case):
    from test_gpu_parity import _prep as prep
    return prep(case)


@pytest.mark.parametrize("name,sdim,icpt", [("two_way", 1, True), ("two_way", 1, False),
                                             ("akm_three_way", 1, True)])
def test_slope_mixed_matches_oracle_and_reference(golden, cases, name, sdim, icpt):
    from golden_cases import slope_values
    _log(f"mixed slopes {name} dim {sdim} intercept {icpt}")
    _, fes, gids, G, V = _prep(cases[name])
    z = slope_values(V.shape[0])
    yg, Xg, info = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1, slopes={sdim: (z, icpt)})
    assert info["converged"] and info["precision_used"] == 1 and not info["precision_fallback"], info
    assert info["certificate_computed"] and info["precision_certified"], info
    Vg = np.column_stack([yg, Xg])
    Vo = oracle.slope_projection(V, gids, G, sdim, z, intercept=icpt)
    tag = f"slope{sdim}_{'icpt' if icpt else 'noicpt'}"
    Vr = np.column_stack([golden[f"{name}/{tag}/y_tilde"], golden[f"{name}/{tag}/X_tilde"]])
    scale = np.abs(V).max()
    err_o = np.abs(Vg - Vo).max() / scale
    err_r = np.abs(Vg - Vr).max() / scale
    _log(f"  iterations {info['iterations']} replacements {info['refinements']} vs oracle {err_o:.2e} "
         f"vs reference {err_r:.2e} certificate {info['abs_residual_rel']:.2e}")
    assert err_o <= 1e-9 and err_r <= 2e-8


def _c2_panel(scale, regressors):
    p = panels.slope_panel(scale=scale, regressors=regressors)
    worker = np.unique(p["fes"][0], return_inverse=True)[1].astype(np.int32)
    firm = np.unique(p["fes"][1], return_inverse=True)[1].astype(np.int32)
    cell = np.unique(p["fes"][2], return_inverse=True)[1].astype(np.int32)
    gids = [worker, firm, cell]
    G = [int(g.max()) + 1 for g in gids]
    return p, gids, G


def test_slope_mixed_c2_shape_matches_fp64_path():
    """The config-2 shape at small scale (worker, firm with intercept and slope
    on x1, firm x year cells; 32 columns): the firm intercept is spanned by the
    cells and left out of the solve; the result equals the FP64 slope path at
    a tight tolerance, every column."""
    _log("mixed slopes, C2 shape")
    p, gids, G = _c2_panel(0.003, 31)
    V = np.column_stack([p["y"], p["X"]])
    z = p["z"]
    ym, Xm, im = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1, slopes={1: (z, True)})
    assert im["converged"] and im["precision_used"] == 1 and im["precision_certified"], im
    y6, X6, i6 = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-13, precision=0, slopes={1: (z, True)})
    assert i6["converged"]
    Vm, V6 = np.column_stack([ym, Xm]), np.column_stack([y6, X6])
    err = np.abs(Vm - V6).max(axis=0) / np.abs(V).max(axis=0)
    _log(f"  n {V.shape[0]} levels {G} iterations {im['iterations']} (FP64 path {i6['iterations']}); "
         f"max column error {err.max():.2e}; certificate {im['abs_residual_rel']:.2e}")
    assert err.max() <= 1e-8
    # the dense projection on a smaller panel, same structure
    q, qg, qG = _c2_panel(0.0002, 3)
    Vq = np.column_stack([q["y"], q["X"]])
    yq, Xq, iq = xfe_amd.absorb(Vq[:, 0], Vq[:, 1:], qg, qG, tol=1e-8, precision=1, slopes={1: (q["z"], True)})
    Vo = oracle.slope_projection(Vq, qg, qG, 1, q["z"], intercept=True)
    assert iq["converged"] and np.abs(np.column_stack([yq, Xq]) - Vo).max() <= 1e-9 * np.abs(Vq).max()


def test_slope_mixed_device_entry_matches_host_entry():
    import torch
    _log("mixed slopes, device entry")
    p, gids, G = _c2_panel(0.002, 7)
    V = np.column_stack([p["y"], p["X"]])
    n, C = V.shape
    ym, Xm, im = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1, slopes={1: (p["z"], True)})
    dev = torch.device("cuda:0")
    dV = torch.from_numpy(np.asfortranarray(V).T.copy()).to(dev)   # C x n: column-major, ld n
    dI = torch.from_numpy(np.stack(gids)).to(dev)
    dz = torch.from_numpy(p["z"]).to(dev)
    info = xfe_amd.absorb_device_slopes(dV.data_ptr(), n, C, n, dI.data_ptr(), G, [0, dz.data_ptr(), 0], [0, 1, 0])
    Vd = dV.cpu().numpy().T
    assert info["converged"] and info["precision_used"] == 1
    assert np.array_equal(Vd, np.column_stack([ym, Xm]))
    # weights are not on the mixed path: the FP64 slope path runs and says so
    yw, Xw, iw = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, weights=np.ones(n), tol=1e-8, precision=1,
                                slopes={1: (p["z"], True)})
    assert iw["precision_used"] == 0 and iw["precision_fallback"]


def test_slope_mixed_fp64_continuation(monkeypatch):
    """Columns the FP32 runs leave unconverged continue with FP64 block-PCG
    from their iterate: with the Jacobi preconditioner in place of the firm
    blocks (XFE_SLOPE_JACOBI) the FP32 runs need several times the 1,500
    iterations allowed here, the FP64 continuation finishes, and the result is
    certified and equal to the FP64 slope path run tight."""
    _log("mixed slopes, FP64 continuation")
    p, gids, G = _c2_panel(0.003, 7)
    V = np.column_stack([p["y"], p["X"]])
    z = p["z"]
    monkeypatch.setenv("XFE_SLOPE_JACOBI", "1")
    ym, Xm, im = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1, slopes={1: (z, True)},
                                max_iter=1500)
    monkeypatch.delenv("XFE_SLOPE_JACOBI")
    y6, X6, i6 = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-13, precision=0, slopes={1: (z, True)})
    _log(f"  continuation: converged {im['converged']} precision_used {im['precision_used']} "
         f"iterations {im['iterations']} sweeps {im['sweeps']}")
    assert im["converged"] and im["precision_used"] == 1 and im["precision_certified"], im
    assert im["iterations"] > 1500  # the FP32 runs stopped at 1,500; the rest is the FP64 continuation
    Vm, V6 = np.column_stack([ym, Xm]), np.column_stack([y6, X6])
    err = np.abs(Vm - V6).max(axis=0) / np.abs(V).max(axis=0)
    assert err.max() <= 1e-8


def test_slope_mixed_falls_back_to_fp64_path(monkeypatch):
    """A mixed run that stops unconverged leaves the columns to the FP64 slope
    path, which absorbs them and says so (precision_used 0, precision_fallback).
    The run is made to stop by the Jacobi preconditioner in place of the firm
    blocks (XFE_SLOPE_JACOBI), which needs several times the 1,500 FP32
    iterations allowed here on this shape, and no FP64 continuation
    (XFE_SLOPE_NO_CONTINUATION); the FP64 path needs about 360."""
    _log("mixed slopes, fallback to the FP64 path")
    p, gids, G = _c2_panel(0.003, 7)
    V = np.column_stack([p["y"], p["X"]])
    z = p["z"]
    monkeypatch.setenv("XFE_SLOPE_JACOBI", "1")
    monkeypatch.setenv("XFE_SLOPE_NO_CONTINUATION", "1")
    ym, Xm, im = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=1, slopes={1: (z, True)},
                                max_iter=1500)
    monkeypatch.delenv("XFE_SLOPE_JACOBI")
    monkeypatch.delenv("XFE_SLOPE_NO_CONTINUATION")
    y6, X6, i6 = xfe_amd.absorb(V[:, 0], V[:, 1:], gids, G, tol=1e-8, precision=0, slopes={1: (z, True)},
                                max_iter=1500)
    _log(f"  fallback: converged {im['converged']} precision_used {im['precision_used']} "
         f"fallback {im['precision_fallback']} iterations {im['iterations']} (FP64 path {i6['iterations']})")
    assert i6["converged"]
    assert im["converged"] and im["precision_used"] == 0 and im["precision_fallback"], im
    # the same FP64 path both times; its level sums use atomics, so two runs
    # agree to the tolerance, not bit for bit
    Vm, V6 = np.column_stack([ym, Xm]), np.column_stack([y6, X6])
    assert np.abs(Vm - V6).max() <= 1e-8 * np.abs(V).max()
