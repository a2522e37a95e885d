"""Python binding (ctypes) of the MI355X absorber library ``libxfe_hip.so``.

The library is the product: HIP kernels for gfx950 behind the C ABI declared
in ``include/xfe_hip.h``.  This module only marshals arguments; it never
computes a result on the CPU.  Loading fails loudly when the library is not
built (``make -C xhdfe-xfe_amd`` or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libxfe_hip.so")
MAX_DIMS = 8

_i = ctypes.c_int
_d = ctypes.c_double
_p = ctypes.c_void_p


class FeInput(ctypes.Structure):
    """xfe_fe_input (mirrors hdfe::detail::GpuFeInput, src/fe_absorption_cuda.hpp:12-21)."""
    _fields_ = [("group_ids", _p), ("num_groups", _i), ("num_levels_present", _i),
                ("weight_sums", _p), ("is_slope", _i), ("slope_has_intercept", _i),
                ("slope_values", _p), ("slope_sum_z", _p), ("slope_sum_zz", _p)]


class Options(ctypes.Structure):
    _fields_ = [("tol", _d), ("tolerance_mode", _i), ("max_iter", _i), ("method", _i),
                ("force_cg", _i), ("precision", _i), ("certify", _i),
                ("refine_tol", _d), ("preconditioner", _i), ("lp_params", _p)]


class AbsorbResult(ctypes.Structure):
    _fields_ = [("iterations", _i), ("converged", _i), ("used_accel", _i), ("used_cg", _i),
                ("polish_sweeps", _i), ("polish_final_max", _d), ("sweep_order_used", _i * MAX_DIMS),
                ("fe_levels", _i * MAX_DIMS), ("sweeps", _i), ("refinements", _i), ("solve_ms", _d),
                ("abs_residual", _d), ("abs_residual_rel", _d), ("precision_certified", _i),
                ("certificate_computed", _i), ("precision_used", _i), ("precision_fallback", _i),
                ("refine_target", _d), ("preconditioner_used", _i), ("precond_fallback", _i),
                ("precond_blocks", _i), ("precond_ms", _d)]


class GroupIndividual(ctypes.Structure):
    _fields_ = [("num_groups", _i), ("num_individuals", _i), ("group_ptr", _p), ("group_individual", _p),
                ("individual_ptr", _p), ("individual_group", _p), ("group_scale", _p)]


class CertShadowResult(ctypes.Structure):
    _fields_ = [("verdict", _i), ("host_like_pass", _i), ("rel_max", _d), ("rel_max_perdim", _d),
                ("r_lower", _d), ("r_upper", _d), ("wsum_maxdiff", _d), ("wsum_tol", _d), ("wsum_ok", _i),
                ("empty_groups", ctypes.c_longlong), ("shadow_begin_ms", _d), ("shadow_finish_ms", _d),
                ("dims", _i), ("rhs", _i)]


class RecoveryResult(ctypes.Structure):
    _fields_ = [("iterations", _i), ("converged", _i), ("max_delta", _d), ("absorb_converged", _i)]


class Certificate(ctypes.Structure):
    _fields_ = [("abs_residual", _d), ("abs_residual_rel", _d), ("precision_certified", _i)]


# Every entry point of include/xfe_hip.h (the CPU test checks that each is exported).
EXPORTS = ("xfe_backend_available", "xfe_absorb", "xfe_absorb_device", "xfe_absorb_device_slopes",
           "xfe_drop_singletons",
           "xfe_bipartite_components", "xfe_fe_dof", "xfe_gram", "xfe_cluster_meat",
           "xfe_recovery_max_delta", "xfe_last_error", "xfe_profile_enable", "xfe_profile_reset",
           "xfe_profile_get", "xfe_release_workspace", "xfe_certify", "xfe_akm_available",
           "xfe_akm_stable_sort_u64", "xfe_akm_create", "xfe_akm_destroy", "xfe_akm_solve_S",
           "xfe_akm_solve_S_multi", "xfe_akm_max_lanes", "xfe_akm_solve_S_wide", "xfe_akm_solve_S_wide_device",
           "xfe_akm_wide_max_lanes", "xfe_boot_create", "xfe_boot_destroy", "xfe_boot_refits", "xfe_boot_last",
           "xfe_akm_wide_last", "xfe_akm_exact_leverages", "xfe_akm_exact_leverages_last_retries", "xfe_akm_pack_buffer", "xfe_schwarz_demean",
           "xfe_akm_jla_pii", "xfe_akm_jla_leverages", "xfe_akm_jla_bii", "xfe_recover_fixed_effects",
           "xfe_absorb_group_individual",
           "xfe_cert_shadow_begin", "xfe_cert_shadow_finish", "xfe_cert_shadow_abort",
           "xfe_articulation_points", "xfe_akm_leave_out_set", "xfe_build_indexer",
           "xfe_build_indexer_device", "xfe_weight_sums", "xfe_build_indexer_device_noalloc",
           "xfe_lp_param_count", "xfe_lp_gnn_eval", "xfe_lp_probe", "xfe_lp_probe_fetch", "xfe_lp_train", "xfe_lp_blocks_apply")

_lib = None


class XfeError(RuntimeError):
    pass


def lib():
    """Load libxfe_hip.so; raises if it is missing (no CPU fallback exists).

    PyTorch-ROCm ships its own libamdhip64.so.7 with the same soname as the
    system runtime this library links; a process can hold only one.  torch is
    imported first (when installed) so that torch and this library share
    torch's runtime; loading the system one first leaves torch without a
    usable device."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise XfeError(f"{LIB_PATH} is not built; run make -C xhdfe-xfe_amd")
        L = ctypes.CDLL(LIB_PATH)
        L.xfe_last_error.restype = ctypes.c_char_p
        L.xfe_profile_get.argtypes = [ctypes.POINTER(_d), ctypes.POINTER(ctypes.c_longlong),
                                      ctypes.POINTER(_d)]
        _lib = L
    return _lib


def last_error():
    return lib().xfe_last_error().decode()


def backend_available():
    return bool(lib().xfe_backend_available())



def _i32(a):
    """int32 copy of an integer array for the C ABI; values outside the int32
    range raise XfeError instead of wrapping (a wrapped id would merge levels)."""
    arr = np.asarray(a)
    if arr.dtype != np.int32 and arr.size:
        if arr.dtype.kind not in "iub":
            if arr.dtype.kind == "f" and np.all(np.isfinite(arr)) and np.all(arr == np.floor(arr)):
                pass
            else:
                raise XfeError(f"integer ids expected, got dtype {arr.dtype}")
        lo, hi = arr.min(), arr.max()
        info = np.iinfo(np.int32)
        if lo < info.min or hi > info.max:
            raise XfeError(f"ids outside the int32 range ([{lo}, {hi}]); compact them first")
    return np.ascontiguousarray(arr, dtype=np.int32)

def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


def _opts(tol, tolerance_mode, max_iter, method, force_cg, precision, certify=True, refine_tol=0.0,
          preconditioner=0, lp_params=None):
    """xfe_options; lp_params (float32 array) must stay alive for the call."""
    return Options(float(tol), int(tolerance_mode), int(max_iter), int(method), int(bool(force_cg)),
                   int(precision), int(bool(certify)), float(refine_tol), int(preconditioner),
                   None if lp_params is None else lp_params.ctypes.data_as(_p))


def _lp_params(preconditioner, lp_params):
    """The network parameters for preconditioner 2: the given ones or the
    trained defaults shipped with the package."""
    if int(preconditioner) != 2:
        return None
    if lp_params is None:
        return default_lp_params()
    a = np.ascontiguousarray(lp_params, dtype=np.float32)
    if a.size != lp_param_count():
        raise XfeError(f"lp_params: {a.size} values, the network has {lp_param_count()}")
    return a


def _info(r, D):
    return dict(iterations=r.iterations, converged=bool(r.converged), used_accel=bool(r.used_accel),
                used_cg=bool(r.used_cg), polish_sweeps=r.polish_sweeps,
                polish_final_max=r.polish_final_max, sweep_order=list(r.sweep_order_used)[:D],
                fe_levels=list(r.fe_levels)[:D], sweeps=r.sweeps, refinements=r.refinements,
                solve_ms=r.solve_ms, abs_residual=r.abs_residual, abs_residual_rel=r.abs_residual_rel,
                precision_certified=bool(r.precision_certified),
                certificate_computed=bool(r.certificate_computed), precision_used=r.precision_used,
                precision_fallback=bool(r.precision_fallback), refine_target=r.refine_target,
                preconditioner_used=r.preconditioner_used, precond_fallback=bool(r.precond_fallback),
                precond_blocks=r.precond_blocks, precond_ms=r.precond_ms)


def absorb(y, X, gids, G, weight_sums=None, weights=None, sweep_order=None, tol=1e-8,
           tolerance_mode=1, max_iter=100000, method=None, force_cg=False, precision=0, certify=True,
           refine_tol=0.0, slopes=None, preconditioner=0, lp_params=None):
    """absorb_fixed_effects_cuda on host arrays: returns (y_tilde, X_tilde, info).

    gids: compact 0-based level ids per dimension; G: level counts.  ``method``
    defaults to the reference's choice (SymmetricGaussSeidel for two FEs).
    slopes: {dimension: (z, with_intercept)} heterogeneous slope terms."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    X = np.asfortranarray(np.asarray(X, dtype=np.float64).reshape(y.size, -1))
    n, k = X.shape
    D = len(gids)
    ids = [_i32(g) for g in gids]
    ws = [None if weight_sums is None else np.ascontiguousarray(weight_sums[d], dtype=np.float64)
          for d in range(D)]
    fes = (FeInput * max(D, 1))()
    for d in range(D):
        fes[d].group_ids = ids[d].ctypes.data
        fes[d].num_groups = int(G[d])
        fes[d].num_levels_present = int(G[d])
        fes[d].weight_sums = None if ws[d] is None else ws[d].ctypes.data
    zs = {}
    for d, (z, icpt) in (slopes or {}).items():
        zs[d] = np.ascontiguousarray(z, dtype=np.float64)
        fes[d].is_slope = 1
        fes[d].slope_has_intercept = int(bool(icpt))
        fes[d].slope_values = zs[d].ctypes.data
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    order = None if sweep_order is None else _i32(sweep_order)
    if method is None:
        method = 2 if D == 2 else 1
    lpp = _lp_params(preconditioner, lp_params)
    o = _opts(tol, tolerance_mode, max_iter, method, force_cg, precision, certify, refine_tol, preconditioner, lpp)
    r = AbsorbResult()
    yo = np.empty(n)
    Xo = np.empty((n, k), order="F")
    ok = lib().xfe_absorb(_ptr(y), _ptr(X), _i(n), _i(k), fes, _i(D), _ptr(w), _ptr(order),
                          ctypes.byref(o), ctypes.byref(r), _ptr(yo), _ptr(Xo))
    if not ok:
        raise XfeError(last_error())
    return yo, Xo, _info(r, D)


def absorb_device_slopes(d_V, n, C, ld, d_ids, G, slope_ptrs, slope_mode, tol=1e-8, tolerance_mode=1,
                         max_iter=100000, method=1, precision=1, certify=True, refine_tol=0.0):
    """xfe_absorb_device_slopes on device pointers: d_V (n x C, ld), d_ids
    (num_fes x n int32), slope_ptrs[d] (device pointer or 0), slope_mode[d]
    (0 standard, 1 slope with intercept, 2 slope only).  Returns info."""
    D = len(G)
    Garr = np.ascontiguousarray(G, dtype=np.int32)
    modes = np.ascontiguousarray(slope_mode, dtype=np.int32)
    zp = (ctypes.c_void_p * D)(*[int(p) if p else None for p in slope_ptrs])
    o = _opts(tol, tolerance_mode, max_iter, method, False, precision, certify, refine_tol)
    r = AbsorbResult()
    L = lib()
    L.xfe_absorb_device_slopes.argtypes = [_p, _i, _i, _i, _p, _p, _i, _p, _p, _p, _p, _p, _p]
    ok = L.xfe_absorb_device_slopes(_p(d_V), n, C, ld, _p(d_ids), _ptr(Garr), D, zp, _ptr(modes), None, None,
                                    ctypes.byref(o), ctypes.byref(r))
    if not ok:
        raise XfeError(last_error())
    return _info(r, D)


def absorb_group_individual(y, X, gids, G, gi, weights=None, sweep_order=None, tol=1e-8, tolerance_mode=1,
                            max_iter=100000, method=1):
    """absorb_fixed_effects_group_individual_cuda on host arrays: returns
    (y_tilde, X_tilde, info).  gi: dict with group_ptr, group_individual,
    individual_ptr, individual_group (int32), group_scale (float64),
    num_individuals."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    X = np.asfortranarray(np.asarray(X, dtype=np.float64).reshape(y.size, -1))
    n, k = X.shape
    D = len(gids)
    ids = [_i32(g) for g in gids]
    fes = (FeInput * max(D, 1))()
    for d in range(D):
        fes[d].group_ids = ids[d].ctypes.data
        fes[d].num_groups = int(G[d])
        fes[d].num_levels_present = int(G[d])
    arrs = {key: _i32(gi[key])
            for key in ("group_ptr", "group_individual", "individual_ptr", "individual_group")}
    scale = np.ascontiguousarray(gi["group_scale"], dtype=np.float64)
    s = GroupIndividual(n, int(gi["num_individuals"]), arrs["group_ptr"].ctypes.data,
                        arrs["group_individual"].ctypes.data, arrs["individual_ptr"].ctypes.data,
                        arrs["individual_group"].ctypes.data, scale.ctypes.data)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    order = None if sweep_order is None else _i32(sweep_order)
    o = _opts(tol, tolerance_mode, max_iter, method, False, 0, True, 0.0)
    r = AbsorbResult()
    yo = np.empty(n)
    Xo = np.empty((n, k), order="F")
    ok = lib().xfe_absorb_group_individual(_ptr(y), _ptr(X), _i(n), _i(k), fes, _i(D), ctypes.byref(s), _ptr(w),
                                           _ptr(order), ctypes.byref(o), ctypes.byref(r), _ptr(yo), _ptr(Xo))
    if not ok:
        raise XfeError(last_error())
    return yo, Xo, _info(r, D)


def build_indexer(raw):
    """build_indexer on the device: (compact ids, num_groups, levels_present, contiguous)."""
    r = _i32(raw)
    gid = np.empty(max(r.size, 1), dtype=np.int32)
    G, lp, cont = _i(), _i(), _i()
    if not lib().xfe_build_indexer(_ptr(r) if r.size else None, _i(r.size), _ptr(gid), ctypes.byref(G),
                                   ctypes.byref(lp), ctypes.byref(cont)):
        raise XfeError(last_error())
    return gid[:r.size], G.value, lp.value, bool(cont.value)


def build_indexer_device(raw_t):
    """build_indexer on a device tensor of int32 ids: (ids tensor, num_groups, levels_present, contiguous)."""
    import torch
    out = torch.empty_like(raw_t)
    G, lp, cont = _i(), _i(), _i()
    if not lib().xfe_build_indexer_device(_p(raw_t.data_ptr()), _i(raw_t.numel()), _p(out.data_ptr()),
                                          ctypes.byref(G), ctypes.byref(lp), ctypes.byref(cont)):
        raise XfeError(last_error())
    return out, G.value, lp.value, bool(cont.value)


def weight_sums(gid, G, weights=None):
    """compute_weight_sums on the device: level sums of the weights (counts when None)."""
    g = _i32(gid)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    out = np.empty(int(G))
    if not lib().xfe_weight_sums(_ptr(g) if g.size else None, _i(g.size), _i(int(G)), _ptr(w), _ptr(out)):
        raise XfeError(last_error())
    return out


def articulation_points(adj_ptr, adj):
    """Cut vertices of an undirected CSR graph (uint8 flag per vertex)."""
    ptr = _i32(adj_ptr)
    a = _i32(adj)
    n = ptr.size - 1
    out = np.zeros(max(n, 1), dtype=np.uint8)
    if not lib().xfe_articulation_points(_i(n), _ptr(ptr), _ptr(a) if a.size else None, _ptr(out)):
        raise XfeError(last_error())
    return out[:n]


class LeaveOutResult(ctypes.Structure):
    _fields_ = [("n_obs_input", ctypes.c_longlong), ("n_obs_connected", ctypes.c_longlong),
                ("n_obs", ctypes.c_longlong), ("n_workers", _i), ("n_firms", _i), ("n_matches", ctypes.c_longlong),
                ("n_movers", _i), ("n_stayers", _i), ("prune_iterations", _i), ("gpu_used", _i)]


def leave_out_set(worker_ids, firm_ids, fweights=None):
    """leave_out_connected_set: (keep mask per input row, info dict)."""
    w = _i32(worker_ids)
    f = _i32(firm_ids)
    fw = None if fweights is None else np.ascontiguousarray(fweights, dtype=np.int64)
    keep = np.zeros(max(w.size, 1), dtype=np.uint8)
    r = LeaveOutResult()
    if not lib().xfe_akm_leave_out_set(_ptr(w), _ptr(f), _i(w.size), _ptr(fw), _ptr(keep), ctypes.byref(r)):
        raise XfeError(last_error())
    return keep[:w.size].astype(bool), {name: getattr(r, name) for name, _ in LeaveOutResult._fields_}


VERDICTS = ("PASS", "FAIL", "INDETERMINATE", "ERROR")


def cert_shadow_raw(y0, x0, y_res, x_res, n, k, ld, gid_ptrs, G, limit, w=None, wsum_ptrs=None):
    """The shadow device certificate on device addresses: original and
    residual columns (y, then X column-major with leading dimension ld),
    per-dimension id and weight-sum addresses.  Returns the result fields as
    a dict ("verdict" as a name)."""
    L = lib()
    D = len(gid_ptrs)
    gp = (_p * D)(*gid_ptrs)
    wsp = None if wsum_ptrs is None else (_p * D)(*wsum_ptrs)
    Garr = (_i * D)(*[int(x) for x in G])
    L.xfe_cert_shadow_begin.restype = _p
    L.xfe_cert_shadow_begin.argtypes = [_p, _p, _i, _i, _i, _p, _p, _p, _p, _i, _p, _p]
    h = L.xfe_cert_shadow_begin(y0, x0 if k else None, n, k, ld, w, gp, wsp, Garr, D, None, None)
    if not h:
        raise XfeError(last_error())
    r = CertShadowResult()
    L.xfe_cert_shadow_finish.argtypes = [_p, _p, _p, _d, ctypes.c_void_p]
    if not L.xfe_cert_shadow_finish(h, y_res, x_res if k else None, float(limit), ctypes.byref(r)):
        raise XfeError(last_error())
    out = {name: getattr(r, name) for name, _ in CertShadowResult._fields_}
    out["verdict"] = VERDICTS[r.verdict]
    return out


def cert_shadow(y0, X0, y_res, X_res, gids, G, limit, weights=None, weight_sums=None):
    """cert_shadow_raw on torch device tensors: y (n), X (k x n, each row one
    column, so ld = n), gids (int32 device tensors), optional weights (n) and
    weight_sums (per dimension)."""
    n = int(y0.numel())
    k = int(X0.shape[0]) if X0 is not None and X0.numel() else 0
    return cert_shadow_raw(y0.data_ptr(), X0.data_ptr() if k else None, y_res.data_ptr(),
                           X_res.data_ptr() if k else None, n, k, n, [g.data_ptr() for g in gids], G, limit,
                           w=weights.data_ptr() if weights is not None else None,
                           wsum_ptrs=None if weight_sums is None else [t.data_ptr() for t in weight_sums])


def recover_fixed_effects(partial, gids, G, weights=None, sweep_order=None, tol=1e-8, fe_tol=1e-6,
                          tolerance_mode=1, max_iter=100000, refine_tol=0.0, contributions=True):
    """recover_fixed_effects_group_ids' GPU branch on host arrays: returns
    (alphas, contributions or None, info); alphas[d] holds the level
    coefficients of dimension d, contributions is D x n."""
    partial = np.ascontiguousarray(partial, dtype=np.float64)
    n, D = partial.size, len(gids)
    ids = [_i32(g) for g in gids]
    fes = (FeInput * max(D, 1))()
    for d in range(D):
        fes[d].group_ids = ids[d].ctypes.data
        fes[d].num_groups = int(G[d])
        fes[d].num_levels_present = int(G[d])
    alphas = [np.empty(int(G[d])) for d in range(D)]
    aptr = (_p * max(D, 1))(*[a.ctypes.data for a in alphas])
    contrib = np.empty((D, n)) if contributions else None
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    order = None if sweep_order is None else _i32(sweep_order)
    o = _opts(tol, tolerance_mode, max_iter, 2 if D == 2 else 1, False, 1, False, refine_tol)
    r = RecoveryResult()
    ok = lib().xfe_recover_fixed_effects(_ptr(partial), _i(n), fes, _i(D), _ptr(w), _ptr(order),
                                         ctypes.byref(o), _d(fe_tol), aptr, _ptr(contrib), ctypes.byref(r))
    if not ok:
        raise XfeError(last_error())
    return alphas, contrib, dict(iterations=r.iterations, converged=bool(r.converged),
                                 max_delta=r.max_delta, absorb_converged=bool(r.absorb_converged))


def absorb_device(V_ptr, n, C, ld, gid_ptr, G, weights_ptr=None, sweep_order=None, tol=1e-8,
                  tolerance_mode=1, max_iter=100000, method=None, force_cg=False, precision=0,
                  certify=True, preconditioner=0, lp_params=None):
    """xfe_absorb_device on device pointers (e.g. torch tensors' data_ptr())."""
    D = len(G)
    Garr = _i32(G)
    order = None if sweep_order is None else _i32(sweep_order)
    if method is None:
        method = 2 if D == 2 else 1
    lpp = _lp_params(preconditioner, lp_params)
    o = _opts(tol, tolerance_mode, max_iter, method, force_cg, precision, certify, 0.0, preconditioner, lpp)
    r = AbsorbResult()
    ok = lib().xfe_absorb_device(_p(V_ptr), _i(n), _i(C), _i(ld), _p(gid_ptr), _ptr(Garr), _i(D),
                                 _p(weights_ptr) if weights_ptr else None, _ptr(order),
                                 ctypes.byref(o), ctypes.byref(r))
    if not ok:
        raise XfeError(last_error())
    return _info(r, D)


def certify(y, X, y_t, X_t, gids, G, weights=None, tol=1e-8, tolerance_mode=1, converged=True):
    """xfe_certify (certify_cuda_absorption_result / certify_absorption_result,
    src/fe_absorption.cpp:14942-15190) on host arrays: dict(abs_residual,
    abs_residual_rel, certified)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    X = np.asfortranarray(np.asarray(X, dtype=np.float64).reshape(y.size, -1))
    yt = np.ascontiguousarray(y_t, dtype=np.float64)
    Xt = np.asfortranarray(np.asarray(X_t, dtype=np.float64).reshape(y.size, -1))
    n, k = X.shape
    D = len(gids)
    ids = [_i32(g) for g in gids]
    fes = (FeInput * max(D, 1))()
    for d in range(D):
        fes[d].group_ids = ids[d].ctypes.data
        fes[d].num_groups = int(G[d])
        fes[d].num_levels_present = int(G[d])
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    o = _opts(tol, tolerance_mode, 0, 0, False, 0, True)
    out = Certificate()
    ok = lib().xfe_certify(_ptr(y), _ptr(X), _ptr(yt), _ptr(Xt), _i(n), _i(k), fes, _i(D), _ptr(w),
                           ctypes.byref(o), _i(int(bool(converged))), ctypes.byref(out))
    if not ok:
        raise XfeError(last_error())
    return dict(abs_residual=out.abs_residual, abs_residual_rel=out.abs_residual_rel,
                certified=bool(out.precision_certified))


def schwarz_demean(Y, X, fe, nlev, tol=1e-8, max_iter=10000, threads=1, collapse=-1):
    """xhdfe::schwarz_demean_raw (src/schwarz_demean.cpp:419-438) through
    xfe_schwarz_demean: demeans Y (n) and X (n x k, column-major) in place
    when it converges.  fe: one int32 id array per dimension, ids in
    [0, nlev[d]).  Returns (converged, iterations), like the reference's bool
    and iters_out."""
    Y = np.asarray(Y)
    if Y.dtype != np.float64 or not Y.flags.c_contiguous:
        raise XfeError("Y must be a contiguous float64 array")
    n = Y.size
    if X is None:
        k = 0
        Xp = None
    else:
        if X.dtype != np.float64 or not X.flags.f_contiguous or X.shape[0] != n:
            raise XfeError("X must be a column-major float64 (n, k) array")
        k = X.shape[1] if X.ndim == 2 else 1
        Xp = _ptr(X)
    ids = [_i32(f) for f in fe]
    ptrs = (_p * max(len(ids), 1))(*[a.ctypes.data for a in ids])
    lev = _i32(nlev)
    it = _i(0)
    L = lib()
    L.xfe_schwarz_demean.argtypes = [_p, _p, ctypes.c_int64, _i, _p, _p, _i, _i, _d, _i, _p, _i]
    ok = L.xfe_schwarz_demean(_ptr(Y), Xp, n, k, ptrs, _ptr(lev), len(ids), int(threads), float(tol),
                              int(max_iter), ctypes.byref(it), int(collapse))
    return bool(ok), it.value


class BootRefits:
    """xfe_boot_create / _refits / _destroy: cluster (worker) bootstrap refits
    of the two-way FE regression of V[:, 0] on V[:, 1:], 64 refits per
    operator pass on the wide solver."""

    def __init__(self, worker, firm, V, n_workers=None, n_firms=None):
        w = _i32(worker)
        f = _i32(firm)
        V = np.asarray(V, dtype=np.float64)
        if V.ndim != 2 or V.shape[0] != w.size or f.size != w.size or not 1 <= V.shape[1] <= 4:
            raise XfeError("BootRefits: V must be n x C with C <= 4 and one row per id")
        self.N = int(n_workers if n_workers is not None else w.max() + 1)
        self.J = int(n_firms if n_firms is not None else f.max() + 1)
        self.C = V.shape[1]
        self._V = np.asfortranarray(V)
        self._w, self._f = w, f
        L = lib()
        L.xfe_boot_create.restype = _p
        L.xfe_boot_create.argtypes = [_i, _i, _i, _p, _p, _p, _i, _i]
        self._ctx = L.xfe_boot_create(w.size, self.N, self.J, _ptr(w), _ptr(f), _ptr(self._V), w.size, self.C)
        if not self._ctx:
            raise XfeError(last_error())

    def refits(self, mult, tol=1e-10, max_iter=5000):
        """mult: (R, n_workers) multiplicities; returns (gram (R, C, C),
        iterations (R,), (fp32 passes, fp64 replacements), converged)."""
        m = np.ascontiguousarray(mult, dtype=np.int32)
        if m.ndim != 2 or m.shape[1] != self.N:
            raise XfeError("refits: mult must be R x n_workers")
        R = m.shape[0]
        gram = np.empty((R, self.C, self.C))
        its = np.empty(R, dtype=np.int32)
        L = lib()
        L.xfe_boot_refits.argtypes = [_p, _i, _p, _d, _i, _p, _p]
        rc = L.xfe_boot_refits(_p(self._ctx), R, _ptr(m), float(tol), int(max_iter), _ptr(gram), _ptr(its))
        if rc < 0:
            raise XfeError(last_error())
        passes, upd = ctypes.c_int(0), ctypes.c_int(0)
        L.xfe_boot_last(_p(self._ctx), ctypes.byref(passes), ctypes.byref(upd))
        return gram, its, (passes.value, upd.value), rc == 0

    @staticmethod
    def betas(gram):
        """beta of each refit from its Gram (y first): G[1:,1:]^-1 G[1:,0]."""
        return np.stack([np.linalg.solve(g[1:, 1:], g[1:, 0]) for g in gram])

    def close(self):
        if self._ctx:
            lib().xfe_boot_destroy(_p(self._ctx))
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AkmSolver:
    """xfe_akm_create / _solve_S(_multi) / _destroy (the reference's akm_cuda_*
    context, include/hdfe/akm_kss_cuda.hpp) from a structure dict with keys
    N, J, m_w, m_f, m_c, w_ptr, f_ptr, f_matches, Dw, Df, inv_diag."""

    def __init__(self, st):
        L = lib()
        L.xfe_akm_create.restype = _p
        L.xfe_akm_pack_buffer.restype = _p
        self._keep = {k: np.ascontiguousarray(st[k], dtype=t) for k, t in (
            ("m_w", np.int32), ("m_f", np.int32), ("m_c", np.float64), ("w_ptr", np.int64),
            ("f_ptr", np.int64), ("f_matches", np.int64), ("Dw", np.float64), ("Df", np.float64),
            ("inv_diag", np.float64))}
        k = self._keep
        self.J = int(st["J"])
        self._ctx = L.xfe_akm_create(_i(int(st["N"])), _i(self.J), _ptr(k["m_w"]), _ptr(k["m_f"]),
                                     _ptr(k["m_c"]), ctypes.c_size_t(k["m_w"].size), _ptr(k["w_ptr"]),
                                     _ptr(k["f_ptr"]), _ptr(k["f_matches"]), _ptr(k["Dw"]), _ptr(k["Df"]),
                                     _ptr(k["inv_diag"]))
        if not self._ctx:
            raise XfeError(last_error())

    def solve(self, rhs, tol=1e-10, max_iter=1000):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        z = np.empty_like(rhs)
        it = lib().xfe_akm_solve_S(_p(self._ctx), _ptr(rhs), _ptr(z), _d(tol), _i(max_iter))
        return z, it

    def solve_multi(self, rhs_lanes, tol=1e-10, max_iter=1000):
        """rhs_lanes: (nb, J) array, one lane per row; returns (z (nb, J), iterations)."""
        rhs = np.ascontiguousarray(rhs_lanes, dtype=np.float64)
        nb = rhs.shape[0]
        z = np.empty_like(rhs)
        its = np.empty(nb, dtype=np.int32)
        rc = lib().xfe_akm_solve_S_multi(_p(self._ctx), _ptr(rhs), _ptr(z), _i(nb), _d(tol), _i(max_iter),
                                         _ptr(its))
        if rc != 0:
            raise XfeError(last_error())
        return z, its

    def solve_wide(self, rhs_lanes, tol=1e-10, max_iter=1000):
        """xfe_akm_solve_S_wide: up to 256 lanes per operator pass, FP32 PCG
        under FP64 residual replacement.  rhs_lanes: (nb, J); returns
        (z (nb, J), iterations, (fp32 passes, fp64 replacements))."""
        rhs = np.ascontiguousarray(rhs_lanes, dtype=np.float64)
        nb = rhs.shape[0]
        z = np.empty_like(rhs)
        its = np.empty(nb, dtype=np.int32)
        L = lib()
        rc = L.xfe_akm_solve_S_wide(_p(self._ctx), _ptr(rhs), _ptr(z), _i(nb), _d(tol), _i(max_iter), _ptr(its))
        if rc != 0:
            raise XfeError(last_error())
        passes, upd = ctypes.c_int(0), ctypes.c_int(0)
        L.xfe_akm_wide_last(_p(self._ctx), ctypes.byref(passes), ctypes.byref(upd))
        return z, its, (passes.value, upd.value)

    def solve_wide_device(self, d_rhs, d_z, nb, tol=1e-10, max_iter=1000):
        """xfe_akm_solve_S_wide_device on device pointers (nb x J lane-major
        FP64); returns (iterations, (fp32 passes, fp64 replacements))."""
        its = np.empty(nb, dtype=np.int32)
        L = lib()
        L.xfe_akm_solve_S_wide_device.argtypes = [_p, _p, _p, _i, _d, _i, _p]
        rc = L.xfe_akm_solve_S_wide_device(_p(self._ctx), _p(d_rhs), _p(d_z), int(nb), float(tol), int(max_iter),
                                           _ptr(its))
        if rc != 0:
            raise XfeError(last_error())
        passes, upd = ctypes.c_int(0), ctypes.c_int(0)
        L.xfe_akm_wide_last(_p(self._ctx), ctypes.byref(passes), ctypes.byref(upd))
        return its, (passes.value, upd.value)

    def exact_leverages(self, tol=1e-10, max_iter=10000, accept=0.0):
        """xfe_akm_exact_leverages: the exact P_ii and B_ii (fe, pe, cov) of
        the matches (in the given order), one solve per match, 256 per wide
        solve; returns (pii, bfe, bpe, bcov, number of unconverged solves)."""
        M = self._keep["m_w"].size
        out = [np.empty(M) for _ in range(4)]
        L = lib()
        L.xfe_akm_exact_leverages.argtypes = [_p, _d, _d, _i, _p, _p, _p, _p]
        rc = L.xfe_akm_exact_leverages(_p(self._ctx), float(tol), float(accept), int(max_iter),
                                       *[_ptr(o) for o in out])
        if rc < 0:
            raise XfeError(last_error())
        self.exact_retries = int(L.xfe_akm_exact_leverages_last_retries())
        return out[0], out[1], out[2], out[3], int(rc)

    def jla_pii(self, draws=200, seed=20260705, tol=1e-10, max_iter=10000):
        """xfe_akm_jla_pii: JLA leverages of the matches (in the given order)
        from `draws` Rademacher draws of the reference's streams."""
        pii = np.empty(self._keep["m_w"].size)
        L = lib()
        L.xfe_akm_jla_pii.argtypes = [_p, _i, ctypes.c_uint64, _d, _i, _p]
        rc = L.xfe_akm_jla_pii(_p(self._ctx), int(draws), int(seed), float(tol), int(max_iter), _ptr(pii))
        if rc < 0:
            raise XfeError(last_error())
        return pii, rc == 0

    def jla_leverages(self, draws=200, seed=20260705, tol=1e-10, max_iter=10000):
        """xfe_akm_jla_leverages: JLA leverages of the matches and the ratio
        estimator's non-linearity correction; returns (pii, corr, converged)."""
        M = self._keep["m_w"].size
        pii = np.empty(M)
        corr = np.empty(M)
        L = lib()
        L.xfe_akm_jla_leverages.argtypes = [_p, _i, ctypes.c_uint64, _d, _i, _p, _p]
        rc = L.xfe_akm_jla_leverages(_p(self._ctx), int(draws), int(seed), float(tol), int(max_iter), _ptr(pii),
                                     _ptr(corr))
        if rc < 0:
            raise XfeError(last_error())
        return pii, corr, rc == 0

    def jla_bii(self, py_count=None, draws=200, seed=20260705, tol=1e-10, max_iter=10000, n_py=0.0):
        """xfe_akm_jla_bii: JLA estimates of the KSS weights (Bfe, Bpe, Bcov)
        of the matches from `draws` draws of the reference's streams.
        py_count: person-year rows per match (default: the match weights,
        which must then be integral)."""
        if py_count is None:
            c = np.asarray(self._keep["m_c"], dtype=np.float64)
            py_count = np.rint(c)
            if np.any(py_count != c):
                raise XfeError("jla_bii: non-integral match weights need py_count")
        cnt = np.ascontiguousarray(py_count, dtype=np.int64)
        M = self._keep["m_w"].size
        if cnt.shape != (M,):
            raise XfeError("jla_bii: py_count needs one entry per match")
        out = [np.empty(M) for _ in range(3)]
        L = lib()
        L.xfe_akm_jla_bii.argtypes = [_p, _i, ctypes.c_uint64, _d, _i, _p, _d, _p, _p, _p]
        rc = L.xfe_akm_jla_bii(_p(self._ctx), int(draws), int(seed), float(tol), int(max_iter), _ptr(cnt),
                               float(n_py), *(_ptr(a) for a in out))
        if rc < 0:
            raise XfeError(last_error())
        return out[0], out[1], out[2], rc == 0

    def close(self):
        if self._ctx:
            lib().xfe_akm_destroy(_p(self._ctx))
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def akm_stable_sort_u64(keys):
    """xfe_akm_stable_sort_u64: stable ascending order of uint64 keys."""
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    out = np.empty(keys.size, dtype=np.int64)
    if not lib().xfe_akm_stable_sort_u64(_ptr(keys), _ptr(out), ctypes.c_size_t(keys.size)):
        raise XfeError(last_error())
    return out


def drop_singletons(fes, max_rounds=100):
    """compute_keep_mask_drop_singletons on the device: (keep mask, dropped, rounds)."""
    ids = np.ascontiguousarray(np.stack([_i32(f) for f in fes]))
    D, n = ids.shape
    keep = np.empty(n, dtype=np.uint8)
    rounds = _i()
    dropped = lib().xfe_drop_singletons(_ptr(ids), _i(n), _i(D), _i(max_rounds), _ptr(keep),
                                        ctypes.byref(rounds))
    if dropped < 0:
        raise XfeError(last_error())
    return keep.astype(bool), dropped, rounds.value


def components(a, b, expected_a=-1, expected_b=-1, labels=True):
    """count_bipartite_components on the device: (count, 1-based labels or None)."""
    a = _i32(a)
    b = _i32(b)
    lab = np.empty(a.size, dtype=np.int32) if labels else None
    c = lib().xfe_bipartite_components(_ptr(a), _ptr(b), _i(a.size), _i(expected_a),
                                       _i(expected_b), _ptr(lab))
    if c < 0:
        raise XfeError(last_error())
    return c, lab


def fe_dof(fes, levels):
    """compute_fe_dof_reghdfe (pairwise) on the device: (df_a, redundant, num_coefs)."""
    ids = np.ascontiguousarray(np.stack([_i32(f) for f in fes]))
    D, n = ids.shape
    lv = _i32(levels)
    red = np.empty(D, dtype=np.int32)
    nc = np.empty(D, dtype=np.int32)
    df = lib().xfe_fe_dof(_ptr(ids), _i(n), _i(D), _ptr(lv), _ptr(red), _ptr(nc))
    if df < 0:
        raise XfeError(last_error())
    return df, red, nc


def gram(V):
    V = np.asfortranarray(V, dtype=np.float64)
    n, C = V.shape
    out = np.empty((C, C), order="F")
    if lib().xfe_gram(_ptr(V), _i(n), _i(C), _ptr(out)) != 0:
        raise XfeError(last_error())
    return out


def cluster_meat(V, beta, cluster_ids, num_clusters):
    V = np.asfortranarray(V, dtype=np.float64)
    n, C = V.shape
    b = np.ascontiguousarray(beta, dtype=np.float64)
    cl = _i32(cluster_ids)
    out = np.empty((C - 1, C - 1), order="F")
    if lib().xfe_cluster_meat(_ptr(V), _i(n), _i(C), _ptr(b), _ptr(cl), _i(num_clusters),
                              _ptr(out)) != 0:
        raise XfeError(last_error())
    return out


def profile_enable(on=True):
    lib().xfe_profile_enable(_i(int(on)))


def profile_reset():
    lib().xfe_profile_reset()


def profile_get():
    ms, n, b = _d(), ctypes.c_longlong(), _d()
    lib().xfe_profile_get(ctypes.byref(ms), ctypes.byref(n), ctypes.byref(b))
    return ms.value, n.value, b.value


_LP_DEFAULT = None
LP_PARAMS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lp_params.f32")


def default_lp_params():
    """The trained parameters of the preconditioner network (lp_params.f32,
    written by train_lp.py)."""
    global _LP_DEFAULT
    if _LP_DEFAULT is None:
        if not os.path.exists(LP_PARAMS_FILE):
            raise XfeError(f"no trained preconditioner parameters at {LP_PARAMS_FILE}")
        a = np.fromfile(LP_PARAMS_FILE, dtype="<f4")
        if a.size != lp_param_count():
            raise XfeError(f"{LP_PARAMS_FILE}: {a.size} values, the network has {lp_param_count()}")
        _LP_DEFAULT = a
    return _LP_DEFAULT


def lp_param_count():
    return int(lib().xfe_lp_param_count())


def lp_gnn_eval(g, theta):
    """The learned preconditioner's network on the device: (loss, grad, u, t)
    for a graph dict (xf, xw, f_ptr, f_idx, w_ptr, w_idx, pf, pg, pq optional,
    u_tgt, u_mask, y_tgt) and float32 parameters."""
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    xf, xw = f32(g["xf"]), f32(g["xw"])
    F, Wm = xf.shape[0], xw.shape[0]
    fp, fi, wp, wi = (_i32(g[k]) for k in ("f_ptr", "f_idx", "w_ptr", "w_idx"))
    pf, pg = _i32(g["pf"]), _i32(g["pg"])
    pq = f32(g["pq"]) if g.get("pq") is not None else None
    ut, um, yt, th = f32(g["u_tgt"]), f32(g["u_mask"]), f32(g["y_tgt"]), f32(theta)
    npair = pf.size
    loss = _d()
    grad = np.empty(lp_param_count(), dtype=np.float32)
    u = np.empty(max(F, 1), dtype=np.float32)
    t = np.empty(max(npair, 1), dtype=np.float32)
    L = lib()
    L.xfe_lp_gnn_eval.argtypes = [_i, _i] + [_p] * 6 + [_i] + [_p] * 3 + [_p] * 4 + [ctypes.POINTER(_d)] + [_p] * 3
    ok = L.xfe_lp_gnn_eval(F, Wm, _ptr(xf), _ptr(xw), _ptr(fp), _ptr(fi), _ptr(wp), _ptr(wi), npair, _ptr(pf),
                           _ptr(pg), _ptr(pq), _ptr(ut), _ptr(um), _ptr(yt), _ptr(th), ctypes.byref(loss),
                           _ptr(grad), _ptr(u), _ptr(t))
    if not ok:
        raise XfeError(last_error())
    return loss.value, grad, u[:F], t[:npair]


class LpPanelSizes(ctypes.Structure):
    _fields_ = [("F", _i), ("Wm", _i), ("E", _i), ("U", _i), ("C", _i), ("inner_iterations", _i),
                ("converged", _i)]


class LpGraph(ctypes.Structure):
    _fields_ = [("F", _i), ("Wm", _i), ("U", _i)] + [(k, _p) for k in (
        "xf", "xw", "f_ptr", "f_idx", "w_ptr", "w_idx", "pf", "pg", "pq", "u_tgt", "u_mask", "y_tgt")]


def lp_probe(y, X, gids, G, probe_iter=8):
    """xfe_lp_probe + xfe_lp_probe_fetch: the absorber's iterate errors at inner
    iteration probe_iter on the firm levels (the second dimension), and the
    preconditioner's pairs, graph and features for this panel (a dict)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    X = np.asfortranarray(np.asarray(X, dtype=np.float64).reshape(y.size, -1))
    n, k = X.shape
    D = len(gids)
    ids = [_i32(g) for g in gids]
    fes = (FeInput * D)()
    for d in range(D):
        fes[d].group_ids = ids[d].ctypes.data
        fes[d].num_groups = int(G[d])
        fes[d].num_levels_present = int(G[d])
    sz = LpPanelSizes()
    L = lib()
    ok = L.xfe_lp_probe(_ptr(y), _ptr(X), _i(n), _i(k), fes, _i(D), _i(probe_iter), ctypes.byref(sz))
    if not ok:
        raise XfeError(last_error())
    F, Wm, E, U, C = sz.F, sz.Wm, sz.E, sz.U, sz.C
    out = dict(xf=np.empty((F, 8), np.float32), xw=np.empty((max(Wm, 1), 8), np.float32),
               f_ptr=np.empty(F + 1, np.int32), f_idx=np.empty(max(E, 1), np.int32),
               w_ptr=np.empty(Wm + 1, np.int32), w_idx=np.empty(max(E, 1), np.int32),
               pf=np.empty(max(U, 1), np.int32), pg=np.empty(max(U, 1), np.int32),
               pq=np.empty(max(U, 1), np.float32), W=np.empty(max(U, 1)), dex=np.empty(F),
               err=np.empty((F, C)))
    ok = L.xfe_lp_probe_fetch(*[_ptr(out[key]) for key in ("xf", "xw", "f_ptr", "f_idx", "w_ptr", "w_idx", "pf",
                                                           "pg", "pq", "W", "dex", "err")])
    if not ok:
        raise XfeError(last_error())
    out["xw"], out["f_idx"], out["w_idx"] = out["xw"][:Wm], out["f_idx"][:E], out["w_idx"][:E]
    for key in ("pf", "pg", "pq", "W"):
        out[key] = out[key][:U]
    out.update(F=F, Wm=Wm, E=E, U=U, inner_iterations=sz.inner_iterations, converged=bool(sz.converged))
    return out


def lp_train(graphs, theta, steps, lr=1e-3):
    """xfe_lp_train: Adam on the device over the graphs (dicts with the arrays
    of lp_gnn_eval), one graph per step in turn.  Returns (theta, losses)."""
    keep = []
    arr = (LpGraph * len(graphs))()
    for i, g in enumerate(graphs):
        cv = {}
        for key in ("xf", "xw", "pq", "u_tgt", "u_mask", "y_tgt"):
            cv[key] = np.ascontiguousarray(g[key], dtype=np.float32)
        for key in ("f_ptr", "f_idx", "w_ptr", "w_idx", "pf", "pg"):
            cv[key] = _i32(g[key])
        keep.append(cv)
        arr[i].F = int(cv["xf"].shape[0])
        arr[i].Wm = int(cv["w_ptr"].size - 1)
        arr[i].U = int(cv["pf"].size)
        for key, v in cv.items():
            setattr(arr[i], key, v.ctypes.data)
    th = np.ascontiguousarray(theta, dtype=np.float32).copy()
    hist = np.zeros(max(int(steps), 1))
    ok = lib().xfe_lp_train(_i(len(graphs)), arr, _i(steps), ctypes.c_float(lr), _ptr(th), _ptr(hist))
    if not ok:
        raise XfeError(last_error())
    return th, hist[:steps]


def lp_blocks_apply(pf, pg, W, dex, score, r, cap=16, rounds=6):
    """xfe_lp_blocks_apply: (agg, z) for host pairs, couplings, diagonal and
    scores; r: F x C."""
    pf, pg = _i32(pf), _i32(pg)
    W = np.ascontiguousarray(W, dtype=np.float64)
    dex = np.ascontiguousarray(dex, dtype=np.float64)
    sc = np.ascontiguousarray(score, dtype=np.float32)
    F = dex.size
    r = np.ascontiguousarray(np.asarray(r, dtype=np.float32).reshape(F, -1))
    C = r.shape[1]
    agg = np.empty(F, np.int32)
    z = np.empty((F, C), np.float32)
    ok = lib().xfe_lp_blocks_apply(_i(F), _i(pf.size), _ptr(pf), _ptr(pg), _ptr(W), _ptr(dex), _ptr(sc), _i(cap),
                                   _i(rounds), _i(C), _ptr(r), _ptr(agg), _ptr(z))
    if not ok:
        raise XfeError(last_error())
    return agg, z
