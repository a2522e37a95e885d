"""The C ABI boundary on the CPU: the library loads, exports every symbol the
header declares, and fails closed (returns 0 / -1 with an error) where no
device is present.  No compute calls run here."""
import os
import re

import numpy as np
import pytest

import xfe_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "xfe_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xfe_[A-Za-z_0-9]+)\s*\(", text)))


def test_header_symbols_exported():
    L = xfe_amd.lib()
    names = _declared()
    assert len(names) >= 10
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(xfe_amd.EXPORTS)


def test_oracle_is_not_linked_by_the_product():
    import subprocess
    out = subprocess.run(["ldd", xfe_amd.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out and "xhdfe_ref" not in out


@pytest.mark.skipif(xfe_amd.backend_available(), reason="a device is present")
def test_fail_closed_without_device():
    """Explicit GPU request, no device: the call fails (status 4 at the caller),
    never silently computes on the CPU (src/fe_absorption.cpp:8799-8811)."""
    n = 16
    y = np.arange(n, dtype=float)
    X = np.ones((n, 1))
    g = (np.arange(n) % 4).astype(np.int32)
    with pytest.raises(xfe_amd.XfeError):
        xfe_amd.absorb(y, X, [g, g[::-1].copy()], [4, 4])
    with pytest.raises(xfe_amd.XfeError):
        xfe_amd.drop_singletons([g])
    assert xfe_amd.last_error() != ""
    # the savefe recovery and the group()/individual() absorber fail closed too
    with pytest.raises(xfe_amd.XfeError):
        xfe_amd.recover_fixed_effects(y, [g, g[::-1].copy()], [4, 4])
    gi = dict(num_individuals=2, group_ptr=np.arange(n + 1, dtype=np.int32),
              group_individual=(np.arange(n) % 2).astype(np.int32),
              individual_ptr=np.array([0, n // 2, n], dtype=np.int32),
              individual_group=np.concatenate([np.arange(0, n, 2), np.arange(1, n, 2)]).astype(np.int32),
              group_scale=np.ones(n))
    with pytest.raises(xfe_amd.XfeError) as e:
        xfe_amd.absorb_group_individual(y, X, [g], [4], gi)
    assert "structure" not in str(e.value)  # a valid structure: the device is what is missing


def test_invalid_arguments_rejected_before_any_device_call():
    y = np.zeros(4)
    X = np.zeros((4, 1))
    with pytest.raises(xfe_amd.XfeError, match="invalid"):
        xfe_amd.absorb(y, X, [np.zeros(4, np.int32)], [0])


def test_new_entry_points_validate_before_the_device():
    """Argument checks of the recovery, group()/individual() and shadow
    entry points run on the host: they fail with a message and no device
    call (also without a device)."""
    y = np.arange(8, dtype=np.float64)
    g = (np.arange(8) % 4).astype(np.int32)
    with pytest.raises(xfe_amd.XfeError, match="weighted recovery"):
        xfe_amd.recover_fixed_effects(y, [g, g[::-1].copy()], [4, 4], weights=np.ones(8))
    bad = dict(num_individuals=2, group_ptr=np.arange(9, dtype=np.int32),
               group_individual=np.full(8, 5, dtype=np.int32),  # individual 5 of 2
               individual_ptr=np.array([0, 4, 8], dtype=np.int32),
               individual_group=np.arange(8, dtype=np.int32), group_scale=np.ones(8))
    with pytest.raises(xfe_amd.XfeError, match="group/individual structure"):
        xfe_amd.absorb_group_individual(y, y.reshape(8, 1), [g], [4], bad)
    import ctypes
    L = xfe_amd.lib()
    L.xfe_cert_shadow_begin.restype = ctypes.c_void_p
    assert not L.xfe_cert_shadow_begin(None, None, 8, 0, 8, None, None, None, None, 1, None, None)
    assert "invalid" in xfe_amd.last_error()


def test_refused_methods_fail_closed():
    """Jacobi, LSMR and MLSMR are not device methods (the reference's
    dispatcher never sends them to its device absorber,
    src/fe_absorption.cpp:9127-9141): the call fails with a message naming the
    method, before any device work."""
    n = 16
    y = np.arange(n, dtype=float)
    X = np.ones((n, 1))
    g = (np.arange(n) % 4).astype(np.int32)
    for method, word in ((3, "Jacobi"), (5, "LSMR"), (6, "LSMR"), (9, "unknown")):
        with pytest.raises(xfe_amd.XfeError, match=word):
            xfe_amd.absorb(y, X, [g, g[::-1].copy()], [4, 4], method=method)


def test_ids_outside_int32_rejected():
    """int64 ids that do not fit int32 raise instead of wrapping onto other levels."""
    w = np.array([0, 1, 2 ** 32 + 5, 3], dtype=np.int64)
    with pytest.raises(xfe_amd.XfeError, match="int32"):
        xfe_amd.leave_out_set(w, np.zeros(4, dtype=np.int64))
    with pytest.raises(xfe_amd.XfeError, match="int32"):
        xfe_amd.build_indexer(w)


def test_articulation_rejects_non_simple_graphs():
    """A repeated edge or a self loop is refused on the host (the device method
    assumes a simple graph)."""
    ptr = np.array([0, 2, 3, 4])
    with pytest.raises(xfe_amd.XfeError, match="repeated edge"):
        xfe_amd.articulation_points(ptr, np.array([1, 1, 0, 0]))
    with pytest.raises(xfe_amd.XfeError, match="self loop"):
        xfe_amd.articulation_points(np.array([0, 1, 2]), np.array([0, 0]))


@pytest.mark.skipif(not os.path.isdir("/root/reference/include"), reason="reference headers not present")
def test_adapters_compile_against_reference_headers():
    """The drop-in adapters (xhdfe-xfe_amd/adapter) compile against the
    reference's own headers with -DHDFE_USE_CUDA, so the signatures they define
    are exactly the ones the reference's callers declare."""
    import subprocess
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "xhdfe-xfe_amd", "adapter"), "check"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]


def test_kss_sums_match_reference_with_dense_leverages():
    """The host's O(rows) variance-component sums (xfe_amd.akm.kss_sums), fed
    with the oracle's dense two-way fit and exact leverages, against the
    reference's akm_kss_decompose with LeverageMethod::Exact (fixture): the
    plug-in, AGSU and KSS components, sigma2_ho, n_py and sigma_i per match.
    With the device's fit and leverages the same sums are the GPU test's."""
    import oracle
    from golden_cases import jla_panel
    from xfe_amd import akm
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "exact_leverages.npz"))
    jp = jla_panel()
    y = jp["y"]
    wu, fu, wi, fi, ridx, mw, mf, c = akm.match_panel(jp["worker"], jp["firm"])
    N, J = wu.size, fu.size
    st = oracle.akm_build(mw, mf, c, N, J)
    rhs, _, tw = akm.two_way_rhs(y, ridx, mw, mf, c, N, J, st["Dw"])
    S = np.column_stack([oracle.akm_apply_S(st, np.eye(J)[f]) for f in range(J)])
    psi = np.zeros(J)
    psi[:J - 1] = np.linalg.solve(S[:J - 1, :J - 1], rhs[:J - 1])
    alpha = (tw - np.bincount(mw, weights=c * psi[mf], minlength=N)) / st["Dw"]
    r = akm.kss_sums(y, wi, fi, ridx, mw, mf, c, alpha, psi, *oracle.akm_exact_leverages(st))
    got = list(r["plugin"]) + list(r["agsu"]) + list(r["kss"]) + [r["sigma2_ho"], r["n_py"]]
    comp = fx["components"]
    for i, (g, w) in enumerate(zip(got, comp)):
        assert abs(g - w) <= 1e-10 * max(abs(w), abs(comp[i % 3]) if 3 <= i < 9 else 0.0), (i, g, w)
    key = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(wu[mw], fu[mf]))}
    order = np.array([key[(int(a), int(b))] for a, b in zip(fx["row_worker"], fx["row_firm"])])
    assert order.size == mw.size
    np.testing.assert_array_equal(c[order], fx["row_weight"])
    assert np.abs(r["sigma_i"][order] - fx["sigma_i"]).max() <= 1e-12 * np.abs(fx["sigma_i"]).max()
    assert r["unit_leverage"] == 0


def _dense_two_way(y, panel):
    """The oracle's dense two-way fit on match_panel's matches: (st, alpha, psi)."""
    import oracle
    from xfe_amd import akm
    wu, fu, wi, fi, ridx, mw, mf, c = panel
    N, J = wu.size, fu.size
    st = oracle.akm_build(mw, mf, c, N, J)
    rhs, _, tw = akm.two_way_rhs(y, ridx, mw, mf, c, N, J, st["Dw"])
    S = np.column_stack([oracle.akm_apply_S(st, np.eye(J)[f]) for f in range(J)])
    psi = np.zeros(J)
    psi[:J - 1] = np.linalg.solve(S[:J - 1, :J - 1], rhs[:J - 1])
    alpha = (tw - np.bincount(mw, weights=c * psi[mf], minlength=N)) / st["Dw"]
    return st, alpha, psi


def test_kss_sums_match_reference_with_jla_draws(golden):
    """kss_sums with the JLA leverages (50 draws, seed 42) and their
    non-linearity correction, fed with the oracle's dense fit, JLA moments
    and B_ii draws on match_panel's rows (the reference's working-row order,
    which the draws' streams are indexed by), against the reference's JLA
    decomposition on the same panel (fixture): the 11 numbers and sigma_i."""
    import oracle
    from golden_cases import JLA_DRAWS, JLA_SEED, jla_panel
    from xfe_amd import akm
    jp = jla_panel()
    panel = akm.match_panel(jp["worker"], jp["firm"])
    wu, fu, wi, fi, ridx, mw, mf, c = panel
    np.testing.assert_array_equal(wu[mw], golden["jla/row_worker"])
    np.testing.assert_array_equal(fu[mf], golden["jla/row_firm"])
    st, alpha, psi = _dense_two_way(jp["y"], panel)
    P, corr = oracle.akm_jla_moments(st, JLA_DRAWS, JLA_SEED)
    assert np.abs(P - golden["jla/pii"]).max() <= 1e-12
    bfe, bpe, bcov = oracle.akm_jla_bii(st, JLA_DRAWS, JLA_SEED, c.astype(np.int64))
    r = akm.kss_sums(jp["y"], wi, fi, ridx, mw, mf, c, alpha, psi, P, bfe, bpe, bcov, corr=corr)
    got = list(r["plugin"]) + list(r["agsu"]) + list(r["kss"]) + [r["sigma2_ho"], r["n_py"]]
    comp = golden["jla/components"]
    for i, (g, w) in enumerate(zip(got, comp)):
        assert abs(g - w) <= 1e-10 * max(abs(w), abs(comp[i % 3]) if 3 <= i < 9 else 0.0), (i, g, w)
    sig = golden["jla/sigma_i"]
    assert np.abs(r["sigma_i"] - sig).max() <= 1e-12 * np.abs(sig).max()
    # the correction is what moves sigma_i off the exact-leverage form
    assert np.abs(corr - 1.0).max() > 1e-3
