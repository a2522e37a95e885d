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


def test_invalid_arguments_rejected_before_any_device_call():
    y = np.zeros(4)
    X = np.zeros((4, 1))
    with pytest.raises(xfe_amd.XfeError, match="invalid"):
        xfe_amd.absorb(y, X, [np.zeros(4, np.int32)], [0])
