"""The allocation-free indexer entry point (xfe_build_indexer_device_noalloc)
against the allocating device entry point (xfe_build_indexer_device), through
the C ABI: two calls in a row, the second with a smaller input that takes the
other path, then the size limit."""
import ctypes
import sys

import numpy as np
import pytest

import xfe_amd

pytestmark = pytest.mark.gpu

MAX_ROWS = 1 << 20


def _noalloc(t):
    import torch
    out = torch.full_like(t, -7)
    G, lp, cont = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    ok = xfe_amd.lib().xfe_build_indexer_device_noalloc(ctypes.c_void_p(t.data_ptr()), ctypes.c_int(t.numel()),
                                                         ctypes.c_void_p(out.data_ptr()), ctypes.byref(G),
                                                         ctypes.byref(lp), ctypes.byref(cont))
    if not ok:
        raise xfe_amd.XfeError(xfe_amd.last_error())
    return out, G.value, lp.value, bool(cont.value)


def test_noalloc_indexer_matches_allocating_entry_point():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(20261204)
    cases = [
        ("dense, gaps kept, maximum size", rng.integers(0, 700_000, MAX_ROWS)),
        ("sparse range, smaller", rng.integers(-2 ** 31, 2 ** 31 - 1, 150_000, dtype=np.int64)),
        ("dense re-ranked, smaller still", rng.integers(0, 20_000, 9_000) * 3),
    ]
    for name, raw in cases:
        print(f"noalloc indexer: {name}, n={raw.size}", flush=True)
        t = torch.from_numpy(raw.astype(np.int32)).to("cuda")
        a, aG, alp, ac = xfe_amd.build_indexer_device(t)
        b, bG, blp, bc = _noalloc(t)
        assert torch.equal(a, b), name
        assert (aG, alp, ac) == (bG, blp, bc), name
    sys.stdout.flush()


def test_noalloc_indexer_rejects_oversize():
    torch = pytest.importorskip("torch")
    t = torch.zeros(MAX_ROWS + 1, dtype=torch.int32, device="cuda")
    with pytest.raises(xfe_amd.XfeError, match="exceed the maximum"):
        _noalloc(t)
