"""The learned preconditioner on an MI355X, through the C ABI.

Network parity: the device's forward pass, loss and gradient against the
oracle's restatement with the same bfloat16 rounding of every matrix-product
operand (oracle/learned.py, whose gradient is pinned by finite differences);
the device accumulates in FP32, the oracle exactly, so the tolerances are
relative: outputs 2e-3 of their scale, loss 1e-4, gradient 5e-3 of each
parameter block's largest entry.
"""
import sys

import numpy as np
import pytest

import xfe_amd
from oracle import learned as L

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("size", [(30, 40), (3000, 5000)])
def test_gnn_matches_oracle(size):
    F, Wm = size
    _log(f"gnn eval F={F} Wm={Wm}")
    g = L.random_graph(11 + F, F=F, Wm=Wm)
    th = L.init_params(5)
    loss, grad, u, t = xfe_amd.lp_gnn_eval(g, th)
    lo, go, out = L.forward_backward(th.astype(np.float64), g, bf16_products=True)
    _log(f"  loss device {loss:.6g} oracle {lo:.6g}")
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert np.abs(u - out["u"]).max() <= 2e-3 * max(1.0, np.abs(out["u"]).max())
    assert np.abs(t - out["t"]).max() <= 2e-3 * max(1.0, np.abs(out["t"]).max())
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in L.LAYOUT])
    for k, (name, _) in enumerate(L.LAYOUT):
        a, b = grad[offs[k]:offs[k + 1]], go[offs[k]:offs[k + 1]]
        scale = max(np.abs(b).max(), 1e-12)
        err = np.abs(a - b).max() / scale
        assert err <= 5e-3, (name, err)
