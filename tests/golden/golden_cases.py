"""Inputs of the golden cases (regenerated from seeds; the outputs live in
ref_cases.npz, written by make_golden.py from the reference's CPU core)."""
from __future__ import annotations

import hashlib

import numpy as np

from xfe_amd import panels


def _two_way():
    return panels.config0(seed=panels.SEED0 + 7, workers=600, years=10, firms=120)


def _akm_three_way():
    return panels.akm_panel(seed=panels.SEED0 + 107, workers=500, firms=80, years=36)


def _sparse_two_way():
    # low mobility: many single-firm workers, several mobility groups
    p = panels.config0(seed=panels.SEED0 + 9, workers=400, years=4, firms=150)
    return p


CASES = {
    "two_way": _two_way,
    "akm_three_way": _akm_three_way,
    "sparse_two_way": _sparse_two_way,
}


def input_checksum(p):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(p["y"]).tobytes())
    h.update(np.asfortranarray(p["X"]).tobytes(order="F"))
    for f in p["fes"]:
        h.update(np.ascontiguousarray(f, dtype=np.int32).tobytes())
    return np.frombuffer(h.digest()[:8], dtype=np.int64)[0]


def certificate_weights(n):
    """Deterministic row weights for the weighted certificate fixtures."""
    return 0.5 + (np.arange(n) % 7) / 4.0


# Edge cases of schwarz_demean_raw (src/schwarz_demean.cpp:290-438), as
# ((y, X, ids, levels), max_iter) on one seeded 2000-row panel:
# 0 the plain two-way case, 1 an empty level in the first dimension, 2 an
# empty level in a third dimension, 3 one dimension only, 4 all-zero columns,
# 5 one zero column among non-zero ones, 6 max_iter 0, 7 k = 15, 8 k = 16.
SCHWARZ_EDGE_NAMES = ("two_way", "empty_level_dim0", "empty_level_extra", "one_dim", "zero_columns",
                      "one_zero_column", "max_iter_0", "k15", "k16")


def schwarz_edge_cases():
    rng = np.random.default_rng(20261203)
    n = 2000
    a = rng.integers(0, 100, n).astype(np.int32)
    b = rng.integers(0, 30, n).astype(np.int32)
    c = rng.integers(0, 7, n).astype(np.int32)
    y = rng.normal(size=n)
    X = np.asfortranarray(rng.normal(size=(n, 2)))
    X1 = np.asfortranarray(np.column_stack([X[:, 0], np.zeros(n)]))
    X15 = np.asfortranarray(rng.normal(size=(n, 15)))
    X16 = np.asfortranarray(rng.normal(size=(n, 16)))
    return [
        ((y, X, [a, b], [100, 30]), 10000),
        ((y, X, [a, b], [101, 30]), 10000),
        ((y, X, [a, b, c], [100, 30, 8]), 10000),
        ((y, X, [a], [100]), 10000),
        ((np.zeros(n), np.zeros((n, 2), order="F"), [a, b], [100, 30]), 10000),
        ((y, X1, [a, b], [100, 30]), 10000),
        ((y, X, [a, b], [100, 30]), 0),
        ((y, X15, [a, b], [100, 30]), 10000),
        ((y, X16, [a, b], [100, 30]), 10000),
    ]


# Heterogeneous-slope cases: (case name, slope dimension, with intercept) on
# the kept rows of the golden panels, slope variable slope_values(n).
SLOPE_CASES = (("two_way", 1, True), ("two_way", 1, False), ("akm_three_way", 1, True))


def slope_values(n):
    """The slope variable of the slope cases: N(1, 1), seeded."""
    return np.random.default_rng(20261213).normal(1.0, 1.0, n)


def jla_panel():
    """A small AKM panel whose worker-firm graph is leave-out connected: 600
    workers x 5 years over 40 firms, a move with probability 0.35 a year."""
    rng = np.random.default_rng(4)
    W, J, T = 600, 40, 5
    worker = np.repeat(np.arange(W), T)
    firm = np.empty(W * T, dtype=np.int64)
    for w in range(W):
        f = rng.integers(0, J)
        for t in range(T):
            if rng.random() < 0.35:
                f = rng.integers(0, J)
            firm[w * T + t] = f
    return dict(y=rng.normal(size=W * T), worker=worker, firm=firm)


JLA_DRAWS, JLA_SEED = 50, 42
