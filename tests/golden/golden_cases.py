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
