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


def indexer_cases():
    """Raw id vectors that take every branch of build_indexer
    (src/fe_absorption.cpp:845-940), each with seeded row weights:
    dense with every slot used, dense with gaps above the 90 % rule, dense
    below it (re-ranked), negative ids, a range wider than 2n (sorted rank),
    ids near the int32 limits, one row, one repeated id, already sorted ids."""
    rng = np.random.default_rng(20261204)
    out = {}
    n = 5000
    out["dense_full"] = rng.permutation(np.arange(n) % 1200)
    v = rng.integers(0, 1000, n)
    out["dense_gaps_kept"] = v[(v % 20) != 3]                    # 95 % of the slots used
    out["dense_reranked"] = rng.integers(0, 2 * n, n) * 1         # about 39 % used -> ranked
    out["negative"] = rng.integers(-3000, 500, n)
    out["sparse_range"] = rng.integers(-10 ** 6, 10 ** 6, n)
    out["int32_limits"] = np.concatenate([rng.integers(2 ** 31 - 40, 2 ** 31, 300),
                                          rng.integers(-2 ** 31, -2 ** 31 + 40, 300)])
    out["one_row"] = np.array([17])
    out["one_id"] = np.full(700, -5)
    out["sorted"] = np.sort(rng.integers(100, 4000, n))
    return {k: (np.asarray(v, dtype=np.int32), 0.25 + rng.uniform(0.0, 2.0, np.asarray(v).size))
            for k, v in out.items()}


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


def gi_panel(seed=20261203 + 500, n=3000, individuals=800, years=20, classes=40):
    """A group()/individual() panel (patents and their inventors): each row
    (group) has 1-3 distinct individuals drawn with log-normal activity
    weights; group_scale = 1 / team size (mean aggregation); standard FEs
    year and class; y = x b + mean of the team's effects + year + class + e."""
    rng = np.random.default_rng(seed)
    act = rng.lognormal(0.0, 1.0, individuals)
    cdf = np.cumsum(act) / act.sum()
    size = rng.integers(1, 4, n)
    gp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(size, out=gp[1:])
    members = []
    for g in range(n):
        team = set()
        while len(team) < size[g]:
            team.add(int(np.searchsorted(cdf, rng.random(), side="right")))
        members.extend(sorted(team))
    gind = np.asarray(members, dtype=np.int32)
    rows = np.repeat(np.arange(n, dtype=np.int32), size)
    order = np.lexsort((rows, gind))
    iptr = np.zeros(individuals + 1, dtype=np.int64)
    np.cumsum(np.bincount(gind, minlength=individuals), out=iptr[1:])
    scale = 1.0 / size.astype(np.float64)
    year = rng.integers(0, years, n).astype(np.int32)
    cls = rng.integers(0, classes, n).astype(np.int32)
    alpha = rng.normal(0.0, 0.5, individuals)
    team_effect = np.bincount(rows, weights=alpha[gind] * scale[rows], minlength=n)
    X = rng.standard_normal((n, 2))
    y = X @ np.array([0.4, -0.3]) + team_effect + rng.normal(0, 0.2, years)[year] + \
        rng.normal(0, 0.3, classes)[cls] + rng.normal(0.0, 0.25, n)
    gi = dict(num_individuals=individuals, group_ptr=gp.astype(np.int32), group_individual=gind,
              individual_ptr=iptr.astype(np.int32), individual_group=rows[order].astype(np.int32),
              group_scale=scale)
    return dict(y=y, X=np.asfortranarray(X), fes=[year, cls], G=[years, classes], gi=gi)


def gi_weights(n):
    """Deterministic row weights for the weighted group()/individual() case."""
    return 0.5 + (np.arange(n) % 5) / 3.0


def leave_out_panels():
    """Panels for the leave-out connected set: the KSS shape at small scale
    with low mobility (many cut-vertex movers, several pruning rounds), the
    same with moderate mobility, and permuted firm ids with gaps."""
    out = {}
    for name, seed, workers, firms, p_move in (("low_mobility", 311, 3000, 400, 0.02),
                                               ("moderate", 312, 4000, 300, 0.08)):
        p = panels.kss_panel(seed=panels.SEED0 + seed, workers=workers, firms=firms, p_move=p_move)
        out[name] = (p["worker"].astype(np.int32), p["firm"].astype(np.int32))
    w, f = out["moderate"]
    perm = np.random.default_rng(5).permutation(300) * 7 + 11
    out["gapped_ids"] = (w * 3 + 2, perm[f].astype(np.int32))
    return out
