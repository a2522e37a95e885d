"""Synthetic worker-firm panels of the benchmark shapes (numpy, fixed seeds).

Seeds are ``20261203 + 100 * config``.  Row order is worker-major, years
ascending inside a worker (the natural layout of an AKM panel).
"""
from __future__ import annotations

import numpy as np

SEED0 = 20261203


def _size_weighted(rng, cdf, size):
    """Draw firm ids with probability proportional to their size weights."""
    u = rng.random(size) * cdf[-1]
    return np.searchsorted(cdf, u, side="right").astype(np.int32)


def _firm_path(rng, cdf, first_row_mask, p_move):
    """Firm of every row: a size-weighted draw at each worker's first row and at
    each move (probability p_move per later row), carried forward otherwise."""
    n = first_row_mask.size
    switch = first_row_mask | (rng.random(n) < p_move)
    draws = _size_weighted(rng, cdf, n)
    idx = np.where(switch, np.arange(n, dtype=np.int64), 0)
    np.maximum.accumulate(idx, out=idx)
    return draws[idx]


def config0(seed=SEED0 + 0, workers=20000, years=10, firms=2000):
    """Two-way FE OLS panel: N = workers x years, log-normal(0,1) firm sizes,
    switch probability 0.1 per worker-year, worker FE ~ N(0, 0.3^2), firm FE ~
    N(0, 0.2^2), three regressors with correlation 0.3 to the worker FE,
    y = x [0.5, -0.2, 0.1] + alpha + psi + eps, eps ~ N(0, 0.25^2)."""
    rng = np.random.default_rng(seed)
    n = workers * years
    size = rng.lognormal(0.0, 1.0, firms)
    cdf = np.cumsum(size)
    worker = np.repeat(np.arange(workers, dtype=np.int32), years)
    year = np.tile(np.arange(years, dtype=np.int32), workers)
    first = np.zeros(n, dtype=bool)
    first[::years] = True
    firm = _firm_path(rng, cdf, first, 0.1)
    alpha = rng.normal(0.0, 0.3, workers)
    psi = rng.normal(0.0, 0.2, firms)
    a = alpha[worker]
    X = np.empty((n, 3), order="F")
    for j in range(3):
        X[:, j] = 0.3 * (a / 0.3) + np.sqrt(1.0 - 0.09) * rng.standard_normal(n)
    beta = np.array([0.5, -0.2, 0.1])
    y = X @ beta + a + psi[firm] + rng.normal(0.0, 0.25, n)
    return dict(y=y, X=X, fes=[worker, firm], cluster=worker, year=year, beta=beta,
                name="c0_two_way_200k")


def _akm_ids(rng, workers, firms, years, tenure_mean, pareto_alpha, p_move):
    """Worker, firm and year ids (and the within-spell index) of the AKM shape."""
    T = np.minimum(rng.geometric(1.0 / tenure_mean, workers), years).astype(np.int64)
    start = (rng.random(workers) * (years - T + 1)).astype(np.int64)
    n = int(T.sum())
    offs = np.zeros(workers + 1, dtype=np.int64)
    np.cumsum(T, out=offs[1:])
    worker = np.repeat(np.arange(workers, dtype=np.int32), T)
    within = np.arange(n, dtype=np.int64) - np.repeat(offs[:-1], T)
    year = (np.repeat(start, T) + within).astype(np.int32)
    first = np.zeros(n, dtype=bool)
    first[offs[:-1]] = True
    size = (1.0 - rng.random(firms)) ** (-1.0 / pareto_alpha)
    cdf = np.cumsum(size)
    firm = _firm_path(rng, cdf, first, p_move)
    return worker, firm, year, within


def akm_panel(seed=SEED0 + 100, workers=5_950_000, firms=800_000, years=36, tenure_mean=9.4,
              pareto_alpha=1.1, p_move=0.05, beta=(0.5, -0.2, 0.1, 0.05)):
    """AKM-shape panel: geometric tenure (mean 9.4, capped at `years`), start
    years spread so every spell fits the window, Pareto(alpha) firm sizes,
    size-weighted moves with probability p_move per year.  Regressors:
    tenure, tenure^2/100 and two N(0,1).  FE = (worker, firm, year)."""
    rng = np.random.default_rng(seed)
    worker, firm, year, within = _akm_ids(rng, workers, firms, years, tenure_mean, pareto_alpha, p_move)
    n = worker.size
    tenure = (within + 1).astype(np.float64)
    X = np.empty((n, 4), order="F")
    X[:, 0] = tenure
    X[:, 1] = tenure * tenure / 100.0
    X[:, 2] = rng.standard_normal(n)
    X[:, 3] = rng.standard_normal(n)
    alpha = rng.normal(0.0, 0.3, workers)
    psi = rng.normal(0.0, 0.2, firms)
    delta = rng.normal(0.0, 0.1, years)
    b = np.asarray(beta, dtype=np.float64)
    y = X @ b + alpha[worker] + psi[firm] + delta[year] + rng.normal(0.0, 0.25, n)
    return dict(y=y, X=X, fes=[worker, firm, year], cluster=worker, beta=b,
                name=f"akm_{n}obs_{workers}w_{firms}f_{years}y")


def config1(seed=SEED0 + 100):
    """The 56M-obs AKM shape: 5.95M workers, 800k firms (Pareto 1.1), 36 years,
    mover probability 0.05, k = 4."""
    return akm_panel(seed=seed)


def akm_small(seed=SEED0 + 100, scale=0.01):
    """A config-1-shaped panel scaled down by `scale` (workers and firms)."""
    return akm_panel(seed=seed, workers=max(10, int(5_950_000 * scale)),
                     firms=max(5, int(800_000 * scale)))


def kss_panel(seed=SEED0 + 300, workers=1_200_000, firms=150_000, years=36, tenure_mean=8.4,
              pareto_alpha=1.3, p_move=0.08):
    """The KSS leverage shape: 1.2M workers, 150k firms with Pareto(1.3)
    sizes, mover probability 0.08 per year, geometric tenure of mean 8.4
    (about 10M person-years).  Returns worker and firm ids per row."""
    rng = np.random.default_rng(seed)
    worker, firm, _, _ = _akm_ids(rng, workers, firms, years, tenure_mean, pareto_alpha, p_move)
    return dict(worker=worker, firm=firm, name=f"kss_{worker.size}obs_{workers}w_{firms}f")


def slope_panel(seed=SEED0 + 200, scale=0.1, regressors=31, p_move=0.08):
    """The heterogeneous-slope shape at `scale` of 30M obs: 3M workers and
    400k firms (Pareto 1.1) times scale, geometric tenure of mean 10 capped at
    36 years, mover probability 0.08; FE = (worker, firm with a slope on x1,
    firm x year); y and `regressors` N(0, 1) columns besides x1."""
    rng = np.random.default_rng(seed)
    workers = max(10, int(3_000_000 * scale))
    firms = max(5, int(400_000 * scale))
    worker, firm, year, _ = _akm_ids(rng, workers, firms, 36, 10.0, 1.1, p_move)
    n = worker.size
    x1 = rng.standard_normal(n)
    cell = firm.astype(np.int64) * 36 + year
    X = np.empty((n, regressors), order="F")
    for j in range(regressors):
        X[:, j] = rng.standard_normal(n)
    psi = rng.normal(0.0, 0.2, firms)
    slope = rng.normal(0.5, 0.1, firms)
    y = X[:, 0] * 0.3 + psi[firm] + slope[firm] * x1 + rng.normal(0.0, 0.25, n)
    return dict(y=y, X=X, z=x1, fes=[worker, firm, cell], name=f"slope_{n}obs_{workers}w_{firms}f")
