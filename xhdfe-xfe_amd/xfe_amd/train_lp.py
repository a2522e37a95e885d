"""Training of the learned preconditioner's network on the GPU.

    python -m xfe_amd.train_lp [--panels 48] [--steps 3000] [--out lp_params.f32]

Training data come from the absorber itself (xfe_lp_probe): on each synthetic
panel the mixed-precision absorber runs Jacobi-PCG on the panel's columns; the
error of its iterate after a few inner iterations against the final solution,
on the firm levels, is the slow error the preconditioner has to remove.  From
it (per firm, over the columns):

* the per-level target u*_f = log(1e-3 + E_f), E_f the firm's share of the
  error energy (sum over columns of S_ff e_f^2 / sum_g S_gg e_g^2, scaled so
  the mean over firms with movers is 1);
* the pair targets y*_fg = 1 when firms f and g fall in one aggregate of the
  teacher aggregation, 0 otherwise.  The teacher is the sequential greedy
  aggregation (union of the pairs in order of decreasing score, at most 16
  firms per aggregate) over the coupling strength weighted by the error
  amplitude at both ends, q_fg (1 + sqrt(E_f) + sqrt(E_g)): it merges first
  the strongly coupled firms that carry slow error.

The network (lp_gnn.hip) learns to predict both from the graph and its local
features alone, so that at solve time the aggregation needs no probe solve.
The training loop (forward, backward, Adam) runs on the device
(xfe_lp_train); the parameters are written as little-endian float32.

Panels (the learned-preconditioner configuration, scaled to the GPU time of
this build): AKM-shaped, 30k - 1M workers, log-uniform (about 0.3M - 9M
rows), plus --big panels of 2M workers (about 18M rows), firms =
workers / 7.4, mover probability U(0.005, 0.10), Pareto firm-size exponent
U(1.05, 2.0), plus chains of small firms joined by single movers and attached
to the main set by one mover (weakly connected chain components).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

import xfe_amd
from xfe_amd import panels

CAP = 16


def chain_rows(rng, first_worker, first_firm, n_chains, length, years=36):
    """Rows of chains of `length` new firms: per firm 2 - 5 stayers (whole
    careers), per link one mover with a segment at each end, and one mover
    from the chain's first firm to an existing firm (ids < first_firm)."""
    W, F, Y = [], [], []
    w = first_worker
    f0 = first_firm
    for _ in range(n_chains):
        for j in range(length):
            for _ in range(int(rng.integers(2, 6))):
                T = int(rng.integers(2, 12))
                a = int(rng.integers(0, years - T + 1))
                W += [w] * T; F += [f0 + j] * T; Y += list(range(a, a + T))
                w += 1
        links = [(f0 + j, f0 + j + 1) for j in range(length - 1)]
        links.append((int(rng.integers(0, first_firm)), f0))
        for fa, fb in links:
            La, Lb = int(rng.integers(1, 8)), int(rng.integers(1, 8))
            a = int(rng.integers(0, years - La - Lb + 1))
            W += [w] * (La + Lb); F += [fa] * La + [fb] * Lb; Y += list(range(a, a + La + Lb))
            w += 1
        f0 += length
    return np.array(W, np.int32), np.array(F, np.int32), np.array(Y, np.int32), f0


def training_panel(seed, workers=None, p_move=None, alpha=None, chains=True, extra_cols=3):
    """One synthetic panel: (V, gids, G, description)."""
    rng = np.random.default_rng(seed)
    workers = int(workers or np.exp(rng.uniform(np.log(30_000), np.log(1_000_000))))
    firms = max(50, int(workers / 7.4))
    p_move = float(p_move if p_move is not None else rng.uniform(0.005, 0.10))
    alpha = float(alpha if alpha is not None else rng.uniform(1.05, 2.0))
    worker, firm, year, within = panels._akm_ids(rng, workers, firms, 36, 9.4, alpha, p_move)
    if chains:
        cw, cf, cy, _ = chain_rows(rng, workers, firms, int(rng.integers(5, 40)), int(rng.integers(3, 12)))
        worker = np.concatenate([worker, cw]); firm = np.concatenate([firm, cf]); year = np.concatenate([year, cy])
        within = np.concatenate([within, np.zeros(cw.size, np.int64)])
    n = worker.size
    cols = [rng.standard_normal(n) + 0.3 * np.sin(worker * 0.01) for _ in range(2 + extra_cols)]
    cols.append((within + 1).astype(np.float64))
    V = np.column_stack(cols)
    fes = [worker, firm, year]
    keep, _, _ = xfe_amd.drop_singletons(fes)
    fes = [f[keep] for f in fes]
    idx = [xfe_amd.build_indexer(f) for f in fes]
    gids, G = [i[0] for i in idx], [i[1] for i in idx]
    desc = f"workers={workers} firms={firms} p_move={p_move:.4f} alpha={alpha:.3f} rows={int(keep.sum())}"
    return V[keep], gids, G, desc


def teacher(pf, pg, score, F, cap=CAP):
    """Sequential greedy aggregation: pairs in order of decreasing score (ties:
    lower index), union when the two aggregates together hold at most cap
    firms.  Returns the aggregate (root) of every firm."""
    parent = np.arange(F)
    size = np.ones(F, dtype=np.int64)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for p in np.lexsort((np.arange(score.size), -score)):
        a, b = find(pf[p]), find(pg[p])
        if a == b or size[a] + size[b] > cap:
            continue
        if b < a:
            a, b = b, a
        parent[b] = a
        size[a] += size[b]
    return np.array([find(a) for a in range(F)])


def targets(pr):
    """u*, mask and y* of one probed panel (see the module docstring)."""
    err, dex = pr["err"], pr["dex"]
    has = dex > 0
    en = dex[:, None] * err ** 2
    tot = en.sum(axis=0)
    share = np.where(tot > 0, en / np.maximum(tot, 1e-300), 0.0).mean(axis=1)
    E = share * max(int(has.sum()), 1)
    u_tgt = np.clip(np.log(1e-3 + E), -7.0, 5.0)
    amp = np.sqrt(E)
    ts = pr["pq"] * (1.0 + amp[pr["pf"]] + amp[pr["pg"]])
    root = teacher(pr["pf"], pr["pg"], ts, pr["F"])
    y_tgt = (root[pr["pf"]] == root[pr["pg"]]).astype(np.float32)
    return u_tgt.astype(np.float32), has.astype(np.float32), y_tgt


def graph_of(pr):
    u_tgt, u_mask, y_tgt = targets(pr)
    return dict(xf=pr["xf"], xw=pr["xw"], f_ptr=pr["f_ptr"], f_idx=pr["f_idx"], w_ptr=pr["w_ptr"],
                w_idx=pr["w_idx"], pf=pr["pf"], pg=pr["pg"], pq=pr["pq"], u_tgt=u_tgt, u_mask=u_mask, y_tgt=y_tgt)


def _layout():
    H = 64
    return ([("Wf0", (8, H)), ("bf0", (H,)), ("Ww0", (8, H)), ("bw0", (H,))]
            + [item for l in range(1, 4)
               for item in ((f"Ww{l}", (2 * H, H)), (f"bw{l}", (H,)), (f"Wf{l}", (2 * H, H)), (f"bf{l}", (H,)))]
            + [("wa", (H,)), ("ba", (1,)), ("we", (H,)), ("be", (1,)), ("ws", (1,))])


def initial_params(seed=0):
    """He-normal weights, zero biases, the pair feature's weight at 4 (the
    scores start as a multiple of the coupling strength); the layout of
    lp_gnn.h."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in _layout():
        if len(shape) == 2:
            parts.append(rng.standard_normal(shape).ravel() * np.sqrt(2.0 / shape[0]))
        elif name in ("wa", "we"):
            parts.append(rng.standard_normal(shape) * np.sqrt(1.0 / 64) * 0.1)
        elif name == "ws":
            parts.append(np.full(shape, 4.0))
        else:
            parts.append(np.zeros(shape))
    th = np.concatenate(parts).astype(np.float32)
    if th.size != xfe_amd.lp_param_count():
        raise RuntimeError(f"parameter layout: {th.size} != {xfe_amd.lp_param_count()}")
    return th


def probe_panel(seed, **kw):
    V, gids, G, desc = training_panel(seed, **kw)
    pr = xfe_amd.lp_probe(V[:, 0], V[:, 1:], gids, G, probe_iter=8)
    return pr, desc


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--panels", type=int, default=48)
    ap.add_argument("--big", type=int, default=4)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--out", default=xfe_amd.LP_PARAMS_FILE)
    a = ap.parse_args(argv)
    t0 = time.time()
    graphs = []
    for k in range(a.panels + a.big):
        pr, desc = probe_panel(a.seed + k, workers=2_000_000 if k >= a.panels else None)
        graphs.append(graph_of(pr))
        print(f"[lp_train] panel {k}: {desc} F={pr['F']} movers={pr['Wm']} pairs={pr['U']} "
              f"inner={pr['inner_iterations']} blocked={int(graphs[-1]['y_tgt'].sum())} "
              f"({time.time() - t0:.0f}s)", flush=True)
    th0 = initial_params(a.seed)
    th, hist = xfe_amd.lp_train(graphs, th0, a.steps, a.lr)
    w = max(1, a.panels)
    print(f"[lp_train] loss first {np.mean(hist[:w]):.4f} last {np.mean(hist[-w:]):.4f} "
          f"({a.steps} steps, {time.time() - t0:.0f}s)", flush=True)
    th.astype("<f4").tofile(a.out)
    print(f"[lp_train] wrote {a.out}", flush=True)
    return th, hist


if __name__ == "__main__":
    sys.exit(0 if main() is not None else 1)
