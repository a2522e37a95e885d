"""Benchmark: seconds to absorb the 56M-obs AKM panel to 1e-8 on one MI355X.

Workload: the AKM-shape panel (5.95M workers with geometric tenure of mean 9.4
capped at 36 years, 800k firms with Pareto(1.1) sizes, 36 year FE, mover
probability 0.05, k = 4 regressors: tenure, tenure^2/100, two N(0,1)),
synthetic and seeded.  Singletons are dropped on the device once (reported,
not timed); one step is the full within transform of [y | X] (C = 5 columns,
3-way FE, tolerance 1e-8) starting from the raw columns already resident in
HBM, through the level sort, the solve and the scatter back to the caller's
layout.  The timed method is the mixed-precision absorber (precision 1: FP32
Schur-complement PCG with FP64 residual replacement, final transform in FP64)
with the learned preconditioner (the trained graph network's aggregation of
firms into blocks solved exactly); the same method with Jacobi and with the
coupling-strength aggregation is timed beside it, and the
plain FP64 absorber (precision 0: Irons-Tuck accelerated Gauss-Seidel with the
reference's gates and stopping rule) is timed in the same run, and the two
results are compared.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "xhdfe-xfe_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
METRIC = "s to absorb 56M-obs AKM at 1e-8 tol: learned+FP32/FP64 vs plain FP64; HBM GB/s"
# dominant kernel of each method and the committed rocprofv3 summary holding its HBM traffic
KERNELS = {
    1: ("seg_firm_kernel (firm-major segment pass of the FP32 Schur operator)",
        os.path.join(ROOT, "profiles", "seg_firm_akm56m.md")),
    0: ("tile_pass_kernel (fused apply / in-tile projection / reduce pass)",
        os.path.join(ROOT, "profiles", "tile_pass_akm56m.md")),
}


def measured_traffic(path):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy of known size),
    or None when it is absent."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def compact(ids):
    """build_indexer's dense rule: id - min, re-ranked without gaps when fewer
    than 90 % of the range is occupied (src/fe_absorption.cpp:873-904)."""
    ids = np.asarray(ids, dtype=np.int64)
    mn = int(ids.min())
    rng = int(ids.max()) - mn + 1
    x = ids - mn
    seen = np.zeros(rng, dtype=bool)
    seen[x] = True
    uniq = int(seen.sum())
    if uniq * 10 < rng * 9:
        remap = np.cumsum(seen) - 1
        return remap[x].astype(np.int32), uniq
    return x.astype(np.int32), rng


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(V_dev, ids_dev, n, C, G, V_gpu, threads, iters_gpu, scale=0.05):
    """The oracle (C restatement of the reference CPU absorber: packed
    Irons-Tuck path, contiguous / CSR / observation-partitioned projections,
    OpenMP; 0.73-0.91x the compiled reference's time on the same panel) on the
    full panel on `threads` threads, from the same rows as the GPU, its result
    compared with the HIP plain FP64 result (outside the GPU's timed region);
    plus one run on 1 thread on the AKM panel at `scale`, scaled to the full
    panel per obs-column-iteration."""
    import oracle
    from xfe_amd import panels
    Vh = V_dev[:, :n].cpu().numpy().T.copy()
    gids = [ids_dev[d].cpu().numpy() for d in range(len(G))]
    t0 = time.perf_counter()
    Vo, info = oracle.absorb(Vh, gids, G, tol=1e-8, threads=threads)
    dt = time.perf_counter() - t0
    err = None
    if V_gpu is not None:
        err = float(np.abs(Vo - V_gpu[:, :n].cpu().numpy().T).max() / np.abs(Vh).max())
    del Vh, Vo, gids
    p = panels.akm_small(scale=scale)
    keep, _, _ = oracle.drop_singletons(p["fes"])
    fes = [f[keep] for f in p["fes"]]
    idx = [oracle.build_indexer(f) for f in fes]
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    del p
    t0 = time.perf_counter()
    _, info1 = oracle.absorb(V, [i[0] for i in idx], [i[1] for i in idx], tol=1e-8, threads=1)
    dt1 = time.perf_counter() - t0
    ns = V.shape[0]
    return dict(value=round(dt, 3), unit="s", cores=threads, kind="port", cpu_model=cpu_model(),
                sample=(f"the full panel ({n} obs, {C} cols, 3-way, tol 1e-8) on {threads} threads: "
                        f"{info['iterations']} iterations (the GPU plain FP64 absorber: {iters_gpu}); 1 thread "
                        f"on the AKM panel at {scale:.0%} ({ns} obs, {info1['iterations']} iterations): {dt1:.2f} s"),
                iterations=info["iterations"], max_abs_diff_vs_hip_plain_over_max_abs_v=err,
                value_1thread_scaled=round(dt1 / (ns * info1["iterations"]) * n * info["iterations"], 2))


AKM_METRIC = "s per 256-RHS grounded firm-Laplacian PCG solve (KSS JLA draws), 10M-obs KSS panel"


def akm_cpu_baseline(st, gpu_iters, lanes):
    """The oracle's single-lane Jacobi-PCG (numpy) on the same structure for a
    bounded number of iterations; value = seconds for all lanes at the GPU's
    iteration counts and the sample's seconds per lane-iteration."""
    import oracle
    rng = np.random.default_rng(1)
    rhs = rng.choice([-1.0, 1.0], size=st["J"])
    t0 = time.perf_counter()
    _, its = oracle.akm_solve_S(st, rhs, 1e-10, 60)
    dt = time.perf_counter() - t0
    per = dt / 60.0
    return dict(value=round(per * float(np.sum(gpu_iters)), 2), unit="s", cores=1, kind="port",
                sample=(f"oracle single-lane Jacobi-PCG (numpy bincount), 60 iterations on the full structure "
                        f"in {dt:.2f} s; value = {per * 1e3:.1f} ms per lane-iteration x the GPU's "
                        f"{int(np.sum(gpu_iters))} lane-iterations over {lanes} lanes"))


def bench_akm(args):
    """The leverage solver (akm_cuda_solve_S_multi drop-in): 256 Rademacher
    right-hand sides, as in the JLA draws, solved at cg_tol 1e-10 in batches of
    32 lanes on the largest mobility group of the KSS-shape panel."""
    import torch  # noqa: F401  (shares torch's HIP runtime)

    import xfe_amd
    from xfe_amd import akm, panels

    if not xfe_amd.backend_available():
        raise SystemExit("no HIP device")
    t0 = time.perf_counter()
    p = panels.kss_panel()
    n = p["worker"].size
    mw, mf, mc, N, J = akm.panel_matches(p["worker"], p["firm"])
    st = akm.build_structure(mw, mf, mc, N, J)
    M = mw.size
    log(f"KSS panel {n} obs -> {N} workers, {J} firms, {M} matches in {time.perf_counter() - t0:.1f} s")
    solver = xfe_amd.AkmSolver(st)
    lanes, batch = 256, 32
    rng = np.random.default_rng(20261503)
    rhs = rng.choice([-1.0, 1.0], size=(lanes, J))
    max_iter = 20000
    # the wide solver: all 256 lanes in one FP32 pass per iteration, FP64
    # residual replacement, right-hand sides and solutions resident in HBM
    # (timed; its pair pass is the roofline kernel)
    import torch
    rhs_d = torch.from_numpy(rhs).to("cuda:0")
    z_d = torch.empty_like(rhs_d)
    torch.cuda.synchronize()
    for w in range(args.warmup):
        t0 = time.perf_counter()
        its, _ = solver.solve_wide_device(rhs_d.data_ptr(), z_d.data_ptr(), lanes, tol=1e-10, max_iter=max_iter)
        log(f"wide warmup {w}: {time.perf_counter() - t0:.3f} s, iterations {its.min()}..{its.max()}")
    xfe_amd.profile_reset()
    xfe_amd.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        its, (passes, updates) = solver.solve_wide_device(rhs_d.data_ptr(), z_d.data_ptr(), lanes, tol=1e-10,
                                                          max_iter=max_iter)
        if (its < 0).any():
            raise SystemExit(f"lanes did not converge: {its}")
        log(f"wide step {k}: {time.perf_counter() - t0:.3f} s so far, {passes} FP32 passes, "
            f"{updates} FP64 replacements")
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    xfe_amd.profile_enable(False)
    pass_ms, launches, pass_bytes = xfe_amd.profile_get()
    z_wide = z_d.cpu().numpy()
    # the host-pack entry point (PCIe transfers of the 256 x J packs included)
    t0 = time.perf_counter()
    z_host, _, _ = solver.solve_wide(rhs, tol=1e-10, max_iter=max_iter)
    wide_host_s = time.perf_counter() - t0
    if not np.array_equal(z_host, z_wide):
        raise SystemExit("host-pack and device entry points differ")
    log(f"wide host-pack entry point (PCIe included): {wide_host_s:.3f} s")
    # the same wide solver with Jacobi alone (no aggregate blocks), timed once
    os.environ["XFE_AKM_JACOBI"] = "1"
    jac = xfe_amd.AkmSolver(st)
    del os.environ["XFE_AKM_JACOBI"]
    jac.solve_wide(rhs[:8], tol=1e-10, max_iter=max_iter)
    t0 = time.perf_counter()
    _, its_j, (passes_j, _) = jac.solve_wide(rhs, tol=1e-10, max_iter=max_iter)
    wide_jacobi_s = time.perf_counter() - t0
    jac.close()
    log(f"wide Jacobi: {wide_jacobi_s:.3f} s, {passes_j} FP32 passes")
    # the reference's pattern: FP64, 32 lanes per launch, 8 batches (timed once)
    t0 = time.perf_counter()
    all_its = []
    err = 0.0
    for b0 in range(0, lanes, batch):
        zb, itb = solver.solve_multi(rhs[b0:b0 + batch], tol=1e-10, max_iter=max_iter)
        if (itb < 0).any():
            raise SystemExit(f"lanes did not converge: {itb}")
        all_its.append(itb)
        err = max(err, float(np.max(np.abs(z_wide[b0:b0 + batch] - zb) / np.abs(zb).max(axis=1, keepdims=True))))
    fp64_32_s = time.perf_counter() - t0
    its64 = np.concatenate(all_its)
    log(f"FP64 32-lane batches: {fp64_32_s:.3f} s, iterations {its64.min()}..{its64.max()}; "
        f"wide vs FP64 max rel diff {err:.2e}")
    # the JLA draws with the reference's default 200 (streams, right-hand
    # sides, solves and per-row sums on the device): P_ii, then the B_ii
    # weights of the three components (two solves per draw)
    t0 = time.perf_counter()
    pii, jla_ok = solver.jla_pii(draws=200, seed=20260705, tol=1e-10, max_iter=max_iter)
    jla_s = time.perf_counter() - t0
    log(f"JLA P_ii, 200 draws: {jla_s:.2f} s, converged {jla_ok}, mean P_ii {pii.mean():.4f}")
    t0 = time.perf_counter()
    bfe, bpe, bcov, bii_ok = solver.jla_bii(None, draws=200, seed=20260705, tol=1e-10, max_iter=max_iter)
    bii_s = time.perf_counter() - t0
    log(f"JLA B_ii, 200 draws: {bii_s:.2f} s, converged {bii_ok}, sums {bfe.sum():.4g} {bpe.sum():.4g} "
        f"{bcov.sum():.4g}")
    solver.close()
    # exact leverages (the reference's LeverageMethod::Exact: one solve per
    # match, here 256 matches per wide solve) and the KSS decomposition built
    # on them, against the 200 JLA draws, on the leave-out connected set (the
    # device's xfe_akm_leave_out_set) of a 200k-obs panel of the same shape
    # (24k workers, 3k firms); y = alpha + psi + N(0, 1), synthetic
    ps = panels.kss_panel(workers=24_000, firms=3_000)
    keep, _ = xfe_amd.leave_out_set(ps["worker"], ps["firm"])
    keep = np.asarray(keep, dtype=bool)
    wk, fk = ps["worker"][keep], ps["firm"][keep]
    rng_y = np.random.default_rng(7)
    yk = (rng_y.normal(size=int(wk.max()) + 1)[wk] + rng_y.normal(0.0, 0.3, size=int(fk.max()) + 1)[fk]
          + rng_y.normal(size=wk.size))
    t0 = time.perf_counter()
    dec = akm.kss_components_exact(yk, wk, fk, tol=1e-10, max_iter=max_iter, accept=1e-8)
    exact_s = time.perf_counter() - t0
    ex_p, ex_fe, ex_pe, ex_cov = dec["pii"], dec["bfe"], dec["bpe"], dec["bcov"]
    t0 = time.perf_counter()
    dj = akm.kss_components(yk, wk, fk, method="jla", draws=200, seed=20260705, tol=1e-10, max_iter=max_iter)
    jla_s = time.perf_counter() - t0
    jp2, jfe, jpe, jcov = dj["pii"], dj["bfe"], dj["bpe"], dj["bcov"]
    N2, J2 = int(np.unique(wk).size), int(np.unique(fk).size)
    mw2 = dec["match_worker"]
    exact_cmp = {
        "obs": int(ps["worker"].size), "workers": int(N2), "firms": int(J2), "matches": int(mw2.size),
        "rows_leave_out_set": int(wk.size),
        "exact_s": round(exact_s, 3), "seconds": {k: round(v, 3) for k, v in dec["seconds"].items()},
        "kss": [float(v) for v in dec["kss"]], "plugin": [float(v) for v in dec["plugin"]],
        "unit_leverage_rows": dec["unit_leverage"],
        "jla200_s": round(jla_s, 3), "jla200_seconds": {k: round(v, 3) for k, v in dj["seconds"].items()},
        "jla200_kss": [float(v) for v in dj["kss"]],
        "tol": 1e-10, "accept_on_stall": 1e-8,
        "sum_pii_minus_rank": float(ex_p.sum() - (N2 + J2 - 1)),
        "jla200_pii_mean_abs_err": float(np.abs(jp2 - ex_p).mean()),
        "jla200_pii_max_abs_err": float(np.abs(jp2 - ex_p).max()),
        "jla200_sum_b_rel_err": [float(abs(j.sum() - e.sum()) / abs(e.sum()))
                                 for j, e in ((jfe, ex_fe), (jpe, ex_pe), (jcov, ex_cov))],
    }
    log(f"KSS decomposition with 200 JLA draws: {jla_s:.2f} s {exact_cmp['jla200_seconds']}; "
        f"KSS {exact_cmp['jla200_kss']}")
    log(f"KSS decomposition with exact leverages, {mw2.size} matches ({wk.size} obs of {ps['worker'].size} in "
        f"the leave-out set): {exact_s:.2f} s {exact_cmp['seconds']}; KSS {exact_cmp['kss']}; "
        f"JLA-200 P_ii mean abs err {exact_cmp['jla200_pii_mean_abs_err']:.3e}, sum B rel err "
        f"{exact_cmp['jla200_sum_b_rel_err']}")
    # the JLA decomposition (200 draws) at the full panel's shape: its leave-out
    # connected set, y = alpha + psi + N(0, 1) synthetic
    t0 = time.perf_counter()
    keep, _ = xfe_amd.leave_out_set(p["worker"], p["firm"])
    keep = np.asarray(keep, dtype=bool)
    lo_s = time.perf_counter() - t0
    wk, fk = p["worker"][keep], p["firm"][keep]
    yk = (rng_y.normal(size=int(wk.max()) + 1)[wk] + rng_y.normal(0.0, 0.3, size=int(fk.max()) + 1)[fk]
          + rng_y.normal(size=wk.size))
    t0 = time.perf_counter()
    dfull = akm.kss_components(yk, wk, fk, method="jla", draws=200, seed=20260705, tol=1e-10, max_iter=max_iter)
    full_s = time.perf_counter() - t0
    kss_full = {"obs": int(n), "rows_leave_out_set": int(wk.size), "matches": int(dfull["match_count"].size),
                "leave_out_set_s": round(lo_s, 3), "decomposition_s": round(full_s, 3),
                "seconds": {k: round(v, 3) for k, v in dfull["seconds"].items()},
                "plugin": [float(v) for v in dfull["plugin"]], "kss": [float(v) for v in dfull["kss"]],
                "agsu": [float(v) for v in dfull["agsu"]], "unit_leverage_rows": dfull["unit_leverage"]}
    log(f"KSS decomposition, 200 JLA draws, full panel: {wk.size} of {n} rows in the leave-out set "
        f"({lo_s:.2f} s), {full_s:.2f} s in all (match table and sums on the host included) {kss_full['seconds']}; "
        f"KSS {kss_full['kss']}")
    achieved = pass_bytes / (pass_ms * 1e-3) / 1e9 if pass_ms > 0 else 0.0
    log(f"{sec:.3f} s per {lanes} lanes; wide_pair_kernel: {launches} launches, "
        f"{pass_ms / max(launches, 1):.4f} ms avg, {achieved:.0f} GB/s algorithmic")
    cpu = None
    if not args.no_cpu:
        cpu = akm_cpu_baseline(st, its64, lanes)
        log(f"cpu baseline {cpu['value']} s (extrapolated)")
    out = {
        "metric": AKM_METRIC, "value": round(sec, 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "method": ("FP32 PCG on the firm-pair form of S, 256 lanes per operator pass, aggregate blocks of "
                   "strongly coupled firms solved exactly (Jacobi elsewhere), FP64 residual replacement "
                   "(x accumulated in FP64), per-lane retire on the FP64 residual"),
        "config": {"workload": "kss_leverage_256rhs", "obs": int(n), "workers": N, "firms": J, "matches": int(M),
                   "lanes": lanes, "lanes_per_launch": lanes, "cg_tol": 1e-10},
        "iterations": {"min": int(its.min()), "max": int(its.max()), "mean": round(float(its.mean()), 1),
                       "fp32_passes": passes, "fp64_replacements": updates},
        "wide_host_pack_s": round(wide_host_s, 4),
        "wide_jacobi_s": round(wide_jacobi_s, 4), "wide_jacobi_fp32_passes": passes_j,
        "fp64_32lane_batches_s": round(fp64_32_s, 4),
        "fp64_32lane_iterations": {"min": int(its64.min()), "max": int(its64.max())},
        "max_rel_diff_vs_fp64_32lane": err,
        "jla_pii_200_draws_s": round(jla_s, 3), "jla_bii_200_draws_s": round(bii_s, 3),
        "exact_vs_jla_200k": exact_cmp,
        "kss_jla200_full": kss_full,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(os.path.join(ROOT, "profiles", "wide_pair_kss10m.md")),
                     "kernel": "wide_pair_kernel<float, 4> (firm-major pass over the firm pairs, 256 lanes)",
                     "avg_launch_ms": round(pass_ms / max(launches, 1), 4),
                     "bytes_per_launch": round(pass_bytes / max(launches, 1), 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


BOOT_METRIC = "s for 2,000 uneven-size cluster-bootstrap refits of a 500k-obs two-way FE regression"


def bench_boot(args):
    """Cluster (worker) bootstrap: 2,000 resamples of the workers of a 500k-obs
    two-way panel (50,000 workers x 10 years, 2,000 firms, the two-way DGP
    with three regressors), resample sizes U(0.7, 1.3) x workers; each refit is
    the two-way FE regression with rows weighted by their worker's
    multiplicity; 64 refits (x 4 columns) share each FP32 operator pass of
    the wide solver, the Grams are FP64.  Timed: the refits from the
    multiplicities on the host to the 2,000 Grams on the host."""
    import xfe_amd
    from xfe_amd import panels

    if not xfe_amd.backend_available():
        raise SystemExit("no HIP device")
    p = panels.config0(seed=panels.SEED0 + 900, workers=50_000, years=10, firms=2_000)
    worker = p["fes"][0]
    firm = np.unique(p["fes"][1], return_inverse=True)[1].astype(np.int32)
    V = np.column_stack([p["y"], p["X"]])
    Nw, J, n = int(worker.max()) + 1, int(firm.max()) + 1, worker.size
    rng = np.random.default_rng(20261209)
    R = 2000
    sizes = (rng.uniform(0.7, 1.3, size=R) * Nw).astype(np.int64)
    mult = np.stack([rng.multinomial(int(k), np.ones(Nw) / Nw) for k in sizes]).astype(np.int32)
    log(f"bootstrap panel: {n} obs, {Nw} workers, {J} firms; {R} resamples of {sizes.min()}..{sizes.max()} workers")
    t0 = time.perf_counter()
    boot = xfe_amd.BootRefits(worker, firm, V)
    log(f"context (matches, moments, firm pairs): {time.perf_counter() - t0:.2f} s")
    for w in range(args.warmup):
        boot.refits(mult[:64], tol=1e-10)
    xfe_amd.profile_reset()
    xfe_amd.profile_enable(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        gram, its, (passes, upd), ok = boot.refits(mult, tol=1e-10)
        if not ok:
            raise SystemExit("some refits did not converge")
        log(f"step {k}: {time.perf_counter() - t0:.3f} s so far, {passes} FP32 passes, {upd} FP64 replacements")
    sec = (time.perf_counter() - t0) / args.steps
    xfe_amd.profile_enable(False)
    pass_ms, launches, pass_bytes = xfe_amd.profile_get()
    betas = xfe_amd.BootRefits.betas(gram)
    se = betas.std(axis=0, ddof=1)
    # check refit 0 against the plain FP64 absorber on the explicitly weighted rows
    w0 = mult[0][worker].astype(np.float64)
    keep = w0 > 0
    wk = np.unique(worker[keep], return_inverse=True)[1].astype(np.int32)
    fk = np.unique(firm[keep], return_inverse=True)[1].astype(np.int32)
    yt, Xt, _ = xfe_amd.absorb(V[keep, 0], V[keep, 1:], [wk, fk], [int(wk.max()) + 1, int(fk.max()) + 1],
                               weights=w0[keep], tol=1e-12)
    sw = np.sqrt(w0[keep])
    b_abs = np.linalg.lstsq(Xt * sw[:, None], yt * sw, rcond=None)[0]
    err = float(np.abs(betas[0] - b_abs).max() / np.abs(b_abs).max())
    log(f"{sec:.3f} s for {R} refits; beta SE {se}; refit 0 vs plain FP64 absorber: {err:.2e}")
    cpu = None
    if not args.no_cpu:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        k_cpu = 3
        for r in range(k_cpu):
            wr = mult[r][worker].astype(np.float64)
            kr = wr > 0
            gw = np.unique(worker[kr], return_inverse=True)[1].astype(np.int32)
            gf = np.unique(firm[kr], return_inverse=True)[1].astype(np.int32)
            oracle.absorb(V[kr], [gw, gf], [int(gw.max()) + 1, int(gf.max()) + 1], weights=wr[kr], tol=1e-8,
                          threads=threads)
        dt = (time.perf_counter() - t0) / k_cpu
        cpu = dict(value=round(dt * R, 2), unit="s", cores=threads, kind="port",
                   sample=(f"the oracle's weighted absorber (C restatement of the reference's CPU absorber, "
                           f"tol 1e-8) on {k_cpu} of the resamples, {dt:.3f} s per refit; value = x {R} refits"))
    achieved = pass_bytes / (pass_ms * 1e-3) / 1e9 if pass_ms > 0 else 0.0
    out = {
        "metric": BOOT_METRIC, "value": round(sec, 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "method": ("per batch of 64 resamples x 4 columns: firm-effect right-hand sides and Jacobi diagonals per "
                   "resample, one FP32 PCG over the per-worker firm-pair entries weighted by each lane's "
                   "multiplicity with FP64 residual replacement, FP64 Grams from the match moments"),
        "config": {"workload": "bootstrap_2000_refits_500k", "obs": n, "workers": Nw, "firms": J, "refits": R,
                   "resample_workers": [int(sizes.min()), int(sizes.max())], "columns": V.shape[1], "tol": 1e-10},
        "iterations": {"min": int(its.min()), "max": int(its.max()), "fp32_passes": passes, "fp64_replacements": upd},
        "beta_bootstrap_se": [float(x) for x in se],
        "refit0_vs_plain_fp64_absorber": err,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "wide_pair_kernel<float, 4, true> (per-worker pair entries, 256 lanes)",
                     "avg_launch_ms": round(pass_ms / max(launches, 1), 4),
                     "bytes_per_launch": round(pass_bytes / max(launches, 1), 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


SLOPE_METRIC = "s to absorb the 30M-obs heterogeneous-slope panel (worker, firm intercept + slope on x1, firm x year; 32 columns) at 1e-8, mixed precision"


def slope_case(scale):
    """The slope shape at `scale` of 30M obs, singletons dropped on the device,
    ids compacted: (V n x C, [worker, firm, cell] ids, levels, z, dropped)."""
    import xfe_amd
    from xfe_amd import panels
    p = panels.slope_panel(scale=scale)
    keep, dropped, _ = xfe_amd.drop_singletons([f.astype(np.int32) for f in p["fes"]])
    comp = [compact(f[keep]) for f in p["fes"]]
    V = np.column_stack([p["y"][keep], p["X"][keep]])
    z = p["z"][keep]
    return V, [c[0] for c in comp], [c[1] for c in comp], z, dropped


def bench_slope(args):
    """Config 2: the heterogeneous-slope shape (worker, firm with intercept and
    slope on x1, firm x year cells; y and 31 regressors, C = 32) at --scale of
    30M obs (default 1.0), singletons dropped on the device; the timed step is
    the mixed-precision absorption (xfe_absorb_device_slopes, precision 1) of
    the columns resident in HBM.  Beside it, on the same shape at 3 %, the
    mixed path and the FP64 slope path (CG on the symmetric projection sweep)
    are timed and compared."""
    import torch

    import xfe_amd

    if not xfe_amd.backend_available():
        raise SystemExit("no HIP device")
    scale = args.scale
    t0 = time.perf_counter()
    V, gids, G, z, dropped = slope_case(scale)
    n, C = V.shape
    log(f"slope panel: n={n} C={C} levels={G} ({dropped} singletons) in {time.perf_counter() - t0:.1f} s")
    dev = torch.device("cuda:0")
    V0 = torch.from_numpy(np.ascontiguousarray(V.T)).to(dev)  # C x n: column-major, ld n
    Vd = torch.empty_like(V0)
    ids = torch.from_numpy(np.ascontiguousarray(np.stack(gids))).to(dev)
    zd = torch.from_numpy(np.ascontiguousarray(z)).to(dev)
    vmax = float(np.abs(V).max())
    del V
    torch.cuda.synchronize()

    def step():
        Vd.copy_(V0)
        return xfe_amd.absorb_device_slopes(Vd.data_ptr(), n, C, n, ids.data_ptr(), G, [0, zd.data_ptr(), 0],
                                            [0, 1, 0], tol=1e-8, precision=1)

    for w in range(args.warmup):
        t0 = time.perf_counter()
        info = step()
        torch.cuda.synchronize()
        log(f"warmup {w}: {time.perf_counter() - t0:.3f} s, iterations {info['iterations']}, "
            f"replacements {info['refinements']}, converged {info['converged']}")
    xfe_amd.profile_reset()
    xfe_amd.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        info = step()
        log(f"step {k}: iterations {info['iterations']} replacements {info['refinements']} "
            f"converged {info['converged']} certificate {info['abs_residual_rel']:.2e}")
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    xfe_amd.profile_enable(False)
    pass_ms, launches, pass_bytes = xfe_amd.profile_get()
    achieved = pass_bytes / (pass_ms * 1e-3) / 1e9 if pass_ms > 0 else 0.0
    log(f"{sec:.3f} s per absorption; sm_apply_kernel<float>: {launches} launches, "
        f"{pass_ms / max(launches, 1):.3f} ms avg, {achieved:.0f} GB/s algorithmic")
    del V0, Vd, ids, zd
    torch.cuda.empty_cache()
    # equal-size comparison with the FP64 slope path (host entry points)
    cmp_scale = args.cmp_scale
    Vs, gs, Gs, zs, _ = slope_case(cmp_scale)
    t0 = time.perf_counter()
    ym, Xm, im = xfe_amd.absorb(Vs[:, 0], Vs[:, 1:], gs, Gs, tol=1e-8, precision=1, slopes={1: (zs, True)})
    mixed_small = time.perf_counter() - t0
    t0 = time.perf_counter()
    y6, X6, i6 = xfe_amd.absorb(Vs[:, 0], Vs[:, 1:], gs, Gs, tol=1e-8, precision=0, slopes={1: (zs, True)})
    fp64_small = time.perf_counter() - t0
    y7, X7, i7 = xfe_amd.absorb(Vs[:, 0], Vs[:, 1:], gs, Gs, tol=1e-13, precision=0, slopes={1: (zs, True)})
    ref = np.column_stack([y7, X7])
    colscale = np.abs(Vs).max(axis=0)
    err_m = float((np.abs(np.column_stack([ym, Xm]) - ref).max(axis=0) / colscale).max())
    err_6 = float((np.abs(np.column_stack([y6, X6]) - ref).max(axis=0) / colscale).max())
    log(f"{cmp_scale:g} shape ({Vs.shape[0]} obs): mixed {mixed_small:.2f} s ({im['iterations']} it, "
        f"precision_used {im['precision_used']}), FP64 path "
        f"{fp64_small:.2f} s ({i6['iterations']} it); max column error vs FP64 tol 1e-13: mixed {err_m:.2e}, "
        f"FP64 tol 1e-8 {err_6:.2e}")
    out = {
        "metric": SLOPE_METRIC, "value": round(sec, 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "method": ("worker dimension eliminated; FP32 PCG on the Schur system of the firm x year cells and "
                   "firm slopes (the firm intercept lies in the cells' span) with one exactly inverted block per "
                   "firm (its cells and slope), 32 columns as lanes, restarted FP64 residual replacement; columns "
                   "the 2000 FP32 iterations leave unconverged continue with FP64 block-PCG from their iterate; "
                   "FP64 final transform and certificate; the FP64 slope path only if that fails; columns "
                   "resident in HBM"),
        "path": ("mixed" if info["precision_used"] == 1 else
                 "FP64 slope path (fallback: the mixed run did not converge)" if info["precision_fallback"]
                 else "FP64 slope path"),
        "config": {"workload": f"slope_{scale:g}x30m_32cols", "obs": int(n), "columns": C, "levels": G,
                   "slope": "firm intercept + slope on x1", "tol": 1e-8, "singletons_dropped": int(dropped)},
        "iterations": info["iterations"], "replacements": info["refinements"], "converged": info["converged"],
        "certificate_rel": info["abs_residual_rel"], "certified": info["precision_certified"],
        "obs_cols_per_s": round(n * C / sec, 1),
        "equal_size_small": {"scale": cmp_scale, "obs": int(Vs.shape[0]), "mixed_s": round(mixed_small, 3),
                             "mixed_precision_used": im["precision_used"],
                            "fp64_slope_path_s": round(fp64_small, 3), "mixed_iterations": im["iterations"],
                            "fp64_iterations": i6["iterations"], "mixed_max_col_err_vs_fp64_tight": err_m,
                            "fp64_tol_1e8_max_col_err_vs_fp64_tight": err_6,
                            "note": "host entry points (PCIe transfers included) for both"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "sm_apply_kernel<float> (per worker: coefficient gathers, worker mean, R rows)",
                     "avg_launch_ms": round(pass_ms / max(launches, 1), 4),
                     "bytes_per_launch": round(pass_bytes / max(launches, 1), 1)},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_recover(args):
    """The savefe FE recovery (xfe_recover_fixed_effects) of y on the AKM
    panel at --scale (1.0 = 56M obs): the host entry point (upload, Schur
    solve, coefficients, expansion, max-delta check, download) timed."""
    import torch  # noqa: F401  (shares torch's HIP runtime)

    import xfe_amd
    from xfe_amd import panels

    if not xfe_amd.backend_available():
        raise SystemExit("no HIP device")
    t0 = time.perf_counter()
    p = panels.config1() if args.scale >= 1.0 else panels.akm_small(scale=args.scale)
    keep, dropped, _ = xfe_amd.drop_singletons(p["fes"])
    comp = [compact(f[keep]) for f in p["fes"]]
    gids, G = [c[0] for c in comp], [c[1] for c in comp]
    y = np.ascontiguousarray(p["y"][keep])
    del p
    n = y.size
    log(f"recovery panel: n={n} levels={G} in {time.perf_counter() - t0:.1f} s")
    for w in range(args.warmup):
        t0 = time.perf_counter()
        _, _, info = xfe_amd.recover_fixed_effects(y, gids, G)
        log(f"warmup {w}: {time.perf_counter() - t0:.3f} s {info}")
    t0 = time.perf_counter()
    for k in range(args.steps):
        alphas, contrib, info = xfe_amd.recover_fixed_effects(y, gids, G)
    sec = (time.perf_counter() - t0) / args.steps
    log(f"{sec:.3f} s per recovery, {info}")
    out = {
        "metric": "s to recover the FE values of y (savefe) on the AKM panel", "value": round(sec, 4),
        "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "method": "Schur solve of the non-primary levels (FP32 PCG under FP64 refinement), primary-level means, "
                  "contributions expanded on the host, one Gauss-Seidel check pass (host entry: PCIe included)",
        "config": {"workload": f"akm_{n}obs_recover", "obs": int(n), "levels": G, "fe_tol": 1e-6,
                   "singletons_dropped": int(dropped)},
        "max_delta": info["max_delta"], "converged": info["converged"], "iterations": info["iterations"],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def gi_big_panel(n=2_000_000, individuals=500_000, seed=20261203 + 600):
    """A group()/individual() panel of n rows (patents) with 1-3 distinct
    individuals each (log-normal activity), mean aggregation, standard FEs
    year (30) and class (600), k = 3 regressors."""
    rng = np.random.default_rng(seed)
    act = rng.lognormal(0.0, 1.0, individuals)
    cdf = np.cumsum(act) / act.sum()
    size = rng.integers(1, 4, n)
    draws = np.searchsorted(cdf, rng.random((n, 3)), side="right").astype(np.int64)
    draws = np.minimum(draws, individuals - 1)
    rows = np.repeat(np.arange(n, dtype=np.int64), 3)
    cand = draws.ravel()
    mask = (np.tile(np.arange(3), n) < np.repeat(size, 3))
    pair = np.unique(rows[mask] * individuals + cand[mask])
    grow, gind = (pair // individuals).astype(np.int64), (pair % individuals).astype(np.int32)
    cnt = np.bincount(grow, minlength=n)
    gp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(cnt, out=gp[1:])
    order = np.lexsort((grow, gind))
    ip = np.zeros(individuals + 1, dtype=np.int64)
    np.cumsum(np.bincount(gind, minlength=individuals), out=ip[1:])
    scale = 1.0 / cnt.astype(np.float64)
    year = rng.integers(0, 30, n).astype(np.int32)
    cls = rng.integers(0, 600, n).astype(np.int32)
    alpha = rng.normal(0.0, 0.5, individuals)
    team = np.bincount(grow, weights=alpha[gind] * scale[grow], minlength=n)
    X = np.asfortranarray(rng.standard_normal((n, 3)))
    y = X @ np.array([0.4, -0.3, 0.2]) + team + rng.normal(0, 0.25, n)
    gi = dict(num_individuals=individuals, group_ptr=gp.astype(np.int32), group_individual=gind,
              individual_ptr=ip.astype(np.int32), individual_group=grow[order].astype(np.int32), group_scale=scale)
    return y, X, [year, cls], [30, 600], gi


def bench_gi(args):
    """The group()/individual() absorber (xfe_absorb_group_individual) on a
    2M-row panel: the host entry point timed (upload, sweeps, certificate,
    download)."""
    import torch  # noqa: F401  (shares torch's HIP runtime)

    import xfe_amd

    if not xfe_amd.backend_available():
        raise SystemExit("no HIP device")
    t0 = time.perf_counter()
    y, X, fes, G, gi = gi_big_panel()
    n = y.size
    log(f"gi panel: n={n} individuals={gi['num_individuals']} incidences={gi['group_ptr'][-1]} "
        f"in {time.perf_counter() - t0:.1f} s")
    for w in range(args.warmup):
        t0 = time.perf_counter()
        _, _, info = xfe_amd.absorb_group_individual(y, X, fes, G, gi)
        log(f"warmup {w}: {time.perf_counter() - t0:.3f} s, sweeps {info['iterations']}")
    t0 = time.perf_counter()
    for k in range(args.steps):
        _, _, info = xfe_amd.absorb_group_individual(y, X, fes, G, gi)
    sec = (time.perf_counter() - t0) / args.steps
    log(f"{sec:.3f} s per absorption, sweeps {info['iterations']}, certificate {info['abs_residual_rel']:.2e}")
    out = {
        "metric": "s to absorb a 2M-row group()/individual() panel at 1e-8", "value": round(sec, 4),
        "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "method": "colour-class Gauss-Seidel sweeps, change-of-norm stop, host certificate (host entry: PCIe included)",
        "config": {"workload": "gi_2m_patents", "obs": int(n), "individuals": gi["num_individuals"],
                   "incidences": int(gi["group_ptr"][-1]), "standard_levels": G, "columns": 4, "tol": 1e-8},
        "iterations": info["iterations"], "converged": info["converged"],
        "certificate_rel": info["abs_residual_rel"], "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def bench_leaveout(args):
    """The leave-out connected set (xfe_akm_leave_out_set) at the KSS shape
    (1.2M workers, 150k firms, about 10M rows; --scale shrinks it): the host
    entry point timed; the reference's own leave_out_connected_set (compiled
    from its sources into the checker library, one thread) beside it when
    that library is present."""
    import torch  # noqa: F401  (shares torch's HIP runtime)

    import xfe_amd
    from xfe_amd import panels

    if not xfe_amd.backend_available():
        raise SystemExit("no HIP device")
    sc = min(args.scale, 1.0)
    p = panels.kss_panel(workers=int(1_200_000 * sc), firms=int(150_000 * sc))
    w, f = p["worker"].astype(np.int32), p["firm"].astype(np.int32)
    log(f"kss panel: n={w.size}")
    for k in range(args.warmup):
        xfe_amd.leave_out_set(w, f)
    t0 = time.perf_counter()
    for k in range(args.steps):
        keep, info = xfe_amd.leave_out_set(w, f)
    sec = (time.perf_counter() - t0) / args.steps
    log(f"{sec:.3f} s per leave-out set, {info}")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        import oracle
        if oracle.ref() is not None:
            t0 = time.perf_counter()
            rk, ri = oracle.ref_leave_out_set(w, f)
            ref_s = time.perf_counter() - t0
            log(f"reference leave_out_connected_set: {ref_s:.3f} s, same rows: {bool((rk == keep).all())}")
            cpu = {"value": round(ref_s, 3), "unit": "s", "cores": 1, "kind": "reference",
                   "sample": "the whole panel", "same_rows": bool((rk == keep).all())}
    out = {
        "metric": "s to build the KSS leave-out connected set (10M rows)", "value": round(sec, 4), "unit": "s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "method": "match-key sort and Tarjan-Vishkin cut-vertex search on the device, host bookkeeping",
        "config": {"workload": f"kss_{w.size}rows_leave_out", "rows": int(w.size)},
        "result": {k: info[k] for k in ("n_obs_connected", "n_obs", "n_workers", "n_firms", "n_movers",
                                        "prune_iterations")},
        "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="panel scale (1.0 = 56M obs)")
    ap.add_argument("--precision", type=int, default=1, help="1 FP32 under FP64 refinement, 0 FP64")
    ap.add_argument("--plain-steps", type=int, default=1, help="timed steps of the plain FP64 absorber")
    ap.add_argument("--preconditioner", type=int, default=2,
                    help="precision 1: 0 Jacobi, 1 strength aggregates, 2 learned aggregates (default)")
    ap.add_argument("--ablation-steps", type=int, default=2,
                    help="timed steps of the mixed method with the other preconditioners")
    ap.add_argument("--cmp-scale", type=float, default=0.03,
                    help="--path slope: scale of the mixed vs FP64 slope path comparison")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--path", choices=("absorb", "akm", "boot", "slope", "recover", "gi", "leaveout"),
                    default="absorb",
                    help="absorb: the 56M-obs absorber (default); akm: the KSS leverage solver; "
                         "boot: 2,000 cluster-bootstrap refits; "
                         "slope: the heterogeneous-slope path; recover: savefe FE recovery; "
                         "gi: the group()/individual() absorber; leaveout: the KSS leave-out connected set")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              f"(the absorber runs on a single device)", file=sys.stderr)
        sys.exit(2)
    if args.path == "akm":
        bench_akm(args)
        return
    if args.path == "boot":
        bench_boot(args)
        return
    if args.path == "slope":
        bench_slope(args)
        return
    if args.path == "recover":
        bench_recover(args)
        return
    if args.path == "gi":
        bench_gi(args)
        return
    if args.path == "leaveout":
        bench_leaveout(args)
        return

    import torch

    import xfe_amd
    from xfe_amd import panels

    if not xfe_amd.backend_available():
        raise SystemExit("no HIP device")
    t0 = time.perf_counter()
    if args.scale >= 1.0:
        p = panels.config1()
    else:
        p = panels.akm_small(scale=args.scale)
    log(f"panel generated: n={p['y'].size} in {time.perf_counter() - t0:.1f} s")

    t0 = time.perf_counter()
    keep, dropped, rounds = xfe_amd.drop_singletons(p["fes"])
    t_single = time.perf_counter() - t0
    log(f"singletons: dropped {dropped} in {rounds} rounds, {t_single:.2f} s (device, incl. upload)")
    fes = [f[keep] for f in p["fes"]]
    comp = [compact(f) for f in fes]
    gids = [c[0] for c in comp]
    G = [c[1] for c in comp]
    y = p["y"][keep]
    X = p["X"][keep]
    del p
    n = y.size
    C = X.shape[1] + 1
    ld = ((n + 63) // 64) * 64
    log(f"n={n} C={C} levels={G}")

    dev = torch.device("cuda:0")
    V0 = torch.zeros((C, ld), dtype=torch.float64, device=dev)
    V0[0, :n] = torch.from_numpy(y).to(dev)
    V0[1:, :n] = torch.from_numpy(np.ascontiguousarray(X.T)).to(dev)
    ids = torch.from_numpy(np.ascontiguousarray(np.stack(gids))).to(dev)
    V = torch.empty_like(V0)
    del X, y
    torch.cuda.synchronize()

    lp_params = xfe_amd.default_lp_params()

    def restore():
        # the transform overwrites its input: a fresh copy of the panel before
        # each step (outside the timed region)
        V.copy_(V0)
        torch.cuda.synchronize()

    def step(precision, pc=0):
        return xfe_amd.absorb_device(V.data_ptr(), n, C, ld, ids.data_ptr(), G, tol=1e-8,
                                     tolerance_mode=1, precision=precision, preconditioner=pc,
                                     lp_params=lp_params if pc == 2 else None)

    def timed(precision, steps, warmup, profile, pc=0):
        for w in range(warmup):
            restore()
            t0 = time.perf_counter()
            info = step(precision, pc)
            torch.cuda.synchronize()
            log(f"precision {precision} preconditioner {pc} warmup {w}: {time.perf_counter() - t0:.3f} s, "
                f"iterations {info['iterations']}, converged {info['converged']}")
        if profile:
            xfe_amd.profile_reset()
            xfe_amd.profile_enable(True)
        elapsed = 0.0
        infos = []
        for k in range(steps):
            restore()
            t0 = time.perf_counter()
            infos.append(step(precision, pc))   # returns after its own device work has finished
            torch.cuda.synchronize()
            elapsed += time.perf_counter() - t0
            log(f"precision {precision} preconditioner {pc} step {k}: iterations {infos[-1]['iterations']} "
                f"refinements {infos[-1]['refinements']} converged {infos[-1]['converged']} "
                f"precond setup {infos[-1]['precond_ms']:.1f} ms")
        if profile:
            xfe_amd.profile_enable(False)
        return elapsed / steps, infos[-1]

    # the plain FP64 absorber at the same tolerance, timed, and once at tol 1e-14
    # (untimed) as the FP64 reference the results are compared with
    plain_s = None
    V_plain = V_ref = None
    plain_info = ref_info = None
    if args.plain_steps > 0 and args.precision != 0:
        plain_s, plain_info = timed(0, args.plain_steps, 1, False)
        V_plain = V.clone()
        log(f"plain FP64: {plain_s:.3f} s/step, {plain_info['iterations']} iterations, "
            f"certificate {plain_info['abs_residual_rel']:.3e}")
        V.copy_(V0)
        ref_info = xfe_amd.absorb_device(V.data_ptr(), n, C, ld, ids.data_ptr(), G, tol=1e-14,
                                         tolerance_mode=1, precision=0)
        V_ref = V.clone()
        log(f"FP64 reference (tol 1e-14): {ref_info['iterations']} iterations, "
            f"certificate {ref_info['abs_residual_rel']:.3e}")

    # the mixed method with the other preconditioners (the method without the
    # learned part: Jacobi; and the aggregation by coupling strength alone)
    ablation = None
    if args.precision == 1 and args.ablation_steps > 0:
        ablation = {}
        for pc, name in ((0, "mixed_jacobi"), (1, "mixed_strength_aggregates")):
            if pc == args.preconditioner:
                continue
            a_s, a_info = timed(1, args.ablation_steps, 1, False, pc)
            ablation[name] = {"s": round(a_s, 4), "iterations": a_info["iterations"],
                              "refinements": a_info["refinements"], "precond_setup_ms": round(a_info["precond_ms"], 2)}
            log(f"ablation {name}: {a_s:.4f} s/step, {a_info['iterations']} inner iterations")
        # the same Schur-complement Jacobi-PCG entirely in FP64 (precision 2): what
        # the precision alone changes, against mixed_jacobi
        a_s, a_info = timed(2, args.ablation_steps, 1, False, 0)
        ablation["schur_fp64_jacobi"] = {"s": round(a_s, 4), "iterations": a_info["iterations"],
                                         "note": "precision 2: mixed_jacobi's algorithm in FP64 throughout"}
        log(f"ablation schur_fp64_jacobi: {a_s:.4f} s/step, {a_info['iterations']} iterations")

    # the timed steps run as in production (the mixed path replays its inner
    # iterations as HIP graphs); the per-launch kernel timing of the roofline
    # comes from one more step with the profile's HIP events, which issues the
    # same launches one by one
    sec, info = timed(args.precision, args.steps, args.warmup, False, args.preconditioner)
    _, _ = timed(args.precision, 1, 0, True, args.preconditioner)
    pass_ms, launches, pass_bytes = xfe_amd.profile_get()
    achieved = pass_bytes / (pass_ms * 1e-3) / 1e9 if pass_ms > 0 else 0.0
    kernel, traffic_file = KERNELS[args.precision]
    log(f"{sec:.3f} s/step; {kernel.split()[0]}: {launches} launches, "
        f"{pass_ms / max(launches, 1):.3f} ms avg, {achieved:.0f} GB/s algorithmic")
    accuracy = None
    V_plain_keep = None
    if V_ref is not None:
        scale = float(V0[:, :n].abs().max())

        def err(A):
            return float((A[:, :n] - V_ref[:, :n]).abs().max()) / scale

        accuracy = {"fp64_reference": "plain FP64 absorber at tol 1e-14",
                    "max_abs_err_over_max_abs_v": err(V),
                    "plain_fp64_default_tol_max_abs_err_over_max_abs_v": err(V_plain),
                    "certificate_rel": info["abs_residual_rel"],
                    "certified": info["precision_certified"],
                    "plain_fp64_certificate_rel": plain_info["abs_residual_rel"],
                    "fp64_reference_certificate_rel": ref_info["abs_residual_rel"]}
        log(f"accuracy: {accuracy}")
        V_plain_keep = V_plain
        del V_plain, V_ref

    # the shadow device certificate of the last step's result (outside the timed region)
    torch.cuda.synchronize()
    shadow = xfe_amd.cert_shadow_raw(V0.data_ptr(), V0.data_ptr() + 8 * ld, V.data_ptr(), V.data_ptr() + 8 * ld,
                                     n, C - 1, ld, [ids.data_ptr() + 4 * n * d for d in range(len(G))], G,
                                     limit=8e-9)
    shadow = {"verdict": shadow["verdict"], "rel_max": shadow["rel_max"], "r_upper": shadow["r_upper"],
              "ms": round(shadow["shadow_begin_ms"] + shadow["shadow_finish_ms"], 2), "limit": 8e-9}
    log(f"shadow certificate: {shadow}")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline on {threads} threads (full panel) ...")
        cpu = cpu_baseline(V0, ids, n, C, G, V_plain_keep, threads,
                           plain_info["iterations"] if plain_info is not None else None)
        log(f"cpu baseline {cpu['value']} s; {cpu['sample']}; vs HIP plain "
            f"{cpu['max_abs_diff_vs_hip_plain_over_max_abs_v']}")

    out = {
        "metric": METRIC,
        "value": round(sec, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == 1 else "f64",
        "data": "synthetic",
        "method": (("mixed precision: FP32 PCG on the Schur complement of the worker FE with FP64 residual "
                    "replacement; preconditioner: ")
                   + {0: "Jacobi", 1: "aggregate blocks from the coupling strengths",
                      2: "learned (graph-network aggregation into exactly solved firm blocks)"}[args.preconditioner])
                  if args.precision == 1 else "plain FP64: Irons-Tuck accelerated Gauss-Seidel",
        "preconditioner_used": info.get("preconditioner_used"),
        "precond_fallback": info.get("precond_fallback"),
        "precond_setup_ms": round(info.get("precond_ms", 0.0), 2),
        "precond_blocked_firms": info.get("precond_blocks"),
        "ablation": ablation,
        "config": {"workload": "akm_56m_3way", "obs": n, "workers": G[0], "firms": G[1],
                   "years": G[2], "columns": C, "tol": 1e-8,
                   "tolerance_mode": "reghdfe-comparable", "singletons_dropped": dropped},
        "obs_cols_per_s": round(n * C / sec, 1),
        "iterations": info["iterations"],
        "refinements": info["refinements"],
        "plain_fp64_s": None if plain_s is None else round(plain_s, 4),
        "speedup_vs_plain_fp64": None if plain_s is None else round(plain_s / sec, 2),
        "timing": ("per step, the absorber call alone (it returns after its device work has finished); "
                   "the input block, which the transform overwrites, is restored from a device copy of "
                   "the panel before each step, untimed"),
        "accuracy": accuracy,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(traffic_file), "kernel": kernel,
                     "avg_launch_ms": round(pass_ms / max(launches, 1), 4),
                     "bytes_per_launch": round(pass_bytes / max(launches, 1), 1)},
        "cpu_baseline": cpu,
        "singleton_s": round(t_single, 3),
        "shadow_certificate": shadow,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
