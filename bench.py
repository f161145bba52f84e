"""Headline benchmark: 3D lid-driven cavity projection steps at 256^3.

Workload (the configuration the headline metric is quoted on): unit cube,
256^3 MAC grid, Re 1000, dt 5e-4 (the reference's 3D default,
src/solver/incompressible_3d.rs:39), central advective predictor, pressure
solved each step to rel-res 1e-8 from a zero guess.  A "step" is one whole
projection step (predictor, RHS, pressure solve, correction, diagnostics) on
device-resident fields; the flow starts from rest (synthetic data: no
dataset is needed).

Variants timed back to back on the same workload:
  plain_fp64_pcg  FP64 Jacobi-PCG (the judge)
  mixed_ir        FP32 stencil CG inside FP64 iterative refinement (FP64
                  residual, dots, alpha/beta and stop test)
  mg_fcg          the same refinement with multigrid-preconditioned flexible
                  CG as the inner solver (plain FP64 PCG fallback)
  learned_fcg     the same flexible CG whose V-cycle takes the U-Net's
                  prediction as its coarse solve on the 64^3 level (weights
                  from flursys_amd/train_unet.py; skipped when none are
                  present)
`value` is the steps/s of learned_fcg, the metric's method; every variant
meets the same FP64 rel-res 1e-8 stop test, and each one's steps/s and CG
iterations/step are top-level fields beside it (`config.variants` holds the
details).  Supporting configurations, one record each: C1 (2D 128^2), C2 (3D
128^3, both precisions, with the per-field error against the FP64 oracle)
and C3 (2M-cell CSR CG / PCG / BiCGSTAB).

Output: one JSON line with steps/s, CG iterations/step, the HBM roofline of
the dominant kernel (timed live with HIP events on the library's stream) and
a CPU baseline (the oracle timed on this host on a bounded sample).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n, re, dt, iters_per_step, sample_iters):
    """Oracle ("port") timed on the host: one step's predictor + RHS +
    correction plus `sample_iters` PCG iterations at the full grid size;
    steps/s is extrapolated to the GPU's measured iterations per step."""
    import oracle as O

    threads = os.cpu_count() or 1
    c = O.Cavity3D(n, reynolds=re, dt=dt, pressure_tolerance=1e-30,
                   stop_rule=O.STOP_REL_L2, initial_guess=O.GUESS_ZERO,
                   pressure_max_iters=sample_iters, threads=threads)
    t0 = time.perf_counter()
    c.predict()
    c.build_rhs()
    t1 = time.perf_counter()
    rep = c.solve_pressure()
    t2 = time.perf_counter()
    c.correct()
    t3 = time.perf_counter()
    per_iter = (t2 - t1) / max(rep.iterations, 1)
    fixed = (t1 - t0) + (t3 - t2)
    step_s = fixed + per_iter * iters_per_step
    return {
        "value": 1.0 / step_s,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"C oracle (OpenMP FP64 restatement of the reference path) at {n}^3: "
                   f"predictor+RHS+correction once and {rep.iterations} Jacobi-PCG "
                   f"iterations timed ({t3 - t0:.1f} s), extrapolated to "
                   f"{iters_per_step:.0f} iterations/step (the plain FP64 GPU count); "
                   f"{per_iter * 1e3:.1f} ms/iteration"),
    }


def run_variant(fa, ctx, name, precision, precond, args, learned_mode=None):
    n = args.n
    cfg = fa.LidDrivenCavity3DConfig(nx=n, ny=n, nz=n, reynolds=args.reynolds, dt=args.dt,
                                     pressure_tolerance=args.tol, stop_rule=fa.STOP_REL_L2,
                                     initial_guess=fa.GUESS_ZERO, pressure_max_iters=50000,
                                     precision=precision, preconditioner=precond,
                                     inner_tolerance=args.inner_tol)
    log(f"[bench] {name}: {n}^3 Re={args.reynolds} dt={args.dt} tol={args.tol}")
    s = fa.LidDrivenCavity3DSolver(cfg, ctx)
    net = None
    if precond == fa.PRECOND_LEARNED:
        from flursys_amd.unet import UNet

        net = UNet(args.unet_n, 1, ctx=ctx)
        net.load(args.weights)
        s.set_unet(net)
        if learned_mode is not None:
            s.set_learned_mode(learned_mode)
    # The last warmup step times every kernel (untimed region; the first
    # step's launches pay one-time costs): its breakdown names the dominant
    # kernel, which alone is event-timed in the timed region.
    for k in range(args.warmup):
        if k == args.warmup - 1:
            s.set_timing(True)
        rep = s.advance_one_step()
        log(f"[bench] {name} warmup {k + 1}: {rep.pressure_iterations} iterations, "
            f"{rep.pressure_refinements} refinements, fallback {rep.pressure_fallback}")
    breakdown = s.kernel_stats() if args.warmup else {}
    # the dominant single HBM-bound kernel: the U-Net entries and the coarse
    # multigrid levels (k_mg_down/up/tail: the 128^3 level's three sweep
    # launches timed as one entry, and cache-resident 64^3-8^3 levels) are
    # groups of launches whose bytes are no HBM roofline
    groups = ("unet_", "k_mg_down", "k_mg_up", "k_mg_tail")
    single = {k: v for k, v in breakdown.items() if not k.startswith(groups)}
    dom = max(single, key=lambda k: single[k]["total_ms"]) if single else None
    s.set_timing(True, dom)
    iters, refs, fbs, warm = [], [], [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        rep = s.advance_one_step()  # returns after its stream has drained
        iters.append(rep.pressure_iterations)
        refs.append(rep.pressure_refinements)
        fbs.append(rep.pressure_fallback)
        warm.append(rep.pressure_warm_start)
        if not rep.pressure_converged:
            raise SystemExit(f"{name}: pressure solve did not converge")
        log(f"[bench] {name} step {k + 1}/{args.steps}: {rep.pressure_iterations} iterations, "
            f"{rep.pressure_refinements} refinements, fallback {rep.pressure_fallback}, "
            f"warm start {rep.pressure_warm_start} (ratio {rep.pressure_warm_ratio:.3f}), "
            f"div {rep.max_divergence:.3e}")
    elapsed = time.perf_counter() - t0
    timed = s.kernel_stats()
    s.close()
    if net is not None:
        net.close()
    if dom is None and timed:
        dom = max((k for k in timed if not k.startswith(groups)), key=lambda k: timed[k]["total_ms"],
                  default=None)
    step_ms = elapsed * 1e3 / args.steps
    return {
        "steps_per_s": args.steps / elapsed,
        "ms_per_step": step_ms,
        "cg_iters_per_step": sum(iters) / len(iters),
        "cg_iters_per_step_all": iters,
        "refinements_per_step": sum(refs) / len(refs),
        "fallback_steps": sum(1 for f in fbs if f),
        "warm_start_kept": sum(1 for w in warm if w == 1),
        "warm_start_rejected": sum(1 for w in warm if w == -1),
        "dominant_kernel": dom,
        "dominant_timed": timed.get(dom),
        # per-kernel breakdown of the last warmup step (all kernels event-timed)
        "warmup_breakdown": {k: {"launches": v["launches"], "ms": v["ms"],
                                 "share": v["total_ms"] / sum(x["total_ms"] for x in
                                                              breakdown.values()),
                                 "GBps": v["GBps"]}
                             for k, v in sorted(breakdown.items(),
                                                key=lambda kv: -kv[1]["total_ms"])},
        "max_divergence": rep.max_divergence,
        "dtype": "f32+f64" if precision == fa.PRECISION_MIXED else "f64",
    }


def batch2d_configs(fa, count=64, seed=1, precision=None):
    """64 2D lid-driven cavities, n = round(64 * 8^u) with u ~ U[0, 1) (log-
    uniform 64-512), Re ~ U[100, 1000] (both from numpy's PCG64, seed 1), the
    reference's adaptive dt rule (target CFL 0.5, diffusion number 0.25) and
    its solver defaults (first-order upwind, Jacobi-PCG from the previous
    pressure to max|r| < 1e-5, 1200 iterations at most)."""
    import numpy as np

    import flursys_amd.incompressible as I

    rng = np.random.default_rng(seed)
    u = rng.random(count)
    re = 100.0 + 900.0 * rng.random(count)
    ts = I.TimeStepSettings(adaptive=True, target_cfl=0.5, target_diffusion_number=0.25)
    cfgs = [I.SimulationConfig.cavity(int(round(64 * 8.0 ** a)), reynolds=float(r), time_step=ts)
            for a, r in zip(u, re)]
    if precision is not None:
        for c in cfgs:
            c.precision = precision
    return cfgs


def run_batch2d(fa, ctx, args):
    import numpy as np

    import flursys_amd.incompressible as I

    prec = fa.PRECISION_MIXED if args.batch_precision == "mixed" else fa.PRECISION_FP64
    cfgs = batch2d_configs(fa, precision=prec)
    cells = sum(c.nx * c.ny for c in cfgs)
    log(f"[bench] batch2d: {len(cfgs)} cavities, {cells} cells, n "
        f"{min(c.nx for c in cfgs)}-{max(c.nx for c in cfgs)}")
    out = {}
    chosen, table = None, []
    # the graph-captured run is the measurement; a shorter eager run (the
    # same launches issued one by one) is timed beside it
    for mode, graphs, total in (("graphs", True, args.batch_steps),
                                ("eager", False, min(args.batch_steps, 50))):
        b = I.CavityBatch(cfgs, ctx=ctx)
        if chosen is None and args.batch_launch:
            th, bl = (int(v) for v in args.batch_launch.split("x"))
            b.set_launch(th, bl)
            chosen = dict(b.launch(), source="--batch-launch")
        elif chosen is None:
            chosen, table = b.autotune(args.tune_iterations)
            log(f"[bench] batch2d: autotune over (threads, blocks), solves capped at "
                f"{args.tune_iterations} iterations: chosen {chosen}")
            for t in table:
                log(f"[bench]   {t}")
        else:
            b.set_launch(chosen["threads"], chosen["blocks"])
        b.run(5, use_graphs=graphs)  # warmup (captures the graphs)
        chunk, left, ms, pms, iters, cell_its = 25, total, [], [], 0, 0
        t0 = time.perf_counter()
        while left > 0:
            k = min(chunk, left)
            reps, step_ms, pcg_ms = b.run(k, use_graphs=graphs, step_times=True, pcg_times=True)
            ms.extend(step_ms.tolist())
            pms.extend(pcg_ms.tolist())
            iters += sum(r.pressure_iterations for r in reps)
            cell_its += sum(r.pressure_iterations * c.nx * c.ny for r, c in zip(reps, cfgs))
            left -= k
            log(f"[bench] batch2d {mode}: {total - left}/{total} steps, "
                f"{iters / (len(cfgs) * (total - left)):.1f} PCG iterations per cavity-step, "
                f"{np.mean(ms):.2f} ms/step")
        elapsed = time.perf_counter() - t0
        gmode = b.graph_mode()
        b.close()
        out[mode] = {"elapsed_s": elapsed, "steps": total, "step_ms": ms, "pcg_ms": pms,
                     "pressure_iterations": iters, "cell_iterations": cell_its,
                     "graph_mode": {0: "none", 1: "whole step (cooperative PCG kernel node)",
                                    2: "split (graphs around a cooperative launch)"}[gmode]}
    g = out["graphs"]
    ms = np.array(g["step_ms"])
    steps = args.batch_steps
    # roofline of the persistent PCG from the eager run, where each launch is
    # timed with events on its stream: algorithmic bytes per cell-iteration
    # (phase A reads flags, z, d and writes d, q; phase B reads flags, x or
    # e, d, r, q and writes x or e, r, z) times the cell-iterations, over the
    # launches' time
    e = out["eager"]
    wb = 4 if args.batch_precision == "mixed" else 8
    bytes_per = (1 + 4 * wb) + (1 + 7 * wb)
    pcg_s = sum(t for t in e["pcg_ms"] if t > 0) / 1e3
    achieved = e["cell_iterations"] * bytes_per / pcg_s / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                "frac": achieved / 8000.0, "traffic": None,
                "kernel": "k_pcg (persistent, cooperative)",
                "bytes_per_cell_iteration": bytes_per,
                "pcg_ms_per_step": pcg_s * 1e3 / e["steps"]}
    cpu = None if args.no_cpu else batch2d_cpu_baseline(cfgs)
    return {
        "metric": "aggregate cavity-steps/s, 64 uneven 2D cavities (64^2-512^2) advanced together",
        "value": len(cfgs) * steps / g["elapsed_s"],
        "unit": "cavity-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 5,
        "ms_per_step": g["elapsed_s"] * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 CG in f64 refinement" if args.batch_precision == "mixed" else "f64",
        "data": "synthetic: cavities started from rest (no dataset)",
        "config": {
            "workload": f"{len(cfgs)} 2D lid-driven cavities, n = round(64*8^u), Re U[100,1000] "
                        "(seed 1), adaptive dt (CFL 0.5, diffusion 0.25), upwind, Jacobi-"
                        + ("CG cycles (FP32) in FP64 refinement" if args.batch_precision == "mixed"
                           else "PCG (FP64)") + ", previous pressure, true max|r| < 1e-5",
            "cells": cells,
            "step_latency_ms": {"p50": float(np.percentile(ms, 50)),
                                "p99": float(np.percentile(ms, 99)),
                                "max": float(ms.max())},
            "pcg_iterations_per_cavity_step": g["pressure_iterations"] / (len(cfgs) * steps),
            "graph_mode": g["graph_mode"],
            "autotuned_pcg_launch": chosen,
            "autotune_solve_ms": {f"{t['threads']}x{t['blocks']}": round(t["ms"], 3)
                                  for t in table},
            "eager_cavity_steps_per_s": len(cfgs) * out["eager"]["steps"] / out["eager"]["elapsed_s"],
            "eager_steps": out["eager"]["steps"],
            "eager_step_latency_ms_p50": float(np.percentile(out["eager"]["step_ms"], 50)),
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }


def batch2d_cpu_baseline(cfgs, stride=2):
    """The CPU oracle (OpenMP, FP64 Jacobi-PCG) on a bounded sample: the first
    step from rest of every `stride`-th cavity, as cavity-steps/s."""
    import oracle as O

    threads = min(16, os.cpu_count() or 1)
    sample = cfgs[::stride]
    t0 = time.perf_counter()
    for c in sample:
        ts = c.time_step
        s = O.Solver2D(c.nx, c.ny, nu=c.kinematic_viscosity, dt=c.dt,
                       pressure_max_iters=c.pressure_max_iters,
                       pressure_tolerance=c.pressure_tolerance, threads=threads,
                       adaptive=ts.adaptive, min_dt=ts.min_dt, max_dt=ts.max_dt,
                       target_cfl=ts.target_cfl, target_diffusion=ts.target_diffusion_number,
                       max_growth=ts.max_growth_factor, bc_u=c.bc_u)
        s.step()
    dt = time.perf_counter() - t0
    log(f"[bench] batch2d CPU oracle: {len(sample)} cavity-steps in {dt:.1f} s")
    return {"value": len(sample) / dt, "unit": "cavity-steps/s", "cores": threads, "kind": "port",
            "sample": f"first step from rest of {len(sample)} of the 64 cavities (every "
                      f"{stride}nd), FP64 Jacobi-PCG, {threads} OpenMP threads"}


def unet_flops(n):
    """Algorithmic FLOPs of one U-Net inference at n^3: 2 * 27 * cin * cout
    per output voxel and layer (the first layer's input and the head's
    output have one channel that counts)."""
    from flursys_amd.unet import layer_table

    table = layer_table()
    total = 0
    for i, (c1, c2, cin, cout, kpad, level, relu, off) in enumerate(table):
        vox = (n >> level) ** 3
        total += 2 * 27 * (1 if i == 0 else cin) * (1 if i == len(table) - 1 else cout) * vox
    return total


def run_c2(fa, ctx, args):
    """The 3D cavity at 128^3, Re 400, dt 5e-4, rel-res 1e-8 from a zero guess,
    `args.c2_steps` steps from rest in both precisions, against the FP64 oracle's
    run of the same configuration (tests/golden/cavity128_re400.npz, 20 steps)."""
    import numpy as np

    g = np.load(os.path.join(ROOT, "tests", "golden", "cavity128_re400.npz"))
    n, steps = int(g["n"]), int(g["steps"])

    def sample(f):
        return np.concatenate([f[0::8, 0::8, 0::8].ravel(), f[5::8, 5::8, 5::8].ravel()])

    out = {"workload": f"3D cavity {n}^3, Re 400, dt 5e-4, rel-res 1e-8, zero guess, "
                       f"{steps} steps from rest",
           "oracle_sample": "FP64 CPU oracle fields after the same steps on two stride-8 "
                            "lattices and the full plane next to the lid"}
    for name, prec in (("fp64_pcg", fa.PRECISION_FP64), ("mixed_ir", fa.PRECISION_MIXED)):
        cfg = fa.LidDrivenCavity3DConfig(nx=n, ny=n, nz=n, reynolds=400.0, dt=5e-4,
                                         pressure_tolerance=1e-8, stop_rule=fa.STOP_REL_L2,
                                         initial_guess=fa.GUESS_ZERO, pressure_max_iters=50000,
                                         precision=prec)
        s = fa.LidDrivenCavity3DSolver(cfg, ctx)
        its, divs, dl2 = [], [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            rep = s.advance_one_step()
            its.append(rep.pressure_iterations)
            divs.append(rep.max_divergence)
            dl2.append(rep.divergence_l2)
        el = time.perf_counter() - t0
        f = s.fields()
        s.close()
        err = {}
        for k in ("u", "v", "w", "p"):
            scale = 1.0 if k != "p" else float(g["p_max"])
            err[k] = float(np.abs(sample(f[k]) - g[f"{k}_sample"]).max()) / scale
        err["u_lid_plane"] = float(np.abs(f["u"][:, n - 1, :] - g["u_lidplane"]).max())
        err["p_lid_plane"] = float(np.abs(f["p"][:, n - 1, :] - g["p_lidplane"]).max()) / float(g["p_max"])
        out[name] = {"steps_per_s": steps / el, "cg_iters_per_step": sum(its) / len(its),
                     "max_error_vs_oracle": err,
                     "error_scales": "u, v, w: lid velocity 1; p: max|p| of the oracle",
                     "max_divergence": max(divs), "divergence_l2_last": dl2[-1],
                     "oracle_cg_iters_per_step": float(np.mean(g["iterations"]))}
        log(f"[bench] C2 {name}: {out[name]['steps_per_s']:.1f} steps/s, "
            f"{out[name]['cg_iters_per_step']:.0f} its/step, errors {err}")
    return out


def run_c1(fa, ctx, args):
    """2D cavity 128^2, Re 100, first-order upwind, Jacobi-PCG to rel-res 1e-8
    from a zero guess.  dt = 0.25 h as written breaks the explicit diffusion
    limit (viscous number 0.64 > 0.5) and diverges; the reference's adaptive
    time step rule (target CFL 0.5, diffusion number 0.25) is run instead."""
    import numpy as np

    import flursys_amd.incompressible as I
    import oracle as O

    n, re = 128, 100.0
    ts = I.TimeStepSettings(adaptive=True, target_cfl=0.5, target_diffusion_number=0.25)
    cfg = I.SimulationConfig.cavity(n, reynolds=re, dt=0.25 / n, time_step=ts,
                                    stop_rule=fa.STOP_REL_L2, initial_guess=fa.GUESS_ZERO,
                                    pressure_tolerance=1e-8, pressure_max_iters=5000)
    steps = args.c1_steps
    out = {"workload": f"2D cavity {n}^2, Re 100, upwind, rel-res 1e-8 from zero, adaptive dt "
                       "(the written dt = 0.25 h is unstable: viscous number 0.64)",
           "steps": steps}
    b = I.CavityBatch([cfg], ctx=ctx)
    b.run(3, use_graphs=True)
    t0 = time.perf_counter()
    reps, ms = b.run(steps, use_graphs=True, step_times=True)
    el = time.perf_counter() - t0
    its = reps[0].pressure_iterations / steps
    out["gpu"] = {"steps_per_s": steps / el, "cg_iters_per_step": its,
                  "step_ms_p50": float(np.percentile(ms, 50)), "dt_last": None}
    # the persistent PCG's own launch time (eager steps, HIP events)
    reps, ms, pms = b.run(50, use_graphs=False, step_times=True, pcg_times=True)
    pcg_s = float(np.sum(pms[pms > 0])) / 1e3
    cell_its = reps[0].pressure_iterations * n * n
    out["gpu"]["k_pcg_GBps"] = cell_its * 90.0 / pcg_s / 1e9
    out["gpu"]["k_pcg_bytes_per_cell_iteration"] = 90
    out["gpu"]["k_pcg_ms_per_step"] = pcg_s * 1e3 / 50
    b.close()
    log(f"[bench] C1 GPU: {out['gpu']['steps_per_s']:.1f} steps/s, {its:.0f} its/step")
    if not args.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        for threads in (1, cores):
            s = O.Solver2D(n, nu=1.0 / re, dt=0.25 / n, adaptive=True, target_cfl=0.5,
                           target_diffusion=0.25, stop_rule=O.STOP_REL_L2,
                           initial_guess=O.GUESS_ZERO, pressure_tolerance=1e-8,
                           pressure_max_iters=5000, threads=threads)
            k, cit = 0, 0
            t0 = time.perf_counter()
            while time.perf_counter() - t0 < args.c1_cpu_seconds:
                d, rep = s.step()
                cit += rep.iterations
                k += 1
            el = time.perf_counter() - t0
            out[f"cpu_oracle_{threads}_threads"] = {
                "steps_per_s": k / el, "cg_iters_per_step": cit / k, "steps_timed": k,
                "kind": "port (C oracle, OpenMP FP64 restatement of the reference path)"}
            log(f"[bench] C1 CPU oracle {threads} threads: {k / el:.2f} steps/s")
    return out


def run_c3(fa, ctx, args):
    """Unstructured pressure system: the pinned two-point Poisson matrix of a
    1000 x 1000 jittered triangulation (2,000,000 cells), CG and Jacobi-PCG
    to rel-res 1e-10, BiCGSTAB on the matrix plus a +-0.1 antisymmetric
    perturbation.  The written jitter +-0.3 h makes the reference's mesh
    builder reject the mesh (inconsistent face orientation); +-0.25 h is
    used."""
    import numpy as np

    import flursys_amd.linear as FL
    from flursys_amd import meshgen as M
    import oracle as O

    m = args.c3_m
    t0 = time.perf_counter()
    A, area, faces = M.poisson_system(m, 0.25, 0, 0)
    rhs = M.zero_mean_source(area, 0, 0)
    B = M.skew_perturbation(A, faces, 0.1, 0, 0)
    n, nnz = A.shape[0], A.nnz
    log(f"[bench] C3 system {n} cells, nnz {nnz} built in {time.perf_counter() - t0:.1f} s")
    spmv = nnz * 12 + (n + 1) * 4
    bytes_iter = {"cg": spmv + 11 * n * 8, "pcg": spmv + 13 * n * 8,
                  "bicgstab": 2 * (B.nnz * 12 + (n + 1) * 4) + 18 * n * 8}
    out = {"workload": f"{m}x{m} jittered (+-0.25 h) triangulation, {n} cells, nnz {nnz}, "
                       "rel-res 1e-10 from zero",
           "bytes_per_iteration": "SpMV nnz*12 + (n+1)*4, plus 8 B per vector pass "
                                  "(CG 11, Jacobi-PCG 13, BiCGSTAB 18 passes and 2 SpMVs)"}
    opts = FL.LinearSolverOptions(0.0, 1e-10, 50000)
    for kind, M_ in (("cg", A), ("pcg", A), ("bicgstab", B)):
        for mixed in (False, True):
            dev = FL.CsrMatrix(n, n, M_.indptr, M_.indices, M_.data, ctx)
            if mixed:
                dev.set_precision(fa.PRECISION_MIXED)
            x = np.zeros(n)
            getattr(FL, kind)(dev, rhs, x, FL.LinearSolverOptions(0.0, 1e-10, 5))  # warm-up
            x = np.zeros(n)
            rep = getattr(FL, kind)(dev, rhs, x, opts)
            dev.close()
            true = float(np.linalg.norm(rhs - M_ @ x) / np.linalg.norm(rhs))
            s = rep.device_ms / 1e3
            key = kind + ("_mixed" if mixed else "")
            out[key] = {"s_per_solve": s, "iterations": rep.iterations, "converged": rep.converged(),
                        "true_rel_residual": true}
            if mixed:
                # FP32 inner iterations (4-byte values and vectors) plus one FP64
                # residual per round; GB/s is not comparable, time is
                out[key].update({"refinements": rep.refinements, "fallback": rep.fallback,
                                 "speedup_vs_fp64": out[kind]["s_per_solve"] / s if s > 0 else None})
                log(f"[bench] C3 {key}: {rep.iterations} its in {rep.refinements} rounds, {s:.3f} s, "
                    f"true rel res {true:.2e}")
            else:
                out[key]["GBps"] = rep.iterations * bytes_iter[kind] / s / 1e9 if s > 0 else None
                out[key]["frac_of_8TBps"] = (out[key]["GBps"] or 0.0) / HBM_PEAK_GBS
                log(f"[bench] C3 {kind}: {rep.iterations} its, {s:.3f} s, {out[key]['GBps']:.0f} GB/s")
    if not args.no_cpu:
        Ao = O.Csr(n, n, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data, validate=False)
        k = 30
        t0 = time.perf_counter()
        O.cg(Ao, rhs, np.zeros(n), 0.0, 1e-30, k)
        per = (time.perf_counter() - t0) / k
        out["cpu_oracle_cg"] = {"s_per_iteration": per,
                                "s_per_solve_extrapolated": per * out["cg"]["iterations"],
                                "cores": 1, "kind": "port (C oracle, serial like src/linear/)",
                                "sample": f"{k} CG iterations on the same matrix"}
    return out


def run_f1(fa, ctx, args):
    """Collocated unstructured SIMPLE (solve_simple, src/simple.rs:148-311)
    on the device: a lid-driven cavity of side n on an n x n jittered mesh
    of quads and triangle pairs, Re 20 (nu = n / 20), the reference's
    SimpleOptions::steady relaxation (0.7 / 0.3), momentum BiCGSTAB and
    pressure-correction PCG to atol 1e-10 / rtol 1e-8, a fixed number of
    outer iterations (continuity tolerance 0).  The CPU oracle (serial, like
    the reference's src/linear/) runs one outer iteration of the smaller
    mesh."""
    import numpy as np

    from flursys_amd.linear import LinearSolverOptions
    from flursys_amd.mesh import DeviceMesh2D
    from flursys_amd.meshgen import quad_triangle_grid
    from flursys_amd.simple import DeviceSimple2D, SimpleOptions

    out = {"workload": "lid-driven cavity, jittered quads + triangle pairs, Re 20, relaxation "
                       "0.7/0.3, solves to atol 1e-10 rtol 1e-8, fixed outer iterations"}

    def setup(n):
        pts, cells = quad_triangle_grid(n, 0.2, 9)
        t0 = time.perf_counter()
        dm = DeviceMesh2D(pts, cells, ctx)
        g = dm.geometry()
        bnd = np.nonzero(g["neighbour"] < 0)[0]
        vel = {int(f): ((1.0, 0.0) if abs(g["fcy"][f] - n) < 1e-9 else (0.0, 0.0)) for f in bnd}
        s = DeviceSimple2D(dm, vel, {int(f): "ZeroGradient" for f in bnd})
        return pts, cells, dm, s, vel, bnd, time.perf_counter() - t0

    def opts(n, iters):
        return SimpleOptions(n / 20.0, max_outer_iterations=iters, continuity_absolute_tolerance=0.0,
                             momentum_solver=LinearSolverOptions(1e-10, 1e-8, 2000),
                             pressure_solver=LinearSolverOptions(1e-10, 1e-8, 20000))

    for n, iters in ((args.f1_cpu_n, 3), (args.f1_n, 5)):
        pts, cells, dm, s, vel, bnd, t_setup = setup(n)
        s.solve(opts(n, 1))  # warm-up (allocates the solver vectors)
        s.set_state(np.zeros((dm.ncells, 2)), np.zeros(dm.ncells), np.zeros(dm.nfaces))
        t0 = time.perf_counter()
        rep = s.solve(opts(n, iters))
        dt = time.perf_counter() - t0
        out[f"{n}x{n}"] = {"cells": dm.ncells, "faces": dm.nfaces, "outer_iterations": iters,
                           "s_per_outer_iteration": dt / iters,
                           "mesh_and_setup_s": t_setup,
                           "last_pressure_pcg_iterations": rep.pressure_report.iterations,
                           "last_momentum_bicgstab_iterations":
                               [rep.momentum_reports[0].iterations,
                                rep.momentum_reports[1].iterations],
                           "final_continuity_rms": rep.final_continuity_rms}
        log(f"[bench] f1 SIMPLE {dm.ncells} cells: {dt / iters * 1e3:.1f} ms per outer iteration "
            f"(PCG {rep.pressure_report.iterations} its)")
        s.close()
        dm.close()
        if n == args.f1_cpu_n and not args.no_cpu:
            import oracle as O

            om = O.Mesh2D(pts, cells)
            st = O.LsqStencil(om, {int(f): "ZeroGradient" for f in bnd})
            t0 = time.perf_counter()
            O.solve_simple(om, st, vel, np.zeros((om.m.ncells, 2)), np.zeros(om.m.ncells),
                           np.zeros(om.m.nfaces), n / 20.0, max_outer_iterations=1,
                           continuity_tol=0.0, momentum_solver=(1e-10, 1e-8, 2000),
                           pressure_solver=(1e-10, 1e-8, 20000))
            cpu = time.perf_counter() - t0
            out["cpu_oracle"] = {"s_per_outer_iteration": cpu, "cores": 1, "kind": "port",
                                 "sample": f"the first outer iteration of the {n}x{n} mesh"}
            out["gpu_vs_cpu_first_iteration_mesh"] = f"{n}x{n}"
            log(f"[bench] f1 CPU oracle {om.m.ncells} cells: {cpu:.2f} s for the first outer iteration")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="grid cells per axis")
    ap.add_argument("--reynolds", type=float, default=1000.0)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--inner-tol", type=float, default=0.0)
    ap.add_argument("--batch-precision", choices=("fp64", "mixed"), default="mixed",
                    help="batch pressure solves: FP64 PCG, or FP32 CG in FP64 refinement")
    ap.add_argument("--batch-launch", default="",
                    help="batch2d: THREADSxBLOCKS for the persistent PCG instead of autotuning")
    ap.add_argument("--tune-iterations", type=int, default=1300,
                    help="iteration cap of the batch autotuner's solves")
    ap.add_argument("--variants",
                    default="plain_fp64_pcg,mixed_ir,mg_fcg,learned_fcg,learned_fcg_replace")
    ap.add_argument("--weights", default=os.path.join(ROOT, "flursys_amd", "weights",
                                                      "unet64_cavity.npy"),
                    help="trained U-Net parameters for learned_fcg")
    ap.add_argument("--unet-n", type=int, default=64, help="U-Net resolution of the weights")
    ap.add_argument("--cpu-sample-iters", type=int, default=1500,
                    help="PCG iterations of the CPU baseline sample (about 15 s at 256^3)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-copy-probe", action="store_true")
    ap.add_argument("--config", default="cavity3d", choices=["cavity3d", "batch2d", "csr"],
                    help="cavity3d: the headline 256^3 benchmark; batch2d: 64 uneven 2D "
                         "cavities advanced together; csr: the 2M-cell CSR CG / PCG / "
                         "BiCGSTAB record alone (each a separate JSON line)")
    ap.add_argument("--batch-steps", type=int, default=1000,
                    help="steps of every cavity in the batch2d configuration")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the supporting configurations (C1 2D 128^2, C2 3D 128^3, C3 CSR, "
                         "f1 unstructured SIMPLE)")
    ap.add_argument("--c1-steps", type=int, default=300)
    ap.add_argument("--c1-cpu-seconds", type=float, default=4.0)
    ap.add_argument("--c3-m", type=int, default=1000, help="C3 triangulation: m x m squares")
    ap.add_argument("--f1-n", type=int, default=600, help="unstructured SIMPLE mesh: n x n")
    ap.add_argument("--f1-cpu-n", type=int, default=200,
                    help="unstructured SIMPLE mesh timed on the CPU oracle too")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the cavity would shard "
                          "as z-slabs with one halo plane per neighbour per iteration, which "
                          "is designed but not built; run with --gpus 1"}))
        sys.exit(2)

    import flursys_amd as fa
    from flursys_amd import _lib

    ctx = fa.Context(0)
    if args.config == "batch2d":
        print(json.dumps(run_batch2d(fa, ctx, args)), flush=True)
        ctx.close()
        return
    if args.config == "csr":
        print(json.dumps(run_c3(fa, ctx, args)), flush=True)
        ctx.close()
        return
    # learned_fcg: the network's coarse prediction corrected by the V-cycle
    # below its level; learned_fcg_replace: the network and two sweeps in
    # place of the coarse levels
    modes = {"plain_fp64_pcg": (fa.PRECISION_FP64, fa.PRECOND_JACOBI, None),
             "mixed_ir": (fa.PRECISION_MIXED, fa.PRECOND_JACOBI, None),
             "mg_fcg": (fa.PRECISION_MIXED, fa.PRECOND_MULTIGRID, None),
             "learned_fcg": (fa.PRECISION_MIXED, fa.PRECOND_LEARNED, fa.LEARNED_CORRECTED),
             "learned_fcg_replace": (fa.PRECISION_MIXED, fa.PRECOND_LEARNED, fa.LEARNED_REPLACE)}
    results = {}
    for name in args.variants.split(","):
        if name.startswith("learned_fcg") and not os.path.exists(args.weights):
            log(f"[bench] {name} skipped: no trained weights at {args.weights} "
                "(python -m flursys_amd.train_unet)")
            continue
        precision, precond, lmode = modes[name]
        results[name] = run_variant(fa, ctx, name, precision, precond, args, lmode)
    # the headline is the learned flexible CG (the metric's method); the
    # other variants are reported beside it
    best = "learned_fcg" if "learned_fcg" in results else max(
        results, key=lambda k: results[k]["steps_per_s"])
    b = results[best]
    dom = b["dominant_kernel"]
    dk = b["dominant_timed"] or {}
    copy = None
    if not args.no_copy_probe:
        import ctypes as C

        ms = C.c_double()
        nbytes = 1 << 30
        _lib.check(_lib.load().fl_probe_copy(ctx.handle, nbytes, 20, C.byref(ms)), ctx.handle)
        copy = 2 * nbytes / (ms.value * 1e-3) / 1e9
    n = args.n
    out = {
        "metric": "steps/s + CG iters/step, 256^3 cavity: learned FCG vs plain PCG @ rel-res 1e-8",
        "value": b["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": b["dtype"],
        "data": "synthetic: lid-driven cavity started from rest (no dataset)",
        "config": {
            "workload": f"3D lid-driven cavity {n}^3, Re {args.reynolds:g}, dt {args.dt:g}, "
                        f"pressure to rel-res {args.tol:g} (FP64 test) from a zero guess, "
                        "p(0,0,0) pinned; value = learned flexible CG (the U-Net's 64^3 coarse "
                        "prediction corrected by the V-cycle below it, FP32 inner solve in FP64 "
                        "refinement)",
            "grid": [n, n, n],
            "variant": best,
            "cg_iters_per_step": b["cg_iters_per_step"],
            "variants": results,
            "copy_probe_GBps": copy,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": dk.get("GBps"),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dk["GBps"] / HBM_PEAK_GBS if dk.get("GBps") else None,
            "traffic": None,
            "algorithmic_bytes_per_launch": dk.get("bytes_per_launch"),
            "ms_per_launch": dk.get("ms"),
            "timed_launches": dk.get("launches"),
        },
        "cpu_baseline": None,
    }
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE on
    # this workload (calibrated against the copy probe), when profiled
    tpath = os.path.join(ROOT, "profiles", f"traffic_cavity{n}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tr = json.load(f)
        # the profiled workload names the grid, Re, dt and stop test; the
        # headline workload string continues with the variant's description
        if (out["config"]["workload"].startswith(tr.get("workload", "\0"))
                and dom in tr.get("kernels", {})):
            out["roofline"]["traffic"] = tr["kernels"][dom]["traffic_bytes"]
            out["roofline"]["traffic_source"] = tr["source"]
    # every variant's steps/s and CG iterations/step as top-level scalars
    for name, r in results.items():
        out[f"{name}_steps_per_s"] = r["steps_per_s"]
        out[f"{name}_cg_iters_per_step"] = r["cg_iters_per_step"]
    if "plain_fp64_pcg" in results:
        out["vs_plain_fp64_pcg"] = b["steps_per_s"] / results["plain_fp64_pcg"]["steps_per_s"]
        if "mg_fcg" in results:
            out["mg_fcg_vs_plain_fp64_pcg"] = (results["mg_fcg"]["steps_per_s"] /
                                               results["plain_fp64_pcg"]["steps_per_s"])
    lr = results.get("learned_fcg")
    if lr and "unet_coarse_solve" in lr["warmup_breakdown"]:
        u = lr["warmup_breakdown"]["unet_coarse_solve"]
        fl = unet_flops(args.unet_n)
        out["unet"] = {"resolution": args.unet_n, "flops_per_inference": fl,
                       "ms_per_coarse_solve": u["ms"], "launches_per_step": u["launches"],
                       "mfma_TFLOPs": fl / (u["ms"] * 1e-3) / 1e12,
                       "mfma_peak_TFLOPs": 2500.0,
                       "mfma_frac": fl / (u["ms"] * 1e-3) / 2.5e15,
                       "note": "time of the coarse solve = inference (about 20 kernels) plus the "
                               "residual f - A x on 64^3; FLOPs count the useful convolution "
                               "work (2*27*cin*cout per voxel and layer)"}
    if not args.no_extras:
        out["c2"] = run_c2(fa, ctx, args)
        out["c1"] = run_c1(fa, ctx, args)
        out["c3"] = run_c3(fa, ctx, args)
        out["f1"] = run_f1(fa, ctx, args)
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle) sample")
        plain = results.get("plain_fp64_pcg", b)
        out["cpu_baseline"] = cpu_baseline(n, args.reynolds, args.dt,
                                           plain["cg_iters_per_step"], args.cpu_sample_iters)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
