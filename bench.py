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
  learned_fcg     mg_fcg started from the U-Net's warm start (weights from
                  flursys_amd/train_unet.py; skipped when none are present)
`value` is the steps/s of the fastest variant that meets the same FP64
rel-res 1e-8 stop test; every variant's figures are in `config.variants`.

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


def run_variant(fa, ctx, name, precision, precond, args):
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
    dom = max(breakdown, key=lambda k: breakdown[k]["total_ms"]) if breakdown else None
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
        dom = max(timed, key=lambda k: timed[k]["total_ms"])
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
    ap.add_argument("--variants", default="plain_fp64_pcg,mixed_ir,mg_fcg,learned_fcg")
    ap.add_argument("--weights", default=os.path.join(ROOT, "flursys_amd", "weights",
                                                      "unet64_cavity.npy"),
                    help="trained U-Net parameters for learned_fcg")
    ap.add_argument("--unet-n", type=int, default=64, help="U-Net resolution of the weights")
    ap.add_argument("--cpu-sample-iters", type=int, default=1500,
                    help="PCG iterations of the CPU baseline sample (about 15 s at 256^3)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-copy-probe", action="store_true")
    ap.add_argument("--config", default="cavity3d", choices=["cavity3d", "batch2d"],
                    help="cavity3d: the headline 256^3 benchmark; batch2d: 64 uneven 2D "
                         "cavities advanced together (a separate JSON line)")
    ap.add_argument("--batch-steps", type=int, default=1000,
                    help="steps of every cavity in the batch2d configuration")
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
    modes = {"plain_fp64_pcg": (fa.PRECISION_FP64, fa.PRECOND_JACOBI),
             "mixed_ir": (fa.PRECISION_MIXED, fa.PRECOND_JACOBI),
             "mg_fcg": (fa.PRECISION_MIXED, fa.PRECOND_MULTIGRID),
             "learned_fcg": (fa.PRECISION_MIXED, fa.PRECOND_LEARNED)}
    results = {}
    for name in args.variants.split(","):
        if name == "learned_fcg" and not os.path.exists(args.weights):
            log(f"[bench] learned_fcg skipped: no trained weights at {args.weights} "
                "(python -m flursys_amd.train_unet)")
            continue
        results[name] = run_variant(fa, ctx, name, *modes[name], args)
    best = max(results, key=lambda k: results[k]["steps_per_s"])
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
                        "p(0,0,0) pinned",
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
        if tr.get("workload") == out["config"]["workload"] and dom in tr.get("kernels", {}):
            out["roofline"]["traffic"] = tr["kernels"][dom]["traffic_bytes"]
            out["roofline"]["traffic_source"] = tr["source"]
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle) sample")
        plain = results.get("plain_fp64_pcg", b)
        out["cpu_baseline"] = cpu_baseline(n, args.reynolds, args.dt,
                                           plain["cg_iters_per_step"], args.cpu_sample_iters)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
