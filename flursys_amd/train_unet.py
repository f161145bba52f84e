"""Trains the U-Net of the learned flexible CG (FL_PRECOND_LEARNED) on
(right-hand side, pressure) pairs that the plain solver generates.

The network is the coarse solve of the flexible CG's V-cycle on the 64^3
level (256^3 grid: level 2; 128^3 grid: level 1) and, where the grid itself
is 64^3, a warm start.  Its pairs (b, p) with A p = b on the 64^3 grid (A the
rediscretised Neumann 7-point Laplacian) are of two kinds, 2048 each:

* divergence pairs: cavity runs from rest at 64^3, 128^3 and 256^3 whose
  pressure is solved every step by the plain FP64 Jacobi-PCG (rel-res 1e-8,
  zero guess); each step's b = -rhs and p, mean-restricted to 64^3;
* coarse-solve pairs: the right-hand side the V-cycle's 64^3 level receives in
  FCG iteration j = step mod 4 of each step of multigrid-FCG cavity runs at
  128^3 and 256^3 (captured with fl_cavity3d_capture_coarse; later
  iterations' residuals are rougher than the first's, and the network sees
  them all in use), each solved on a 64^3 unit cube by the same plain FP64
  Jacobi-PCG to rel-res 1e-8.

The benchmark's flow (Re 1000) is held out of training and used for
evaluation.  Samples are augmented by one of the 48 symmetries of the cube
(the Neumann Laplacian on a cube commutes with them) and a random sign (the
solution operator is linear).

    python -m flursys_amd.train_unet --out flursys_amd/weights/unet64_cavity.npy

writes the FP32 parameters (numpy .npy, loadable with allow_pickle=False)
and a JSON record of the run beside them.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import flursys_amd as fa  # noqa: E402
from flursys_amd.unet import UNet  # noqa: E402


def log(msg):
    print(msg, flush=True)


def cavity(ctx, grid, re, dt, tol, precision, precond):
    cfg = fa.LidDrivenCavity3DConfig(nx=grid, ny=grid, nz=grid, reynolds=re, dt=dt,
                                     pressure_tolerance=tol, stop_rule=fa.STOP_REL_L2,
                                     initial_guess=fa.GUESS_ZERO, pressure_max_iters=50000,
                                     precision=precision, preconditioner=precond)
    return fa.LidDrivenCavity3DSolver(cfg, ctx)


def divergence_pairs(ctx, n_net, runs, dt, tol):
    """[(b, p, diag)] from plain FP64 Jacobi-PCG cavity runs [(grid, Re, steps)]."""
    pairs = []
    for grid, re, steps in runs:
        s = cavity(ctx, grid, re, dt, tol, fa.PRECISION_FP64, fa.PRECOND_JACOBI)
        t0 = time.perf_counter()
        for k in range(steps):
            rep = s.advance_one_step()
            if not rep.pressure_converged:
                raise RuntimeError(f"pressure solve did not converge ({grid}^3, Re {re})")
            b, p, diag = s.coarse_pair(n_net)
            pairs.append((b.reshape((n_net,) * 3), p.reshape((n_net,) * 3), diag))
            if (k + 1) % 64 == 0:
                log(f"[train] divergence pairs {grid}^3 Re {re:g}: {k + 1}/{steps} "
                    f"({time.perf_counter() - t0:.1f} s)")
        s.close()
    return pairs


def coarse_pairs(ctx, n_net, runs, dt, tol):
    """[(f, p, diag)]: the right-hand side of V-cycle (step mod 4) of each
    step's first inner solve on the n_net^3 level of multigrid-FCG cavity
    runs, solved by the plain FP64 PCG."""
    plain = cavity(ctx, n_net, 100.0, dt, tol, fa.PRECISION_FP64, fa.PRECOND_JACOBI)
    pairs = []
    for grid, re, steps in runs:
        s = cavity(ctx, grid, re, dt, tol, fa.PRECISION_MIXED, fa.PRECOND_MULTIGRID)
        t0 = time.perf_counter()
        for k in range(steps):
            s.capture_coarse(n_net, k % 4)
            rep = s.advance_one_step()
            if not rep.pressure_converged:
                raise RuntimeError(f"pressure solve did not converge ({grid}^3, Re {re})")
            f, diag = s.captured_rhs()
            f = f.astype(np.float64).reshape((n_net,) * 3)
            p = np.zeros_like(f)
            # the cavity solver solves Lap p = rhs, i.e. A p = -rhs, pinned at
            # cell 0; for a compatible f that is the Neumann solution shifted
            r = plain.pressure_solve(-f, p)
            if not r.converged:
                raise RuntimeError("plain PCG did not converge on a coarse pair")
            p -= p.mean()
            pairs.append((f.astype(np.float32), p.astype(np.float32), diag))
            if (k + 1) % 128 == 0:
                log(f"[train] coarse-solve pairs {grid}^3 Re {re:g}: {k + 1}/{steps} "
                    f"({time.perf_counter() - t0:.1f} s)")
        s.close()
    plain.close()
    return pairs


def augment(rng, b, p):
    """One random cube symmetry (axis permutation and flips) and sign."""
    perm = rng.permutation(3)
    flips = rng.integers(0, 2, 3)
    sign = 1.0 if rng.integers(0, 2) else -1.0

    def f(a):
        a = np.transpose(a, perm)
        for ax in range(3):
            if flips[ax]:
                a = a[(slice(None),) * ax + (slice(None, None, -1),)]
        return sign * a

    return f(b), f(p)


def train(net, pairs, steps, batch, lr, seed, warmup=50):
    """Adam with linear warmup and cosine decay; returns the loss history."""
    rng = np.random.default_rng(seed)
    n3 = net.n ** 3
    xb = np.empty((batch, n3), np.float32)
    tb = np.empty((batch, n3), np.float32)
    diag = pairs[0][2]
    hist = []
    t0 = time.perf_counter()
    for k in range(steps):
        for i in range(batch):
            b, p, d = pairs[rng.integers(len(pairs))]
            if abs(d - diag) > 1e-9 * diag:
                raise ValueError("all pairs must share the network's resolution")
            ab, ap = augment(rng, b, p)
            xb[i] = ab.ravel()
            tb[i] = ap.ravel()
        frac = k / max(steps - 1, 1)
        rate = lr * min(1.0, (k + 1) / warmup) * 0.5 * (1.0 + math.cos(math.pi * frac))
        hist.append(net.train_step(xb, tb, diag, rate))
        if (k + 1) % 100 == 0 or k == 0:
            log(f"[train] step {k + 1}/{steps}: loss {np.mean(hist[-100:]):.4f} "
                f"({time.perf_counter() - t0:.1f} s)")
    return hist


def typical_gain(pairs):
    """Median over pairs of rms(p - mean p) diag / rms(b): the output gain
    that makes the normalised targets O(1)."""
    g = [float(np.sqrt(np.mean((p - p.mean()) ** 2)) * d / np.sqrt(np.mean(b.astype(np.float64) ** 2)))
         for b, p, d in pairs if np.any(b)]
    return float(np.median(g))


def evaluate(net, pairs):
    """Mean relative L2 error of the zero-mean prediction on pairs."""
    errs = []
    for b, p, diag in pairs:
        pred = net.infer(b.ravel()[None], diag)[0]
        pz = p.ravel() - p.mean()
        errs.append(float(np.linalg.norm(pred - pz) / np.linalg.norm(pz)))
    return float(np.mean(errs))


def solver_check(ctx, net, grid, re, steps, dt, tol):
    """Iterations and fallbacks of the learned FCG with `net` as its coarse
    solve on a held-out flow."""
    s = cavity(ctx, grid, re, dt, tol, fa.PRECISION_MIXED, fa.PRECOND_LEARNED)
    s.set_unet(net)
    its, fbs = [], 0
    for _ in range(steps):
        rep = s.advance_one_step()
        its.append(rep.pressure_iterations)
        fbs += rep.pressure_fallback
    s.close()
    return {"grid": grid, "reynolds": re, "steps": steps, "iterations": its, "fallback_steps": fbs}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  "weights", "unet64_cavity.npy"))
    ap.add_argument("--n", type=int, default=64, help="network resolution")
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--quick", action="store_true", help="small data set (a smoke run)")
    args = ap.parse_args(argv)

    ctx = fa.Context(0)
    n = args.n
    tol = 1e-8
    if args.quick:
        div_runs = [(n, 400.0, 8), (2 * n, 400.0, 4)]
        cs_runs = [(2 * n, 400.0, 8)]
    else:
        # 2048 + 2048 pairs; Re 1000 (the benchmark's flow) held out; 256^3 at
        # Re 100 breaks the explicit diffusion limit at dt 5e-4
        div_runs = ([(n, re, 256) for re in (100.0, 400.0, 2000.0, 5000.0)] +
                    [(2 * n, re, 192) for re in (100.0, 400.0, 2000.0, 5000.0)] +
                    [(4 * n, re, 128) for re in (400.0, 2000.0)])
        cs_runs = ([(2 * n, re, 256) for re in (100.0, 400.0, 2000.0, 5000.0)] +
                   [(4 * n, re, 342) for re in (400.0, 2000.0)] + [(4 * n, 5000.0, 340)])
    t0 = time.perf_counter()
    pairs = divergence_pairs(ctx, n, div_runs, args.dt, tol)
    n_div = len(pairs)
    pairs += coarse_pairs(ctx, n, cs_runs, args.dt, tol)
    log(f"[train] {n_div} divergence + {len(pairs) - n_div} coarse-solve pairs in "
        f"{time.perf_counter() - t0:.0f} s")
    held_div = divergence_pairs(ctx, n, [(n, 1000.0, 6)], args.dt, tol)
    held_cs = coarse_pairs(ctx, n, [(4 * n, 1000.0, 6)], args.dt, tol)
    net = UNet(n, args.batch, seed=args.seed, ctx=ctx)
    net.set_gain(typical_gain(pairs))
    log(f"[train] output gain {net.gain:.4g}")
    before = (evaluate(net, held_div), evaluate(net, held_cs))
    hist = train(net, pairs, args.steps, args.batch, args.lr, args.seed)
    after = (evaluate(net, held_div), evaluate(net, held_cs))
    log(f"[train] held-out relative error: divergence pairs {before[0]:.3f} -> {after[0]:.3f}, "
        f"coarse-solve pairs {before[1]:.3f} -> {after[1]:.3f}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    net.save(args.out)
    check = {}
    if not args.quick:
        infer_net = UNet(n, 1, ctx=ctx)
        infer_net.load(args.out)
        check["trained"] = solver_check(ctx, infer_net, 4 * n, 1000.0, 4, args.dt, tol)
        infer_net.close()
        untrained = UNet(n, 1, seed=args.seed, ctx=ctx)
        untrained.set_gain(net.gain)
        check["untrained"] = solver_check(ctx, untrained, 4 * n, 1000.0, 4, args.dt, tol)
        untrained.close()
        s = cavity(ctx, 4 * n, 1000.0, args.dt, tol, fa.PRECISION_MIXED, fa.PRECOND_MULTIGRID)
        check["multigrid"] = {"iterations": [s.advance_one_step().pressure_iterations for _ in range(4)]}
        s.close()
        log(f"[train] learned FCG at {4 * n}^3 Re 1000: {check}")
    meta = {
        "network": {"resolution": n, "levels": 4, "channels": [16, 32, 48, 64],
                    "params": net.num_params, "output_gain": net.gain},
        "data": {"divergence_runs": [{"grid": g, "reynolds": r, "steps": s} for g, r, s in div_runs],
                 "coarse_solve_runs": [{"grid": g, "reynolds": r, "steps": s} for g, r, s in cs_runs],
                 "pairs": len(pairs), "divergence_pairs": n_div,
                 "coarse_solve_pairs": len(pairs) - n_div, "dt": args.dt,
                 "solver": "plain FP64 Jacobi-PCG, rel-res 1e-8, zero guess",
                 "augmentation": "48 cube symmetries x sign"},
        "training": {"steps": args.steps, "batch": args.batch, "lr": args.lr,
                     "schedule": "linear warmup 50, cosine decay", "seed": args.seed,
                     "loss_first_100": float(np.mean(hist[:100])),
                     "loss_last_100": float(np.mean(hist[-100:]))},
        "held_out": {"flow": "Re 1000 (the benchmark's)",
                     "divergence_pairs": {"grid": n, "steps": 6, "rel_error_before": before[0],
                                          "rel_error_after": after[0]},
                     "coarse_solve_pairs": {"grid": 4 * n, "steps": 6, "rel_error_before": before[1],
                                            "rel_error_after": after[1]}},
        "solver_check": check,
    }
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(f"[train] wrote {args.out}")
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
