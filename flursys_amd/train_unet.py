"""Trains the learned warm start (the U-Net of the cavity solver's learned
preconditioner) on pressure solves of the 3D lid-driven cavity.

Training pairs (b, p) come from the library's own converged solves: cavity
runs from rest on 64^3, 128^3 and 256^3 grids at several Reynolds numbers,
each step's right-hand side b = -rhs and pressure p (solved to rel-res
1e-8 by multigrid FCG) mean-restricted to the network's 64^3 grid.  The
benchmark's flow (Re 1000) is left out of the training set.  Each sample is
augmented by one of the 48 symmetries of the cube (the Neumann Laplacian on a
cube commutes with them) and a random sign (the solution operator is linear).

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


def generate_pairs(ctx, n_net, runs, dt, tol):
    """[(b, p, diag)] from cavity runs [(grid, reynolds, steps)]."""
    pairs = []
    for grid, re, steps in runs:
        cfg = fa.LidDrivenCavity3DConfig(nx=grid, ny=grid, nz=grid, reynolds=re, dt=dt,
                                         pressure_tolerance=tol, stop_rule=fa.STOP_REL_L2,
                                         initial_guess=fa.GUESS_ZERO, pressure_max_iters=20000,
                                         precision=fa.PRECISION_MIXED,
                                         preconditioner=fa.PRECOND_MULTIGRID)
        s = fa.LidDrivenCavity3DSolver(cfg, ctx)
        t0 = time.perf_counter()
        for _ in range(steps):
            rep = s.advance_one_step()
            if not rep.pressure_converged:
                raise RuntimeError(f"pressure solve did not converge ({grid}^3, Re {re})")
            b, p, diag = s.coarse_pair(n_net)
            pairs.append((b.reshape(n_net, n_net, n_net), p.reshape(n_net, n_net, n_net), diag))
        s.close()
        log(f"[train] pairs from {grid}^3 Re {re:g}: {steps} steps in "
            f"{time.perf_counter() - t0:.1f} s")
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
            if d != diag:
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


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  "weights", "unet64_cavity.npy"))
    ap.add_argument("--n", type=int, default=64, help="network resolution")
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--quick", action="store_true", help="small data set (a smoke run)")
    args = ap.parse_args(argv)

    ctx = fa.Context(0)
    res = (100.0, 400.0, 2000.0, 5000.0)  # Re 1000 (the benchmark's flow) left out
    if args.quick:
        runs = [(args.n, re, 10) for re in res[:2]]
    else:
        # 256^3 at Re 100 breaks the explicit diffusion limit at dt 5e-4
        runs = ([(args.n, re, 50) for re in res] + [(2 * args.n, re, 30) for re in res] +
                [(4 * args.n, re, 25) for re in res[1:]])
    pairs = generate_pairs(ctx, args.n, runs, args.dt, 1e-8)
    held = [(args.n, 1000.0, 6)]  # evaluation only
    test_pairs = generate_pairs(ctx, args.n, held, args.dt, 1e-8)
    net = UNet(args.n, args.batch, seed=args.seed, ctx=ctx)
    net.set_gain(typical_gain(pairs))
    log(f"[train] output gain {net.gain:.4g}")
    before = evaluate(net, test_pairs)
    hist = train(net, pairs, args.steps, args.batch, args.lr, args.seed)
    after = evaluate(net, test_pairs)
    log(f"[train] held-out relative error of the prediction: {before:.3f} -> {after:.3f}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    net.save(args.out)
    meta = {
        "network": {"resolution": args.n, "levels": 4, "channels": [16, 32, 48, 64],
                    "params": net.num_params, "output_gain": net.gain},
        "data": {"runs": [{"grid": g, "reynolds": r, "steps": s} for g, r, s in runs],
                 "pairs": len(pairs), "dt": args.dt, "solver": "multigrid FCG, rel-res 1e-8",
                 "augmentation": "48 cube symmetries x sign"},
        "training": {"steps": args.steps, "batch": args.batch, "lr": args.lr,
                     "schedule": "linear warmup 50, cosine decay", "seed": args.seed,
                     "loss_first_100": float(np.mean(hist[:100])),
                     "loss_last_100": float(np.mean(hist[-100:]))},
        "held_out": {"runs": [{"grid": g, "reynolds": r, "steps": s} for g, r, s in held],
                     "rel_error_before": before, "rel_error_after": after},
    }
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(f"[train] wrote {args.out}")
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
