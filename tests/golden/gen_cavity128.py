"""Writes ``cavity128_re400.npz``: the FP64 oracle's 3D lid-driven cavity at
128^3 (unit cube, lid 1, Re 400, dt 5e-4, pinned Jacobi-PCG to rel-res 1e-8
from a zero guess) after 20 projection steps from rest.  Run from the repo
root: ``python tests/golden/gen_cavity128.py`` (about 4 minutes on 8 cores).

The oracle takes ~10 s per step at this size, too long to run beside every
GPU test, so the fixture keeps what the GPU parity test compares against:

* per step: PCG iterations, max divergence (incompressible_3d.rs:504-517)
  and max speed (:519-532);
* after the last step: every field on a stride-8 lattice in each axis at
  two offsets (0 and 5, so boundary-adjacent and interior cells are both
  sampled), the full plane next to the lid (j = ny - 1) of u and p where the
  flow is strongest, and per-field max |f| and ||f||_2.
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
N, STEPS = 128, 20


def sample(f):
    """Stride-8 lattice samples at offsets 0 and 5 (f indexed [k, j, i])."""
    return np.concatenate([f[0::8, 0::8, 0::8].ravel(), f[5::8, 5::8, 5::8].ravel()])


def main():
    sys.path.insert(0, ROOT)
    import oracle as O

    c = O.Cavity3D(N, reynolds=400.0, dt=5e-4, pressure_tolerance=1e-8,
                   stop_rule=O.STOP_REL_L2, initial_guess=O.GUESS_ZERO,
                   pressure_max_iters=20000)
    iters, div, speed = [], [], []
    t0 = time.time()
    for k in range(STEPS):
        rep = c.step()
        iters.append(rep.iterations)
        div.append(c.max_divergence())
        speed.append(c.max_speed())
        print(f"step {k + 1}: {rep.iterations} iterations, max div {div[-1]:.3e} "
              f"({time.time() - t0:.0f} s)", flush=True)
    out = dict(n=N, steps=STEPS, reynolds=400.0, dt=5e-4, iterations=np.array(iters),
               max_divergence=np.array(div), max_speed=np.array(speed))
    for name in ("u", "v", "w", "p"):
        f = c.field(name)
        out[f"{name}_sample"] = sample(f)
        out[f"{name}_max"] = float(np.abs(f).max())
        out[f"{name}_l2"] = float(np.linalg.norm(f))
    out["u_lidplane"] = c.field("u")[:, N - 1, :].copy()
    out["p_lidplane"] = c.field("p")[:, N - 1, :].copy()
    np.savez_compressed(os.path.join(HERE, "cavity128_re400.npz"), **out)


if __name__ == "__main__":
    main()
