"""Writes the golden fixtures under tests/golden/ (run from the repo root:
``python tests/golden/gen_golden.py``).

Nothing here runs the reference (a Rust crate that cannot be built in this
container).  The fixtures are:

* ``neumann_dct.npz`` -- exact discrete solutions of the cell-centred
  pure-Neumann 5-point (2D) and 7-point (3D) Poisson problems, from the DCT-II
  diagonalisation of the operator (scipy.fft).  With A = -Lap_h (ghost =
  centre, src/solver/incompressible.rs:1216-1255, incompressible_3d.rs:450-474)
  the eigenvalues are lambda_k = sum_d (2 - 2 cos(pi k_d / n_d)) / h_d^2.
  - 2D zero-mean case (the reference's project_zero_mean gauge): b zero-mean,
    p = DCT^-1(DCT(b) / lambda), k = 0 removed.
  - 3D pinned case (p(0,0,0) = 0, row 0 eliminated): rows k != 0 must hold
    exactly, so the Neumann problem is solved with b_0 replaced by
    -sum_{k!=0} b_k (making it compatible) and the result shifted by -p_0.
* ``csr_known_answers.json`` -- the reference's own hand-calculated CSR tests
  (src/linear/mod.rs:97-285), as data.
* ``mesh_small.npz`` -- the 20 x 20 jittered triangulation's pinned Poisson
  system together with scipy.sparse.linalg.spsolve's solution for a seeded
  zero-mean source (a library routine, not the reference).
"""

import json
import os
import sys

import numpy as np
import scipy.fft as sfft
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def neumann_eigs(shape, spacing):
    lam = np.zeros(shape)
    for axis, (n, h) in enumerate(zip(shape, spacing)):
        k = np.arange(n)
        e = (2.0 - 2.0 * np.cos(np.pi * k / n)) / (h * h)
        sh = [1] * len(shape)
        sh[axis] = n
        lam = lam + e.reshape(sh)
    return lam


def neumann_solve(b, spacing):
    """Zero-mean solution of (-Lap_h) p = b for zero-mean b."""
    bh = sfft.dctn(b, type=2, norm="ortho")
    lam = neumann_eigs(b.shape, spacing)
    ph = np.zeros_like(bh)
    mask = lam > 0
    ph[mask] = bh[mask] / lam[mask]
    return sfft.idctn(ph, type=2, norm="ortho")


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for (ny, nx) in [(16, 16), (32, 32), (24, 40)]:
        b = rng.standard_normal((ny, nx))
        b -= b.mean()
        h = (1.0 / ny, 1.0 / nx)
        out[f"b2d_{ny}x{nx}"] = b
        out[f"p2d_{ny}x{nx}"] = neumann_solve(b, h)
    for (nz, ny, nx) in [(8, 8, 8), (16, 16, 16), (12, 10, 14)]:
        rhs = rng.standard_normal((nz, ny, nx))
        rhs[0, 0, 0] = 0.0
        h = (1.0 / nz, 1.0 / ny, 1.0 / nx)
        b = -rhs.copy()  # A p = b with A = -Lap, b = -rhs
        b[0, 0, 0] = -(b.sum() - b[0, 0, 0])
        p = neumann_solve(b, h)
        p -= p[0, 0, 0]
        out[f"rhs3d_{nz}x{ny}x{nx}"] = rhs
        out[f"p3d_{nz}x{ny}x{nx}"] = p
    np.savez_compressed(os.path.join(HERE, "neumann_dct.npz"), **out)

    kat = {
        "builder": {"rows": [0, 0, 1, 1, 1, 2, 2], "cols": [1, 0, 2, 1, 0, 2, 1],
                    "vals": [-1.0, 4.0, -1.0, 4.0, -1.0, 3.0, -1.0],
                    "offsets": [0, 2, 5, 7], "out_cols": [0, 1, 0, 1, 2, 1, 2],
                    "out_vals": [4.0, -1.0, -1.0, 4.0, -1.0, -1.0, 3.0],
                    "diagonal": [4.0, 4.0, 3.0]},
        "duplicates": {"rows": [0, 0, 0, 1, 1], "cols": [0, 0, 1, 0, 1],
                       "vals": [2.0, 3.0, -1.0, 1.0, 4.0], "nnz": 4, "a00": 5.0,
                       "spmv": [[[1.0, 2.0], [3.0, 9.0]], [[3.0, -1.0], [16.0, -1.0]]],
                       "residual": {"x": [1.0, 2.0], "rhs": [5.0, 10.0], "r": [2.0, 1.0]}},
        "spd": {"a": [[4.0, 1.0], [1.0, 3.0]], "x": [1.0 / 11.0, 7.0 / 11.0]},
        "nonsymmetric": {"a": [[4.0, 1.0, 0.0], [2.0, 3.0, 1.0], [0.0, 1.0, 2.0]],
                         "x": [1.0, -2.0, 3.0]},
        "poisson1d": {"n": 32},
        "options": {"atol": 1.0e-12, "rtol": 1.0e-12, "max_iter": 1000},
    }
    with open(os.path.join(HERE, "csr_known_answers.json"), "w") as f:
        json.dump(kat, f, indent=1)

    sys.path.insert(0, ROOT)
    import oracle as O

    mesh = O.mesh_poisson(20, 0.3, 0, 0)
    n = len(mesh["volumes"])
    src = np.random.default_rng(0).standard_normal(n)
    src -= (src * mesh["volumes"]).sum() / mesh["volumes"].sum()
    rhs = src * mesh["volumes"]
    rhs[0] = 0.0
    A = sp.csr_matrix((mesh["vals"], mesh["cols"], mesh["offsets"]), shape=(n, n))
    x = spla.spsolve(A.tocsc(), rhs)
    np.savez_compressed(os.path.join(HERE, "mesh_small.npz"), offsets=mesh["offsets"],
                        cols=mesh["cols"], vals=mesh["vals"], volumes=mesh["volumes"],
                        rhs=rhs, x=x)


if __name__ == "__main__":
    main()
