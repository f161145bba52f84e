"""Synthetic unstructured pressure systems for the CSR path (host side).

The CSR benchmark configuration is the pinned two-point Poisson matrix of a
jittered structured triangulation of the unit square: m x m squares, each
split into T0 = (v00, v10, v11) and T1 = (v00, v11, v01) (cell 2 (a + m b)
+ t); interior vertices jittered by U(-jitter h, +jitter h) in x then y,
vertices visited row by row, from a splitmix64 stream.  The matrix is what
assemble_poisson builds for it (src/poisson.rs:146-263: a_f = S_f.d / |d|^2
with Gamma = 1, ZeroGradient boundaries) with cell `ref` pinned to 0
(apply_reference, :303-337), in CSR with FP64 values.

This module builds that system vectorised in numpy (the 2,000,000-cell case
in a few seconds), as input for the GPU solvers; the tests check it against
the CPU oracle's own assembly.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64_uniform(seed: int, count: int) -> np.ndarray:
    """The first `count` values of the splitmix64 stream seeded with `seed`,
    as doubles in [0, 1) from the top 53 bits."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + _GOLDEN * np.arange(1, count + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def triangulation(m: int, jitter: float, seed: int):
    """(points [(m+1)^2, 2], triangles [2 m^2, 3])."""
    h = 1.0 / m
    b, a = np.meshgrid(np.arange(m + 1), np.arange(m + 1), indexing="ij")
    x, y = a * h, b * h
    inner = (a > 0) & (a < m) & (b > 0) & (b < m)
    u = splitmix64_uniform(seed, 2 * int(inner.sum()))
    x = x.astype(np.float64)
    y = y.astype(np.float64)
    x[inner] += (2.0 * u[0::2] - 1.0) * jitter * h
    y[inner] += (2.0 * u[1::2] - 1.0) * jitter * h
    pts = np.stack([x.ravel(), y.ravel()], 1)
    bb, aa = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    v00 = (aa + (m + 1) * bb).ravel()
    v10, v11, v01 = v00 + 1, v00 + m + 2, v00 + m + 1
    tri = np.empty((2 * m * m, 3), np.int64)
    tri[0::2] = np.stack([v00, v10, v11], 1)
    tri[1::2] = np.stack([v00, v11, v01], 1)
    return pts, tri


def interior_faces(m: int):
    """(owner, neighbour, vertex a, vertex b) of every interior edge."""
    bb, aa = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    sq = aa + m * bb
    v00 = aa + (m + 1) * bb
    # the diagonal of every square: T0 | T1
    own = [2 * sq.ravel()]
    nei = [2 * sq.ravel() + 1]
    va, vb = [v00.ravel()], [(v00 + m + 2).ravel()]
    # horizontal edges: T1 of square (a, b - 1) | T0 of square (a, b)
    sel = bb > 0
    own.append((2 * (sq - m) + 1)[sel])
    nei.append((2 * sq)[sel])
    va.append(v00[sel])
    vb.append((v00 + 1)[sel])
    # vertical edges: T0 of square (a - 1, b) | T1 of square (a, b)
    sel = aa > 0
    own.append((2 * (sq - 1))[sel])
    nei.append((2 * sq + 1)[sel])
    va.append(v00[sel])
    vb.append((v00 + m + 1)[sel])
    return (np.concatenate(own), np.concatenate(nei), np.concatenate(va), np.concatenate(vb))


def poisson_system(m: int, jitter: float = 0.25, seed: int = 0, ref: int = 0):
    """The pinned two-point Poisson CSR matrix (scipy csr, FP64), the cell
    areas and the interior faces (owner, neighbour, coefficient)."""
    pts, tri = triangulation(m, jitter, seed)
    p = pts[tri]  # [cells, 3, 2]
    centre = p.mean(axis=1)  # a triangle's centroid is its vertex average
    area = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) -
                        (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    own, nei, va, vb = interior_faces(m)
    e = pts[vb] - pts[va]
    s = np.stack([e[:, 1], -e[:, 0]], 1)  # edge normal with |s| = edge length
    # oriented away from the owner's vertex average, as oriented_face does
    # (src/mesh.rs:541-559); a triangle's vertex average is its centroid
    mid = 0.5 * (pts[va] + pts[vb])
    s *= np.where((s * (mid - centre[own])).sum(1) < 0.0, -1.0, 1.0)[:, None]
    d = centre[nei] - centre[own]
    coef = (s * d).sum(1) / (d * d).sum(1)
    if not np.all(np.isfinite(coef)) or np.any(coef <= 0.0):
        raise ValueError("invalid non-orthogonal geometry: a two-point coefficient is not positive")
    n = 2 * m * m
    rows = np.concatenate([own, own, nei, nei])
    cols = np.concatenate([own, nei, nei, own])
    vals = np.concatenate([coef, -coef, coef, -coef])
    keep = (rows != ref) & (cols != ref)
    rows = np.concatenate([rows[keep], [ref]])
    cols = np.concatenate([cols[keep], [ref]])
    vals = np.concatenate([vals[keep], [1.0]])
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return A, area, (own, nei, coef)


def zero_mean_source(area: np.ndarray, seed: int = 0, ref: int = 0) -> np.ndarray:
    """rhs = q V with q standard normal (numpy PCG64, `seed`) projected so
    that sum(q V) = 0, and rhs[ref] = 0 (the pinned value)."""
    q = np.random.default_rng(seed).standard_normal(len(area))
    q -= (q * area).sum() / area.sum()
    rhs = q * area
    rhs[ref] = 0.0
    return rhs


def skew_perturbation(A, faces, amplitude: float = 0.1, seed: int = 0, ref: int = 0):
    """A + the antisymmetric convection-like perturbation of every interior
    face: A[P, N] += e a_f, A[N, P] -= e a_f with e ~ U(-amplitude, amplitude)
    (numpy PCG64, `seed`); faces touching the pinned cell keep their
    eliminated entries out."""
    own, nei, coef = faces
    e = np.random.default_rng(seed).uniform(-amplitude, amplitude, len(coef))
    keep = (own != ref) & (nei != ref)
    D = sp.csr_matrix(((e * coef)[keep], (own[keep], nei[keep])), shape=A.shape)
    out = (A + D - D.T).tocsr()
    out.sort_indices()
    return out


def quad_triangle_grid(n: int, jitter: float = 0.2, seed: int = 9):
    """An n x n square of side n: quads, the squares with x + y even split
    into two triangles (v00, v10, v11), (v00, v11, v01); interior vertices
    moved by U(-jitter, +jitter) in x and y (numpy generator `seed`).  Every
    cell keeps at least two face neighbours, so the least-squares stencil of
    all-ZeroGradient walls is regular.  Returns (points (N, 2), polygon list)."""
    rng = np.random.default_rng(seed)
    ys, xs = np.meshgrid(np.arange(n + 1, dtype=np.float64), np.arange(n + 1, dtype=np.float64),
                         indexing="ij")
    inner = (xs > 0) & (xs < n) & (ys > 0) & (ys < n)
    jx = rng.uniform(-jitter, jitter, xs.shape)
    jy = rng.uniform(-jitter, jitter, xs.shape)
    pts = np.stack([(xs + inner * jx).ravel(), (ys + inner * jy).ravel()], 1)
    cells = []
    for y in range(n):
        for x in range(n):
            a = x + (n + 1) * y
            b, c, d = a + 1, a + n + 2, a + n + 1
            if (x + y) % 2 == 0:
                cells += [[a, b, c], [a, c, d]]
            else:
                cells.append([a, b, c, d])
    return pts, cells
