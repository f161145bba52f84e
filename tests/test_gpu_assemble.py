"""Device assembly (fl_mesh2d_*, flursys_amd/mesh.py) against the oracle's
restatement of UnstructuredMesh::from_cells (src/mesh.rs:317-423) and
assemble_poisson + apply_reference (src/poisson.rs:146-263, 303-337) with
CsrBuilder::finalize (src/linear/csr.rs:159-192):

* face numbering, owner/neighbour, face and cell geometry: bitwise;
* CSR offsets, columns, values and right-hand side: bitwise (duplicates
  summed in insertion order on both sides), on triangles, quads and mixed
  polygons, with sources, FixedValue faces, gamma != 1 and a reference;
* the config-3 triangulation (1000 x 1000 squares, +-0.25 h, 2 M cells):
  the device CSR equals the oracle's bitwise, and the solve on it matches;
* MeshError / PoissonError kinds against the oracle's.
"""

import time

import numpy as np
import pytest

import flursys_amd.linear as FL
from flursys_amd import meshgen
from flursys_amd.mesh import DeviceMesh2D, MeshError, PoissonError
import oracle as O

pytestmark = pytest.mark.gpu


def quads_and_triangles(n, seed):
    """n x n squares, the even ones split into two triangles, jittered."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(0, 1, n + 1)
    X, Y = np.meshgrid(xs, xs)
    inner = (X > 0) & (X < 1) & (Y > 0) & (Y < 1)
    h = 1.0 / n
    X = X + inner * rng.uniform(-0.15 * h, 0.15 * h, X.shape)
    Y = Y + inner * rng.uniform(-0.15 * h, 0.15 * h, Y.shape)
    pts = np.stack([X.ravel(), Y.ravel()], 1)
    cells = []
    for j in range(n):
        for i in range(n):
            a = i + (n + 1) * j
            if (i + j) % 2:
                cells.append([a, a + 1, a + n + 2, a + n + 1])
            else:
                cells += [[a, a + 1, a + n + 2], [a, a + n + 2, a + n + 1]]
    return pts, cells


def compare_geometry(dm, om):
    g = dm.geometry()
    assert np.array_equal(g["owner"], om.owner) and np.array_equal(g["neighbour"], om.neighbour)
    assert np.array_equal(g["fcx"], om.face_centers[:, 0]) and np.array_equal(g["fcy"], om.face_centers[:, 1])
    assert np.array_equal(g["sx"], om.area_vectors[:, 0]) and np.array_equal(g["sy"], om.area_vectors[:, 1])
    assert np.array_equal(g["cx"], om.centers[:, 0]) and np.array_equal(g["cy"], om.centers[:, 1])
    assert np.array_equal(g["volume"], om.volumes)


def compare_system(A, rhs, Ao, rhs_o):
    off, cols, vals = A.arrays()
    assert np.array_equal(off, Ao.offsets) and np.array_equal(cols, Ao.cols)
    assert np.array_equal(vals, Ao.vals)
    assert np.array_equal(rhs, rhs_o)


def jittered_triangles(n, jitter, seed):
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(xs, xs)
    inner = (X > 0) & (X < 1) & (Y > 0) & (Y < 1)
    h = 1.0 / n
    X = X + inner * rng.uniform(-jitter * h, jitter * h, X.shape)
    Y = Y + inner * rng.uniform(-jitter * h, jitter * h, Y.shape)
    pts = np.stack([X.ravel(), Y.ravel()], 1)
    cells = []
    for j in range(n):
        for i in range(n):
            a = i + (n + 1) * j
            cells += [[a, a + 1, a + n + 2], [a, a + n + 2, a + n + 1]]
    return pts, cells


@pytest.mark.parametrize("kind", ["triangles", "mixed"])
def test_mesh_and_poisson_match_oracle_bitwise(gpu_ctx, kind):
    print(f"[gpu] device assembly {kind}", flush=True)
    pts, cells = jittered_triangles(30, 0.2, 7) if kind == "triangles" else quads_and_triangles(24, 3)
    om = O.Mesh2D(pts, cells)
    dm = DeviceMesh2D(pts, cells, gpu_ctx)
    assert dm.nfaces == om.m.nfaces and dm.ncells == om.m.ncells
    compare_geometry(dm, om)
    rng = np.random.default_rng(11)
    src = rng.standard_normal(dm.ncells)
    bnd = om.boundary_faces
    boundary = {int(f): ("FixedValue", float(v)) for f, v in zip(bnd[::3], rng.standard_normal(len(bnd[::3])))}
    for gamma, ref in ((0.7, None), (1.0, (5, 0.25))):
        A, rhs = dm.assemble_poisson(gamma, src, boundary, ref)
        Ao, rhs_o = O.assemble_poisson(om, gamma, src, boundary, ref)
        compare_system(A, rhs, Ao, rhs_o)
        # the device matrix solves like an uploaded copy of the oracle's
        x1, x2 = np.zeros(dm.ncells), np.zeros(dm.ncells)
        r1 = FL.pcg(A, rhs, x1, FL.LinearSolverOptions(0.0, 1e-10, 5000))
        B = FL.CsrMatrix(Ao.nrows, Ao.ncols, Ao.offsets, Ao.cols, Ao.vals, gpu_ctx)
        r2 = FL.pcg(B, rhs_o, x2, FL.LinearSolverOptions(0.0, 1e-10, 5000))
        assert r1.converged() and r1.iterations == r2.iterations and np.array_equal(x1, x2)
    # a pure-Neumann system with a compatible source and a reference
    src0 = src - (src * om.volumes).sum() / om.volumes.sum()
    A, rhs = dm.assemble_poisson(1.0, src0, None, (0, 0.0))
    Ao, rhs_o = O.assemble_poisson(om, 1.0, src0, None, (0, 0.0))
    compare_system(A, rhs, Ao, rhs_o)


def test_config3_triangulation_bitwise_and_timed(gpu_ctx):
    """1000 x 1000 squares jittered by +-0.25 h (seed 0): 2 M triangles."""
    print("[gpu] 2M-cell device assembly", flush=True)
    pts, tri = meshgen.triangulation(1000, 0.25, 0)
    t0 = time.perf_counter()
    dm = DeviceMesh2D(pts, tri, gpu_ctx)
    A, rhs = dm.assemble_poisson(1.0, None, None, (0, 0.0))
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    ref = O.mesh_poisson(1000, 0.25, 0, 0)
    t_cpu = time.perf_counter() - t0
    print(f"[gpu]   device mesh + assembly {t_dev:.3f} s, oracle {t_cpu:.3f} s, "
          f"{dm.ncells} cells, {A.nnz} entries", flush=True)
    off, cols, vals = A.arrays()
    assert np.array_equal(off, ref["offsets"]) and np.array_equal(cols, ref["cols"])
    assert np.array_equal(vals, ref["vals"])
    assert not rhs.any()
    assert np.array_equal(dm.geometry()["volume"], ref["volumes"])


def test_mesh_and_poisson_errors_match_oracle(gpu_ctx):
    sq = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [2, 0], [2, 1], [1, 2]], float)
    cases = [
        (sq, [[0, 1]]),                            # InvalidCell
        (sq, [[0, 1, 9]]),                         # InvalidPointIndex
        (np.array([[0, 0], [0, 0], [1, 1.0]]), [[0, 1, 2]]),   # DegenerateFace
        (sq, [[0, 1, 2], [0, 2, 1], [0, 1, 2]]),   # NonManifoldFace
        (np.array([[0, 0], [1, 0], [2, 0.0]]), [[0, 1, 2]]),   # ZeroVolume
        (np.array([[0, 0], [np.nan, 0], [1, 1]]), [[0, 1, 2]]),  # NonFinitePoint
    ]
    for pts, cells in cases:
        with pytest.raises(O.MeshError) as eo:
            O.Mesh2D(pts, cells)
        with pytest.raises(MeshError) as ed:
            DeviceMesh2D(pts, cells, gpu_ctx)
        assert ed.value.kind == eo.value.kind and str(ed.value) == str(eo.value)
    pts, cells = quads_and_triangles(4, 1)
    dm = DeviceMesh2D(pts, cells, gpu_ctx)
    om = O.Mesh2D(pts, cells)
    with pytest.raises(PoissonError) as e:
        dm.assemble_poisson(1.0, np.ones(dm.ncells), None, (0, 0.0))
    assert e.value.kind == "IncompatiblePureNeumannSource"
    with pytest.raises(O.PoissonError):
        O.assemble_poisson(om, 1.0, np.ones(dm.ncells), None, (0, 0.0))
    with pytest.raises(PoissonError) as e:
        dm.assemble_poisson(1.0, None, None, None)
    assert e.value.kind == "MissingReferenceForPureNeumann"
    with pytest.raises(PoissonError) as e:
        dm.assemble_poisson(1.0, None, None, (dm.ncells, 0.0))
    assert e.value.kind == "InvalidReferenceCell"
