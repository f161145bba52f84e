"""Device collocated SIMPLE (fl_simple2d_*, flursys_amd/simple.py) against the
oracle's restatement of solve_simple (src/simple.rs:148-311, oracle/
simple2d.c).

Bars: the reference's own hand tests of src/simple.rs reproduced through the
C ABI (two-cell cavity, imbalance removed in one iteration, leaking wall,
option errors, 3 x 3 cavities straight and skewed); on larger jittered
polygon meshes the same number of outer iterations and velocity, pressure and
face flux within 1e-9 of their scales after every outer iteration count
tested (the device sums its dot products in another order than the oracle,
so agreement is to round-off, not bitwise); the open channel anchored by a
pressure-correction boundary; the continuity history.
"""

import time

import numpy as np
import pytest

from flursys_amd.linear import LinearSolverOptions
from flursys_amd.mesh import DeviceMesh2D
from flursys_amd.simple import DeviceSimple2D, SimpleError, SimpleOptions
import oracle as O

pytestmark = pytest.mark.gpu


def grid_cells(nx, ny, skew=0.0, jitter=0.0, seed=0, triangles=False):
    rng = np.random.default_rng(seed)
    pts = []
    for y in range(ny + 1):
        for x in range(nx + 1):
            px, py = x + skew * y, float(y)
            if 0 < x < nx and 0 < y < ny and jitter:
                px += rng.uniform(-jitter, jitter)
                py += rng.uniform(-jitter, jitter)
            pts.append((px, py))
    p = lambda x, y: x + (nx + 1) * y  # noqa: E731
    cells = []
    for y in range(ny):
        for x in range(nx):
            a, b, c, d = p(x, y), p(x + 1, y), p(x + 1, y + 1), p(x, y + 1)
            if triangles and (x + y) % 2 == 0:
                cells += [[a, b, c], [a, c, d]]
            else:
                cells.append([a, b, c, d])
    return np.array(pts), cells


def boundary(om):
    return [f for f in range(om.m.nfaces) if om.neighbour[f] < 0]


def both(pts, cells, ctx):
    return O.Mesh2D(pts, cells), DeviceMesh2D(pts, cells, ctx)


def run_both(ctx, pts, cells, vel, pbc, corr=None, flux0=None, **kw):
    om, dm = both(pts, cells, ctx)
    nc, nf = om.m.ncells, om.m.nfaces
    flux0 = np.zeros(nf) if flux0 is None else flux0
    st = O.LsqStencil(om, pbc)
    okw = dict(kw)
    mom = okw.pop("momentum_solver", (1e-10, 1e-8, 1000))
    pre = okw.pop("pressure_solver", (1e-10, 1e-8, 1000))
    orep, ov, op, ophi, ohist = O.solve_simple(om, st, vel, np.zeros((nc, 2)), np.zeros(nc), flux0,
                                               okw.pop("viscosity"), correction_boundary=corr,
                                               momentum_solver=mom, pressure_solver=pre, **okw)
    s = DeviceSimple2D(dm, vel, pbc, corr)
    s.set_state(np.zeros((nc, 2)), np.zeros(nc), flux0)
    opts = SimpleOptions(kw["viscosity"], momentum_solver=LinearSolverOptions(*mom),
                         pressure_solver=LinearSolverOptions(*pre))
    for k, v in kw.items():
        if k not in ("viscosity", "momentum_solver", "pressure_solver"):
            setattr(opts, "continuity_absolute_tolerance" if k == "continuity_tol" else k, v)
    drep = s.solve(opts)
    return (orep, ov, op, ophi, ohist), drep, s.state(), s


def close(a, b, tol):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= tol * scale


def lid(om, top):
    return {f: ((1.0, 0.0) if abs(om.face_centers[f, 1] - top) < 1e-12 else (0.0, 0.0))
            for f in boundary(om)}


def test_two_cell_hand_tests_on_the_device(gpu_ctx):
    print("[gpu] SIMPLE two-cell hand tests", flush=True)
    pts = np.array([[0, 0], [1, 0], [2, 0], [0, 1], [1, 1], [2, 1]], float)
    cells = [[0, 1, 4, 3], [1, 2, 5, 4]]
    om, dm = both(pts, cells, gpu_ctx)
    walls = {f: ("FixedValue", 0.0) for f in boundary(om)}
    vel = {f: (0.0, 0.0) for f in boundary(om)}
    s = DeviceSimple2D(dm, vel, walls)
    # zero cavity stays zero (:1201-1234)
    rep = s.solve(SimpleOptions(1.0))
    st = s.state()
    assert rep.converged and rep.final_continuity_rms < 1e-12
    assert not st["velocity"].any() and not st["pressure"].any() and not st["face_flux"].any()
    # the interior imbalance is removed in one iteration (:1236-1281)
    flux = np.where(om.neighbour >= 0, 3.0, 0.0)
    s.set_state(np.zeros((2, 2)), np.zeros(2), flux)
    rep = s.solve(SimpleOptions(1.0, velocity_relaxation=1.0))
    assert rep.converged and rep.outer_iterations == 1 and rep.initial_continuity_rms == 3.0
    assert rep.final_continuity_rms < 1e-12 and np.abs(s.state()["face_flux"]).max() < 1e-12
    assert rep.momentum_reports[2].iterations == 0 and rep.momentum_reports[2].converged
    # a leaking wall cannot converge (:1283-1328)
    flux = np.zeros(om.m.nfaces)
    flux[boundary(om)[0]] = 1.0
    s.set_state(np.zeros((2, 2)), np.zeros(2), flux)
    rep = s.solve(SimpleOptions(1.0, max_outer_iterations=1, continuity_absolute_tolerance=1e-12))
    assert not rep.converged and rep.outer_iterations == 1 and rep.final_continuity_rms > 1e-12
    # option errors (:1330-1358)
    for kw, kind in (({"pressure_relaxation": 0.0}, "InvalidPressureRelaxation"),
                     ({"max_outer_iterations": 0}, "InvalidOuterIterations"),
                     ({"reference_cell": 2}, "InvalidPressureReference")):
        with pytest.raises(SimpleError) as e:
            s.solve(SimpleOptions(1.0, **kw))
        assert e.value.kind == kind


@pytest.mark.parametrize("skew,iters,tol", [(0.0, 300, 1e-9), (0.2, 500, 1e-8)])
def test_three_by_three_cavities_match_the_oracle(gpu_ctx, skew, iters, tol):
    """low_reynolds / skewed closed cavity regressions (:1406-1537)."""
    print(f"[gpu] SIMPLE 3x3 cavity skew={skew}", flush=True)
    pts, cells = grid_cells(3, 3, skew)
    om = O.Mesh2D(pts, cells)
    zg = {f: "ZeroGradient" for f in boundary(om)}
    (orep, ov, op, ophi, ohist), drep, st, _ = run_both(
        gpu_ctx, pts, cells, lid(om, 3.0), zg, viscosity=0.1, max_outer_iterations=iters,
        continuity_tol=tol, momentum_solver=(1e-12, 1e-8, 1000), pressure_solver=(1e-12, 1e-8, 1000))
    assert drep.converged and orep["converged"]
    assert abs(drep.outer_iterations - orep["outer_iterations"]) <= 1
    assert np.abs(st["face_flux"][boundary(om)]).max() < 1e-12
    assert close(st["velocity"], ov, 1e-8) and close(st["pressure"], op, 1e-7)


@pytest.mark.parametrize("triangles", [False, True])
def test_jittered_cavity_tracks_the_oracle(gpu_ctx, triangles):
    """24 x 20 jittered quads (or quads and triangles), Re 40 lid cavity:
    12 outer iterations with a zero continuity tolerance, so both run all of
    them; fields within 1e-9 of their scales, the same continuity history."""
    print(f"[gpu] SIMPLE jittered cavity triangles={triangles}", flush=True)
    pts, cells = grid_cells(24, 20, 0.05, 0.2, 5, triangles)
    om = O.Mesh2D(pts, cells)
    zg = {f: "ZeroGradient" for f in boundary(om)}
    (orep, ov, op, ophi, ohist), drep, st, _ = run_both(
        gpu_ctx, pts, cells, lid(om, 20.0), zg, viscosity=0.5, max_outer_iterations=12,
        continuity_tol=0.0, momentum_solver=(1e-13, 1e-12, 2000),
        pressure_solver=(1e-13, 1e-12, 4000))
    assert not drep.converged and drep.outer_iterations == 12
    assert close(st["velocity"], ov, 1e-9) and close(st["pressure"], op, 1e-9)
    assert close(st["face_flux"], ophi, 1e-9)
    # with tight pressure solves each iteration leaves only round-off imbalance
    assert drep.continuity_history[0] == 0.0 and len(drep.continuity_history) == 13
    assert drep.continuity_history[1:].max() < 1e-11 and ohist[1:].max() < 1e-11


def test_open_channel_with_a_correction_boundary(gpu_ctx):
    print("[gpu] SIMPLE open channel", flush=True)
    pts, cells = grid_cells(16, 6, 0.0, 0.15, 2)
    om = O.Mesh2D(pts, cells)
    left = [f for f in boundary(om) if abs(om.face_centers[f, 0]) < 1e-12]
    right = [f for f in boundary(om) if abs(om.face_centers[f, 0] - 16.0) < 1e-12]
    vel = {f: (0.0, 0.0) for f in boundary(om)}
    corr = {f: "ZeroGradient" for f in boundary(om)}
    pbc = {f: "ZeroGradient" for f in boundary(om)}
    for f in left:
        vel[f] = (1.0, 0.0)
    for f in right:
        vel[f], corr[f], pbc[f] = "ZeroGradient", ("FixedValue", 0.0), ("FixedValue", 0.0)
    wall = np.zeros(om.m.nfaces)
    for f in left:
        wall[f] = om.area_vectors[f, 0]
    flux0 = O.initial_face_flux(om, np.zeros((om.m.ncells, 2)), wall)
    # the first projection already meets any continuity tolerance the solves
    # allow, so run a fixed 15 iterations (tolerance 0) and compare fields
    (orep, ov, op, ophi, ohist), drep, st, s = run_both(
        gpu_ctx, pts, cells, vel, pbc, corr, flux0, viscosity=0.2, max_outer_iterations=15,
        continuity_tol=0.0, momentum_solver=(1e-13, 1e-12, 2000),
        pressure_solver=(1e-13, 1e-12, 4000))
    assert drep.outer_iterations == 15 and orep["outer_iterations"] == 15
    assert drep.initial_continuity_rms == pytest.approx(orep["initial_continuity_rms"], rel=1e-14)
    assert close(st["velocity"], ov, 1e-9) and close(st["pressure"], op, 1e-9)
    assert close(st["face_flux"], ophi, 1e-9)
    phi = st["face_flux"]
    assert -phi[left].sum() == pytest.approx(6.0) and phi[right].sum() == pytest.approx(6.0, rel=1e-8)
    # the device initial flux equals the oracle's
    assert np.array_equal(s.initial_face_flux(np.zeros((om.m.ncells, 2)), wall), flux0)


def test_missing_boundary_and_singular_stencil(gpu_ctx):
    pts, cells = grid_cells(3, 2)
    om, dm = both(pts, cells, gpu_ctx)
    bnd = boundary(om)
    vel = {f: (0.0, 0.0) for f in bnd}
    with pytest.raises(SimpleError) as e:
        DeviceSimple2D(dm, vel, {f: "ZeroGradient" for f in bnd[1:]})
    assert e.value.kind == "MissingBoundaryCondition" and e.value.index == bnd[0]
    one = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float)
    dm1 = DeviceMesh2D(one, [[0, 1, 2, 3]], gpu_ctx)
    with pytest.raises(SimpleError) as e:
        DeviceSimple2D(dm1, {f: (0.0, 0.0) for f in range(4)}, {f: "ZeroGradient" for f in range(4)})
    assert e.value.kind == "SingularLeastSquaresStencil"


def test_large_mesh_outer_iterations_timed(gpu_ctx):
    """A 400 x 400 quad/triangle lid cavity (240 000 cells): 3 outer
    iterations on the device, reported with their wall time; the same 3
    iterations of the oracle on a 100 x 100 cut give the CPU rate."""
    print("[gpu] SIMPLE large mesh timing", flush=True)
    pts, cells = grid_cells(400, 400, 0.0, 0.2, 9, True)
    dm = DeviceMesh2D(pts, cells, gpu_ctx)
    g = dm.geometry()
    bnd = np.nonzero(g["neighbour"] < 0)[0]
    vel = {int(f): ((1.0, 0.0) if abs(g["fcy"][f] - 400.0) < 1e-9 else (0.0, 0.0)) for f in bnd}
    s = DeviceSimple2D(dm, vel, {int(f): "ZeroGradient" for f in bnd})
    opts = SimpleOptions(20.0, max_outer_iterations=3, continuity_absolute_tolerance=0.0,
                         momentum_solver=LinearSolverOptions(1e-10, 1e-8, 2000),
                         pressure_solver=LinearSolverOptions(1e-10, 1e-8, 20000))
    t0 = time.perf_counter()
    rep = s.solve(opts)
    dt = time.perf_counter() - t0
    print(f"[gpu]   {dm.ncells} cells: 3 outer iterations in {dt:.3f} s "
          f"(pressure PCG {rep.pressure_report.iterations} its, momentum "
          f"{rep.momentum_reports[0].iterations}/{rep.momentum_reports[1].iterations})", flush=True)
    assert rep.outer_iterations == 3 and np.isfinite(rep.final_continuity_rms)
    assert rep.final_continuity_rms < rep.initial_continuity_rms or rep.initial_continuity_rms == 0.0
