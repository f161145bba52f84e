"""Device-fed outputs (flursys_amd/output.py over fl_grid2d_cell_fields /
fl_cavity3d_cell_fields and the fl_write_* writers):

* a 2D cavity run with the reference's cadence (history every 5 steps,
  frames every 10) streamed from a background writer: the files are the
  oracle's text of the GPU's own fields byte for byte, the streamed files
  equal the synchronous ones, and field.csv's numbers agree with the CPU
  oracle's run to the 2D parity bar (1e-9 of the velocity scale);
* the 3D VTK: cell-centred velocity averaged on the device equals the
  oracle's text from the GPU's staggered fields byte for byte.
"""

import os

import numpy as np
import pytest

import flursys_amd as fa
import flursys_amd.incompressible as I
from flursys_amd import output as W
import oracle as O
from oracle import output as OO

pytestmark = pytest.mark.gpu


def run_2d(ctx, out_dir, stream, steps=20):
    cfg = I.SimulationConfig.cavity(32, reynolds=100.0, dt=1e-3, pressure_tolerance=1e-10,
                                    stop_rule=fa.STOP_REL_L2, initial_guess=fa.GUESS_ZERO,
                                    pressure_max_iters=5000)
    s = I.IncompressibleSolver(cfg, ctx=ctx)
    out = W.Outputs2D(s, str(out_dir), W.CAVITY, output_every=5, frame_every=10, stream=stream)
    reps = []
    for _ in range(steps):
        r = s.advance()
        reps.append(r)
        out.after_step(r)
    out.finish()
    cells = s.field_update()
    s.close()
    return cfg, reps, cells


def test_streamed_2d_outputs(gpu_ctx, tmp_path):
    print("[gpu] 2D cavity outputs, streamed and synchronous", flush=True)
    cfg, reps, cells = run_2d(gpu_ctx, tmp_path / "a", True)
    run_2d(gpu_ctx, tmp_path / "b", False)
    a, b = tmp_path / "a", tmp_path / "b"
    names = sorted(f for f in os.listdir(a) if (a / f).is_file()) + \
        ["frames/" + f for f in sorted(os.listdir(a / "frames"))]
    assert "field.csv" in names and "field.vtk" in names and "history.csv" in names
    assert "centerline_u_vertical.csv" in names and "case_summary.txt" in names
    # frames at steps 0 (prepare_output), 10, 20
    assert sorted(os.listdir(a / "frames")) == [f"{k}_{i:05d}.ppm" for k in ("speed", "vorticity")
                                                for i in range(3)]
    for n in names:
        assert (a / n).read_bytes() == (b / n).read_bytes(), n
    nx = ny = 32
    h = 1.0 / 32
    p, uc, vc = cells["pressure"], cells["velocity_x"], cells["velocity_y"]
    assert (a / "field.csv").read_text() == OO.field_csv(nx, ny, h, h, None, p, uc, vc,
                                                          cells["vorticity"])
    assert (a / "frames" / "speed_00002.ppm").read_bytes() == OO.ppm(nx, ny, None,
                                                                     cells["speed"], False)
    hist = (a / "history.csv").read_text().splitlines()
    assert hist[0].startswith(OO.HISTORY_HEADER) and len(hist) == 1 + 4
    r = reps[19]
    row = OO.history_row(r.iteration, r.time, r.pressure_residual, r.pressure_iterations,
                         r.continuity_residual, r.momentum_residual, r.max_speed, 0.0, 0.0, 0.0,
                         r.momentum_cfl, r.viscous_diffusion_number, r.time_step)
    assert hist[4].startswith(row + ",")
    # the physics: field.csv against the CPU oracle's run of the same steps
    o = O.Solver2D(nx, ny, nu=cfg.kinematic_viscosity, dt=cfg.dt,
                   pressure_max_iters=cfg.pressure_max_iters,
                   pressure_tolerance=cfg.pressure_tolerance, stop_rule=O.STOP_REL_L2,
                   initial_guess=O.GUESS_ZERO, bc_u=cfg.bc_u)
    for _ in range(20):
        o.step()
    o.cell_fields()
    data = np.loadtxt(a / "field.csv", delimiter=",", skiprows=1)
    assert np.abs(data[:, 6] - o.field("ucell").ravel()).max() < 1e-9
    assert np.abs(data[:, 7] - o.field("vcell").ravel()).max() < 1e-9


def test_3d_vtk_from_device_cell_fields(gpu_ctx, tmp_path):
    print("[gpu] 3D cavity VTK", flush=True)
    cfg = fa.LidDrivenCavity3DConfig(nx=10, ny=8, nz=6, reynolds=400.0, dt=5e-4,
                                     pressure_tolerance=1e-8, stop_rule=fa.STOP_REL_L2,
                                     initial_guess=fa.GUESS_ZERO, pressure_max_iters=2000)
    s = fa.LidDrivenCavity3DSolver(cfg, gpu_ctx)
    for _ in range(3):
        s.advance_one_step()
    path = tmp_path / "field.vtk"
    W.write_cavity3d_vtk(s, str(path))
    f = s.fields()
    expect = OO.vtk_3d("FLURSYS 3D lid-driven cavity", 10, 8, 6, 0.1, 0.125, 1 / 6, f["p"],
                       f["u"], f["v"], f["w"])
    assert path.read_text() == expect
