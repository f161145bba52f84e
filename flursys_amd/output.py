"""On-disk outputs of the GPU solvers in the reference's formats.

Mirrors src/output.rs and the output half of the 2D solver's run loop
(append_histories, write_snapshot, write_final_outputs,
write_case_specific_profiles, write_case_summary; src/solver/
incompressible.rs:1744-1978) and the 3D solver's final VTK
(src/solver/incompressible_3d.rs:136-162).  The fields come from the
device (cell-centred velocity, speed and vorticity formed by GPU kernels and
copied out once per snapshot); the text formatting runs in the library's C++
writers (include/flursys_hip.h, fl_write_*), on a background thread when
streaming, so the device keeps stepping while a snapshot is written.
"""

from __future__ import annotations

import ctypes as C
import math
import os
import queue
import threading
from decimal import Decimal

import numpy as np

from . import _lib
from ._lib import HistoryRow, load

CAVITY, CYLINDER, BACKWARD_STEP, CHANNEL = "LidDrivenCavity", "CylinderRe100", \
    "BackwardFacingStep", "Channel"
CASE_NAMES = {CAVITY: "Lid-Driven Cavity", CYLINDER: "Cylinder Flow Re=100",
              BACKWARD_STEP: "Backward-Facing Step Flow",
              CHANNEL: "Plane Poiseuille Channel Flow"}  # Case::name (src/cases/mod.rs:66-73)


class OutputError(RuntimeError):
    """The reference's String error of a failed write (e.g. "Cannot create ...")."""


def _call(fn, *args):
    err = C.create_string_buffer(512)
    st = fn(*args, err, 512)
    if st != _lib.OK:
        raise OutputError(err.value.decode() or _lib.STATUS_NAMES.get(st, str(st)))


def _p(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _mask(a, n):
    if a is None:
        return None
    m = np.ascontiguousarray(a, dtype=np.uint8).reshape(-1)
    if m.size != n:
        raise ValueError("solid mask must hold nx*ny cells")
    return m


def ensure_output_tree(root: str) -> None:
    """ensure_output_tree (output.rs:24-29): root and root/frames."""
    try:
        os.makedirs(root, exist_ok=True)
    except OSError as e:
        raise OutputError(f"Cannot create {root}: {e}") from e
    try:
        os.makedirs(os.path.join(root, "frames"), exist_ok=True)
    except OSError as e:
        raise OutputError(f"Cannot create frames directory: {e}") from e


def write_field_csv(path, nx, ny, dx, dy, solid, p, u_cell, v_cell, vorticity, temperature=None):
    """write_field_csv (output.rs:31-67)."""
    a = [_f64(x) for x in (p, u_cell, v_cell, vorticity, temperature)]
    _call(load().fl_write_field_csv, os.fsencode(path), nx, ny, dx, dy, _p(_mask(solid, nx * ny)),
          *[_p(x) for x in a])


def write_legacy_vtk(path, title, nx, ny, dx, dy, solid, p, u_cell, v_cell, vorticity,
                     temperature=None):
    """write_legacy_vtk (output.rs:69-125)."""
    a = [_f64(x) for x in (p, u_cell, v_cell, vorticity, temperature)]
    _call(load().fl_write_vtk_2d, os.fsencode(path), title.encode(), nx, ny, dx, dy,
          _p(_mask(solid, nx * ny)), *[_p(x) for x in a])


def write_legacy_vtk_3d(path, title, nx, ny, nz, dx, dy, dz, p, u_cell, v_cell, w_cell):
    """write_legacy_vtk_3d (output.rs:127-183), cell-centred velocity given."""
    a = [_f64(x) for x in (p, u_cell, v_cell, w_cell)]
    _call(load().fl_write_vtk_3d, os.fsencode(path), title.encode(), nx, ny, nz, dx, dy, dz,
          *[_p(x) for x in a])


def write_ppm_frame(path, nx, ny, solid, field, symmetric):
    """write_ppm_frame (output.rs:201-255)."""
    _call(load().fl_write_ppm, os.fsencode(path), nx, ny, _p(_mask(solid, nx * ny)),
          _p(_f64(field)), int(bool(symmetric)))


def write_columns(path, header, columns, fmt):
    """A CSV of equal-length columns, column c formatted by fmt[c] ('f' {:.12}, 'e' {:.12e})."""
    cols = [_f64(c) for c in columns]
    arr = (C.c_void_p * len(cols))(*[_p(c) for c in cols])
    _call(load().fl_write_csv_columns, os.fsencode(path), header.encode(), len(cols),
          len(cols[0]) if cols else 0, arr, fmt.encode())


def append_history(path, row: HistoryRow, extended=True):
    """append_history (output.rs:185-199) of one append_histories row."""
    _call(load().fl_append_history, os.fsencode(path), C.byref(row), int(bool(extended)))


def rust_display(x: float) -> str:
    """Rust's `{}` for f64: the shortest round-trip digits, positional (no
    exponent), "-0", "NaN", "inf" -- the format of the case summary."""
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    t = format(Decimal(repr(float(x))), "f")
    if "." in t:
        t = t.rstrip("0").rstrip(".")
    return t


# ------------------------------------------------------------ diagnostics
def cylinder_force_coefficients(p, u_cell, v_cell, solid, dx, dy, mu, rho, u_inf, diameter):
    """cylinder_force_coefficients (incompressible.rs:1634-1676): pressure and
    wall-shear forces on the solid faces, divided by 1/2 rho U^2 D.  Arrays
    are (ny, nx); the sums run in the reference's cell order."""
    ny, nx = p.shape
    s = np.asarray(solid, dtype=bool).reshape(ny, nx)
    fx = fy = 0.0
    for j in range(ny):
        for i in range(nx):
            if s[j, i]:
                continue
            pc, uc, vc = p[j, i], u_cell[j, i], v_cell[j, i]
            if i + 1 < nx and s[j, i + 1]:
                fx += pc * dy
                fy += mu * vc / (0.5 * dx) * dy
            if i > 0 and s[j, i - 1]:
                fx -= pc * dy
                fy += mu * vc / (0.5 * dx) * dy
            if j + 1 < ny and s[j + 1, i]:
                fy += pc * dx
                fx += mu * uc / (0.5 * dy) * dx
            if j > 0 and s[j - 1, i]:
                fy -= pc * dx
                fx += mu * uc / (0.5 * dy) * dx
    q = 0.5 * rho * u_inf * u_inf * diameter
    return (fx / q, fy / q) if q > 0.0 else (0.0, 0.0)


def reattachment_length_ratio(u_cell, solid, dx, dy, step_x, step_height, mu):
    """reattachment_length_ratio (incompressible.rs:1678-1705): the first
    sign change of the bottom-wall shear behind the step, linearly
    interpolated, as (x_r - step_x) / h; NaN when there is none."""
    ny, nx = u_cell.shape
    s = np.asarray(solid, dtype=bool).reshape(ny, nx) if solid is not None else np.zeros((ny, nx), bool)
    start = min(int(math.floor(step_x / dx)), nx - 2)
    prev_tau, prev_x = None, 0.0
    for i in range(start + 1, nx - 1):
        if s[0, i]:
            continue
        tau = mu * u_cell[0, i] / (0.5 * dy)
        x = (i + 0.5) * dx
        if prev_tau is not None and prev_tau <= 0.0 and tau > 0.0:
            fraction = min(max(-prev_tau / (tau - prev_tau), 0.0), 1.0)
            xr = prev_x + fraction * (x - prev_x)
            return (xr - step_x) / step_height
        prev_tau, prev_x = tau, x
    return float("nan")


# ------------------------------------------------------------- streaming
class SnapshotStreamer:
    """Runs output writes on one background thread, in submission order.
    The device copy-out happens on the caller's thread (it is the solver's
    stream); only the host formatting and file writes are deferred.  The
    first failure is raised by the next submit() or by close()."""

    def __init__(self, max_pending=4):
        self._q: queue.Queue = queue.Queue(max_pending)
        self._error: BaseException | None = None
        self._t = threading.Thread(target=self._run, daemon=True)
        self._t.start()

    def _run(self):
        while True:
            job = self._q.get()
            if job is None:
                return
            if self._error is None:
                try:
                    job()
                except BaseException as e:  # reported on the caller's thread
                    self._error = e

    def _raise(self):
        if self._error is not None:
            e, self._error = self._error, None
            raise e

    def submit(self, job):
        self._raise()
        self._q.put(job)

    def close(self):
        if self._t.is_alive():
            self._q.put(None)
            self._t.join()
        self._raise()


class _Sync:
    def submit(self, job):
        job()

    def close(self):
        pass


class Outputs2D:
    """The 2D solver's outputs at the reference's cadence: a history row
    every `output_every` steps, speed and vorticity frames every
    `frame_every` steps, case_summary.txt at the start and field.csv,
    field.vtk and the case's profile file at the end.  `case` is one of
    CAVITY, CYLINDER, BACKWARD_STEP, CHANNEL; `cylinder` = (mu, rho, u_inf,
    diameter) and `step` = (step_x, step_height, mu) feed the Cd/Cl and
    reattachment columns.  With `extended` the history also carries
    cg_iters, hbm_GBps, roofline_frac and fallback_count."""

    def __init__(self, solver, output_dir, case=CAVITY, output_every=100, frame_every=500,
                 solid=None, cylinder=None, step=None, extended=True, stream=True):
        self.s = solver
        self.dir = output_dir
        self.case = case
        self.output_every, self.frame_every = output_every, frame_every
        cfg = solver.cfg
        self.nx, self.ny = cfg.nx, cfg.ny
        self.dx, self.dy = cfg.length / cfg.nx, cfg.height / cfg.ny
        self.solid = None if solid is None else np.asarray(solid, np.uint8).reshape(self.ny, self.nx)
        self.cylinder, self.step_geom = cylinder, step
        self.extended = extended
        self.frame_index = 0
        self.fallbacks = 0
        self.w = SnapshotStreamer() if stream else _Sync()
        ensure_output_tree(output_dir)
        self.write_case_summary()
        self.snapshot()

    # the cell fields from the device
    def _cells(self):
        f = self.s.field_update()
        return f["pressure"], f["velocity_x"], f["velocity_y"], f["speed"], f["vorticity"]

    def snapshot(self):
        """write_snapshot (incompressible.rs:1769-1789)."""
        _, _, _, speed, vort = self._cells()
        k = self.frame_index
        self.frame_index += 1
        frames = os.path.join(self.dir, "frames")
        nx, ny, solid = self.nx, self.ny, self.solid

        def job():
            write_ppm_frame(os.path.join(frames, f"speed_{k:05d}.ppm"), nx, ny, solid, speed, False)
            write_ppm_frame(os.path.join(frames, f"vorticity_{k:05d}.ppm"), nx, ny, solid, vort, True)
        self.w.submit(job)

    def after_step(self, rep, cg_iters=None, hbm_GBps=float("nan"), roofline_frac=float("nan"),
                   fallback=False):
        """The run loop's cadence (incompressible.rs:439-448) after one step."""
        self.fallbacks += int(bool(fallback))
        if rep.iteration % self.output_every == 0:
            cd = cl = xr = 0.0
            if self.case in (CYLINDER, BACKWARD_STEP):
                p, uc, vc, _, _ = self._cells()
                if self.case == CYLINDER and self.cylinder:
                    mu, rho, uinf, diam = self.cylinder
                    cd, cl = cylinder_force_coefficients(p, uc, vc, self._solid_or_zero(),
                                                         self.dx, self.dy, mu, rho, uinf, diam)
                if self.case == BACKWARD_STEP and self.step_geom:
                    sx, sh, mu = self.step_geom
                    xr = reattachment_length_ratio(uc, self.solid, self.dx, self.dy, sx, sh, mu)
            row = HistoryRow(rep.iteration, rep.time, rep.pressure_residual, rep.pressure_iterations,
                             rep.continuity_residual, rep.momentum_residual, rep.max_speed, cd, cl,
                             xr, rep.momentum_cfl, rep.viscous_diffusion_number, rep.time_step,
                             rep.pressure_iterations if cg_iters is None else cg_iters, hbm_GBps,
                             roofline_frac, self.fallbacks)
            path, ext = os.path.join(self.dir, "history.csv"), self.extended
            self.w.submit(lambda: append_history(path, row, ext))
        if rep.iteration % self.frame_every == 0:
            self.snapshot()

    def _solid_or_zero(self):
        return self.solid if self.solid is not None else np.zeros((self.ny, self.nx), np.uint8)

    def finish(self):
        """write_final_outputs (incompressible.rs:1791-1819) and the case's
        profile (write_case_specific_profiles :1821-1913); waits for every
        queued write."""
        p, uc, vc, _, vort = self._cells()
        nx, ny, dx, dy, solid, d = self.nx, self.ny, self.dx, self.dy, self.solid, self.dir
        title = CASE_NAMES[self.case]
        case, cyl, stp = self.case, self.cylinder, self.step_geom
        s = self._solid_or_zero().astype(bool)

        def job():
            write_field_csv(os.path.join(d, "field.csv"), nx, ny, dx, dy, solid, p, uc, vc, vort)
            write_legacy_vtk(os.path.join(d, "field.vtk"), title, nx, ny, dx, dy, solid, p, uc, vc,
                             vort)
            xs = (np.arange(nx) + 0.5) * dx
            ys = (np.arange(ny) + 0.5) * dy
            if case == CAVITY:
                write_columns(os.path.join(d, "centerline_u_vertical.csv"), "y,u",
                              [ys, uc[:, nx // 2]], "fe")
                write_columns(os.path.join(d, "centerline_v_horizontal.csv"), "x,v",
                              [xs, vc[ny // 2, :]], "fe")
            elif case == CYLINDER:
                keep = ~s[ny // 2, :]
                j = ny // 2
                write_columns(os.path.join(d, "wake_centerline.csv"), "x,u,v,p",
                              [xs[keep], uc[j, keep], vc[j, keep], p[j, keep]], "feee")
            elif case == BACKWARD_STEP:
                mu = stp[2] if stp else 0.0
                keep = ~s[0, :]
                tau = mu * uc[0, :] / (0.5 * dy)
                write_columns(os.path.join(d, "bottom_wall_shear.csv"), "x,tau_wall",
                              [xs[keep], tau[keep]], "fe")
            elif case == CHANNEL:
                i = nx - 1
                write_columns(os.path.join(d, "outlet_velocity_profile.csv"), "y,u,v,p",
                              [ys, uc[:, i], vc[:, i], p[:, i]], "feee")
        self.w.submit(job)
        self.w.close()

    def write_case_summary(self):
        """write_case_summary (incompressible.rs:1915-1978) for the projection
        coupling with the configuration's values; the host runs one thread."""
        cfg = self.s.cfg
        ts = cfg.time_step
        conv = "FirstOrderUpwind" if cfg.convection == 0 else "Central"
        solver = "Sor" if cfg.pressure_solver == _lib.SOLVER_SOR else "Pcg"
        nu = cfg.kinematic_viscosity
        lines = [f"Case: {CASE_NAMES[self.case]}",
                 "Method: finite volume, staggered grid, Projection",
                 f"Grid: {cfg.nx} x {cfg.ny}",
                 f"Domain: {rust_display(cfg.length)} x {rust_display(cfg.height)}",
                 f"dx: {rust_display(self.dx)}", f"dy: {rust_display(self.dy)}",
                 f"rho: {rust_display(cfg.density)}", f"nu: {rust_display(nu)}",
                 f"dt: {rust_display(cfg.dt)}"]
        if ts.adaptive:
            lines.append(
                "time_stepping: adaptive, initial_dt={}, min_dt={}, max_dt={}, target_cfl={}, "
                "target_diffusion={}, max_growth={}".format(*map(rust_display, (
                    cfg.dt, ts.min_dt, ts.max_dt, ts.target_cfl, ts.target_diffusion_number,
                    ts.max_growth_factor))))
        else:
            lines.append(f"time_stepping: fixed, dt={rust_display(cfg.dt)}")
        lines += [f"t_end: {rust_display(cfg.t_end)}", f"max_steps: {getattr(cfg, 'max_steps', 0)}",
                  "cpu_threads: 1", f"convection: {conv}", "coupling: Projection",
                  f"pressure_solver: {solver}", "velocity_relaxation: 0.7",
                  "pressure_relaxation: 0.3", "energy_model: Off", "buoyancy_model: Off"]
        path = os.path.join(self.dir, "case_summary.txt")
        try:
            with open(path, "w") as f:
                f.write("\n".join(lines) + "\n")
        except OSError as e:
            raise OutputError(f"Cannot write case summary: {e}") from e


def write_cavity3d_vtk(solver, path, title="FLURSYS 3D lid-driven cavity"):
    """The 3D run's final output (incompressible_3d.rs:145-153): pressure and
    the cell-centred velocity, averaged from the faces on the device."""
    cfg = solver.cfg
    n = (cfg.nz, cfg.ny, cfg.nx)
    p, uc, vc, wc = (np.empty(n) for _ in range(4))
    _lib.check(load().fl_cavity3d_cell_fields(solver.handle, _p(p), _p(uc), _p(vc), _p(wc)),
               solver.ctx.handle)
    write_legacy_vtk_3d(path, title, cfg.nx, cfg.ny, cfg.nz, cfg.length / cfg.nx,
                        cfg.height / cfg.ny, cfg.depth / cfg.nz, p, uc, vc, wc)
