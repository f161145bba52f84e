"""Host-side mirror of LidDrivenCavity3DSolver over the C ABI.

Same names, knobs and error behaviour as src/solver/incompressible_3d.rs:
``LidDrivenCavity3DConfig`` (:8-47, defaults 32^3, dt 5e-4, tol 1e-6),
``LidDrivenCavity3DSolver.new/run`` (:117-162) and ``advance_one_step``
(:164-172).  The fields stay resident on the GPU; numpy copies are made only
on request.  Errors raise ``FlursysError`` carrying the reference's messages.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (GUESS_PREVIOUS, PRECISION_FP64, PRECOND_JACOBI, STOP_ABS_INF, Cavity3DConfig,
                   Cavity3DStepReport, SolveReport, check, load, ptr)


@dataclass
class LidDrivenCavity3DConfig:
    """LidDrivenCavity3DConfig (incompressible_3d.rs:8-47) plus the GPU knobs."""

    nx: int = 32
    ny: int = 32
    nz: int = 32
    length: float = 1.0
    height: float = 1.0
    depth: float = 1.0
    density: float = 1.0
    lid_velocity: float = 1.0
    reynolds: float = 100.0
    dt: float = 5.0e-4
    max_steps: int = 10_000
    pressure_max_iters: int = 2_000
    pressure_tolerance: float = 1.0e-6
    stop_rule: int = STOP_ABS_INF
    initial_guess: int = GUESS_PREVIOUS
    precision: int = PRECISION_FP64
    preconditioner: int = PRECOND_JACOBI
    inner_tolerance: float = 0.0
    pressure_solver: int = 0  # SOLVER_PCG; SOLVER_SOR is the reference's solver
    pressure_omega: float = 1.7

    def to_c(self) -> Cavity3DConfig:
        return Cavity3DConfig(self.nx, self.ny, self.nz, self.length, self.height, self.depth,
                              self.density, self.lid_velocity, self.reynolds, self.dt,
                              self.pressure_max_iters, self.pressure_tolerance, self.stop_rule,
                              self.initial_guess, self.precision,
                              self.preconditioner, self.inner_tolerance,
                              self.pressure_solver, 0, self.pressure_omega)

    def viscosity(self) -> float:
        """viscosity (incompressible_3d.rs:86-88)."""
        return self.lid_velocity * self.length / self.reynolds


@dataclass
class RunSummary3D:
    """RunSummary3D (incompressible_3d.rs:91-99) plus the pressure figures."""

    steps: int
    final_time: float
    max_divergence: float
    pressure_residual: float
    max_speed: float
    divergence_l2: float
    pressure_iterations: int


class LidDrivenCavity3DSolver:
    """GPU LidDrivenCavity3DSolver; the whole step runs on device."""

    def __init__(self, cfg: LidDrivenCavity3DConfig, ctx: _lib.Context | None = None):
        self.cfg = cfg
        self.ctx = ctx or _lib.default_context()
        lib = load()
        h = C.c_void_p()
        c = cfg.to_c()
        check(lib.fl_cavity3d_create(self.ctx.handle, C.byref(c), C.byref(h)), self.ctx.handle)
        self.handle = h
        self.last_report = Cavity3DStepReport()

    def close(self):
        if getattr(self, "handle", None):
            load().fl_cavity3d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shapes(self):
        nx, ny, nz = self.cfg.nx, self.cfg.ny, self.cfg.nz
        return {"u": (nz, ny, nx + 1), "v": (nz, ny + 1, nx), "w": (nz + 1, ny, nx),
                "p": (nz, ny, nx)}

    def advance_one_step(self) -> Cavity3DStepReport:
        rep = Cavity3DStepReport()
        check(load().fl_cavity3d_step(self.handle, C.byref(rep)), self.ctx.handle)
        self.last_report = rep
        return rep

    def run(self, steps: int | None = None) -> RunSummary3D:
        """run (incompressible_3d.rs:136-162) without the VTK write."""
        steps = self.cfg.max_steps if steps is None else steps
        rep = Cavity3DStepReport()
        check(load().fl_cavity3d_run(self.handle, steps, C.byref(rep)), self.ctx.handle)
        self.last_report = rep
        return RunSummary3D(rep.step, rep.time, rep.max_divergence, rep.pressure_residual,
                            rep.max_speed, rep.divergence_l2, rep.pressure_iterations)

    def pressure_solve(self, rhs: np.ndarray, p: np.ndarray) -> SolveReport:
        """solve_pressure_poisson (incompressible_3d.rs:293-332) on host arrays;
        `p` is the initial guess (GUESS_PREVIOUS) and receives the solution."""
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        if not (p.dtype == np.float64 and p.flags["C_CONTIGUOUS"]):
            raise ValueError("p must be a C-contiguous float64 array")
        n = self.cfg.nx * self.cfg.ny * self.cfg.nz
        if rhs.size != n or p.size != n:
            raise ValueError(f"expected {n} cells")
        rep = SolveReport()
        check(load().fl_cavity3d_pressure_solve(self.handle, ptr(rhs), ptr(p), C.byref(rep)),
              self.ctx.handle)
        return rep

    def set_unet(self, net) -> None:
        """Attach a ``flursys_amd.unet.UNet`` (or None) as the warm start of
        the learned preconditioner; the solver keeps a reference to it."""
        check(load().fl_cavity3d_set_unet(self.handle, net.handle if net is not None else None),
              self.ctx.handle)
        self._unet = net

    def set_learned_mode(self, mode: int) -> None:
        """_lib.LEARNED_CORRECTED (default: the V-cycle corrects the network's
        coarse prediction) or _lib.LEARNED_REPLACE (the prediction and two
        damped-Jacobi sweeps replace the coarse levels)."""
        check(load().fl_cavity3d_set_learned_mode(self.handle, mode), self.ctx.handle)

    def coarse_pair(self, n: int):
        """(b, p, diag): the last step's right-hand side b = -rhs and pressure,
        mean-restricted to n^3, as float32 (training pairs of the U-Net)."""
        b = np.empty(n ** 3, np.float32)
        p = np.empty(n ** 3, np.float32)
        diag = C.c_double()
        check(load().fl_cavity3d_coarse_pair(self.handle, n, ptr(b), ptr(p), C.byref(diag)),
              self.ctx.handle)
        return b, p, diag.value

    def capture_coarse(self, n: int, iteration: int = 0) -> None:
        """Arm (n > 0) or stop (0) the capture of the right-hand side V-cycle
        `iteration` of each pressure solve's first inner FCG solve receives on
        the level of n^3 cells."""
        self._capture_n = n
        check(load().fl_cavity3d_capture_coarse(self.handle, n, iteration), self.ctx.handle)

    def captured_rhs(self):
        """(f, diag): the last captured coarse right-hand side (float32, n^3)
        and the rediscretised Laplacian diagonal at that resolution."""
        n = self._capture_n
        f = np.empty(n ** 3, np.float32)
        diag = C.c_double()
        check(load().fl_cavity3d_captured_rhs(self.handle, ptr(f), C.byref(diag)), self.ctx.handle)
        return f, diag.value

    def fields(self):
        s = self.shapes
        out = {k: np.empty(v, np.float64) for k, v in s.items()}
        check(load().fl_cavity3d_get_fields(self.handle, ptr(out["u"]), ptr(out["v"]),
                                            ptr(out["w"]), ptr(out["p"])), self.ctx.handle)
        return out

    def set_fields(self, u=None, v=None, w=None, p=None):
        arrs = []
        for name, a in (("u", u), ("v", v), ("w", w), ("p", p)):
            if a is None:
                arrs.append(None)
                continue
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.size != int(np.prod(self.shapes[name])):
                raise ValueError(f"{name} has the wrong size")
            arrs.append(a)
        check(load().fl_cavity3d_set_fields(self.handle, *[ptr(a) for a in arrs]),
              self.ctx.handle)

    def intermediate(self):
        s = self.shapes
        out = {"us": np.empty(s["u"]), "vs": np.empty(s["v"]), "ws": np.empty(s["w"]),
               "rhs": np.empty(s["p"])}
        check(load().fl_cavity3d_get_intermediate(self.handle, ptr(out["us"]), ptr(out["vs"]),
                                                  ptr(out["ws"]), ptr(out["rhs"])),
              self.ctx.handle)
        return out

    def set_timing(self, enable: bool, kernel: str | None = None):
        """Time every kernel (or only `kernel`) with HIP events from now on;
        resets the statistics."""
        name = kernel.encode() if kernel else None
        check(load().fl_cavity3d_set_timing(self.handle, int(enable), name), self.ctx.handle)

    def kernel_stats(self) -> dict:
        """{kernel name: {launches, total_ms, bytes, ms, GBps}} of the kernels
        timed since set_timing; `ms` is the mean per launch and `GBps` the
        algorithmic HBM bytes over device time."""
        lib = load()
        out = {}
        k = 0
        st = _lib.KernelStat()
        while lib.fl_cavity3d_kernel_stats(self.handle, k, C.byref(st)) == _lib.OK:
            if st.launches:
                ms = st.total_ms / st.launches
                out[st.name.decode()] = {
                    "launches": st.launches, "total_ms": st.total_ms, "bytes": st.bytes,
                    "ms": ms, "bytes_per_launch": st.bytes / st.launches,
                    "GBps": st.bytes / (st.total_ms * 1e-3) / 1e9 if st.total_ms > 0 else None,
                }
            k += 1
        return out
