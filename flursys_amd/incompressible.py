"""Host-side mirror of the 2D IncompressibleSolver (projection or SIMPLE-style
coupling) over the C ABI, plus the batched uneven-cavity driver.

Names follow src/solver/incompressible.rs: ``SimulationConfig`` (:58-90, the
flow-only knobs), ``TimeStepSettings`` (:35-56), ``ConvectionScheme``
(:14-18), ``IncompressibleSolver.advance`` (:523-555) returning a
``SolverStep`` (:196-214), ``field_update`` (:563-586).  Boundaries are
SolverBoundaryOverrides-style per-side kinds with constant velocities
(src/preprocess.rs:1362-1407); the lid-driven cavity is the default.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import (GUESS_PREVIOUS, STOP_ABS_INF, Grid2DConfig, Grid2DStepReport, SolveReport,
                   check, load, ptr)

FIRST_ORDER_UPWIND, CENTRAL = 0, 1
BC_VELOCITY, BC_SYMMETRY, BC_PRESSURE_OUTLET = 0, 1, 2
LEFT, RIGHT, BOTTOM, TOP = 0, 1, 2, 3
INLET_UNIFORM, INLET_PARABOLIC = 0, 1
SOLVER_PCG, SOLVER_SOR = _lib.SOLVER_PCG, _lib.SOLVER_SOR
# PressureVelocityCoupling (incompressible.rs:26-33)
PROJECTION, SIMPLE = 0, 1


@dataclass
class TimeStepSettings:
    """TimeStepSettings (incompressible.rs:35-56)."""

    adaptive: bool = False
    min_dt: float = 1.0e-6
    max_dt: float = 1.0e-2
    target_cfl: float = 0.5
    target_diffusion_number: float = 0.25
    max_growth_factor: float = 1.1


@dataclass
class SimulationConfig:
    """SimulationConfig (incompressible.rs:58-90), flow-only knobs.

    The default case is the lid-driven cavity (src/cases/cavity.rs): unit
    square, lid u = 1 on top, nu = U L / Re."""

    nx: int = 64
    ny: int = 64
    length: float = 1.0
    height: float = 1.0
    density: float = 1.0
    kinematic_viscosity: float = 0.01
    dt: float = 1.0e-3
    t_end: float = 1.0e30
    time_step: TimeStepSettings = field(default_factory=TimeStepSettings)
    convection: int = FIRST_ORDER_UPWIND
    pressure_max_iters: int = 1200
    pressure_tolerance: float = 1.0e-5
    stop_rule: int = STOP_ABS_INF
    initial_guess: int = GUESS_PREVIOUS
    bc_kind: tuple = (BC_VELOCITY,) * 4
    bc_u: tuple = (0.0, 0.0, 0.0, 1.0)
    bc_v: tuple = (0.0, 0.0, 0.0, 0.0)
    outlet_pressure: tuple = (0.0, 0.0, 0.0, 0.0)
    # GPU path: PRECISION_FP64 (the reference's PCG) or PRECISION_MIXED
    # (FP32 CG cycles inside FP64 iterative refinement)
    precision: int = _lib.PRECISION_FP64
    inner_tolerance: float = 0.0
    # PressureSolverKind: SOLVER_PCG (default) or SOLVER_SOR with pressure_omega
    pressure_solver: int = _lib.SOLVER_PCG
    pressure_omega: float = 1.7
    # left Velocity side: INLET_UNIFORM (bc_u) or INLET_PARABOLIC (the Channel /
    # BackwardFacingStep inlet profile, mean bc_u[LEFT], zero below
    # inlet_step_height)
    inlet_profile: int = 0
    inlet_step_height: float = 0.0
    # PROJECTION (transient) or SIMPLE (steady pseudo-steps of size dt with
    # the relaxation factors below, incompressible.rs:659-712)
    coupling: int = PROJECTION
    velocity_relaxation: float = 0.7
    pressure_relaxation: float = 0.3

    @staticmethod
    def cavity(n, reynolds=100.0, lid_velocity=1.0, **kw):
        """The lid-driven cavity (CavityCase, src/cases/cavity.rs:11-25)."""
        kw.setdefault("kinematic_viscosity", lid_velocity * 1.0 / reynolds)
        kw.setdefault("bc_u", (0.0, 0.0, 0.0, lid_velocity))
        return SimulationConfig(nx=n, ny=n, **kw)

    def to_c(self) -> Grid2DConfig:
        c = Grid2DConfig()
        c.nx, c.ny, c.length, c.height = self.nx, self.ny, self.length, self.height
        c.density, c.kinematic_viscosity = self.density, self.kinematic_viscosity
        c.dt, c.t_end = self.dt, self.t_end
        c.pressure_max_iters, c.pressure_tolerance = self.pressure_max_iters, self.pressure_tolerance
        c.convection, c.stop_rule, c.initial_guess = self.convection, self.stop_rule, self.initial_guess
        ts = self.time_step
        c.adaptive = int(ts.adaptive)
        for s in range(4):
            c.bc_kind[s] = self.bc_kind[s]
            c.bc_u[s] = self.bc_u[s]
            c.bc_v[s] = self.bc_v[s]
            c.outlet_pressure[s] = self.outlet_pressure[s]
        c.min_dt, c.max_dt, c.target_cfl = ts.min_dt, ts.max_dt, ts.target_cfl
        c.target_diffusion_number, c.max_growth_factor = (ts.target_diffusion_number,
                                                          ts.max_growth_factor)
        c.precision, c.inner_tolerance = self.precision, self.inner_tolerance
        c.pressure_solver, c.pressure_omega = self.pressure_solver, self.pressure_omega
        c.inlet_profile, c.inlet_step_height = self.inlet_profile, self.inlet_step_height
        # the C struct reads 0 as "the default"; the reference rejects 0
        for name in ("velocity_relaxation", "pressure_relaxation"):
            if getattr(self, name) == 0.0:
                raise _lib.FlursysError(_lib.ERR_CONFIG,
                                        f"{name} must lie between 0 (exclusive) and 1")
        c.coupling = self.coupling
        c.velocity_relaxation, c.pressure_relaxation = (self.velocity_relaxation,
                                                        self.pressure_relaxation)
        return c


@dataclass
class SolverStep:
    """SolverStep (incompressible.rs:196-214), flow-only fields."""

    iteration: int
    time: float
    time_step: float
    continuity_residual: float
    pressure_residual: float
    pressure_iterations: int
    momentum_residual: float
    max_speed: float
    momentum_cfl: float
    viscous_diffusion_number: float
    pressure_converged: bool
    reached_end_time: bool

    @staticmethod
    def from_c(r: Grid2DStepReport) -> "SolverStep":
        return SolverStep(r.iteration, r.time, r.time_step, r.continuity_residual,
                          r.pressure_residual, r.pressure_iterations, r.momentum_residual,
                          r.max_speed, r.momentum_cfl, r.viscous_diffusion_number,
                          bool(r.pressure_converged), bool(r.reached_end_time))


class IncompressibleSolver:
    """GPU IncompressibleSolver (projection or SIMPLE coupling; PCG or SOR
    pressure)."""

    def __init__(self, cfg: SimulationConfig, solid=None, ctx: _lib.Context | None = None):
        self.cfg = cfg
        self.ctx = ctx or _lib.default_context()
        self._solid = None
        if solid is not None:
            self._solid = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
            if self._solid.size != cfg.nx * cfg.ny:
                raise ValueError("solid mask must hold nx*ny cells")
        h = C.c_void_p()
        c = cfg.to_c()
        check(load().fl_grid2d_create(self.ctx.handle, C.byref(c), ptr(self._solid), C.byref(h)),
              self.ctx.handle)
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            load().fl_grid2d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self) -> SolverStep:
        r = Grid2DStepReport()
        check(load().fl_grid2d_step(self.handle, C.byref(r)), self.ctx.handle)
        return SolverStep.from_c(r)

    def pressure_solve(self, rhs, p):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        n = self.cfg.nx * self.cfg.ny
        if rhs.size != n or p.size != n or p.dtype != np.float64 or not p.flags["C_CONTIGUOUS"]:
            raise ValueError("rhs and p must hold nx*ny float64 cells")
        rep = SolveReport()
        check(load().fl_grid2d_pressure_solve(self.handle, ptr(rhs), ptr(p), C.byref(rep)),
              self.ctx.handle)
        return rep

    def fields(self):
        nx, ny = self.cfg.nx, self.cfg.ny
        u, v, p = np.empty((ny, nx + 1)), np.empty((ny + 1, nx)), np.empty((ny, nx))
        check(load().fl_grid2d_get_fields(self.handle, ptr(u), ptr(v), ptr(p)), self.ctx.handle)
        return {"u": u, "v": v, "p": p}

    def set_fields(self, u=None, v=None, p=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, p)]
        check(load().fl_grid2d_set_fields(self.handle, *[ptr(a) for a in arrs]), self.ctx.handle)

    def field_update(self):
        """Cell-centred fields (FieldUpdate, incompressible.rs:226-239)."""
        nx, ny = self.cfg.nx, self.cfg.ny
        out = {k: np.empty((ny, nx)) for k in ("velocity_x", "velocity_y", "speed", "vorticity")}
        check(load().fl_grid2d_cell_fields(self.handle, ptr(out["velocity_x"]),
                                           ptr(out["velocity_y"]), ptr(out["speed"]),
                                           ptr(out["vorticity"])), self.ctx.handle)
        out["pressure"] = self.fields()["p"]
        return out


class CavityBatch:
    """Many independent 2D problems advanced together on one GPU."""

    def __init__(self, cfgs, ctx: _lib.Context | None = None):
        self.cfgs = list(cfgs)
        self.ctx = ctx or _lib.default_context()
        arr = (Grid2DConfig * len(self.cfgs))(*[c.to_c() for c in self.cfgs])
        h = C.c_void_p()
        check(load().fl_batch2d_create(self.ctx.handle, len(self.cfgs), arr, C.byref(h)),
              self.ctx.handle)
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            load().fl_batch2d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self):
        reps = (Grid2DStepReport * len(self.cfgs))()
        check(load().fl_batch2d_step(self.handle, reps), self.ctx.handle)
        return [SolverStep.from_c(r) for r in reps]

    def fields(self, index):
        c = self.cfgs[index]
        u, v, p = np.empty((c.ny, c.nx + 1)), np.empty((c.ny + 1, c.nx)), np.empty((c.ny, c.nx))
        check(load().fl_batch2d_get_fields(self.handle, index, ptr(u), ptr(v), ptr(p)),
              self.ctx.handle)
        return {"u": u, "v": v, "p": p}

    def run(self, steps: int, use_graphs: bool = True, step_times: bool = False,
            pcg_times: bool = False):
        """`steps` device-driven steps (one host synchronisation).  Returns
        (per-grid run reports, per-step device milliseconds or None), plus
        the per-step persistent-PCG milliseconds when `pcg_times` (-1 where
        the step replays as one graph); raises FlursysError with the first
        failing grid's message."""
        reps = (_lib.Batch2DRunReport * len(self.cfgs))()
        ms = np.zeros(max(steps, 1)) if step_times else None
        pms = np.zeros(max(steps, 1)) if pcg_times else None
        st = load().fl_batch2d_run(self.handle, steps, int(use_graphs), reps,
                                   ptr(ms) if step_times else None,
                                   ptr(pms) if pcg_times else None)
        self.last_run = list(reps)
        check(st, self.ctx.handle)
        out = (self.last_run, (ms[:steps] if step_times else None))
        return out + (pms[:steps],) if pcg_times else out

    def autotune(self, iterations: int = 20):
        """Chooses the persistent PCG's launch shape; returns (chosen, table),
        each entry a dict with threads, blocks and ms."""
        best = _lib.Batch2DTuning()
        tab = (_lib.Batch2DTuning * 64)()
        n = C.c_int32()
        check(load().fl_batch2d_autotune(self.handle, iterations, C.byref(best), 64, tab,
                                         C.byref(n)), self.ctx.handle)
        return _tuning(best), [_tuning(tab[i]) for i in range(n.value)]

    def graph_mode(self) -> int:
        """How graph runs replay a step: _lib.GRAPH_WHOLE_STEP (one graph with
        the persistent PCG as a cooperative kernel node), GRAPH_SPLIT (graphs
        around a cooperative launch) or GRAPH_NONE (not captured yet)."""
        m = C.c_int32()
        check(load().fl_batch2d_graph_mode(self.handle, C.byref(m)), self.ctx.handle)
        return m.value

    def launch(self):
        """The persistent PCG's current launch shape."""
        t = _lib.Batch2DTuning()
        check(load().fl_batch2d_get_launch(self.handle, C.byref(t)), self.ctx.handle)
        return _tuning(t)

    def set_launch(self, threads: int, blocks: int):
        t = _lib.Batch2DTuning(threads, blocks)
        check(load().fl_batch2d_set_launch(self.handle, C.byref(t)), self.ctx.handle)


def _tuning(t) -> dict:
    return {"threads": t.threads, "blocks": t.blocks, "ms": t.ms}
