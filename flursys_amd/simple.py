"""Collocated unstructured SIMPLE on the GPU for 2D polygon meshes.

Mirrors src/simple.rs: ``SimpleOptions.steady`` (:115-131), ``solve_simple``
(:148-311) over a ``SimpleState`` (velocity, pressure, authoritative face
flux, :64-91), ``initial_face_flux`` (:719-744), and the SimpleError kinds
(:17-29), over the C ABI (include/flursys_hip.h, fl_simple2d_*).  Every outer
iteration runs on the device: the least-squares gradient, the momentum and
pressure-correction CSR assemblies, BiCGSTAB and Jacobi-PCG, Rhie-Chow and
the corrections; the host reads one error word and the continuity RMS per
iteration.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import LinOpts, Simple2DOptions, Simple2DReport, check, load, ptr
from .linear import LinearSolverOptions
from .mesh import DeviceMesh2D

# fl_simple2d_report.error_kind -> SimpleError / NumericsError / MomentumError variant
SIMPLE_KINDS = {1: "InvalidMomentumDiagonal", 2: "InvalidPressureFaceCoefficient",
                3: "InvalidPressureReference", 4: "InvalidPressureRelaxation",
                5: "InvalidOuterIterations", 6: "MomentumLinearDidNotConverge",
                7: "PressureLinearDidNotConverge", 8: "SingularLeastSquaresStencil",
                9: "DegenerateDistance", 10: "MissingBoundaryCondition", 11: "Backflow",
                12: "NonFiniteFlux", 13: "Linear", 14: "InvalidOptions"}
ZERO_GRADIENT, FIXED_VALUE = 0, 1
_SIMPLE_STATUS = 16


class SimpleError(ValueError):
    def __init__(self, kind, message, index=0, value=0.0):
        super().__init__(f"{kind}: {message}")
        self.kind, self.index, self.value = kind, index, value


@dataclass
class SimpleOptions:
    """SimpleOptions::steady (src/simple.rs:115-131) with a constant viscosity."""

    viscosity: float
    reference_cell: int = 0
    max_outer_iterations: int = 200
    continuity_absolute_tolerance: float = 1.0e-10
    momentum_solver: LinearSolverOptions = field(default_factory=LinearSolverOptions)
    pressure_solver: LinearSolverOptions = field(default_factory=LinearSolverOptions)
    velocity_relaxation: float = 0.7
    pressure_relaxation: float = 0.3

    def to_c(self) -> Simple2DOptions:
        def lin(o):
            return LinOpts(o.absolute_tolerance, o.relative_tolerance, o.max_iterations)
        return Simple2DOptions(self.viscosity, self.reference_cell, self.max_outer_iterations,
                               self.continuity_absolute_tolerance, lin(self.momentum_solver),
                               lin(self.pressure_solver), self.velocity_relaxation,
                               self.pressure_relaxation)


@dataclass
class SimpleReport:
    """SimpleReport (src/simple.rs:133-142)."""

    converged: bool
    outer_iterations: int
    initial_continuity_rms: float
    final_continuity_rms: float
    continuity_history: np.ndarray
    momentum_reports: list
    pressure_report: object
    device_ms: float


def _scalar(nf, boundary, default=None):
    kind = np.full(nf, -1 if default is None else default, np.int32)
    value = np.zeros(nf)
    for f, bc in (boundary or {}).items():
        if bc == "ZeroGradient":
            kind[f] = ZERO_GRADIENT
        else:
            kind[f], value[f] = FIXED_VALUE, bc[1]
    return kind, value


class DeviceSimple2D:
    """A SIMPLE solver bound to a DeviceMesh2D.

    velocity_boundary: {boundary face: (ux, uy) (FixedVelocity) | "ZeroGradient"};
    pressure_boundary: {boundary face: "ZeroGradient" | ("FixedValue", p)}, the
    LSQ stencil's boundary; correction_boundary: the same form or None (then
    ``reference_cell`` pins the correction)."""

    def __init__(self, mesh: DeviceMesh2D, velocity_boundary, pressure_boundary,
                 correction_boundary=None):
        self.mesh = mesh
        self.ctx = mesh.ctx
        nf = mesh.nfaces
        vk = np.full(nf, -1, np.int32)
        vx, vy = np.zeros(nf), np.zeros(nf)
        for f, bc in velocity_boundary.items():
            if bc == "ZeroGradient":
                vk[f] = ZERO_GRADIENT
            else:
                vk[f], vx[f], vy[f] = FIXED_VALUE, bc[0], bc[1]
        pk, pv = _scalar(nf, pressure_boundary)
        ck = cv = None
        if correction_boundary is not None:
            ck, cv = _scalar(nf, correction_boundary)
        h = C.c_void_p()
        rep = Simple2DReport()
        st = load().fl_simple2d_create(mesh.handle, ptr(vk), ptr(vx), ptr(vy), ptr(pk), ptr(pv),
                                       ptr(ck), ptr(cv), C.byref(h), C.byref(rep))
        self._raise(st, rep)
        self.handle = h

    def _raise(self, st, rep):
        if st == _SIMPLE_STATUS:
            raw = load().fl_last_error(self.ctx.handle)
            raise SimpleError(SIMPLE_KINDS.get(rep.error_kind, str(rep.error_kind)),
                              raw.decode() if raw else "", rep.error_index, rep.error_value)
        check(st, self.ctx.handle)

    def close(self):
        if getattr(self, "handle", None):
            load().fl_simple2d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initial_face_flux(self, velocity, wall_flux):
        """initial_face_flux (src/simple.rs:719-744) on the device."""
        v = np.ascontiguousarray(velocity, dtype=np.float64)
        ux, uy = np.ascontiguousarray(v[:, 0]), np.ascontiguousarray(v[:, 1])
        out = np.zeros(self.mesh.nfaces)
        check(load().fl_simple2d_initial_face_flux(self.handle, ptr(ux), ptr(uy),
                                                   ptr(np.ascontiguousarray(wall_flux, dtype=np.float64)),
                                                   ptr(out)), self.ctx.handle)
        return out

    def set_state(self, velocity=None, pressure=None, face_flux=None):
        ux = uy = None
        if velocity is not None:
            v = np.ascontiguousarray(velocity, dtype=np.float64)
            ux, uy = np.ascontiguousarray(v[:, 0]), np.ascontiguousarray(v[:, 1])
        p = None if pressure is None else np.ascontiguousarray(pressure, dtype=np.float64)
        phi = None if face_flux is None else np.ascontiguousarray(face_flux, dtype=np.float64)
        check(load().fl_simple2d_set_state(self.handle, ptr(ux), ptr(uy), ptr(p), ptr(phi)),
              self.ctx.handle)

    def state(self):
        nc, nf = self.mesh.ncells, self.mesh.nfaces
        ux, uy, p, phi = np.zeros(nc), np.zeros(nc), np.zeros(nc), np.zeros(nf)
        check(load().fl_simple2d_get_state(self.handle, ptr(ux), ptr(uy), ptr(p), ptr(phi)),
              self.ctx.handle)
        return {"velocity": np.stack([ux, uy], 1), "pressure": p, "face_flux": phi}

    def solve(self, options: SimpleOptions) -> SimpleReport:
        """solve_simple (src/simple.rs:148-311) from the current state."""
        o = options.to_c()
        rep = Simple2DReport()
        hist = np.zeros(max(options.max_outer_iterations, 0) + 1)
        st = load().fl_simple2d_solve(self.handle, C.byref(o), C.byref(rep), ptr(hist))
        self._raise(st, rep)
        n = rep.outer_iterations if rep.converged else options.max_outer_iterations
        mom = [_lib.LinReport.from_buffer_copy(rep.momentum[k]) for k in range(3)]
        return SimpleReport(bool(rep.converged), rep.outer_iterations if rep.converged
                            else options.max_outer_iterations, rep.initial_continuity_rms,
                            rep.final_continuity_rms, hist[:n + 1], mom,
                            _lib.LinReport.from_buffer_copy(rep.pressure), rep.device_ms)
