"""ctypes front-end of the C oracle (liboracle.so built by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: the checker for the GPU path and the timed CPU
baseline.  Each wrapper names the C function it calls; the C sources cite the
reference lines they restate.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

STOP_ABS_INF, STOP_REL_L2 = 0, 1
GUESS_PREVIOUS, GUESS_ZERO = 0, 1
SOLVER_PCG, SOLVER_SOR = 0, 1
COUPLING_PROJECTION, COUPLING_SIMPLE = 0, 1
UPWIND, CENTRAL = 0, 1
BC_VELOCITY, BC_SYMMETRY, BC_OUTLET = 0, 1, 2
LEFT, RIGHT, BOTTOM, TOP = 0, 1, 2, 3

_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class SolveReport(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("residual_abs_inf", C.c_double),
        ("residual_l2", C.c_double),
        ("initial_l2", C.c_double),
        ("converged", C.c_int32),
        ("status", C.c_int32),
        ("message", C.c_char * 160),
    ]


class Config3D(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("length", C.c_double), ("height", C.c_double), ("depth", C.c_double),
        ("density", C.c_double), ("lid_velocity", C.c_double), ("reynolds", C.c_double),
        ("dt", C.c_double),
        ("pressure_max_iters", C.c_int64),
        ("pressure_tolerance", C.c_double),
        ("pressure_omega", C.c_double),
        ("solver", C.c_int32), ("stop_rule", C.c_int32), ("initial_guess", C.c_int32),
        ("threads", C.c_int32),
    ]


class State3D(C.Structure):
    _fields_ = [
        ("cfg", Config3D),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("n", C.c_int64),
        ("p", C.POINTER(C.c_double)), ("rhs", C.POINTER(C.c_double)),
        ("u", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)),
        ("w", C.POINTER(C.c_double)), ("us", C.POINTER(C.c_double)),
        ("vs", C.POINTER(C.c_double)), ("ws", C.POINTER(C.c_double)),
        ("x", C.POINTER(C.c_double)), ("b", C.POINTER(C.c_double)),
        ("r", C.POINTER(C.c_double)), ("z", C.POINTER(C.c_double)),
        ("d", C.POINTER(C.c_double)), ("q", C.POINTER(C.c_double)),
        ("step", C.c_int64),
    ]


class Config2D(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64),
        ("length", C.c_double), ("height", C.c_double),
        ("density", C.c_double), ("nu", C.c_double), ("dt", C.c_double),
        ("pressure_max_iters", C.c_int64),
        ("pressure_tolerance", C.c_double),
        ("pressure_omega", C.c_double),
        ("convection", C.c_int32), ("solver", C.c_int32),
        ("stop_rule", C.c_int32), ("initial_guess", C.c_int32), ("threads", C.c_int32),
        ("bc_kind", C.c_int32 * 4),
        ("bc_u", C.c_double * 4), ("bc_v", C.c_double * 4),
        ("outlet_pressure", C.c_double * 4),
        ("adaptive", C.c_int32),
        ("min_dt", C.c_double), ("max_dt", C.c_double), ("target_cfl", C.c_double),
        ("target_diffusion", C.c_double), ("max_growth", C.c_double),
        ("t_end", C.c_double),
        ("inlet_profile", C.c_int32), ("inlet_step_height", C.c_double),
        ("coupling", C.c_int32), ("velocity_relaxation", C.c_double),
        ("pressure_relaxation", C.c_double),
    ]


class Diag2D(C.Structure):
    _fields_ = [
        ("time_step", C.c_double), ("pressure_residual", C.c_double),
        ("max_divergence", C.c_double), ("max_speed", C.c_double),
        ("momentum_cfl", C.c_double), ("viscous_diffusion_number", C.c_double),
        ("velocity_change", C.c_double), ("pressure_iterations", C.c_int64),
    ]


class State2D(C.Structure):
    _fields_ = [
        ("cfg", Config2D),
        ("dx", C.c_double), ("dy", C.c_double),
        ("n", C.c_int64),
        ("solid", C.POINTER(C.c_uint8)),
    ] + [(name, C.POINTER(C.c_double)) for name in (
        "p", "rhs", "u", "v", "us", "vs", "uold", "vold", "ucell", "vcell", "speed", "vort",
        "x", "b", "r", "z", "d", "q", "pold")] + [
        ("step", C.c_int64), ("time", C.c_double), ("current_dt", C.c_double),
    ]


class LinReport(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("converged", C.c_int32), ("iterations", C.c_int64),
        ("initial_residual", C.c_double), ("final_residual", C.c_double),
        ("relative_residual", C.c_double),
        ("err_index", C.c_int64), ("err_expected", C.c_int64), ("err_actual", C.c_int64),
        ("err_row", C.c_int64), ("err_col", C.c_int64),
        ("reason", C.c_char * 96),
    ]


class Mesh2DStruct(C.Structure):
    _fields_ = [("ncells", C.c_int64), ("nfaces", C.c_int64)] + [
        (name, C.POINTER(C.c_double)) for name in ("cx", "cy", "vol")] + [
        (name, C.POINTER(C.c_int64)) for name in ("owner", "neigh")] + [
        (name, C.POINTER(C.c_double)) for name in ("fcx", "fcy", "sx", "sy")] + [
        (name, C.POINTER(C.c_int64)) for name in ("fv0", "fv1", "cell_off", "cell_faces")]


class Lsq2DStruct(C.Structure):
    _fields_ = [("ncells", C.c_int64), ("kind", C.POINTER(C.c_int32)),
                ("nb", C.POINTER(C.c_int64))] + [
        (name, C.POINTER(C.c_double)) for name in ("val", "rcx", "rcy", "inv")]


class SimpleOpts(C.Structure):
    _fields_ = [("viscosity", C.c_double), ("reference_cell", C.c_int64),
                ("max_outer_iterations", C.c_int64), ("continuity_tol", C.c_double),
                ("mom_atol", C.c_double), ("mom_rtol", C.c_double), ("mom_max", C.c_int64),
                ("p_atol", C.c_double), ("p_rtol", C.c_double), ("p_max", C.c_int64),
                ("alpha_u", C.c_double), ("alpha_p", C.c_double)]


class SimpleReportC(C.Structure):
    _fields_ = [("converged", C.c_int32), ("pad", C.c_int32), ("outer_iterations", C.c_int64),
                ("initial_rms", C.c_double), ("final_rms", C.c_double),
                ("mom_iters", C.c_int64 * 3), ("mom_converged", C.c_int32 * 3),
                ("p_converged", C.c_int32), ("mom_final", C.c_double * 3),
                ("p_iters", C.c_int64), ("p_final", C.c_double), ("err_index", C.c_int64),
                ("err_value", C.c_double)]


def build() -> None:
    """Compile liboracle.so from the C sources (make, in-tree)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or3d_init.argtypes = [C.POINTER(State3D), C.POINTER(Config3D)]
        L.or3d_free.argtypes = [C.POINTER(State3D)]
        for name in ("or3d_predict", "or3d_build_rhs", "or3d_correct"):
            getattr(L, name).argtypes = [C.POINTER(State3D)]
        L.or3d_solve_pressure.argtypes = [C.POINTER(State3D), C.POINTER(SolveReport)]
        L.or3d_step.argtypes = [C.POINTER(State3D), C.POINTER(SolveReport)]
        for name in ("or3d_max_divergence", "or3d_max_speed", "or3d_pressure_residual"):
            getattr(L, name).argtypes = [C.POINTER(State3D)]
            getattr(L, name).restype = C.c_double
        L.or3d_apply_operator.argtypes = [C.POINTER(State3D), _f64p, _f64p]
        L.or2d_init.argtypes = [C.POINTER(State2D), C.POINTER(Config2D), C.c_void_p]
        L.or2d_free.argtypes = [C.POINTER(State2D)]
        L.or2d_step.argtypes = [C.POINTER(State2D), C.POINTER(Diag2D), C.POINTER(SolveReport)]
        L.or2d_reached_end.argtypes = [C.POINTER(State2D)]
        L.or2d_solve_pressure.argtypes = [C.POINTER(State2D), C.POINTER(SolveReport)]
        L.or2d_predict.argtypes = [C.POINTER(State2D), C.c_double]
        L.or2d_build_rhs.argtypes = [C.POINTER(State2D), C.c_double]
        L.or2d_correct.argtypes = [C.POINTER(State2D), C.c_double]
        L.or2d_apply_velocity_boundaries.argtypes = [
            C.POINTER(State2D), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or2d_cell_fields.argtypes = [C.POINTER(State2D)]
        for name in ("or2d_max_divergence", "or2d_pressure_residual"):
            getattr(L, name).argtypes = [C.POINTER(State2D)]
            getattr(L, name).restype = C.c_double
        L.or2d_apply_operator.argtypes = [C.POINTER(State2D), _f64p, _f64p]
        L.or2d_stability_numbers.argtypes = [C.c_double] * 6 + [C.POINTER(C.c_double)] * 2
        L.or2d_adaptive_dt.argtypes = [C.c_double] * 12 + [
            C.POINTER(C.c_double), C.c_char_p, C.c_int]
        L.or_csr_validate.argtypes = [C.c_int64, C.c_int64, _i64p, _i64p, _f64p, C.c_int64,
                                      C.POINTER(LinReport)]
        L.or_csr_spmv.argtypes = [C.c_int64, _i64p, _i64p, _f64p, _f64p, _f64p,
                                  C.POINTER(LinReport)]
        L.or_csr_finalize.argtypes = [C.c_int64, C.c_int64, C.c_int64, _i64p, _i64p, _f64p,
                                      _i64p, _i64p, _f64p, C.POINTER(C.c_int64),
                                      C.POINTER(LinReport)]
        L.or_jacobi_inverse.argtypes = [C.c_int64, _i64p, _i64p, _f64p, _f64p,
                                        C.POINTER(LinReport)]
        L.or_cg.argtypes = [C.c_int64, _i64p, _i64p, _f64p, _f64p, _f64p, C.c_double,
                            C.c_double, C.c_int64, C.c_int32, C.POINTER(LinReport)]
        L.or_bicgstab.argtypes = [C.c_int64, _i64p, _i64p, _f64p, _f64p, _f64p, C.c_double,
                                  C.c_double, C.c_int64, C.POINTER(LinReport)]
        P64 = C.POINTER(C.POINTER(C.c_int64))
        PD = C.POINTER(C.POINTER(C.c_double))
        L.or_mesh_poisson.argtypes = [
            C.c_int64, C.c_double, C.c_uint64, C.c_int64, C.POINTER(C.c_int64),
            C.POINTER(C.c_int64), P64, P64, PD, PD, C.POINTER(C.c_int64), P64, P64, PD,
            C.c_char_p, C.c_int]
        L.or_mesh2d_from_cells.argtypes = [C.c_int64, _f64p, C.c_int64, _i64p, _i64p,
                                           C.POINTER(Mesh2DStruct), C.c_char_p, C.c_int]
        L.or_mesh2d_free.argtypes = [C.POINTER(Mesh2DStruct)]
        L.or_assemble_poisson.argtypes = [
            C.POINTER(Mesh2DStruct), C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
            C.c_int64, C.c_double, C.POINTER(C.c_int64), P64, P64, PD, PD,
            C.POINTER(C.c_double), C.c_char_p, C.c_int]
        L.or_face_weights.argtypes = [C.POINTER(Mesh2DStruct), _f64p, C.c_char_p, C.c_int]
        L.or_pressure_face_coefficients.argtypes = [C.POINTER(Mesh2DStruct), _f64p, _f64p,
                                                    C.c_char_p, C.c_int]
        L.or_assemble_pressure_correction.argtypes = [
            C.POINTER(Mesh2DStruct), _f64p, _f64p, C.c_int64, C.POINTER(C.c_int64), P64, P64,
            PD, PD, C.c_char_p, C.c_int]
        L.or_assemble_momentum.argtypes = [
            C.POINTER(Mesh2DStruct), C.c_double, _f64p, C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_void_p, C.c_double, C.c_void_p, C.c_double, C.POINTER(C.c_int64), P64, P64, PD,
            PD, PD, C.c_char_p, C.c_int]
        L.or_lsq2d_build.argtypes = [C.POINTER(Mesh2DStruct), C.c_void_p, C.c_void_p,
                                     C.POINTER(Lsq2DStruct), C.POINTER(C.c_int64), C.c_char_p,
                                     C.c_int]
        L.or_lsq2d_free.argtypes = [C.POINTER(Lsq2DStruct)]
        L.or_lsq2d_gradient.argtypes = [C.POINTER(Mesh2DStruct), C.POINTER(Lsq2DStruct), _f64p,
                                        _f64p, _f64p]
        L.or_initial_face_flux.argtypes = [C.POINTER(Mesh2DStruct), _f64p, _f64p, _f64p, _f64p,
                                           C.POINTER(C.c_int64), C.c_char_p, C.c_int]
        L.or_continuity_rms.argtypes = [C.POINTER(Mesh2DStruct), _f64p]
        L.or_continuity_rms.restype = C.c_double
        L.or_rhie_chow_flux.argtypes = [C.POINTER(Mesh2DStruct), C.POINTER(Lsq2DStruct), _f64p,
                                        _f64p, _f64p, _f64p, _f64p, _f64p, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_double), C.c_char_p, C.c_int]
        L.or_correct_face_flux.argtypes = [C.POINTER(Mesh2DStruct), _f64p, _f64p, _f64p,
                                           C.c_void_p, C.c_void_p, _f64p, C.POINTER(C.c_int64),
                                           C.POINTER(C.c_double), C.c_char_p, C.c_int]
        L.or_correct_cell_velocity.argtypes = [C.c_int64, _f64p, _f64p, _f64p, _f64p, _f64p,
                                               _f64p, _f64p, C.POINTER(C.c_int64),
                                               C.POINTER(C.c_double), C.c_char_p, C.c_int]
        L.or_simple2d_solve.argtypes = [
            C.POINTER(Mesh2DStruct), C.POINTER(Lsq2DStruct), C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_void_p, C.c_void_p, C.POINTER(SimpleOpts), _f64p, _f64p, _f64p, _f64p, _f64p,
            C.POINTER(SimpleReportC), C.c_char_p, C.c_int]
        L.or_free.argtypes = [C.c_void_p]
        L.or_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.or_splitmix64.restype = C.c_uint64
        L.or_uniform.argtypes = [C.POINTER(C.c_uint64)]
        L.or_uniform.restype = C.c_double
        L.or_normal.argtypes = [C.POINTER(C.c_uint64)]
        L.or_normal.restype = C.c_double
        _lib = L
    return _lib


def _arr(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,))


# ----------------------------------------------------------------------------
# 3D cavity
# ----------------------------------------------------------------------------
class Cavity3D:
    """Oracle mirror of LidDrivenCavity3DSolver (src/solver/incompressible_3d.rs)."""

    def __init__(self, nx, ny=None, nz=None, *, length=1.0, height=1.0, depth=1.0,
                 density=1.0, lid_velocity=1.0, reynolds=100.0, dt=5.0e-4,
                 pressure_max_iters=2000, pressure_tolerance=1.0e-6, pressure_omega=1.7,
                 solver=SOLVER_PCG, stop_rule=STOP_ABS_INF, initial_guess=GUESS_PREVIOUS,
                 threads=0):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        self.cfg = Config3D(nx, ny, nz, length, height, depth, density, lid_velocity,
                            reynolds, dt, pressure_max_iters, pressure_tolerance,
                            pressure_omega, solver, stop_rule, initial_guess, threads)
        self.s = State3D()
        st = lib().or3d_init(C.byref(self.s), C.byref(self.cfg))
        if st != 0:
            raise ValueError(f"invalid 3D cavity configuration (status {st})")
        self.nx, self.ny, self.nz = nx, ny, nz

    def __del__(self):
        try:
            lib().or3d_free(C.byref(self.s))
        except Exception:
            pass

    # field views (live, writable)
    def field(self, name):
        nx, ny, nz = self.nx, self.ny, self.nz
        shapes = {"p": (nz, ny, nx), "rhs": (nz, ny, nx), "u": (nz, ny, nx + 1),
                  "us": (nz, ny, nx + 1), "v": (nz, ny + 1, nx), "vs": (nz, ny + 1, nx),
                  "w": (nz + 1, ny, nx), "ws": (nz + 1, ny, nx)}
        shp = shapes[name]
        return _arr(getattr(self.s, name), int(np.prod(shp))).reshape(shp)

    def step(self):
        rep = SolveReport()
        st = lib().or3d_step(C.byref(self.s), C.byref(rep))
        if st != 0:
            raise RuntimeError(rep.message.decode())
        return rep

    def predict(self):
        lib().or3d_predict(C.byref(self.s))

    def build_rhs(self):
        lib().or3d_build_rhs(C.byref(self.s))

    def solve_pressure(self):
        rep = SolveReport()
        st = lib().or3d_solve_pressure(C.byref(self.s), C.byref(rep))
        if st != 0:
            raise RuntimeError(rep.message.decode())
        return rep

    def correct(self):
        lib().or3d_correct(C.byref(self.s))

    def apply_operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.empty_like(x)
        lib().or3d_apply_operator(C.byref(self.s), x, y)
        return y

    def max_divergence(self):
        return lib().or3d_max_divergence(C.byref(self.s))

    def max_speed(self):
        return lib().or3d_max_speed(C.byref(self.s))

    def pressure_residual(self):
        return lib().or3d_pressure_residual(C.byref(self.s))


# ----------------------------------------------------------------------------
# 2D solver
# ----------------------------------------------------------------------------
class Solver2D:
    """Oracle mirror of IncompressibleSolver (src/solver/incompressible.rs),
    projection or SIMPLE-style coupling."""

    def __init__(self, nx, ny=None, *, length=1.0, height=1.0, density=1.0, nu=0.01,
                 dt=1.0e-3, t_end=1.0e30, pressure_max_iters=800, pressure_tolerance=1.0e-6,
                 pressure_omega=1.7, convection=UPWIND, solver=SOLVER_PCG,
                 stop_rule=STOP_ABS_INF, initial_guess=GUESS_PREVIOUS, threads=0,
                 bc_kind=(BC_VELOCITY,) * 4, bc_u=(0.0, 0.0, 0.0, 1.0), bc_v=(0.0,) * 4,
                 outlet_pressure=(0.0,) * 4, adaptive=False, min_dt=1.0e-6, max_dt=1.0e-2,
                 target_cfl=0.5, target_diffusion=0.25, max_growth=1.1, solid=None,
                 inlet_profile=0, inlet_step_height=0.0, coupling=COUPLING_PROJECTION,
                 velocity_relaxation=0.7, pressure_relaxation=0.3):
        ny = nx if ny is None else ny
        cfg = Config2D()
        cfg.nx, cfg.ny, cfg.length, cfg.height = nx, ny, length, height
        cfg.density, cfg.nu, cfg.dt = density, nu, dt
        cfg.pressure_max_iters, cfg.pressure_tolerance = pressure_max_iters, pressure_tolerance
        cfg.pressure_omega, cfg.convection, cfg.solver = pressure_omega, convection, solver
        cfg.stop_rule, cfg.initial_guess, cfg.threads = stop_rule, initial_guess, threads
        for i in range(4):
            cfg.bc_kind[i] = bc_kind[i]
            cfg.bc_u[i] = bc_u[i]
            cfg.bc_v[i] = bc_v[i]
            cfg.outlet_pressure[i] = outlet_pressure[i]
        cfg.adaptive = int(adaptive)
        cfg.min_dt, cfg.max_dt, cfg.target_cfl = min_dt, max_dt, target_cfl
        cfg.target_diffusion, cfg.max_growth, cfg.t_end = target_diffusion, max_growth, t_end
        cfg.inlet_profile, cfg.inlet_step_height = inlet_profile, inlet_step_height
        cfg.coupling, cfg.velocity_relaxation = coupling, velocity_relaxation
        cfg.pressure_relaxation = pressure_relaxation
        self.cfg = cfg
        self.s = State2D()
        self._solid = None
        sp = None
        if solid is not None:
            self._solid = np.ascontiguousarray(solid, dtype=np.uint8).ravel()
            sp = self._solid.ctypes.data
        st = lib().or2d_init(C.byref(self.s), C.byref(cfg), sp)
        if st != 0:
            raise ValueError(f"invalid 2D configuration (status {st})")
        self.nx, self.ny = nx, ny

    def __del__(self):
        try:
            lib().or2d_free(C.byref(self.s))
        except Exception:
            pass

    def field(self, name):
        nx, ny = self.nx, self.ny
        shp = {"u": (ny, nx + 1), "us": (ny, nx + 1), "uold": (ny, nx + 1),
               "v": (ny + 1, nx), "vs": (ny + 1, nx), "vold": (ny + 1, nx)}.get(name, (ny, nx))
        return _arr(getattr(self.s, name), shp[0] * shp[1]).reshape(shp)

    def step(self):
        d, rep = Diag2D(), SolveReport()
        st = lib().or2d_step(C.byref(self.s), C.byref(d), C.byref(rep))
        if st != 0:
            raise RuntimeError(rep.message.decode())
        return d, rep

    def solve_pressure(self):
        rep = SolveReport()
        st = lib().or2d_solve_pressure(C.byref(self.s), C.byref(rep))
        if st != 0:
            raise RuntimeError(rep.message.decode())
        return rep

    def predict(self, dt):
        """predict_momentum (projection coupling): u, v -> us, vs."""
        lib().or2d_predict(C.byref(self.s), dt)

    def apply_boundaries(self):
        """apply_velocity_boundaries on the current u, v."""
        lib().or2d_apply_velocity_boundaries(C.byref(self.s), self.s.u, self.s.v)

    def cell_fields(self):
        """compute_cell_fields: ucell, vcell, speed, vort from u, v."""
        lib().or2d_cell_fields(C.byref(self.s))

    def apply_operator(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.empty_like(x)
        lib().or2d_apply_operator(C.byref(self.s), x, y)
        return y

    def max_divergence(self):
        return lib().or2d_max_divergence(C.byref(self.s))

    def pressure_residual(self):
        return lib().or2d_pressure_residual(C.byref(self.s))

    @property
    def time(self):
        return self.s.time

    def reached_end(self):
        return bool(lib().or2d_reached_end(C.byref(self.s)))


def stability_numbers(dt, nu, dx, dy, umax, vmax):
    """momentum_stability_numbers (src/solver/incompressible.rs:2294-2306)."""
    a, b = C.c_double(), C.c_double()
    lib().or2d_stability_numbers(dt, nu, dx, dy, umax, vmax, C.byref(a), C.byref(b))
    return a.value, b.value


def adaptive_time_step(current_dt, umax, vmax, dx, dy, nu, thermal_diffusivity=None, *,
                       min_dt, max_dt, target_cfl, target_diffusion, max_growth):
    """adaptive_time_step (src/solver/incompressible.rs:2308-2336)."""
    out = C.c_double()
    msg = C.create_string_buffer(256)
    td = -1.0 if thermal_diffusivity is None else thermal_diffusivity
    st = lib().or2d_adaptive_dt(current_dt, umax, vmax, dx, dy, nu, td, min_dt, max_dt,
                                target_cfl, target_diffusion, max_growth, C.byref(out), msg,
                                256)
    if st != 0:
        raise RuntimeError(msg.value.decode())
    return out.value


# ----------------------------------------------------------------------------
# CSR
# ----------------------------------------------------------------------------
class LinearAlgebraError(Exception):
    def __init__(self, kind, rep):
        self.kind = kind
        self.reason = rep.reason.decode()
        self.index = rep.err_index
        self.row = rep.err_row
        super().__init__(f"{kind}: {self.reason}")


_KINDS = {1: "DimensionMismatch", 2: "InvalidCsr", 3: "IndexOutOfBounds",
          4: "NonFiniteValue", 5: "MissingDiagonal", 6: "ZeroDiagonal", 7: "Breakdown"}


def _check(st, rep):
    if st != 0:
        raise LinearAlgebraError(_KINDS.get(st, str(st)), rep)


class Csr:
    def __init__(self, nrows, ncols, offsets, cols, vals, validate=True):
        self.nrows, self.ncols = nrows, ncols
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.cols = np.ascontiguousarray(cols, dtype=np.int64)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)
        if validate:
            rep = LinReport()
            if len(self.offsets) != nrows + 1:
                raise LinearAlgebraError("InvalidCsr", rep)
            if len(self.cols) != len(self.vals):
                raise LinearAlgebraError("InvalidCsr", rep)
            _check(lib().or_csr_validate(nrows, ncols, self.offsets, self.cols, self.vals,
                                         len(self.vals), C.byref(rep)), rep)

    @staticmethod
    def from_triplets(nrows, ncols, rows, cols, vals):
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        cols = np.ascontiguousarray(cols, dtype=np.int64)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        n = len(vals)
        off = np.zeros(nrows + 1, np.int64)
        oc = np.zeros(max(n, 1), np.int64)
        ov = np.zeros(max(n, 1), np.float64)
        nnz = C.c_int64()
        rep = LinReport()
        _check(lib().or_csr_finalize(nrows, ncols, n, rows, cols, vals, off, oc, ov,
                                     C.byref(nnz), C.byref(rep)), rep)
        return Csr(nrows, ncols, off, oc[:nnz.value].copy(), ov[:nnz.value].copy())

    @staticmethod
    def from_dense(rows):
        a = np.asarray(rows, dtype=np.float64)
        r, c = np.nonzero(a)
        return Csr.from_triplets(a.shape[0], a.shape[1], r, c, a[r, c])

    @property
    def nnz(self):
        return len(self.vals)

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.nrows)
        rep = LinReport()
        _check(lib().or_csr_spmv(self.nrows, self.offsets, self.cols, self.vals, x, y,
                                 C.byref(rep)), rep)
        return y

    def jacobi_inverse(self):
        inv = np.zeros(self.nrows)
        rep = LinReport()
        _check(lib().or_jacobi_inverse(self.nrows, self.offsets, self.cols, self.vals, inv,
                                       C.byref(rep)), rep)
        return inv


def _solve(kind, A, rhs, x, atol, rtol, max_iter):
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    rep = LinReport()
    if kind == "bicgstab":
        st = lib().or_bicgstab(A.nrows, A.offsets, A.cols, A.vals, rhs, x, atol, rtol,
                               max_iter, C.byref(rep))
    else:
        st = lib().or_cg(A.nrows, A.offsets, A.cols, A.vals, rhs, x, atol, rtol, max_iter,
                         1 if kind == "pcg" else 0, C.byref(rep))
    _check(st, rep)
    return rep


def cg(A, rhs, x, atol=1e-10, rtol=1e-8, max_iter=1000):
    """cg (src/linear/solver.rs:43-50); x is updated in place."""
    return _solve("cg", A, rhs, x, atol, rtol, max_iter)


def pcg(A, rhs, x, atol=1e-10, rtol=1e-8, max_iter=1000):
    """pcg with JacobiPreconditioner (src/linear/solver.rs:52-60)."""
    return _solve("pcg", A, rhs, x, atol, rtol, max_iter)


def bicgstab(A, rhs, x, atol=1e-10, rtol=1e-8, max_iter=1000):
    """bicgstab (src/linear/solver.rs:155-256)."""
    return _solve("bicgstab", A, rhs, x, atol, rtol, max_iter)


def mesh_poisson(m, jitter=0.3, seed=0, ref_cell=0):
    """Jittered triangulation + pinned two-point Poisson CSR (oracle/mesh.c)."""
    L = lib()
    nc, nnz, nf = C.c_int64(), C.c_int64(), C.c_int64()
    po, pc, pfo, pfn = (C.POINTER(C.c_int64)() for _ in range(4))
    pv, pvol, pfc = (C.POINTER(C.c_double)() for _ in range(3))
    msg = C.create_string_buffer(256)
    st = L.or_mesh_poisson(m, jitter, seed, ref_cell, C.byref(nc), C.byref(nnz), C.byref(po),
                           C.byref(pc), C.byref(pv), C.byref(pvol), C.byref(nf), C.byref(pfo),
                           C.byref(pfn), C.byref(pfc), msg, 256)
    if st != 0:
        raise RuntimeError(msg.value.decode())
    n, z, f = nc.value, nnz.value, nf.value
    out = dict(
        offsets=_arr(po, n + 1).copy(), cols=_arr(pc, z).copy(), vals=_arr(pv, z).copy(),
        volumes=_arr(pvol, n).copy(), face_owner=_arr(pfo, f).copy(),
        face_neigh=_arr(pfn, f).copy(), face_coef=_arr(pfc, f).copy())
    for p in (po, pc, pv, pvol, pfo, pfn, pfc):
        L.or_free(p)
    out["A"] = Csr(n, n, out["offsets"], out["cols"], out["vals"], validate=False)
    return out


_MESH_KINDS = {20: "EmptyMesh", 21: "NonFinitePoint", 22: "InvalidCellDefinition",
               23: "InvalidPointIndex", 24: "DegenerateFace", 25: "NonManifoldFace",
               26: "ZeroVolumeCell", 27: "InconsistentFaceOrientation"}
_POISSON_KINDS = {30: "MissingReferenceForPureNeumann", 31: "IncompatiblePureNeumannSource",
                  32: "InvalidReferenceCell", 33: "InvalidReferenceValue",
                  34: "NonFiniteSource", 35: "DegenerateOwnerNeighbourDistance",
                  36: "InvalidNonOrthogonalGeometry", 37: "DegenerateOwnerFaceDistance",
                  38: "MissingBoundaryCondition", 39: "Numerics", 40: "Linear"}


class MeshError(Exception):
    def __init__(self, kind, message):
        self.kind = kind
        super().__init__(f"{kind}: {message}")


class PoissonError(Exception):
    def __init__(self, kind, message, detail=0.0):
        self.kind = kind
        self.detail = detail
        super().__init__(f"{kind}: {message}")


ZERO_GRADIENT, FIXED_VALUE = 0, 1


class Mesh2D:
    """UnstructuredMesh::from_cells for 2D polygons (oracle/fvmesh.c)."""

    def __init__(self, points, cells):
        xy = np.ascontiguousarray(np.asarray(points, dtype=np.float64)[:, :2]).ravel()
        off = np.zeros(len(cells) + 1, np.int64)
        off[1:] = np.cumsum([len(c) for c in cells])
        vtx = np.ascontiguousarray(np.concatenate([np.asarray(c, np.int64) for c in cells])
                                   if cells else np.zeros(1, np.int64))
        self.m = Mesh2DStruct()
        msg = C.create_string_buffer(256)
        st = lib().or_mesh2d_from_cells(len(xy) // 2, xy, len(cells), off, vtx,
                                        C.byref(self.m), msg, 256)
        if st != 0:
            raise MeshError(_MESH_KINDS.get(st, str(st)), msg.value.decode())
        nc, nf = self.m.ncells, self.m.nfaces
        self.centers = np.stack([_arr(self.m.cx, nc), _arr(self.m.cy, nc)], 1).copy()
        self.volumes = _arr(self.m.vol, nc).copy()
        self.owner = _arr(self.m.owner, nf).copy()
        self.neighbour = _arr(self.m.neigh, nf).copy()
        self.face_centers = np.stack([_arr(self.m.fcx, nf), _arr(self.m.fcy, nf)], 1).copy()
        self.area_vectors = np.stack([_arr(self.m.sx, nf), _arr(self.m.sy, nf)], 1).copy()

    def __del__(self):
        try:
            lib().or_mesh2d_free(C.byref(self.m))
        except Exception:
            pass

    @property
    def boundary_faces(self):
        return np.nonzero(self.neighbour < 0)[0]


def assemble_poisson(mesh, gamma=1.0, source=None, boundary=None, reference=None):
    """assemble_poisson (src/poisson.rs:146-263) + apply_reference (:303-337).

    `boundary` maps a boundary face index to ("FixedValue", value) or
    "ZeroGradient" (default for every face: ZeroGradient); `reference` is
    (cell, value) or None.  Returns (Csr, rhs)."""
    L = lib()
    nf = mesh.m.nfaces
    kind = np.zeros(nf, np.int32)
    value = np.zeros(nf, np.float64)
    for f, bc in (boundary or {}).items():
        if bc == "ZeroGradient":
            continue
        kind[f], value[f] = FIXED_VALUE, bc[1]
    src = None if source is None else np.ascontiguousarray(source, dtype=np.float64)
    nnz = C.c_int64()
    po, pc = C.POINTER(C.c_int64)(), C.POINTER(C.c_int64)()
    pv, pr = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    detail = C.c_double()
    msg = C.create_string_buffer(256)
    has_ref = reference is not None
    rc, rv = (reference if has_ref else (0, 0.0))
    st = L.or_assemble_poisson(C.byref(mesh.m), gamma,
                               None if src is None else src.ctypes.data, kind.ctypes.data,
                               value.ctypes.data, int(has_ref), rc, rv, C.byref(nnz),
                               C.byref(po), C.byref(pc), C.byref(pv), C.byref(pr),
                               C.byref(detail), msg, 256)
    if st != 0:
        raise PoissonError(_POISSON_KINDS.get(st, str(st)), msg.value.decode(), detail.value)
    n, z = mesh.m.ncells, nnz.value
    off, cols = _arr(po, n + 1).copy(), _arr(pc, z).copy()
    vals, rhs = _arr(pv, z).copy(), _arr(pr, n).copy()
    for p in (po, pc, pv, pr):
        L.or_free(p)
    return Csr(n, n, off, cols, vals, validate=False), rhs


class Rng:
    """splitmix64 stream shared with the library's input generators."""

    def __init__(self, seed):
        self.state = C.c_uint64(seed)

    def next_u64(self):
        return lib().or_splitmix64(C.byref(self.state))

    def uniform(self, n):
        return np.array([lib().or_uniform(C.byref(self.state)) for _ in range(n)])

    def normal(self, n):
        return np.array([lib().or_normal(C.byref(self.state)) for _ in range(n)])


class CallerSystemError(RuntimeError):
    """SimpleError / MomentumError of the caller-shaped assemblies (oracle/simple.c)."""

    def __init__(self, code, message):
        super().__init__(f"{code}: {message}")
        self.code = code


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def face_weights(mesh):
    """Interior-face interpolation weights (src/numerics.rs:294-303); 0 on boundary faces."""
    out = np.zeros(mesh.m.nfaces)
    msg = C.create_string_buffer(256)
    st = lib().or_face_weights(C.byref(mesh.m), out, msg, 256)
    if st != 0:
        raise CallerSystemError(st, msg.value.decode())
    return out


def pressure_face_coefficients(mesh, r_au):
    """pressure_face_coefficients (src/simple.rs:342-370): d_f per face."""
    out = np.zeros(mesh.m.nfaces)
    msg = C.create_string_buffer(256)
    st = lib().or_pressure_face_coefficients(C.byref(mesh.m), _f64(r_au), out, msg, 256)
    if st != 0:
        raise CallerSystemError(st, msg.value.decode())
    return out


def _take_csr(n, nnz, po, pc, pv, extra):
    L = lib()
    off, cols = _arr(po, n + 1).copy(), _arr(pc, nnz).copy()
    vals = _arr(pv, nnz).copy()
    outs = [_arr(p, n).copy() for p in extra]
    for p in (po, pc, pv, *extra):
        L.or_free(p)
    return Csr(n, n, off, cols, vals, validate=False), outs


def assemble_pressure_correction(mesh, flux, coef, reference_cell):
    """assemble_pressure_correction (src/simple.rs:518-579): (Csr, rhs)."""
    nnz = C.c_int64()
    po, pc = C.POINTER(C.c_int64)(), C.POINTER(C.c_int64)()
    pv, pr = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    msg = C.create_string_buffer(256)
    st = lib().or_assemble_pressure_correction(C.byref(mesh.m), _f64(flux), _f64(coef),
                                               reference_cell, C.byref(nnz), C.byref(po),
                                               C.byref(pc), C.byref(pv), C.byref(pr), msg, 256)
    if st != 0:
        raise CallerSystemError(st, msg.value.decode())
    A, (rhs,) = _take_csr(mesh.m.ncells, nnz.value, po, pc, pv, [pr])
    return A, rhs


def assemble_momentum(mesh, viscosity, flux, boundary, source=None, old=None, dt=0.0,
                      relaxation_field=None, under_relaxation=1.0):
    """assemble_momentum_component (src/momentum.rs:180-341) with a constant
    viscosity.  `boundary` maps every boundary face to ("FixedValue", value)
    or "ZeroGradient".  Returns (Csr, rhs, unrelaxed diagonal)."""
    nf = mesh.m.nfaces
    kind = np.full(nf, -1, np.int32)
    value = np.zeros(nf)
    for f, bc in boundary.items():
        if bc == "ZeroGradient":
            kind[f] = 0
        else:
            kind[f], value[f] = FIXED_VALUE, bc[1]
    keep = [None if a is None else _f64(a) for a in (source, old, relaxation_field)]
    ptr = [None if a is None else a.ctypes.data for a in keep]
    nnz = C.c_int64()
    po, pc = C.POINTER(C.c_int64)(), C.POINTER(C.c_int64)()
    pv, pr, pd = C.POINTER(C.c_double)(), C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    msg = C.create_string_buffer(256)
    st = lib().or_assemble_momentum(C.byref(mesh.m), viscosity, _f64(flux), kind.ctypes.data,
                                    value.ctypes.data, ptr[0], ptr[1], dt, ptr[2],
                                    under_relaxation, C.byref(nnz), C.byref(po), C.byref(pc),
                                    C.byref(pv), C.byref(pr), C.byref(pd), msg, 256)
    if st != 0:
        raise CallerSystemError(st, msg.value.decode())
    A, (rhs, diag) = _take_csr(mesh.m.ncells, nnz.value, po, pc, pv, [pr, pd])
    return A, rhs, diag


def triangulated_square(n, jitter=0.2, seed=0):
    """Mesh2D of the unit square: n x n squares, each cut into two triangles
    along its rising diagonal, interior vertices moved uniformly by up to
    `jitter` h in x and y (numpy generator `seed`)."""
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
    return Mesh2D(pts, cells)


# ----------------------------------------------------------------------------
# collocated unstructured SIMPLE (oracle/simple2d.c)
# ----------------------------------------------------------------------------
SIMPLE_ERRORS = {50: "InvalidPressureReference", 51: "InvalidPressureFaceCoefficient",
                 53: "Linear", 54: "InvalidMomentumDiagonal", 55: "InvalidPressureRelaxation",
                 56: "InvalidOuterIterations", 57: "MomentumLinearDidNotConverge",
                 58: "PressureLinearDidNotConverge", 80: "SingularLeastSquaresStencil",
                 81: "DegenerateOwnerNeighbourDistance", 82: "DegenerateOwnerFaceDistance",
                 83: "MissingBoundaryCondition"}


class SimpleError(RuntimeError):
    """SimpleError / NumericsError of the SIMPLE oracle (kind, index, value)."""

    def __init__(self, code, message, index=0, value=0.0):
        self.code, self.kind = code, SIMPLE_ERRORS.get(code, str(code))
        self.index, self.value = index, value
        super().__init__(f"{self.kind}: {message}")


def _scalar_boundary(mesh, boundary):
    """Per-face (kind, value) arrays of a {face: "ZeroGradient" | ("FixedValue", v)}
    map; faces not in the map are -1 (missing)."""
    nf = mesh.m.nfaces
    kind = np.full(nf, -1, np.int32)
    value = np.zeros(nf)
    for f, bc in boundary.items():
        if bc == "ZeroGradient":
            kind[f] = ZERO_GRADIENT
        else:
            kind[f], value[f] = FIXED_VALUE, bc[1]
    return kind, value


class LsqStencil:
    """LeastSquaresGradientStencil::new (src/numerics.rs:106-200) on a Mesh2D
    with a {face: "ZeroGradient" | ("FixedValue", v)} boundary."""

    def __init__(self, mesh, boundary):
        self.mesh = mesh
        self.s = Lsq2DStruct()
        kind, value = _scalar_boundary(mesh, boundary)
        bad = C.c_int64()
        msg = C.create_string_buffer(256)
        st = lib().or_lsq2d_build(C.byref(mesh.m), kind.ctypes.data, value.ctypes.data,
                                  C.byref(self.s), C.byref(bad), msg, 256)
        if st != 0:
            raise SimpleError(st, msg.value.decode(), bad.value)

    def __del__(self):
        try:
            lib().or_lsq2d_free(C.byref(self.s))
        except Exception:
            pass

    def gradient(self, values):
        nc = self.mesh.m.ncells
        gx, gy = np.zeros(nc), np.zeros(nc)
        lib().or_lsq2d_gradient(C.byref(self.mesh.m), C.byref(self.s), _f64(values), gx, gy)
        return np.stack([gx, gy], 1)


def initial_face_flux(mesh, velocity, wall_flux):
    """initial_face_flux (src/simple.rs:719-744); velocity is (ncells, 2)."""
    v = _f64(velocity)
    out = np.zeros(mesh.m.nfaces)
    bad = C.c_int64()
    msg = C.create_string_buffer(256)
    st = lib().or_initial_face_flux(C.byref(mesh.m), _f64(v[:, 0]), _f64(v[:, 1]),
                                    _f64(wall_flux), out, C.byref(bad), msg, 256)
    if st != 0:
        raise SimpleError(st, msg.value.decode(), bad.value)
    return out


def continuity_rms(mesh, flux):
    """continuity_rms (src/simple.rs:703-717)."""
    return lib().or_continuity_rms(C.byref(mesh.m), _f64(flux))


def rhie_chow_flux(mesh, stencil, predictor, pressure, r_au, wall_flux):
    """rhie_chow_predicted_flux (src/simple.rs:746-805)."""
    v = _f64(predictor)
    out = np.zeros(mesh.m.nfaces)
    bad, badv = C.c_int64(), C.c_double()
    msg = C.create_string_buffer(256)
    st = lib().or_rhie_chow_flux(C.byref(mesh.m), C.byref(stencil.s), _f64(v[:, 0]),
                                 _f64(v[:, 1]), _f64(pressure), _f64(r_au), _f64(wall_flux), out,
                                 C.byref(bad), C.byref(badv), msg, 256)
    if st != 0:
        raise SimpleError(st, msg.value.decode(), bad.value, badv.value)
    return out


def correct_face_flux(mesh, predicted, coefficients, correction, correction_boundary=None):
    """correct_face_flux(_with_boundary) (src/simple.rs:407-484)."""
    out = np.zeros(mesh.m.nfaces)
    kind = value = None
    if correction_boundary is not None:
        kind, value = _scalar_boundary(mesh, correction_boundary)
    bad, badv = C.c_int64(), C.c_double()
    msg = C.create_string_buffer(256)
    st = lib().or_correct_face_flux(C.byref(mesh.m), _f64(predicted), _f64(coefficients),
                                    _f64(correction), None if kind is None else kind.ctypes.data,
                                    None if value is None else value.ctypes.data, out,
                                    C.byref(bad), C.byref(badv), msg, 256)
    if st != 0:
        raise SimpleError(st, msg.value.decode(), bad.value, badv.value)
    return out


def correct_cell_velocity(predictor, r_au, gradient):
    """correct_cell_velocity_from_gradient (src/simple.rs:670-701)."""
    v, g = _f64(predictor), _f64(gradient)
    n = len(v)
    ux, uy = np.zeros(n), np.zeros(n)
    bad, badv = C.c_int64(), C.c_double()
    msg = C.create_string_buffer(256)
    st = lib().or_correct_cell_velocity(n, _f64(v[:, 0]), _f64(v[:, 1]), _f64(r_au),
                                        _f64(g[:, 0]), _f64(g[:, 1]), ux, uy, C.byref(bad),
                                        C.byref(badv), msg, 256)
    if st != 0:
        raise SimpleError(st, msg.value.decode(), bad.value, badv.value)
    return np.stack([ux, uy], 1)


def solve_simple(mesh, stencil, velocity_boundary, velocity, pressure, flux, viscosity,
                 correction_boundary=None, reference_cell=0, max_outer_iterations=200,
                 continuity_tol=1.0e-10, momentum_solver=(1.0e-10, 1.0e-8, 1000),
                 pressure_solver=(1.0e-10, 1.0e-8, 1000), velocity_relaxation=0.7,
                 pressure_relaxation=0.3):
    """solve_simple (src/simple.rs:148-311) with SimpleOptions::steady defaults.
    `velocity_boundary` maps every boundary face to (ux, uy) (FixedVelocity)
    or "ZeroGradient".  Returns (report dict, velocity, pressure, flux,
    continuity history)."""
    nf = mesh.m.nfaces
    vk = np.full(nf, -1, np.int32)
    vx, vy = np.zeros(nf), np.zeros(nf)
    for f, bc in velocity_boundary.items():
        if bc == "ZeroGradient":
            vk[f] = ZERO_GRADIENT
        else:
            vk[f], vx[f], vy[f] = FIXED_VALUE, bc[0], bc[1]
    ck = cv = None
    if correction_boundary is not None:
        ck, cv = _scalar_boundary(mesh, correction_boundary)
    o = SimpleOpts(viscosity, reference_cell, max_outer_iterations, continuity_tol,
                   *momentum_solver, *pressure_solver, velocity_relaxation, pressure_relaxation)
    v = _f64(velocity).copy()
    ux, uy = _f64(v[:, 0]).copy(), _f64(v[:, 1]).copy()
    p, phi = _f64(pressure).copy(), _f64(flux).copy()
    hist = np.zeros(max(max_outer_iterations, 0) + 1)
    rep = SimpleReportC()
    msg = C.create_string_buffer(256)
    st = lib().or_simple2d_solve(C.byref(mesh.m), C.byref(stencil.s), vk.ctypes.data,
                                 vx.ctypes.data, vy.ctypes.data,
                                 None if ck is None else ck.ctypes.data,
                                 None if cv is None else cv.ctypes.data, C.byref(o), ux, uy, p,
                                 phi, hist, C.byref(rep), msg, 256)
    if st != 0:
        raise SimpleError(st, msg.value.decode(), rep.err_index, rep.err_value)
    report = {"converged": bool(rep.converged), "outer_iterations": rep.outer_iterations,
              "initial_continuity_rms": rep.initial_rms, "final_continuity_rms": rep.final_rms,
              "momentum_iterations": list(rep.mom_iters),
              "momentum_converged": [bool(c) for c in rep.mom_converged],
              "pressure_iterations": rep.p_iters, "pressure_converged": bool(rep.p_converged)}
    done = rep.outer_iterations if rep.converged else max_outer_iterations
    return report, np.stack([ux, uy], 1), p, phi, hist[:done + 1]
