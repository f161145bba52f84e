"""Loader and ctypes declarations of libflursys_hip.so (the C ABI in
include/flursys_hip.h).

The shared library is built in-tree (flursys_amd/csrc/Makefile).  There is no
fallback: when the library is missing or no HIP device is present, every entry
point raises.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflursys_hip.so")

# status codes (fl_status)
OK = 0
ERR_CONFIG = 9
STATUS_NAMES = {
    1: "DimensionMismatch", 2: "InvalidCsr", 3: "IndexOutOfBounds", 4: "NonFiniteValue",
    5: "MissingDiagonal", 6: "ZeroDiagonal", 7: "Breakdown", 8: "SolverError",
    9: "ConfigError", 10: "InvalidArgument", 11: "HipError", 12: "OutOfMemory",
    13: "IoError", 14: "MeshError", 15: "PoissonError", 16: "SimpleError",
}
STOP_ABS_INF, STOP_REL_L2 = 0, 1
GUESS_PREVIOUS, GUESS_ZERO = 0, 1
PRECISION_FP64, PRECISION_MIXED = 0, 1
PRECOND_JACOBI, PRECOND_MULTIGRID, PRECOND_LEARNED = 0, 1, 2
SOLVER_PCG, SOLVER_SOR = 0, 1
LEARNED_CORRECTED, LEARNED_REPLACE = 0, 1
# fl_solve_report.fallback_reason
FALLBACK_NONE, FALLBACK_BUDGET, FALLBACK_STAGNATION, FALLBACK_BREAKDOWN = 0, 1, 2, 3


class FlursysError(RuntimeError):
    """A non-OK fl_status, with the context's error text."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.kind = STATUS_NAMES.get(status, str(status))
        self.message = message
        super().__init__(f"{self.kind}: {message}")


class SolveReport(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64), ("refinements", C.c_int64),
        ("fallback_iterations", C.c_int64), ("fallback", C.c_int32), ("warm_start", C.c_int32),
        ("residual_abs_inf", C.c_double), ("residual_l2", C.c_double),
        ("initial_l2", C.c_double), ("converged", C.c_int32), ("status", C.c_int32),
        ("warm_start_ratio", C.c_double), ("fallback_reason", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Cavity3DConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("length", C.c_double), ("height", C.c_double), ("depth", C.c_double),
        ("density", C.c_double), ("lid_velocity", C.c_double), ("reynolds", C.c_double),
        ("dt", C.c_double),
        ("pressure_max_iters", C.c_int64), ("pressure_tolerance", C.c_double),
        ("stop_rule", C.c_int32), ("initial_guess", C.c_int32), ("precision", C.c_int32),
        ("preconditioner", C.c_int32), ("inner_tolerance", C.c_double),
        ("pressure_solver", C.c_int32), ("reserved", C.c_int32), ("pressure_omega", C.c_double),
    ]


class Cavity3DStepReport(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("time", C.c_double),
        ("pressure_iterations", C.c_int64), ("pressure_refinements", C.c_int64),
        ("pressure_residual", C.c_double), ("pressure_residual_l2", C.c_double),
        ("max_divergence", C.c_double), ("divergence_l2", C.c_double),
        ("max_speed", C.c_double), ("pressure_converged", C.c_int32),
        ("pressure_fallback", C.c_int32), ("pressure_warm_start", C.c_int32),
        ("pressure_fallback_reason", C.c_int32), ("pressure_warm_ratio", C.c_double),
    ]


class Batch2DRunReport(C.Structure):
    _fields_ = [("steps", C.c_int64), ("time", C.c_double), ("pressure_iterations", C.c_int64),
                ("last_pressure_residual", C.c_double), ("max_divergence", C.c_double),
                ("max_speed", C.c_double), ("reached_end_time", C.c_int32),
                ("status", C.c_int32)]


GRAPH_NONE, GRAPH_WHOLE_STEP, GRAPH_SPLIT = 0, 1, 2


class Batch2DTuning(C.Structure):
    _fields_ = [("threads", C.c_int32), ("blocks", C.c_int32), ("reserved", C.c_int32 * 2),
                ("ms", C.c_double)]


class UNetConfig(C.Structure):
    _fields_ = [("n", C.c_int32), ("max_batch", C.c_int32), ("seed", C.c_uint64)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("bytes", C.c_double)]


class LinOpts(C.Structure):
    _fields_ = [("absolute_tolerance", C.c_double), ("relative_tolerance", C.c_double),
                ("max_iterations", C.c_int64)]


class LinReport(C.Structure):
    _fields_ = [("converged", C.c_int32), ("refinements", C.c_int32), ("iterations", C.c_int64),
                ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("relative_residual", C.c_double), ("device_ms", C.c_double),
                ("fallback", C.c_int32), ("reserved", C.c_int32)]


class ErrDetail(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("index", C.c_int64),
                ("expected", C.c_int64), ("actual", C.c_int64), ("row", C.c_int64),
                ("column", C.c_int64), ("context", C.c_char * 96)]


class Simple2DOptions(C.Structure):
    """fl_simple2d_options: SimpleOptions (src/simple.rs:93-131)."""
    _fields_ = [("viscosity", C.c_double), ("reference_cell", C.c_int64),
                ("max_outer_iterations", C.c_int64), ("continuity_absolute_tolerance", C.c_double),
                ("momentum_solver", LinOpts), ("pressure_solver", LinOpts),
                ("velocity_relaxation", C.c_double), ("pressure_relaxation", C.c_double)]


class Simple2DReport(C.Structure):
    """fl_simple2d_report: SimpleReport (src/simple.rs:133-142) and the error item."""
    _fields_ = [("converged", C.c_int32), ("error_kind", C.c_int32),
                ("outer_iterations", C.c_int64), ("initial_continuity_rms", C.c_double),
                ("final_continuity_rms", C.c_double), ("momentum", LinReport * 3),
                ("pressure", LinReport), ("error_index", C.c_int64), ("error_value", C.c_double),
                ("device_ms", C.c_double)]


class Grid2DConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("length", C.c_double), ("height", C.c_double),
        ("density", C.c_double), ("kinematic_viscosity", C.c_double), ("dt", C.c_double),
        ("t_end", C.c_double), ("pressure_max_iters", C.c_int64),
        ("pressure_tolerance", C.c_double), ("convection", C.c_int32),
        ("stop_rule", C.c_int32), ("initial_guess", C.c_int32), ("adaptive", C.c_int32),
        ("bc_kind", C.c_int32 * 4), ("bc_u", C.c_double * 4), ("bc_v", C.c_double * 4),
        ("outlet_pressure", C.c_double * 4), ("min_dt", C.c_double), ("max_dt", C.c_double),
        ("target_cfl", C.c_double), ("target_diffusion_number", C.c_double),
        ("max_growth_factor", C.c_double), ("precision", C.c_int32), ("reserved", C.c_int32),
        ("inner_tolerance", C.c_double), ("pressure_solver", C.c_int32),
        ("inlet_profile", C.c_int32), ("pressure_omega", C.c_double),
        ("inlet_step_height", C.c_double), ("coupling", C.c_int32), ("reserved2", C.c_int32),
        ("velocity_relaxation", C.c_double), ("pressure_relaxation", C.c_double),
    ]


class Grid2DStepReport(C.Structure):
    _fields_ = [
        ("iteration", C.c_int64), ("time", C.c_double), ("time_step", C.c_double),
        ("continuity_residual", C.c_double), ("pressure_residual", C.c_double),
        ("pressure_iterations", C.c_int64), ("momentum_residual", C.c_double),
        ("max_speed", C.c_double), ("momentum_cfl", C.c_double),
        ("viscous_diffusion_number", C.c_double), ("pressure_converged", C.c_int32),
        ("reached_end_time", C.c_int32),
    ]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_LIN =[_P, _P, _P, C.POINTER(LinOpts), C.POINTER(LinReport), C.POINTER(ErrDetail)]

# name -> (restype, argtypes); every symbol include/flursys_hip.h declares.
class HistoryRow(C.Structure):
    """fl_history_row: one history.csv row (incompressible.rs:1744-1767) plus
    the GPU path's cg_iters, hbm_GBps, roofline_frac and fallback_count."""
    _fields_ = [("step", C.c_int64), ("time", C.c_double), ("pressure_residual", C.c_double),
                ("pressure_iterations", C.c_int64)] + [
        (name, C.c_double) for name in (
            "max_divergence", "velocity_change", "max_speed", "cd", "cl", "reattachment",
            "momentum_cfl", "viscous_diffusion_number", "time_step")] + [
        ("cg_iters", C.c_int64), ("hbm_GBps", C.c_double), ("roofline_frac", C.c_double),
        ("fallback_count", C.c_int64)]


SIGNATURES = {
    "fl_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "fl_destroy": (None, [_P]),
    "fl_last_error": (C.c_char_p, [_P]),
    "fl_version": (C.c_char_p, []),
    "fl_probe_copy": (C.c_int, [_P, C.c_int64, C.c_int32, _D]),
    "fl_cavity3d_create": (C.c_int, [_P, C.POINTER(Cavity3DConfig), C.POINTER(_P)]),
    "fl_cavity3d_destroy": (None, [_P]),
    "fl_cavity3d_step": (C.c_int, [_P, C.POINTER(Cavity3DStepReport)]),
    "fl_cavity3d_run": (C.c_int, [_P, C.c_int64, C.POINTER(Cavity3DStepReport)]),
    "fl_cavity3d_pressure_solve": (C.c_int, [_P, _P, _P, C.POINTER(SolveReport)]),
    "fl_cavity3d_get_fields": (C.c_int, [_P, _P, _P, _P, _P]),
    "fl_cavity3d_cell_fields": (C.c_int, [_P, _P, _P, _P, _P]),
    "fl_mesh2d_create": (C.c_int, [_P, C.c_int64, _P, C.c_int64, _P, _P, C.POINTER(_P),
                                   C.POINTER(C.c_int32)]),
    "fl_mesh2d_destroy": (None, [_P]),
    "fl_mesh2d_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "fl_mesh2d_get": (C.c_int, [_P] + [_P] * 9),
    "fl_mesh2d_assemble_poisson": (C.c_int, [_P, C.c_double, _P, _P, _P, C.c_int32, C.c_int64,
                                             C.c_double, C.POINTER(_P), _P,
                                             C.POINTER(C.c_int32), _D]),
    "fl_csr_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                              C.POINTER(C.c_int64)]),
    "fl_csr_download": (C.c_int, [_P, _P, _P, _P]),
    "fl_simple2d_create": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, C.POINTER(_P), _P]),
    "fl_simple2d_destroy": (None, [_P]),
    "fl_simple2d_initial_face_flux": (C.c_int, [_P, _P, _P, _P, _P]),
    "fl_simple2d_set_state": (C.c_int, [_P, _P, _P, _P, _P]),
    "fl_simple2d_get_state": (C.c_int, [_P, _P, _P, _P, _P]),
    "fl_simple2d_solve": (C.c_int, [_P, _P, _P, _P]),
    "fl_write_field_csv": (C.c_int, [C.c_char_p, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                     _P, _P, _P, _P, _P, _P, C.c_char_p, C.c_int32]),
    "fl_write_vtk_2d": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int64, C.c_int64, C.c_double,
                                  C.c_double, _P, _P, _P, _P, _P, _P, C.c_char_p, C.c_int32]),
    "fl_write_vtk_3d": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int64, C.c_int64, C.c_int64,
                                  C.c_double, C.c_double, C.c_double, _P, _P, _P, _P,
                                  C.c_char_p, C.c_int32]),
    "fl_write_ppm": (C.c_int, [C.c_char_p, C.c_int64, C.c_int64, _P, _P, C.c_int32, C.c_char_p,
                               C.c_int32]),
    "fl_write_csv_columns": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int32, C.c_int64, _P,
                                       C.c_char_p, C.c_char_p, C.c_int32]),
    "fl_append_history": (C.c_int, [C.c_char_p, C.POINTER(HistoryRow), C.c_int32, C.c_char_p,
                                    C.c_int32]),
    "fl_cavity3d_set_fields": (C.c_int, [_P, _P, _P, _P, _P]),
    "fl_cavity3d_get_intermediate": (C.c_int, [_P, _P, _P, _P, _P]),
    "fl_cavity3d_set_timing": (C.c_int, [_P, C.c_int32, C.c_char_p]),
    "fl_cavity3d_kernel_stats": (C.c_int, [_P, C.c_int32, C.POINTER(KernelStat)]),
    "fl_grid2d_create": (C.c_int, [_P, C.POINTER(Grid2DConfig), _P, C.POINTER(_P)]),
    "fl_grid2d_destroy": (None, [_P]),
    "fl_grid2d_step": (C.c_int, [_P, C.POINTER(Grid2DStepReport)]),
    "fl_grid2d_pressure_solve": (C.c_int, [_P, _P, _P, C.POINTER(SolveReport)]),
    "fl_grid2d_get_fields": (C.c_int, [_P, _P, _P, _P]),
    "fl_grid2d_set_fields": (C.c_int, [_P, _P, _P, _P]),
    "fl_grid2d_cell_fields": (C.c_int, [_P, _P, _P, _P, _P]),
    "fl_batch2d_create": (C.c_int, [_P, C.c_int32, C.POINTER(Grid2DConfig), C.POINTER(_P)]),
    "fl_batch2d_destroy": (None, [_P]),
    "fl_batch2d_step": (C.c_int, [_P, C.POINTER(Grid2DStepReport)]),
    "fl_batch2d_get_fields": (C.c_int, [_P, C.c_int32, _P, _P, _P]),
    "fl_batch2d_run": (C.c_int, [_P, C.c_int64, C.c_int32, C.POINTER(Batch2DRunReport), _P, _P]),
    "fl_batch2d_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(Batch2DTuning), C.c_int32,
                                      C.POINTER(Batch2DTuning), C.POINTER(C.c_int32)]),
    "fl_batch2d_set_launch": (C.c_int, [_P, C.POINTER(Batch2DTuning)]),
    "fl_batch2d_graph_mode": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "fl_batch2d_get_launch": (C.c_int, [_P, C.POINTER(Batch2DTuning)]),
    "fl_csr_upload": (C.c_int,[_P, C.c_int64, C.c_int64, _P, _P, _P, C.c_int64,
                                C.POINTER(_P), C.POINTER(ErrDetail)]),
    "fl_csr_destroy": (None, [_P]),
    "fl_spmv": (C.c_int, [_P, _P, _P, C.POINTER(ErrDetail)]),
    "fl_cg": (C.c_int, _LIN),
    "fl_pcg_jacobi": (C.c_int, _LIN),
    "fl_bicgstab": (C.c_int, _LIN),
    "fl_pcg_jacobi_noalloc": (C.c_int, _LIN),
    "fl_csr_set_precision": (C.c_int, [_P, C.c_int32, C.c_double]),
    "fl_unet_create": (C.c_int, [_P, C.POINTER(UNetConfig), C.POINTER(_P)]),
    "fl_unet_destroy": (None, [_P]),
    "fl_unet_num_params": (C.c_int64, [_P]),
    "fl_unet_get_params": (C.c_int, [_P, _P]),
    "fl_unet_set_params": (C.c_int, [_P, _P]),
    "fl_unet_train_step": (C.c_int, [_P, C.c_int32, _P, _P, C.c_double, C.c_float,
                                     C.POINTER(C.c_float)]),
    "fl_unet_infer": (C.c_int, [_P, C.c_int32, _P, C.c_double, _P]),
    "fl_unet_forward_raw": (C.c_int, [_P, C.c_int32, _P, _P]),
    "fl_unet_grads": (C.c_int, [_P, C.c_int32, _P, _P, C.c_double, _P, C.POINTER(C.c_float)]),
    "fl_probe_mfma_bf16": (C.c_int, [_P, _P, _P, _P]),
    "fl_cavity3d_set_unet": (C.c_int, [_P, _P]),
    "fl_cavity3d_set_learned_mode": (C.c_int, [_P, C.c_int32]),
    "fl_cavity3d_coarse_pair": (C.c_int, [_P, C.c_int32, _P, _P, C.POINTER(C.c_double)]),
    "fl_cavity3d_capture_coarse": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "fl_cavity3d_captured_rhs": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
}

_lib = None


def load():
    """Load libflursys_hip.so and declare its signatures (raises if absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FlursysError(
                11, f"{LIB_PATH} is not built; run __graft_entry__.build() "
                "(make -C flursys_amd/csrc)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status: int, ctx=None) -> None:
    if status != OK:
        msg = ""
        if ctx is not None:
            raw = load().fl_last_error(ctx)
            msg = raw.decode() if raw else ""
        raise FlursysError(status, msg)


def ptr(array):
    """Raw data pointer of a C-contiguous numpy array (None passes NULL)."""
    if array is None:
        return None
    if not array.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return array.ctypes.data


class Context:
    """fl_ctx: a device binding with its own HIP stream."""

    def __init__(self, device: int = 0):
        lib = load()
        h = _P()
        st = lib.fl_create(device, C.byref(h))
        if st != OK:
            raise FlursysError(st, f"fl_create(device={device}) failed (is a GPU present?)")
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            load().fl_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
