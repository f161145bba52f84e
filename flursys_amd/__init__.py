"""MI355X-native pressure-projection path of FLURSYS.

The numerics run in libflursys_hip.so (hand-written gfx950 HIP kernels behind
the C ABI in include/flursys_hip.h); this package is the host-side mirror of
the reference's solver interfaces over that ABI.
"""

from ._lib import (FALLBACK_BREAKDOWN, FALLBACK_BUDGET, FALLBACK_NONE, FALLBACK_STAGNATION,
                   GUESS_PREVIOUS, GUESS_ZERO, LEARNED_CORRECTED, LEARNED_REPLACE,
                   PRECISION_FP64, PRECISION_MIXED,
                   PRECOND_JACOBI, PRECOND_LEARNED, PRECOND_MULTIGRID, SOLVER_PCG, SOLVER_SOR,
                   STOP_ABS_INF, STOP_REL_L2, Context, FlursysError, default_context, load)
from .cavity3d import LidDrivenCavity3DConfig, LidDrivenCavity3DSolver, RunSummary3D

__all__ = [
    "Context", "FlursysError", "default_context", "load",
    "LidDrivenCavity3DConfig", "LidDrivenCavity3DSolver", "RunSummary3D",
    "STOP_ABS_INF", "STOP_REL_L2", "GUESS_PREVIOUS", "GUESS_ZERO",
    "PRECISION_FP64", "PRECISION_MIXED", "PRECOND_JACOBI", "PRECOND_MULTIGRID",
    "PRECOND_LEARNED", "SOLVER_PCG", "SOLVER_SOR",
    "FALLBACK_NONE", "FALLBACK_BUDGET", "FALLBACK_STAGNATION", "FALLBACK_BREAKDOWN",
    "LEARNED_CORRECTED", "LEARNED_REPLACE",
]
