"""Host-side mirror of the reference's sparse linear algebra over the C ABI.

Names and behaviour follow src/linear/: ``CsrMatrix`` (csr.rs:4-125, validated
like validate_csr :195-251), ``cg`` / ``pcg`` / ``bicgstab``
(solver.rs:43-256) returning a ``LinearSolveReport`` (:29-41) with
MaxIterations as a status, and ``LinearAlgebraError`` (mod.rs:17-51) raised
with the reference's kinds and static texts.  The solution array is updated
in place and its initial contents are the initial guess.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ErrDetail, LinOpts, LinReport, load, ptr


class LinearAlgebraError(_lib.FlursysError):
    """LinearAlgebraError: `kind` is the variant name, `context` the static str."""

    def __init__(self, status, detail: ErrDetail, message: str = ""):
        super().__init__(status, message)
        self.context = detail.context.decode()
        self.index = detail.index
        self.row = detail.row
        self.column = detail.column
        self.expected = detail.expected
        self.actual = detail.actual


def _raise(status, detail, ctx):
    raw = load().fl_last_error(ctx)
    raise LinearAlgebraError(status, detail, raw.decode() if raw else "")


@dataclass
class LinearSolverOptions:
    """LinearSolverOptions (solver.rs:5-20)."""

    absolute_tolerance: float = 1.0e-10
    relative_tolerance: float = 1.0e-8
    max_iterations: int = 1000


@dataclass
class LinearSolveReport:
    """LinearSolveReport (solver.rs:29-41); status "Converged" / "MaxIterations"."""

    status: str
    iterations: int
    initial_residual: float
    final_residual: float
    relative_residual: float
    device_ms: float
    refinements: int = 0  # mixed precision: FP64 residual rounds
    fallback: bool = False  # mixed precision: finished on the FP64 solver

    def converged(self) -> bool:
        return self.status == "Converged"


class CsrMatrix:
    """CsrMatrix resident on the GPU (int32 indices, FP64 values)."""

    def __init__(self, nrows, ncols, row_offsets, column_indices, values,
                 ctx: _lib.Context | None = None):
        self.ctx = ctx or _lib.default_context()
        off = np.ascontiguousarray(row_offsets, dtype=np.int64)
        cols = np.ascontiguousarray(column_indices, dtype=np.int64)
        vals = np.ascontiguousarray(values, dtype=np.float64)
        det = ErrDetail()
        if off.size != nrows + 1:
            det.kind = 2
            det.context = b"row_offsets length must equal nrows + 1"
            raise LinearAlgebraError(2, det, "row_offsets length must equal nrows + 1")
        if cols.size != vals.size:
            det.kind = 2
            det.context = b"column and value counts must match"
            raise LinearAlgebraError(2, det, "column and value counts must match")
        h = C.c_void_p()
        st = load().fl_csr_upload(self.ctx.handle, nrows, ncols, ptr(off), ptr(cols),
                                  ptr(vals), vals.size, C.byref(h), C.byref(det))
        if st != 0:
            _raise(st, det, self.ctx.handle)
        self.handle = h
        self.nrows, self.ncols, self.nnz = nrows, ncols, vals.size

    @classmethod
    def adopt(cls, handle, ctx):
        """Wraps a matrix the device built (fl_mesh2d_assemble_poisson)."""
        m = cls.__new__(cls)
        m.ctx, m.handle = ctx, handle
        n, nc, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        load().fl_csr_info(handle, C.byref(n), C.byref(nc), C.byref(nnz))
        m.nrows, m.ncols, m.nnz = n.value, nc.value, nnz.value
        return m

    def arrays(self):
        """Host copies (row_offsets, column_indices, values)."""
        off = np.empty(self.nrows + 1, np.int64)
        cols = np.empty(max(self.nnz, 1), np.int64)
        vals = np.empty(max(self.nnz, 1), np.float64)
        _lib.check(load().fl_csr_download(self.handle, ptr(off), ptr(cols), ptr(vals)),
                   self.ctx.handle)
        return off, cols[:self.nnz], vals[:self.nnz]

    def close(self):
        if getattr(self, "handle", None):
            load().fl_csr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, precision: int, inner_tolerance: float = 0.0):
        """FL_PRECISION_FP64 (the reference's solvers) or FL_PRECISION_MIXED
        (FP32 solver inside FP64 iterative refinement) for this matrix's
        cg / pcg / bicgstab (no counterpart in the reference)."""
        _lib.check(load().fl_csr_set_precision(self.handle, int(precision), float(inner_tolerance)),
                   self.ctx.handle)

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size != self.ncols:
            det = ErrDetail()
            det.expected, det.actual = self.ncols, x.size
            raise LinearAlgebraError(1, det, "dimension mismatch")
        y = np.empty(self.nrows)
        det = ErrDetail()
        st = load().fl_spmv(self.handle, ptr(x), ptr(y), C.byref(det))
        if st != 0:
            _raise(st, det, self.ctx.handle)
        return y


def _solve(fn, matrix: CsrMatrix, rhs, solution, options):
    options = options or LinearSolverOptions()
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    for arr, expected in ((rhs, matrix.nrows), (solution, matrix.ncols)):
        if arr.size != expected:
            det = ErrDetail()
            det.expected, det.actual = expected, arr.size
            raise LinearAlgebraError(1, det, "dimension mismatch")
    if not (solution.dtype == np.float64 and solution.flags["C_CONTIGUOUS"]):
        raise ValueError("solution must be a C-contiguous float64 array")
    o = LinOpts(options.absolute_tolerance, options.relative_tolerance, options.max_iterations)
    rep, det = LinReport(), ErrDetail()
    st = getattr(load(), fn)(matrix.handle, ptr(rhs), ptr(solution), C.byref(o), C.byref(rep),
                             C.byref(det))
    if st != 0:
        _raise(st, det, matrix.ctx.handle)
    return LinearSolveReport("Converged" if rep.converged else "MaxIterations", rep.iterations,
                             rep.initial_residual, rep.final_residual, rep.relative_residual,
                             rep.device_ms, rep.refinements, bool(rep.fallback))


def cg(matrix, rhs, solution, options=None) -> LinearSolveReport:
    """cg (solver.rs:43-50)."""
    return _solve("fl_cg", matrix, rhs, solution, options)


def pcg(matrix, rhs, solution, options=None) -> LinearSolveReport:
    """pcg with JacobiPreconditioner::new(matrix) (solver.rs:52-60)."""
    return _solve("fl_pcg_jacobi", matrix, rhs, solution, options)


def bicgstab(matrix, rhs, solution, options=None) -> LinearSolveReport:
    """bicgstab (solver.rs:155-256)."""
    return _solve("fl_bicgstab", matrix, rhs, solution, options)


def pcg_noalloc(matrix, rhs, solution, options=None) -> LinearSolveReport:
    """pcg on the library's static per-GPU buffers: no device allocation
    (no reference counterpart; at most 1048576 rows; calls that overlap on
    one GPU race)."""
    return _solve("fl_pcg_jacobi_noalloc", matrix, rhs, solution, options)
