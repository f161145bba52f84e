"""The allocation-free Jacobi-PCG entry point (fl_pcg_jacobi_noalloc) against
fl_pcg_jacobi through the C ABI: two calls in a row, the second on a smaller
system, must give bitwise the same solution and report (the two share their
kernels and launch shapes; only the work buffers' home differs)."""

import ctypes as C

import numpy as np
import pytest

import flursys_amd.linear as FL
from flursys_amd import _lib
from flursys_amd._lib import ErrDetail, LinOpts, LinReport, ptr
import oracle as O

pytestmark = pytest.mark.gpu


def _noalloc(lib):
    return lib.fl_pcg_jacobi_noalloc


def _system(m, seed):
    mesh = O.mesh_poisson(m, 0.25, seed, 0)
    n = len(mesh["volumes"])
    src = np.random.default_rng(seed).standard_normal(n)
    src -= (src * mesh["volumes"]).sum() / mesh["volumes"].sum()
    rhs = src * mesh["volumes"]
    rhs[0] = 0.0
    return mesh, rhs


def _call(fn, A, rhs, x, opts):
    rep, det = LinReport(), ErrDetail()
    st = fn(A.handle, ptr(rhs), ptr(x), C.byref(opts), C.byref(rep), C.byref(det))
    return st, rep


def test_noalloc_matches_allocating_entry_point(gpu_ctx):
    lib = _lib.load()
    fn = _noalloc(lib)
    opts = LinOpts(0.0, 1e-10, 20000)
    for m, seed in ((120, 3), (70, 4)):  # second call: a smaller system
        print(f"[gpu] no-alloc PCG vs fl_pcg_jacobi, {2 * m * m} cells", flush=True)
        mesh, rhs = _system(m, seed)
        n = len(rhs)
        A = FL.CsrMatrix(n, n, mesh["offsets"], mesh["cols"], mesh["vals"], ctx=gpu_ctx)
        x0 = np.zeros(n)
        x0[1::7] = 0.25  # a non-zero initial guess, honoured by both
        xa = x0.copy()
        st_a, rep_a = _call(lib.fl_pcg_jacobi, A, rhs, xa, opts)
        xb = x0.copy()
        st_b, rep_b = _call(fn, A, rhs, xb, opts)
        assert st_a == 0 and st_b == 0
        assert rep_a.converged == rep_b.converged == 1
        assert rep_a.iterations == rep_b.iterations > 0
        assert rep_a.final_residual == rep_b.final_residual
        assert rep_a.initial_residual == rep_b.initial_residual
        assert np.array_equal(xa, xb)
        A.close()


def test_noalloc_rejects_oversized_matrices(gpu_ctx):
    lib = _lib.load()
    fn = _noalloc(lib)
    n = (1 << 20) + 1
    A = FL.CsrMatrix(n, n, np.arange(n + 1), np.arange(n), np.ones(n), ctx=gpu_ctx)
    rhs, x = np.ones(n), np.zeros(n)
    st, _ = _call(fn, A, rhs, x, LinOpts(0.0, 1e-10, 10))
    assert st == 10  # FL_ERR_INVALID_ARGUMENT
    assert b"allocation-free" in lib.fl_last_error(gpu_ctx.handle)
    assert not x.any()
    A.close()


def test_noalloc_host_mirror(gpu_ctx):
    """The host mirror's pcg_noalloc keeps pcg's contract (report, in-place
    solution) on the reference's known-answer system (linear/mod.rs:184-219)."""
    a = np.array([[4.0, 1.0], [1.0, 3.0]])
    A = FL.CsrMatrix(2, 2, [0, 2, 4], [0, 1, 0, 1], a.ravel(), ctx=gpu_ctx)
    x = np.zeros(2)
    rep = FL.pcg_noalloc(A, a @ np.array([1 / 11, 7 / 11]), x,
                         FL.LinearSolverOptions(1e-12, 1e-12, 1000))
    assert rep.converged() and np.allclose(x, [1 / 11, 7 / 11], atol=1e-12, rtol=0)
    A.close()
