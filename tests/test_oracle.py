"""Pins the CPU oracle (CPU only, no GPU).

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a sign, a transposed index, a wrong stop rule) fails it:
* the reference's own known-answer tests (src/linear/mod.rs:97-285,
  src/solver/incompressible.rs:2418-2583, tests/smoke.rs:57-156), restated as
  data in tests/golden/csr_known_answers.json or inline;
* closed forms: DCT-II exact solutions of the pure-Neumann operators
  (tests/golden/neumann_dct.npz), scipy spsolve on the small jittered mesh
  (tests/golden/mesh_small.npz);
* invariants: operator symmetry, telescoping sum of the cavity RHS, gauge.
"""

import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def kat():
    with open(os.path.join(GOLD, "csr_known_answers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def dct():
    return np.load(os.path.join(GOLD, "neumann_dct.npz"))


# --------------------------------------------------------------- CSR (linear/)
def test_csr_builder_known_answer(kat):
    k = kat["builder"]
    A = O.Csr.from_triplets(3, 3, k["rows"], k["cols"], k["vals"])
    assert A.offsets.tolist() == k["offsets"]
    assert A.cols.tolist() == k["out_cols"]
    assert A.vals.tolist() == k["out_vals"]
    assert (1.0 / A.jacobi_inverse()).tolist() == k["diagonal"]


def test_duplicates_spmv_residual(kat):
    k = kat["duplicates"]
    A = O.Csr.from_triplets(2, 2, k["rows"], k["cols"], k["vals"])
    assert A.nnz == k["nnz"]
    assert A.vals[0] == k["a00"]
    for x, y in k["spmv"]:
        assert A.spmv(x).tolist() == y
    r = np.array(k["residual"]["rhs"]) - A.spmv(k["residual"]["x"])
    assert r.tolist() == k["residual"]["r"]


def test_exact_zero_sums_are_dropped():
    A = O.Csr.from_triplets(2, 2, [0, 0, 1, 1], [0, 1, 1, 1], [1.0, 0.0, 2.0, -2.0])
    assert A.offsets.tolist() == [0, 1, 1]


def test_malformed_csr_rejected():
    with pytest.raises(O.LinearAlgebraError) as e:
        O.Csr(1, 1, [1, 1], [], [])
    assert e.value.kind == "InvalidCsr"
    with pytest.raises(O.LinearAlgebraError) as e:
        O.Csr(1, 1, [0, 1], [1], [1.0])
    assert e.value.kind == "InvalidCsr"
    with pytest.raises(O.LinearAlgebraError) as e:
        O.Csr(1, 1, [0, 1], [0], [np.nan])
    assert e.value.kind == "NonFiniteValue"
    with pytest.raises(O.LinearAlgebraError) as e:
        O.Csr(2, 2, [0, 2, 2], [1, 0], [1.0, 1.0])
    assert e.value.kind == "InvalidCsr"


def test_jacobi_missing_and_zero_diagonal():
    missing = O.Csr.from_dense([[0.0, 1.0], [1.0, 1.0]])
    with pytest.raises(O.LinearAlgebraError) as e:
        missing.jacobi_inverse()
    assert e.value.kind == "MissingDiagonal" and e.value.row == 0
    zero = O.Csr(2, 2, [0, 2, 4], [0, 1, 0, 1], [0.0, 1.0, 1.0, 2.0])
    with pytest.raises(O.LinearAlgebraError) as e:
        zero.jacobi_inverse()
    assert e.value.kind == "ZeroDiagonal"


def test_cg_pcg_bicgstab_known_solutions(kat):
    o = kat["options"]
    a = np.array(kat["spd"]["a"])
    ex = np.array(kat["spd"]["x"])
    A = O.Csr.from_dense(a)
    x = np.zeros(2)
    rc = O.cg(A, a @ ex, x, o["atol"], o["rtol"], o["max_iter"])
    assert rc.converged and np.allclose(x, ex, atol=1e-10, rtol=0)
    assert rc.final_residual <= o["atol"] + o["rtol"] * rc.initial_residual
    x = np.zeros(2)
    rp = O.pcg(A, a @ ex, x, o["atol"], o["rtol"], o["max_iter"])
    assert rp.converged and rp.iterations <= rc.iterations
    assert np.allclose(x, ex, atol=1e-10, rtol=0)
    a = np.array(kat["nonsymmetric"]["a"])
    ex = np.array(kat["nonsymmetric"]["x"])
    x = np.zeros(3)
    rb = O.bicgstab(O.Csr.from_dense(a), a @ ex, x, o["atol"], o["rtol"], o["max_iter"])
    assert rb.converged and np.allclose(x, ex, atol=1e-10, rtol=0)


def _poisson1d(n):
    rows, cols, vals = [], [], []
    for r in range(n):
        rows.append(r); cols.append(r); vals.append(2.0)
        if r > 0:
            rows.append(r); cols.append(r - 1); vals.append(-1.0)
        if r + 1 < n:
            rows.append(r); cols.append(r + 1); vals.append(-1.0)
    return O.Csr.from_triplets(n, n, rows, cols, vals)


def test_solver_edge_cases(kat):
    o = kat["options"]
    n = kat["poisson1d"]["n"]
    A = _poisson1d(n)
    ex = (np.arange(n) + 1.0) / n
    rhs = A.spmv(ex)
    x = np.zeros(n)
    assert O.cg(A, rhs, x, o["atol"], o["rtol"], o["max_iter"]).converged
    assert np.allclose(x, ex, atol=1e-10, rtol=0)
    # honoured non-zero initial guess
    r = O.bicgstab(A, np.zeros(n), x.copy(), o["atol"], o["rtol"], o["max_iter"])
    assert r.iterations > 0
    # exact initial guess -> zero iterations
    assert O.cg(A, rhs, ex.copy(), o["atol"], o["rtol"], o["max_iter"]).iterations == 0
    r = O.bicgstab(A, np.zeros(n), np.zeros(n), o["atol"], o["rtol"], o["max_iter"])
    assert r.converged and r.iterations == 0
    # MaxIterations is a status, not an error
    r = O.cg(A, rhs, np.zeros(n), o["atol"], o["rtol"], 1)
    assert not r.converged and r.iterations == 1
    with pytest.raises(O.LinearAlgebraError) as e:
        O.bicgstab(O.Csr.from_dense([[0.0]]), np.array([1.0]), np.zeros(1), o["atol"],
                   o["rtol"], o["max_iter"])
    assert e.value.kind == "Breakdown"
    with pytest.raises(O.LinearAlgebraError) as e:
        O.cg(A, np.full(n, np.inf), np.zeros(n), o["atol"], o["rtol"], o["max_iter"])
    assert e.value.kind == "NonFiniteValue"


# ------------------------------------------------- stability numbers and dt
def test_stability_numbers_known_values():
    cfl, dnu = O.stability_numbers(0.01, 0.1, 0.5, 0.25, 2.0, 1.0)
    assert abs(cfl - 0.08) < 1e-12 and abs(dnu - 0.02) < 1e-12
    cfl, dnu = O.stability_numbers(0.1, 0.25, 0.5, 0.25, 0.0, 0.0)
    assert cfl == 0.0 and abs(dnu - 0.5) < 1e-12
    a, _ = O.stability_numbers(0.1, 0.2, 0.5, 0.25, 2.0, 1.0)
    b, _ = O.stability_numbers(0.1, 0.2, 0.25, 0.5, 1.0, 2.0)
    assert abs(a - b) < 1e-12


def test_adaptive_time_step_known_values():
    f = O.adaptive_time_step
    assert abs(f(0.1, 2.0, 1.0, 0.5, 0.25, 1e-3, min_dt=1e-3, max_dt=1.0, target_cfl=0.5,
                 target_diffusion=0.5, max_growth=10.0) - 0.0625) < 1e-12
    assert abs(f(0.5, 0.0, 0.0, 0.5, 0.25, 0.1, min_dt=1e-3, max_dt=1.0, target_cfl=1.0,
                 target_diffusion=0.25, max_growth=10.0) - 0.125) < 1e-12
    assert abs(f(0.5, 0.0, 0.0, 0.5, 0.25, 0.1, 0.4, min_dt=1e-3, max_dt=1.0, target_cfl=1.0,
                 target_diffusion=0.25, max_growth=10.0) - 0.03125) < 1e-12
    assert abs(f(0.1, 0.0, 0.0, 1.0, 1.0, 1e-3, min_dt=1e-3, max_dt=0.8, target_cfl=1.0,
                 target_diffusion=0.25, max_growth=1.1) - 0.11) < 1e-12
    with pytest.raises(RuntimeError, match="required stable time step"):
        f(0.5, 10.0, 10.0, 0.1, 0.1, 1.0, min_dt=0.2, max_dt=1.0, target_cfl=0.5,
          target_diffusion=0.25, max_growth=1.1)


# ------------------------------------------------------------- closed forms
@pytest.mark.parametrize("key", ["16x16", "32x32", "24x40"])
def test_2d_pcg_matches_dct_solution(dct, key):
    ny, nx = map(int, key.split("x"))
    s = O.Solver2D(nx, ny, pressure_tolerance=1e-13, stop_rule=O.STOP_REL_L2,
                   initial_guess=O.GUESS_ZERO, pressure_max_iters=10000)
    s.field("rhs")[:] = -dct[f"b2d_{key}"]  # A p = b with b = -rhs
    rep = s.solve_pressure()
    assert rep.converged
    ex = dct[f"p2d_{key}"]
    p = s.field("p")
    assert np.abs(p - ex).max() <= 1e-10 * np.abs(ex).max()
    assert abs(p.mean()) < 1e-12  # zero-mean gauge


@pytest.mark.parametrize("key", ["8x8x8", "16x16x16", "12x10x14"])
def test_3d_pinned_pcg_matches_dct_solution(dct, key):
    nz, ny, nx = map(int, key.split("x"))
    c = O.Cavity3D(nx, ny, nz, pressure_tolerance=1e-13, stop_rule=O.STOP_REL_L2,
                   initial_guess=O.GUESS_ZERO, pressure_max_iters=10000)
    c.field("rhs")[:] = dct[f"rhs3d_{key}"]
    rep = c.solve_pressure()
    assert rep.converged
    ex = dct[f"p3d_{key}"]
    p = c.field("p")
    assert np.abs(p - ex).max() <= 1e-10 * np.abs(ex).max()
    assert p[0, 0, 0] == 0.0


def test_3d_sor_reaches_the_pcg_solution(dct):
    key = "8x8x8"
    c = O.Cavity3D(8, solver=O.SOLVER_SOR, pressure_tolerance=1e-11, pressure_max_iters=20000)
    c.field("rhs")[:] = dct[f"rhs3d_{key}"]
    rep = c.solve_pressure()
    assert rep.converged
    ex = dct[f"p3d_{key}"]
    assert np.abs(c.field("p") - ex).max() <= 1e-9 * np.abs(ex).max()


def test_operators_symmetric_and_consistent():
    rng = np.random.default_rng(1)
    c = O.Cavity3D(6, 5, 7)
    x, y = rng.standard_normal((2, 7 * 5 * 6))
    x[0] = y[0] = 0.0
    assert abs(y @ c.apply_operator(x) - x @ c.apply_operator(y)) < 1e-9 * abs(x @ c.apply_operator(x))
    # -Lap of a constant vanishes away from the pinned cell's neighbours
    ones = np.ones(7 * 5 * 6)
    ones[0] = 0.0
    Ao = c.apply_operator(ones).reshape(7, 5, 6)
    assert np.abs(Ao[3:, 3:, 3:]).max() == 0.0
    s = O.Solver2D(9, 7)
    x, y = rng.standard_normal((2, 63))
    assert abs(y @ s.apply_operator(x) - x @ s.apply_operator(y)) < 1e-9 * abs(x @ s.apply_operator(x))
    assert np.abs(s.apply_operator(np.ones(63))).max() == 0.0


def test_mesh_poisson_against_spsolve():
    m = np.load(os.path.join(GOLD, "mesh_small.npz"))
    n = len(m["volumes"])
    A = O.Csr(n, n, m["offsets"], m["cols"], m["vals"])
    assert abs(m["volumes"].sum() - 1.0) < 1e-12
    fresh = O.mesh_poisson(20, 0.3, 0, 0)
    assert np.array_equal(fresh["offsets"], m["offsets"])
    assert np.array_equal(fresh["vals"], m["vals"])
    x = np.zeros(n)
    r = O.cg(A, m["rhs"], x, 0.0, 1e-13, 5000)
    assert r.converged
    assert np.abs(x - m["x"]).max() <= 1e-9 * np.abs(m["x"]).max()
    x = np.zeros(n)
    r = O.pcg(A, m["rhs"], x, 0.0, 1e-13, 5000)
    assert r.converged and np.abs(x - m["x"]).max() <= 1e-9 * np.abs(m["x"]).max()


# ------------------------------------------------------------ whole steps
def test_2d_cavity_smoke_threshold():
    """tests/smoke.rs:57-69: 16^2 cavity, one step, max divergence < 1e-5."""
    s = O.Solver2D(16, dt=1e-3, t_end=1e-3, pressure_max_iters=800, pressure_tolerance=1e-6)
    d, rep = s.step()
    assert np.isfinite(d.max_divergence) and d.max_divergence < 1e-5
    assert abs(s.time - 1e-3) < 1e-12
    with pytest.raises(RuntimeError, match="already reached t_end"):
        s.step()


def test_3d_cavity_smoke_threshold():
    """tests/smoke.rs:113-130: 8^3, two steps, divergence < 1e-5, speed > 0."""
    for solver in (O.SOLVER_SOR, O.SOLVER_PCG):
        c = O.Cavity3D(8, dt=5e-4, solver=solver)
        for _ in range(2):
            c.step()
        assert c.max_divergence() < 1e-5 and c.max_speed() > 0.0


def test_threads_agree():
    """tests/smoke.rs:132-156: 1 vs 4 threads agree to 1e-10."""
    out = []
    for threads in (1, 4):
        s = O.Solver2D(32, dt=1e-3, t_end=3e-3, threads=threads)
        for _ in range(3):
            d, rep = s.step()
        out.append((d.max_divergence, d.pressure_residual))
    assert abs(out[0][0] - out[1][0]) < 1e-10 and abs(out[0][1] - out[1][1]) < 1e-10


def test_closed_cavity_rhs_telescopes():
    c = O.Cavity3D(10, 8, 6, reynolds=400.0)
    for _ in range(3):
        c.step()
    c.predict()
    c.build_rhs()
    us, vs, ws = c.field("us"), c.field("vs"), c.field("ws")
    div = ((us[:, :, 1:] - us[:, :, :-1]) * 10 + (vs[:, 1:, :] - vs[:, :-1, :]) * 8
           + (ws[1:] - ws[:-1]) * 6)
    # boundary-normal faces are zero, so the divergence telescopes to zero
    assert us[:, :, 0].any() == 0 and ws[0].any() == 0
    assert abs(div.sum()) < 1e-10 * np.abs(div).sum()
    rhs = c.field("rhs")
    expect = div / 5e-4
    assert rhs[0, 0, 0] == 0.0
    assert np.allclose(rhs.ravel()[1:], expect.ravel()[1:], rtol=1e-9,
                       atol=1e-9 * np.abs(expect).max())


def test_config1_as_written_is_unstable():
    """128^2, Re 100, dt = 0.25 h: viscous number 0.64 > 0.5 (the reference's
    explicit limit, incompressible.rs:411-413), so the run must blow up."""
    s = O.Solver2D(128, nu=0.01, dt=0.25 / 128, t_end=1e9, pressure_tolerance=1e-8,
                   stop_rule=O.STOP_REL_L2, initial_guess=O.GUESS_ZERO,
                   pressure_max_iters=5000)
    with pytest.raises(RuntimeError, match="Non-finite"):
        for _ in range(200):
            d, _ = s.step()
            assert abs(d.viscous_diffusion_number - 0.64) < 1e-12
