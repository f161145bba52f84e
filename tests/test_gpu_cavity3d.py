"""GPU parity of the 3D cavity path (through the C ABI) against the oracle.

Bars (FP64 path):
* pressure solve: same iteration count as the oracle's pinned PCG to within
  1 iteration, p within 1e-9 of max|p| (the two differ only by FMA
  contraction and reduction order), and within 1e-10 of the DCT closed form
  when solved to rel 1e-13;
* whole steps: u, v, w, p within 1e-9 relative (north_star asks 1e-6 in FP64
  mode), u* and rhs within 1e-12;
* smoke thresholds of tests/smoke.rs:113-130;
* size-independent properties at 128^3: true relative residual and gauge.
"""

import os

import numpy as np
import pytest

import flursys_amd as fa
import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def make(ctx, nx, ny, nz, **kw):
    cfg = fa.LidDrivenCavity3DConfig(nx=nx, ny=ny, nz=nz, **kw)
    return fa.LidDrivenCavity3DSolver(cfg, ctx)


@pytest.mark.parametrize("key", ["8x8x8", "16x16x16", "12x10x14"])
def test_pressure_solve_matches_dct(gpu_ctx, key):
    d = np.load(os.path.join(GOLD, "neumann_dct.npz"))
    nz, ny, nx = map(int, key.split("x"))
    print(f"[gpu] pressure solve vs DCT {key}", flush=True)
    s = make(gpu_ctx, nx, ny, nz, pressure_tolerance=1e-13, stop_rule=fa.STOP_REL_L2,
             initial_guess=fa.GUESS_ZERO, pressure_max_iters=10000)
    p = np.zeros((nz, ny, nx))
    rep = s.pressure_solve(d[f"rhs3d_{key}"], p)
    assert rep.converged and rep.status == 0
    assert rel(p, d[f"p3d_{key}"]) < 1e-10
    assert p[0, 0, 0] == 0.0


@pytest.mark.parametrize("stop,tol,guess", [(fa.STOP_ABS_INF, 1e-6, fa.GUESS_PREVIOUS),
                                            (fa.STOP_REL_L2, 1e-8, fa.GUESS_ZERO)])
# (130, 7, 18) and (260, 6, 35): two and three x-tiles of the plane-tile passes
# (128 FP64 cells wide) with a partial last tile, a partial 4-row y tile and a
# partial 16-plane z chunk; (33, 20, 9) has odd x and takes the column march.
@pytest.mark.parametrize("shape", [(16, 16, 16), (14, 10, 12), (33, 20, 9), (130, 7, 18),
                                   (260, 6, 35)])
def test_pressure_solve_matches_oracle(gpu_ctx, shape, stop, tol, guess):
    nx, ny, nz = shape
    print(f"[gpu] pressure solve vs oracle {shape} stop={stop}", flush=True)
    rng = np.random.default_rng(7)
    rhs = rng.standard_normal((nz, ny, nx)) * 50.0
    rhs[0, 0, 0] = 0.0
    p0 = rng.standard_normal((nz, ny, nx)) * 0.01
    p0[0, 0, 0] = 0.0
    o = O.Cavity3D(nx, ny, nz, pressure_tolerance=tol, stop_rule=stop, initial_guess=guess,
                   pressure_max_iters=5000)
    o.field("rhs")[:] = rhs
    o.field("p")[:] = p0
    orep = o.solve_pressure()
    s = make(gpu_ctx, nx, ny, nz, pressure_tolerance=tol, stop_rule=stop, initial_guess=guess,
             pressure_max_iters=5000)
    p = p0.copy()
    rep = s.pressure_solve(rhs, p)
    assert rep.converged and orep.converged
    assert abs(rep.iterations - orep.iterations) <= 1
    assert rel(p, o.field("p")) < 1e-9
    assert abs(rep.initial_l2 - orep.initial_l2) <= 1e-12 * orep.initial_l2


@pytest.mark.parametrize("shape,re", [((16, 16, 16), 400.0), ((12, 10, 14), 1000.0)])
@pytest.mark.parametrize("stop,tol,guess", [(fa.STOP_ABS_INF, 1e-6, fa.GUESS_PREVIOUS),
                                            (fa.STOP_REL_L2, 1e-8, fa.GUESS_ZERO)])
def test_steps_match_oracle(gpu_ctx, shape, re, stop, tol, guess):
    nx, ny, nz = shape
    print(f"[gpu] cavity steps vs oracle {shape} Re={re} stop={stop}", flush=True)
    kw = dict(reynolds=re, dt=5e-4, pressure_tolerance=tol, stop_rule=stop,
              initial_guess=guess, pressure_max_iters=5000)
    o = O.Cavity3D(nx, ny, nz, **kw)
    s = make(gpu_ctx, nx, ny, nz, **kw)
    for step in range(6):
        orep = o.step()
        rep = s.advance_one_step()
        assert rep.pressure_converged
        assert abs(rep.pressure_iterations - orep.iterations) <= 1
    f = s.fields()
    # velocity components share one scale (w is ~1e-10 of u in the first
    # steps, so a per-component relative bar would measure round-off of 0)
    uscale = max(np.abs(o.field(c)).max() for c in ("u", "v", "w"))
    for name in ("u", "v", "w"):
        assert np.abs(f[name] - o.field(name)).max() < 1e-9 * uscale, name
    assert rel(f["p"], o.field("p")) < 1e-9
    inter = s.intermediate()
    sscale = max(np.abs(o.field(c)).max() for c in ("us", "vs", "ws"))
    for name in ("us", "vs", "ws"):
        assert np.abs(inter[name] - o.field(name)).max() < 1e-12 * sscale, name
    assert rel(inter["rhs"], o.field("rhs")) < 1e-11
    assert abs(rep.max_divergence - o.max_divergence()) <= 1e-9 * max(o.max_divergence(), 1e-12) + 1e-12
    assert abs(rep.max_speed - o.max_speed()) <= 1e-12 * o.max_speed()


def test_smoke_threshold(gpu_ctx):
    """tests/smoke.rs:113-130 on the GPU path."""
    s = make(gpu_ctx, 8, 8, 8, dt=5e-4)
    summary = s.run(2)
    assert summary.steps == 2
    assert np.isfinite(summary.max_divergence) and summary.max_divergence < 1e-5
    assert summary.max_speed > 0.0


def test_gpu_matches_reference_sor_solution(gpu_ctx):
    """The reference's own 3D solver is SOR; at a tight tolerance both solvers
    reach the same pinned solution, so a 2-step run agrees with it."""
    o = O.Cavity3D(8, solver=O.SOLVER_SOR, pressure_tolerance=1e-11, pressure_max_iters=50000)
    s = make(gpu_ctx, 8, 8, 8, pressure_tolerance=1e-11, pressure_max_iters=50000)
    for _ in range(2):
        o.step()
        s.advance_one_step()
    f = s.fields()
    uscale = max(np.abs(o.field(c)).max() for c in ("u", "v", "w"))
    for name in ("u", "v", "w"):
        assert np.abs(f[name] - o.field(name)).max() < 1e-8 * uscale, name
    assert rel(f["p"], o.field("p")) < 1e-8


@pytest.mark.parametrize("key", ["8x8x8", "12x10x14"])
def test_sor_matches_dct_and_reference_sor(gpu_ctx, key):
    """PressureSolverKind::Sor on the GPU (red-black order) reaches the
    reference solver's fixed point: the DCT closed form of the pinned
    system, and the oracle's lexicographic SOR at the same tolerance (the two
    orders differ in their iterates and counts, not in the solution).  Bar:
    1e-9 of max|p| at max-residual tolerance 1e-11."""
    d = np.load(os.path.join(GOLD, "neumann_dct.npz"))
    nz, ny, nx = map(int, key.split("x"))
    print(f"[gpu] SOR pressure solve {key}", flush=True)
    rhs = d[f"rhs3d_{key}"]
    s = make(gpu_ctx, nx, ny, nz, pressure_solver=fa.SOLVER_SOR, pressure_omega=1.7,
             pressure_tolerance=1e-11, pressure_max_iters=50000)
    p = np.zeros((nz, ny, nx))
    rep = s.pressure_solve(rhs, p)
    assert rep.converged and rep.residual_abs_inf < 1e-11
    assert rel(p, d[f"p3d_{key}"]) < 1e-9
    assert p[0, 0, 0] == 0.0
    o = O.Cavity3D(nx, ny, nz, solver=O.SOLVER_SOR, pressure_omega=1.7, pressure_tolerance=1e-11,
                   pressure_max_iters=50000)
    o.field("rhs")[:] = rhs
    orep = o.solve_pressure()
    assert orep.converged
    assert rel(p, o.field("p")) < 1e-9
    print(f"[gpu] SOR sweeps: red-black {rep.iterations}, lexicographic {orep.iterations}", flush=True)


def test_sor_steps_match_reference_sor(gpu_ctx):
    """Two cavity steps (tests/smoke.rs:113-130 setting, 8^3) with SOR on both
    sides to max residual 1e-10: fields agree to 1e-8 of their scale."""
    kw = dict(dt=5e-4, pressure_tolerance=1e-10, pressure_max_iters=50000)
    o = O.Cavity3D(8, solver=O.SOLVER_SOR, pressure_omega=1.7, **kw)
    s = make(gpu_ctx, 8, 8, 8, pressure_solver=fa.SOLVER_SOR, pressure_omega=1.7, **kw)
    for _ in range(2):
        o.step()
        rep = s.advance_one_step()
        assert rep.pressure_converged and rep.max_divergence < 1e-5
    f = s.fields()
    uscale = max(np.abs(o.field(c)).max() for c in ("u", "v", "w"))
    for name in ("u", "v", "w"):
        assert np.abs(f[name] - o.field(name)).max() < 1e-8 * uscale, name
    assert rel(f["p"], o.field("p")) < 1e-8


def test_sor_config_errors(gpu_ctx):
    with pytest.raises(fa.FlursysError, match="pressure_omega must lie between 0 and 2"):
        make(gpu_ctx, 8, 8, 8, pressure_solver=fa.SOLVER_SOR, pressure_omega=2.0)
    with pytest.raises(fa.FlursysError, match="max-residual stop rule"):
        make(gpu_ctx, 8, 8, 8, pressure_solver=fa.SOLVER_SOR, stop_rule=fa.STOP_REL_L2)


def test_repeatable_bitwise(gpu_ctx):
    out = []
    for _ in range(2):
        s = make(gpu_ctx, 20, 18, 16, reynolds=400.0)
        s.run(3)
        out.append(s.fields())
    for name in ("u", "v", "w", "p"):
        assert np.array_equal(out[0][name], out[1][name]), name


def test_max_iterations_is_a_status(gpu_ctx):
    s = make(gpu_ctx, 16, 16, 16, pressure_max_iters=3, pressure_tolerance=1e-12,
             stop_rule=fa.STOP_REL_L2)
    rhs = np.random.default_rng(3).standard_normal((16, 16, 16))
    p = np.zeros_like(rhs)
    rep = s.pressure_solve(rhs, p)
    assert rep.iterations == 3 and not rep.converged and rep.status == 0


def test_error_paths(gpu_ctx):
    with pytest.raises(fa.FlursysError, match="at least 4") as e:
        make(gpu_ctx, 3, 8, 8)
    assert e.value.kind == "ConfigError"
    with pytest.raises(fa.FlursysError, match="dt must be finite"):
        make(gpu_ctx, 8, 8, 8, dt=float("nan"))
    s = make(gpu_ctx, 8, 8, 8)
    rhs = np.zeros((8, 8, 8))
    rhs[2, 3, 4] = np.inf
    with pytest.raises(fa.FlursysError) as e:
        s.pressure_solve(rhs, np.zeros_like(rhs))
    assert e.value.kind == "NonFiniteValue"


def test_zero_rhs_takes_zero_iterations(gpu_ctx):
    s = make(gpu_ctx, 8, 8, 8, stop_rule=fa.STOP_REL_L2, initial_guess=fa.GUESS_ZERO)
    p = np.ones((8, 8, 8))
    rep = s.pressure_solve(np.zeros((8, 8, 8)), p)
    assert rep.iterations == 0 and rep.converged and np.all(p == 0.0)


def test_large_grid_properties(gpu_ctx):
    """128^3 at rel 1e-8: the true residual of the returned p (recomputed on
    the CPU) meets the tolerance and the gauge holds."""
    n = 128
    print("[gpu] 128^3 property solve", flush=True)
    s = make(gpu_ctx, n, n, n, pressure_tolerance=1e-8, stop_rule=fa.STOP_REL_L2,
             initial_guess=fa.GUESS_ZERO, pressure_max_iters=5000)
    rng = np.random.default_rng(11)
    rhs = rng.standard_normal((n, n, n))
    rhs[0, 0, 0] = 0.0
    p = np.zeros_like(rhs)
    rep = s.pressure_solve(rhs, p)
    assert rep.converged
    o = O.Cavity3D(n, n, n)
    b = -rhs.ravel()
    b[0] = 0.0
    r = b - o.apply_operator(p.ravel())
    assert np.linalg.norm(r) <= 1.5e-8 * np.linalg.norm(b)
    assert p[0, 0, 0] == 0.0


# ----------------------------------------------------------- mixed precision
# FP32 stencil CG inside FP64 iterative refinement.  Stated tolerance: the
# outer FP64 stop test is the same rel-res 1e-8, so the pressure agrees with
# the FP64 solution to ~cond(A) * 1e-8 relative; the bars below are 1e-6 for
# p and 1e-7 of the velocity scale for u, v, w after 6 steps.
# (32, 32, 32) and (264, 9, 19) run the FP32 plane-tile passes (256 cells wide:
# a partial second x-tile); (30, 10, 8) has nx % 4 != 0 and takes the column march.
@pytest.mark.parametrize("shape", [(32, 32, 32), (264, 9, 19), (30, 10, 8)])
def test_mixed_pressure_solve_meets_fp64_tolerance(gpu_ctx, shape):
    nx, ny, nz = shape
    print(f"[gpu] mixed pressure solve {shape}", flush=True)
    rng = np.random.default_rng(5)
    rhs = rng.standard_normal((nz, ny, nx)) * 100.0
    rhs[0, 0, 0] = 0.0
    out = {}
    for prec in (fa.PRECISION_FP64, fa.PRECISION_MIXED):
        s = make(gpu_ctx, nx, ny, nz, pressure_tolerance=1e-8, stop_rule=fa.STOP_REL_L2,
                 initial_guess=fa.GUESS_ZERO, pressure_max_iters=20000, precision=prec)
        p = np.zeros_like(rhs)
        rep = s.pressure_solve(rhs, p)
        assert rep.converged
        out[prec] = (p, rep)
    p64, r64 = out[fa.PRECISION_FP64]
    p32, r32 = out[fa.PRECISION_MIXED]
    assert r32.refinements >= 1
    assert r32.residual_l2 <= 1e-8 * r32.initial_l2
    o = O.Cavity3D(nx, ny, nz)
    b = -rhs.ravel()
    b[0] = 0.0
    true = np.linalg.norm(b - o.apply_operator(p32.ravel())) / np.linalg.norm(b)
    assert true <= 1.5e-8
    assert rel(p32, p64) < 1e-6
    assert p32[0, 0, 0] == 0.0


def test_mixed_steps_match_oracle(gpu_ctx):
    nx, ny, nz = 16, 14, 12
    kw = dict(reynolds=400.0, dt=5e-4, pressure_tolerance=1e-8, stop_rule=fa.STOP_REL_L2,
              initial_guess=fa.GUESS_ZERO, pressure_max_iters=5000)
    o = O.Cavity3D(nx, ny, nz, **kw)
    s = make(gpu_ctx, nx, ny, nz, precision=fa.PRECISION_MIXED, **kw)
    for _ in range(6):
        o.step()
        rep = s.advance_one_step()
        assert rep.pressure_converged and rep.pressure_refinements >= 1
    f = s.fields()
    uscale = max(np.abs(o.field(c)).max() for c in ("u", "v", "w"))
    for name in ("u", "v", "w"):
        assert np.abs(f[name] - o.field(name)).max() < 1e-7 * uscale, name
    assert rel(f["p"], o.field("p")) < 1e-6
    assert rep.max_divergence < 1e-5


def _c2_sample(f):
    return np.concatenate([f[0::8, 0::8, 0::8].ravel(), f[5::8, 5::8, 5::8].ravel()])


@pytest.mark.parametrize("prec", [fa.PRECISION_FP64, fa.PRECISION_MIXED])
def test_cavity128_re400_20_steps_match_oracle(gpu_ctx, prec):
    """The 3D cavity at 128^3, Re 400, dt 5e-4, rel-res 1e-8 from a zero
    guess, 20 steps from rest, against the FP64 oracle's run of the same
    configuration (tests/golden/cavity128_re400.npz: every field on two
    stride-8 lattices, the full plane next to the lid, norms, per-step
    iterations and divergence).  Stated tolerances: FP64 u, v, w within
    1e-9 of the lid velocity and p within 1e-8 of max|p|; FP32-in-FP64
    refinement u, v, w within 1e-7 and p within 1e-6; max divergence below
    the reference's smoke bar 1e-5 (tests/smoke.rs:127) at every step."""
    g = np.load(os.path.join(GOLD, "cavity128_re400.npz"))
    n, steps = int(g["n"]), int(g["steps"])
    print(f"[gpu] 128^3 Re 400 {steps} steps precision={prec}", flush=True)
    s = make(gpu_ctx, n, n, n, reynolds=400.0, dt=5e-4, pressure_tolerance=1e-8,
             stop_rule=fa.STOP_REL_L2, initial_guess=fa.GUESS_ZERO,
             pressure_max_iters=20000, precision=prec)
    divs = []
    for k in range(steps):
        rep = s.advance_one_step()
        assert rep.pressure_converged
        divs.append(rep.max_divergence)
        if prec == fa.PRECISION_FP64:
            assert abs(rep.pressure_iterations - int(g["iterations"][k])) <= 1
    f = s.fields()
    vel_tol, p_tol = (1e-9, 1e-8) if prec == fa.PRECISION_FP64 else (1e-7, 1e-6)
    errs = {}
    for name in ("u", "v", "w", "p"):
        scale = 1.0 if name != "p" else float(g["p_max"])
        errs[name] = float(np.abs(_c2_sample(f[name]) - g[f"{name}_sample"]).max()) / scale
    errs["u_lid"] = float(np.abs(f["u"][:, n - 1, :] - g["u_lidplane"]).max())
    errs["p_lid"] = float(np.abs(f["p"][:, n - 1, :] - g["p_lidplane"]).max()) / float(g["p_max"])
    print(f"[gpu] 128^3 errors vs oracle: {errs}, max div {max(divs):.3e}", flush=True)
    for name in ("u", "v", "w", "u_lid"):
        assert errs[name] < vel_tol, name
    for name in ("p", "p_lid"):
        assert errs[name] < p_tol, name
    for name in ("u", "v", "w", "p"):
        assert abs(np.linalg.norm(f[name]) - float(g[f"{name}_l2"])) <= (
            vel_tol if name != "p" else p_tol) * float(g[f"{name}_l2"]), name
    assert max(divs) < 1e-5
    assert abs(rep.max_speed - float(g["max_speed"][-1])) <= vel_tol


# ------------------------------------------- multigrid-preconditioned FCG
def _mg(ctx, n, **kw):
    kw.setdefault("pressure_tolerance", 1e-8)
    kw.setdefault("stop_rule", fa.STOP_REL_L2)
    kw.setdefault("initial_guess", fa.GUESS_ZERO)
    kw.setdefault("pressure_max_iters", 20000)
    return make(ctx, n if isinstance(n, int) else n[0], n if isinstance(n, int) else n[1],
                n if isinstance(n, int) else n[2], precision=fa.PRECISION_MIXED,
                preconditioner=fa.PRECOND_MULTIGRID, **kw)


@pytest.mark.parametrize("shape", [16, 32, (32, 16, 24), 64, (96, 64, 80)])
def test_multigrid_fcg_meets_fp64_tolerance(gpu_ctx, shape):
    print(f"[gpu] MG-FCG {shape}", flush=True)
    dims = (shape,) * 3 if isinstance(shape, int) else shape
    nx, ny, nz = dims
    rng = np.random.default_rng(9)
    rhs = rng.standard_normal((nz, ny, nx)) * 10.0
    rhs[0, 0, 0] = 0.0
    s = _mg(gpu_ctx, dims)
    p = np.zeros_like(rhs)
    rep = s.pressure_solve(rhs, p)
    assert rep.converged and not rep.fallback
    assert rep.iterations < 60, rep.iterations
    o = O.Cavity3D(nx, ny, nz)
    b = -rhs.ravel()
    b[0] = 0.0
    true = np.linalg.norm(b - o.apply_operator(p.ravel())) / np.linalg.norm(b)
    assert true <= 1.5e-8
    assert p[0, 0, 0] == 0.0
    ref = make(gpu_ctx, nx, ny, nz, pressure_tolerance=1e-12, stop_rule=fa.STOP_REL_L2,
               initial_guess=fa.GUESS_ZERO, pressure_max_iters=20000)
    pr = np.zeros_like(rhs)
    ref.pressure_solve(rhs, pr)
    assert rel(p, pr) < 1e-6


@pytest.mark.parametrize("size,scale", [(40.0, 1.0), (0.02, 1.0), (1.0, 1e25), (1.0, 1e-6),
                                        (3.0, 1e12)])
def test_multigrid_fcg_scaled_problems(gpu_ctx, size, scale):
    """The level-0 iterates are stored in FP16 scaled from max|r| and the
    domain size: large and small domains and right-hand sides far from 1
    still meet the FP64 tolerance without the FP64 fallback.  (Much smaller
    right-hand sides trip the solvers' absolute 1e-30 breakdown thresholds,
    the reference's, before any storage format matters.)"""
    print(f"[gpu] MG-FCG scaled problem size={size} scale={scale}", flush=True)
    n = 32
    rng = np.random.default_rng(5)
    rhs = rng.standard_normal((n, n, n)) * scale
    rhs[0, 0, 0] = 0.0
    s = _mg(gpu_ctx, n, length=size, height=size, depth=size)
    p = np.zeros_like(rhs)
    rep = s.pressure_solve(rhs, p)
    assert rep.converged and not rep.fallback, (rep.iterations, rep.fallback)
    assert rep.iterations < 60, rep.iterations
    # the reference on the unscaled problem (the solution is linear in rhs)
    ref = make(gpu_ctx, n, n, n, length=size, height=size, depth=size, pressure_tolerance=1e-12,
               stop_rule=fa.STOP_REL_L2, initial_guess=fa.GUESS_ZERO, pressure_max_iters=20000)
    pr = np.zeros_like(rhs)
    ref.pressure_solve(rhs / scale, pr)
    assert rel(p, pr * scale) < 1e-6


def test_multigrid_fcg_matches_dct(gpu_ctx):
    d = np.load(os.path.join(GOLD, "neumann_dct.npz"))
    s = _mg(gpu_ctx, 16, pressure_tolerance=1e-10)
    p = np.zeros((16, 16, 16))
    rep = s.pressure_solve(d["rhs3d_16x16x16"], p)
    assert rep.converged and rel(p, d["p3d_16x16x16"]) < 1e-8


def test_multigrid_steps_match_oracle(gpu_ctx):
    n = 32
    kw = dict(reynolds=1000.0, dt=5e-4, pressure_tolerance=1e-8, stop_rule=fa.STOP_REL_L2,
              initial_guess=fa.GUESS_ZERO, pressure_max_iters=5000)
    o = O.Cavity3D(n, **kw)
    s = _mg(gpu_ctx, n, reynolds=1000.0, dt=5e-4)
    for _ in range(4):
        o.step()
        rep = s.advance_one_step()
        assert rep.pressure_converged and not rep.pressure_fallback
        assert rep.pressure_iterations < 60
    f = s.fields()
    uscale = max(np.abs(o.field(c)).max() for c in ("u", "v", "w"))
    for name in ("u", "v", "w"):
        assert np.abs(f[name] - o.field(name)).max() < 1e-7 * uscale, name
    assert rel(f["p"], o.field("p")) < 1e-6


def test_multigrid_fallback_to_plain_pcg(gpu_ctx):
    """An inner budget too small to converge trips the fallback; the plain
    FP64 PCG then meets the same rel-res tolerance (the judge)."""
    n = 32
    rng = np.random.default_rng(2)
    rhs = rng.standard_normal((n, n, n))
    rhs[0, 0, 0] = 0.0
    s = _mg(gpu_ctx, n, inner_tolerance=1e-12)  # FP32 cannot reach 1e-12: stagnates
    p = np.zeros_like(rhs)
    rep = s.pressure_solve(rhs, p)
    assert rep.fallback == 1 and rep.converged
    # at 1e-12 the FP32 inner solve either stops halving its residual or
    # meets a singular search direction at round-off level
    assert rep.fallback_reason in (fa.FALLBACK_STAGNATION, fa.FALLBACK_BREAKDOWN)
    assert rep.residual_l2 <= 1e-8 * rep.initial_l2 * 1.0000001
    # the same solve at the default inner tolerance needs no fallback
    s = _mg(gpu_ctx, n)
    p = np.zeros_like(rhs)
    rep = s.pressure_solve(rhs, p)
    assert rep.fallback == 0 and rep.fallback_reason == fa.FALLBACK_NONE and rep.converged
    # an iteration cap below the multigrid's need: the inner budget runs out
    # and no iterations are left for the fallback
    s = _mg(gpu_ctx, n, pressure_max_iters=2)
    p = np.zeros_like(rhs)
    rep = s.pressure_solve(rhs, p)
    assert rep.fallback_reason == fa.FALLBACK_BUDGET
    assert rep.fallback == 0 and not rep.converged and rep.iterations == 2


def test_multigrid_zero_guess_ignores_stale_pressure(gpu_ctx):
    """A zero-guess solve on the mixed multigrid path does not clear p first
    (its first correction writes p): a second solve on the same solver, with
    the first solution still in device memory, gives the same bits, and a
    zero right-hand side leaves p = 0 (no correction is ever added)."""
    print("[gpu] MG-FCG stale p", flush=True)
    n = 32
    rng = np.random.default_rng(5)
    rhs = rng.standard_normal((n, n, n))
    rhs[0, 0, 0] = 0.0
    s = _mg(gpu_ctx, n)
    sols = []
    for start in (np.zeros((n, n, n)), 1e3 * rng.standard_normal((n, n, n))):
        p = start.copy()
        rep = s.pressure_solve(rhs, p)
        assert rep.converged and rep.fallback == 0
        sols.append(p)
    assert np.array_equal(sols[0], sols[1])
    p = np.ones((n, n, n))
    rep = s.pressure_solve(np.zeros((n, n, n)), p)
    assert rep.iterations == 0 and rep.converged and np.all(p == 0.0)


def test_multigrid_rejects_unsupported_grids(gpu_ctx):
    with pytest.raises(fa.FlursysError, match="halve"):
        _mg(gpu_ctx, (33, 20, 9))
    with pytest.raises(fa.FlursysError, match="mixed-precision"):
        make(gpu_ctx, 16, 16, 16, preconditioner=fa.PRECOND_MULTIGRID)
