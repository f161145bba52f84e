"""GPU parity of the 2D path (IncompressibleSolver with projection or
SIMPLE-style coupling, batched cavities) through the C ABI against the oracle.

Bars: pressure solves within 1 iteration of the oracle and p within 1e-9 of
max|p| (the device applies the reference's zero-mean projections through the
reductions, so the two differ only by round-off); whole steps within 1e-9 of
the velocity scale; the DCT closed form to 1e-10; the reference's smoke
thresholds and error texts.
"""

import os

import numpy as np
import pytest

import flursys_amd as fa
import flursys_amd.incompressible as I
import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def oracle_for(cfg: I.SimulationConfig, solid=None, threads=0):
    ts = cfg.time_step
    return O.Solver2D(cfg.nx, cfg.ny, length=cfg.length, height=cfg.height,
                      solver=O.SOLVER_SOR if cfg.pressure_solver == I.SOLVER_SOR else O.SOLVER_PCG,
                      pressure_omega=cfg.pressure_omega, inlet_profile=cfg.inlet_profile,
                      inlet_step_height=cfg.inlet_step_height,
                      density=cfg.density, nu=cfg.kinematic_viscosity, dt=cfg.dt,
                      t_end=cfg.t_end, pressure_max_iters=cfg.pressure_max_iters,
                      pressure_tolerance=cfg.pressure_tolerance, convection=cfg.convection,
                      stop_rule=cfg.stop_rule, initial_guess=cfg.initial_guess,
                      threads=threads, bc_kind=cfg.bc_kind, bc_u=cfg.bc_u, bc_v=cfg.bc_v,
                      outlet_pressure=cfg.outlet_pressure, adaptive=ts.adaptive,
                      min_dt=ts.min_dt, max_dt=ts.max_dt, target_cfl=ts.target_cfl,
                      target_diffusion=ts.target_diffusion_number,
                      max_growth=ts.max_growth_factor, solid=solid, coupling=cfg.coupling,
                      velocity_relaxation=cfg.velocity_relaxation,
                      pressure_relaxation=cfg.pressure_relaxation)


def channel_cfg(nx, ny, **kw):
    """Velocity inlet left (u = 1), pressure outlet right, walls top/bottom."""
    return I.SimulationConfig(nx=nx, ny=ny, length=nx / ny, height=1.0,
                              bc_kind=(I.BC_VELOCITY, I.BC_PRESSURE_OUTLET, I.BC_VELOCITY,
                                       I.BC_VELOCITY),
                              bc_u=(1.0, 0.0, 0.0, 0.0), **kw)


def block_mask(nx, ny):
    m = np.zeros((ny, nx), np.uint8)
    m[ny // 3: ny // 2, nx // 4: nx // 3] = 1
    return m


@pytest.mark.parametrize("key", ["16x16", "32x32", "24x40"])
def test_pressure_solve_matches_dct(gpu_ctx, key):
    d = np.load(os.path.join(GOLD, "neumann_dct.npz"))
    ny, nx = map(int, key.split("x"))
    cfg = I.SimulationConfig(nx=nx, ny=ny, pressure_tolerance=1e-13, stop_rule=fa.STOP_REL_L2,
                             initial_guess=fa.GUESS_ZERO, pressure_max_iters=10000)
    s = I.IncompressibleSolver(cfg, ctx=gpu_ctx)
    p = np.zeros((ny, nx))
    rep = s.pressure_solve(-d[f"b2d_{key}"], p)
    assert rep.converged
    assert rel(p, d[f"p2d_{key}"]) < 1e-10
    assert abs(p.mean()) < 1e-12


@pytest.mark.parametrize("case", ["cavity", "obstacle", "channel"])
@pytest.mark.parametrize("stop,tol,guess", [(fa.STOP_ABS_INF, 1e-6, fa.GUESS_PREVIOUS),
                                            (fa.STOP_REL_L2, 1e-8, fa.GUESS_ZERO)])
def test_pressure_solve_matches_oracle(gpu_ctx, case, stop, tol, guess):
    print(f"[gpu] 2D pressure solve {case} stop={stop}", flush=True)
    nx, ny = 40, 24
    kw = dict(pressure_tolerance=tol, stop_rule=stop, initial_guess=guess,
              pressure_max_iters=5000)
    solid = None
    if case == "channel":
        cfg = channel_cfg(nx, ny, **kw)
    else:
        cfg = I.SimulationConfig(nx=nx, ny=ny, **kw)
        if case == "obstacle":
            solid = block_mask(nx, ny)
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal((ny, nx)) * 30.0
    p0 = rng.standard_normal((ny, nx)) * 0.01
    if solid is not None:
        rhs[solid.astype(bool)] = 0.0
    if case == "channel":
        rhs[:, -1] = 0.0
    o = oracle_for(cfg, solid)
    o.field("rhs")[:] = rhs
    o.field("p")[:] = p0
    orep = o.solve_pressure()
    s = I.IncompressibleSolver(cfg, solid=solid, ctx=gpu_ctx)
    p = p0.copy()
    rep = s.pressure_solve(rhs, p)
    assert rep.converged and orep.converged
    assert abs(rep.iterations - orep.iterations) <= 1
    assert rel(p, o.field("p")) < 1e-9


@pytest.mark.parametrize("scheme", [I.FIRST_ORDER_UPWIND, I.CENTRAL])
@pytest.mark.parametrize("case", ["cavity", "obstacle", "channel"])
def test_steps_match_oracle(gpu_ctx, scheme, case):
    print(f"[gpu] 2D steps {case} scheme={scheme}", flush=True)
    nx, ny = 32, 24
    kw = dict(dt=2e-3, convection=scheme, pressure_tolerance=1e-8, stop_rule=fa.STOP_REL_L2,
              initial_guess=fa.GUESS_ZERO, pressure_max_iters=5000, kinematic_viscosity=0.01)
    solid = None
    if case == "channel":
        cfg = channel_cfg(nx, ny, **kw)
    else:
        cfg = I.SimulationConfig(nx=nx, ny=ny, **kw)
        if case == "obstacle":
            solid = block_mask(nx, ny)
    o = oracle_for(cfg, solid)
    s = I.IncompressibleSolver(cfg, solid=solid, ctx=gpu_ctx)
    for _ in range(5):
        d, orep = o.step()
        st = s.advance()
        assert st.pressure_converged
        assert abs(st.pressure_iterations - orep.iterations) <= 1
    f = s.fields()
    uscale = max(np.abs(o.field("u")).max(), np.abs(o.field("v")).max())
    assert np.abs(f["u"] - o.field("u")).max() < 1e-9 * uscale
    assert np.abs(f["v"] - o.field("v")).max() < 1e-9 * uscale
    assert rel(f["p"], o.field("p")) < 1e-8
    assert abs(st.continuity_residual - d.max_divergence) <= 1e-9 * uscale / cfg.length * cfg.nx
    assert abs(st.momentum_residual - d.velocity_change) <= 1e-9 * uscale
    assert abs(st.max_speed - d.max_speed) <= 1e-9 * uscale
    fu = s.field_update()
    assert np.abs(fu["speed"] - o.field("speed")).max() < 1e-9 * uscale
    assert np.abs(fu["vorticity"] - o.field("vort")).max() < 1e-7 * max(np.abs(o.field("vort")).max(), 1)


def test_smoke_threshold_and_t_end(gpu_ctx):
    """tests/smoke.rs:57-69: 16^2 cavity, one step, divergence < 1e-5; a
    second step past t_end is an error (incompressible.rs:623-625)."""
    cfg = I.SimulationConfig(nx=16, ny=16, dt=1e-3, t_end=1e-3, pressure_max_iters=800,
                             pressure_tolerance=1e-6)
    s = I.IncompressibleSolver(cfg, ctx=gpu_ctx)
    st = s.advance()
    assert st.continuity_residual < 1e-5 and st.reached_end_time
    assert abs(st.time - 1e-3) < 1e-12
    with pytest.raises(fa.FlursysError, match="already reached t_end"):
        s.advance()


def test_adaptive_time_stepping_matches_oracle(gpu_ctx):
    cfg = I.SimulationConfig(nx=32, ny=32, dt=1e-3, kinematic_viscosity=0.01,
                             time_step=I.TimeStepSettings(adaptive=True, min_dt=1e-6,
                                                          max_dt=1e-2, target_cfl=0.5,
                                                          target_diffusion_number=0.25,
                                                          max_growth_factor=1.1))
    o = oracle_for(cfg)
    s = I.IncompressibleSolver(cfg, ctx=gpu_ctx)
    for _ in range(6):
        d, _ = o.step()
        st = s.advance()
        assert abs(st.time_step - d.time_step) <= 1e-12 * d.time_step
    assert abs(st.time - o.time) < 1e-12


def test_config1_as_written_blows_up(gpu_ctx):
    """128^2, Re 100, dt = 0.25 h: viscous number 0.64 > 0.5 -> non-finite."""
    cfg = I.SimulationConfig.cavity(128, reynolds=100.0, dt=0.25 / 128, pressure_tolerance=1e-8,
                                    stop_rule=fa.STOP_REL_L2, initial_guess=fa.GUESS_ZERO,
                                    pressure_max_iters=5000)
    s = I.IncompressibleSolver(cfg, ctx=gpu_ctx)
    with pytest.raises(fa.FlursysError, match="Non-finite"):
        for _ in range(200):
            st = s.advance()
            assert abs(st.viscous_diffusion_number - 0.64) < 1e-12


def test_config_validation_texts(gpu_ctx):
    with pytest.raises(fa.FlursysError, match="nx and ny must be at least 4"):
        I.IncompressibleSolver(I.SimulationConfig(nx=3, ny=8), ctx=gpu_ctx)
    with pytest.raises(fa.FlursysError, match="t_end must be positive for projection coupling"):
        I.IncompressibleSolver(I.SimulationConfig(t_end=0.0), ctx=gpu_ctx)
    bad = I.SimulationConfig(dt=1.0, time_step=I.TimeStepSettings(adaptive=True))
    with pytest.raises(fa.FlursysError, match="adaptive initial dt must lie between"):
        I.IncompressibleSolver(bad, ctx=gpu_ctx)


def test_batch_matches_individual_runs(gpu_ctx):
    cfgs = [I.SimulationConfig.cavity(n, reynolds=re, dt=dt, pressure_tolerance=1e-8,
                                      stop_rule=fa.STOP_REL_L2, initial_guess=fa.GUESS_ZERO,
                                      pressure_max_iters=5000)
            for n, re, dt in ((16, 100.0, 2e-3), (40, 400.0, 1e-3), (27, 1000.0, 1.5e-3))]
    b = I.CavityBatch(cfgs, ctx=gpu_ctx)
    oracles = [oracle_for(c) for c in cfgs]
    for _ in range(4):
        reps = b.step()
        for o, r in zip(oracles, reps):
            d, orep = o.step()
            assert abs(r.pressure_iterations - orep.iterations) <= 1
    for k, o in enumerate(oracles):
        f = b.fields(k)
        uscale = max(np.abs(o.field("u")).max(), np.abs(o.field("v")).max())
        assert np.abs(f["u"] - o.field("u")).max() < 1e-9 * uscale
        assert rel(f["p"], o.field("p")) < 1e-8


def test_batch_respects_each_t_end(gpu_ctx):
    cfgs = [I.SimulationConfig.cavity(16, dt=1e-3, t_end=2e-3),
            I.SimulationConfig.cavity(20, dt=1e-3, t_end=4e-3)]
    b = I.CavityBatch(cfgs, ctx=gpu_ctx)
    for _ in range(4):
        reps = b.step()
    assert abs(reps[1].time - 4e-3) < 1e-12
    assert reps[1].iteration == 4


def uneven_cfgs():
    """Uneven cavities with the reference's adaptive dt rule and defaults
    (warm start, max|r| < 1e-5), as in the batched benchmark."""
    ts = I.TimeStepSettings(adaptive=True, target_cfl=0.5, target_diffusion_number=0.25)
    return [I.SimulationConfig.cavity(n, reynolds=re, time_step=ts)
            for n, re in ((16, 100.0), (37, 400.0), (64, 1000.0), (24, 250.0))]


def all_fields(b):
    return [b.fields(k) for k in range(len(b.cfgs))]


def test_device_run_matches_host_steps(gpu_ctx):
    print("[gpu] batch device run vs host steps", flush=True)
    host = I.CavityBatch(uneven_cfgs(), ctx=gpu_ctx)
    for _ in range(6):
        host.step()
    dev = I.CavityBatch(uneven_cfgs(), ctx=gpu_ctx)
    reps, ms = dev.run(6, use_graphs=False, step_times=True)
    assert all(r.steps == 6 and r.status == 0 for r in reps)
    assert len(ms) == 6 and np.all(ms > 0)
    for fh, fd in zip(all_fields(host), all_fields(dev)):
        for name in ("u", "v", "p"):
            assert rel(fd[name], fh[name]) < 1e-12, name


def test_graph_replays_match_eager_runs(gpu_ctx):
    """Captured step halves replayed with eager work and host reads between
    the replays give the eager run's fields after every replay."""
    print("[gpu] batch graph replays", flush=True)
    eager = I.CavityBatch(uneven_cfgs(), ctx=gpu_ctx)
    graph = I.CavityBatch(uneven_cfgs(), ctx=gpu_ctx)
    assert graph.graph_mode() == fa._lib.GRAPH_NONE
    for rnd in range(3):
        eager.run(2, use_graphs=False)
        graph.run(2, use_graphs=True)
        print(f"[gpu] graph mode {graph.graph_mode()}", flush=True)
        assert graph.graph_mode() in (fa._lib.GRAPH_WHOLE_STEP, fa._lib.GRAPH_SPLIT)
        for fe, fg in zip(all_fields(eager), all_fields(graph)):
            for name in ("u", "v", "p"):
                assert np.array_equal(fe[name], fg[name]), (rnd, name)
    # eager host-driven steps in between keep both in step
    eager.step()
    graph.step()
    eager.run(1, use_graphs=False)
    graph.run(1, use_graphs=True)
    for fe, fg in zip(all_fields(eager), all_fields(graph)):
        assert np.array_equal(fe["p"], fg["p"])
    # a launch-shape change re-captures the step
    graph.set_launch(256, 3)
    assert graph.graph_mode() == fa._lib.GRAPH_NONE
    eager.run(1, use_graphs=False)
    graph.run(1, use_graphs=True)
    for fe, fg in zip(all_fields(eager), all_fields(graph)):
        assert np.array_equal(fe["p"], fg["p"])


def test_device_run_reports_the_first_error_and_continues(gpu_ctx):
    ts_bad = I.TimeStepSettings(adaptive=True, min_dt=0.5, max_dt=1.0)  # stable dt < min_dt
    cfgs = uneven_cfgs()[:2] + [I.SimulationConfig.cavity(20, dt=0.6, time_step=ts_bad)]
    b = I.CavityBatch(cfgs, ctx=gpu_ctx)
    with pytest.raises(fa.FlursysError) as e:
        b.run(3)
    assert "below configured min_dt" in e.value.message
    reps = b.last_run
    assert reps[2].status != 0 and reps[2].steps == 0
    assert reps[0].status == 0 and reps[0].steps == 3


def test_autotune_keeps_results(gpu_ctx):
    print("[gpu] batch autotune", flush=True)
    ref = I.CavityBatch(uneven_cfgs(), ctx=gpu_ctx)
    ref.run(4)
    b = I.CavityBatch(uneven_cfgs(), ctx=gpu_ctx)
    chosen, table = b.autotune(10)
    assert chosen["blocks"] >= 1 and chosen["ms"] > 0
    assert chosen in table and min(t["ms"] for t in table) == chosen["ms"]
    assert {t["threads"] for t in table} == {256, 512, 1024}
    assert b.launch() == dict(chosen, ms=0.0)
    b.run(4)
    for fr, fb in zip(all_fields(ref), all_fields(b)):
        assert np.array_equal(fr["p"], fb["p"])


def test_launch_shapes_give_identical_results(gpu_ctx):
    """Partials are per 64-cell sub-tile and folded in a fixed order, so the
    persistent PCG's block size and block count change nothing."""
    print("[gpu] batch launch shapes", flush=True)
    cfgs = uneven_cfgs() + [I.SimulationConfig.cavity(96, reynolds=700.0)]
    ref = I.CavityBatch(cfgs, ctx=gpu_ctx)
    ref.run(3)
    for threads, blocks in ((256, 1), (512, 7), (1024, 3), (256, 40)):
        b = I.CavityBatch(cfgs, ctx=gpu_ctx)
        b.set_launch(threads, blocks)
        assert b.launch()["threads"] == threads
        b.run(3)
        for fr, fb in zip(all_fields(ref), all_fields(b)):
            for name in ("u", "v", "p"):
                assert np.array_equal(fr[name], fb[name]), (threads, blocks, name)
        b.close()


def test_launch_shape_validation(gpu_ctx):
    b = I.CavityBatch(uneven_cfgs()[:1], ctx=gpu_ctx)
    with pytest.raises(fa.FlursysError, match="256, 512 or 1024 threads"):
        b.set_launch(128, 4)
    too_many = [I.SimulationConfig.cavity(8)] * 257
    with pytest.raises(fa.FlursysError, match="at most 256 grids"):
        I.CavityBatch(too_many, ctx=gpu_ctx)


def mixed(cfg, **kw):
    """The same case with FP32 CG cycles inside FP64 iterative refinement."""
    cfg.precision = fa.PRECISION_MIXED
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def test_mixed_batch_meets_the_tolerance_and_matches_the_oracle(gpu_ctx):
    """Mixed precision: every accepted pressure meets the same FP64 relative
    residual (the stop test runs on true FP64 residuals), so the fields
    agree with the FP64 oracle to the level that tolerance allows."""
    print("[gpu] mixed batch vs oracle", flush=True)
    base = dict(pressure_tolerance=1e-8, stop_rule=fa.STOP_REL_L2, initial_guess=fa.GUESS_ZERO,
                pressure_max_iters=5000)
    cfgs = [mixed(I.SimulationConfig.cavity(n, reynolds=re, dt=dt, **base))
            for n, re, dt in ((16, 100.0, 2e-3), (40, 400.0, 1e-3), (27, 1000.0, 1.5e-3))]
    b = I.CavityBatch(cfgs, ctx=gpu_ctx)
    oracles = [oracle_for(c) for c in cfgs]
    for _ in range(4):
        reps = b.step()
        for o, r in zip(oracles, reps):
            _, orep = o.step()
            assert r.pressure_converged
            assert r.pressure_iterations <= 2 * orep.iterations + 10
    for k, o in enumerate(oracles):
        f = b.fields(k)
        uscale = max(np.abs(o.field("u")).max(), np.abs(o.field("v")).max())
        assert np.abs(f["u"] - o.field("u")).max() < 1e-7 * uscale
        assert rel(f["p"], o.field("p")) < 1e-6


def test_mixed_reference_defaults_meet_max_residual(gpu_ctx):
    """The reference's defaults (previous pressure, max|r| < 1e-5): mixed
    steps report a true FP64 residual under the tolerance and stay close to
    the FP64 batch."""
    print("[gpu] mixed batch, reference defaults", flush=True)
    f64 = I.CavityBatch(uneven_cfgs(), ctx=gpu_ctx)
    mix = I.CavityBatch([mixed(c) for c in uneven_cfgs()], ctx=gpu_ctx)
    for _ in range(5):
        r64, rmx = f64.step(), mix.step()
        for a, m in zip(r64, rmx):
            assert m.pressure_converged and m.pressure_residual < 1e-5
            assert m.pressure_iterations <= 2 * a.pressure_iterations + 10
    for fa64, fmx in zip(all_fields(f64), all_fields(mix)):
        uscale = max(np.abs(fa64["u"]).max(), np.abs(fa64["v"]).max())
        assert np.abs(fmx["u"] - fa64["u"]).max() < 1e-4 * uscale


def test_mixed_launch_shapes_and_graphs_give_identical_results(gpu_ctx):
    print("[gpu] mixed batch launch shapes and graphs", flush=True)
    cfgs = [mixed(c) for c in uneven_cfgs()]
    ref = I.CavityBatch(cfgs, ctx=gpu_ctx)
    ref.run(3, use_graphs=False)
    for threads, blocks, graphs in ((256, 5, True), (1024, 2, False)):
        b = I.CavityBatch(cfgs, ctx=gpu_ctx)
        b.set_launch(threads, blocks)
        b.run(3, use_graphs=graphs)
        for fr, fb in zip(all_fields(ref), all_fields(b)):
            assert np.array_equal(fr["p"], fb["p"])
        b.close()


def test_precision_validation(gpu_ctx):
    bad = I.SimulationConfig.cavity(16)
    bad.precision = 7
    with pytest.raises(fa.FlursysError, match="unknown precision"):
        I.CavityBatch([bad], ctx=gpu_ctx)
    with pytest.raises(fa.FlursysError, match="share one precision"):
        I.CavityBatch([I.SimulationConfig.cavity(16), mixed(I.SimulationConfig.cavity(16))],
                      ctx=gpu_ctx)
    with pytest.raises(fa.FlursysError, match="inner_tolerance"):
        I.CavityBatch([mixed(I.SimulationConfig.cavity(16), inner_tolerance=2.0)], ctx=gpu_ctx)


def test_run_step_and_pcg_times(gpu_ctx):
    b = I.CavityBatch(uneven_cfgs(), ctx=gpu_ctx)
    _, ms, pms = b.run(2, use_graphs=False, step_times=True, pcg_times=True)
    assert np.all(ms > 0) and np.all(pms > 0) and np.all(pms <= ms)
    _, ms, pms = b.run(2, use_graphs=True, step_times=True, pcg_times=True)
    assert np.all(ms > 0)
    if b.graph_mode() == fa._lib.GRAPH_WHOLE_STEP:
        assert np.all(pms == -1.0)


# ------------------------------------------------------------------ SOR
@pytest.mark.parametrize("case", ["cavity", "obstacle", "channel"])
def test_sor_pressure_solve_matches_oracle_sor(gpu_ctx, case):
    """PressureSolverKind::Sor (incompressible.rs:979-1046) on the GPU in
    red-black order against the oracle's lexicographic SOR, both to max
    residual 1e-11: the same fixed point (the first fluid cell's gauge, or
    the outlet pressure) to 1e-8 of max|p|.  Sweep counts differ by order."""
    nx, ny = 24, 20
    solid = block_mask(nx, ny) if case == "obstacle" else None
    kw = dict(pressure_solver=I.SOLVER_SOR, pressure_omega=1.7, pressure_tolerance=1e-11,
              pressure_max_iters=50000)
    cfg = channel_cfg(nx, ny, **kw) if case == "channel" else I.SimulationConfig(nx=nx, ny=ny, **kw)
    print(f"[gpu] 2D SOR {case}", flush=True)
    rng = np.random.default_rng(3)
    rhs = rng.standard_normal((ny, nx)) * 20.0
    if case != "channel":
        fluid = np.ones((ny, nx), bool) if solid is None else solid == 0
        rhs[fluid] -= rhs[fluid].mean()  # compatible Neumann right-hand side
    if solid is not None:
        rhs[solid == 1] = 0.0
    o = oracle_for(cfg, solid)
    o.field("rhs")[:] = rhs
    orep = o.solve_pressure()
    s = I.IncompressibleSolver(cfg, solid=solid, ctx=gpu_ctx)
    p = np.zeros((ny, nx))
    rep = s.pressure_solve(rhs, p)
    assert rep.converged and orep.converged
    assert rep.residual_abs_inf < 1e-11
    assert rel(p, o.field("p")) < 1e-8
    print(f"[gpu] SOR sweeps: red-black {rep.iterations}, lexicographic {orep.iterations}", flush=True)


def test_sor_steps_match_oracle(gpu_ctx):
    """Whole cavity steps (tests/smoke.rs:57-69 setting, 16^2) with SOR to max
    residual 1e-10 on both sides: fields within 1e-8 of the velocity scale,
    and the smoke threshold."""
    cfg = I.SimulationConfig.cavity(16, dt=1e-3, pressure_solver=I.SOLVER_SOR,
                                    pressure_omega=1.7, pressure_tolerance=1e-10,
                                    pressure_max_iters=50000)
    o = oracle_for(cfg)
    s = I.IncompressibleSolver(cfg, ctx=gpu_ctx)
    for _ in range(3):
        od, _ = o.step()
        r = s.advance()
        assert r.pressure_converged
    assert r.continuity_residual < 1e-5
    f = s.fields()
    for name in ("u", "v"):
        assert np.abs(f[name] - o.field(name)).max() < 1e-8, name
    assert rel(f["p"], o.field("p")) < 1e-7


# ---------------------------------------------- Symmetry sides, inlet profile
@pytest.mark.parametrize("kinds", [
    (I.BC_SYMMETRY, I.BC_SYMMETRY, I.BC_VELOCITY, I.BC_VELOCITY),
    (I.BC_VELOCITY, I.BC_PRESSURE_OUTLET, I.BC_SYMMETRY, I.BC_SYMMETRY)])
def test_symmetry_sides_match_oracle(gpu_ctx, kinds):
    """Symmetry sides (apply_velocity_boundaries :1996-2023: zero normal face;
    tangential ghost = centre, :2176, :2197) on the GPU against the oracle:
    a lid-driven box with symmetric side walls, and a channel with symmetric
    top and bottom (the CylinderRe100 case's sides, src/cases/mod.rs:121-127).
    The channel takes the parabolic inlet: with a uniform one, uniform flow is
    the exact solution, the second step's right-hand side is round-off and
    the PCG breaks down on both sides."""
    nx, ny = 28, 20
    box = kinds[0] == I.BC_SYMMETRY
    bc_u = (0.0, 0.0, 0.0, 1.0) if box else (1.0, 0.0, 0.0, 0.0)
    cfg = I.SimulationConfig(nx=nx, ny=ny, length=1.4, height=1.0, bc_kind=kinds, bc_u=bc_u,
                             inlet_profile=I.INLET_UNIFORM if box else I.INLET_PARABOLIC,
                             dt=1e-3, pressure_tolerance=1e-8, stop_rule=fa.STOP_REL_L2,
                             initial_guess=fa.GUESS_ZERO, pressure_max_iters=5000,
                             convection=I.CENTRAL)
    o = oracle_for(cfg)
    s = I.IncompressibleSolver(cfg, ctx=gpu_ctx)
    for _ in range(4):
        od, orep = o.step()
        r = s.advance()
        assert r.pressure_converged
        assert abs(r.pressure_iterations - orep.iterations) <= 1
    f = s.fields()
    for name in ("u", "v"):
        assert np.abs(f[name] - o.field(name)).max() < 1e-9, name
    assert rel(f["p"], o.field("p")) < 1e-9
    # the symmetric sides' normal faces are zero
    if kinds[0] == I.BC_SYMMETRY:
        assert not f["u"][:, 0].any() and not f["u"][:, nx].any()
    else:
        assert not f["v"][0].any() and not f["v"][ny].any()


@pytest.mark.parametrize("step_h", [0.0, 0.5])
def test_parabolic_inlet_matches_oracle(gpu_ctx, step_h):
    """The Channel (step 0) and BackwardFacingStep (step 0.5) inlets of
    Case::boundary_velocity (src/cases/mod.rs:141-170): u(y) = 6 ubar eta
    (1 - eta) above the step, 0 below; the inlet faces against the hand
    formula and whole steps against the oracle."""
    nx, ny, ubar = 32, 16, 0.8
    cfg = I.SimulationConfig(nx=nx, ny=ny, length=2.0, height=1.0,
                             bc_kind=(I.BC_VELOCITY, I.BC_PRESSURE_OUTLET, I.BC_VELOCITY,
                                      I.BC_VELOCITY),
                             bc_u=(ubar, 0.0, 0.0, 0.0), inlet_profile=I.INLET_PARABOLIC,
                             inlet_step_height=step_h, dt=1e-3, pressure_tolerance=1e-8,
                             stop_rule=fa.STOP_REL_L2, initial_guess=fa.GUESS_ZERO,
                             pressure_max_iters=5000)
    o = oracle_for(cfg)
    s = I.IncompressibleSolver(cfg, ctx=gpu_ctx)
    y = (np.arange(ny) + 0.5) / ny
    eta = np.clip((y - step_h) / (1.0 - step_h), 0.0, 1.0)
    inlet = np.where(y <= step_h, 0.0, 6.0 * ubar * eta * (1.0 - eta))
    assert np.abs(s.fields()["u"][:, 0] - inlet).max() < 1e-15
    for _ in range(3):
        o.step()
        r = s.advance()
        assert r.pressure_converged
    f = s.fields()
    assert np.abs(f["u"][:, 0] - inlet).max() < 1e-15
    for name in ("u", "v"):
        assert np.abs(f[name] - o.field(name)).max() < 1e-9, name
    assert rel(f["p"], o.field("p")) < 1e-9


# ------------------------------------------------ SIMPLE-style coupling
@pytest.mark.parametrize("case,guess", [("obstacle", fa.GUESS_ZERO), ("obstacle", fa.GUESS_PREVIOUS),
                                        ("channel", fa.GUESS_PREVIOUS)])
def test_simple_steps_match_oracle(gpu_ctx, case, guess):
    """SIMPLE pseudo-steps (incompressible.rs:659-712) with the reference's
    relaxation defaults: the p' solve starts from zero whatever the warm-start
    setting, fields within 1e-9 of the scale, time = step * dt, no end time
    (t_end = 0 is valid under SIMPLE)."""
    print(f"[gpu] 2D SIMPLE {case} guess={guess}", flush=True)
    nx, ny = 36, 28
    kw = dict(dt=4e-3, pressure_tolerance=1e-10, stop_rule=fa.STOP_REL_L2, initial_guess=guess,
              pressure_max_iters=5000, kinematic_viscosity=0.01, coupling=I.SIMPLE, t_end=0.0)
    solid = None
    if case == "channel":
        cfg = channel_cfg(nx, ny, inlet_profile=I.INLET_PARABOLIC, **kw)
    else:
        cfg = I.SimulationConfig(nx=nx, ny=ny, **kw)
        solid = block_mask(nx, ny)
    o = oracle_for(cfg, solid)
    s = I.IncompressibleSolver(cfg, solid=solid, ctx=gpu_ctx)
    for k in range(1, 16):
        d, orep = o.step()
        st = s.advance()
        assert st.pressure_converged and abs(st.pressure_iterations - orep.iterations) <= 1
        assert st.iteration == k and st.time == k * cfg.dt and not st.reached_end_time
        assert st.time_step == cfg.dt
    f = s.fields()
    uscale = max(np.abs(o.field("u")).max(), np.abs(o.field("v")).max())
    assert np.abs(f["u"] - o.field("u")).max() < 1e-9 * uscale
    assert np.abs(f["v"] - o.field("v")).max() < 1e-9 * uscale
    assert rel(f["p"], o.field("p")) < 1e-8
    assert abs(st.momentum_residual - d.velocity_change) <= 1e-9 * uscale


def test_simple_sor_steps_match_oracle(gpu_ctx):
    cfg = I.SimulationConfig.cavity(16, dt=2e-3, pressure_solver=I.SOLVER_SOR, pressure_omega=1.7,
                                    pressure_tolerance=1e-10, pressure_max_iters=50000,
                                    coupling=I.SIMPLE, velocity_relaxation=0.8,
                                    pressure_relaxation=0.5)
    o = oracle_for(cfg)
    s = I.IncompressibleSolver(cfg, ctx=gpu_ctx)
    for _ in range(4):
        o.step()
        assert s.advance().pressure_converged
    f = s.fields()
    for name in ("u", "v"):
        assert np.abs(f[name] - o.field(name)).max() < 1e-8, name
    assert rel(f["p"], o.field("p")) < 1e-7


def test_simple_and_projection_grids_in_one_device_run(gpu_ctx):
    """A batch holding SIMPLE and projection grids, stepped on the device
    (captured graphs) and on the host: each grid equals its oracle run."""
    print("[gpu] SIMPLE + projection batch device run", flush=True)
    common = dict(pressure_tolerance=1e-10, stop_rule=fa.STOP_REL_L2, initial_guess=fa.GUESS_ZERO,
                  pressure_max_iters=5000)
    cfgs = [I.SimulationConfig.cavity(24, reynolds=100.0, dt=3e-3, coupling=I.SIMPLE, **common),
            I.SimulationConfig.cavity(20, reynolds=400.0, dt=1e-3, t_end=5e-3, **common),
            I.SimulationConfig.cavity(30, reynolds=200.0, dt=2e-3, coupling=I.SIMPLE,
                                      velocity_relaxation=0.5, pressure_relaxation=0.9, **common)]
    dev = I.CavityBatch(cfgs, ctx=gpu_ctx)
    reps, _ = dev.run(8, use_graphs=True)
    host = I.CavityBatch(cfgs, ctx=gpu_ctx)
    for _ in range(8):
        host.step()
    assert [r.steps for r in reps] == [8, 5, 8]
    assert not reps[0].reached_end_time and reps[1].reached_end_time
    assert reps[0].time == 8 * 3e-3 and reps[2].time == 8 * 2e-3
    for k, c in enumerate(cfgs):
        o = oracle_for(c)
        for _ in range(reps[k].steps):
            o.step()
        for f in (dev.fields(k), host.fields(k)):
            uscale = max(np.abs(o.field("u")).max(), np.abs(o.field("v")).max())
            assert np.abs(f["u"] - o.field("u")).max() < 1e-9 * uscale
            assert rel(f["p"], o.field("p")) < 1e-8


def test_simple_validation_texts(gpu_ctx):
    bad = I.SimulationConfig(coupling=I.SIMPLE, time_step=I.TimeStepSettings(adaptive=True))
    with pytest.raises(fa.FlursysError, match="adaptive time stepping requires transient projection coupling"):
        I.IncompressibleSolver(bad, ctx=gpu_ctx)
    for name, val in (("velocity_relaxation", 0.0), ("velocity_relaxation", 1.5),
                      ("pressure_relaxation", -0.1), ("pressure_relaxation", float("nan"))):
        cfg = I.SimulationConfig(coupling=I.SIMPLE, **{name: val})
        with pytest.raises(fa.FlursysError, match=f"{name} must lie between 0 \\(exclusive\\) and 1"):
            I.IncompressibleSolver(cfg, ctx=gpu_ctx)
