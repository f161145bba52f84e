"""The oracle's SIMPLE-style coupling of the 2D solver (advance_one_step with
PressureVelocityCoupling::Simple, src/solver/incompressible.rs:659-712;
predict_momentum :822-826 / :848-853, build_pressure_rhs :936-941,
correct_velocity :1322-1323, reached_end_time :557-561), pinned by what the
algebra of one step fixes rather than by retyping it:

* relaxation factors 1 turn a SIMPLE step into a projection step (p' = p -
  p_old because the divergence of the face gradient is the pressure
  operator), so the two runs agree to the solver tolerance;
* with a velocity relaxation a the velocity is that of projection steps of
  size a dt (the factor sits in the predictor, the right-hand side and the
  correction alike: a factor missing from one of them fails this), whatever
  the pressure relaxation b, since the old pressure's gradient is projected
  out again; the pressure follows p_k = p_(k-1) + b (P_k - p_(k-1)) with P_k
  the projection's pressure;
* from rest, the pressure relaxation only scales the first pressure, bitwise;
* the steady state SIMPLE converges to is the projection's;
* time = step * dt, no end time, the reference's validation texts.
"""

import numpy as np
import pytest

import oracle as O


def cavity(coupling, n=12, **kw):
    kw.setdefault("pressure_tolerance", 1e-13)
    kw.setdefault("pressure_max_iters", 4000)
    return O.Solver2D(n, n + 2, length=1.0, height=1.0 + 2.0 / n, nu=0.02, dt=kw.pop("dt", 4e-3),
                      coupling=coupling, stop_rule=O.STOP_REL_L2, initial_guess=O.GUESS_ZERO,
                      **kw)


def fields(s):
    return [s.field(k).copy() for k in ("u", "v", "p")]


def close(a, b, tol):
    for x, y in zip(a, b):
        scale = max(np.abs(y).max(), 1e-300)
        assert np.abs(x - y).max() <= tol * scale


def test_unit_relaxation_is_a_projection_step():
    solid = np.zeros((14, 12), np.uint8)
    solid[5:8, 4:6] = 1
    a = cavity(O.COUPLING_SIMPLE, velocity_relaxation=1.0, pressure_relaxation=1.0, solid=solid)
    b = cavity(O.COUPLING_PROJECTION, solid=solid)
    for k in range(12):
        a.step()
        b.step()
        close(fields(a), fields(b), 1e-9)
    # and the agreement is not vacuous: the pressure moved and differs from zero
    assert np.abs(b.field("p")).max() > 1e-3


def test_relaxed_steps_follow_the_projection():
    a = cavity(O.COUPLING_SIMPLE, velocity_relaxation=0.6, pressure_relaxation=0.3, dt=5e-3)
    b = cavity(O.COUPLING_PROJECTION, dt=0.6 * 5e-3)
    p = np.zeros_like(a.field("p"))
    for k in range(10):
        da, _ = a.step()
        b.step()
        close(fields(a)[:2], fields(b)[:2], 1e-9)
        p = p + 0.3 * (b.field("p") - p)
        close([a.field("p")], [p], 1e-9)
        assert da.time_step == 5e-3
    # a relaxation of 1 with the same dt does not match: the check sees a
    c = cavity(O.COUPLING_SIMPLE, velocity_relaxation=1.0, pressure_relaxation=1.0, dt=5e-3)
    for k in range(10):
        c.step()
    assert np.abs(c.field("u") - b.field("u")).max() > 1e-4


def test_pressure_relaxation_scales_the_first_pressure():
    # one thread: the OpenMP reductions' combining order would differ run to run
    a = cavity(O.COUPLING_SIMPLE, velocity_relaxation=0.7, pressure_relaxation=0.3, threads=1)
    b = cavity(O.COUPLING_SIMPLE, velocity_relaxation=0.7, pressure_relaxation=1.0, threads=1)
    a.step()
    b.step()
    ua, va, pa = fields(a)
    ub, vb, pb = fields(b)
    assert np.array_equal(ua, ub) and np.array_equal(va, vb)
    assert np.array_equal(pa, 0.3 * pb)


def test_simple_reaches_the_projection_steady_state():
    """16 x 16 cavity at Re = 50 driven to a steady state both ways."""
    def run(s, steps):
        last = None
        for _ in range(steps):
            d, _ = s.step()
            last = d.velocity_change
        return last
    # the reference's max-residual rule: near the steady state SIMPLE's p'
    # right-hand side vanishes and a relative rule would chase round-off
    kw = dict(nu=0.02, dt=0.02, pressure_tolerance=1e-12, pressure_max_iters=4000)
    a = O.Solver2D(16, coupling=O.COUPLING_SIMPLE, **kw)
    b = O.Solver2D(16, **kw)
    assert run(a, 3000) < 1e-11 and run(b, 3000) < 1e-11
    close(fields(a)[:2], fields(b)[:2], 1e-7)
    close(fields(a)[2:], fields(b)[2:], 1e-6)


def test_time_counter_end_time_and_validation():
    # simple_steps_do_not_report_a_physical_end_time (incompressible.rs:2639-2651)
    s = O.Solver2D(8, dt=1e-3, t_end=0.0, coupling=O.COUPLING_SIMPLE, pressure_max_iters=50,
                   pressure_tolerance=1e-4)
    for k in range(1, 4):
        s.step()
        assert s.s.step == k and s.s.time == k * 1e-3
        assert not s.reached_end()
    with pytest.raises(ValueError):
        O.Solver2D(8, t_end=0.0)
    for bad in ({"velocity_relaxation": 0.0}, {"pressure_relaxation": 1.5},
                {"velocity_relaxation": float("nan")}, {"adaptive": True}):
        with pytest.raises(ValueError):
            O.Solver2D(8, coupling=O.COUPLING_SIMPLE, **bad)
