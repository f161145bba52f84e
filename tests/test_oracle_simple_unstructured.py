"""The oracle's collocated unstructured SIMPLE (oracle/simple2d.c) pinned on
the reference's own hand tests of src/simple.rs (:989-1689) and on what the
mathematics fixes:

* weighted least squares reproduces the gradient of a linear field exactly
  on any mesh whose boundary carries the field's values;
* the reference's two-cell hand values: velocity correction sign, continuity
  RMS, initial face flux, the one-iteration removal of an interior imbalance,
  a leaking wall that cannot converge, the zero cavity that stays zero;
* Rhie-Chow: no flux from a constant pressure, the consistent flux for a
  linear pressure, a checkerboard the direct pressure term sees;
* the 3 x 3 lid-driven cavity (straight and skewed) converges without wall
  leakage, as the reference's regression tests require.
"""

import numpy as np
import pytest

import oracle as O


def two_cells():
    pts = np.array([[0, 0], [1, 0], [2, 0], [0, 1], [1, 1], [2, 1]], float)
    return O.Mesh2D(pts, [[0, 1, 4, 3], [1, 2, 5, 4]])


def grid(nx, ny, skew=0.0):
    pts = np.array([[x + skew * y, y] for y in range(ny + 1) for x in range(nx + 1)], float)
    p = lambda x, y: x + (nx + 1) * y  # noqa: E731
    cells = [[p(x, y), p(x + 1, y), p(x + 1, y + 1), p(x, y + 1)]
             for y in range(ny) for x in range(nx)]
    return O.Mesh2D(pts, cells)


def boundary_faces(m):
    return [f for f in range(m.m.nfaces) if m.neighbour[f] < 0]


def walls(m, value=0.0):
    return {f: ("FixedValue", value) for f in boundary_faces(m)}


def no_slip(m):
    return {f: (0.0, 0.0) for f in boundary_faces(m)}


def test_lsq_gradient_exact_for_linear_fields():
    m = O.triangulated_square(6, 0.25, 3)
    fn = lambda x, y: 2.5 * x - 1.25 * y + 0.75  # noqa: E731
    bnd = {f: ("FixedValue", fn(*m.face_centers[f])) for f in boundary_faces(m)}
    st = O.LsqStencil(m, bnd)
    g = st.gradient(fn(m.centers[:, 0], m.centers[:, 1]))
    assert np.abs(g - [2.5, -1.25]).max() < 1e-12
    # ZeroGradient faces contribute no equation: interior cells keep the exact
    # gradient (a quad grid: a triangle with two wall edges has one equation
    # and a singular 2D stencil, as in the reference)
    m = grid(5, 4, 0.1)
    st2 = O.LsqStencil(m, {f: "ZeroGradient" for f in boundary_faces(m)})
    g2 = st2.gradient(fn(m.centers[:, 0], m.centers[:, 1]))
    interior = [c for c in range(m.m.ncells)
                if all(m.neighbour[f] >= 0 for f in range(m.m.nfaces)
                       if m.owner[f] == c or m.neighbour[f] == c)]
    assert interior and np.abs(g2[interior] - [2.5, -1.25]).max() < 1e-12


def test_lsq_stencil_errors():
    m = two_cells()
    with pytest.raises(O.SimpleError) as e:
        O.LsqStencil(m, {})
    assert e.value.kind == "MissingBoundaryCondition"
    # a single cell with every face ZeroGradient has no equation: singular
    one = O.Mesh2D(np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float), [[0, 1, 2, 3]])
    with pytest.raises(O.SimpleError) as e:
        O.LsqStencil(one, {f: "ZeroGradient" for f in range(4)})
    assert e.value.kind == "SingularLeastSquaresStencil"


def test_velocity_correction_has_the_negative_gradient_sign():
    """velocity_pressure_correction_has_the_derived_negative_gradient_sign (:1139-1157), x and y."""
    out = O.correct_cell_velocity([[4.0, -1.0], [1.0, 3.0]], [0.5, 2.0], [[2.0, -4.0], [-1.0, 0.5]])
    assert np.array_equal(out, [[3.0, 1.0], [3.0, 2.0]])
    with pytest.raises(O.SimpleError) as e:
        O.correct_cell_velocity([[0.0, 0.0]], [0.0], [[0.0, 0.0]])
    assert e.value.kind == "InvalidMomentumDiagonal"


def test_continuity_rms_and_initial_flux_by_hand():
    """continuity_rms_is_the_direct_integrated_flux_imbalance_norm (:1159-1174) and
    initial_face_flux_interpolates_cell_velocity_but_preserves_wall_fluxes (:1176-1199)."""
    m = two_cells()
    flux = np.where(m.neighbour >= 0, 3.0, 0.0)
    assert O.continuity_rms(m, flux) == 3.0
    wall = np.where(m.neighbour < 0, 4.0, 0.0)
    out = O.initial_face_flux(m, [[1.0, 0.0], [1.0, 0.0]], wall)
    for f in range(m.m.nfaces):
        expect = m.area_vectors[f, 0] if m.neighbour[f] >= 0 else 4.0
        assert out[f] == expect


def simple_two_cells(flux, **kw):
    m = two_cells()
    st = O.LsqStencil(m, walls(m, 0.0))
    z = np.zeros((2, 2))
    return m, O.solve_simple(m, st, no_slip(m), z, np.zeros(2), flux, 1.0, **kw)


def test_zero_cavity_stays_zero():
    """simple_converges_a_zero_velocity_closed_cavity_without_creating_pressure_or_flux (:1201-1234)."""
    m, (rep, v, p, phi, hist) = simple_two_cells(np.zeros(7))
    assert rep["converged"] and rep["final_continuity_rms"] < 1e-12
    assert not v.any() and not p.any() and not phi.any()


def test_interior_imbalance_removed_in_one_iteration():
    """simple_uses_the_corrected_face_flux_to_remove_a_nonzero_initial_two_cell_imbalance (:1236-1281)."""
    m = two_cells()
    flux = np.where(m.neighbour >= 0, 3.0, 0.0)
    _, (rep, v, p, phi, hist) = simple_two_cells(flux, velocity_relaxation=1.0)
    assert rep["converged"] and rep["outer_iterations"] == 1
    assert rep["initial_continuity_rms"] == 3.0 and rep["final_continuity_rms"] < 1e-12
    assert np.abs(phi).max() < 1e-12
    assert len(hist) == 2 and hist[0] == 3.0


def test_leaking_wall_reports_max_iterations():
    """simple_reports_max_iterations_when_an_immutable_boundary_flux_prevents_continuity (:1283-1328)."""
    m = two_cells()
    flux = np.zeros(m.m.nfaces)
    flux[boundary_faces(m)[0]] = 1.0
    _, (rep, *_rest) = simple_two_cells(flux, max_outer_iterations=1, continuity_tol=1e-12)
    assert not rep["converged"] and rep["outer_iterations"] == 1
    assert rep["final_continuity_rms"] > 1e-12


def test_invalid_options_are_rejected():
    """simple_rejects_an_invalid_pressure_relaxation (:1330-1358)."""
    m = two_cells()
    for kw, kind in (({"pressure_relaxation": 0.0}, "InvalidPressureRelaxation"),
                     ({"pressure_relaxation": 1.5}, "InvalidPressureRelaxation"),
                     ({"max_outer_iterations": 0}, "InvalidOuterIterations"),
                     ({"continuity_tol": -1.0}, "InvalidPressureRelaxation"),
                     ({"reference_cell": 2}, "InvalidPressureReference")):
        with pytest.raises(O.SimpleError) as e:
            simple_two_cells(np.zeros(m.m.nfaces), **kw)
        assert e.value.kind == kind


def lid_cavity(m, top):
    vel = {}
    for f in boundary_faces(m):
        vel[f] = (1.0, 0.0) if abs(m.face_centers[f, 1] - top) < 1e-12 else (0.0, 0.0)
    return vel


@pytest.mark.parametrize("skew,iters,tol", [(0.0, 300, 1e-9), (0.2, 500, 1e-8)])
def test_low_reynolds_cavity_converges_without_wall_leakage(skew, iters, tol):
    """low_reynolds_closed_cavity_simple_converges_without_wall_leakage (:1406-1473) and
    skewed_low_reynolds_closed_cavity_has_finite_converged_simple_coupling (:1475-1537)."""
    m = grid(3, 3, skew)
    st = O.LsqStencil(m, {f: "ZeroGradient" for f in boundary_faces(m)})
    z = np.zeros((m.m.ncells, 2))
    rep, v, p, phi, hist = O.solve_simple(m, st, lid_cavity(m, 3.0), z, np.zeros(m.m.ncells),
                                          np.zeros(m.m.nfaces), 0.1, max_outer_iterations=iters,
                                          continuity_tol=tol, momentum_solver=(1e-12, 1e-8, 1000),
                                          pressure_solver=(1e-12, 1e-8, 1000))
    assert rep["converged"] and rep["final_continuity_rms"] < tol
    assert (v ** 2).sum(1).max() > 1e-12 and np.isfinite(v).all() and np.isfinite(p).all()
    assert np.abs(phi[boundary_faces(m)]).max() < 1e-12
    # the history is the RMS after each outer iteration, ending at the reported value
    assert hist[-1] == rep["final_continuity_rms"] and len(hist) == rep["outer_iterations"] + 1


def test_rhie_chow_constant_linear_and_checkerboard():
    """rhie_chow_* (:1600-1688)."""
    m = two_cells()
    st = O.LsqStencil(m, walls(m, 7.0))
    flux = O.rhie_chow_flux(m, st, np.zeros((2, 2)), np.full(2, 7.0), np.ones(2), np.zeros(7))
    assert np.abs(flux).max() < 1e-12
    g = grid(3, 3)
    fn = lambda x, y: 2.0 * x - 3.0 * y + 5.0  # noqa: E731
    st = O.LsqStencil(g, {f: ("FixedValue", fn(*g.face_centers[f])) for f in boundary_faces(g)})
    p = fn(g.centers[:, 0], g.centers[:, 1])
    flux = O.rhie_chow_flux(g, st, np.tile([-2.0, 3.0], (9, 1)), p, np.ones(9), np.zeros(g.m.nfaces))
    inner = g.neighbour >= 0
    expect = -2.0 * g.area_vectors[:, 0] + 3.0 * g.area_vectors[:, 1]
    assert np.abs(flux[inner] - expect[inner]).max() < 1e-12
    st = O.LsqStencil(g, walls(g, 0.0))
    cb = np.array([1.0 if (c % 3 + c // 3) % 2 == 0 else -1.0 for c in range(9)])
    flux = O.rhie_chow_flux(g, st, np.zeros((9, 2)), cb, np.ones(9), np.zeros(g.m.nfaces))
    assert np.abs(flux[inner]).max() > 1e-8


def test_pressure_correction_boundary_anchors_an_open_channel():
    """With a FixedValue(0) correction face (a pressure outlet,
    unstructured_incompressible.rs:158-166) the correction needs no reference
    cell; an inlet flux through the opposite wall leaves through the outlet."""
    m = grid(4, 2)
    left = [f for f in boundary_faces(m) if abs(m.face_centers[f, 0]) < 1e-12]
    right = [f for f in boundary_faces(m) if abs(m.face_centers[f, 0] - 4.0) < 1e-12]
    vel = {f: (0.0, 0.0) for f in boundary_faces(m)}
    for f in left:
        vel[f] = (1.0, 0.0)
    for f in right:
        vel[f] = "ZeroGradient"
    corr = {f: "ZeroGradient" for f in boundary_faces(m)}
    for f in right:
        corr[f] = ("FixedValue", 0.0)
    pb = {f: "ZeroGradient" for f in boundary_faces(m)}
    for f in right:
        pb[f] = ("FixedValue", 0.0)
    st = O.LsqStencil(m, pb)
    wall = np.zeros(m.m.nfaces)
    for f in left:
        wall[f] = np.dot([1.0, 0.0], m.area_vectors[f])
    flux0 = O.initial_face_flux(m, np.zeros((8, 2)), wall)
    rep, v, p, phi, hist = O.solve_simple(m, st, vel, np.zeros((8, 2)), np.zeros(8), flux0, 0.5,
                                          correction_boundary=corr, max_outer_iterations=400,
                                          continuity_tol=1e-9, momentum_solver=(1e-12, 1e-8, 1000),
                                          pressure_solver=(1e-12, 1e-8, 1000))
    assert rep["converged"]
    inflow = -phi[left].sum()
    assert inflow == pytest.approx(2.0) and phi[right].sum() == pytest.approx(inflow, rel=1e-8)
