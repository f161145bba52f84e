"""Pins of the oracle's caller-shaped CSR assemblies (oracle/simple.c): the
SIMPLE pressure correction (src/simple.rs:342-370, 518-579) and one momentum
component (src/momentum.rs:180-341), against the reference's own hand-computed
tests (src/simple.rs:995-1088, src/momentum.rs:719-765, 767-783, 870-886,
1054-1119, 1476-1510), against the textbook properties of the systems on a
skewed triangulation (symmetry and conservation, upwind M-matrix signs), and
with planted mistakes that must fail them."""

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O


def two_cells():
    """The two unit squares of the reference tests (simple.rs:817-837, momentum.rs:484-513)."""
    pts = [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)]
    return O.Mesh2D(np.array(pts, float), [[0, 1, 4, 3], [1, 2, 5, 4]])


def channel(ny):
    """momentum.rs:530-560: one column of ny unit-wide cells."""
    pts = []
    for y in range(ny + 1):
        pts += [(0.0, y / ny), (1.0, y / ny)]
    cells = [[2 * y, 2 * y + 1, 2 * (y + 1) + 1, 2 * (y + 1)] for y in range(ny)]
    return O.Mesh2D(np.array(pts), cells)


def dense(A):
    return sp.csr_matrix((A.vals, A.cols, A.offsets), shape=(A.nrows, A.ncols)).toarray()


def internal(mesh):
    return int(np.nonzero(mesh.neighbour >= 0)[0][0])


def all_boundary(mesh, bc):
    return {int(f): bc for f in mesh.boundary_faces}


# ------------------------------------------------------------------ SIMPLE
def test_pressure_face_coefficient_two_cell_hand_value():
    """simple.rs:995-1011: rAU_f = (0.5 + 1.5) / 2 = 1, |S| = |d| = 1."""
    m = two_cells()
    c = O.pressure_face_coefficients(m, [0.5, 1.5])
    f = internal(m)
    assert c[f] == 1.0
    assert np.all(np.delete(c, f) == 0.0)


def test_pressure_correction_two_cell_matrix_and_rhs():
    """simple.rs:1054-1088: pinned cell 0, flux 3 and d_f 2 on the internal
    face give A = diag(1, 2) with no off-diagonal entries and rhs = [0, 3]."""
    m = two_cells()
    f = internal(m)
    flux = np.zeros(m.m.nfaces)
    coef = np.zeros(m.m.nfaces)
    flux[f], coef[f] = 3.0, 2.0
    A, rhs = O.assemble_pressure_correction(m, flux, coef, 0)
    assert np.array_equal(dense(A), [[1.0, 0.0], [0.0, 2.0]])
    assert A.offsets[-1] == 2  # the pinned row and column carry no entries
    assert list(rhs) == [0.0, 3.0]


def test_pressure_correction_rejects_bad_reference_and_coefficients():
    m = two_cells()
    flux, coef = np.zeros(m.m.nfaces), np.zeros(m.m.nfaces)
    with pytest.raises(O.CallerSystemError):
        O.assemble_pressure_correction(m, flux, coef, 2)  # InvalidPressureReference
    with pytest.raises(O.CallerSystemError):
        O.assemble_pressure_correction(m, flux, coef, 0)  # internal d_f = 0 is invalid


# ---------------------------------------------------------------- momentum
@pytest.mark.parametrize("phi", [2.0, -2.0])
def test_momentum_upwind_coefficients_two_cell_hand_values(phi):
    """momentum.rs:719-765: nu = 1 (D = 1), ZeroGradient walls, internal flux
    +-2: a_PP = 1 + max(phi, 0), a_NN = 1 - min(phi, 0),
    a_PN = -1 + min(phi, 0), a_NP = -1 - max(phi, 0)."""
    m = two_cells()
    f = internal(m)
    P, N = int(m.owner[f]), int(m.neighbour[f])
    flux = np.zeros(m.m.nfaces)
    flux[f] = phi
    A, rhs, _ = O.assemble_momentum(m, 1.0, flux, all_boundary(m, "ZeroGradient"))
    D = dense(A)
    assert D[P, P] == 1.0 + max(phi, 0.0)
    assert D[N, N] == 1.0 - min(phi, 0.0)
    assert D[P, N] == -1.0 + min(phi, 0.0)
    assert D[N, P] == -1.0 - max(phi, 0.0)
    assert np.all(rhs == 0.0)


def test_momentum_backward_euler_and_source_two_cell_hand_values():
    """momentum.rs:767-783: V = 1, dt = 0.5 give a_t = 2; b = a_t u_old + s V."""
    m = two_cells()
    A, rhs, _ = O.assemble_momentum(m, 0.0, np.zeros(m.m.nfaces),
                                    all_boundary(m, ("FixedValue", 0.0)), source=[4.0, 4.0],
                                    old=[3.0, -2.0], dt=0.5)
    assert np.array_equal(np.diag(dense(A)), [2.0, 2.0])
    assert list(rhs) == [10.0, 0.0]


def test_momentum_under_relaxation_two_cell_hand_values():
    """momentum.rs:870-886: a_P = 2 relaxed by alpha 0.5 -> 4; b = 2*3 + 2*3 = 12."""
    m = two_cells()
    A, rhs, diag = O.assemble_momentum(m, 0.0, np.zeros(m.m.nfaces),
                                       all_boundary(m, ("FixedValue", 0.0)), old=[3.0, 3.0],
                                       dt=0.5, relaxation_field=[3.0, 3.0],
                                       under_relaxation=0.5)
    assert list(diag) == [2.0, 2.0]
    assert dense(A)[0, 0] == 4.0
    assert list(rhs) == [12.0, 12.0]


def test_momentum_fixed_inflow_and_zero_gradient_outflow_hand_values():
    """momentum.rs:1054-1090 (inflow u_b = 5 with phi = -2: b = 10, a = 1)
    and :1092-1119 (ZeroGradient outflow phi = 2: a = 1 + 2, b = 0)."""
    m = channel(1)
    left = [int(f) for f in m.boundary_faces if m.face_centers[f, 0] < 1e-12][0]
    right = [int(f) for f in m.boundary_faces if abs(m.face_centers[f, 0] - 1) < 1e-12][0]
    bc = all_boundary(m, ("FixedValue", 0.0))
    bc[left] = ("FixedValue", 5.0)
    flux = np.zeros(m.m.nfaces)
    flux[left] = -2.0
    A, rhs, _ = O.assemble_momentum(m, 0.0, flux, bc, old=[0.0], dt=1.0)
    assert list(rhs) == [10.0] and dense(A)[0, 0] == 1.0
    flux = np.zeros(m.m.nfaces)
    flux[right] = 2.0
    A, rhs, _ = O.assemble_momentum(m, 0.0, flux, all_boundary(m, "ZeroGradient"), old=[0.0],
                                    dt=1.0)
    assert dense(A)[0, 0] == 3.0 and list(rhs) == [0.0]
    flux[right] = -2.0  # BackflowOnZeroGradient (:1121-1143)
    with pytest.raises(O.CallerSystemError):
        O.assemble_momentum(m, 0.0, flux, all_boundary(m, "ZeroGradient"), old=[0.0], dt=1.0)


def test_momentum_fixed_boundary_diffusion_hand_values():
    """momentum.rs:1476-1510: four boundary faces with D_b = 2, u_b = 3 on
    the left only: a_PP = 8, b = 6."""
    m = channel(1)
    left = [int(f) for f in m.boundary_faces if m.face_centers[f, 0] < 1e-12][0]
    bc = all_boundary(m, ("FixedValue", 0.0))
    bc[left] = ("FixedValue", 3.0)
    A, rhs, _ = O.assemble_momentum(m, 1.0, np.zeros(m.m.nfaces), bc)
    assert dense(A)[0, 0] == 8.0 and list(rhs) == [6.0]


# ------------------------------------------------- properties on a skewed mesh
def skewed_triangulation(n, seed):
    return O.triangulated_square(n, 0.2, seed)


def test_pressure_correction_properties_on_a_skewed_mesh():
    """SPD-shaped: symmetric, positive diagonal, non-positive off-diagonals,
    rows other than the pinned one summing to zero except through the pinned
    column; rhs = -div(Phi*) (sums to zero before the pin)."""
    m = skewed_triangulation(6, 3)
    rng = np.random.default_rng(1)
    r_au = rng.uniform(0.5, 2.0, m.m.ncells)
    coef = O.pressure_face_coefficients(m, r_au)
    lam = O.face_weights(m)
    f = internal(m)
    P, N = int(m.owner[f]), int(m.neighbour[f])
    d = m.centers[N] - m.centers[P]
    hand = ((1 - lam[f]) * r_au[P] + lam[f] * r_au[N]) * (m.area_vectors[f] @ d) / (d @ d)
    assert coef[f] == pytest.approx(hand, rel=1e-15)
    flux = rng.standard_normal(m.m.nfaces)
    A, rhs = O.assemble_pressure_correction(m, flux, coef, 5)
    D = dense(A)
    assert np.array_equal(D, D.T)
    assert np.all(np.diag(D) > 0)
    off = D - np.diag(np.diag(D))
    assert np.all(off <= 0)
    div = np.zeros(m.m.ncells)
    np.add.at(div, m.owner, flux)
    inner = m.neighbour >= 0
    np.add.at(div, m.neighbour[inner], -flux[inner])
    expect = -div
    expect[5] = 0.0
    assert np.allclose(rhs, expect, atol=1e-14)
    rows = np.delete(np.arange(m.m.ncells), 5)
    # a full row of the unpinned operator sums to zero: here only the pinned
    # neighbour's entry (dropped) is missing
    full = np.zeros((m.m.ncells, m.m.ncells))
    for g in np.nonzero(inner)[0]:
        a, b, c = m.owner[g], m.neighbour[g], coef[g]
        full[a, a] += c; full[b, b] += c; full[a, b] -= c; full[b, a] -= c
    assert np.allclose(D[np.ix_(rows, rows)], full[np.ix_(rows, rows)], atol=1e-13)


def test_momentum_properties_on_a_skewed_mesh():
    """Upwind convection + two-point diffusion on a skewed triangulation:
    positive diagonal equal to the reported unrelaxed diagonal, non-positive
    off-diagonals (an M-matrix shape for any flux sign), and with zero flux
    the diffusion operator maps the uniform field onto the boundary terms
    (a uniform field is the solution for uniform FixedValue walls)."""
    m = skewed_triangulation(6, 4)
    rng = np.random.default_rng(2)
    flux = np.zeros(m.m.nfaces)
    inner = np.nonzero(m.neighbour >= 0)[0]
    flux[inner] = 0.3 * rng.standard_normal(len(inner))
    bc = all_boundary(m, ("FixedValue", 1.0))
    A, rhs, diag = O.assemble_momentum(m, 0.05, flux, bc)
    D = dense(A)
    assert np.all(np.diag(D) > 0) and np.array_equal(np.diag(D), diag)
    assert np.all(D - np.diag(np.diag(D)) <= 0)
    A0, rhs0, _ = O.assemble_momentum(m, 0.05, np.zeros(m.m.nfaces), bc)
    assert np.allclose(dense(A0) @ np.ones(m.m.ncells), rhs0, rtol=1e-13, atol=1e-13)


def test_face_weights_are_one_half_on_the_two_cell_mesh():
    """The shared edge's centre sits midway between the two square centres."""
    m = two_cells()
    lam = O.face_weights(m)
    assert lam[internal(m)] == 0.5
    assert np.all(np.delete(lam, internal(m)) == 0.0)


def test_planted_mistakes_fail_the_hand_pins():
    """A transposed upwind pair or a dropped transient term must fail the
    hand values above (the checks pin the formula, not just its shape)."""
    m = two_cells()
    f = internal(m)
    P, N = int(m.owner[f]), int(m.neighbour[f])
    flux = np.zeros(m.m.nfaces)
    flux[f] = 2.0
    A, _, _ = O.assemble_momentum(m, 1.0, flux, all_boundary(m, "ZeroGradient"))
    D = dense(A)
    assert D[P, N] != D.T[P, N]  # a transposed assembly would swap these
    A, rhs, _ = O.assemble_momentum(m, 0.0, np.zeros(m.m.nfaces),
                                    all_boundary(m, ("FixedValue", 0.0)), source=[4.0, 4.0],
                                    old=[0.0, 0.0], dt=0.5)
    assert list(rhs) != [10.0, 0.0]  # a dropped a_t u_old term fails the pin
    # with no transient and no viscosity the rows have no diagonal at all:
    # CsrBuilder::finalize drops zero sums (src/linear/csr.rs:182-185) and
    # CsrMatrix::diagonal reports MissingDiagonal (:61-76)
    with pytest.raises(O.CallerSystemError, match="missing diagonal"):
        O.assemble_momentum(m, 0.0, np.zeros(m.m.nfaces), all_boundary(m, ("FixedValue", 0.0)),
                            source=[4.0, 4.0])
