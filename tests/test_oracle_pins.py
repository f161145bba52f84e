"""Independent checks of the oracle's discretisation and assembly (CPU only).

The GPU kernels are compared with the oracle to round-off; these tests pin
the oracle itself, each with values that are derived here by hand or from a
closed form, never by re-running the oracle's own formula:

* momentum predictors (3D advective central form, incompressible_3d.rs:174-273;
  2D flux form with upwind/central transport, incompressible.rs:748-922):
  on low-order polynomial fields the central first and second differences and
  the 2- and 4-point face averages are exact (or carry a known O(h^2) term),
  so interior faces must equal the continuum expression; first-order upwind
  carries its known O(h) term with the sign set by the flow direction;
* wall ghosts (lid 2U - c, no-slip -c, symmetry/outlet c, blocked face -c;
  incompressible_3d.rs:366-447, incompressible.rs:2089-2199) on constant
  fields, where every term but the wall's is zero;
* velocity boundaries (incompressible.rs:1981-2044) on a hand-sized grid;
* cell fields (incompressible.rs:1381-1439) on a solid-body rotation:
  vorticity 2 omega, speed omega r;
* Poisson assembly (poisson.rs:146-337) against the reference's own unit
  tests (poisson.rs:627-748) and the hand-computed two-point coefficients of
  the unjittered triangulation (3 on diagonal faces, 6/5 on axis faces).
"""

import numpy as np
import pytest

import oracle as O


def close(a, b, tol=1e-12):
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= tol * scale, np.abs(a - b).max() / scale


# ------------------------------------------------------------- 3D predictor
NX, NY, NZ = 7, 6, 5
LX, LY, LZ = 1.3, 0.9, 1.1
RE, DT = 250.0, 1.0e-3
A = [0.3, 0.7, -0.4, 0.25, 0.9, -0.6, 0.45, 0.35, -0.2, 0.15]  # quadratic u
B = [-0.2, 0.5, 0.3, -0.35]  # affine v
Cw = [0.1, -0.45, 0.2, 0.6]  # affine w


def _u(x, y, z):
    a = A
    return (a[0] + a[1] * x + a[2] * y + a[3] * z + a[4] * x * x + a[5] * y * y
            + a[6] * z * z + a[7] * x * y + a[8] * y * z + a[9] * x * z)


def _aff(c, x, y, z):
    return c[0] + c[1] * x + c[2] * y + c[3] * z


def _coords(shape, offs, h):
    """Face-centre coordinates of a (nz, ny, nx)-shaped staggered array."""
    k, j, i = np.meshgrid(*[np.arange(n, dtype=float) for n in shape], indexing="ij")
    return (i + offs[0]) * h[0], (j + offs[1]) * h[1], (k + offs[2]) * h[2]


def _cavity3d():
    return O.Cavity3D(NX, NY, NZ, length=LX, height=LY, depth=LZ, reynolds=RE, dt=DT)


def test_3d_predictor_exact_on_polynomials():
    c = _cavity3d()
    h = (LX / NX, LY / NY, LZ / NZ)
    nu = 1.0 * LX / RE  # LidDrivenCavity3DConfig::viscosity, U L / Re
    xu, yu, zu = _coords((NZ, NY, NX + 1), (0.0, 0.5, 0.5), h)
    xv, yv, zv = _coords((NZ, NY + 1, NX), (0.5, 0.0, 0.5), h)
    xw, yw, zw = _coords((NZ + 1, NY, NX), (0.5, 0.5, 0.0), h)
    c.field("u")[:] = _u(xu, yu, zu)
    c.field("v")[:] = _aff(B, xv, yv, zv)
    c.field("w")[:] = _aff(Cw, xw, yw, zw)
    c.predict()
    a = A
    # u* on faces whose whole stencil is interior (j, k not next to a wall)
    x, y, z = xu, yu, zu
    ux = a[1] + 2 * a[4] * x + a[7] * y + a[9] * z
    uy = a[2] + 2 * a[5] * y + a[7] * x + a[8] * z
    uz = a[3] + 2 * a[6] * z + a[8] * y + a[9] * x
    lap = 2 * (a[4] + a[5] + a[6])
    U = _u(x, y, z)
    ex = U + DT * (-U * ux - _aff(B, x, y, z) * uy - _aff(Cw, x, y, z) * uz + nu * lap)
    sl = (slice(1, NZ - 1), slice(1, NY - 1), slice(1, NX))
    close(c.field("us")[sl], ex[sl])
    # v*: u-bar averages the quadratic u over an (x, y) square of half-sides
    # (dx/2, dy/2), which adds a4 (dx/2)^2 + a5 (dy/2)^2 to u at the centre
    x, y, z = xv, yv, zv
    ubar = _u(x, y, z) + a[4] * (h[0] / 2) ** 2 + a[5] * (h[1] / 2) ** 2
    V = _aff(B, x, y, z)
    ex = V + DT * (-ubar * B[1] - V * B[2] - _aff(Cw, x, y, z) * B[3])
    sl = (slice(1, NZ - 1), slice(1, NY), slice(1, NX - 1))
    close(c.field("vs")[sl], ex[sl])
    # w*: the (x, z) square adds a4 (dx/2)^2 + a6 (dz/2)^2
    x, y, z = xw, yw, zw
    ubar = _u(x, y, z) + a[4] * (h[0] / 2) ** 2 + a[6] * (h[2] / 2) ** 2
    W = _aff(Cw, x, y, z)
    ex = W + DT * (-ubar * Cw[1] - _aff(B, x, y, z) * Cw[2] - W * Cw[3])
    sl = (slice(1, NZ), slice(1, NY - 1), slice(1, NX - 1))
    close(c.field("ws")[sl], ex[sl])
    # boundary-normal faces are never predicted (left at zero)
    assert not c.field("us")[:, :, 0].any() and not c.field("us")[:, :, NX].any()
    assert not c.field("vs")[:, 0].any() and not c.field("ws")[0].any()


@pytest.mark.parametrize("vbar", [0.0, 0.4])
def test_3d_predictor_wall_ghosts(vbar):
    """u = c0, v = vbar, w = 0: only wall ghosts make terms non-zero.  Lid
    (y = top) ghost 2U - c, every other wall -c; v-bar carries the ghost
    into the y advection term."""
    c = _cavity3d()
    dx, dy, dz = LX / NX, LY / NY, LZ / NZ
    nu, c0, lid = 1.0 * LX / RE, 0.8, 1.0
    c.field("u")[:] = c0
    c.field("v")[:] = vbar
    c.field("w")[:] = 0.0
    c.predict()
    k, j = np.meshgrid(np.arange(NZ), np.arange(NY), indexing="ij")
    north = np.where(j == NY - 1, 2 * lid - c0, c0)
    south = np.where(j == 0, -c0, c0)
    front = np.where(k == NZ - 1, -c0, c0)
    back = np.where(k == 0, -c0, c0)
    ex = c0 + DT * (-vbar * (north - south) / (2 * dy)
                    + nu * ((north - 2 * c0 + south) / dy**2 + (front - 2 * c0 + back) / dz**2))
    got = c.field("us")[:, :, 1:NX]
    close(got, np.repeat(ex[:, :, None], NX - 1, axis=2))
    # corners of the lid row: top (2U - c) and a z wall (-c) together
    assert abs(got[0, NY - 1, 0] - (c0 + DT * (-vbar * (2 * lid - 2 * c0) / (2 * dy)
                                               + nu * ((2 * lid - 2 * c0) / dy**2
                                                       - 2 * c0 / dz**2)))) < 1e-12
    # v = c0 alone: no-slip ghosts on the x and z walls (the lid is u only)
    c.field("u")[:] = 0.0
    c.field("v")[:] = c0
    c.predict()
    kk, ii = np.meshgrid(np.arange(NZ), np.arange(NX), indexing="ij")
    ex = c0 + DT * nu * ((np.where(ii == 0, -2 * c0, 0) + np.where(ii == NX - 1, -2 * c0, 0)) / dx**2
                         + (np.where(kk == 0, -2 * c0, 0) + np.where(kk == NZ - 1, -2 * c0, 0)) / dz**2)
    for jj in range(1, NY):
        close(c.field("vs")[:, jj, :], ex)
    # w = c0 alone: no-slip ghosts on the x and y walls (also at the lid)
    c.field("v")[:] = 0.0
    c.field("w")[:] = c0
    c.predict()
    jj, ii = np.meshgrid(np.arange(NY), np.arange(NX), indexing="ij")
    ex = c0 + DT * nu * ((np.where(ii == 0, -2 * c0, 0) + np.where(ii == NX - 1, -2 * c0, 0)) / dx**2
                         + (np.where(jj == 0, -2 * c0, 0) + np.where(jj == NY - 1, -2 * c0, 0)) / dy**2)
    for kz in range(1, NZ):
        close(c.field("ws")[kz], ex)


# ------------------------------------------------------------- 2D predictor
NX2, NY2, L2, H2, NU2, DT2 = 8, 7, 1.2, 0.8, 0.02, 1.0e-3


def _solver2d(**kw):
    kw.setdefault("bc_kind", (O.BC_VELOCITY,) * 4)
    kw.setdefault("bc_u", (0.0,) * 4)
    return O.Solver2D(NX2, NY2, length=L2, height=H2, nu=NU2, dt=DT2, **kw)


def _faces2d():
    dx, dy = L2 / NX2, H2 / NY2
    ju, iu = np.meshgrid(np.arange(NY2), np.arange(NX2 + 1), indexing="ij")
    jv, iv = np.meshgrid(np.arange(NY2 + 1), np.arange(NX2), indexing="ij")
    return (iu * dx, (ju + 0.5) * dy), ((iv + 0.5) * dx, jv * dy), dx, dy


def test_2d_central_flux_form_exact_on_linear_fields():
    """Central transport of linear u, v: the fluxes ue*phi_e etc. are exact
    face values of the quadratic products, so the flux differences are the
    exact derivatives d(uu)/dx + d(vu)/dy; diffusion of linear fields is 0."""
    s = _solver2d(convection=O.CENTRAL)
    (xu, yu), (xv, yv), dx, dy = _faces2d()
    a0, a1, a2, b0, b1, b2 = 0.4, 0.6, -0.3, -0.1, 0.25, 0.5
    s.field("u")[:] = a0 + a1 * xu + a2 * yu
    s.field("v")[:] = b0 + b1 * xv + b2 * yv
    s.predict(DT2)
    U = a0 + a1 * xu + a2 * yu
    V = b0 + b1 * xu + b2 * yu
    ex = U + DT2 * (-(2 * U * a1) - (a2 * V + U * b2))
    close(s.field("us")[1:NY2 - 1, 1:NX2], ex[1:NY2 - 1, 1:NX2])
    U = a0 + a1 * xv + a2 * yv
    V = b0 + b1 * xv + b2 * yv
    ex = V + DT2 * (-(a1 * V + U * b1) - 2 * V * b2)
    close(s.field("vs")[1:NY2, 1:NX2 - 1], ex[1:NY2, 1:NX2 - 1])


@pytest.mark.parametrize("scheme", [O.UPWIND, O.CENTRAL])
def test_2d_diffusion_exact_on_quadratic(scheme):
    """u = 0, v = c x^2 (independent of y): no convection, v* = v + dt nu v_xx."""
    s = _solver2d(convection=scheme)
    (xu, yu), (xv, yv), dx, dy = _faces2d()
    cq = 0.7
    s.field("u")[:] = 0.0
    s.field("v")[:] = 0.2 + cq * xv**2
    s.predict(DT2)
    ex = 0.2 + cq * xv**2 + DT2 * NU2 * 2 * cq
    close(s.field("vs")[1:NY2, 1:NX2 - 1], ex[1:NY2, 1:NX2 - 1])
    assert not s.field("us").any()


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_2d_upwind_first_order_term(sign):
    """u = a + b x of one sign, v = 0.  First-order upwind takes the upstream
    value, so d(uu)/dx becomes 2 b u -+ b^2 dx / 2 (minus for flow in +x,
    plus for flow in -x) instead of the central 2 b u."""
    s = _solver2d(convection=O.UPWIND)
    (xu, yu), _, dx, dy = _faces2d()
    a, b = sign * 2.0, 0.5
    s.field("u")[:] = a + b * xu
    s.field("v")[:] = 0.0
    s.predict(DT2)
    u0 = a + b * xu
    ex = u0 - DT2 * (2 * b * u0 - sign * b * b * dx / 2)
    close(s.field("us")[1:NY2 - 1, 1:NX2], ex[1:NY2 - 1, 1:NX2])
    # the same along y for v
    s.field("u")[:] = 0.0
    _, (xv, yv), _, _ = _faces2d()
    s.field("v")[:] = a + b * yv
    s.predict(DT2)
    v0 = a + b * yv
    ex = v0 - DT2 * (2 * b * v0 - sign * b * b * dy / 2)
    close(s.field("vs")[1:NY2, 1:NX2 - 1], ex[1:NY2, 1:NX2 - 1])


def test_2d_wall_ghosts_and_blocked_faces():
    """Constant fields: a Velocity wall's ghost is 2 u_wall - c, a Symmetry
    (or PressureOutlet) wall's is c, a blocked neighbour face's is -c."""
    dx, dy = L2 / NX2, H2 / NY2
    c0, ub, vl = 0.8, 0.3, 0.2
    kinds = (O.BC_VELOCITY, O.BC_SYMMETRY, O.BC_VELOCITY, O.BC_SYMMETRY)
    s = _solver2d(convection=O.CENTRAL, bc_kind=kinds, bc_u=(0.0, 0.0, ub, 0.0),
                  bc_v=(vl, 0.0, 0.0, 0.0))
    s.field("u")[:] = c0
    s.field("v")[:] = 0.0
    s.predict(DT2)
    us = s.field("us")
    close(us[0, 1:NX2], np.full(NX2 - 1, c0 + DT2 * NU2 * (2 * ub - 2 * c0) / dy**2))
    close(us[NY2 - 1, 1:NX2], np.full(NX2 - 1, c0))  # symmetry top: ghost = c
    close(us[1:NY2 - 1, 1:NX2], np.full((NY2 - 2, NX2 - 1), c0))
    s.field("u")[:] = 0.0
    s.field("v")[:] = c0
    s.predict(DT2)
    vs = s.field("vs")
    close(vs[1:NY2, 0], np.full(NY2 - 1, c0 + DT2 * NU2 * (2 * vl - 2 * c0) / dx**2))
    close(vs[1:NY2, NX2 - 1], np.full(NY2 - 1, c0))  # symmetry right
    # one solid cell (3, 3): its faces are skipped, neighbours see -c
    solid = np.zeros((NY2, NX2), np.uint8)
    solid[3, 3] = 1
    s = _solver2d(convection=O.CENTRAL, solid=solid)
    s.field("u")[:] = c0
    s.field("v")[:] = 0.0
    s.predict(DT2)
    us = s.field("us")
    assert us[3, 3] == 0.0 and us[3, 4] == 0.0
    close(np.array([us[4, 3], us[2, 3]]), np.full(2, c0 - DT2 * NU2 * 2 * c0 / dy**2))
    # face (2, 3): east neighbour blocked -> -c0 in both transport and diffusion
    # ue = (c0 - c0)/2 = 0, uw = c0, phi_w = c0: convection -c0^2/dx
    assert abs(us[3, 2] - (c0 + DT2 * (c0 * c0 / dx - NU2 * 2 * c0 / dx**2))) < 1e-12
    assert abs(us[3, 5] - (c0 + DT2 * (-c0 * c0 / dx - NU2 * 2 * c0 / dx**2))) < 1e-12


# -------------------------------------------------------- velocity boundaries
def test_2d_velocity_boundaries_hand_grid():
    rng = np.random.default_rng(3)
    solid = np.zeros((5, 6), np.uint8)
    solid[2, 3] = 1
    s = O.Solver2D(6, 5, bc_kind=(O.BC_VELOCITY, O.BC_OUTLET, O.BC_SYMMETRY, O.BC_VELOCITY),
                   bc_u=(0.3, 0.0, 0.0, 1.0), bc_v=(0.0, 0.0, 0.0, 0.2), solid=solid)
    u0, v0 = rng.standard_normal((5, 7)), rng.standard_normal((6, 6))
    s.field("u")[:] = u0
    s.field("v")[:] = v0
    s.apply_boundaries()
    u, v = s.field("u"), s.field("v")
    assert (u[:, 0] == 0.3).all()                 # Velocity: the side's u
    assert (u[:, 6] == u0[:, 5]).all()            # PressureOutlet: copy the inner face
    assert (v[0, :] == 0.0).all()                 # Symmetry: zero normal velocity
    assert (v[5, :] == 0.2).all()                 # Velocity: the side's v
    assert u[2, 3] == 0.0 and u[2, 4] == 0.0      # faces of the solid cell
    assert v[2, 3] == 0.0 and v[3, 3] == 0.0
    keep_u = np.ones_like(u0, bool)
    keep_u[:, [0, 6]] = False
    keep_u[2, [3, 4]] = False
    assert (u[keep_u] == u0[keep_u]).all()
    keep_v = np.ones_like(v0, bool)
    keep_v[[0, 5], :] = False
    keep_v[[2, 3], 3] = False
    assert (v[keep_v] == v0[keep_v]).all()
    s = O.Solver2D(6, 5, bc_kind=(O.BC_SYMMETRY, O.BC_VELOCITY, O.BC_OUTLET, O.BC_SYMMETRY),
                   bc_u=(0.0, -0.1, 0.0, 0.0))
    s.field("u")[:] = u0
    s.field("v")[:] = v0
    s.apply_boundaries()
    u, v = s.field("u"), s.field("v")
    assert (u[:, 0] == 0.0).all() and (u[:, 6] == -0.1).all()
    assert (v[0, :] == v0[1, :]).all() and (v[5, :] == 0.0).all()


# ------------------------------------------------------------- cell fields
def test_cell_fields_solid_body_rotation():
    """u = -w (y - yc), v = w (x - xc): interior and one-sided differences of
    linear fields are exact, so vorticity = 2 w and speed = w r everywhere
    except where a cell has no open neighbour along an axis (that axis's
    derivative is 0, leaving w)."""
    nx, ny, lx, ly, om, xc, yc = 9, 7, 1.5, 1.0, 1.7, 0.6, 0.45
    solid = np.zeros((ny, nx), np.uint8)
    solid[3, 2] = solid[3, 4] = 1
    s = O.Solver2D(nx, ny, length=lx, height=ly, solid=solid)
    dx, dy = lx / nx, ly / ny
    ju, iu = np.meshgrid(np.arange(ny), np.arange(nx + 1), indexing="ij")
    jv, iv = np.meshgrid(np.arange(ny + 1), np.arange(nx), indexing="ij")
    s.field("u")[:] = -om * ((ju + 0.5) * dy - yc)
    s.field("v")[:] = om * ((iv + 0.5) * dx - xc)
    s.cell_fields()
    jc, ic = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    xcell, ycell = (ic + 0.5) * dx - xc, (jc + 0.5) * dy - yc
    fluid = solid == 0
    close(s.field("speed")[fluid], (om * np.hypot(xcell, ycell))[fluid])
    close(s.field("ucell")[fluid], (-om * ycell)[fluid])
    vort = s.field("vort")
    expect = np.where(fluid, 2 * om, 0.0)
    expect[3, 3] = om  # both x neighbours solid: d(v)/dx = 0
    close(vort, expect)
    assert s.field("speed")[3, 2] == 0.0


# ------------------------------------------------------------ Poisson assembly
def _square():
    return O.Mesh2D([(0, 0), (1, 0), (1, 1), (0, 1)], [[0, 1, 2, 3]])


def _two_squares():
    return O.Mesh2D([(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1)],
                    [[0, 1, 4, 3], [1, 2, 5, 4]])


def _grid(nx, ny):
    pts = [(x / nx, y / ny) for y in range(ny + 1) for x in range(nx + 1)]
    idx = lambda x, y: x + (nx + 1) * y  # noqa: E731
    cells = [[idx(x, y), idx(x + 1, y), idx(x + 1, y + 1), idx(x, y + 1)]
             for y in range(ny) for x in range(nx)]
    return O.Mesh2D(pts, cells)


def _solve(A, rhs):
    x = np.zeros(A.nrows)
    rep = O.pcg(A, rhs, x, 1e-13, 1e-13, 10000)
    assert rep.converged
    assert np.abs(rhs - A.spmv(x)).max() < 1e-10
    return x


def _dense(A):
    d = np.zeros((A.nrows, A.nrows))
    for r in range(A.nrows):
        for k in range(A.offsets[r], A.offsets[r + 1]):
            d[r, A.cols[k]] = A.vals[k]
    return d


def test_poisson_single_cell_dirichlet_and_source():
    """poisson.rs:627-639: four unit faces at distance 1/2 give a_b = 2, so
    A_00 = 8 and b_0 = 8 * 3 + 2 * 1 = 26."""
    m = _square()
    bc = {f: ("FixedValue", 3.0) for f in m.boundary_faces}
    A, rhs = O.assemble_poisson(m, 1.0, source=[2.0], boundary=bc)
    assert _dense(A)[0, 0] == 8.0 and rhs.tolist() == [26.0]


def test_poisson_two_point_block():
    """poisson.rs:642-654: the internal face couples the two cells by -1."""
    m = _two_squares()
    bc = {f: ("FixedValue", 0.0) for f in m.boundary_faces}
    A, _ = O.assemble_poisson(m, 1.0, boundary=bc)
    d = _dense(A)
    assert d[0, 1] == -1.0 and d[1, 0] == -1.0 and d[0, 0] > 0 and (d == d.T).all()


@pytest.mark.parametrize("fn", [lambda x, y: 7.25 + 0 * x, lambda x, y: 2 * x - 3 * y + 5])
def test_poisson_constant_and_linear_dirichlet(fn):
    """poisson.rs:657-674: grid 4 x 3 with every boundary face its own
    FixedValue patch at the face centre; two-point flux is exact for linear
    fields on this orthogonal grid."""
    m = _grid(4, 3)
    bc = {f: ("FixedValue", float(fn(*m.face_centers[f]))) for f in m.boundary_faces}
    A, rhs = O.assemble_poisson(m, 1.0, boundary=bc)
    x = _solve(A, rhs)
    assert np.abs(x - fn(m.centers[:, 0], m.centers[:, 1])).max() < 1e-10


def test_poisson_pure_neumann_reference_and_compatibility():
    """poisson.rs:677-707."""
    m = _two_squares()
    with pytest.raises(O.PoissonError) as e:
        O.assemble_poisson(m, 1.0)
    assert e.value.kind == "MissingReferenceForPureNeumann"
    with pytest.raises(O.PoissonError) as e:
        O.assemble_poisson(m, 1.0, source=[1.0, 1.0], reference=(0, 3.2))
    assert e.value.kind == "IncompatiblePureNeumannSource" and e.value.detail == 2.0
    A, rhs = O.assemble_poisson(m, 1.0, reference=(0, 3.2))
    d = _dense(A)
    assert (d == d.T).all()
    assert _solve(A, rhs).tolist() == [3.2, 3.2]
    with pytest.raises(O.PoissonError) as e:
        O.assemble_poisson(m, 1.0, reference=(2, 0.0))
    assert e.value.kind == "InvalidReferenceCell"
    with pytest.raises(O.PoissonError) as e:
        O.assemble_poisson(m, 1.0, reference=(0, np.nan))
    assert e.value.kind == "InvalidReferenceValue"


def test_poisson_zero_gradient_and_mixed():
    """poisson.rs:710-722 (zero gradient + pin: A = [1], b = [0]) and
    :724-748 (one FixedValue(4) face, others ZeroGradient -> [4])."""
    m = _square()
    A, rhs = O.assemble_poisson(m, 1.0, reference=(0, 0.0))
    assert _dense(A).tolist() == [[1.0]] and rhs.tolist() == [0.0]
    bc = {int(m.boundary_faces[0]): ("FixedValue", 4.0)}
    A, rhs = O.assemble_poisson(m, 1.0, boundary=bc)
    assert _solve(A, rhs).tolist() == [4.0]
    with pytest.raises(O.PoissonError) as e:
        O.assemble_poisson(m, 1.0, source=[np.inf], reference=(0, 0.0))
    assert e.value.kind == "NonFiniteSource"


def test_unjittered_triangulation_coefficients():
    """Right isosceles triangles with legs h: centroids differ by (h/3, -h/3)
    across the diagonal (|S| = h sqrt 2, S.d / |d|^2 = 3) and by (2h/3, h/3)
    across an axis face (S.d / |d|^2 = (2/3) / (5/9) = 6/5)."""
    m = 5
    mesh = O.mesh_poisson(m, 0.0, 0, 0)
    coef = mesh["face_coef"]
    assert np.isclose(coef, 3.0, rtol=0, atol=1e-12).sum() == m * m
    assert np.isclose(coef, 1.2, rtol=0, atol=1e-12).sum() == 2 * m * (m - 1)
    assert len(coef) == m * m + 2 * m * (m - 1)
    A = _dense(mesh["A"])
    # away from the pinned cell 0 each row sums to zero (pure Neumann rows)
    nb0 = set(mesh["face_neigh"][mesh["face_owner"] == 0]) | set(
        mesh["face_owner"][mesh["face_neigh"] == 0])
    rows = [r for r in range(1, 2 * m * m) if r not in nb0]
    assert np.abs(A[rows].sum(axis=1)).max() < 1e-12
    # a neighbour of the pin keeps its diagonal but loses the coupling to 0
    for r in nb0:
        assert abs(A[r].sum() - float(coef[(mesh["face_owner"] == 0) & (mesh["face_neigh"] == r)
                                           | (mesh["face_owner"] == r) & (mesh["face_neigh"] == 0)][0])) < 1e-12
    # the pin: row 0 is the identity, column 0 is dropped
    assert A[0].tolist() == [1.0] + [0.0] * (2 * m * m - 1) and not A[1:, 0].any()
    assert np.abs(mesh["volumes"] - 0.5 / (m * m)).max() < 1e-15


def test_reference_value_shifts_the_neumann_solution():
    """apply_reference with value v (poisson.rs:303-337): the pinned
    pure-Neumann solution moves by exactly v, and the general assembly of the
    triangulation reproduces the benchmark producer's matrix bit for bit."""
    m, jitter = 6, 0.2
    rng = np.random.default_rng(4)
    h = 1.0 / m
    pts = []
    for b in range(m + 1):
        for a in range(m + 1):
            x, y = a * h, b * h
            if 0 < a < m and 0 < b < m:
                x, y = x + rng.uniform(-jitter, jitter) * h, y + rng.uniform(-jitter, jitter) * h
            pts.append((x, y))
    cells = []
    for b in range(m):
        for a in range(m):
            v00 = a + (m + 1) * b
            cells += [[v00, v00 + 1, v00 + m + 2], [v00, v00 + m + 2, v00 + m + 1]]
    mesh = O.Mesh2D(pts, cells)
    src = rng.standard_normal(len(cells))
    src -= (src * mesh.volumes).sum() / mesh.volumes.sum()
    A0, b0 = O.assemble_poisson(mesh, 1.0, source=src, reference=(5, 0.0))
    A1, b1 = O.assemble_poisson(mesh, 1.0, source=src, reference=(5, 2.5))
    assert np.array_equal(A0.vals, A1.vals)
    x0, x1 = _solve(A0, b0), _solve(A1, b1)
    assert abs(x0[5]) < 1e-12 and abs(x1[5] - 2.5) < 1e-12
    assert np.abs((x1 - x0) - 2.5).max() < 1e-9
    # the producer of the benchmark system is this assembly on its triangulation
    ref = O.mesh_poisson(m, 0.0, 0, 0)
    flat = O.Mesh2D([(a * h, b * h) for b in range(m + 1) for a in range(m + 1)], cells)
    A, rhs = O.assemble_poisson(flat, 1.0, reference=(0, 0.0))
    assert np.array_equal(A.offsets, ref["offsets"]) and np.array_equal(A.vals, ref["vals"])
    assert not rhs.any()
