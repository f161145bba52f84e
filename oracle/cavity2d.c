/*
 * Oracle: 2D structured projection step on a staggered MAC grid (FP64).
 * TEST INFRASTRUCTURE ONLY -- the checker, never the shipped path.
 *
 * Restates src/solver/incompressible.rs (projection coupling, flow only):
 *   advance_one_step          :622-746
 *   predict_momentum          :748-922  (+ MomentumStencil :2089-2157,
 *                                         tangential ghosts :2159-2199,
 *                                         transported :2201-2212)
 *   build_pressure_rhs        :924-970
 *   solve_pressure_sor        :979-1046
 *   solve_pressure_pcg        :1048-1214
 *   apply_pressure_operator   :1216-1255
 *   jacobi preconditioner     :1257-1294
 *   correct_velocity          :1322-1379
 *   compute_cell_fields       :1381-1439 (+ derivative_x/y_slice :2228-2285)
 *   pressure_residual         :1532-1567
 *   max_divergence            :1569-1582
 *   ensure_finite             :1707-1742
 *   apply_velocity_boundaries :1981-2044
 *   u/v_face_open             :2046-2064
 *   stability numbers / dt    :2294-2336
 *   project_zero_mean         :2338-2377
 * Boundary data follow SolverBoundaryOverrides (src/preprocess.rs:1362-1407):
 * per side a kind (Velocity / Symmetry / PressureOutlet) and constant (u, v).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define UI(i, j) ((i) + (nx + 1) * (j))
#define VI(i, j) ((i) + nx * (j))
#define PI_(i, j) ((i) + nx * (j))
#define SOLID(i, j) (s->solid[(i) + nx * (j)])

static void set_threads(int32_t t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

static inline int u_open(const or2d_state *s, int64_t i, int64_t j) {
    const int64_t nx = s->cfg.nx;
    if (i == 0) return !SOLID(0, j);
    if (i == nx) return !SOLID(nx - 1, j);
    return !SOLID(i - 1, j) && !SOLID(i, j);
}
static inline int v_open(const or2d_state *s, int64_t i, int64_t j) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    if (j == 0) return !SOLID(i, 0);
    if (j == ny) return !SOLID(i, ny - 1);
    return !SOLID(i, j - 1) && !SOLID(i, j);
}
static inline int right_outlet(const or2d_state *s) {
    return s->cfg.bc_kind[OR_SIDE_RIGHT] == OR_BC_OUTLET;
}
/* pressure_outlet_cell, incompressible.rs:2066-2082 */
static inline int outlet_cell(const or2d_state *s, int64_t i, int64_t j) {
    const int64_t nx = s->cfg.nx;
    return i + 1 == nx && !SOLID(i, j) && right_outlet(s);
}
static inline double outlet_pressure(const or2d_state *s) {
    return right_outlet(s) ? s->cfg.outlet_pressure[OR_SIDE_RIGHT] : 0.0;
}

int or2d_init(or2d_state *s, const or2d_config *cfg, const uint8_t *solid) {
    memset(s, 0, sizeof(*s));
    /* SimulationConfig::validate (incompressible.rs:93-191), the parts this path uses */
    if (cfg->nx < 4 || cfg->ny < 4) return OR_ERR_CONFIG;
    if (!(isfinite(cfg->dt) && cfg->dt > 0)) return OR_ERR_CONFIG;
    if (!(isfinite(cfg->t_end) && cfg->t_end >= 0)) return OR_ERR_CONFIG;
    const int simple = cfg->coupling == OR_COUPLING_SIMPLE;
    if (cfg->coupling != OR_COUPLING_PROJECTION && !simple) return OR_ERR_CONFIG;
    if (!simple && cfg->t_end == 0) return OR_ERR_CONFIG;
    if (!(isfinite(cfg->velocity_relaxation) && cfg->velocity_relaxation > 0 &&
          cfg->velocity_relaxation <= 1))
        return OR_ERR_CONFIG;
    if (!(isfinite(cfg->pressure_relaxation) && cfg->pressure_relaxation > 0 &&
          cfg->pressure_relaxation <= 1))
        return OR_ERR_CONFIG;
    if (simple && cfg->adaptive) return OR_ERR_CONFIG;
    if (cfg->pressure_max_iters <= 0) return OR_ERR_CONFIG;
    if (!(cfg->pressure_omega > 0 && cfg->pressure_omega < 2)) return OR_ERR_CONFIG;
    if (!(isfinite(cfg->density) && cfg->density > 0 && isfinite(cfg->nu) && cfg->nu > 0))
        return OR_ERR_CONFIG;
    s->cfg = *cfg;
    const int64_t nx = cfg->nx, ny = cfg->ny;
    s->dx = cfg->length / (double)nx;
    s->dy = cfg->height / (double)ny;
    s->n = nx * ny;
    s->solid = calloc(s->n, 1);
    if (solid) memcpy(s->solid, solid, s->n);
    const int64_t nu = (nx + 1) * ny, nv = nx * (ny + 1);
    s->p = calloc(s->n, 8); s->rhs = calloc(s->n, 8);
    s->u = calloc(nu, 8); s->v = calloc(nv, 8);
    s->us = calloc(nu, 8); s->vs = calloc(nv, 8);
    s->uold = calloc(nu, 8); s->vold = calloc(nv, 8);
    s->ucell = calloc(s->n, 8); s->vcell = calloc(s->n, 8);
    s->speed = calloc(s->n, 8); s->vort = calloc(s->n, 8);
    s->x = calloc(s->n, 8); s->b = calloc(s->n, 8); s->r = calloc(s->n, 8);
    s->z = calloc(s->n, 8); s->d = calloc(s->n, 8); s->q = calloc(s->n, 8);
    s->pold = calloc(s->n, 8);
    s->current_dt = cfg->dt;
    /* initialize_velocity (incompressible.rs:588-620) with the cavity's zero
     * initial field: blocked faces 0, then the boundary velocities at t = 0. */
    or2d_apply_velocity_boundaries(s, s->u, s->v);
    or2d_cell_fields(s);
    return OR_OK;
}

void or2d_free(or2d_state *s) {
    void *f[] = {s->solid, s->p, s->rhs, s->u, s->v, s->us, s->vs, s->uold, s->vold, s->ucell,
                 s->vcell, s->speed, s->vort, s->x, s->b, s->r, s->z, s->d, s->q, s->pold};
    for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); ++i) free(f[i]);
    memset(s, 0, sizeof(*s));
}

/* apply_velocity_boundaries, incompressible.rs:1981-2044 */
void or2d_apply_velocity_boundaries(or2d_state *s, double *u, double *v) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    const int32_t *kind = s->cfg.bc_kind;
    for (int64_t j = 0; j < ny; ++j) {
        /* Case::boundary_velocity (src/cases/mod.rs:141-170): a Channel or
         * BackwardFacingStep inlet is parabolic in y = u_face_y(j) */
        double ul = s->cfg.bc_u[OR_SIDE_LEFT];
        if (s->cfg.inlet_profile == 1) {
            const double y = ((double)j + 0.5) * s->dy, sh = s->cfg.inlet_step_height;
            double eta = (y - sh) / (s->cfg.height - sh);
            eta = eta < 0.0 ? 0.0 : (eta > 1.0 ? 1.0 : eta);
            ul = y <= sh ? 0.0 : 6.0 * s->cfg.bc_u[OR_SIDE_LEFT] * eta * (1.0 - eta);
        }
        switch (kind[OR_SIDE_LEFT]) {
            case OR_BC_VELOCITY: u[UI(0, j)] = ul; break;
            case OR_BC_SYMMETRY: u[UI(0, j)] = 0.0; break;
            default: u[UI(0, j)] = u[UI(1, j)]; break;
        }
        switch (kind[OR_SIDE_RIGHT]) {
            case OR_BC_VELOCITY: u[UI(nx, j)] = s->cfg.bc_u[OR_SIDE_RIGHT]; break;
            case OR_BC_SYMMETRY: u[UI(nx, j)] = 0.0; break;
            default: u[UI(nx, j)] = u[UI(nx - 1, j)]; break;
        }
    }
    for (int64_t i = 0; i < nx; ++i) {
        switch (kind[OR_SIDE_BOTTOM]) {
            case OR_BC_VELOCITY: v[VI(i, 0)] = s->cfg.bc_v[OR_SIDE_BOTTOM]; break;
            case OR_BC_SYMMETRY: v[VI(i, 0)] = 0.0; break;
            default: v[VI(i, 0)] = v[VI(i, 1)]; break;
        }
        switch (kind[OR_SIDE_TOP]) {
            case OR_BC_VELOCITY: v[VI(i, ny)] = s->cfg.bc_v[OR_SIDE_TOP]; break;
            case OR_BC_SYMMETRY: v[VI(i, ny)] = 0.0; break;
            default: v[VI(i, ny)] = v[VI(i, ny - 1)]; break;
        }
    }
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 1; i < nx; ++i)
            if (!u_open(s, i, j)) u[UI(i, j)] = 0.0;
    for (int64_t j = 1; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i)
            if (!v_open(s, i, j)) v[VI(i, j)] = 0.0;
}

/* MomentumStencil::u_neighbor (incompressible.rs:2096-2126) with
 * tangential_ghost_u (:2159-2178). */
static inline double u_nb(const or2d_state *s, const double *u, int64_t i, int64_t j, int di,
                          int dj, double c) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    int64_t a = i + di, b = j + dj;
    if (a >= 0 && a <= nx && b >= 0 && b < ny) return u_open(s, a, b) ? u[UI(a, b)] : -c;
    int side = -1;
    if (b < 0) side = OR_SIDE_BOTTOM;
    else if (b >= ny) side = OR_SIDE_TOP;
    if (side < 0) return c;
    return s->cfg.bc_kind[side] == OR_BC_VELOCITY ? 2.0 * s->cfg.bc_u[side] - c : c;
}
/* MomentumStencil::v_neighbor (incompressible.rs:2128-2157) with
 * tangential_ghost_v (:2180-2199). */
static inline double v_nb(const or2d_state *s, const double *v, int64_t i, int64_t j, int di,
                          int dj, double c) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    int64_t a = i + di, b = j + dj;
    if (a >= 0 && a < nx && b >= 0 && b <= ny) return v_open(s, a, b) ? v[VI(a, b)] : -c;
    int side = -1;
    if (a < 0) side = OR_SIDE_LEFT;
    else if (a >= nx) side = OR_SIDE_RIGHT;
    if (side < 0) return c;
    return s->cfg.bc_kind[side] == OR_BC_VELOCITY ? 2.0 * s->cfg.bc_v[side] - c : c;
}
/* transported, incompressible.rs:2201-2212 */
static inline double transported(int scheme, double fv, double left, double right) {
    if (scheme == OR_UPWIND) return fv >= 0.0 ? left : right;
    return 0.5 * (left + right);
}

/* predict_momentum, incompressible.rs:748-922 (projection: no pressure
 * gradient, velocity relaxation 1, no buoyancy). */
/* predict_momentum, incompressible.rs:748-922; SIMPLE adds the relaxed
 * pressure gradient of the current pressure (:822-826, :848-853). */
void or2d_predict(or2d_state *s, double dt) {
    set_threads(s->cfg.threads);
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    const double dx = s->dx, dy = s->dy, nu = s->cfg.nu;
    const int sc = s->cfg.convection;
    const double *u = s->u, *v = s->v, *p = s->p;
    const int simple = s->cfg.coupling == OR_COUPLING_SIMPLE;
    const double rho = s->cfg.density, au = s->cfg.velocity_relaxation;
    memset(s->us, 0, sizeof(double) * (nx + 1) * ny);
    memset(s->vs, 0, sizeof(double) * nx * (ny + 1));
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 1; i < nx; ++i) {
            if (!u_open(s, i, j)) continue;
            double up = u[UI(i, j)];
            double ue_nb = u_nb(s, u, i, j, 1, 0, up), uw_nb = u_nb(s, u, i, j, -1, 0, up);
            double un_nb = u_nb(s, u, i, j, 0, 1, up), us_nb = u_nb(s, u, i, j, 0, -1, up);
            double ue = 0.5 * (up + ue_nb), uw = 0.5 * (uw_nb + up);
            double vn = 0.5 * (v[VI(i - 1, j + 1)] + v[VI(i, j + 1)]);
            double vs = 0.5 * (v[VI(i - 1, j)] + v[VI(i, j)]);
            double pe = transported(sc, ue, up, ue_nb), pw = transported(sc, uw, uw_nb, up);
            double pn = transported(sc, vn, up, un_nb), ps = transported(sc, vs, us_nb, up);
            double conv = (ue * pe - uw * pw) / dx + (vn * pn - vs * ps) / dy;
            double diff = (ue_nb - 2.0 * up + uw_nb) / (dx * dx) + (un_nb - 2.0 * up + us_nb) / (dy * dy);
            s->us[UI(i, j)] =
                simple ? up + au * dt * (-conv + nu * diff - (p[PI_(i, j)] - p[PI_(i - 1, j)]) / dx / rho)
                       : up + dt * (-conv + nu * diff);
        }
#pragma omp parallel for schedule(static)
    for (int64_t j = 1; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            if (!v_open(s, i, j)) continue;
            double vp = v[VI(i, j)];
            double ve_nb = v_nb(s, v, i, j, 1, 0, vp), vw_nb = v_nb(s, v, i, j, -1, 0, vp);
            double vn_nb = v_nb(s, v, i, j, 0, 1, vp), vs_nb = v_nb(s, v, i, j, 0, -1, vp);
            double ue = 0.5 * (u[UI(i + 1, j - 1)] + u[UI(i + 1, j)]);
            double uw = 0.5 * (u[UI(i, j - 1)] + u[UI(i, j)]);
            double vn = 0.5 * (vp + vn_nb), vs = 0.5 * (vs_nb + vp);
            double pe = transported(sc, ue, vp, ve_nb), pw = transported(sc, uw, vw_nb, vp);
            double pn = transported(sc, vn, vp, vn_nb), ps = transported(sc, vs, vs_nb, vp);
            double conv = (ue * pe - uw * pw) / dx + (vn * pn - vs * ps) / dy;
            double diff = (ve_nb - 2.0 * vp + vw_nb) / (dx * dx) + (vn_nb - 2.0 * vp + vs_nb) / (dy * dy);
            s->vs[VI(i, j)] =
                simple ? vp + au * dt * (-conv + nu * diff - (p[PI_(i, j)] - p[PI_(i, j - 1)]) / dy / rho)
                       : vp + dt * (-conv + nu * diff);
        }
}

/* build_pressure_rhs, incompressible.rs:924-970 */
void or2d_build_rhs(or2d_state *s, double dt) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    const double scale = s->cfg.coupling == OR_COUPLING_SIMPLE ? s->cfg.velocity_relaxation : 1.0;
    const double rod = s->cfg.density / (scale * dt);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            if (SOLID(i, j) || outlet_cell(s, i, j)) {
                s->rhs[PI_(i, j)] = 0.0;
                continue;
            }
            double div = (s->us[UI(i + 1, j)] - s->us[UI(i, j)]) / s->dx +
                         (s->vs[VI(i, j + 1)] - s->vs[VI(i, j)]) / s->dy;
            s->rhs[PI_(i, j)] = rod * div;
        }
}

/* apply_pressure_operator, incompressible.rs:1216-1255 */
void or2d_apply_operator(const or2d_state *s, const double *x, double *y) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    const double idx2 = 1.0 / (s->dx * s->dx), idy2 = 1.0 / (s->dy * s->dy);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t k = PI_(i, j);
            if (SOLID(i, j) || outlet_cell(s, i, j)) {
                y[k] = x[k];
                continue;
            }
            double diag = 0.0, sum = 0.0;
            if (i + 1 < nx && !SOLID(i + 1, j)) { diag += idx2; sum += idx2 * x[k + 1]; }
            if (i > 0 && !SOLID(i - 1, j)) { diag += idx2; sum += idx2 * x[k - 1]; }
            if (j + 1 < ny && !SOLID(i, j + 1)) { diag += idy2; sum += idy2 * x[k + nx]; }
            if (j > 0 && !SOLID(i, j - 1)) { diag += idy2; sum += idy2 * x[k - nx]; }
            y[k] = diag * x[k] - sum;
        }
}

/* apply_jacobi_preconditioner, incompressible.rs:1257-1294 */
static void jacobi2d(const or2d_state *s, const double *r, double *z) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    const double idx2 = 1.0 / (s->dx * s->dx), idy2 = 1.0 / (s->dy * s->dy);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t k = PI_(i, j);
            if (SOLID(i, j) || outlet_cell(s, i, j)) {
                z[k] = r[k];
                continue;
            }
            double diag = 0.0;
            if (i + 1 < nx && !SOLID(i + 1, j)) diag += idx2;
            if (i > 0 && !SOLID(i - 1, j)) diag += idx2;
            if (j + 1 < ny && !SOLID(i, j + 1)) diag += idy2;
            if (j > 0 && !SOLID(i, j - 1)) diag += idy2;
            z[k] = diag > 0.0 ? r[k] / diag : r[k];
        }
}

/* project_zero_mean, incompressible.rs:2338-2377: row sums reduced, then the
 * mean over non-solid, non-Dirichlet cells is subtracted. */
static void zero_mean(const or2d_state *s, double *x) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    double sum = 0.0;
    int64_t count = 0;
    for (int64_t j = 0; j < ny; ++j) {
        double rs = 0.0;
        int64_t rc = 0;
        for (int64_t i = 0; i < nx; ++i)
            if (!SOLID(i, j) && !outlet_cell(s, i, j)) {
                rs += x[PI_(i, j)];
                rc++;
            }
        sum += rs;
        count += rc;
    }
    if (count == 0) return;
    double mean = sum / (double)count;
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i)
            if (!SOLID(i, j) && !outlet_cell(s, i, j)) x[PI_(i, j)] -= mean;
}

static double dot2(const double *a, const double *b, int64_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static double maxabs2(const double *a, int64_t n) {
    double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (int64_t i = 0; i < n; ++i) m = fmax(m, fabs(a[i]));
    return m;
}

#define FAIL(msg)                                                   \
    do {                                                            \
        rep->status = OR_ERR_SOLVER;                                \
        snprintf(rep->message, sizeof rep->message, "%s", msg);     \
        return rep->status;                                         \
    } while (0)

/* solve_pressure_pcg, incompressible.rs:1048-1214.  The stop rule is the
 * reference's max|r| < tol or, for OR_STOP_REL_L2, ||r||_2 <= tol*||r0||_2. */
static int pcg2d(or2d_state *s, or_solve_report *rep) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, n = s->n;
    double *x = s->x, *b = s->b, *r = s->r, *z = s->z, *d = s->d, *q = s->q;
    const int zm = !right_outlet(s);
    const int rel = s->cfg.stop_rule == OR_STOP_REL_L2;
    const double tol = s->cfg.pressure_tolerance, pout = outlet_pressure(s);
    if (s->cfg.initial_guess == OR_GUESS_ZERO) memset(x, 0, 8 * n);
    else memcpy(x, s->p, 8 * n);
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t k = PI_(i, j);
            if (SOLID(i, j) || outlet_cell(s, i, j)) {
                double val = outlet_cell(s, i, j) ? pout : 0.0;
                x[k] = val;
                b[k] = val;
            } else {
                b[k] = -s->rhs[k];
            }
        }
    if (zm) { zero_mean(s, x); zero_mean(s, b); }
    or2d_apply_operator(s, x, q);
    for (int64_t k = 0; k < n; ++k) r[k] = b[k] - q[k];
    if (zm) zero_mean(s, r);
    double residual = maxabs2(r, n);
    double rr = dot2(r, r, n);
    rep->initial_l2 = sqrt(rr);
    const double thresh = tol * rep->initial_l2;
    if (rel ? (sqrt(rr) <= thresh) : (residual < tol)) {
        memcpy(s->p, x, 8 * n);
        rep->iterations = 0;
        rep->residual_abs_inf = residual;
        rep->residual_l2 = sqrt(rr);
        rep->converged = 1;
        return OR_OK;
    }
    jacobi2d(s, r, z);
    if (zm) zero_mean(s, z);
    memcpy(d, z, 8 * n);
    double rz_old = dot2(r, z, n);
    int64_t it = 0;
    for (int64_t iter = 0; iter < s->cfg.pressure_max_iters; ++iter) {
        it = iter + 1;
        or2d_apply_operator(s, d, q);
        double den = dot2(d, q, n);
        if (fabs(den) < 1.0e-30 || !isfinite(den))
            FAIL("PCG pressure solver encountered a singular search direction");
        double alpha = rz_old / den;
        if (!isfinite(alpha)) FAIL("PCG pressure solver produced a non-finite step length");
        for (int64_t k = 0; k < n; ++k) {
            x[k] += alpha * d[k];
            r[k] -= alpha * q[k];
        }
        if (zm) { zero_mean(s, x); zero_mean(s, r); }
        residual = maxabs2(r, n);
        rr = dot2(r, r, n);
        if (!isfinite(residual)) FAIL("PCG pressure solver produced a non-finite residual");
        if (rel ? (sqrt(rr) <= thresh) : (residual < tol)) {
            rep->converged = 1;
            break;
        }
        jacobi2d(s, r, z);
        if (zm) zero_mean(s, z);
        double rz_new = dot2(r, z, n);
        if (fabs(rz_old) < 1.0e-30) break;
        double beta = rz_new / rz_old;
        for (int64_t k = 0; k < n; ++k) d[k] = z[k] + beta * d[k];
        if (zm) zero_mean(s, d);
        rz_old = rz_new;
    }
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t k = PI_(i, j);
            s->p[k] = SOLID(i, j) ? 0.0 : (outlet_cell(s, i, j) ? pout : x[k]);
        }
    rep->iterations = it;
    rep->residual_abs_inf = residual;
    rep->residual_l2 = sqrt(rr);
    return OR_OK;
}

/* pressure_residual, incompressible.rs:1532-1567 */
double or2d_pressure_residual(const or2d_state *s) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    const double idx2 = 1.0 / (s->dx * s->dx), idy2 = 1.0 / (s->dy * s->dy);
    double m = 0.0;
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            if (SOLID(i, j) || outlet_cell(s, i, j)) continue;
            int64_t k = PI_(i, j);
            double p0 = s->p[k];
            double pe = (i + 1 < nx && !SOLID(i + 1, j)) ? s->p[k + 1] : p0;
            double pw = (i > 0 && !SOLID(i - 1, j)) ? s->p[k - 1] : p0;
            double pn = (j + 1 < ny && !SOLID(i, j + 1)) ? s->p[k + nx] : p0;
            double ps = (j > 0 && !SOLID(i, j - 1)) ? s->p[k - nx] : p0;
            double lap = (pe - 2.0 * p0 + pw) * idx2 + (pn - 2.0 * p0 + ps) * idy2;
            m = fmax(m, fabs(lap - s->rhs[k]));
        }
    return m;
}

/* solve_pressure_sor, incompressible.rs:979-1046 */
static int sor2d(or2d_state *s, or_solve_report *rep) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    const double idx2 = 1.0 / (s->dx * s->dx), idy2 = 1.0 / (s->dy * s->dy);
    const double omega = s->cfg.pressure_omega;
    double residual = INFINITY, *p = s->p;
    int64_t it = 0;
    for (int64_t iter = 0; iter < s->cfg.pressure_max_iters; ++iter) {
        it = iter + 1;
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int64_t k = PI_(i, j);
                if (SOLID(i, j)) { p[k] = 0.0; continue; }
                if (outlet_cell(s, i, j)) { p[k] = outlet_pressure(s); continue; }
                double ap = 0.0, sum = 0.0;
                if (i + 1 < nx && !SOLID(i + 1, j)) { ap += idx2; sum += idx2 * p[k + 1]; }
                if (i > 0 && !SOLID(i - 1, j)) { ap += idx2; sum += idx2 * p[k - 1]; }
                if (j + 1 < ny && !SOLID(i, j + 1)) { ap += idy2; sum += idy2 * p[k + nx]; }
                if (j > 0 && !SOLID(i, j - 1)) { ap += idy2; sum += idy2 * p[k - nx]; }
                if (ap > 0.0) {
                    double cand = (sum - s->rhs[k]) / ap;
                    double old = p[k];
                    p[k] = old + omega * (cand - old);
                }
            }
        if (!right_outlet(s)) { /* find_reference_cell :1614-1623 */
            int64_t ref = -1;
            for (int64_t k = 0; k < s->n && ref < 0; ++k)
                if (!s->solid[k]) ref = k;
            if (ref >= 0) {
                double pr = p[ref];
                for (int64_t k = 0; k < s->n; ++k)
                    if (!s->solid[k]) p[k] -= pr;
            }
        }
        residual = or2d_pressure_residual(s);
        if (!isfinite(residual)) FAIL("Pressure solver produced a non-finite residual");
        if (residual < s->cfg.pressure_tolerance) { rep->converged = 1; break; }
    }
    rep->iterations = it;
    rep->residual_abs_inf = residual;
    return OR_OK;
}

int or2d_solve_pressure(or2d_state *s, or_solve_report *rep) {
    memset(rep, 0, sizeof(*rep));
    set_threads(s->cfg.threads);
    return s->cfg.solver == OR_SOLVER_SOR ? sor2d(s, rep) : pcg2d(s, rep);
}

/* correct_velocity, incompressible.rs:1322-1379 (projection: scale 1) */
void or2d_correct(or2d_state *s, double dt) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    const double scale = s->cfg.coupling == OR_COUPLING_SIMPLE ? s->cfg.velocity_relaxation : 1.0;
    const double dor = scale * dt / s->cfg.density;
    memcpy(s->u, s->us, 8 * (nx + 1) * ny);
    memcpy(s->v, s->vs, 8 * nx * (ny + 1));
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 1; i < nx; ++i)
            s->u[UI(i, j)] = u_open(s, i, j)
                                 ? s->us[UI(i, j)] - dor * ((s->p[PI_(i, j)] - s->p[PI_(i - 1, j)]) / s->dx)
                                 : 0.0;
#pragma omp parallel for schedule(static)
    for (int64_t j = 1; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i)
            s->v[VI(i, j)] = v_open(s, i, j)
                                 ? s->vs[VI(i, j)] - dor * ((s->p[PI_(i, j)] - s->p[PI_(i, j - 1)]) / s->dy)
                                 : 0.0;
}

/* derivative_x_slice / derivative_y_slice, incompressible.rs:2228-2285 */
static double ddx(const or2d_state *s, const double *f, int64_t i, int64_t j) {
    const int64_t nx = s->cfg.nx;
    double c = f[PI_(i, j)];
    int eo = i + 1 < nx && !SOLID(i + 1, j), wo = i > 0 && !SOLID(i - 1, j);
    double e = eo ? f[PI_(i + 1, j)] : c, w = wo ? f[PI_(i - 1, j)] : c;
    if (eo && wo) return (e - w) / (2.0 * s->dx);
    if (eo) return (e - c) / s->dx;
    if (wo) return (c - w) / s->dx;
    return 0.0;
}
static double ddy(const or2d_state *s, const double *f, int64_t i, int64_t j) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    double c = f[PI_(i, j)];
    int no = j + 1 < ny && !SOLID(i, j + 1), so = j > 0 && !SOLID(i, j - 1);
    double n = no ? f[PI_(i, j + 1)] : c, sv = so ? f[PI_(i, j - 1)] : c;
    if (no && so) return (n - sv) / (2.0 * s->dy);
    if (no) return (n - c) / s->dy;
    if (so) return (c - sv) / s->dy;
    return 0.0;
}

/* compute_cell_fields, incompressible.rs:1381-1439 */
void or2d_cell_fields(or2d_state *s) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            int64_t k = PI_(i, j);
            if (SOLID(i, j)) { s->ucell[k] = s->vcell[k] = s->speed[k] = 0.0; continue; }
            double uc = 0.5 * (s->u[UI(i, j)] + s->u[UI(i + 1, j)]);
            double vc = 0.5 * (s->v[VI(i, j)] + s->v[VI(i, j + 1)]);
            s->ucell[k] = uc;
            s->vcell[k] = vc;
            s->speed[k] = sqrt(uc * uc + vc * vc);
        }
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i)
            s->vort[PI_(i, j)] = SOLID(i, j) ? 0.0 : ddx(s, s->vcell, i, j) - ddy(s, s->ucell, i, j);
}

/* max_divergence, incompressible.rs:1569-1582 */
double or2d_max_divergence(const or2d_state *s) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    double m = 0.0;
    for (int64_t j = 0; j < ny; ++j)
        for (int64_t i = 0; i < nx; ++i) {
            if (SOLID(i, j)) continue;
            double div = (s->u[UI(i + 1, j)] - s->u[UI(i, j)]) / s->dx +
                         (s->v[VI(i, j + 1)] - s->v[VI(i, j)]) / s->dy;
            m = fmax(m, fabs(div));
        }
    return m;
}

/* momentum_stability_numbers, incompressible.rs:2294-2306 */
void or2d_stability_numbers(double dt, double nu, double dx, double dy, double umax,
                            double vmax, double *cfl, double *dnu) {
    *cfl = dt * (umax / dx + vmax / dy);
    *dnu = nu * dt * (1.0 / (dx * dx) + 1.0 / (dy * dy));
}

/* adaptive_time_step, incompressible.rs:2308-2336 (thermal_diffusivity < 0 = None) */
int or2d_adaptive_dt(double current_dt, double umax, double vmax, double dx, double dy,
                     double nu, double thermal_diffusivity, double min_dt, double max_dt,
                     double target_cfl, double target_diffusion, double max_growth,
                     double *out_dt, char *message, int cap) {
    double rate = umax / dx + vmax / dy;
    double cfl_dt = rate > 2.220446049250313e-16 ? target_cfl / rate : max_dt;
    double inv = 1.0 / (dx * dx) + 1.0 / (dy * dy);
    double eff = thermal_diffusivity >= 0.0 ? fmax(thermal_diffusivity, nu) : nu;
    double diff_dt = target_diffusion / (eff * inv);
    double stable = fmin(fmin(max_dt, cfl_dt), diff_dt);
    if (stable < min_dt) {
        if (message)
            snprintf(message, cap,
                     "required stable time step %.3e is below configured min_dt %.3e; CFL "
                     "constraint=%.3e, diffusion constraint=%.3e",
                     stable, min_dt, cfl_dt, diff_dt);
        return OR_ERR_SOLVER;
    }
    double g = fmin(stable, current_dt * max_growth);
    *out_dt = g < min_dt ? min_dt : (g > max_dt ? max_dt : g);
    return OR_OK;
}

static double maxabs_n(const double *a, int64_t n) {
    double m = 0.0;
    for (int64_t i = 0; i < n; ++i) m = fmax(m, fabs(a[i]));
    return m;
}
static double maxdiff(const double *a, const double *b, int64_t n) {
    double m = 0.0;
    for (int64_t i = 0; i < n; ++i) m = fmax(m, fabs(a[i] - b[i]));
    return m;
}

/* reached_end_time, incompressible.rs:557-561: SIMPLE never ends on time. */
int or2d_reached_end(const or2d_state *s) {
    const double tend = s->cfg.t_end;
    const double ttol = 2.220446049250313e-16 * fmax(fabs(tend), 1.0) * 16.0;
    return s->cfg.coupling == OR_COUPLING_PROJECTION && s->time >= tend - ttol;
}

/* advance_one_step, incompressible.rs:622-746, both couplings. */
int or2d_step(or2d_state *s, or2d_diag *diag, or_solve_report *rep) {
    memset(rep, 0, sizeof(*rep));
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny;
    const int64_t nu_ = (nx + 1) * ny, nv_ = nx * (ny + 1);
    const double tend = s->cfg.t_end;
    const int simple = s->cfg.coupling == OR_COUPLING_SIMPLE;
    if (or2d_reached_end(s)) {
        rep->status = OR_ERR_SOLVER;
        snprintf(rep->message, sizeof rep->message, "projection simulation has already reached t_end");
        return rep->status;
    }
    double dt = s->cfg.dt;
    if (s->cfg.adaptive) {
        int st = or2d_adaptive_dt(s->current_dt, maxabs_n(s->u, nu_), maxabs_n(s->v, nv_), s->dx,
                                  s->dy, s->cfg.nu, -1.0, s->cfg.min_dt, s->cfg.max_dt,
                                  s->cfg.target_cfl, s->cfg.target_diffusion, s->cfg.max_growth,
                                  &dt, rep->message, sizeof rep->message);
        if (st) { rep->status = st; return st; }
    }
    if (!simple) {
        double remaining = tend - s->time;
        dt = fmin(dt, remaining);
        if (dt <= 0.0) {
            rep->status = OR_ERR_SOLVER;
            snprintf(rep->message, sizeof rep->message, "projection simulation has already reached t_end");
            return rep->status;
        }
    }
    if (s->cfg.adaptive) s->current_dt = dt;
    memcpy(s->uold, s->u, 8 * nu_);
    memcpy(s->vold, s->v, 8 * nv_);
    or2d_predict(s, dt);
    or2d_apply_velocity_boundaries(s, s->us, s->vs);
    if (simple) {
        /* the solve is for p', not the absolute pressure (:669-673) */
        memcpy(s->pold, s->p, 8 * s->n);
        memset(s->p, 0, 8 * s->n);
    }
    or2d_build_rhs(s, dt);
    int st = or2d_solve_pressure(s, rep);
    if (st) return st;
    or2d_correct(s, dt);
    if (simple) {
        /* p = p_old + alpha_p p' (:683-694) */
        const double ap = s->cfg.pressure_relaxation;
        for (int64_t k = 0; k < s->n; ++k) s->p[k] = s->pold[k] + ap * s->p[k];
    }
    or2d_apply_velocity_boundaries(s, s->u, s->v);
    s->step += 1;
    if (simple)
        s->time = (double)s->step * s->cfg.dt;
    else
        s->time += dt;
    or2d_cell_fields(s);
    /* ensure_finite, incompressible.rs:1707-1742 */
    struct { const char *name; const double *f; int64_t n; } fs[] = {
        {"pressure", s->p, s->n}, {"u", s->u, nu_}, {"v", s->v, nv_}, {"speed", s->speed, s->n}};
    for (int a = 0; a < 4; ++a)
        for (int64_t i = 0; i < fs[a].n; ++i)
            if (!isfinite(fs[a].f[i])) {
                rep->status = OR_ERR_SOLVER;
                snprintf(rep->message, sizeof rep->message,
                         "Non-finite %s value at flat index %lld: %g (step %lld)", fs[a].name,
                         (long long)i, fs[a].f[i], (long long)s->step);
                return rep->status;
            }
    if (diag) {
        double umax = maxabs_n(s->u, nu_), vmax = maxabs_n(s->v, nv_);
        diag->time_step = dt;
        diag->pressure_residual = rep->residual_abs_inf;
        diag->pressure_iterations = rep->iterations;
        or2d_stability_numbers(dt, s->cfg.nu, s->dx, s->dy, umax, vmax, &diag->momentum_cfl,
                               &diag->viscous_diffusion_number);
        diag->max_divergence = or2d_max_divergence(s);
        diag->max_speed = maxabs_n(s->speed, s->n);
        diag->velocity_change = fmax(maxdiff(s->u, s->uold, nu_), maxdiff(s->v, s->vold, nv_));
    }
    return OR_OK;
}
