/*
 * Oracle: 3D lid-driven cavity projection step (FP64).
 * TEST INFRASTRUCTURE ONLY -- the checker, never the shipped path.
 *
 * Restates src/solver/incompressible_3d.rs:
 *   predictor            :174-273  (+ u/v/w_neighbor :366-447)
 *   pressure RHS         :275-291
 *   SOR Poisson          :293-332  (+ pressure_neighbor :450-474, residual :476-502)
 *   velocity correction  :334-363
 *   diagnostics          :504-554
 * and adds the pinned Jacobi-PCG (not in the reference): the A7 loop of
 * src/solver/incompressible.rs:1048-1214 applied to the B4 operator with cell
 * (0,0,0) eliminated as src/poisson.rs:303-337 eliminates a reference cell.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define U3(i, j, k) ((i) + (nx + 1) * ((j) + ny * (k)))
#define V3(i, j, k) ((i) + nx * ((j) + (ny + 1) * (k)))
#define W3(i, j, k) ((i) + nx * ((j) + ny * (k)))
#define P3(i, j, k) ((i) + nx * ((j) + ny * (k)))

static void set_threads(int32_t t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

int or3d_init(or3d_state *s, const or3d_config *cfg) {
    memset(s, 0, sizeof(*s));
    /* LidDrivenCavity3DConfig::validate, incompressible_3d.rs:49-84, with
     * UniformGrid3D::new's minimum of 4 cells per axis (src/grid.rs:74-84) */
    if (cfg->nx < 4 || cfg->ny < 4 || cfg->nz < 4) return OR_ERR_CONFIG;
    if (!(cfg->length > 0 && cfg->height > 0 && cfg->depth > 0)) return OR_ERR_CONFIG;
    if (!(isfinite(cfg->density) && cfg->density > 0)) return OR_ERR_CONFIG;
    if (!(isfinite(cfg->lid_velocity) && cfg->lid_velocity > 0)) return OR_ERR_CONFIG;
    if (!(isfinite(cfg->reynolds) && cfg->reynolds > 0)) return OR_ERR_CONFIG;
    if (!(isfinite(cfg->dt) && cfg->dt > 0)) return OR_ERR_CONFIG;
    if (cfg->pressure_max_iters <= 0) return OR_ERR_CONFIG;
    if (!(isfinite(cfg->pressure_tolerance) && cfg->pressure_tolerance > 0)) return OR_ERR_CONFIG;
    if (!(cfg->pressure_omega > 0 && cfg->pressure_omega < 2)) return OR_ERR_CONFIG;
    s->cfg = *cfg;
    const int64_t nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    s->dx = cfg->length / (double)nx;
    s->dy = cfg->height / (double)ny;
    s->dz = cfg->depth / (double)nz;
    s->n = nx * ny * nz;
    const int64_t nu = (nx + 1) * ny * nz, nv = nx * (ny + 1) * nz, nw = nx * ny * (nz + 1);
    s->p = calloc(s->n, 8);
    s->rhs = calloc(s->n, 8);
    s->u = calloc(nu, 8);
    s->v = calloc(nv, 8);
    s->w = calloc(nw, 8);
    s->us = calloc(nu, 8);
    s->vs = calloc(nv, 8);
    s->ws = calloc(nw, 8);
    s->x = calloc(s->n, 8);
    s->b = calloc(s->n, 8);
    s->r = calloc(s->n, 8);
    s->z = calloc(s->n, 8);
    s->d = calloc(s->n, 8);
    s->q = calloc(s->n, 8);
    if (!s->p || !s->rhs || !s->u || !s->v || !s->w || !s->us || !s->vs || !s->ws || !s->x ||
        !s->b || !s->r || !s->z || !s->d || !s->q) {
        or3d_free(s);
        return OR_ERR_CONFIG;
    }
    return OR_OK;
}

void or3d_free(or3d_state *s) {
    double **f[] = {&s->p, &s->rhs, &s->u, &s->v, &s->w, &s->us, &s->vs,
                    &s->ws, &s->x, &s->b, &s->r, &s->z, &s->d, &s->q};
    for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); ++i) {
        free(*f[i]);
        *f[i] = NULL;
    }
}

/* u/v/w_neighbor, incompressible_3d.rs:366-447: interior value, else the lid
 * ghost 2U - c (u only, j >= ny) or the no-slip ghost -c. */
static inline double u_nb(const or3d_state *s, int64_t i, int64_t j, int64_t k, int di, int dj,
                          int dk, double c) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    int64_t a = i + di, b = j + dj, e = k + dk;
    if (a >= 0 && a <= nx && b >= 0 && b < ny && e >= 0 && e < nz) return s->u[U3(a, b, e)];
    if (b >= ny) return 2.0 * s->cfg.lid_velocity - c;
    return -c;
}
static inline double v_nb(const or3d_state *s, int64_t i, int64_t j, int64_t k, int di, int dj,
                          int dk, double c) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    int64_t a = i + di, b = j + dj, e = k + dk;
    if (a >= 0 && a < nx && b >= 0 && b <= ny && e >= 0 && e < nz) return s->v[V3(a, b, e)];
    return -c;
}
static inline double w_nb(const or3d_state *s, int64_t i, int64_t j, int64_t k, int di, int dj,
                          int dk, double c) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    int64_t a = i + di, b = j + dj, e = k + dk;
    if (a >= 0 && a < nx && b >= 0 && b < ny && e >= 0 && e <= nz) return s->w[W3(a, b, e)];
    return -c;
}

/* predict_momentum, incompressible_3d.rs:174-273 (central advective form). */
void or3d_predict(or3d_state *s) {
    set_threads(s->cfg.threads);
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    const double nu = s->cfg.lid_velocity * s->cfg.length / s->cfg.reynolds; /* :86-88 */
    const double dt = s->cfg.dt, dx = s->dx, dy = s->dy, dz = s->dz;
    const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    memset(s->us, 0, sizeof(double) * (nx + 1) * ny * nz);
    memset(s->vs, 0, sizeof(double) * nx * (ny + 1) * nz);
    memset(s->ws, 0, sizeof(double) * nx * ny * (nz + 1));
    const double *u = s->u, *v = s->v, *w = s->w;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 1; i < nx; ++i) {
                double c = u[U3(i, j, k)];
                double e = u_nb(s, i, j, k, 1, 0, 0, c), wv = u_nb(s, i, j, k, -1, 0, 0, c);
                double n = u_nb(s, i, j, k, 0, 1, 0, c), so = u_nb(s, i, j, k, 0, -1, 0, c);
                double f = u_nb(s, i, j, k, 0, 0, 1, c), bk = u_nb(s, i, j, k, 0, 0, -1, c);
                double vv = 0.25 * (v[V3(i - 1, j, k)] + v[V3(i, j, k)] + v[V3(i - 1, j + 1, k)] +
                                    v[V3(i, j + 1, k)]);
                double ww = 0.25 * (w[W3(i - 1, j, k)] + w[W3(i, j, k)] + w[W3(i - 1, j, k + 1)] +
                                    w[W3(i, j, k + 1)]);
                s->us[U3(i, j, k)] =
                    c + dt * (-c * (e - wv) / (2.0 * dx) - vv * (n - so) / (2.0 * dy) -
                              ww * (f - bk) / (2.0 * dz) +
                              nu * ((e - 2.0 * c + wv) / dx2 + (n - 2.0 * c + so) / dy2 +
                                    (f - 2.0 * c + bk) / dz2));
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 1; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                double c = v[V3(i, j, k)];
                double e = v_nb(s, i, j, k, 1, 0, 0, c), wv = v_nb(s, i, j, k, -1, 0, 0, c);
                double n = v_nb(s, i, j, k, 0, 1, 0, c), so = v_nb(s, i, j, k, 0, -1, 0, c);
                double f = v_nb(s, i, j, k, 0, 0, 1, c), bk = v_nb(s, i, j, k, 0, 0, -1, c);
                double uu = 0.25 * (u[U3(i, j - 1, k)] + u[U3(i + 1, j - 1, k)] + u[U3(i, j, k)] +
                                    u[U3(i + 1, j, k)]);
                double ww = 0.25 * (w[W3(i, j - 1, k)] + w[W3(i, j, k)] + w[W3(i, j - 1, k + 1)] +
                                    w[W3(i, j, k + 1)]);
                s->vs[V3(i, j, k)] =
                    c + dt * (-uu * (e - wv) / (2.0 * dx) - c * (n - so) / (2.0 * dy) -
                              ww * (f - bk) / (2.0 * dz) +
                              nu * ((e - 2.0 * c + wv) / dx2 + (n - 2.0 * c + so) / dy2 +
                                    (f - 2.0 * c + bk) / dz2));
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 1; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                double c = w[W3(i, j, k)];
                double e = w_nb(s, i, j, k, 1, 0, 0, c), wv = w_nb(s, i, j, k, -1, 0, 0, c);
                double n = w_nb(s, i, j, k, 0, 1, 0, c), so = w_nb(s, i, j, k, 0, -1, 0, c);
                double f = w_nb(s, i, j, k, 0, 0, 1, c), bk = w_nb(s, i, j, k, 0, 0, -1, c);
                double uu = 0.25 * (u[U3(i, j, k - 1)] + u[U3(i + 1, j, k - 1)] + u[U3(i, j, k)] +
                                    u[U3(i + 1, j, k)]);
                double vv = 0.25 * (v[V3(i, j, k - 1)] + v[V3(i, j + 1, k - 1)] + v[V3(i, j, k)] +
                                    v[V3(i, j + 1, k)]);
                s->ws[W3(i, j, k)] =
                    c + dt * (-uu * (e - wv) / (2.0 * dx) - vv * (n - so) / (2.0 * dy) -
                              c * (f - bk) / (2.0 * dz) +
                              nu * ((e - 2.0 * c + wv) / dx2 + (n - 2.0 * c + so) / dy2 +
                                    (f - 2.0 * c + bk) / dz2));
            }
}

/* build_pressure_rhs, incompressible_3d.rs:275-291. */
void or3d_build_rhs(or3d_state *s) {
    set_threads(s->cfg.threads);
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    const double scale = s->cfg.density / s->cfg.dt;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i)
                s->rhs[P3(i, j, k)] =
                    scale * (s->us[U3(i + 1, j, k)] - s->us[U3(i, j, k)]) / s->dx +
                    scale * (s->vs[V3(i, j + 1, k)] - s->vs[V3(i, j, k)]) / s->dy +
                    scale * (s->ws[W3(i, j, k + 1)] - s->ws[W3(i, j, k)]) / s->dz;
    s->rhs[0] = 0.0;
}

/* -Laplacian with Neumann ghosts (ghost = centre, pressure_neighbor :450-474),
 * written as diag*x_k - sum_open c*x_nb; row 0 is the identity (pin). */
void or3d_apply_operator(const or3d_state *s, const double *x, double *y) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    const double ix2 = 1.0 / (s->dx * s->dx), iy2 = 1.0 / (s->dy * s->dy),
                 iz2 = 1.0 / (s->dz * s->dz);
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int64_t c = P3(i, j, k);
                if (c == 0) {
                    y[0] = x[0];
                    continue;
                }
                double diag = 0.0, sum = 0.0;
                if (i + 1 < nx) { diag += ix2; sum += ix2 * x[c + 1]; }
                if (i > 0) { diag += ix2; sum += ix2 * x[c - 1]; }
                if (j + 1 < ny) { diag += iy2; sum += iy2 * x[c + nx]; }
                if (j > 0) { diag += iy2; sum += iy2 * x[c - nx]; }
                if (k + 1 < nz) { diag += iz2; sum += iz2 * x[c + nx * ny]; }
                if (k > 0) { diag += iz2; sum += iz2 * x[c - nx * ny]; }
                y[c] = diag * x[c] - sum;
            }
}

static void jacobi3d(const or3d_state *s, const double *r, double *z) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    const double ix2 = 1.0 / (s->dx * s->dx), iy2 = 1.0 / (s->dy * s->dy),
                 iz2 = 1.0 / (s->dz * s->dz);
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int64_t c = P3(i, j, k);
                if (c == 0) {
                    z[0] = r[0];
                    continue;
                }
                double diag = 0.0;
                if (i + 1 < nx) diag += ix2;
                if (i > 0) diag += ix2;
                if (j + 1 < ny) diag += iy2;
                if (j > 0) diag += iy2;
                if (k + 1 < nz) diag += iz2;
                if (k > 0) diag += iz2;
                z[c] = r[c] / diag;
            }
}

static double dot(const double *a, const double *b, int64_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static double max_abs(const double *a, int64_t n) {
    double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (int64_t i = 0; i < n; ++i) m = fmax(m, fabs(a[i]));
    return m;
}

/* pressure_residual, incompressible_3d.rs:476-502: max|Lap p - rhs|, (0,0,0) excluded. */
double or3d_pressure_residual(const or3d_state *s) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    const double ix2 = 1.0 / (s->dx * s->dx), iy2 = 1.0 / (s->dy * s->dy),
                 iz2 = 1.0 / (s->dz * s->dz);
    double m = 0.0;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int64_t c = P3(i, j, k);
                if (c == 0) continue;
                double pc = s->p[c];
                double pe = i + 1 < nx ? s->p[c + 1] : pc, pw = i > 0 ? s->p[c - 1] : pc;
                double pn = j + 1 < ny ? s->p[c + nx] : pc, ps = j > 0 ? s->p[c - nx] : pc;
                double pf = k + 1 < nz ? s->p[c + nx * ny] : pc,
                       pb = k > 0 ? s->p[c - nx * ny] : pc;
                double lap = ix2 * (pe - 2.0 * pc + pw) + iy2 * (pn - 2.0 * pc + ps) +
                             iz2 * (pf - 2.0 * pc + pb);
                /* fmax ignores a NaN operand exactly like Rust's f64::max. */
                m = fmax(m, fabs(lap - s->rhs[c]));
            }
    return m;
}

/* SOR, incompressible_3d.rs:293-332. */
static int sor3d(or3d_state *s, or_solve_report *rep) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    const double ix2 = 1.0 / (s->dx * s->dx), iy2 = 1.0 / (s->dy * s->dy),
                 iz2 = 1.0 / (s->dz * s->dz);
    const double diagonal = 2.0 * (ix2 + iy2 + iz2), omega = s->cfg.pressure_omega;
    double residual = INFINITY;
    int64_t it;
    double *p = s->p;
    for (it = 0; it < s->cfg.pressure_max_iters; ++it) {
        for (int64_t k = 0; k < nz; ++k)
            for (int64_t j = 0; j < ny; ++j)
                for (int64_t i = 0; i < nx; ++i) {
                    int64_t c = P3(i, j, k);
                    if (c == 0) {
                        p[0] = 0.0;
                        continue;
                    }
                    double pc = p[c];
                    double sum = ix2 * ((i + 1 < nx ? p[c + 1] : pc) + (i > 0 ? p[c - 1] : pc)) +
                                 iy2 * ((j + 1 < ny ? p[c + nx] : pc) + (j > 0 ? p[c - nx] : pc)) +
                                 iz2 * ((k + 1 < nz ? p[c + nx * ny] : pc) +
                                        (k > 0 ? p[c - nx * ny] : pc));
                    double cand = (sum - s->rhs[c]) / diagonal;
                    p[c] += omega * (cand - p[c]);
                }
        residual = or3d_pressure_residual(s);
        if (!isfinite(residual)) {
            rep->status = OR_ERR_SOLVER;
            snprintf(rep->message, sizeof rep->message,
                     "3D pressure solver produced a non-finite residual");
            return rep->status;
        }
        if (residual < s->cfg.pressure_tolerance) {
            ++it;
            rep->converged = 1;
            break;
        }
    }
    rep->iterations = it;
    rep->residual_abs_inf = residual;
    return OR_OK;
}

/* Pinned Jacobi-PCG.  Loop structure and error texts follow
 * src/solver/incompressible.rs:1048-1214 (A7); the stop rule is either the
 * reference's absolute max|r| < tol or ||r||_2 <= tol*||r0||_2 (the form of
 * src/linear/solver.rs:84-85 with atol = 0). */
static int pcg3d(or3d_state *s, or_solve_report *rep) {
    const int64_t n = s->n;
    double *x = s->x, *b = s->b, *r = s->r, *z = s->z, *d = s->d, *q = s->q;
    const double tol = s->cfg.pressure_tolerance;
    const int rel = s->cfg.stop_rule == OR_STOP_REL_L2;
    if (s->cfg.initial_guess == OR_GUESS_ZERO)
        memset(x, 0, sizeof(double) * n);
    else
        memcpy(x, s->p, sizeof(double) * n);
    x[0] = 0.0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) b[i] = -s->rhs[i];
    b[0] = 0.0;
    or3d_apply_operator(s, x, q);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
    double rr = dot(r, r, n);
    double residual = max_abs(r, n);
    rep->initial_l2 = sqrt(rr);
    const double thresh = tol * rep->initial_l2;
    int done = rel ? (sqrt(rr) <= thresh) : (residual < tol);
    if (done) {
        memcpy(s->p, x, sizeof(double) * n);
        rep->iterations = 0;
        rep->residual_abs_inf = residual;
        rep->residual_l2 = sqrt(rr);
        rep->converged = 1;
        return OR_OK;
    }
    jacobi3d(s, r, z);
    memcpy(d, z, sizeof(double) * n);
    double rz_old = dot(r, z, n);
    int64_t it = 0;
    for (int64_t iter = 0; iter < s->cfg.pressure_max_iters; ++iter) {
        it = iter + 1;
        or3d_apply_operator(s, d, q);
        double den = dot(d, q, n);
        if (fabs(den) < 1.0e-30 || !isfinite(den)) {
            rep->status = OR_ERR_SOLVER;
            snprintf(rep->message, sizeof rep->message,
                     "PCG pressure solver encountered a singular search direction");
            return rep->status;
        }
        double alpha = rz_old / den;
        if (!isfinite(alpha)) {
            rep->status = OR_ERR_SOLVER;
            snprintf(rep->message, sizeof rep->message,
                     "PCG pressure solver produced a non-finite step length");
            return rep->status;
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * d[i];
            r[i] -= alpha * q[i];
        }
        residual = max_abs(r, n);
        rr = dot(r, r, n);
        if (!isfinite(residual)) {
            rep->status = OR_ERR_SOLVER;
            snprintf(rep->message, sizeof rep->message,
                     "PCG pressure solver produced a non-finite residual");
            return rep->status;
        }
        if (rel ? (sqrt(rr) <= thresh) : (residual < tol)) {
            rep->converged = 1;
            break;
        }
        jacobi3d(s, r, z);
        double rz_new = dot(r, z, n);
        if (fabs(rz_old) < 1.0e-30) break;
        double beta = rz_new / rz_old;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) d[i] = z[i] + beta * d[i];
        rz_old = rz_new;
    }
    memcpy(s->p, x, sizeof(double) * n);
    s->p[0] = 0.0;
    rep->iterations = it;
    rep->residual_abs_inf = residual;
    rep->residual_l2 = sqrt(rr);
    return OR_OK;
}

int or3d_solve_pressure(or3d_state *s, or_solve_report *rep) {
    memset(rep, 0, sizeof(*rep));
    set_threads(s->cfg.threads);
    if (s->cfg.solver == OR_SOLVER_SOR) return sor3d(s, rep);
    return pcg3d(s, rep);
}

/* correct_velocity, incompressible_3d.rs:334-363. */
void or3d_correct(or3d_state *s) {
    set_threads(s->cfg.threads);
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    const double scale = s->cfg.dt / s->cfg.density;
    const double *p = s->p;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 1; i < nx; ++i)
                s->u[U3(i, j, k)] =
                    s->us[U3(i, j, k)] - scale * (p[P3(i, j, k)] - p[P3(i - 1, j, k)]) / s->dx;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 1; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i)
                s->v[V3(i, j, k)] =
                    s->vs[V3(i, j, k)] - scale * (p[P3(i, j, k)] - p[P3(i, j - 1, k)]) / s->dy;
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t k = 1; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i)
                s->w[W3(i, j, k)] =
                    s->ws[W3(i, j, k)] - scale * (p[P3(i, j, k)] - p[P3(i, j, k - 1)]) / s->dz;
}

/* ensure_finite, incompressible_3d.rs:534-554. */
int or3d_ensure_finite(const or3d_state *s, char *message, int cap) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    struct { const char *name; const double *f; int64_t n; } fs[] = {
        {"pressure", s->p, nx * ny * nz},
        {"u", s->u, (nx + 1) * ny * nz},
        {"v", s->v, nx * (ny + 1) * nz},
        {"w", s->w, nx * ny * (nz + 1)},
    };
    for (int a = 0; a < 4; ++a)
        for (int64_t i = 0; i < fs[a].n; ++i)
            if (!isfinite(fs[a].f[i])) {
                if (message)
                    snprintf(message, cap, "Non-finite 3D %s value at flat index %lld: %g",
                             fs[a].name, (long long)i, fs[a].f[i]);
                return OR_ERR_SOLVER;
            }
    return OR_OK;
}

/* advance_one_step, incompressible_3d.rs:164-172. */
int or3d_step(or3d_state *s, or_solve_report *rep) {
    or3d_predict(s);
    or3d_build_rhs(s);
    int st = or3d_solve_pressure(s, rep);
    if (st != OR_OK) return st;
    or3d_correct(s);
    s->step += 1;
    st = or3d_ensure_finite(s, rep->message, sizeof rep->message);
    if (st != OR_OK) rep->status = st;
    return st;
}

/* max_divergence, incompressible_3d.rs:504-517. */
double or3d_max_divergence(const or3d_state *s) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    double m = 0.0;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                double div = (s->u[U3(i + 1, j, k)] - s->u[U3(i, j, k)]) / s->dx +
                             (s->v[V3(i, j + 1, k)] - s->v[V3(i, j, k)]) / s->dy +
                             (s->w[W3(i, j, k + 1)] - s->w[W3(i, j, k)]) / s->dz;
                m = fmax(m, fabs(div));
            }
    return m;
}

/* max_speed, incompressible_3d.rs:519-532. */
double or3d_max_speed(const or3d_state *s) {
    const int64_t nx = s->cfg.nx, ny = s->cfg.ny, nz = s->cfg.nz;
    double m = 0.0;
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                double a = 0.5 * (s->u[U3(i, j, k)] + s->u[U3(i + 1, j, k)]);
                double b = 0.5 * (s->v[V3(i, j, k)] + s->v[V3(i, j + 1, k)]);
                double c = 0.5 * (s->w[W3(i, j, k)] + s->w[W3(i, j, k + 1)]);
                m = fmax(m, sqrt(a * a + b * b + c * c));
            }
    return m;
}
