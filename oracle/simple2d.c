/*
 * Oracle: the collocated unstructured SIMPLE outer iteration on the 2D
 * polygon meshes of fvmesh.c (FP64), with the weighted least-squares
 * gradient it needs.
 * TEST INFRASTRUCTURE ONLY -- the device SIMPLE path is checked against it;
 * it is never part of the shipped path.
 *
 * Restates:
 *   LeastSquaresGradientStencil::new (2D)     src/numerics.rs:106-200
 *   inverse_normal_matrix (TwoD)              src/numerics.rs:227-254
 *   least_squares_gradient_into               src/numerics.rs:650-682
 *   interpolate_vector_into (OwnerValue)      src/numerics.rs:531-551
 *   integrated_divergence_into                src/numerics.rs:573-588
 *   pressure_gradient_source (density 1)      src/momentum.rs:403-416
 *   solve_momentum_velocity                   src/momentum.rs:445-475
 *   solve_simple                              src/simple.rs:148-311
 *   momentum_inverse_diagonal                 src/simple.rs:316-331
 *   pressure_face_coefficients(_with_boundary) src/simple.rs:342-405
 *   correct_face_flux(_with_boundary)         src/simple.rs:407-484
 *   assemble_pressure_correction_with_boundary src/simple.rs:583-638
 *   solve_pressure_correction                 src/simple.rs:643-668
 *   correct_cell_velocity_from_gradient       src/simple.rs:670-701
 *   continuity_rms                            src/simple.rs:703-717
 *   initial_face_flux                         src/simple.rs:719-744
 *   rhie_chow_predicted_flux                  src/simple.rs:746-805
 * The velocity is (x, y) with z = 0: on a 2D mesh the z momentum system has a
 * zero right-hand side and a zero initial guess, so the reference's BiCGSTAB
 * returns it unchanged after 0 iterations; the report says so.  Vec3 sums with
 * a zero z term are the same numbers as the two-term sums below.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define F64_EPSILON 2.220446049250313e-16

static int fail(char *message, int cap, int code, const char *fmt, long long a) {
    if (message && cap > 0) snprintf(message, cap, fmt, a);
    return code;
}

/* ------------------------------------------------------------ LSQ gradient */
int or_lsq2d_build(const or_mesh2d *m, const int32_t *bc_kind, const double *bc_value,
                   or_lsq2d *st, int64_t *bad, char *message, int cap) {
    memset(st, 0, sizeof(*st));
    const int64_t nc = m->ncells, ns = m->cell_off[nc];
    st->ncells = nc;
    st->kind = calloc(ns ? ns : 1, sizeof(int32_t));
    st->nb = calloc(ns ? ns : 1, sizeof(int64_t));
    st->val = calloc(ns ? ns : 1, sizeof(double));
    st->rcx = calloc(ns ? ns : 1, sizeof(double));
    st->rcy = calloc(ns ? ns : 1, sizeof(double));
    st->inv = calloc(4 * (nc ? nc : 1), sizeof(double));
    int rc = OR_OK;
    for (int64_t c = 0; c < nc && rc == OR_OK; ++c) {
        double n0 = 0.0, n1 = 0.0, n3 = 0.0, n4 = 0.0;
        for (int64_t e = m->cell_off[c]; e < m->cell_off[c + 1]; ++e) {
            const int64_t f = m->cell_faces[e];
            double ox, oy;
            if (m->neigh[f] >= 0) {
                const int64_t other = m->owner[f] == c ? m->neigh[f] : m->owner[f];
                ox = m->cx[other] - m->cx[c];
                oy = m->cy[other] - m->cy[c];
                st->kind[e] = OR_LSQ_NEIGHBOUR;
                st->nb[e] = other;
            } else {
                if (!bc_kind || (bc_kind[f] != OR_SCALAR_FIXED_VALUE && bc_kind[f] != OR_SCALAR_ZERO_GRADIENT)) {
                    *bad = f;
                    rc = fail(message, cap, OR_NUM_MISSING_BOUNDARY, "missing boundary condition at face %lld", f);
                    break;
                }
                if (bc_kind[f] == OR_SCALAR_ZERO_GRADIENT) {
                    st->kind[e] = OR_LSQ_SKIP;
                    continue;
                }
                ox = m->fcx[f] - m->cx[c];
                oy = m->fcy[f] - m->cy[c];
                st->kind[e] = OR_LSQ_FIXED;
                st->val[e] = bc_value[f];
            }
            const double d2 = ox * ox + oy * oy;
            if (d2 <= F64_EPSILON) {
                *bad = f;
                rc = m->neigh[f] >= 0
                         ? fail(message, cap, OR_NUM_DEGENERATE_ON, "degenerate owner-neighbour distance at face %lld", f)
                         : fail(message, cap, OR_NUM_DEGENERATE_OF, "degenerate owner-face distance at face %lld", f);
                break;
            }
            const double w = 1.0 / d2;
            const double rx = ox * w, ry = oy * w;
            n0 += rx * ox;
            n1 += rx * oy;
            n3 += ry * ox;
            n4 += ry * oy;
            st->rcx[e] = rx;
            st->rcy[e] = ry;
        }
        if (rc) break;
        /* inverse_normal_matrix, TwoD */
        if (!(isfinite(n0) && isfinite(n1) && isfinite(n3) && isfinite(n4))) {
            *bad = c;
            rc = fail(message, cap, OR_NUM_SINGULAR_LSQ, "singular least-squares stencil at cell %lld", c);
            break;
        }
        const double det = n0 * n4 - n1 * n3;
        const double scale = fmax(fmax(fabs(n0), fabs(n4)), 1.0);
        if (fabs(det) <= F64_EPSILON * scale * scale) {
            *bad = c;
            rc = fail(message, cap, OR_NUM_SINGULAR_LSQ, "singular least-squares stencil at cell %lld", c);
            break;
        }
        st->inv[4 * c + 0] = n4 / det;
        st->inv[4 * c + 1] = -n1 / det;
        st->inv[4 * c + 2] = -n3 / det;
        st->inv[4 * c + 3] = n0 / det;
    }
    if (rc) or_lsq2d_free(st);
    return rc;
}

void or_lsq2d_free(or_lsq2d *st) {
    void *f[] = {st->kind, st->nb, st->val, st->rcx, st->rcy, st->inv};
    for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); ++i) free(f[i]);
    memset(st, 0, sizeof(*st));
}

void or_lsq2d_gradient(const or_mesh2d *m, const or_lsq2d *st, const double *values, double *gx,
                       double *gy) {
    for (int64_t c = 0; c < m->ncells; ++c) {
        double rx = 0.0, ry = 0.0;
        for (int64_t e = m->cell_off[c]; e < m->cell_off[c + 1]; ++e) {
            if (st->kind[e] == OR_LSQ_SKIP) continue;
            const double adj = st->kind[e] == OR_LSQ_NEIGHBOUR ? values[st->nb[e]] : st->val[e];
            const double diff = adj - values[c];
            rx += st->rcx[e] * diff;
            ry += st->rcy[e] * diff;
        }
        const double *inv = st->inv + 4 * c;
        gx[c] = inv[0] * rx + inv[1] * ry;
        gy[c] = inv[2] * rx + inv[3] * ry;
    }
}

/* ------------------------------------------------------------ face helpers */
/* weight (numerics.rs:294-303): ((x_f - x_P) . d) / |d|^2, unclamped; 0 when
 * the owner-neighbour distance is degenerate (the caller reports it) */
static int weight(const or_mesh2d *m, int64_t f, double *l) {
    const int64_t P = m->owner[f], N = m->neigh[f];
    const double dx = m->cx[N] - m->cx[P], dy = m->cy[N] - m->cy[P];
    const double d2 = dx * dx + dy * dy;
    if (d2 <= F64_EPSILON) return 0;
    *l = ((m->fcx[f] - m->cx[P]) * dx + (m->fcy[f] - m->cy[P]) * dy) / d2;
    return 1;
}

int or_initial_face_flux(const or_mesh2d *m, const double *ux, const double *uy,
                         const double *wall_flux, double *flux, int64_t *bad, char *message, int cap) {
    for (int64_t f = 0; f < m->nfaces; ++f) {
        const int64_t P = m->owner[f], N = m->neigh[f];
        if (N < 0) {
            flux[f] = wall_flux[f];
            continue;
        }
        double l;
        if (!weight(m, f, &l)) {
            *bad = f;
            return fail(message, cap, OR_NUM_DEGENERATE_ON, "degenerate owner-neighbour distance at face %lld", f);
        }
        const double vx = ux[P] * (1.0 - l) + ux[N] * l, vy = uy[P] * (1.0 - l) + uy[N] * l;
        flux[f] = vx * m->sx[f] + vy * m->sy[f];
    }
    return OR_OK;
}

double or_continuity_rms(const or_mesh2d *m, const double *flux) {
    double *imb = calloc(m->ncells ? m->ncells : 1, sizeof(double));
    for (int64_t f = 0; f < m->nfaces; ++f) {
        imb[m->owner[f]] += flux[f];
        if (m->neigh[f] >= 0) imb[m->neigh[f]] -= flux[f];
    }
    double s = 0.0;
    for (int64_t c = 0; c < m->ncells; ++c) s += imb[c] * imb[c];
    free(imb);
    return sqrt(s / (double)m->ncells);
}

/* pressure_face_coefficients(_with_boundary): d_f on interior faces, and on
 * FixedValue correction faces when corr_kind is given; 0 elsewhere. */
static int face_coefficients(const or_mesh2d *m, const double *r_au, const int32_t *corr_kind,
                             double *coef, int64_t *bad, double *badv, char *message, int cap) {
    for (int64_t f = 0; f < m->nfaces; ++f) {
        const int64_t P = m->owner[f], N = m->neigh[f];
        coef[f] = 0.0;
        if (N >= 0) {
            double l;
            if (!weight(m, f, &l)) {
                *bad = f;
                return fail(message, cap, OR_NUM_DEGENERATE_ON, "degenerate owner-neighbour distance at face %lld", f);
            }
            const double fr = (1.0 - l) * r_au[P] + l * r_au[N];
            const double dx = m->cx[N] - m->cx[P], dy = m->cy[N] - m->cy[P];
            const double d2 = dx * dx + dy * dy;
            if (!isfinite(d2) || d2 <= F64_EPSILON) {
                *bad = f;
                return fail(message, cap, OR_NUM_DEGENERATE_ON, "degenerate owner-neighbour distance at face %lld", f);
            }
            const double c = fr * (m->sx[f] * dx + m->sy[f] * dy) / d2;
            if (!isfinite(c) || c <= 0.0) {
                *bad = f;
                *badv = c;
                return fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", f);
            }
            coef[f] = c;
        }
    }
    if (!corr_kind) return OR_OK;
    for (int64_t f = 0; f < m->nfaces; ++f) {
        if (m->neigh[f] >= 0 || corr_kind[f] != OR_SCALAR_FIXED_VALUE) continue;
        const int64_t P = m->owner[f];
        const double dx = m->fcx[f] - m->cx[P], dy = m->fcy[f] - m->cy[P];
        const double d2 = dx * dx + dy * dy;
        if (!isfinite(d2) || d2 <= F64_EPSILON) {
            *bad = f;
            return fail(message, cap, OR_NUM_DEGENERATE_OF, "degenerate owner-face distance at face %lld", f);
        }
        const double c = r_au[P] * (m->sx[f] * dx + m->sy[f] * dy) / d2;
        if (!isfinite(c) || c <= 0.0) {
            *bad = f;
            *badv = c;
            return fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", f);
        }
        coef[f] = c;
    }
    return OR_OK;
}

int or_rhie_chow_flux(const or_mesh2d *m, const or_lsq2d *pst, const double *usx, const double *usy,
                      const double *p, const double *r_au, const double *wall_flux, double *flux,
                      int64_t *bad, double *badv, char *message, int cap) {
    const int64_t nc = m->ncells, nf = m->nfaces;
    double *gx = malloc(sizeof(double) * nc), *gy = malloc(sizeof(double) * nc);
    double *coef = malloc(sizeof(double) * (nf ? nf : 1));
    or_lsq2d_gradient(m, pst, p, gx, gy);
    int rc = face_coefficients(m, r_au, NULL, coef, bad, badv, message, cap);
    for (int64_t f = 0; f < nf && rc == OR_OK; ++f) {
        const int64_t P = m->owner[f], N = m->neigh[f];
        if (N < 0) {
            flux[f] = wall_flux[f];
            continue;
        }
        double l;
        if (!weight(m, f, &l)) {
            *bad = f;
            rc = fail(message, cap, OR_NUM_DEGENERATE_ON, "degenerate owner-neighbour distance at face %lld", f);
            break;
        }
        const double vx = usx[P] * (1.0 - l) + usx[N] * l, vy = usy[P] * (1.0 - l) + usy[N] * l;
        const double rxP = gx[P] * r_au[P], ryP = gy[P] * r_au[P];
        const double rxN = gx[N] * r_au[N], ryN = gy[N] * r_au[N];
        const double qx = rxP * (1.0 - l) + rxN * l, qy = ryP * (1.0 - l) + ryN * l;
        flux[f] = (vx * m->sx[f] + vy * m->sy[f]) + (qx * m->sx[f] + qy * m->sy[f]) -
                  coef[f] * (p[N] - p[P]);
        if (!isfinite(flux[f])) {
            *bad = f;
            *badv = flux[f];
            rc = fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", f);
        }
    }
    free(gx); free(gy); free(coef);
    return rc;
}

/* correct_face_flux(_with_boundary) */
static int correct_flux(const or_mesh2d *m, const double *coef, const double *pc, const int32_t *corr_kind,
                        const double *corr_value, double *flux, int64_t *bad, double *badv, char *message,
                        int cap) {
    for (int64_t f = 0; f < m->nfaces; ++f) {
        const int64_t P = m->owner[f], N = m->neigh[f];
        double other;
        if (N >= 0) {
            other = pc[N];
        } else {
            if (!corr_kind || corr_kind[f] != OR_SCALAR_FIXED_VALUE) continue;
            other = corr_value[f];
        }
        const double c = coef[f];
        if (!isfinite(c) || c <= 0.0) {
            *bad = f;
            *badv = c;
            return fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", f);
        }
        flux[f] -= c * (other - pc[P]);
        if (!isfinite(flux[f])) {
            *bad = f;
            *badv = flux[f];
            return fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", f);
        }
    }
    return OR_OK;
}

int or_correct_face_flux(const or_mesh2d *m, const double *predicted, const double *coef,
                         const double *pc, const int32_t *corr_kind, const double *corr_value,
                         double *out, int64_t *bad, double *badv, char *message, int cap) {
    memcpy(out, predicted, sizeof(double) * m->nfaces);
    return correct_flux(m, coef, pc, corr_kind, corr_value, out, bad, badv, message, cap);
}

int or_correct_cell_velocity(int64_t nc, const double *usx, const double *usy, const double *r_au,
                             const double *gx, const double *gy, double *ux, double *uy, int64_t *bad,
                             double *badv, char *message, int cap) {
    for (int64_t c = 0; c < nc; ++c) {
        const double r = r_au[c];
        if (!isfinite(r) || r <= 0.0) {
            *bad = c;
            *badv = r;
            return fail(message, cap, OR_SIMPLE_INVALID_MOM_DIAG, "invalid momentum diagonal in cell %lld", c);
        }
        ux[c] = usx[c] - gx[c] * r;
        uy[c] = usy[c] - gy[c] * r;
        if (!(isfinite(ux[c]) && isfinite(uy[c]))) {
            *bad = c;
            *badv = r;
            return fail(message, cap, OR_SIMPLE_INVALID_MOM_DIAG, "invalid momentum diagonal in cell %lld", c);
        }
    }
    return OR_OK;
}

/* assemble_pressure_correction_with_boundary (simple.rs:583-638) */
static int assemble_pc_boundary(const or_mesh2d *m, const double *flux, const double *coef,
                                const int32_t *corr_kind, const double *corr_value, int64_t *nnz_out,
                                int64_t **off_out, int64_t **col_out, double **val_out, double **rhs_out,
                                int64_t *bad, double *badv, char *message, int cap) {
    const int64_t nc = m->ncells, nf = m->nfaces;
    double *rhs = calloc(nc, sizeof(double));
    int64_t *tr = malloc(sizeof(int64_t) * (4 * nf + 1)), *tc = malloc(sizeof(int64_t) * (4 * nf + 1));
    double *tv = malloc(sizeof(double) * (4 * nf + 1));
    int64_t k = 0;
    int anchored = 0, rc = OR_OK;
    for (int64_t f = 0; f < nf && rc == OR_OK; ++f) {
        const double phi = flux[f];
        if (!isfinite(phi)) {
            *bad = f;
            *badv = phi;
            rc = fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", f);
            break;
        }
        const int64_t P = m->owner[f], N = m->neigh[f];
        rhs[P] -= phi;
        if (N >= 0) {
            rhs[N] += phi;
            const double c = coef[f];
            if (!isfinite(c) || c <= 0.0) {
                *bad = f;
                *badv = c;
                rc = fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", f);
                break;
            }
            tr[k] = P; tc[k] = P; tv[k++] = c;
            tr[k] = P; tc[k] = N; tv[k++] = -c;
            tr[k] = N; tc[k] = P; tv[k++] = -c;
            tr[k] = N; tc[k] = N; tv[k++] = c;
        } else if (corr_kind[f] == OR_SCALAR_FIXED_VALUE) {
            const double c = coef[f], value = corr_value[f];
            if (!isfinite(c) || c <= 0.0 || !isfinite(value)) {
                *bad = f;
                *badv = c;
                rc = fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", f);
                break;
            }
            tr[k] = P; tc[k] = P; tv[k++] = c;
            rhs[P] += c * value;
            anchored = 1;
        }
    }
    if (rc == OR_OK && !anchored)
        rc = fail(message, cap, OR_SIMPLE_INVALID_REFERENCE, "invalid pressure reference %lld", 0);
    int64_t *off = NULL, *col = NULL;
    double *val = NULL;
    int64_t nnz = 0;
    if (rc == OR_OK) {
        off = malloc(sizeof(int64_t) * (nc + 1));
        col = malloc(sizeof(int64_t) * (k ? k : 1));
        val = malloc(sizeof(double) * (k ? k : 1));
        or_lin_report lr;
        if (or_csr_finalize(nc, nc, k, tr, tc, tv, off, col, val, &nnz, &lr) != OR_OK)
            rc = fail(message, cap, OR_SIMPLE_LINEAR, "pressure-correction matrix: linear algebra error%lld", 0);
    }
    free(tr); free(tc); free(tv);
    if (rc == OR_OK) {
        *nnz_out = nnz;
        *off_out = off;
        *col_out = col;
        *val_out = val;
        *rhs_out = rhs;
    } else {
        free(off); free(col); free(val); free(rhs);
    }
    return rc;
}

/* ------------------------------------------------------------ solve_simple */
int or_simple2d_solve(const or_mesh2d *m, const or_lsq2d *pst, const int32_t *vel_kind,
                      const double *vel_x, const double *vel_y, const int32_t *corr_kind,
                      const double *corr_value, const or_simple_opts *o, double *ux, double *uy,
                      double *p, double *flux, double *history, or_simple_report *rep,
                      char *message, int cap) {
    memset(rep, 0, sizeof(*rep));
    const int64_t nc = m->ncells, nf = m->nfaces;
    if (o->max_outer_iterations <= 0)
        return fail(message, cap, OR_SIMPLE_INVALID_OUTER, "invalid outer iterations %lld", 0);
    if (!(isfinite(o->alpha_p) && 0.0 < o->alpha_p && o->alpha_p <= 1.0)) {
        rep->err_value = o->alpha_p;
        return fail(message, cap, OR_SIMPLE_INVALID_RELAX, "invalid pressure relaxation%lld", 0);
    }
    if (!(isfinite(o->continuity_tol) && o->continuity_tol >= 0.0)) {
        rep->err_value = o->continuity_tol;
        return fail(message, cap, OR_SIMPLE_INVALID_RELAX, "invalid pressure relaxation%lld", 0);
    }
    if (!corr_kind && (o->reference_cell < 0 || o->reference_cell >= nc))
        return fail(message, cap, OR_SIMPLE_INVALID_REFERENCE, "invalid pressure reference %lld", o->reference_cell);
    const double initial = or_continuity_rms(m, flux);
    history[0] = initial;
    rep->initial_rms = initial;
    rep->final_rms = initial;
    double *gx = malloc(sizeof(double) * nc), *gy = malloc(sizeof(double) * nc);
    double *sx = malloc(sizeof(double) * nc), *sy = malloc(sizeof(double) * nc);
    double *psx = malloc(sizeof(double) * nc), *psy = malloc(sizeof(double) * nc);
    double *r_au = malloc(sizeof(double) * nc), *pc = malloc(sizeof(double) * nc);
    double *pred = malloc(sizeof(double) * (nf ? nf : 1)), *coef = malloc(sizeof(double) * (nf ? nf : 1));
    int32_t *mk = malloc(sizeof(int32_t) * (nf ? nf : 1));
    double *mvx = calloc(nf ? nf : 1, sizeof(double)), *mvy = calloc(nf ? nf : 1, sizeof(double));
    for (int64_t f = 0; f < nf; ++f) {
        mk[f] = m->neigh[f] >= 0 ? -1 : vel_kind[f];
        if (m->neigh[f] < 0) {
            mvx[f] = vel_x[f];
            mvy[f] = vel_y[f];
        }
    }
    const int relaxed = o->alpha_u < 1.0;
    int rc = OR_OK;
    int64_t it;
    for (it = 1; it <= o->max_outer_iterations && rc == OR_OK; ++it) {
        /* pressure_gradient_source: -grad(p) / 1 */
        or_lsq2d_gradient(m, pst, p, gx, gy);
        for (int64_t c = 0; c < nc; ++c) {
            sx[c] = -gx[c] / 1.0;
            sy[c] = -gy[c] / 1.0;
        }
        int64_t nnz[2];
        int64_t *off[2] = {NULL, NULL}, *col[2] = {NULL, NULL};
        double *val[2] = {NULL, NULL}, *rhs[2] = {NULL, NULL}, *diag[2] = {NULL, NULL};
        const double *src[2] = {sx, sy}, *cur[2] = {ux, uy}, *bv[2] = {mvx, mvy};
        for (int a = 0; a < 2 && rc == OR_OK; ++a)
            rc = or_assemble_momentum(m, o->viscosity, flux, mk, bv[a], src[a], NULL, 0.0,
                                      relaxed ? cur[a] : NULL, o->alpha_u, &nnz[a], &off[a], &col[a],
                                      &val[a], &rhs[a], &diag[a], message, cap);
        /* solve_momentum_velocity: BiCGSTAB from the current velocity */
        if (rc == OR_OK) {
            memcpy(psx, ux, sizeof(double) * nc);
            memcpy(psy, uy, sizeof(double) * nc);
            double *px[2] = {psx, psy};
            for (int a = 0; a < 2 && rc == OR_OK; ++a) {
                or_lin_report lr;
                if (or_bicgstab(nc, off[a], col[a], val[a], rhs[a], px[a], o->mom_atol, o->mom_rtol,
                                o->mom_max, &lr) != OR_OK) {
                    rc = fail(message, cap, OR_SIMPLE_LINEAR, "momentum solve: %lld", lr.status);
                    break;
                }
                rep->mom_iters[a] = lr.iterations;
                rep->mom_converged[a] = lr.converged;
                rep->mom_final[a] = lr.final_residual;
            }
            rep->mom_iters[2] = 0;
            rep->mom_converged[2] = 1;
            rep->mom_final[2] = 0.0;
            if (rc == OR_OK && !(rep->mom_converged[0] && rep->mom_converged[1]))
                rc = fail(message, cap, OR_SIMPLE_MOM_NOT_CONVERGED, "momentum linear solve did not converge%lld", 0);
        }
        /* momentum_inverse_diagonal of the assembled (relaxed) x diagonal */
        if (rc == OR_OK) {
            for (int64_t c = 0; c < nc && rc == OR_OK; ++c) {
                double d = 0.0;
                for (int64_t e = off[0][c]; e < off[0][c + 1]; ++e)
                    if (col[0][e] == c) d = val[0][e];
                if (!isfinite(d) || d <= 0.0) {
                    rep->err_index = c;
                    rep->err_value = d;
                    rc = fail(message, cap, OR_SIMPLE_INVALID_MOM_DIAG, "invalid momentum diagonal in cell %lld", c);
                    break;
                }
                r_au[c] = 1.0 / d;
            }
        }
        for (int a = 0; a < 2; ++a) {
            free(off[a]); free(col[a]); free(val[a]); free(rhs[a]); free(diag[a]);
        }
        if (rc) break;
        rc = or_rhie_chow_flux(m, pst, psx, psy, p, r_au, flux, pred, &rep->err_index, &rep->err_value,
                               message, cap);
        if (rc) break;
        if (corr_kind)
            for (int64_t f = 0; f < nf; ++f)
                if (m->neigh[f] < 0 && corr_kind[f] == OR_SCALAR_FIXED_VALUE) {
                    const int64_t P = m->owner[f];
                    pred[f] = psx[P] * m->sx[f] + psy[P] * m->sy[f];
                }
        rc = face_coefficients(m, r_au, corr_kind, coef, &rep->err_index, &rep->err_value, message, cap);
        if (rc) break;
        int64_t pnnz;
        int64_t *poff = NULL, *pcol = NULL;
        double *pval = NULL, *prhs = NULL;
        if (corr_kind)
            rc = assemble_pc_boundary(m, pred, coef, corr_kind, corr_value, &pnnz, &poff, &pcol, &pval,
                                      &prhs, &rep->err_index, &rep->err_value, message, cap);
        else
            rc = or_assemble_pressure_correction(m, pred, coef, o->reference_cell, &pnnz, &poff, &pcol,
                                                 &pval, &prhs, message, cap);
        if (rc) break;
        memset(pc, 0, sizeof(double) * nc);
        or_lin_report lr;
        if (or_cg(nc, poff, pcol, pval, prhs, pc, o->p_atol, o->p_rtol, o->p_max, 1, &lr) != OR_OK)
            rc = fail(message, cap, OR_SIMPLE_LINEAR, "pressure-correction solve: %lld", lr.status);
        free(poff); free(pcol); free(pval); free(prhs);
        if (rc) break;
        rep->p_iters = lr.iterations;
        rep->p_converged = lr.converged;
        rep->p_final = lr.final_residual;
        if (!lr.converged) {
            rc = fail(message, cap, OR_SIMPLE_P_NOT_CONVERGED, "pressure linear solve did not converge%lld", 0);
            break;
        }
        for (int64_t c = 0; c < nc; ++c) p[c] += o->alpha_p * pc[c];
        or_lsq2d_gradient(m, pst, pc, gx, gy);
        rc = or_correct_cell_velocity(nc, psx, psy, r_au, gx, gy, ux, uy, &rep->err_index, &rep->err_value,
                                      message, cap);
        if (rc) break;
        memcpy(flux, pred, sizeof(double) * nf);
        rc = correct_flux(m, coef, pc, corr_kind, corr_value, flux, &rep->err_index, &rep->err_value,
                          message, cap);
        if (rc) break;
        const double cont = or_continuity_rms(m, flux);
        if (!isfinite(cont)) {
            rep->err_value = cont;
            rc = fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", 0);
            break;
        }
        history[it] = cont;
        rep->final_rms = cont;
        rep->outer_iterations = it;
        if (cont <= o->continuity_tol) {
            rep->converged = 1;
            break;
        }
    }
    free(gx); free(gy); free(sx); free(sy); free(psx); free(psy); free(r_au); free(pc);
    free(pred); free(coef); free(mk); free(mvx); free(mvy);
    if (rc == OR_OK && !rep->converged) rep->outer_iterations = o->max_outer_iterations;
    return rc;
}
