/*
 * Oracle: the caller-shaped CSR systems of the reference's unstructured
 * solvers -- the SIMPLE pressure correction and one momentum component --
 * assembled on the 2D polygon mesh of fvmesh.c (FP64).
 * TEST INFRASTRUCTURE ONLY -- produces the systems the GPU CSR solvers are
 * checked on; it is never part of the shipped path.
 *
 * Restates:
 *   weight (interior-face interpolation)    src/numerics.rs:294-303
 *   interpolate_scalar_into (OwnerValue)    src/numerics.rs:500-521
 *   interpolate_diffusivity_into (Linear,   src/numerics.rs:383-418 (the default
 *     constant viscosity)                   DiffusionOptions, numerics.rs:41-49)
 *   pressure_face_coefficients              src/simple.rs:342-370
 *   assemble_pressure_correction            src/simple.rs:518-579
 *   assemble_momentum_component             src/momentum.rs:180-341 (no explicit
 *                                           non-orthogonal correction: the
 *                                           default NonOrthogonalCorrection::None
 *                                           makes add_explicit_nonorthogonal_rhs
 *                                           a no-op, :343-350)
 *   CsrBuilder::finalize                    src/linear/csr.rs:159-192 (or_csr_finalize)
 * Vec3::dot is x x' + y y' + z z' (src/mesh.rs:35-37); in 2D the z term adds
 * an exact zero, so the two-term sums below are the same numbers.
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

/* lambda = ((x_f - x_P) . d) / |d|^2, d = x_N - x_P (numerics.rs:294-303) */
static int face_weight(const or_mesh2d *m, int64_t f, double *lambda) {
    const int64_t P = m->owner[f], N = m->neigh[f];
    const double dx = m->cx[N] - m->cx[P], dy = m->cy[N] - m->cy[P];
    const double d2 = dx * dx + dy * dy;
    if (d2 <= F64_EPSILON) return 0;
    *lambda = ((m->fcx[f] - m->cx[P]) * dx + (m->fcy[f] - m->cy[P]) * dy) / d2;
    return 1;
}

int or_face_weights(const or_mesh2d *m, double *lambda, char *message, int cap) {
    for (int64_t f = 0; f < m->nfaces; ++f) {
        lambda[f] = 0.0;
        if (m->neigh[f] >= 0 && !face_weight(m, f, &lambda[f]))
            return fail(message, cap, OR_SIMPLE_DEGENERATE, "degenerate owner-neighbour distance at face %lld", f);
    }
    return OR_OK;
}

int or_pressure_face_coefficients(const or_mesh2d *m, const double *r_au, double *coef, char *message,
                                  int cap) {
    for (int64_t f = 0; f < m->nfaces; ++f) {
        coef[f] = 0.0;
        const int64_t P = m->owner[f], N = m->neigh[f];
        if (N < 0) continue;  /* boundary coefficients stay zero (simple.rs:351-353) */
        double lambda;
        if (!face_weight(m, f, &lambda))
            return fail(message, cap, OR_SIMPLE_DEGENERATE, "degenerate owner-neighbour distance at face %lld", f);
        const double face_r_au = (1.0 - lambda) * r_au[P] + lambda * r_au[N];
        const double dx = m->cx[N] - m->cx[P], dy = m->cy[N] - m->cy[P];
        const double d2 = dx * dx + dy * dy;
        if (!isfinite(d2) || d2 <= F64_EPSILON)
            return fail(message, cap, OR_SIMPLE_DEGENERATE, "degenerate owner-neighbour distance at face %lld", f);
        const double c = face_r_au * (m->sx[f] * dx + m->sy[f] * dy) / d2;
        if (!isfinite(c) || c <= 0.0)
            return fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", f);
        coef[f] = c;
    }
    return OR_OK;
}

static int finalize(int64_t n, int64_t k, const int64_t *tr, const int64_t *tc, const double *tv,
                    int64_t *nnz_out, int64_t **off_out, int64_t **col_out, double **val_out,
                    char *message, int cap, int code) {
    int64_t *off = malloc(sizeof(int64_t) * (n + 1)), *col = malloc(sizeof(int64_t) * (k ? k : 1));
    double *val = malloc(sizeof(double) * (k ? k : 1));
    or_lin_report rep;
    memset(&rep, 0, sizeof rep);
    int64_t nnz = 0;
    if (or_csr_finalize(n, n, k, tr, tc, tv, off, col, val, &nnz, &rep) != OR_OK) {
        if (message) snprintf(message, cap, "CSR finalize failed: %s", rep.reason);
        free(off); free(col); free(val);
        return code;
    }
    *nnz_out = nnz;
    *off_out = off;
    *col_out = col;
    *val_out = val;
    return OR_OK;
}

int or_assemble_pressure_correction(const or_mesh2d *m, const double *flux, const double *coef,
                                    int64_t ref_cell, int64_t *nnz_out, int64_t **off_out,
                                    int64_t **col_out, double **val_out, double **rhs_out,
                                    char *message, int cap) {
    const int64_t nc = m->ncells, nf = m->nfaces;
    if (ref_cell < 0 || ref_cell >= nc)
        return fail(message, cap, OR_SIMPLE_INVALID_REFERENCE, "invalid pressure reference cell %lld", ref_cell);
    double *rhs = calloc(nc, sizeof(double));
    const int64_t tcap = 4 * nf + 1;
    int64_t *tr = malloc(sizeof(int64_t) * tcap), *tc = malloc(sizeof(int64_t) * tcap);
    double *tv = malloc(sizeof(double) * tcap);
    int64_t k = 0;
    int rc = OR_OK;
    for (int64_t f = 0; f < nf; ++f) {
        const double phi = flux[f];
        if (!isfinite(phi)) {
            rc = fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", f);
            break;
        }
        const int64_t P = m->owner[f], N = m->neigh[f];
        rhs[P] -= phi;
        if (N < 0) continue;
        rhs[N] += phi;
        const double c = coef[f];
        if (!isfinite(c) || c <= 0.0) {
            rc = fail(message, cap, OR_SIMPLE_INVALID_FACE_COEF, "invalid pressure face coefficient at face %lld", f);
            break;
        }
        /* the four entries in the reference's order (simple.rs:553-558), the pinned
         * row and column dropped (:560-567) */
        const int64_t r4[4] = {P, P, N, N}, c4[4] = {P, N, P, N};
        const double v4[4] = {c, -c, -c, c};
        for (int e = 0; e < 4; ++e) {
            if (r4[e] == ref_cell || c4[e] == ref_cell) continue;
            tr[k] = r4[e]; tc[k] = c4[e]; tv[k++] = v4[e];
        }
    }
    if (rc == OR_OK) {
        rhs[ref_cell] = 0.0;
        tr[k] = ref_cell; tc[k] = ref_cell; tv[k++] = 1.0;
        rc = finalize(nc, k, tr, tc, tv, nnz_out, off_out, col_out, val_out, message, cap, OR_SIMPLE_LINEAR);
    }
    free(tr); free(tc); free(tv);
    if (rc == OR_OK) *rhs_out = rhs;
    else free(rhs);
    return rc;
}

int or_assemble_momentum(const or_mesh2d *m, double nu, const double *flux, const int32_t *bc_kind,
                         const double *bc_value, const double *source, const double *old, double dt,
                         const double *relax_field, double alpha, int64_t *nnz_out, int64_t **off_out,
                         int64_t **col_out, double **val_out, double **rhs_out, double **diag_out,
                         char *message, int cap) {
    const int64_t nc = m->ncells, nf = m->nfaces;
    if (!(isfinite(alpha) && 0.0 < alpha && alpha <= 1.0))
        return fail(message, cap, OR_MOM_INVALID_RELAX, "invalid under-relaxation%lld", 0);
    if (old && !(isfinite(dt) && dt > 0.0))
        return fail(message, cap, OR_MOM_INVALID_DT, "invalid time step%lld", 0);
    if (!isfinite(nu) || nu < 0.0)
        return fail(message, cap, OR_MOM_FACE_DIFFUSIVITY, "invalid face diffusivity at face %lld", 0);
    double *rhs = calloc(nc, sizeof(double));
    const int64_t tcap = 4 * nf + nc + 1;
    int64_t *tr = malloc(sizeof(int64_t) * tcap), *tc = malloc(sizeof(int64_t) * tcap);
    double *tv = malloc(sizeof(double) * tcap);
    int64_t k = 0;
    int rc = OR_OK;
    /* sources and transient storage per cell (momentum.rs:224-240) */
    for (int64_t c = 0; c < nc && rc == OR_OK; ++c) {
        if (source) {
            if (!isfinite(source[c])) {
                rc = fail(message, cap, OR_MOM_NONFINITE_SOURCE, "non-finite source in cell %lld", c);
                break;
            }
            rhs[c] += source[c] * m->vol[c];
        }
        if (old) {
            const double coefficient = m->vol[c] / dt;
            tr[k] = c; tc[k] = c; tv[k++] = coefficient;
            rhs[c] += coefficient * old[c];
        }
    }
    /* faces (momentum.rs:241-302); the face viscosity is the Linear interpolation of the
     * constant, (1 - lambda) nu + lambda nu, on interior faces and the owner
     * value on boundary faces (numerics.rs:393-412) */
    for (int64_t f = 0; f < nf && rc == OR_OK; ++f) {
        const double phi = flux[f];
        if (!isfinite(phi)) {
            rc = fail(message, cap, OR_MOM_NONFINITE_FLUX, "non-finite flux at face %lld", f);
            break;
        }
        const int64_t P = m->owner[f], N = m->neigh[f];
        if (N >= 0) {
            double lambda;
            if (!face_weight(m, f, &lambda)) {
                rc = fail(message, cap, OR_MOM_DEGENERATE, "degenerate owner-neighbour distance at face %lld", f);
                break;
            }
            const double nu_f = (1.0 - lambda) * nu + lambda * nu;
            const double dx = m->cx[N] - m->cx[P], dy = m->cy[N] - m->cy[P];
            const double d2 = dx * dx + dy * dy;
            if (!isfinite(d2) || d2 <= F64_EPSILON) {
                rc = fail(message, cap, OR_MOM_DEGENERATE, "degenerate owner-neighbour distance at face %lld", f);
                break;
            }
            const double diffusion = nu_f * (m->sx[f] * dx + m->sy[f] * dy) / d2;
            if (!isfinite(diffusion) || diffusion < 0.0) {
                rc = fail(message, cap, OR_MOM_GEOMETRY, "invalid non-orthogonal geometry at face %lld", f);
                break;
            }
            const double pmax = phi > 0.0 ? phi : 0.0, pmin = phi < 0.0 ? phi : 0.0;
            tr[k] = P; tc[k] = P; tv[k++] = diffusion + pmax;
            tr[k] = P; tc[k] = N; tv[k++] = -diffusion + pmin;
            tr[k] = N; tc[k] = P; tv[k++] = -diffusion - pmax;
            tr[k] = N; tc[k] = N; tv[k++] = diffusion - pmin;
        } else {
            if (!bc_kind || (bc_kind[f] != OR_SCALAR_FIXED_VALUE && bc_kind[f] != OR_SCALAR_ZERO_GRADIENT)) {
                rc = fail(message, cap, OR_MOM_MISSING_BOUNDARY, "missing boundary condition at face %lld", f);
                break;
            }
            const double dx = m->fcx[f] - m->cx[P], dy = m->fcy[f] - m->cy[P];
            const double d2 = dx * dx + dy * dy;
            if (!isfinite(d2) || d2 <= F64_EPSILON) {
                rc = fail(message, cap, OR_MOM_DEGENERATE, "degenerate owner-face distance at face %lld", f);
                break;
            }
            const double diffusion = nu * (m->sx[f] * dx + m->sy[f] * dy) / d2;
            if (bc_kind[f] == OR_SCALAR_FIXED_VALUE) {
                const double value = bc_value[f];
                if (!isfinite(value)) {
                    rc = fail(message, cap, OR_MOM_NONFINITE_SOURCE, "non-finite source in cell %lld", P);
                    break;
                }
                const double pmax = phi > 0.0 ? phi : 0.0, pmin = phi < 0.0 ? phi : 0.0;
                tr[k] = P; tc[k] = P; tv[k++] = diffusion + pmax;
                rhs[P] += diffusion * value - pmin * value;
            } else {
                if (phi < 0.0) {
                    rc = fail(message, cap, OR_MOM_BACKFLOW, "backflow on a zero-gradient face %lld", f);
                    break;
                }
                tr[k] = P; tc[k] = P; tv[k++] = phi;
            }
        }
    }
    int64_t nnz = 0, *off = NULL, *col = NULL;
    double *val = NULL, *diag = NULL;
    if (rc == OR_OK)
        rc = finalize(nc, k, tr, tc, tv, &nnz, &off, &col, &val, message, cap, OR_MOM_LINEAR);
    free(tr); free(tc); free(tv);
    if (rc == OR_OK) {
        /* unrelaxed diagonal (CsrMatrix::diagonal: a missing one is an error) */
        diag = malloc(sizeof(double) * (nc ? nc : 1));
        for (int64_t r = 0; r < nc && rc == OR_OK; ++r) {
            int found = 0;
            for (int64_t e = off[r]; e < off[r + 1]; ++e)
                if (col[e] == r) {
                    diag[r] = val[e];
                    found = 1;
                }
            if (!found) rc = fail(message, cap, OR_MOM_LINEAR, "missing diagonal in row %lld", r);
        }
    }
    if (rc == OR_OK && alpha != 1.0) {
        /* equation under-relaxation (momentum.rs:318-340): a_P / alpha on the diagonal,
         * rhs += (1 - alpha) / alpha a_P x_current; rebuilding through the
         * builder keeps the same (sorted, duplicate-free) entries */
        if (!relax_field) {
            rc = fail(message, cap, OR_MOM_MISSING_RELAX_FIELD, "missing velocity for under-relaxation%lld", 0);
        } else {
            for (int64_t r = 0; r < nc; ++r) {
                for (int64_t e = off[r]; e < off[r + 1]; ++e)
                    if (col[e] == r) val[e] /= alpha;
                rhs[r] += (1.0 - alpha) / alpha * diag[r] * relax_field[r];
            }
        }
    }
    if (rc == OR_OK) {
        *nnz_out = nnz;
        *off_out = off;
        *col_out = col;
        *val_out = val;
        *rhs_out = rhs;
        *diag_out = diag;
    } else {
        free(off); free(col); free(val); free(diag); free(rhs);
    }
    return rc;
}
