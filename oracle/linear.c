/*
 * Oracle: serial CSR linear algebra (FP64).
 * TEST INFRASTRUCTURE ONLY -- the checker, never the shipped path.
 *
 * Restates src/linear/:
 *   validate_csr                 csr.rs:195-251
 *   spmv_into                    csr.rs:89-106
 *   CsrBuilder::finalize         csr.rs:159-192
 *   JacobiPreconditioner::new    preconditioner.rs:9-28 (+ diagonal_entry csr.rs:68-84)
 *   solve_cg (cg / pcg)          solver.rs:62-153
 *   bicgstab                     solver.rs:155-256
 *   validate_solve_inputs        solver.rs:258-300
 *   dot / norm_l2 / residual     vector.rs:4-32,112-138
 * Error kinds and "reason" strings are the reference's, verbatim.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define BREAKDOWN_EPSILON 1.0e-30

static int set_err(or_lin_report *rep, int status, const char *reason) {
    rep->status = status;
    if (reason) snprintf(rep->reason, sizeof rep->reason, "%s", reason);
    return status;
}
static int nonfinite(or_lin_report *rep, const char *ctx, int64_t index) {
    rep->err_index = index;
    return set_err(rep, OR_ERR_NON_FINITE, ctx);
}

int or_csr_validate(int64_t nrows, int64_t ncols, const int64_t *offsets, const int64_t *cols,
                    const double *vals, int64_t nnz, or_lin_report *rep) {
    if (offsets[0] != 0) return set_err(rep, OR_ERR_INVALID_CSR, "first row offset must be zero");
    if (offsets[nrows] != nnz)
        return set_err(rep, OR_ERR_INVALID_CSR, "last row offset must equal value count");
    for (int64_t r = 0; r < nrows; ++r) {
        if (offsets[r] > offsets[r + 1])
            return set_err(rep, OR_ERR_INVALID_CSR, "row offsets must be monotonic");
        for (int64_t e = offsets[r] + 1; e < offsets[r + 1]; ++e)
            if (cols[e - 1] >= cols[e])
                return set_err(rep, OR_ERR_INVALID_CSR, "row columns must be strictly ascending");
    }
    for (int64_t e = 0; e < nnz; ++e)
        if (cols[e] < 0 || cols[e] >= ncols) {
            rep->err_index = e;
            return set_err(rep, OR_ERR_INVALID_CSR, "column exceeds ncols");
        }
    for (int64_t e = 0; e < nnz; ++e)
        if (!isfinite(vals[e])) return nonfinite(rep, "matrix", e);
    return OR_OK;
}

int or_csr_spmv(int64_t n, const int64_t *offsets, const int64_t *cols, const double *vals,
                const double *x, double *y, or_lin_report *rep) {
    for (int64_t r = 0; r < n; ++r) {
        double v = 0.0;
        for (int64_t e = offsets[r]; e < offsets[r + 1]; ++e) v += vals[e] * x[cols[e]];
        if (!isfinite(v)) return nonfinite(rep, "SpMV output", r);
        y[r] = v;
    }
    return OR_OK;
}

typedef struct {
    int64_t row, col;
    double val;
    int64_t seq;
} trip;

static int trip_cmp(const void *a, const void *b) {
    const trip *x = a, *y = b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    /* sort_unstable_by_key leaves equal keys in an unspecified order; the
     * summation order of duplicates is fixed here by insertion sequence. */
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

int or_csr_finalize(int64_t nrows, int64_t ncols, int64_t ntrip, const int64_t *trow,
                    const int64_t *tcol, const double *tval, int64_t *offsets, int64_t *cols,
                    double *vals, int64_t *nnz_out, or_lin_report *rep) {
    trip *t = malloc(sizeof(trip) * (ntrip > 0 ? ntrip : 1));
    for (int64_t i = 0; i < ntrip; ++i) {
        if (trow[i] < 0 || trow[i] >= nrows || tcol[i] < 0 || tcol[i] >= ncols) {
            rep->err_row = trow[i];
            rep->err_col = tcol[i];
            free(t);
            return set_err(rep, OR_ERR_INDEX_OOB, "matrix index is out of bounds");
        }
        if (!isfinite(tval[i])) {
            free(t);
            return nonfinite(rep, "matrix", i);
        }
        t[i].row = trow[i];
        t[i].col = tcol[i];
        t[i].val = tval[i];
        t[i].seq = i;
    }
    qsort(t, ntrip, sizeof(trip), trip_cmp);
    int64_t e = 0, nnz = 0;
    offsets[0] = 0;
    for (int64_t r = 0; r < nrows; ++r) {
        while (e < ntrip && t[e].row == r) {
            int64_t c = t[e].col;
            double v = 0.0;
            while (e < ntrip && t[e].row == r && t[e].col == c) v += t[e++].val;
            if (!isfinite(v)) {
                free(t);
                return nonfinite(rep, "assembled matrix", nnz);
            }
            if (v != 0.0) {
                cols[nnz] = c;
                vals[nnz] = v;
                nnz++;
            }
        }
        offsets[r + 1] = nnz;
    }
    free(t);
    *nnz_out = nnz;
    return OR_OK;
}

int or_jacobi_inverse(int64_t n, const int64_t *offsets, const int64_t *cols,
                      const double *vals, double *inv, or_lin_report *rep) {
    for (int64_t r = 0; r < n; ++r) {
        int64_t found = -1;
        for (int64_t e = offsets[r]; e < offsets[r + 1]; ++e)
            if (cols[e] == r) found = e;
        if (found < 0) {
            rep->err_row = r;
            return set_err(rep, OR_ERR_MISSING_DIAGONAL, "missing diagonal entry");
        }
        double d = vals[found];
        if (!isfinite(d)) return nonfinite(rep, "diagonal", r);
        if (d == 0.0) {
            rep->err_row = r;
            return set_err(rep, OR_ERR_ZERO_DIAGONAL, "zero diagonal entry");
        }
        double iv = 1.0 / d;
        if (!isfinite(iv)) return nonfinite(rep, "inverse diagonal", r);
        inv[r] = iv;
    }
    return OR_OK;
}

static int vdot(const double *a, const double *b, int64_t n, double *out, or_lin_report *rep) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    if (!isfinite(s)) return nonfinite(rep, "dot product", 0);
    *out = s;
    return OR_OK;
}
static int vnorm(const double *a, int64_t n, double *out, or_lin_report *rep) {
    double s;
    int st = vdot(a, a, n, &s, rep);
    if (st) return st;
    *out = sqrt(s);
    if (!isfinite(*out)) return nonfinite(rep, "L2 norm", 0);
    return OR_OK;
}
static int residual_into(int64_t n, const int64_t *o, const int64_t *c, const double *v,
                         const double *x, const double *rhs, double *out, or_lin_report *rep) {
    int st = or_csr_spmv(n, o, c, v, x, out, rep);
    if (st) return st;
    for (int64_t i = 0; i < n; ++i) {
        if (!isfinite(rhs[i])) return nonfinite(rep, "right-hand side", i);
        out[i] = rhs[i] - out[i];
        if (!isfinite(out[i])) return nonfinite(rep, "residual", i);
    }
    return OR_OK;
}
static int validate_inputs(int64_t n, const double *rhs, const double *x, double atol,
                           double rtol, or_lin_report *rep) {
    if (!isfinite(atol) || atol < 0 || !isfinite(rtol) || rtol < 0)
        return set_err(rep, OR_ERR_BREAKDOWN, "invalid solver tolerance");
    for (int64_t i = 0; i < n; ++i)
        if (!isfinite(rhs[i])) return nonfinite(rep, "solver vector", i);
    for (int64_t i = 0; i < n; ++i)
        if (!isfinite(x[i])) return nonfinite(rep, "solver vector", n + i);
    return OR_OK;
}
static void finish(or_lin_report *rep, int converged, int64_t it, double init, double fin) {
    rep->status = OR_OK;
    rep->converged = converged;
    rep->iterations = it;
    rep->initial_residual = init;
    rep->final_residual = fin;
    rep->relative_residual = init == 0.0 ? 0.0 : fin / init;
}

#define TRY(expr)                \
    do {                         \
        int st_ = (expr);        \
        if (st_) goto fail;      \
    } while (0)

int or_cg(int64_t n, const int64_t *o, const int64_t *c, const double *v, const double *rhs,
          double *x, double atol, double rtol, int64_t max_iter, int32_t jacobi,
          or_lin_report *rep) {
    memset(rep, 0, sizeof(*rep));
    int st = validate_inputs(n, rhs, x, atol, rtol, rep);
    if (st) return st;
    double *r = calloc(n, 8), *d = calloc(n, 8), *q = calloc(n, 8), *z = calloc(n, 8),
           *inv = jacobi ? calloc(n, 8) : NULL;
    if (jacobi) TRY(or_jacobi_inverse(n, o, c, v, inv, rep));
    TRY(residual_into(n, o, c, v, x, rhs, r, rep));
    double initial, rn;
    TRY(vnorm(r, n, &initial, rep));
    double thresh = atol + rtol * initial;
    if (initial <= thresh) {
        finish(rep, 1, 0, initial, initial);
        goto done;
    }
    if (jacobi) {
        for (int64_t i = 0; i < n; ++i) {
            z[i] = r[i] * inv[i];
            if (!isfinite(z[i])) {
                nonfinite(rep, "preconditioned residual", i);
                goto fail;
            }
        }
    } else {
        memcpy(z, r, 8 * n);
    }
    memcpy(d, z, 8 * n);
    double rho;
    TRY(vdot(r, z, n, &rho, rep));
    for (int64_t it = 1; it <= max_iter; ++it) {
        TRY(or_csr_spmv(n, o, c, v, d, q, rep));
        double den;
        TRY(vdot(d, q, n, &den, rep));
        if (!isfinite(den) || fabs(den) <= BREAKDOWN_EPSILON) {
            set_err(rep, OR_ERR_BREAKDOWN, "zero CG search denominator");
            goto fail;
        }
        double alpha = rho / den;
        if (!isfinite(alpha)) {
            set_err(rep, OR_ERR_BREAKDOWN, "non-finite CG step");
            goto fail;
        }
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * d[i];
            r[i] -= alpha * q[i];
        }
        TRY(vnorm(r, n, &rn, rep));
        if (rn <= thresh) {
            finish(rep, 1, it, initial, rn);
            goto done;
        }
        if (jacobi) {
            for (int64_t i = 0; i < n; ++i) {
                z[i] = r[i] * inv[i];
                if (!isfinite(z[i])) {
                    nonfinite(rep, "preconditioned residual", i);
                    goto fail;
                }
            }
        } else {
            memcpy(z, r, 8 * n);
        }
        double rho_new;
        TRY(vdot(r, z, n, &rho_new, rep));
        if (!isfinite(rho_new) || fabs(rho) <= BREAKDOWN_EPSILON) {
            set_err(rep, OR_ERR_BREAKDOWN, "zero CG recurrence denominator");
            goto fail;
        }
        double beta = rho_new / rho;
        if (!isfinite(beta)) {
            set_err(rep, OR_ERR_BREAKDOWN, "non-finite CG recurrence");
            goto fail;
        }
        for (int64_t i = 0; i < n; ++i) d[i] = z[i] + beta * d[i];
        rho = rho_new;
    }
    TRY(vnorm(r, n, &rn, rep));
    finish(rep, 0, max_iter, initial, rn);
done:
    free(r); free(d); free(q); free(z); free(inv);
    return OR_OK;
fail:
    free(r); free(d); free(q); free(z); free(inv);
    return rep->status;
}

int or_bicgstab(int64_t n, const int64_t *o, const int64_t *c, const double *v,
                const double *rhs, double *x, double atol, double rtol, int64_t max_iter,
                or_lin_report *rep) {
    memset(rep, 0, sizeof(*rep));
    int st = validate_inputs(n, rhs, x, atol, rtol, rep);
    if (st) return st;
    double *r = calloc(n, 8), *rh = calloc(n, 8), *p = calloc(n, 8), *vv = calloc(n, 8),
           *s = calloc(n, 8), *t = calloc(n, 8);
    TRY(residual_into(n, o, c, v, x, rhs, r, rep));
    memcpy(rh, r, 8 * n);
    double initial, rn;
    TRY(vnorm(r, n, &initial, rep));
    double thresh = atol + rtol * initial;
    if (initial <= thresh) {
        finish(rep, 1, 0, initial, initial);
        goto done;
    }
    double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
    for (int64_t it = 1; it <= max_iter; ++it) {
        double rho;
        TRY(vdot(rh, r, n, &rho, rep));
        if (!isfinite(rho) || fabs(rho) <= BREAKDOWN_EPSILON) {
            set_err(rep, OR_ERR_BREAKDOWN, "zero BiCGSTAB rho");
            goto fail;
        }
        double beta = (rho / rho_prev) * (alpha / omega);
        if (!isfinite(beta)) {
            set_err(rep, OR_ERR_BREAKDOWN, "non-finite BiCGSTAB beta");
            goto fail;
        }
        for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * vv[i]);
        TRY(or_csr_spmv(n, o, c, v, p, vv, rep));
        double den;
        TRY(vdot(rh, vv, n, &den, rep));
        if (!isfinite(den) || fabs(den) <= BREAKDOWN_EPSILON) {
            set_err(rep, OR_ERR_BREAKDOWN, "zero BiCGSTAB alpha denominator");
            goto fail;
        }
        alpha = rho / den;
        if (!isfinite(alpha)) {
            set_err(rep, OR_ERR_BREAKDOWN, "non-finite BiCGSTAB alpha");
            goto fail;
        }
        for (int64_t i = 0; i < n; ++i) s[i] = r[i] - alpha * vv[i];
        double sn;
        TRY(vnorm(s, n, &sn, rep));
        if (sn <= thresh) {
            for (int64_t i = 0; i < n; ++i) x[i] += alpha * p[i];
            finish(rep, 1, it, initial, sn);
            goto done;
        }
        TRY(or_csr_spmv(n, o, c, v, s, t, rep));
        double tt, ts;
        TRY(vdot(t, t, n, &tt, rep));
        if (!isfinite(tt) || fabs(tt) <= BREAKDOWN_EPSILON) {
            set_err(rep, OR_ERR_BREAKDOWN, "zero BiCGSTAB omega denominator");
            goto fail;
        }
        TRY(vdot(t, s, n, &ts, rep));
        omega = ts / tt;
        if (!isfinite(omega) || fabs(omega) <= BREAKDOWN_EPSILON) {
            set_err(rep, OR_ERR_BREAKDOWN, "zero BiCGSTAB omega");
            goto fail;
        }
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i] + omega * s[i];
            r[i] = s[i] - omega * t[i];
        }
        TRY(vnorm(r, n, &rn, rep));
        if (rn <= thresh) {
            finish(rep, 1, it, initial, rn);
            goto done;
        }
        rho_prev = rho;
    }
    TRY(vnorm(r, n, &rn, rep));
    finish(rep, 0, max_iter, initial, rn);
done:
    free(r); free(rh); free(p); free(vv); free(s); free(t);
    return OR_OK;
fail:
    free(r); free(rh); free(p); free(vv); free(s); free(t);
    return rep->status;
}

/* ---- deterministic PRNG (shared with the GPU-side input generators) ---- */
uint64_t or_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double or_uniform(uint64_t *state) {
    return (double)(or_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}
double or_normal(uint64_t *state) {
    double u1 = or_uniform(state), u2 = or_uniform(state);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void or_free(void *ptr) { free(ptr); }
