/*
 * Oracle: jittered structured triangulation of the unit square and the
 * two-point finite-volume Poisson matrix on it (FP64).
 * TEST INFRASTRUCTURE ONLY -- produces the unstructured CSR systems the GPU
 * solvers are checked on; it is never part of the shipped path.
 *
 * Restates:
 *   UnstructuredMesh::from_cells     src/mesh.rs:317-423 (face creation order,
 *                                    owner = first cell, neighbour = second)
 *   oriented_face / face_area_vector src/mesh.rs:540-582 (2D edge normal
 *                                    (e.y, -e.x), flipped to point away from
 *                                    the owner's vertex average)
 *   cell_geometry (TwoD)             src/mesh.rs:584-609 (shoelace centroid)
 *   weight                           src/numerics.rs:294-303
 *   interpolate_diffusivity (Linear) src/numerics.rs:383-428 with Gamma = 1
 *   assemble_poisson                 src/poisson.rs:146-263 (ZeroGradient
 *                                    boundaries: no boundary entries)
 *   apply_reference                  src/poisson.rs:303-337
 *   CsrBuilder::finalize             src/linear/csr.rs:159-192
 * Triangles per square (a, b): T0 = (v00, v10, v11), T1 = (v00, v11, v01);
 * cell index 2*(a + m*b) + t.  Interior vertices are jittered by
 * U(-jitter*h, +jitter*h) in x then y, vertices visited row by row, from a
 * splitmix64 stream seeded with `seed`.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GEOMETRY_EPSILON 1.0e-14

typedef struct { double x, y; } vec2;

static inline int64_t hedge(int64_t m, int64_t a, int64_t b) { return a + m * b; }             /* (a,b)-(a+1,b) */
static inline int64_t vedge(int64_t m, int64_t a, int64_t b) { return m * (m + 1) + a + (m + 1) * b; } /* (a,b)-(a,b+1) */
static inline int64_t dedge(int64_t m, int64_t a, int64_t b) { return 2 * m * (m + 1) + a + m * b; }   /* (a,b)-(a+1,b+1) */

int or_mesh_poisson(int64_t m, double jitter, uint64_t seed, int64_t ref_cell,
                    int64_t *ncells_out, int64_t *nnz_out, int64_t **offsets_out,
                    int64_t **cols_out, double **vals_out, double **volumes_out,
                    int64_t *nfaces_out, int64_t **face_owner_out, int64_t **face_neigh_out,
                    double **face_coef_out, char *message, int cap) {
    const int64_t nv = (m + 1) * (m + 1), nc = 2 * m * m, nedge = 3 * m * m + 2 * m;
    const double h = 1.0 / (double)m;
    vec2 *pt = malloc(sizeof(vec2) * nv);
    uint64_t st = seed;
    for (int64_t b = 0; b <= m; ++b)
        for (int64_t a = 0; a <= m; ++a) {
            vec2 p = {a * h, b * h};
            if (a > 0 && a < m && b > 0 && b < m) {
                p.x += (2.0 * or_uniform(&st) - 1.0) * jitter * h;
                p.y += (2.0 * or_uniform(&st) - 1.0) * jitter * h;
            }
            pt[a + (m + 1) * b] = p;
        }
    vec2 *cc = malloc(sizeof(vec2) * nc), *cref = malloc(sizeof(vec2) * nc);
    double *vol = malloc(sizeof(double) * nc);
    int64_t *tri = malloc(sizeof(int64_t) * 3 * nc);
    int64_t *tedge = malloc(sizeof(int64_t) * 3 * nc);
    for (int64_t b = 0; b < m; ++b)
        for (int64_t a = 0; a < m; ++a) {
            int64_t v00 = a + (m + 1) * b, v10 = v00 + 1, v11 = v00 + m + 2, v01 = v00 + m + 1;
            int64_t c0 = 2 * (a + m * b), c1 = c0 + 1;
            tri[3 * c0] = v00; tri[3 * c0 + 1] = v10; tri[3 * c0 + 2] = v11;
            tedge[3 * c0] = hedge(m, a, b); tedge[3 * c0 + 1] = vedge(m, a + 1, b); tedge[3 * c0 + 2] = dedge(m, a, b);
            tri[3 * c1] = v00; tri[3 * c1 + 1] = v11; tri[3 * c1 + 2] = v01;
            tedge[3 * c1] = dedge(m, a, b); tedge[3 * c1 + 1] = hedge(m, a, b + 1); tedge[3 * c1 + 2] = vedge(m, a, b);
        }
    char err[128] = {0};
    for (int64_t c = 0; c < nc; ++c) {
        vec2 P[3];
        for (int t = 0; t < 3; ++t) P[t] = pt[tri[3 * c + t]];
        /* vertex average (from_cells reference point) */
        vec2 r = {(P[0].x + P[1].x + P[2].x) / 3.0, (P[0].y + P[1].y + P[2].y) / 3.0};
        double twice = 0.0, cx = 0.0, cy = 0.0;
        for (int t = 0; t < 3; ++t) {
            vec2 A = P[t], B = P[(t + 1) % 3];
            double cr = A.x * B.y - B.x * A.y;
            twice += cr;
            cx += (A.x + B.x) * cr;
            cy += (A.y + B.y) * cr;
        }
        double signed_area = twice * 0.5;
        if (!(fabs(signed_area) > GEOMETRY_EPSILON)) {
            snprintf(err, sizeof err, "zero-volume cell %lld", (long long)c);
            break;
        }
        cc[c].x = cx / (3.0 * twice);
        cc[c].y = cy / (3.0 * twice);
        cref[c] = r;
        vol[c] = fabs(signed_area);
    }
    int64_t *face_of_edge = malloc(sizeof(int64_t) * nedge);
    for (int64_t e = 0; e < nedge; ++e) face_of_edge[e] = -1;
    int64_t *fo = malloc(sizeof(int64_t) * nedge), *fn = malloc(sizeof(int64_t) * nedge);
    vec2 *fc = malloc(sizeof(vec2) * nedge), *fs = malloc(sizeof(vec2) * nedge);
    int64_t nf = 0;
    for (int64_t c = 0; c < nc && !err[0]; ++c)
        for (int t = 0; t < 3; ++t) {
            int64_t e = tedge[3 * c + t];
            if (face_of_edge[e] >= 0) {
                fn[face_of_edge[e]] = c;
                continue;
            }
            vec2 A = pt[tri[3 * c + t]], B = pt[tri[3 * c + (t + 1) % 3]];
            vec2 ctr = {(A.x + B.x) / 2.0, (A.y + B.y) / 2.0};
            vec2 S = {B.y - A.y, -(B.x - A.x)};
            if ((S.x * (ctr.x - cref[c].x) + S.y * (ctr.y - cref[c].y)) < 0.0) {
                ctr.x = (B.x + A.x) / 2.0;
                ctr.y = (B.y + A.y) / 2.0;
                S.x = -S.x;
                S.y = -S.y;
            }
            face_of_edge[e] = nf;
            fo[nf] = c;
            fn[nf] = -1;
            fc[nf] = ctr;
            fs[nf] = S;
            nf++;
        }
    /* two-point coefficients, entries in face order (poisson.rs:183-206) */
    int64_t nint = 0;
    for (int64_t f = 0; f < nf; ++f) nint += fn[f] >= 0;
    int64_t ntrip = 4 * nint;
    int64_t *tr = malloc(sizeof(int64_t) * (ntrip + 1)), *tc = malloc(sizeof(int64_t) * (ntrip + 1));
    double *tv = malloc(sizeof(double) * (ntrip + 1));
    int64_t *ofo = malloc(sizeof(int64_t) * (nint ? nint : 1)), *ofn = malloc(sizeof(int64_t) * (nint ? nint : 1));
    double *ofc = malloc(sizeof(double) * (nint ? nint : 1));
    int64_t k = 0, kk = 0;
    for (int64_t f = 0; f < nf && !err[0]; ++f) {
        if (fn[f] < 0) continue;
        int64_t P = fo[f], N = fn[f];
        vec2 d = {cc[N].x - cc[P].x, cc[N].y - cc[P].y};
        double d2 = d.x * d.x + d.y * d.y;
        if (!isfinite(d2) || d2 <= 2.220446049250313e-16) {
            snprintf(err, sizeof err, "degenerate owner-neighbour distance at face %lld", (long long)f);
            break;
        }
        double lambda = ((fc[f].x - cc[P].x) * d.x + (fc[f].y - cc[P].y) * d.y) / d2;
        double gamma = (1.0 - lambda) * 1.0 + lambda * 1.0;
        double coef = gamma * (fs[f].x * d.x + fs[f].y * d.y) / d2;
        if (!isfinite(coef) || coef <= 0.0) {
            snprintf(err, sizeof err, "invalid non-orthogonal geometry at face %lld", (long long)f);
            break;
        }
        ofo[kk] = P; ofn[kk] = N; ofc[kk] = coef; kk++;
        tr[k] = P; tc[k] = P; tv[k++] = coef;
        tr[k] = P; tc[k] = N; tv[k++] = -coef;
        tr[k] = N; tc[k] = N; tv[k++] = coef;
        tr[k] = N; tc[k] = P; tv[k++] = -coef;
    }
    int rc = OR_OK;
    if (!err[0]) {
        /* apply_reference: drop the pinned row, drop its column (value 0 moves
         * nothing to the rhs), add the unit diagonal. */
        int64_t w = 0;
        for (int64_t i = 0; i < k; ++i) {
            if (tr[i] == ref_cell || tc[i] == ref_cell) continue;
            tr[w] = tr[i]; tc[w] = tc[i]; tv[w] = tv[i]; w++;
        }
        tr[w] = ref_cell; tc[w] = ref_cell; tv[w] = 1.0; w++;
        int64_t *off = malloc(sizeof(int64_t) * (nc + 1)), *col = malloc(sizeof(int64_t) * w);
        double *val = malloc(sizeof(double) * w);
        int64_t nnz = 0;
        or_lin_report rep;
        memset(&rep, 0, sizeof rep);
        rc = or_csr_finalize(nc, nc, w, tr, tc, tv, off, col, val, &nnz, &rep);
        if (rc) {
            snprintf(err, sizeof err, "CSR finalize failed: %s", rep.reason);
            free(off); free(col); free(val);
        } else {
            *ncells_out = nc;
            *nnz_out = nnz;
            *offsets_out = off;
            *cols_out = col;
            *vals_out = val;
            *volumes_out = vol;
            vol = NULL;
            *nfaces_out = kk;
            *face_owner_out = ofo;
            *face_neigh_out = ofn;
            *face_coef_out = ofc;
            ofo = ofn = NULL;
            ofc = NULL;
        }
    }
    if (err[0]) {
        rc = rc ? rc : OR_ERR_CONFIG;
        if (message) snprintf(message, cap, "%s", err);
    }
    free(pt); free(cc); free(cref); free(vol); free(tri); free(tedge); free(face_of_edge);
    free(fo); free(fn); free(fc); free(fs); free(tr); free(tc); free(tv);
    free(ofo); free(ofn); free(ofc);
    return rc;
}
