/*
 * Oracle: 2D unstructured finite-volume mesh topology/geometry and the
 * two-point Poisson assembly on it (FP64).
 * TEST INFRASTRUCTURE ONLY -- produces the CSR pressure systems the GPU
 * solvers are checked on and pins them against the reference's own
 * assembly tests; it is never part of the shipped path.
 *
 * Restates:
 *   CellDefinition::faces (TwoD polygon)  src/mesh.rs:147-155  (edges (v_i, v_i+1))
 *   UnstructuredMesh::from_cells          src/mesh.rs:317-423  (faces in creation
 *                                         order, keyed by the sorted vertex pair;
 *                                         owner = first cell, neighbour = second;
 *                                         a third cell is NonManifoldFace)
 *   average_point / oriented_face /       src/mesh.rs:533-582  (2D edge normal
 *   face_area_vector                      (e.y, -e.x), reversed when it points
 *                                         towards the owner's vertex average)
 *   cell_geometry (TwoD)                  src/mesh.rs:584-609  (shoelace centroid)
 *   orientation checks                    src/mesh.rs:396-408
 *   assemble_poisson                      src/poisson.rs:146-263 (constant Gamma,
 *                                         interpolated Linear to interior faces,
 *                                         (1 - lambda) G + lambda G, as
 *                                         interpolate_diffusivity does; the owner
 *                                         value G on boundary faces)
 *   apply_reference                       src/poisson.rs:303-337
 *   CsrBuilder::finalize                  src/linear/csr.rs:159-192 (or_csr_finalize)
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GEOMETRY_EPSILON 1.0e-14
#define F64_EPSILON 2.220446049250313e-16

static int fail(char *message, int cap, int code, const char *fmt, long long a) {
    if (message) snprintf(message, cap, fmt, a);
    return code;
}

void or_mesh2d_free(or_mesh2d *m) {
    void *f[] = {m->cx, m->cy, m->vol, m->owner, m->neigh, m->fcx, m->fcy, m->sx, m->sy,
                 m->fv0, m->fv1, m->cell_off, m->cell_faces};
    for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); ++i) free(f[i]);
    memset(m, 0, sizeof(*m));
}

/* open-addressing table from the sorted vertex pair to the face index */
typedef struct {
    int64_t a, b, face;
} slot;

static uint64_t pair_hash(int64_t a, int64_t b) {
    uint64_t h = (uint64_t)a * 0x9E3779B97F4A7C15ull ^ ((uint64_t)b + 0x632BE59BD9B4E019ull);
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    return h ^ (h >> 29);
}

int or_mesh2d_from_cells(int64_t npts, const double *xy, int64_t ncells, const int64_t *cell_off,
                         const int64_t *cell_vtx, or_mesh2d *m, char *message, int cap) {
    memset(m, 0, sizeof(*m));
    if (ncells <= 0) return fail(message, cap, OR_MESH_EMPTY, "mesh has no cells%lld", 0);
    for (int64_t p = 0; p < npts; ++p)
        if (!isfinite(xy[2 * p]) || !isfinite(xy[2 * p + 1]))
            return fail(message, cap, OR_MESH_NON_FINITE_POINT, "non-finite point %lld", p);
    const int64_t nedge_max = cell_off[ncells];
    int64_t cap_slots = 16;
    while (cap_slots < 2 * nedge_max) cap_slots <<= 1;
    slot *tab = malloc(sizeof(slot) * cap_slots);
    for (int64_t i = 0; i < cap_slots; ++i) tab[i].face = -1;
    m->cx = malloc(sizeof(double) * ncells);
    m->cy = malloc(sizeof(double) * ncells);
    m->vol = malloc(sizeof(double) * ncells);
    m->owner = malloc(sizeof(int64_t) * nedge_max);
    m->neigh = malloc(sizeof(int64_t) * nedge_max);
    m->fcx = malloc(sizeof(double) * nedge_max);
    m->fcy = malloc(sizeof(double) * nedge_max);
    m->sx = malloc(sizeof(double) * nedge_max);
    m->sy = malloc(sizeof(double) * nedge_max);
    m->fv0 = malloc(sizeof(int64_t) * nedge_max);
    m->fv1 = malloc(sizeof(int64_t) * nedge_max);
    m->cell_off = malloc(sizeof(int64_t) * (ncells + 1));
    m->cell_faces = malloc(sizeof(int64_t) * (nedge_max ? nedge_max : 1));
    memcpy(m->cell_off, cell_off, sizeof(int64_t) * (ncells + 1));
    m->ncells = ncells;
    int rc = OR_OK;
    int64_t nf = 0;
    for (int64_t c = 0; c < ncells && rc == OR_OK; ++c) {
        const int64_t o = cell_off[c], nvtx = cell_off[c + 1] - o;
        if (nvtx < 3) {
            rc = fail(message, cap, OR_MESH_INVALID_CELL,
                      "a polygon requires at least three vertices (cell %lld)", c);
            break;
        }
        /* unique_vertices + average_point (the vertex average is the orientation
         * reference; a polygon's vertices are distinct) */
        double rx = 0.0, ry = 0.0;
        for (int64_t t = 0; t < nvtx; ++t) {
            const int64_t p = cell_vtx[o + t];
            if (p < 0 || p >= npts) {
                rc = fail(message, cap, OR_MESH_INVALID_POINT_INDEX, "invalid point index in cell %lld", c);
                break;
            }
            rx += xy[2 * p];
            ry += xy[2 * p + 1];
        }
        if (rc) break;
        rx /= (double)nvtx;
        ry /= (double)nvtx;
        for (int64_t t = 0; t < nvtx; ++t) {
            int64_t a = cell_vtx[o + t], b = cell_vtx[o + (t + 1) % nvtx];
            double ccx = (xy[2 * a] + xy[2 * b]) / 2.0, ccy = (xy[2 * a + 1] + xy[2 * b + 1]) / 2.0;
            double sx = xy[2 * b + 1] - xy[2 * a + 1], sy = -(xy[2 * b] - xy[2 * a]);
            if (!(sqrt(sx * sx + sy * sy) > GEOMETRY_EPSILON)) {
                rc = fail(message, cap, OR_MESH_DEGENERATE_FACE, "degenerate face in cell %lld", c);
                break;
            }
            if (sx * (ccx - rx) + sy * (ccy - ry) < 0.0) {
                const int64_t tmp = a;
                a = b;
                b = tmp;
                ccx = (xy[2 * a] + xy[2 * b]) / 2.0;
                ccy = (xy[2 * a + 1] + xy[2 * b + 1]) / 2.0;
                sx = xy[2 * b + 1] - xy[2 * a + 1];
                sy = -(xy[2 * b] - xy[2 * a]);
            }
            const int64_t ka = a < b ? a : b, kb = a < b ? b : a;
            uint64_t h = pair_hash(ka, kb) & (uint64_t)(cap_slots - 1);
            while (tab[h].face >= 0 && !(tab[h].a == ka && tab[h].b == kb))
                h = (h + 1) & (uint64_t)(cap_slots - 1);
            if (tab[h].face >= 0) {
                const int64_t f = tab[h].face;
                if (m->neigh[f] >= 0) {
                    rc = fail(message, cap, OR_MESH_NON_MANIFOLD, "non-manifold face at cell %lld", c);
                    break;
                }
                m->neigh[f] = c;
                m->cell_faces[o + t] = f;
            } else {
                tab[h].a = ka;
                tab[h].b = kb;
                tab[h].face = nf;
                m->owner[nf] = c;
                m->neigh[nf] = -1;
                m->fcx[nf] = ccx;
                m->fcy[nf] = ccy;
                m->sx[nf] = sx;
                m->sy[nf] = sy;
                m->fv0[nf] = a;
                m->fv1[nf] = b;
                m->cell_faces[o + t] = nf;
                nf++;
            }
        }
    }
    free(tab);
    m->nfaces = nf;
    for (int64_t c = 0; c < ncells && rc == OR_OK; ++c) {
        /* cell_geometry (TwoD): the polygon's vertices in definition order */
        const int64_t o = cell_off[c], nvtx = cell_off[c + 1] - o;
        double twice = 0.0, cx = 0.0, cy = 0.0;
        for (int64_t t = 0; t < nvtx; ++t) {
            const int64_t a = cell_vtx[o + t], b = cell_vtx[o + (t + 1) % nvtx];
            const double ax = xy[2 * a], ay = xy[2 * a + 1], bx = xy[2 * b], by = xy[2 * b + 1];
            const double cr = ax * by - bx * ay;
            twice += cr;
            cx += (ax + bx) * cr;
            cy += (ay + by) * cr;
        }
        const double signed_area = twice * 0.5;
        if (!(fabs(signed_area) > GEOMETRY_EPSILON)) {
            rc = fail(message, cap, OR_MESH_ZERO_VOLUME, "zero-volume cell %lld", c);
            break;
        }
        m->cx[c] = cx / (3.0 * twice);
        m->cy[c] = cy / (3.0 * twice);
        m->vol[c] = fabs(signed_area);
    }
    for (int64_t f = 0; f < nf && rc == OR_OK; ++f) {
        const double area = sqrt(m->sx[f] * m->sx[f] + m->sy[f] * m->sy[f]);
        const double nx = m->sx[f] / area, ny = m->sy[f] / area;
        const int64_t P = m->owner[f], N = m->neigh[f];
        if ((m->fcx[f] - m->cx[P]) * nx + (m->fcy[f] - m->cy[P]) * ny <= GEOMETRY_EPSILON ||
            (N >= 0 && (m->cx[N] - m->fcx[f]) * nx + (m->cy[N] - m->fcy[f]) * ny <= GEOMETRY_EPSILON))
            rc = fail(message, cap, OR_MESH_INCONSISTENT_ORIENTATION,
                      "face %lld has inconsistent owner/neighbour orientation", f);
    }
    if (rc) or_mesh2d_free(m);
    return rc;
}

int or_assemble_poisson(const or_mesh2d *m, double gamma, const double *source,
                        const int32_t *bc_kind, const double *bc_value, int32_t has_reference,
                        int64_t ref_cell, double ref_value, int64_t *nnz_out,
                        int64_t **offsets_out, int64_t **cols_out, double **vals_out,
                        double **rhs_out, double *detail, char *message, int cap) {
    const int64_t nc = m->ncells, nf = m->nfaces;
    if (!isfinite(gamma))
        return fail(message, cap, OR_POISSON_NUMERICS, "non-finite diffusivity%lld", 0);
    double *rhs = calloc(nc, sizeof(double));
    const int64_t tcap = 4 * nf + 1;
    int64_t *tr = malloc(sizeof(int64_t) * tcap), *tc = malloc(sizeof(int64_t) * tcap);
    double *tv = malloc(sizeof(double) * tcap);
    int64_t k = 0;
    int rc = OR_OK, has_fixed = 0;
    /* source * volume (poisson.rs:171-182) */
    for (int64_t c = 0; c < nc && source; ++c) {
        if (!isfinite(source[c])) {
            rc = fail(message, cap, OR_POISSON_NON_FINITE_SOURCE, "source value in cell %lld is not finite", c);
            break;
        }
        rhs[c] += source[c] * m->vol[c];
    }
    /* faces in mesh order (poisson.rs:184-246) */
    for (int64_t f = 0; f < nf && rc == OR_OK; ++f) {
        const int64_t P = m->owner[f], N = m->neigh[f];
        if (N >= 0) {
            const double dx = m->cx[N] - m->cx[P], dy = m->cy[N] - m->cy[P];
            const double d2 = dx * dx + dy * dy;
            if (!isfinite(d2) || d2 <= F64_EPSILON) {
                rc = fail(message, cap, OR_POISSON_DEGENERATE_DISTANCE,
                          "degenerate owner-neighbour distance at face %lld", f);
                break;
            }
            /* interpolate_diffusivity, Linear (numerics.rs:393-412): the
             * constant interpolated with the face weight (numerics.rs:294-303) */
            const double lambda = ((m->fcx[f] - m->cx[P]) * dx + (m->fcy[f] - m->cy[P]) * dy) / d2;
            const double gamma_f = (1.0 - lambda) * gamma + lambda * gamma;
            const double coef = gamma_f * (m->sx[f] * dx + m->sy[f] * dy) / d2;
            if (!isfinite(coef) || coef <= 0.0) {
                rc = fail(message, cap, OR_POISSON_INVALID_GEOMETRY,
                          "invalid non-orthogonal geometry at face %lld", f);
                break;
            }
            tr[k] = P; tc[k] = P; tv[k++] = coef;
            tr[k] = P; tc[k] = N; tv[k++] = -coef;
            tr[k] = N; tc[k] = N; tv[k++] = coef;
            tr[k] = N; tc[k] = P; tv[k++] = -coef;
        } else if (bc_kind && bc_kind[f] == OR_SCALAR_FIXED_VALUE) {
            const double value = bc_value[f];
            if (!isfinite(value)) {
                rc = fail(message, cap, OR_POISSON_NON_FINITE_SOURCE, "boundary value at cell %lld is not finite", P);
                break;
            }
            has_fixed = 1;
            const double dx = m->fcx[f] - m->cx[P], dy = m->fcy[f] - m->cy[P];
            const double d2 = dx * dx + dy * dy;
            if (!isfinite(d2) || d2 <= F64_EPSILON) {
                rc = fail(message, cap, OR_POISSON_DEGENERATE_FACE_DISTANCE,
                          "degenerate owner-face distance at face %lld", f);
                break;
            }
            const double coef = gamma * (m->sx[f] * dx + m->sy[f] * dy) / d2;
            if (!isfinite(coef) || coef <= 0.0) {
                rc = fail(message, cap, OR_POISSON_INVALID_GEOMETRY,
                          "invalid non-orthogonal geometry at face %lld", f);
                break;
            }
            tr[k] = P; tc[k] = P; tv[k++] = coef;
            rhs[P] += coef * value;
        } else if (!bc_kind || bc_kind[f] != OR_SCALAR_ZERO_GRADIENT) {
            rc = fail(message, cap, OR_POISSON_MISSING_BOUNDARY, "missing boundary condition at face %lld", f);
            break;
        }
    }
    /* pure-Neumann compatibility and reference (poisson.rs:248-257) */
    if (rc == OR_OK && !has_fixed) {
        double integrated = 0.0, scale = 0.0;
        for (int64_t c = 0; c < nc; ++c) integrated += rhs[c];
        for (int64_t c = 0; c < nc; ++c) scale += fabs(rhs[c]);
        if (fabs(integrated) > 1.0e-12 * (scale > 1.0 ? scale : 1.0)) {
            if (detail) *detail = integrated;
            rc = fail(message, cap, OR_POISSON_INCOMPATIBLE_SOURCE,
                      "pure-Neumann source is incompatible%lld", 0);
        } else if (!has_reference) {
            rc = fail(message, cap, OR_POISSON_MISSING_REFERENCE,
                      "pure ZeroGradient Poisson systems require an explicit reference cell%lld", 0);
        }
    }
    /* apply_reference (poisson.rs:303-337): drop the pinned row, move the
     * pinned column to the right-hand side, add the unit diagonal */
    if (rc == OR_OK && has_reference) {
        if (ref_cell < 0 || ref_cell >= nc) {
            rc = fail(message, cap, OR_POISSON_INVALID_REFERENCE_CELL,
                      "reference cell %lld is outside the mesh", ref_cell);
        } else if (!isfinite(ref_value)) {
            rc = fail(message, cap, OR_POISSON_INVALID_REFERENCE_VALUE, "reference value is not finite%lld", 0);
        } else {
            int64_t w = 0;
            for (int64_t i = 0; i < k; ++i) {
                if (tr[i] == ref_cell) continue;
                if (tc[i] == ref_cell) {
                    rhs[tr[i]] -= tv[i] * ref_value;
                    continue;
                }
                tr[w] = tr[i]; tc[w] = tc[i]; tv[w] = tv[i]; w++;
            }
            rhs[ref_cell] = ref_value;
            tr[w] = ref_cell; tc[w] = ref_cell; tv[w] = 1.0; w++;
            k = w;
        }
    }
    if (rc == OR_OK) {
        int64_t *off = malloc(sizeof(int64_t) * (nc + 1)), *col = malloc(sizeof(int64_t) * (k ? k : 1));
        double *val = malloc(sizeof(double) * (k ? k : 1));
        or_lin_report rep;
        memset(&rep, 0, sizeof rep);
        int64_t nnz = 0;
        if (or_csr_finalize(nc, nc, k, tr, tc, tv, off, col, val, &nnz, &rep) != OR_OK) {
            rc = OR_POISSON_LINEAR;
            if (message) snprintf(message, cap, "CSR finalize failed: %s", rep.reason);
            free(off); free(col); free(val);
        } else {
            *nnz_out = nnz;
            *offsets_out = off;
            *cols_out = col;
            *vals_out = val;
            *rhs_out = rhs;
            rhs = NULL;
        }
    }
    free(tr); free(tc); free(tv); free(rhs);
    return rc;
}
