/*
 * Oracle: the jittered structured triangulation of the unit square and its
 * pinned two-point Poisson matrix (FP64) -- the unstructured pressure system
 * of the CSR benchmark configuration.
 * TEST INFRASTRUCTURE ONLY -- produces the CSR systems the GPU solvers are
 * checked on; it is never part of the shipped path.
 *
 * The points and triangles are built here; topology, geometry and assembly
 * are the general restatements in fvmesh.c (UnstructuredMesh::from_cells,
 * src/mesh.rs:317-423; assemble_poisson with ZeroGradient boundaries and
 * Gamma = 1, src/poisson.rs:146-263; apply_reference, :303-337).
 * Triangles per square (a, b): T0 = (v00, v10, v11), T1 = (v00, v11, v01);
 * cell index 2*(a + m*b) + t.  Interior vertices are jittered by
 * U(-jitter*h, +jitter*h) in x then y, vertices visited row by row, from a
 * splitmix64 stream seeded with `seed`.
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int or_mesh_poisson(int64_t m, double jitter, uint64_t seed, int64_t ref_cell,
                    int64_t *ncells_out, int64_t *nnz_out, int64_t **offsets_out,
                    int64_t **cols_out, double **vals_out, double **volumes_out,
                    int64_t *nfaces_out, int64_t **face_owner_out, int64_t **face_neigh_out,
                    double **face_coef_out, char *message, int cap) {
    const int64_t nv = (m + 1) * (m + 1), nc = 2 * m * m;
    const double h = 1.0 / (double)m;
    double *xy = malloc(sizeof(double) * 2 * nv);
    uint64_t st = seed;
    for (int64_t b = 0; b <= m; ++b)
        for (int64_t a = 0; a <= m; ++a) {
            double x = a * h, y = b * h;
            if (a > 0 && a < m && b > 0 && b < m) {
                x += (2.0 * or_uniform(&st) - 1.0) * jitter * h;
                y += (2.0 * or_uniform(&st) - 1.0) * jitter * h;
            }
            xy[2 * (a + (m + 1) * b)] = x;
            xy[2 * (a + (m + 1) * b) + 1] = y;
        }
    int64_t *off = malloc(sizeof(int64_t) * (nc + 1)), *vtx = malloc(sizeof(int64_t) * 3 * nc);
    for (int64_t b = 0; b < m; ++b)
        for (int64_t a = 0; a < m; ++a) {
            const int64_t v00 = a + (m + 1) * b, v10 = v00 + 1, v11 = v00 + m + 2, v01 = v00 + m + 1;
            const int64_t c0 = 2 * (a + m * b);
            vtx[3 * c0] = v00; vtx[3 * c0 + 1] = v10; vtx[3 * c0 + 2] = v11;
            vtx[3 * c0 + 3] = v00; vtx[3 * c0 + 4] = v11; vtx[3 * c0 + 5] = v01;
        }
    for (int64_t c = 0; c <= nc; ++c) off[c] = 3 * c;
    or_mesh2d mesh;
    int rc = or_mesh2d_from_cells(nv, xy, nc, off, vtx, &mesh, message, cap);
    free(xy);
    free(off);
    free(vtx);
    if (rc) return rc;
    int32_t *kind = calloc(mesh.nfaces, sizeof(int32_t)); /* all ZeroGradient */
    double *rhs = NULL;
    rc = or_assemble_poisson(&mesh, 1.0, NULL, kind, NULL, 1, ref_cell, 0.0, nnz_out, offsets_out,
                             cols_out, vals_out, &rhs, NULL, message, cap);
    free(kind);
    free(rhs);
    if (rc == OR_OK) {
        /* interior faces with their two-point coefficients (poisson.rs:193) */
        int64_t nint = 0;
        for (int64_t f = 0; f < mesh.nfaces; ++f) nint += mesh.neigh[f] >= 0;
        int64_t *fo = malloc(sizeof(int64_t) * (nint ? nint : 1));
        int64_t *fn = malloc(sizeof(int64_t) * (nint ? nint : 1));
        double *fc = malloc(sizeof(double) * (nint ? nint : 1));
        int64_t k = 0;
        for (int64_t f = 0; f < mesh.nfaces; ++f) {
            const int64_t P = mesh.owner[f], N = mesh.neigh[f];
            if (N < 0) continue;
            const double dx = mesh.cx[N] - mesh.cx[P], dy = mesh.cy[N] - mesh.cy[P];
            fo[k] = P;
            fn[k] = N;
            fc[k] = 1.0 * (mesh.sx[f] * dx + mesh.sy[f] * dy) / (dx * dx + dy * dy);
            k++;
        }
        *ncells_out = nc;
        *volumes_out = mesh.vol;
        mesh.vol = NULL;
        *nfaces_out = nint;
        *face_owner_out = fo;
        *face_neigh_out = fn;
        *face_coef_out = fc;
    }
    or_mesh2d_free(&mesh);
    return rc;
}
