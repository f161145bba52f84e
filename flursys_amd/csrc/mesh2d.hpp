// The device 2D polygon mesh shared by the assembly (assemble.hip) and the
// SIMPLE outer iteration (simple2d.hip).
#pragma once

#include "common.hpp"

struct fl_mesh2d {
    fl_ctx *ctx = nullptr;
    int64_t npts = 0, ncells = 0, nfaces = 0, nemit = 0;
    double *cx = nullptr, *cy = nullptr, *vol = nullptr;  // cell centroid and area
    int *owner = nullptr, *neigh = nullptr;               // neigh = -1 on boundary faces
    double *fcx = nullptr, *fcy = nullptr, *sx = nullptr, *sy = nullptr;  // face centre, area vector
    // cell c's edges are emissions off[c] .. off[c+1]-1 in polygon order;
    // emit_face maps each to its face (Cell::faces, src/mesh.rs:353-377)
    int *off = nullptr, *emit_face = nullptr;
    void release() {
        void *b[] = {cx, cy, vol, owner, neigh, fcx, fcy, sx, sy, off, emit_face};
        for (void *p : b)
            if (p) (void)hipFree(p);
    }
};
