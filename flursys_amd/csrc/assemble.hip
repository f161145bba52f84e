// Device assembly of unstructured finite-volume systems: the 2D polygon mesh
// (UnstructuredMesh::from_cells, src/mesh.rs:317-423, 533-609), the two-point
// Poisson matrix (assemble_poisson, src/poisson.rs:146-263, with
// apply_reference :303-337) and its CSR finalisation (CsrBuilder::finalize,
// src/linear/csr.rs:159-192) -- the setup the reference runs serially on the
// host (a HashMap face lookup and a triplet sort) as sorts and scans over
// millions of edges and entries.
//
// Face numbering.  The reference creates a face at the first cell edge that
// names its (unordered) vertex pair, in cell order then edge order; the
// second edge with that pair makes the second cell its neighbour, a third is
// a non-manifold face.  Here every cell edge ("emission" e = cell_off[c] + t)
// is keyed by its sorted vertex pair; a stable radix sort groups the
// emissions of one face in emission order; the first emission of each group
// is flagged, and an exclusive scan of the flags over emission order numbers
// the faces exactly as the reference's creation order does.
//
// Entry order.  CsrBuilder::finalize sums duplicate (row, column) entries
// after an unstable sort, so the reference leaves their summation order
// open; here (as in the oracle) they are summed in insertion order: the
// triplets are keyed row * n + col and a stable radix sort keeps insertion
// order inside each key.  Sums that are exactly zero are dropped.
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "common.hpp"
#include "mesh2d.hpp"

// No fused multiply-adds in this file: the reference's geometry and
// coefficients (Rust, which never contracts) are products and sums rounded
// one at a time, and the assembled values must equal them bit for bit.
#pragma clang fp contract(off)

namespace fl {
namespace asm2 {

constexpr double kGeomEps = 1.0e-14;       // GEOMETRY_EPSILON (src/mesh.rs)
constexpr double kF64Eps = 2.220446049250313e-16;
constexpr int kMaxPolygon = 16;            // vertices of one polygon cell

// Mesh error kinds (MeshError, src/mesh.rs:270-310) in check order
enum { ME_NONE = 0, ME_INVALID_CELL = 1, ME_INVALID_POINT = 2, ME_DEGENERATE_FACE = 3, ME_NON_MANIFOLD = 4 };

// first error of the topology pass: (cell << 24 | position << 4 | kind),
// position 0 for the cell's own checks, 1 + t for its edge t
__device__ inline unsigned long long topo_key(uint32_t cell, uint32_t pos, uint32_t kind) {
    return ((unsigned long long)cell << 24) | ((unsigned long long)pos << 4) | kind;
}

__global__ void k_point_check(int64_t npts, const double *__restrict__ xy, unsigned long long *first) {
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npts; p += (int64_t)gridDim.x * blockDim.x)
        if (!isfinite(xy[2 * p]) || !isfinite(xy[2 * p + 1])) atomicMin(first, (unsigned long long)p);
}

// Per cell: vertex count and point indices, the vertex average (the
// orientation reference of its faces), and the cell of each emission.
__global__ void k_cells(int ncells, int64_t npts, const double *__restrict__ xy, const int *__restrict__ off,
                        const int *__restrict__ vtx, double *__restrict__ ref, int *__restrict__ ecell,
                        unsigned long long *first) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
        const int o = off[c], nv = off[c + 1] - o;
        for (int t = 0; t < nv; ++t) ecell[o + t] = c;
        if (nv < 3) {
            atomicMin(first, topo_key(c, 0, ME_INVALID_CELL));
            continue;
        }
        double rx = 0.0, ry = 0.0;
        bool bad = false;
        for (int t = 0; t < nv; ++t) {
            const int p = vtx[o + t];
            if (p < 0 || p >= npts) {
                bad = true;
                break;
            }
            rx += xy[2 * p];
            ry += xy[2 * p + 1];
        }
        if (bad) {
            atomicMin(first, topo_key(c, 0, ME_INVALID_POINT));
            continue;
        }
        ref[2 * c] = rx / (double)nv;
        ref[2 * c + 1] = ry / (double)nv;
    }
}

// Per emission: the edge oriented away from the cell's vertex average, its
// centre and area vector, and the sort key of its vertex pair.
__global__ void k_emit(int64_t nemit, int64_t npts, const double *__restrict__ xy, const int *__restrict__ off,
                       const int *__restrict__ vtx, const int *__restrict__ ecell, const double *__restrict__ ref,
                       unsigned long long *__restrict__ key, uint32_t *__restrict__ idx, double *__restrict__ geo,
                       unsigned long long *first) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nemit; e += (int64_t)gridDim.x * blockDim.x) {
        const int c = ecell[e], o = off[c], nv = off[c + 1] - o, t = (int)(e - o);
        idx[e] = (uint32_t)e;
        key[e] = ~0ull;  // cells that failed their own checks take no part
        if (nv < 3 || nv > kMaxPolygon) continue;
        int a = vtx[o + t], b = vtx[o + (t + 1) % nv];
        if (a < 0 || a >= npts || b < 0 || b >= npts) continue;
        const double rx = ref[2 * c], ry = ref[2 * c + 1];
        double ccx = (xy[2 * a] + xy[2 * b]) / 2.0, ccy = (xy[2 * a + 1] + xy[2 * b + 1]) / 2.0;
        double sx = xy[2 * b + 1] - xy[2 * a + 1], sy = -(xy[2 * b] - xy[2 * a]);
        if (!(sqrt(sx * sx + sy * sy) > kGeomEps)) {
            atomicMin(first, topo_key(c, 1 + t, ME_DEGENERATE_FACE));
            continue;
        }
        if (sx * (ccx - rx) + sy * (ccy - ry) < 0.0) {
            const int tmp = a;
            a = b;
            b = tmp;
            ccx = (xy[2 * a] + xy[2 * b]) / 2.0;
            ccy = (xy[2 * a + 1] + xy[2 * b + 1]) / 2.0;
            sx = xy[2 * b + 1] - xy[2 * a + 1];
            sy = -(xy[2 * b] - xy[2 * a]);
        }
        const uint32_t ka = (uint32_t)min(a, b), kb = (uint32_t)max(a, b);
        key[e] = ((unsigned long long)ka << 32) | kb;
        geo[4 * e] = ccx;
        geo[4 * e + 1] = ccy;
        geo[4 * e + 2] = sx;
        geo[4 * e + 3] = sy;
    }
}

// On the sorted emissions: flag the first of each group (the face's
// creating emission), record its partner (the second), report a third.
__global__ void k_groups(int64_t nemit, const unsigned long long *__restrict__ skey, const uint32_t *__restrict__ sidx,
                         const int *__restrict__ ecell, const int *__restrict__ off, uint32_t *__restrict__ flag,
                         int *__restrict__ partner, unsigned long long *first) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nemit; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long k = skey[i];
        if (k == ~0ull) continue;
        const bool head = i == 0 || skey[i - 1] != k;
        if (head) {
            const uint32_t e0 = sidx[i];
            flag[e0] = 1;
            partner[e0] = (i + 1 < nemit && skey[i + 1] == k) ? (int)sidx[i + 1] : -1;
        } else if (i >= 2 && skey[i - 2] == k) {  // third (or later) emission of one face
            const uint32_t e = sidx[i];
            const int c = ecell[e];
            atomicMin(first, topo_key((uint32_t)c, 1 + (uint32_t)(e - off[c]), ME_NON_MANIFOLD));
        }
    }
}

// Faces from their creating emissions; every emission learns its face.
__global__ void k_faces(int64_t nemit, const uint32_t *__restrict__ flag, const uint32_t *__restrict__ fid,
                        const int *__restrict__ partner, const int *__restrict__ ecell,
                        const double *__restrict__ geo, int *__restrict__ owner, int *__restrict__ neigh,
                        double *__restrict__ fcx, double *__restrict__ fcy, double *__restrict__ sx,
                        double *__restrict__ sy, int *__restrict__ emit_face) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nemit; e += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[e]) continue;
        const int f = (int)fid[e], q = partner[e];
        owner[f] = ecell[e];
        neigh[f] = q >= 0 ? ecell[q] : -1;
        fcx[f] = geo[4 * e];
        fcy[f] = geo[4 * e + 1];
        sx[f] = geo[4 * e + 2];
        sy[f] = geo[4 * e + 3];
        emit_face[e] = f;
        if (q >= 0) emit_face[q] = f;
    }
}

// cell_geometry (TwoD): shoelace centroid and area over the vertices in
// definition order; the first zero-area cell.
__global__ void k_cell_geometry(int ncells, const double *__restrict__ xy, const int *__restrict__ off,
                                const int *__restrict__ vtx, double *__restrict__ cx, double *__restrict__ cy,
                                double *__restrict__ vol, unsigned long long *first) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
        const int o = off[c], nv = off[c + 1] - o;
        double twice = 0.0, sx = 0.0, sy = 0.0;
        for (int t = 0; t < nv; ++t) {
            const int a = vtx[o + t], b = vtx[o + (t + 1) % nv];
            const double ax = xy[2 * a], ay = xy[2 * a + 1], bx = xy[2 * b], by = xy[2 * b + 1];
            const double cr = ax * by - bx * ay;
            twice += cr;
            sx += (ax + bx) * cr;
            sy += (ay + by) * cr;
        }
        const double signed_area = twice * 0.5;
        if (!(fabs(signed_area) > kGeomEps)) {
            atomicMin(first, (unsigned long long)c);
            continue;
        }
        cx[c] = sx / (3.0 * twice);
        cy[c] = sy / (3.0 * twice);
        vol[c] = fabs(signed_area);
    }
}

__global__ void k_orientation(int nfaces, const int *__restrict__ owner, const int *__restrict__ neigh,
                              const double *__restrict__ fcx, const double *__restrict__ fcy,
                              const double *__restrict__ sx, const double *__restrict__ sy,
                              const double *__restrict__ cx, const double *__restrict__ cy,
                              unsigned long long *first) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nfaces; f += gridDim.x * blockDim.x) {
        const double area = sqrt(sx[f] * sx[f] + sy[f] * sy[f]);
        const double nx = sx[f] / area, ny = sy[f] / area;
        const int P = owner[f], N = neigh[f];
        if ((fcx[f] - cx[P]) * nx + (fcy[f] - cy[P]) * ny <= kGeomEps ||
            (N >= 0 && (cx[N] - fcx[f]) * nx + (cy[N] - fcy[f]) * ny <= kGeomEps))
            atomicMin(first, (unsigned long long)f);
    }
}

// ------------------------------------------------------------- Poisson
// Poisson error phases, in the reference's check order: the cell sources,
// then the faces in order (kind by face), then the reference.
enum {
    PE_NONE = 0, PE_SOURCE = 1, PE_DEGENERATE = 2, PE_GEOMETRY = 3, PE_BOUNDARY_VALUE = 4,
    PE_FACE_DISTANCE = 5, PE_MISSING_BOUNDARY = 6
};
__device__ inline unsigned long long poisson_key(uint32_t phase, uint32_t index, uint32_t kind) {
    return ((unsigned long long)phase << 40) | ((unsigned long long)index << 4) | kind;
}

// Per face: the two-point coefficient gamma_f (S.d) / |d|^2 with the
// Linear interpolation of the constant diffusivity (interpolate_diffusivity,
// numerics.rs:383-418: (1 - lambda) gamma + lambda gamma on interior faces,
// the owner value on boundary faces) and its triplets in the reference's
// order -- slots 4 f .. 4 f + 3 -- minus the pinned row and column.
__global__ void k_poisson_faces(int nfaces, double gamma, const int *__restrict__ owner,
                                const int *__restrict__ neigh, const double *__restrict__ fcx,
                                const double *__restrict__ fcy, const double *__restrict__ sx,
                                const double *__restrict__ sy, const double *__restrict__ cx,
                                const double *__restrict__ cy, const int *__restrict__ bc_kind,
                                const double *__restrict__ bc_value, int has_ref, int ref_cell, int n,
                                double *__restrict__ coef, unsigned long long *__restrict__ key,
                                uint32_t *__restrict__ idx, double *__restrict__ tval,
                                unsigned long long *first) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < nfaces; f += gridDim.x * blockDim.x) {
        const int P = owner[f], N = neigh[f];
        for (int j = 0; j < 4; ++j) {
            key[4 * (size_t)f + j] = ~0ull;
            idx[4 * (size_t)f + j] = 4u * (uint32_t)f + j;
        }
        coef[f] = 0.0;
        if (N >= 0) {
            const double dx = cx[N] - cx[P], dy = cy[N] - cy[P];
            const double d2 = dx * dx + dy * dy;
            if (!isfinite(d2) || d2 <= kF64Eps) {
                atomicMin(first, poisson_key(2, (uint32_t)f, PE_DEGENERATE));
                continue;
            }
            const double lambda = ((fcx[f] - cx[P]) * dx + (fcy[f] - cy[P]) * dy) / d2;
            const double gf = (1.0 - lambda) * gamma + lambda * gamma;
            const double c = gf * (sx[f] * dx + sy[f] * dy) / d2;
            if (!isfinite(c) || c <= 0.0) {
                atomicMin(first, poisson_key(2, (uint32_t)f, PE_GEOMETRY));
                continue;
            }
            coef[f] = c;
            const int rows[4] = {P, P, N, N}, cols[4] = {P, N, N, P};
            const double vals[4] = {c, -c, c, -c};
            for (int j = 0; j < 4; ++j) {
                if (has_ref && (rows[j] == ref_cell || cols[j] == ref_cell)) continue;
                key[4 * (size_t)f + j] = (unsigned long long)rows[j] * (unsigned long long)n + cols[j];
                tval[4 * (size_t)f + j] = vals[j];
            }
        } else {
            const int kind = bc_kind ? bc_kind[f] : 0;
            if (kind == 1) {
                if (!isfinite(bc_value[f])) {
                    atomicMin(first, poisson_key(2, (uint32_t)f, PE_BOUNDARY_VALUE));
                    continue;
                }
                const double dx = fcx[f] - cx[P], dy = fcy[f] - cy[P];
                const double d2 = dx * dx + dy * dy;
                if (!isfinite(d2) || d2 <= kF64Eps) {
                    atomicMin(first, poisson_key(2, (uint32_t)f, PE_FACE_DISTANCE));
                    continue;
                }
                const double c = gamma * (sx[f] * dx + sy[f] * dy) / d2;
                if (!isfinite(c) || c <= 0.0) {
                    atomicMin(first, poisson_key(2, (uint32_t)f, PE_GEOMETRY));
                    continue;
                }
                coef[f] = c;
                if (!(has_ref && P == ref_cell)) {
                    key[4 * (size_t)f] = (unsigned long long)P * (unsigned long long)n + P;
                    tval[4 * (size_t)f] = c;
                }
            } else if (kind != 0) {
                atomicMin(first, poisson_key(2, (uint32_t)f, PE_MISSING_BOUNDARY));
            }
        }
    }
}

// Per cell, in the reference's order of updates to rhs[c]: the source times
// the volume, the fixed-value boundary faces in face order, then (with a
// reference) the moved pinned-column entries in triplet order.
__global__ void k_poisson_rhs(int ncells, const double *__restrict__ source, const double *__restrict__ vol,
                              const int *__restrict__ off, const int *__restrict__ emit_face,
                              const int *__restrict__ owner, const int *__restrict__ neigh,
                              const int *__restrict__ bc_kind, const double *__restrict__ bc_value,
                              const double *__restrict__ coef, int has_ref, int ref_cell, double ref_value,
                              double *__restrict__ rhs, unsigned long long *first) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
        double r = 0.0;
        if (source) {
            if (!isfinite(source[c])) atomicMin(first, poisson_key(1, (uint32_t)c, PE_SOURCE));
            r += source[c] * vol[c];
        }
        const int o = off[c], nv = off[c + 1] - o;
        int fs[kMaxPolygon];
        for (int t = 0; t < nv; ++t) fs[t] = emit_face[o + t];
        for (int a = 1; a < nv; ++a)  // the cell's faces in face order
            for (int b = a; b > 0 && fs[b - 1] > fs[b]; --b) {
                const int tmp = fs[b];
                fs[b] = fs[b - 1];
                fs[b - 1] = tmp;
            }
        for (int t = 0; t < nv; ++t) {
            const int f = fs[t];
            if (neigh[f] < 0 && bc_kind && bc_kind[f] == 1) r += coef[f] * bc_value[f];
        }
        if (has_ref && c != ref_cell)
            for (int t = 0; t < nv; ++t) {
                const int f = fs[t];
                const int other = owner[f] == c ? neigh[f] : owner[f];
                if (neigh[f] >= 0 && other == ref_cell) r -= -coef[f] * ref_value;
            }
        rhs[c] = (has_ref && c == ref_cell) ? ref_value : r;
    }
}

// Sum each run of equal keys in insertion order; keep non-zero sums.
__global__ void k_reduce_runs(int64_t m, const unsigned long long *__restrict__ skey,
                              const uint32_t *__restrict__ sidx, const double *__restrict__ tval,
                              double *__restrict__ sum, uint32_t *__restrict__ keep, unsigned long long *bad) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        keep[i] = 0;
        const unsigned long long k = skey[i];
        if (k == ~0ull || (i > 0 && skey[i - 1] == k)) continue;
        double v = 0.0;
        for (int64_t j = i; j < m && skey[j] == k; ++j) v += tval[sidx[j]];
        sum[i] = v;
        if (!isfinite(v)) atomicMin(bad, (unsigned long long)i);
        keep[i] = v != 0.0;
    }
}

__global__ void k_compact(int64_t m, int n, const unsigned long long *__restrict__ skey,
                          const double *__restrict__ sum, const uint32_t *__restrict__ keep,
                          const uint32_t *__restrict__ pos, int *__restrict__ col, double *__restrict__ val,
                          int *__restrict__ rowcount) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        if (!keep[i]) continue;
        const unsigned long long k = skey[i];
        const int row = (int)(k / (unsigned long long)n), c = (int)(k % (unsigned long long)n);
        col[pos[i]] = c;
        val[pos[i]] = sum[i];
        atomicAdd(&rowcount[row], 1);
    }
}

__global__ void k_fill_u32(int64_t m, uint32_t *p, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_i32(int64_t m, int *p, int v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// A scratch allocation list freed on every exit path.
struct Scratch {
    std::vector<void *> bufs;
    ~Scratch() {
        for (void *b : bufs) (void)hipFree(b);
    }
    template <typename T>
    bool get(T **p, size_t count) {
        *p = nullptr;
        if (hipMalloc((void **)p, (count ? count : 1) * sizeof(T)) != hipSuccess) {
            *p = nullptr;
            return false;
        }
        bufs.push_back(*p);
        return true;
    }
};

}  // namespace asm2
}  // namespace fl

using namespace fl;
using namespace fl::asm2;


namespace {

inline uint32_t blocks_for(fl_ctx *ctx, int64_t n) {
    const int64_t b = (n + kBlock - 1) / kBlock;
    return (uint32_t)std::max<int64_t>(1, std::min<int64_t>(b, (int64_t)ctx->num_cu * 8));
}

// rocPRIM wrappers with their own temporary storage
template <typename K, typename V>
hipError_t sort_pairs(Scratch &sc, const K *kin, K *kout, const V *vin, V *vout, size_t m, unsigned end_bit,
                      hipStream_t s) {
    size_t bytes = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, m, 0, end_bit, s);
    if (e != hipSuccess) return e;
    void *tmp = nullptr;
    if (!sc.get((char **)&tmp, bytes)) return hipErrorOutOfMemory;
    return rocprim::radix_sort_pairs(tmp, bytes, kin, kout, vin, vout, m, 0, end_bit, s);
}
template <typename T>
hipError_t exclusive_scan(Scratch &sc, const T *in, T *out, size_t m, hipStream_t s) {
    size_t bytes = 0;
    hipError_t e = rocprim::exclusive_scan(nullptr, bytes, in, out, (T)0, m, rocprim::plus<T>(), s);
    if (e != hipSuccess) return e;
    void *tmp = nullptr;
    if (!sc.get((char **)&tmp, bytes)) return hipErrorOutOfMemory;
    return rocprim::exclusive_scan(tmp, bytes, in, out, (T)0, m, rocprim::plus<T>(), s);
}

}  // namespace

#define ASM_TRY(call)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) return set_error(ctx, FL_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

extern "C" {

fl_status fl_mesh2d_create(fl_ctx *ctx, int64_t npts, const double *xy, int64_t ncells, const int64_t *cell_off,
                           const int64_t *cell_vtx, fl_mesh2d **out, int32_t *mesh_error) {
    if (mesh_error) *mesh_error = FL_MESH_OK;
    if (!ctx || !out || (npts > 0 && !xy) || ncells < 0 || (ncells > 0 && (!cell_off || !cell_vtx)))
        return set_error(ctx, FL_ERR_INVALID_ARGUMENT, "null or negative argument");
    *out = nullptr;
    auto mesh_fail = [&](int kind, const char *fmt, long long a) {
        if (mesh_error) *mesh_error = kind;
        return set_error(ctx, FL_ERR_MESH, fmt, a);
    };
    if (ncells == 0) return mesh_fail(FL_MESH_EMPTY, "mesh has no cells%.0lld", 0);
    if (cell_off[0] != 0)
        return set_error(ctx, FL_ERR_INVALID_ARGUMENT, "cell offsets must start at zero%.0lld", 0ll);
    for (int64_t c = 0; c < ncells; ++c)
        if (cell_off[c + 1] < cell_off[c])
            return set_error(ctx, FL_ERR_INVALID_ARGUMENT, "cell offsets must not decrease (cell %lld)",
                             (long long)c);
    const int64_t nemit = cell_off[ncells];
    for (int64_t c = 0; c < ncells; ++c)
        if (cell_off[c + 1] - cell_off[c] > kMaxPolygon)
            return set_error(ctx, FL_ERR_INVALID_ARGUMENT,
                             "cell %lld: polygons of more than 16 vertices are not supported on the device",
                             (long long)c);
    if (npts >= ((int64_t)1 << 31) || nemit >= ((int64_t)1 << 31) || ncells >= ((int64_t)1 << 31))
        return set_error(ctx, FL_ERR_INVALID_ARGUMENT, "mesh too large for 32-bit device indices");
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    std::vector<int> hoff(ncells + 1), hvtx(nemit > 0 ? nemit : 1);
    for (int64_t c = 0; c <= ncells; ++c) hoff[c] = (int)cell_off[c];
    for (int64_t e = 0; e < nemit; ++e) {
        const int64_t p = cell_vtx[e];
        hvtx[e] = (p < 0 || p >= npts) ? -1 : (int)p;
    }
    Scratch sc;
    double *dxy, *ref, *geo;
    int *dvtx, *ecell, *partner;
    unsigned long long *key, *skey, *first;
    uint32_t *idx, *sidx, *flag, *fid;
    fl_mesh2d *m = new (std::nothrow) fl_mesh2d();
    if (!m) return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "host allocation failed");
    m->ctx = ctx;
    m->npts = npts;
    m->ncells = ncells;
    m->nemit = nemit;
    auto fail_oom = [&]() {
        m->release();
        delete m;
        return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "device allocation failed");
    };
    if (!sc.get(&dxy, 2 * npts) || !sc.get(&ref, 2 * ncells) || !sc.get(&geo, 4 * nemit) || !sc.get(&dvtx, nemit) ||
        !sc.get(&ecell, nemit) || !sc.get(&partner, nemit) || !sc.get(&key, nemit) || !sc.get(&skey, nemit) ||
        !sc.get(&first, 4) || !sc.get(&idx, nemit) || !sc.get(&sidx, nemit) || !sc.get(&flag, nemit + 1) ||
        !sc.get(&fid, nemit + 1))
        return fail_oom();
    if (hipMalloc(&m->off, (ncells + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&m->emit_face, (nemit ? nemit : 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&m->cx, ncells * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->cy, ncells * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->vol, ncells * sizeof(double)) != hipSuccess)
        return fail_oom();
    const unsigned long long none[4] = {~0ull, ~0ull, ~0ull, ~0ull};
    ASM_TRY(hipMemcpyAsync(first, none, sizeof none, hipMemcpyHostToDevice, s));
    if (npts) ASM_TRY(hipMemcpyAsync(dxy, xy, 2 * npts * sizeof(double), hipMemcpyHostToDevice, s));
    ASM_TRY(hipMemcpyAsync(m->off, hoff.data(), (ncells + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    if (nemit) ASM_TRY(hipMemcpyAsync(dvtx, hvtx.data(), nemit * sizeof(int), hipMemcpyHostToDevice, s));
    // points, cells, emissions, faces
    k_point_check<<<blocks_for(ctx, npts), kBlock, 0, s>>>(npts, dxy, first);
    k_cells<<<blocks_for(ctx, ncells), kBlock, 0, s>>>((int)ncells, npts, dxy, m->off, dvtx, ref, ecell, first + 1);
    k_emit<<<blocks_for(ctx, nemit), kBlock, 0, s>>>(nemit, npts, dxy, m->off, dvtx, ecell, ref, key, idx, geo,
                                                     first + 1);
    ASM_TRY(hipGetLastError());
    ASM_TRY(sort_pairs(sc, key, skey, idx, sidx, (size_t)nemit, 64, s));
    k_fill_u32<<<blocks_for(ctx, nemit + 1), kBlock, 0, s>>>(nemit + 1, flag, 0u);
    k_groups<<<blocks_for(ctx, nemit), kBlock, 0, s>>>(nemit, skey, sidx, ecell, m->off, flag, partner, first + 1);
    ASM_TRY(hipGetLastError());
    ASM_TRY(exclusive_scan(sc, flag, fid, (size_t)nemit + 1, s));
    uint32_t nf = 0;
    unsigned long long h[4];
    ASM_TRY(hipMemcpyAsync(&nf, fid + nemit, sizeof nf, hipMemcpyDeviceToHost, s));
    ASM_TRY(hipMemcpyAsync(h, first, sizeof h, hipMemcpyDeviceToHost, s));
    ASM_TRY(hipStreamSynchronize(s));
    auto first_error = [&]() -> fl_status {
        if (h[0] != ~0ull) return mesh_fail(FL_MESH_NON_FINITE_POINT, "non-finite point %lld", (long long)h[0]);
        if (h[1] != ~0ull) {
            const long long c = (long long)(h[1] >> 24);
            switch ((int)(h[1] & 15)) {
                case ME_INVALID_CELL:
                    return mesh_fail(FL_MESH_INVALID_CELL, "a polygon requires at least three vertices (cell %lld)", c);
                case ME_INVALID_POINT:
                    return mesh_fail(FL_MESH_INVALID_POINT_INDEX, "invalid point index in cell %lld", c);
                case ME_DEGENERATE_FACE:
                    return mesh_fail(FL_MESH_DEGENERATE_FACE, "degenerate face in cell %lld", c);
                default:
                    return mesh_fail(FL_MESH_NON_MANIFOLD, "non-manifold face at cell %lld", c);
            }
        }
        return FL_OK;
    };
    fl_status st = first_error();
    if (st != FL_OK) {
        m->release();
        delete m;
        return st;
    }
    m->nfaces = nf;
    if (hipMalloc(&m->owner, nf * sizeof(int)) != hipSuccess || hipMalloc(&m->neigh, nf * sizeof(int)) != hipSuccess ||
        hipMalloc(&m->fcx, nf * sizeof(double)) != hipSuccess || hipMalloc(&m->fcy, nf * sizeof(double)) != hipSuccess ||
        hipMalloc(&m->sx, nf * sizeof(double)) != hipSuccess || hipMalloc(&m->sy, nf * sizeof(double)) != hipSuccess)
        return fail_oom();
    k_faces<<<blocks_for(ctx, nemit), kBlock, 0, s>>>(nemit, flag, fid, partner, ecell, geo, m->owner, m->neigh,
                                                      m->fcx, m->fcy, m->sx, m->sy, m->emit_face);
    k_cell_geometry<<<blocks_for(ctx, ncells), kBlock, 0, s>>>((int)ncells, dxy, m->off, dvtx, m->cx, m->cy,
                                                               m->vol, first + 2);
    k_orientation<<<blocks_for(ctx, nf), kBlock, 0, s>>>((int)nf, m->owner, m->neigh, m->fcx, m->fcy, m->sx,
                                                         m->sy, m->cx, m->cy, first + 3);
    ASM_TRY(hipGetLastError());
    ASM_TRY(hipMemcpyAsync(h, first, sizeof h, hipMemcpyDeviceToHost, s));
    ASM_TRY(hipStreamSynchronize(s));
    if (h[2] != ~0ull) st = mesh_fail(FL_MESH_ZERO_VOLUME, "zero-volume cell %lld", (long long)h[2]);
    else if (h[3] != ~0ull)
        st = mesh_fail(FL_MESH_INCONSISTENT_ORIENTATION, "face %lld has inconsistent owner/neighbour orientation",
                       (long long)h[3]);
    if (st != FL_OK) {
        m->release();
        delete m;
        return st;
    }
    ctx_retain(ctx);
    *out = m;
    return FL_OK;
}

void fl_mesh2d_destroy(fl_mesh2d *m) {
    if (!m) return;
    fl_ctx *ctx = m->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    m->release();
    delete m;
    ctx_release(ctx);
}

fl_status fl_mesh2d_info(fl_mesh2d *m, int64_t *ncells, int64_t *nfaces) {
    if (!m) return FL_ERR_INVALID_ARGUMENT;
    if (ncells) *ncells = m->ncells;
    if (nfaces) *nfaces = m->nfaces;
    return FL_OK;
}

fl_status fl_mesh2d_get(fl_mesh2d *m, double *cx, double *cy, double *vol, int64_t *owner, int64_t *neigh,
                        double *fcx, double *fcy, double *sx, double *sy) {
    if (!m) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = m->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    auto d2h = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
        return dst ? hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) : hipSuccess;
    };
    const size_t nc = (size_t)m->ncells, nf = (size_t)m->nfaces;
    ASM_TRY(d2h(cx, m->cx, nc * 8));
    ASM_TRY(d2h(cy, m->cy, nc * 8));
    ASM_TRY(d2h(vol, m->vol, nc * 8));
    ASM_TRY(d2h(fcx, m->fcx, nf * 8));
    ASM_TRY(d2h(fcy, m->fcy, nf * 8));
    ASM_TRY(d2h(sx, m->sx, nf * 8));
    ASM_TRY(d2h(sy, m->sy, nf * 8));
    std::vector<int> t(nf ? nf : 1);
    if (owner) {
        ASM_TRY(hipMemcpy(t.data(), m->owner, nf * sizeof(int), hipMemcpyDeviceToHost));
        for (size_t f = 0; f < nf; ++f) owner[f] = t[f];
    }
    if (neigh) {
        ASM_TRY(hipMemcpy(t.data(), m->neigh, nf * sizeof(int), hipMemcpyDeviceToHost));
        for (size_t f = 0; f < nf; ++f) neigh[f] = t[f];
    }
    return FL_OK;
}

fl_status fl_mesh2d_assemble_poisson(fl_mesh2d *m, double gamma, const double *source, const int32_t *bc_kind,
                                     const double *bc_value, int32_t has_reference, int64_t ref_cell,
                                     double ref_value, fl_csr **matrix, double *rhs_out, int32_t *poisson_error,
                                     double *detail) {
    if (poisson_error) *poisson_error = FL_POISSON_OK;
    if (!m || !matrix || !rhs_out) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = m->ctx;
    *matrix = nullptr;
    auto pfail = [&](int kind, const char *fmt, long long a) {
        if (poisson_error) *poisson_error = kind;
        return set_error(ctx, FL_ERR_POISSON, fmt, a);
    };
    if (!std::isfinite(gamma)) return pfail(FL_POISSON_NUMERICS, "non-finite diffusivity%.0lld", 0);
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int64_t nc = m->ncells, nf = m->nfaces, nt = 4 * nf + 1;
    Scratch sc;
    double *dsrc = nullptr, *dval_bc = nullptr, *coef, *tval, *rhs, *sum;
    int *dkind = nullptr;
    unsigned long long *key, *skey, *first;
    uint32_t *idx, *sidx, *keep, *pos;
    if ((source && !sc.get(&dsrc, nc)) || (bc_kind && (!sc.get(&dkind, nf) || !sc.get(&dval_bc, nf))) ||
        !sc.get(&coef, nf) || !sc.get(&tval, nt) || !sc.get(&rhs, nc) || !sc.get(&sum, nt) || !sc.get(&key, nt) ||
        !sc.get(&skey, nt) || !sc.get(&first, 2) || !sc.get(&idx, nt) || !sc.get(&sidx, nt) ||
        !sc.get(&keep, nt + 1) || !sc.get(&pos, nt + 1))
        return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "device allocation failed");
    const unsigned long long none[2] = {~0ull, ~0ull};
    ASM_TRY(hipMemcpyAsync(first, none, sizeof none, hipMemcpyHostToDevice, s));
    if (source) ASM_TRY(hipMemcpyAsync(dsrc, source, nc * 8, hipMemcpyHostToDevice, s));
    std::vector<int> hk;
    if (bc_kind) {
        hk.assign(bc_kind, bc_kind + nf);
        ASM_TRY(hipMemcpyAsync(dkind, hk.data(), nf * sizeof(int), hipMemcpyHostToDevice, s));
        ASM_TRY(hipMemcpyAsync(dval_bc, bc_value, nf * 8, hipMemcpyHostToDevice, s));
    }
    const int href = has_reference ? 1 : 0;
    k_poisson_faces<<<blocks_for(ctx, nf), kBlock, 0, s>>>((int)nf, gamma, m->owner, m->neigh, m->fcx, m->fcy,
                                                           m->sx, m->sy, m->cx, m->cy, dkind, dval_bc, href,
                                                           (int)ref_cell, (int)nc, coef, key, idx, tval, first);
    // the appended (ref, ref, 1) after every face's slots
    const unsigned long long last_key = href ? (unsigned long long)ref_cell * (unsigned long long)nc + ref_cell : ~0ull;
    const uint32_t last_idx = (uint32_t)(4 * nf);
    const double one = 1.0;
    ASM_TRY(hipMemcpyAsync(key + 4 * nf, &last_key, 8, hipMemcpyHostToDevice, s));
    ASM_TRY(hipMemcpyAsync(idx + 4 * nf, &last_idx, 4, hipMemcpyHostToDevice, s));
    ASM_TRY(hipMemcpyAsync(tval + 4 * nf, &one, 8, hipMemcpyHostToDevice, s));
    k_poisson_rhs<<<blocks_for(ctx, nc), kBlock, 0, s>>>((int)nc, dsrc, m->vol, m->off, m->emit_face, m->owner,
                                                         m->neigh, dkind, dval_bc, coef, href, (int)ref_cell,
                                                         ref_value, rhs, first);
    ASM_TRY(hipGetLastError());
    unsigned long long h[2];
    ASM_TRY(hipMemcpyAsync(h, first, sizeof h, hipMemcpyDeviceToHost, s));
    ASM_TRY(hipMemcpyAsync(rhs_out, rhs, nc * 8, hipMemcpyDeviceToHost, s));
    ASM_TRY(hipStreamSynchronize(s));
    if (h[0] != ~0ull) {
        const long long i = (long long)((h[0] >> 4) & 0xFFFFFFFFFull);
        switch ((int)(h[0] & 15)) {
            case PE_SOURCE: return pfail(FL_POISSON_NON_FINITE_SOURCE, "source value in cell %lld is not finite", i);
            case PE_DEGENERATE:
                return pfail(FL_POISSON_DEGENERATE_DISTANCE, "degenerate owner-neighbour distance at face %lld", i);
            case PE_GEOMETRY:
                return pfail(FL_POISSON_INVALID_GEOMETRY, "invalid non-orthogonal geometry at face %lld", i);
            case PE_BOUNDARY_VALUE:
                return pfail(FL_POISSON_NON_FINITE_SOURCE, "boundary value at face %lld is not finite", i);
            case PE_FACE_DISTANCE:
                return pfail(FL_POISSON_DEGENERATE_FACE_DISTANCE, "degenerate owner-face distance at face %lld", i);
            default: return pfail(FL_POISSON_MISSING_BOUNDARY, "missing boundary condition at face %lld", i);
        }
    }
    // pure-Neumann compatibility (poisson.rs:248-257), summed in cell order
    bool has_fixed = false;
    if (bc_kind)
        for (int64_t f = 0; f < nf; ++f) has_fixed = has_fixed || bc_kind[f] == 1;
    if (!has_fixed) {
        // the check reads the right-hand side before the reference is applied
        std::vector<double> pre(rhs_out, rhs_out + nc);
        if (href) {
            // undo the reference's rhs changes for the check: recompute source * volume only
            std::vector<double> vol(nc);
            ASM_TRY(hipMemcpy(vol.data(), m->vol, nc * 8, hipMemcpyDeviceToHost));
            for (int64_t c = 0; c < nc; ++c) pre[c] = source ? 0.0 + source[c] * vol[c] : 0.0;
        }
        double integrated = 0.0, scale = 0.0;
        for (int64_t c = 0; c < nc; ++c) integrated += pre[c];
        for (int64_t c = 0; c < nc; ++c) scale += std::fabs(pre[c]);
        if (std::fabs(integrated) > 1.0e-12 * (scale > 1.0 ? scale : 1.0)) {
            if (detail) *detail = integrated;
            return pfail(FL_POISSON_INCOMPATIBLE_SOURCE, "pure-Neumann source is incompatible%.0lld", 0);
        }
        if (!href)
            return pfail(FL_POISSON_MISSING_REFERENCE,
                         "pure ZeroGradient Poisson systems require an explicit reference cell%.0lld", 0);
    }
    // apply_reference's own checks (poisson.rs:303-337)
    if (href && (ref_cell < 0 || ref_cell >= nc))
        return pfail(FL_POISSON_INVALID_REFERENCE_CELL, "reference cell %lld is outside the mesh", (long long)ref_cell);
    if (href && !std::isfinite(ref_value))
        return pfail(FL_POISSON_INVALID_REFERENCE_VALUE, "reference value is not finite%.0lld", 0);
    // CsrBuilder::finalize: stable sort by (row, col), run sums, compaction
    ASM_TRY(sort_pairs(sc, key, skey, idx, sidx, (size_t)nt, 64, s));
    k_reduce_runs<<<blocks_for(ctx, nt), kBlock, 0, s>>>(nt, skey, sidx, tval, sum, keep, first + 1);
    k_fill_u32<<<1, 1, 0, s>>>(1, keep + nt, 0u);
    ASM_TRY(hipGetLastError());
    ASM_TRY(exclusive_scan(sc, keep, pos, (size_t)nt + 1, s));
    uint32_t nnz = 0;
    unsigned long long bad = ~0ull;
    ASM_TRY(hipMemcpyAsync(&nnz, pos + nt, 4, hipMemcpyDeviceToHost, s));
    ASM_TRY(hipMemcpyAsync(&bad, first + 1, 8, hipMemcpyDeviceToHost, s));
    ASM_TRY(hipStreamSynchronize(s));
    if (bad != ~0ull) return pfail(FL_POISSON_LINEAR, "CSR finalize failed: non-finite assembled matrix%.0lld", 0);
    int *doff = nullptr, *dcol = nullptr, *count = nullptr;
    double *dval = nullptr;
    if (!sc.get(&count, nc + 1)) return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "device allocation failed");
    if (hipMalloc(&doff, (nc + 1) * sizeof(int)) != hipSuccess || hipMalloc(&dcol, (nnz ? nnz : 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&dval, (nnz ? nnz : 1) * sizeof(double)) != hipSuccess) {
        (void)hipFree(doff);
        (void)hipFree(dcol);
        (void)hipFree(dval);
        return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "device allocation failed");
    }
    k_fill_i32<<<blocks_for(ctx, nc + 1), kBlock, 0, s>>>(nc + 1, count, 0);
    k_compact<<<blocks_for(ctx, nt), kBlock, 0, s>>>(nt, (int)nc, skey, sum, keep, pos, dcol, dval, count);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = exclusive_scan(sc, count, doff, (size_t)nc + 1, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipFree(doff);
        (void)hipFree(dcol);
        (void)hipFree(dval);
        return set_error(ctx, FL_ERR_HIP, "CSR finalize failed: %s", hipGetErrorString(e));
    }
    return csr_adopt(ctx, (uint32_t)nc, (uint32_t)nc, nnz, doff, dcol, dval, matrix);
}

}  // extern "C"
