// Collocated unstructured SIMPLE on the device (solve_simple, src/simple.rs:
// 148-311) for 2D polygon meshes built by fl_mesh2d_create.
//
// Every piece of an outer iteration is a kernel over cells or faces; the
// host only sequences them, runs the two CSR solvers on device vectors and
// reads an error word and the continuity RMS once per iteration:
//
//   LSQ gradient of p (numerics.rs:650-682)              k_lsq_grad (per cell)
//   momentum rows + rhs (x, y) + rAU (momentum.rs:180-   k_momentum (per row)
//     341, simple.rs:316-331)
//   BiCGSTAB x, y (momentum.rs:445-475)                  csr_solve_device
//   Rhie-Chow flux + face coefficients (simple.rs:       k_rhie_chow (per face)
//     342-405, 746-805)
//   pressure-correction rows + rhs (simple.rs:518-638)   k_pcorr (per row)
//   Jacobi-PCG (simple.rs:643-668)                       csr_solve_device
//   p += a_p p', LSQ gradient of p', U = U* - rAU g      k_update_p, k_lsq_grad,
//     (simple.rs:670-701)                                  k_correct_velocity
//   face-flux correction (simple.rs:407-484)             k_correct_flux (per face)
//   continuity RMS (simple.rs:703-717)                   k_continuity + fold
//
// Rows are assembled per cell from the cell's faces taken in increasing face
// index, which is the order the reference's CsrBuilder receives (and, as in
// the oracle, sums) the duplicate entries of a row; with no contraction of
// products into fused multiply-adds the assembled values equal the oracle's
// bit for bit.  The momentum and pressure-correction matrices share one
// sparsity (the diagonal and the face neighbours, columns ascending), built
// once per mesh.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "mesh2d.hpp"

#pragma clang fp contract(off)

namespace fl {
namespace s2 {

constexpr double kF64Eps = 2.220446049250313e-16;
constexpr int kMaxFaces = 16;  // faces of one polygon cell (fl_mesh2d_create's limit)
enum : int8_t { BC_ZERO_GRADIENT = 0, BC_FIXED = 1, BC_NONE = -1 };
enum : int8_t { EQ_SKIP = 0, EQ_NEIGHBOUR = 1, EQ_FIXED = 2 };

// error words: the lowest failing index of each check, ~0 for none
enum {
    E_SINGULAR = 0,     // LSQ stencil (cell)
    E_MOM_FACE,         // momentum assembly: face << 1 | (0 non-finite flux, 1 backflow)
    E_SOURCE,           // non-finite momentum source (cell)
    E_RAU,              // invalid momentum diagonal (cell)
    E_COEF_IN,          // invalid interior pressure face coefficient (face)
    E_RC_FLUX,          // non-finite Rhie-Chow flux (face)
    E_COEF_BD,          // invalid boundary pressure face coefficient (face)
    E_VELOCITY,         // non-finite corrected velocity (cell)
    E_FLUX,             // non-finite corrected flux (face)
    E_COUNT
};

struct Geo {
    int nc, nf;
    const double *cx, *cy, *vol;
    const int *owner, *neigh;
    const double *fcx, *fcy, *sx, *sy;
    const int *off, *ef;
};

__device__ inline void flag(unsigned long long *err, int slot, unsigned long long key) {
    atomicMin(err + slot, key);
}

// Per face: the interpolation weight (interior), and (S . d) and |d|^2 with d
// from the owner to the neighbour's centre (interior) or to the face centre.
__global__ void k_face_geom(Geo g, double *lam, double *dot, double *d2, unsigned long long *bad) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < g.nf; f += gridDim.x * blockDim.x) {
        const int P = g.owner[f], N = g.neigh[f];
        double dx, dy;
        if (N >= 0) {
            dx = g.cx[N] - g.cx[P];
            dy = g.cy[N] - g.cy[P];
        } else {
            dx = g.fcx[f] - g.cx[P];
            dy = g.fcy[f] - g.cy[P];
        }
        const double dd = dx * dx + dy * dy;
        d2[f] = dd;
        dot[f] = g.sx[f] * dx + g.sy[f] * dy;
        lam[f] = 0.0;
        if (!(dd > kF64Eps)) {
            atomicMin(bad, (unsigned long long)f);
            continue;
        }
        if (N >= 0) lam[f] = ((g.fcx[f] - g.cx[P]) * dx + (g.fcy[f] - g.cy[P]) * dy) / dd;
    }
}

// LeastSquaresGradientStencil::new (2D): one equation per cell face in
// Cell::faces order, ZeroGradient faces skipped, the inverse 2x2 normal matrix.
__global__ void k_lsq_build(Geo g, const int8_t *pkind, const double *pval, int8_t *ekind, int *enb,
                            double *eval, double *rcx, double *rcy, double *inv,
                            unsigned long long *err) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g.nc; c += gridDim.x * blockDim.x) {
        double n0 = 0.0, n1 = 0.0, n3 = 0.0, n4 = 0.0;
        for (int e = g.off[c]; e < g.off[c + 1]; ++e) {
            const int f = g.ef[e];
            double ox, oy;
            if (g.neigh[f] >= 0) {
                const int other = g.owner[f] == c ? g.neigh[f] : g.owner[f];
                ox = g.cx[other] - g.cx[c];
                oy = g.cy[other] - g.cy[c];
                ekind[e] = EQ_NEIGHBOUR;
                enb[e] = other;
                eval[e] = 0.0;
            } else if (pkind[f] == BC_FIXED) {
                ox = g.fcx[f] - g.cx[c];
                oy = g.fcy[f] - g.cy[c];
                ekind[e] = EQ_FIXED;
                enb[e] = c;
                eval[e] = pval[f];
            } else {
                ekind[e] = EQ_SKIP;
                enb[e] = c;
                eval[e] = 0.0;
                rcx[e] = 0.0;
                rcy[e] = 0.0;
                continue;
            }
            const double w = 1.0 / (ox * ox + oy * oy);
            const double rx = ox * w, ry = oy * w;
            n0 += rx * ox;
            n1 += rx * oy;
            n3 += ry * ox;
            n4 += ry * oy;
            rcx[e] = rx;
            rcy[e] = ry;
        }
        const double det = n0 * n4 - n1 * n3;
        const double scale = fmax(fmax(fabs(n0), fabs(n4)), 1.0);
        const bool finite = isfinite(n0) && isfinite(n1) && isfinite(n3) && isfinite(n4);
        if (!finite || fabs(det) <= kF64Eps * scale * scale) {
            flag(err, E_SINGULAR, (unsigned long long)c);
            inv[4 * c] = inv[4 * c + 1] = inv[4 * c + 2] = inv[4 * c + 3] = 0.0;
            continue;
        }
        inv[4 * c + 0] = n4 / det;
        inv[4 * c + 1] = -n1 / det;
        inv[4 * c + 2] = -n3 / det;
        inv[4 * c + 3] = n0 / det;
    }
}

// Sparsity: per cell the diagonal and the distinct face neighbours.
__device__ inline int cell_columns(const Geo &g, int c, int *cols) {
    int k = 0;
    cols[k++] = c;
    for (int e = g.off[c]; e < g.off[c + 1]; ++e) {
        const int f = g.ef[e];
        if (g.neigh[f] < 0) continue;
        const int other = g.owner[f] == c ? g.neigh[f] : g.owner[f];
        bool seen = false;
        for (int t = 0; t < k; ++t) seen = seen || cols[t] == other;
        if (!seen) cols[k++] = other;
    }
    // insertion sort (at most 17 entries)
    for (int i = 1; i < k; ++i) {
        const int v = cols[i];
        int j = i - 1;
        while (j >= 0 && cols[j] > v) {
            cols[j + 1] = cols[j];
            --j;
        }
        cols[j + 1] = v;
    }
    return k;
}
__global__ void k_pattern_count(Geo g, int *count) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g.nc; c += gridDim.x * blockDim.x) {
        int cols[kMaxFaces + 1];
        count[c] = cell_columns(g, c, cols);
    }
}
// Columns of each row, the diagonal's position, and the cell's faces sorted
// by face index (the summation order of its rows).
__global__ void k_pattern_fill(Geo g, const int *roff, int *col, int *dpos, int *sfaces) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g.nc; c += gridDim.x * blockDim.x) {
        int cols[kMaxFaces + 1];
        const int k = cell_columns(g, c, cols);
        for (int t = 0; t < k; ++t) {
            col[roff[c] + t] = cols[t];
            if (cols[t] == c) dpos[c] = roff[c] + t;
        }
        const int o = g.off[c], n = g.off[c + 1] - o;
        int fs[kMaxFaces];
        for (int t = 0; t < n; ++t) fs[t] = g.ef[o + t];
        for (int i = 1; i < n; ++i) {
            const int v = fs[i];
            int j = i - 1;
            while (j >= 0 && fs[j] > v) {
                fs[j + 1] = fs[j];
                --j;
            }
            fs[j + 1] = v;
        }
        for (int t = 0; t < n; ++t) sfaces[o + t] = fs[t];
    }
}

// least_squares_gradient_into; `neg` gives pressure_gradient_source's
// -grad(p) / 1 instead.
struct Lsq {
    const int8_t *kind;
    const int *nb;
    const double *val, *rcx, *rcy, *inv;
};
__global__ void k_lsq_grad(Geo g, Lsq L, const double *__restrict__ v, double *__restrict__ gx,
                           double *__restrict__ gy) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g.nc; c += gridDim.x * blockDim.x) {
        double rx = 0.0, ry = 0.0;
        const double vc = v[c];
        for (int e = g.off[c]; e < g.off[c + 1]; ++e) {
            const int8_t k = L.kind[e];
            if (k == EQ_SKIP) continue;
            const double adj = k == EQ_NEIGHBOUR ? v[L.nb[e]] : L.val[e];
            const double diff = adj - vc;
            rx += L.rcx[e] * diff;
            ry += L.rcy[e] * diff;
        }
        const double *iv = L.inv + 4 * c;
        gx[c] = iv[0] * rx + iv[1] * ry;
        gy[c] = iv[2] * rx + iv[3] * ry;
    }
}

struct Pattern {
    const int *roff, *col, *dpos, *sfaces;
};
struct FaceGeo {
    const double *lam, *dot, *d2;
};

// assemble_momentum_component for x and y (one matrix: the components share
// it), sources -grad(p) * V, wall terms, equation relaxation, and rAU.
struct MomArgs {
    Geo g;
    Pattern pt;
    FaceGeo fg;
    double nu, alpha;
    int relaxed;
    const int8_t *vkind;
    const double *vx, *vy, *flux, *gx, *gy, *ux, *uy;
    double *val, *rhsx, *rhsy, *rau;
    unsigned long long *err;
};
__global__ void k_momentum(MomArgs a) {
    const Geo &g = a.g;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g.nc; c += gridDim.x * blockDim.x) {
        const double sxc = -a.gx[c] / 1.0, syc = -a.gy[c] / 1.0;
        if (!isfinite(sxc) || !isfinite(syc)) flag(a.err, E_SOURCE, (unsigned long long)c);
        double rx = 0.0, ry = 0.0, diag = 0.0;
        rx += sxc * g.vol[c];
        ry += syc * g.vol[c];
        const int o = g.off[c], n = g.off[c + 1] - o;
        for (int t = 0; t < n; ++t) {
            const int f = a.pt.sfaces[o + t];
            const double phi = a.flux[f];
            if (!isfinite(phi)) flag(a.err, E_MOM_FACE, (unsigned long long)f << 1);
            const double pmax = phi > 0.0 ? phi : 0.0, pmin = phi < 0.0 ? phi : 0.0;
            const int N = g.neigh[f];
            if (N >= 0) {
                const double l = a.fg.lam[f];
                const double nuf = (1.0 - l) * a.nu + l * a.nu;
                const double D = nuf * a.fg.dot[f] / a.fg.d2[f];
                diag += g.owner[f] == c ? D + pmax : D - pmin;
            } else {
                const double D = a.nu * a.fg.dot[f] / a.fg.d2[f];
                if (a.vkind[f] == BC_FIXED) {
                    diag += D + pmax;
                    rx += D * a.vx[f] - pmin * a.vx[f];
                    ry += D * a.vy[f] - pmin * a.vy[f];
                } else {
                    if (phi < 0.0) flag(a.err, E_MOM_FACE, (unsigned long long)f << 1 | 1);
                    diag += phi;
                }
            }
        }
        double dr = diag;
        if (a.relaxed) {
            dr = diag / a.alpha;
            rx += (1.0 - a.alpha) / a.alpha * diag * a.ux[c];
            ry += (1.0 - a.alpha) / a.alpha * diag * a.uy[c];
        }
        a.rhsx[c] = rx;
        a.rhsy[c] = ry;
        if (!isfinite(dr) || dr <= 0.0) {
            flag(a.err, E_RAU, (unsigned long long)c);
            a.rau[c] = 0.0;
        } else {
            a.rau[c] = 1.0 / dr;
        }
        // row entries: the diagonal, and per neighbour its faces' terms in face order
        for (int e = a.pt.roff[c]; e < a.pt.roff[c + 1]; ++e) {
            const int j = a.pt.col[e];
            if (j == c) {
                a.val[e] = dr;
                continue;
            }
            double v = 0.0;
            for (int t = 0; t < n; ++t) {
                const int f = a.pt.sfaces[o + t];
                const int N = g.neigh[f];
                if (N < 0) continue;
                const int P = g.owner[f];
                if ((P == c ? N : P) != j) continue;
                const double phi = a.flux[f];
                const double pmax = phi > 0.0 ? phi : 0.0, pmin = phi < 0.0 ? phi : 0.0;
                const double l = a.fg.lam[f];
                const double nuf = (1.0 - l) * a.nu + l * a.nu;
                const double D = nuf * a.fg.dot[f] / a.fg.d2[f];
                v += P == c ? -D + pmin : -D - pmax;
            }
            a.val[e] = v;
        }
    }
}

// Rhie-Chow predicted flux and the pressure face coefficients
// (pressure_face_coefficients(_with_boundary)) of every face.
struct RcArgs {
    Geo g;
    FaceGeo fg;
    const double *usx, *usy, *p, *gx, *gy, *rau, *flux;
    const int8_t *ckind;  // NULL: no correction boundary
    double *pred, *coef;
    unsigned long long *err;
};
__global__ void k_rhie_chow(RcArgs a) {
    const Geo &g = a.g;
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < g.nf; f += gridDim.x * blockDim.x) {
        const int P = g.owner[f], N = g.neigh[f];
        if (N >= 0) {
            const double l = a.fg.lam[f];
            const double fr = (1.0 - l) * a.rau[P] + l * a.rau[N];
            const double cf = fr * a.fg.dot[f] / a.fg.d2[f];
            if (!isfinite(cf) || cf <= 0.0) flag(a.err, E_COEF_IN, (unsigned long long)f);
            const double vx = a.usx[P] * (1.0 - l) + a.usx[N] * l;
            const double vy = a.usy[P] * (1.0 - l) + a.usy[N] * l;
            const double rxP = a.gx[P] * a.rau[P], ryP = a.gy[P] * a.rau[P];
            const double rxN = a.gx[N] * a.rau[N], ryN = a.gy[N] * a.rau[N];
            const double qx = rxP * (1.0 - l) + rxN * l, qy = ryP * (1.0 - l) + ryN * l;
            const double phi = (vx * g.sx[f] + vy * g.sy[f]) + (qx * g.sx[f] + qy * g.sy[f]) -
                               cf * (a.p[N] - a.p[P]);
            if (!isfinite(phi)) flag(a.err, E_RC_FLUX, (unsigned long long)f);
            a.pred[f] = phi;
            a.coef[f] = cf;
        } else if (a.ckind && a.ckind[f] == BC_FIXED) {
            a.pred[f] = a.usx[P] * g.sx[f] + a.usy[P] * g.sy[f];
            const double cf = a.rau[P] * a.fg.dot[f] / a.fg.d2[f];
            if (!isfinite(cf) || cf <= 0.0) flag(a.err, E_COEF_BD, (unsigned long long)f);
            a.coef[f] = cf;
        } else {
            a.pred[f] = a.flux[f];
            a.coef[f] = 0.0;
        }
    }
}

// assemble_pressure_correction (reference cell ref >= 0: its row and column
// become explicit zeros, 1 on its diagonal, rhs 0) or _with_boundary (ref < 0).
struct PcArgs {
    Geo g;
    Pattern pt;
    const double *pred, *coef;
    const int8_t *ckind;
    const double *cval;
    int ref;
    double *val, *rhs;
};
__global__ void k_pcorr(PcArgs a) {
    const Geo &g = a.g;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g.nc; c += gridDim.x * blockDim.x) {
        const int o = g.off[c], n = g.off[c + 1] - o;
        if (c == a.ref) {
            for (int e = a.pt.roff[c]; e < a.pt.roff[c + 1]; ++e) a.val[e] = a.pt.col[e] == c ? 1.0 : 0.0;
            a.rhs[c] = 0.0;
            continue;
        }
        double r = 0.0, diag = 0.0;
        for (int t = 0; t < n; ++t) {
            const int f = a.pt.sfaces[o + t];
            const double phi = a.pred[f];
            if (g.owner[f] == c) r -= phi;
            else r += phi;
            if (g.neigh[f] >= 0) {
                diag += a.coef[f];
            } else if (a.ckind && a.ckind[f] == BC_FIXED) {
                diag += a.coef[f];
                r += a.coef[f] * a.cval[f];
            }
        }
        a.rhs[c] = r;
        for (int e = a.pt.roff[c]; e < a.pt.roff[c + 1]; ++e) {
            const int j = a.pt.col[e];
            if (j == c) {
                a.val[e] = diag;
                continue;
            }
            double v = 0.0;
            if (j != a.ref)
                for (int t = 0; t < n; ++t) {
                    const int f = a.pt.sfaces[o + t];
                    const int N = g.neigh[f];
                    if (N < 0) continue;
                    const int P = g.owner[f];
                    if ((P == c ? N : P) == j) v += -a.coef[f];
                }
            a.val[e] = v;
        }
    }
}

__global__ void k_update_p(int nc, double ap, const double *__restrict__ pc, double *__restrict__ p) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) p[c] += ap * pc[c];
}

// correct_cell_velocity_from_gradient
__global__ void k_correct_velocity(int nc, const double *usx, const double *usy, const double *rau,
                                   const double *gx, const double *gy, double *ux, double *uy,
                                   unsigned long long *err) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        const double r = rau[c];
        const double a = usx[c] - gx[c] * r, b = usy[c] - gy[c] * r;
        ux[c] = a;
        uy[c] = b;
        if (!(isfinite(a) && isfinite(b))) flag(err, E_VELOCITY, (unsigned long long)c);
    }
}

// correct_face_flux(_with_boundary)
__global__ void k_correct_flux(Geo g, const double *pred, const double *coef, const double *pc,
                               const int8_t *ckind, const double *cval, double *flux,
                               unsigned long long *err) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < g.nf; f += gridDim.x * blockDim.x) {
        const int P = g.owner[f], N = g.neigh[f];
        double phi = pred[f];
        if (N >= 0) phi -= coef[f] * (pc[N] - pc[P]);
        else if (ckind && ckind[f] == BC_FIXED) phi -= coef[f] * (cval[f] - pc[P]);
        flux[f] = phi;
        if (!isfinite(phi)) flag(err, E_FLUX, (unsigned long long)f);
    }
}

// initial_face_flux
__global__ void k_initial_flux(Geo g, FaceGeo fg, const double *ux, const double *uy, const double *wall,
                               double *flux) {
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < g.nf; f += gridDim.x * blockDim.x) {
        const int P = g.owner[f], N = g.neigh[f];
        if (N < 0) {
            flux[f] = wall[f];
            continue;
        }
        const double l = fg.lam[f];
        const double vx = ux[P] * (1.0 - l) + ux[N] * l, vy = uy[P] * (1.0 - l) + uy[N] * l;
        flux[f] = vx * g.sx[f] + vy * g.sy[f];
    }
}

// continuity_rms: the integrated divergence of each cell (its faces in face
// order), squared and summed per block in a fixed tree, then over blocks.
__global__ void __launch_bounds__(kBlock) k_continuity(Geo g, Pattern pt, const double *flux, double *part) {
    __shared__ double red[kBlock];
    double s = 0.0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g.nc; c += gridDim.x * blockDim.x) {
        double imb = 0.0;
        for (int e = g.off[c]; e < g.off[c + 1]; ++e) {
            const int f = pt.sfaces[e];
            if (g.owner[f] == c) imb += flux[f];
            else imb -= flux[f];
        }
        s += imb * imb;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(kBlock) k_continuity_fold(int nparts, int nc, const double *part, double *out) {
    __shared__ double red[kBlock];
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kBlock) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sqrt(red[0] / (double)nc);
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_zero(double *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

}  // namespace s2
}  // namespace fl

using namespace fl;
using namespace fl::s2;

struct fl_simple2d {
    fl_ctx *ctx = nullptr;
    fl_mesh2d *mesh = nullptr;
    int nc = 0, nf = 0;
    int64_t nnz = 0;
    uint32_t blocks_c = 1, blocks_f = 1;
    bool has_corr = false, anchored = false;
    // boundary data and geometry
    int8_t *vkind = nullptr, *ckind = nullptr, *pkind = nullptr;
    double *vx = nullptr, *vy = nullptr, *cval = nullptr, *pval = nullptr;
    double *lam = nullptr, *dot = nullptr, *d2 = nullptr;
    // LSQ stencil
    int8_t *ekind = nullptr;
    int *enb = nullptr;
    double *eval = nullptr, *rcx = nullptr, *rcy = nullptr, *inv = nullptr;
    // sparsity (the momentum matrix owns its copy, the correction its own)
    int *dpos = nullptr, *sfaces = nullptr;
    const int *roff = nullptr, *col = nullptr;
    fl_csr *mom = nullptr, *pcm = nullptr;
    // state and work
    double *ux = nullptr, *uy = nullptr, *p = nullptr, *flux = nullptr;
    double *gx = nullptr, *gy = nullptr, *usx = nullptr, *usy = nullptr, *rhsx = nullptr, *rhsy = nullptr;
    double *rau = nullptr, *pred = nullptr, *coef = nullptr, *pcx = nullptr, *prhs = nullptr;
    double *part = nullptr, *rms = nullptr, *wall = nullptr;
    unsigned long long *err = nullptr, *h_err = nullptr;
    std::vector<void *> bufs;
    template <typename T>
    bool get(T **ptr, size_t count) {
        if (hipMalloc((void **)ptr, (count ? count : 1) * sizeof(T)) != hipSuccess) {
            *ptr = nullptr;
            return false;
        }
        bufs.push_back(*ptr);
        return true;
    }
    ~fl_simple2d() {
        for (void *b : bufs) (void)hipFree(b);
        if (h_err) (void)hipHostFree(h_err);
        if (mom) fl_csr_destroy(mom);
        if (pcm) fl_csr_destroy(pcm);
    }
};

namespace {

Geo geo(const fl_mesh2d *m) {
    return Geo{(int)m->ncells, (int)m->nfaces, m->cx, m->cy, m->vol, m->owner, m->neigh,
               m->fcx, m->fcy, m->sx, m->sy, m->off, m->emit_face};
}
uint32_t blocks(fl_ctx *ctx, int64_t n) {
    const int64_t b = (n + kBlock - 1) / kBlock;
    return (uint32_t)std::max<int64_t>(1, std::min<int64_t>(b, (int64_t)ctx->num_cu * 8));
}
fl_status simple_error(fl_ctx *ctx, fl_simple2d_report *rep, int kind, int64_t index, double value,
                       const char *text) {
    if (rep) {
        rep->error_kind = kind;
        rep->error_index = index;
        rep->error_value = value;
    }
    return set_error(ctx, FL_ERR_SIMPLE, "%s (index %lld)", text, (long long)index);
}
fl_status read_errors(fl_simple2d *s) {
    fl_ctx *ctx = s->ctx;
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->h_err, s->err, E_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   ctx->stream));
    FL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return FL_OK;
}
double fetch(fl_simple2d *s, const double *arr, int64_t i) {
    double v = 0.0;
    (void)hipMemcpy(&v, arr + i, sizeof(double), hipMemcpyDeviceToHost);
    return v;
}
Lsq lsq(const fl_simple2d *s) { return Lsq{s->ekind, s->enb, s->eval, s->rcx, s->rcy, s->inv}; }
Pattern pattern(const fl_simple2d *s) { return Pattern{s->roff, s->col, s->dpos, s->sfaces}; }
FaceGeo facegeo(const fl_simple2d *s) { return FaceGeo{s->lam, s->dot, s->d2}; }

fl_status continuity(fl_simple2d *s, double *out) {
    fl_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    k_continuity<<<s->blocks_c, kBlock, 0, st>>>(geo(s->mesh), pattern(s), s->flux, s->part);
    k_continuity_fold<<<1, kBlock, 0, st>>>((int)s->blocks_c, s->nc, s->part, s->rms);
    FL_HIP_TRY(ctx, hipGetLastError());
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->h_err, s->err, E_COUNT * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    FL_HIP_TRY(ctx, hipMemcpyAsync(out, s->rms, sizeof(double), hipMemcpyDeviceToHost, st));
    FL_HIP_TRY(ctx, hipStreamSynchronize(st));
    return FL_OK;
}

fl_status linear_error(fl_simple2d *s, fl_simple2d_report *rep, const char *what) {
    return simple_error(s->ctx, rep, FL_SIMPLE_LINEAR, 0, 0.0, what);
}

}  // namespace

extern "C" {

fl_status fl_simple2d_create(fl_mesh2d *mesh, const int32_t *velocity_kind, const double *velocity_x,
                             const double *velocity_y, const int32_t *pressure_kind,
                             const double *pressure_value, const int32_t *correction_kind,
                             const double *correction_value, fl_simple2d **out, fl_simple2d_report *erep) {
    if (!mesh || !out) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = mesh->ctx;
    *out = nullptr;
    if (erep) std::memset(erep, 0, sizeof(*erep));
    if (!velocity_kind || !velocity_x || !velocity_y || !pressure_kind || !pressure_value ||
        (correction_kind && !correction_value))
        return set_error(ctx, FL_ERR_INVALID_ARGUMENT, "null boundary array");
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const int nc = (int)mesh->ncells, nf = (int)mesh->nfaces;
    // host copies of the topology for the boundary checks
    std::vector<int> hneigh(nf);
    FL_HIP_TRY(ctx, hipMemcpy(hneigh.data(), mesh->neigh, nf * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int8_t> vk(nf, BC_NONE), pk(nf, BC_NONE), ck(nf, BC_NONE);
    bool anchored = false;
    for (int f = 0; f < nf; ++f) {
        if (hneigh[f] >= 0) continue;
        const int32_t kinds[3] = {velocity_kind[f], pressure_kind[f], correction_kind ? correction_kind[f] : 0};
        for (int k = 0; k < 3; ++k)
            if (kinds[k] != 0 && kinds[k] != 1)
                return simple_error(ctx, erep, FL_SIMPLE_MISSING_BOUNDARY, f, 0.0, "missing boundary condition");
        vk[f] = (int8_t)velocity_kind[f];
        pk[f] = (int8_t)pressure_kind[f];
        ck[f] = (int8_t)kinds[2];
        if ((vk[f] == BC_FIXED && !(std::isfinite(velocity_x[f]) && std::isfinite(velocity_y[f]))) ||
            (pk[f] == BC_FIXED && !std::isfinite(pressure_value[f])) ||
            (correction_kind && ck[f] == BC_FIXED && !std::isfinite(correction_value[f])))
            return simple_error(ctx, erep, FL_SIMPLE_MISSING_BOUNDARY, f, 0.0, "non-finite boundary value");
        anchored = anchored || (correction_kind && ck[f] == BC_FIXED);
    }
    fl_simple2d *s = new (std::nothrow) fl_simple2d();
    if (!s) return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "host allocation failed");
    s->ctx = ctx;
    s->mesh = mesh;
    s->nc = nc;
    s->nf = nf;
    s->has_corr = correction_kind != nullptr;
    s->anchored = anchored;
    s->blocks_c = blocks(ctx, nc);
    s->blocks_f = blocks(ctx, nf);
    const int ns = (int)mesh->nemit;
    const size_t NC = nc, NF = nf;
    bool ok = s->get(&s->vkind, NF) && s->get(&s->pkind, NF) && s->get(&s->ckind, NF) && s->get(&s->vx, NF) &&
              s->get(&s->vy, NF) && s->get(&s->cval, NF) && s->get(&s->pval, NF) && s->get(&s->lam, NF) &&
              s->get(&s->dot, NF) && s->get(&s->d2, NF) && s->get(&s->ekind, (size_t)ns) &&
              s->get(&s->enb, (size_t)ns) && s->get(&s->eval, (size_t)ns) && s->get(&s->rcx, (size_t)ns) &&
              s->get(&s->rcy, (size_t)ns) && s->get(&s->inv, 4 * NC) && s->get(&s->dpos, NC) &&
              s->get(&s->sfaces, (size_t)ns) && s->get(&s->ux, NC) && s->get(&s->uy, NC) && s->get(&s->p, NC) &&
              s->get(&s->flux, NF) && s->get(&s->gx, NC) && s->get(&s->gy, NC) && s->get(&s->usx, NC) &&
              s->get(&s->usy, NC) && s->get(&s->rhsx, NC) && s->get(&s->rhsy, NC) && s->get(&s->rau, NC) &&
              s->get(&s->pred, NF) && s->get(&s->coef, NF) && s->get(&s->pcx, NC) && s->get(&s->prhs, NC) &&
              s->get(&s->part, (size_t)s->blocks_c) && s->get(&s->rms, 1) && s->get(&s->wall, NF) &&
              s->get(&s->err, E_COUNT) &&
              hipHostMalloc(&s->h_err, E_COUNT * sizeof(unsigned long long), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        delete s;
        return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "device allocation failed");
    }
    hipStream_t st = ctx->stream;
    std::vector<double> hv(NF, 0.0);
    auto up8 = [&](int8_t *dst, const std::vector<int8_t> &src) {
        return hipMemcpyAsync(dst, src.data(), NF, hipMemcpyHostToDevice, st) == hipSuccess;
    };
    ok = up8(s->vkind, vk) && up8(s->pkind, pk) && up8(s->ckind, ck) &&
         hipMemcpyAsync(s->vx, velocity_x, NF * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(s->vy, velocity_y, NF * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(s->pval, pressure_value, NF * 8, hipMemcpyHostToDevice, st) == hipSuccess &&
         hipMemcpyAsync(s->cval, correction_kind ? correction_value : hv.data(), NF * 8, hipMemcpyHostToDevice,
                        st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    if (!ok) {
        delete s;
        return set_error(ctx, FL_ERR_HIP, "upload failed");
    }
    const Geo g = geo(mesh);
    // face geometry, LSQ stencil and sparsity
    k_fill_u64<<<1, 64, 0, st>>>(s->err, E_COUNT, ~0ull);
    k_face_geom<<<s->blocks_f, kBlock, 0, st>>>(g, s->lam, s->dot, s->d2, s->err + E_MOM_FACE);
    k_lsq_build<<<s->blocks_c, kBlock, 0, st>>>(g, s->pkind, s->pval, s->ekind, s->enb, s->eval, s->rcx, s->rcy,
                                                s->inv, s->err);
    int *count = nullptr;
    ok = s->get(&count, NC);
    if (ok) {
        k_pattern_count<<<s->blocks_c, kBlock, 0, st>>>(g, count);
        ok = hipGetLastError() == hipSuccess;
    }
    std::vector<int> hcount(NC), hoff(NC + 1, 0);
    ok = ok && hipMemcpyAsync(hcount.data(), count, NC * sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipMemcpyAsync(s->h_err, s->err, E_COUNT * 8, hipMemcpyDeviceToHost, st) == hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    if (!ok) {
        delete s;
        return set_error(ctx, FL_ERR_HIP, "SIMPLE setup failed");
    }
    if (s->h_err[E_MOM_FACE] != ~0ull) {
        const int64_t f = (int64_t)s->h_err[E_MOM_FACE];
        delete s;
        return simple_error(ctx, erep, FL_SIMPLE_DEGENERATE_DISTANCE, f, 0.0, "degenerate owner-neighbour distance");
    }
    if (s->h_err[E_SINGULAR] != ~0ull) {
        const int64_t c = (int64_t)s->h_err[E_SINGULAR];
        delete s;
        return simple_error(ctx, erep, FL_SIMPLE_SINGULAR_STENCIL, c, 0.0, "singular least-squares stencil");
    }
    for (size_t c = 0; c < NC; ++c) hoff[c + 1] = hoff[c] + hcount[c];
    const int64_t nnz = hoff[NC];
    s->nnz = nnz;
    // two copies of the sparsity: the solver matrices each own theirs
    int *off1 = nullptr, *col1 = nullptr, *off2 = nullptr, *col2 = nullptr;
    double *val1 = nullptr, *val2 = nullptr;
    ok = hipMalloc(&off1, (NC + 1) * sizeof(int)) == hipSuccess && hipMalloc(&col1, nnz * sizeof(int)) == hipSuccess &&
         hipMalloc(&val1, nnz * sizeof(double)) == hipSuccess &&
         hipMalloc(&off2, (NC + 1) * sizeof(int)) == hipSuccess && hipMalloc(&col2, nnz * sizeof(int)) == hipSuccess &&
         hipMalloc(&val2, nnz * sizeof(double)) == hipSuccess;
    if (ok) {
        ok = hipMemcpyAsync(off1, hoff.data(), (NC + 1) * sizeof(int), hipMemcpyHostToDevice, st) == hipSuccess;
        k_pattern_fill<<<s->blocks_c, kBlock, 0, st>>>(g, off1, col1, s->dpos, s->sfaces);
        // a diagonal-only start value keeps the adopted matrices' Jacobi setup valid
        k_zero<<<blocks(ctx, nnz), kBlock, 0, st>>>(val1, (int)nnz);
        k_zero<<<blocks(ctx, nnz), kBlock, 0, st>>>(val2, (int)nnz);
        ok = ok && hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(off2, off1, (NC + 1) * sizeof(int), hipMemcpyDeviceToDevice, st) == hipSuccess &&
             hipMemcpyAsync(col2, col1, nnz * sizeof(int), hipMemcpyDeviceToDevice, st) == hipSuccess &&
             hipStreamSynchronize(st) == hipSuccess;
    }
    if (!ok) {
        for (void *b : {(void *)off1, (void *)col1, (void *)val1, (void *)off2, (void *)col2, (void *)val2})
            if (b) (void)hipFree(b);
        delete s;
        return set_error(ctx, FL_ERR_HIP, "SIMPLE sparsity setup failed");
    }
    fl_status r1 = csr_adopt(ctx, (uint32_t)nc, (uint32_t)nc, nnz, off1, col1, val1, &s->mom);
    fl_status r2 = r1 == FL_OK ? csr_adopt(ctx, (uint32_t)nc, (uint32_t)nc, nnz, off2, col2, val2, &s->pcm)
                               : FL_ERR_OUT_OF_MEMORY;
    if (r1 != FL_OK || r2 != FL_OK) {
        if (r1 != FL_OK) {
            (void)hipFree(off2);
            (void)hipFree(col2);
            (void)hipFree(val2);
        }
        delete s;
        return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "solver matrix setup failed");
    }
    s->roff = off1;
    s->col = col1;
    // zero state
    for (double *v : {s->ux, s->uy, s->p, s->gx, s->gy, s->pcx}) k_zero<<<s->blocks_c, kBlock, 0, st>>>(v, nc);
    k_zero<<<s->blocks_f, kBlock, 0, st>>>(s->flux, nf);
    FL_HIP_TRY(ctx, hipStreamSynchronize(st));
    ctx_retain(ctx);
    *out = s;
    return FL_OK;
}

void fl_simple2d_destroy(fl_simple2d *s) {
    if (!s) return;
    fl_ctx *ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete s;
    ctx_release(ctx);
}

fl_status fl_simple2d_initial_face_flux(fl_simple2d *s, const double *ux, const double *uy,
                                        const double *wall_flux, double *face_flux) {
    if (!s || !ux || !uy || !wall_flux || !face_flux) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->gx, ux, s->nc * 8, hipMemcpyHostToDevice, st));
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->gy, uy, s->nc * 8, hipMemcpyHostToDevice, st));
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->wall, wall_flux, s->nf * 8, hipMemcpyHostToDevice, st));
    k_initial_flux<<<s->blocks_f, kBlock, 0, st>>>(geo(s->mesh), facegeo(s), s->gx, s->gy, s->wall, s->pred);
    FL_HIP_TRY(ctx, hipGetLastError());
    FL_HIP_TRY(ctx, hipMemcpyAsync(face_flux, s->pred, s->nf * 8, hipMemcpyDeviceToHost, st));
    FL_HIP_TRY(ctx, hipStreamSynchronize(st));
    return FL_OK;
}

fl_status fl_simple2d_set_state(fl_simple2d *s, const double *ux, const double *uy, const double *p,
                                const double *face_flux) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const double *src[4] = {ux, uy, p, face_flux};
    double *dst[4] = {s->ux, s->uy, s->p, s->flux};
    const size_t n[4] = {(size_t)s->nc, (size_t)s->nc, (size_t)s->nc, (size_t)s->nf};
    for (int k = 0; k < 4; ++k) {
        if (!src[k]) continue;
        for (size_t i = 0; i < n[k]; ++i)
            if (!std::isfinite(src[k][i]))
                return set_error(ctx, FL_ERR_NON_FINITE, "non-finite SIMPLE state value (array %d, index %zu)", k, i);
        FL_HIP_TRY(ctx, hipMemcpyAsync(dst[k], src[k], n[k] * 8, hipMemcpyHostToDevice, st));
    }
    FL_HIP_TRY(ctx, hipStreamSynchronize(st));
    return FL_OK;
}

fl_status fl_simple2d_get_state(fl_simple2d *s, double *ux, double *uy, double *p, double *face_flux) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    double *dst[4] = {ux, uy, p, face_flux};
    const double *src[4] = {s->ux, s->uy, s->p, s->flux};
    const size_t n[4] = {(size_t)s->nc, (size_t)s->nc, (size_t)s->nc, (size_t)s->nf};
    for (int k = 0; k < 4; ++k)
        if (dst[k]) FL_HIP_TRY(ctx, hipMemcpyAsync(dst[k], src[k], n[k] * 8, hipMemcpyDeviceToHost, st));
    FL_HIP_TRY(ctx, hipStreamSynchronize(st));
    return FL_OK;
}

fl_status fl_simple2d_solve(fl_simple2d *s, const fl_simple2d_options *o, fl_simple2d_report *rep,
                            double *history) {
    if (!s || !o || !rep) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = s->ctx;
    hipStream_t st = ctx->stream;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::memset(rep, 0, sizeof(*rep));
    // solve_simple's checks (simple.rs:153-173), then the ones of the first
    // momentum and pressure-correction assemblies
    if (o->max_outer_iterations <= 0)
        return simple_error(ctx, rep, FL_SIMPLE_INVALID_OUTER_ITERATIONS, 0, 0.0, "invalid outer iterations");
    if (!(std::isfinite(o->pressure_relaxation) && o->pressure_relaxation > 0.0 && o->pressure_relaxation <= 1.0))
        return simple_error(ctx, rep, FL_SIMPLE_INVALID_PRESSURE_RELAXATION, 0, o->pressure_relaxation,
                            "invalid pressure relaxation");
    if (!(std::isfinite(o->continuity_absolute_tolerance) && o->continuity_absolute_tolerance >= 0.0))
        return simple_error(ctx, rep, FL_SIMPLE_INVALID_PRESSURE_RELAXATION, 0, o->continuity_absolute_tolerance,
                            "invalid continuity tolerance");
    if (!(std::isfinite(o->velocity_relaxation) && o->velocity_relaxation > 0.0 && o->velocity_relaxation <= 1.0) ||
        !(std::isfinite(o->viscosity) && o->viscosity >= 0.0))
        return simple_error(ctx, rep, FL_SIMPLE_INVALID_OPTIONS, 0, o->velocity_relaxation,
                            "invalid velocity relaxation or viscosity");
    if (s->has_corr && !s->anchored)
        return simple_error(ctx, rep, FL_SIMPLE_INVALID_PRESSURE_REFERENCE, 0, 0.0,
                            "no FixedValue pressure-correction face anchors the correction");
    if (!s->has_corr && (o->reference_cell < 0 || o->reference_cell >= s->nc))
        return simple_error(ctx, rep, FL_SIMPLE_INVALID_PRESSURE_REFERENCE, o->reference_cell, 0.0,
                            "invalid pressure reference cell");
    const Geo g = geo(s->mesh);
    const Pattern pt = pattern(s);
    const FaceGeo fg = facegeo(s);
    const Lsq L = lsq(s);
    const uint32_t Bc = s->blocks_c, Bf = s->blocks_f;
    const auto t0 = std::chrono::steady_clock::now();
    fl_status stt;
    double rms0 = 0.0;
    FL_HIP_TRY(ctx, hipMemsetAsync(s->err, 0xff, E_COUNT * 8, st));
    if ((stt = continuity(s, &rms0)) != FL_OK) return stt;
    if (!std::isfinite(rms0))
        return simple_error(ctx, rep, FL_SIMPLE_INVALID_PRESSURE_FACE_COEFFICIENT, 0, rms0, "non-finite continuity");
    rep->initial_continuity_rms = rms0;
    rep->final_continuity_rms = rms0;
    if (history) history[0] = rms0;
    const bool relaxed = o->velocity_relaxation < 1.0;
    const int ref = s->has_corr ? -1 : (int)o->reference_cell;
    double *mval = csr_values(s->mom), *pval = csr_values(s->pcm);
    for (int64_t it = 1; it <= o->max_outer_iterations; ++it) {
        FL_HIP_TRY(ctx, hipMemsetAsync(s->err, 0xff, E_COUNT * 8, st));
        // pressure gradient (the momentum source and the Rhie-Chow response)
        k_lsq_grad<<<Bc, kBlock, 0, st>>>(g, L, s->p, s->gx, s->gy);
        MomArgs ma{g, pt, fg, o->viscosity, o->velocity_relaxation, relaxed ? 1 : 0, s->vkind, s->vx, s->vy,
                   s->flux, s->gx, s->gy, s->ux, s->uy, mval, s->rhsx, s->rhsy, s->rau, s->err};
        k_momentum<<<Bc, kBlock, 0, st>>>(ma);
        FL_HIP_TRY(ctx, hipMemcpyAsync(s->usx, s->ux, s->nc * 8, hipMemcpyDeviceToDevice, st));
        FL_HIP_TRY(ctx, hipMemcpyAsync(s->usy, s->uy, s->nc * 8, hipMemcpyDeviceToDevice, st));
        FL_HIP_TRY(ctx, hipGetLastError());
        if ((stt = read_errors(s)) != FL_OK) return stt;
        if (s->h_err[E_SOURCE] != ~0ull)
            return simple_error(ctx, rep, FL_SIMPLE_NONFINITE_FLUX, (int64_t)s->h_err[E_SOURCE], 0.0,
                                "non-finite momentum source");
        if (s->h_err[E_MOM_FACE] != ~0ull) {
            const unsigned long long k = s->h_err[E_MOM_FACE];
            const int64_t f = (int64_t)(k >> 1);
            if (k & 1)
                return simple_error(ctx, rep, FL_SIMPLE_BACKFLOW, f, fetch(s, s->flux, f), "backflow on a zero-gradient face");
            return simple_error(ctx, rep, FL_SIMPLE_NONFINITE_FLUX, f, fetch(s, s->flux, f), "non-finite face flux");
        }
        // momentum solves (the reference's bicgstab errors come before the
        // convergence check and the diagonal check)
        fl_err_detail d{};
        if (csr_solve_device(s->mom, 2, s->rhsx, s->usx, &o->momentum_solver, &rep->momentum[0], &d) != FL_OK ||
            csr_solve_device(s->mom, 2, s->rhsy, s->usy, &o->momentum_solver, &rep->momentum[1], &d) != FL_OK)
            return linear_error(s, rep, "momentum BiCGSTAB failed");
        rep->momentum[2] = fl_lin_report{1, 0, 0, 0.0, 0.0, 0.0, 0.0};
        if (!rep->momentum[0].converged || !rep->momentum[1].converged)
            return simple_error(ctx, rep, FL_SIMPLE_MOMENTUM_DID_NOT_CONVERGE, 0, 0.0,
                                "momentum linear solve did not converge");
        if (s->h_err[E_RAU] != ~0ull) {
            const int64_t c = (int64_t)s->h_err[E_RAU];
            double dr = 0.0;
            int dp = 0;
            (void)hipMemcpy(&dp, s->dpos + c, sizeof(int), hipMemcpyDeviceToHost);
            dr = fetch(s, mval, dp);
            return simple_error(ctx, rep, FL_SIMPLE_INVALID_MOMENTUM_DIAGONAL, c, dr, "invalid momentum diagonal");
        }
        // Rhie-Chow flux, face coefficients, pressure-correction system
        RcArgs ra{g, fg, s->usx, s->usy, s->p, s->gx, s->gy, s->rau, s->flux, s->has_corr ? s->ckind : nullptr,
                  s->pred, s->coef, s->err};
        k_rhie_chow<<<Bf, kBlock, 0, st>>>(ra);
        PcArgs pa{g, pt, s->pred, s->coef, s->has_corr ? s->ckind : nullptr, s->cval, ref, pval, s->prhs};
        k_pcorr<<<Bc, kBlock, 0, st>>>(pa);
        k_zero<<<Bc, kBlock, 0, st>>>(s->pcx, s->nc);
        FL_HIP_TRY(ctx, hipGetLastError());
        if ((stt = read_errors(s)) != FL_OK) return stt;
        for (int slot : {(int)E_COEF_IN, (int)E_RC_FLUX, (int)E_COEF_BD})
            if (s->h_err[slot] != ~0ull) {
                const int64_t f = (int64_t)s->h_err[slot];
                return simple_error(ctx, rep, FL_SIMPLE_INVALID_PRESSURE_FACE_COEFFICIENT, f,
                                    fetch(s, slot == E_RC_FLUX ? s->pred : s->coef, f),
                                    "invalid pressure face coefficient");
            }
        if ((stt = csr_refresh_jacobi(s->pcm)) != FL_OK) return stt;
        if (csr_solve_device(s->pcm, 1, s->prhs, s->pcx, &o->pressure_solver, &rep->pressure, &d) != FL_OK)
            return linear_error(s, rep, "pressure-correction PCG failed");
        if (!rep->pressure.converged)
            return simple_error(ctx, rep, FL_SIMPLE_PRESSURE_DID_NOT_CONVERGE, 0, 0.0,
                                "pressure linear solve did not converge");
        // corrections
        k_update_p<<<Bc, kBlock, 0, st>>>(s->nc, o->pressure_relaxation, s->pcx, s->p);
        k_lsq_grad<<<Bc, kBlock, 0, st>>>(g, L, s->pcx, s->gx, s->gy);
        k_correct_velocity<<<Bc, kBlock, 0, st>>>(s->nc, s->usx, s->usy, s->rau, s->gx, s->gy, s->ux, s->uy, s->err);
        k_correct_flux<<<Bf, kBlock, 0, st>>>(g, s->pred, s->coef, s->pcx, s->has_corr ? s->ckind : nullptr, s->cval,
                                              s->flux, s->err);
        FL_HIP_TRY(ctx, hipGetLastError());
        double rms = 0.0;
        if ((stt = continuity(s, &rms)) != FL_OK) return stt;
        if (s->h_err[E_VELOCITY] != ~0ull) {
            const int64_t c = (int64_t)s->h_err[E_VELOCITY];
            return simple_error(ctx, rep, FL_SIMPLE_INVALID_MOMENTUM_DIAGONAL, c, fetch(s, s->rau, c),
                                "non-finite corrected velocity");
        }
        if (s->h_err[E_FLUX] != ~0ull) {
            const int64_t f = (int64_t)s->h_err[E_FLUX];
            return simple_error(ctx, rep, FL_SIMPLE_INVALID_PRESSURE_FACE_COEFFICIENT, f, fetch(s, s->flux, f),
                                "non-finite corrected flux");
        }
        if (!std::isfinite(rms))
            return simple_error(ctx, rep, FL_SIMPLE_INVALID_PRESSURE_FACE_COEFFICIENT, 0, rms, "non-finite continuity");
        if (history) history[it] = rms;
        rep->final_continuity_rms = rms;
        rep->outer_iterations = it;
        if (rms <= o->continuity_absolute_tolerance) {
            rep->converged = 1;
            break;
        }
    }
    rep->device_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FL_OK;
}

}  // extern "C"
