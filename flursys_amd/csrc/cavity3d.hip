// 3D lid-driven cavity projection step on gfx950.
//
// Replaces LidDrivenCavity3DSolver (src/solver/incompressible_3d.rs):
//   predict_momentum :174-273   -> k_predict
//   build_pressure_rhs :275-291 -> k_rhs
//   solve_pressure_poisson :293-332 -> pinned Jacobi-PCG (k_pcg_*); the
//       reference only has SOR here, the PCG loop is the 2D solver's
//       (src/solver/incompressible.rs:1048-1214) on the 7-point operator with
//       cell (0,0,0) eliminated (src/poisson.rs:303-337 style)
//   correct_velocity :334-363   -> k_correct
//   max_divergence / max_speed / ensure_finite :504-554 -> k_diag, k_finite
//
// Data layout in HBM: the reference's own (src/field.rs:104-107), one flat
// array per field, x fastest.  The PCG vectors x (= p), r, d, q live beside
// the fields; z = D^-1 r is never stored (recomputed from r where needed).
//
// Per PCG iteration three bandwidth-bound kernels run, each a grid-stride
// loop over 256-cell tiles with an XCD-aware slab mapping and a one-pass
// grid reduction finished by the last block:
//   k_pcg_apply   q = A d,  d.q           -> alpha       (reads d, writes q)
//   k_pcg_update  x += a d, r -= a q,
//                 r.z, r.r, max|r|        -> stop test, beta (reads x d r q, writes x r)
//   k_pcg_direction  d = D^-1 r + beta d                     (reads r d, writes d)
// 11 vector passes per iteration; all scalars (alpha, beta, r.z, the stop
// test) stay on the device, so the host only polls a small state block.

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"

namespace fl {
namespace c3 {

struct Grid {
    uint32_t nx, ny, nz, nxy, n;
    double ix2, iy2, iz2;
};

__device__ inline void cell_ijk(const Grid &g, uint32_t c, uint32_t &i, uint32_t &j,
                                uint32_t &k) {
    k = c / g.nxy;
    const uint32_t rem = c - k * g.nxy;
    j = rem / g.nx;
    i = rem - j * g.nx;
}

// Neumann operator row at an interior (non-pinned) cell: the reference's
// ghost = centre rule (pressure_neighbor, incompressible_3d.rs:450-474) is
// the same as dropping the missing neighbour from both diagonal and sum.
template <typename T>
__device__ inline T apply_row(const Grid &g, const T *__restrict__ x, uint32_t c, uint32_t i,
                              uint32_t j, uint32_t k, T xc, T &diag) {
    const T cx = (T)g.ix2, cy = (T)g.iy2, cz = (T)g.iz2;
    T dg = 0, s = 0;
    if (i + 1 < g.nx) { dg += cx; s += cx * x[c + 1]; }
    if (i > 0) { dg += cx; s += cx * x[c - 1]; }
    if (j + 1 < g.ny) { dg += cy; s += cy * x[c + g.nx]; }
    if (j > 0) { dg += cy; s += cy * x[c - g.nx]; }
    if (k + 1 < g.nz) { dg += cz; s += cz * x[c + g.nxy]; }
    if (k > 0) { dg += cz; s += cz * x[c - g.nxy]; }
    diag = dg;
    return dg * xc - s;
}

template <typename T>
__device__ inline T diag_at(const Grid &g, uint32_t i, uint32_t j, uint32_t k) {
    const T cx = (T)g.ix2, cy = (T)g.iy2, cz = (T)g.iz2;
    T dg = 0;
    if (i + 1 < g.nx) dg += cx;
    if (i > 0) dg += cx;
    if (j + 1 < g.ny) dg += cy;
    if (j > 0) dg += cy;
    if (k + 1 < g.nz) dg += cz;
    if (k > 0) dg += cz;
    return dg;
}

// Device-resident scalars of one PCG solve.
struct PcgState {
    double rz, alpha, beta;
    double rr0, thresh, tol;
    double rmax, rr;
    long long iter, max_iter;
    int done, converged, status, stop_rule;
    unsigned counters[4];
};
enum { ST_OK = 0, ST_SINGULAR = 1, ST_ALPHA = 2, ST_RESIDUAL = 3 };

// r = b - A x with b = -rhs (b_0 = 0), z = D^-1 r, d = z; r.r, r.z, max|r|.
__global__ void __launch_bounds__(kBlock) k_pcg_init(Grid g, const double *__restrict__ rhs,
                                                     const double *__restrict__ x,
                                                     double *__restrict__ r,
                                                     double *__restrict__ d, double *partials,
                                                     PcgState *st) {
    double s[2] = {0.0, 0.0}, m[1] = {0.0};
    const uint32_t ntiles = (g.n + kBlock - 1) / kBlock;
    TileSched ts(ntiles);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t c = t * kBlock + threadIdx.x;
        if (c >= g.n) continue;
        double rc = 0.0, z = 0.0;
        if (c != 0) {
            uint32_t i, j, k;
            cell_ijk(g, c, i, j, k);
            double dg;
            const double ax = apply_row<double>(g, x, c, i, j, k, x[c], dg);
            rc = -rhs[c] - ax;
            z = rc / dg;
        }
        r[c] = rc;
        d[c] = z;
        s[0] += rc * rc;
        s[1] += rc * z;
        m[0] = fmax(m[0], fabs(rc));
    }
    if (grid_reduce<2, 1>(s, m, partials, &st->counters[0]) && threadIdx.x == 0) {
        st->rr0 = s[0];
        st->rr = s[0];
        st->rz = s[1];
        st->rmax = m[0];
        st->iter = 0;
        st->thresh = st->tol * sqrt(s[0]);
        const bool conv = st->stop_rule == FL_STOP_REL_L2 ? (sqrt(s[0]) <= st->thresh)
                                                          : (m[0] < st->tol);
        st->converged = conv;
        st->done = conv;
        st->status = ST_OK;
    }
}

// q = A d and d.q; the last block forms alpha = (r.z) / (d.q).
__global__ void __launch_bounds__(kBlock) k_pcg_apply(Grid g, const double *__restrict__ d,
                                                      double *__restrict__ q, double *partials,
                                                      PcgState *st) {
    if (st->done) return;
    double s[1] = {0.0}, m[1] = {0.0};
    const uint32_t ntiles = (g.n + kBlock - 1) / kBlock;
    TileSched ts(ntiles);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t c = t * kBlock + threadIdx.x;
        if (c >= g.n) continue;
        const double dc = d[c];
        double qc = dc;
        if (c != 0) {
            uint32_t i, j, k;
            cell_ijk(g, c, i, j, k);
            double dg;
            qc = apply_row<double>(g, d, c, i, j, k, dc, dg);
        }
        q[c] = qc;
        s[0] += dc * qc;
    }
    if (grid_reduce<1, 1>(s, m, partials, &st->counters[1]) && threadIdx.x == 0) {
        const double den = s[0];
        if (fabs(den) < 1.0e-30 || !isfinite(den)) {
            st->status = ST_SINGULAR;
            st->done = 1;
            return;
        }
        const double alpha = st->rz / den;
        if (!isfinite(alpha)) {
            st->status = ST_ALPHA;
            st->done = 1;
            return;
        }
        st->alpha = alpha;
    }
}

// x += alpha d, r -= alpha q; r.z, r.r, max|r|; the last block runs the stop
// test and forms beta, in the reference's order (incompressible.rs:1141-1185).
__global__ void __launch_bounds__(kBlock) k_pcg_update(Grid g, double *__restrict__ x,
                                                       double *__restrict__ r,
                                                       const double *__restrict__ d,
                                                       const double *__restrict__ q,
                                                       double *partials, PcgState *st) {
    if (st->done) return;
    const double alpha = st->alpha;
    double s[2] = {0.0, 0.0}, m[1] = {0.0};
    const uint32_t ntiles = (g.n + kBlock - 1) / kBlock;
    TileSched ts(ntiles);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t c = t * kBlock + threadIdx.x;
        if (c >= g.n) continue;
        x[c] += alpha * d[c];
        const double rc = r[c] - alpha * q[c];
        r[c] = rc;
        double z = rc;
        if (c != 0) {
            uint32_t i, j, k;
            cell_ijk(g, c, i, j, k);
            z = rc / diag_at<double>(g, i, j, k);
        }
        s[0] += rc * z;
        s[1] += rc * rc;
        m[0] = fmax(m[0], fabs(rc));
    }
    if (grid_reduce<2, 1>(s, m, partials, &st->counters[2]) && threadIdx.x == 0) {
        st->iter += 1;
        st->rmax = m[0];
        st->rr = s[1];
        if (!isfinite(m[0])) {
            st->status = ST_RESIDUAL;
            st->done = 1;
            return;
        }
        const bool conv = st->stop_rule == FL_STOP_REL_L2 ? (sqrt(s[1]) <= st->thresh)
                                                          : (m[0] < st->tol);
        if (conv) {
            st->converged = 1;
            st->done = 1;
            return;
        }
        if (fabs(st->rz) < 1.0e-30) {  // the reference's silent break (:1171-1173)
            st->done = 1;
            return;
        }
        st->beta = s[0] / st->rz;
        st->rz = s[0];
        if (st->iter >= st->max_iter) st->done = 1;
    }
}

// d = D^-1 r + beta d.
__global__ void __launch_bounds__(kBlock) k_pcg_direction(Grid g, const double *__restrict__ r,
                                                          double *__restrict__ d,
                                                          const PcgState *st) {
    if (st->done) return;
    const double beta = st->beta;
    const uint32_t ntiles = (g.n + kBlock - 1) / kBlock;
    TileSched ts(ntiles);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t c = t * kBlock + threadIdx.x;
        if (c >= g.n) continue;
        double z = r[c];
        if (c != 0) {
            uint32_t i, j, k;
            cell_ijk(g, c, i, j, k);
            z = z / diag_at<double>(g, i, j, k);
        }
        d[c] = z + beta * d[c];
    }
}

// ---------------------------------------------------------------------------
// Projection-step kernels (one thread per face or cell, grid-stride).
// ---------------------------------------------------------------------------
struct StepParams {
    uint32_t nx, ny, nz;
    double dx, dy, dz, dt, nu, lid, rho;
};

#define U3(i, j, k) ((i) + (P.nx + 1) * ((j) + P.ny * (k)))
#define V3(i, j, k) ((i) + P.nx * ((j) + (P.ny + 1) * (k)))
#define W3(i, j, k) ((i) + P.nx * ((j) + P.ny * (k)))

// u/v/w_neighbor (incompressible_3d.rs:366-447).
__device__ inline double u_nb(const StepParams &P, const double *u, int i, int j, int k, int di,
                              int dj, int dk, double c) {
    const int a = i + di, b = j + dj, e = k + dk;
    if (a >= 0 && a <= (int)P.nx && b >= 0 && b < (int)P.ny && e >= 0 && e < (int)P.nz)
        return u[U3(a, b, e)];
    return b >= (int)P.ny ? 2.0 * P.lid - c : -c;
}
__device__ inline double v_nb(const StepParams &P, const double *v, int i, int j, int k, int di,
                              int dj, int dk, double c) {
    const int a = i + di, b = j + dj, e = k + dk;
    if (a >= 0 && a < (int)P.nx && b >= 0 && b <= (int)P.ny && e >= 0 && e < (int)P.nz)
        return v[V3(a, b, e)];
    return -c;
}
__device__ inline double w_nb(const StepParams &P, const double *w, int i, int j, int k, int di,
                              int dj, int dk, double c) {
    const int a = i + di, b = j + dj, e = k + dk;
    if (a >= 0 && a < (int)P.nx && b >= 0 && b < (int)P.ny && e >= 0 && e <= (int)P.nz)
        return w[W3(a, b, e)];
    return -c;
}

// predict_momentum (incompressible_3d.rs:174-273): central, advective form.
// Boundary-normal faces are written 0 (the reference zero-fills u*, v*, w*).
__global__ void __launch_bounds__(kBlock) k_predict(StepParams P, const double *__restrict__ u,
                                                    const double *__restrict__ v,
                                                    const double *__restrict__ w,
                                                    double *__restrict__ us,
                                                    double *__restrict__ vs,
                                                    double *__restrict__ ws) {
    const uint32_t nu = (P.nx + 1) * P.ny * P.nz, nv = P.nx * (P.ny + 1) * P.nz,
                   nw = P.nx * P.ny * (P.nz + 1);
    const double dx2 = P.dx * P.dx, dy2 = P.dy * P.dy, dz2 = P.dz * P.dz;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nu + nv + nw;
         t += gridDim.x * blockDim.x) {
        if (t < nu) {
            const int i = t % (P.nx + 1), j = (t / (P.nx + 1)) % P.ny, k = t / ((P.nx + 1) * P.ny);
            if (i == 0 || i == (int)P.nx) { us[t] = 0.0; continue; }
            const double c = u[t];
            const double e = u_nb(P, u, i, j, k, 1, 0, 0, c), wv = u_nb(P, u, i, j, k, -1, 0, 0, c);
            const double n = u_nb(P, u, i, j, k, 0, 1, 0, c), so = u_nb(P, u, i, j, k, 0, -1, 0, c);
            const double f = u_nb(P, u, i, j, k, 0, 0, 1, c), bk = u_nb(P, u, i, j, k, 0, 0, -1, c);
            const double vv = 0.25 * (v[V3(i - 1, j, k)] + v[V3(i, j, k)] + v[V3(i - 1, j + 1, k)] +
                                      v[V3(i, j + 1, k)]);
            const double ww = 0.25 * (w[W3(i - 1, j, k)] + w[W3(i, j, k)] + w[W3(i - 1, j, k + 1)] +
                                      w[W3(i, j, k + 1)]);
            us[t] = c + P.dt * (-c * (e - wv) / (2.0 * P.dx) - vv * (n - so) / (2.0 * P.dy) -
                                ww * (f - bk) / (2.0 * P.dz) +
                                P.nu * ((e - 2.0 * c + wv) / dx2 + (n - 2.0 * c + so) / dy2 +
                                        (f - 2.0 * c + bk) / dz2));
        } else if (t < nu + nv) {
            const uint32_t s = t - nu;
            const int i = s % P.nx, j = (s / P.nx) % (P.ny + 1), k = s / (P.nx * (P.ny + 1));
            if (j == 0 || j == (int)P.ny) { vs[s] = 0.0; continue; }
            const double c = v[s];
            const double e = v_nb(P, v, i, j, k, 1, 0, 0, c), wv = v_nb(P, v, i, j, k, -1, 0, 0, c);
            const double n = v_nb(P, v, i, j, k, 0, 1, 0, c), so = v_nb(P, v, i, j, k, 0, -1, 0, c);
            const double f = v_nb(P, v, i, j, k, 0, 0, 1, c), bk = v_nb(P, v, i, j, k, 0, 0, -1, c);
            const double uu = 0.25 * (u[U3(i, j - 1, k)] + u[U3(i + 1, j - 1, k)] + u[U3(i, j, k)] +
                                      u[U3(i + 1, j, k)]);
            const double ww = 0.25 * (w[W3(i, j - 1, k)] + w[W3(i, j, k)] + w[W3(i, j - 1, k + 1)] +
                                      w[W3(i, j, k + 1)]);
            vs[s] = c + P.dt * (-uu * (e - wv) / (2.0 * P.dx) - c * (n - so) / (2.0 * P.dy) -
                                ww * (f - bk) / (2.0 * P.dz) +
                                P.nu * ((e - 2.0 * c + wv) / dx2 + (n - 2.0 * c + so) / dy2 +
                                        (f - 2.0 * c + bk) / dz2));
        } else {
            const uint32_t s = t - nu - nv;
            const int i = s % P.nx, j = (s / P.nx) % P.ny, k = s / (P.nx * P.ny);
            if (k == 0 || k == (int)P.nz) { ws[s] = 0.0; continue; }
            const double c = w[s];
            const double e = w_nb(P, w, i, j, k, 1, 0, 0, c), wv = w_nb(P, w, i, j, k, -1, 0, 0, c);
            const double n = w_nb(P, w, i, j, k, 0, 1, 0, c), so = w_nb(P, w, i, j, k, 0, -1, 0, c);
            const double f = w_nb(P, w, i, j, k, 0, 0, 1, c), bk = w_nb(P, w, i, j, k, 0, 0, -1, c);
            const double uu = 0.25 * (u[U3(i, j, k - 1)] + u[U3(i + 1, j, k - 1)] + u[U3(i, j, k)] +
                                      u[U3(i + 1, j, k)]);
            const double vv = 0.25 * (v[V3(i, j, k - 1)] + v[V3(i, j + 1, k - 1)] + v[V3(i, j, k)] +
                                      v[V3(i, j + 1, k)]);
            ws[s] = c + P.dt * (-uu * (e - wv) / (2.0 * P.dx) - vv * (n - so) / (2.0 * P.dy) -
                                c * (f - bk) / (2.0 * P.dz) +
                                P.nu * ((e - 2.0 * c + wv) / dx2 + (n - 2.0 * c + so) / dy2 +
                                        (f - 2.0 * c + bk) / dz2));
        }
    }
}

// build_pressure_rhs (incompressible_3d.rs:275-291), rhs(0,0,0) = 0.
__global__ void __launch_bounds__(kBlock) k_rhs(StepParams P, const double *__restrict__ us,
                                                const double *__restrict__ vs,
                                                const double *__restrict__ ws,
                                                double *__restrict__ rhs) {
    const uint32_t n = P.nx * P.ny * P.nz;
    const double scale = P.rho / P.dt;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const int i = c % P.nx, j = (c / P.nx) % P.ny, k = c / (P.nx * P.ny);
        rhs[c] = c == 0 ? 0.0
                        : scale * (us[U3(i + 1, j, k)] - us[U3(i, j, k)]) / P.dx +
                              scale * (vs[V3(i, j + 1, k)] - vs[V3(i, j, k)]) / P.dy +
                              scale * (ws[W3(i, j, k + 1)] - ws[W3(i, j, k)]) / P.dz;
    }
}

// correct_velocity (incompressible_3d.rs:334-363): interior faces only.
__global__ void __launch_bounds__(kBlock) k_correct(StepParams P, const double *__restrict__ p,
                                                    const double *__restrict__ us,
                                                    const double *__restrict__ vs,
                                                    const double *__restrict__ ws,
                                                    double *__restrict__ u,
                                                    double *__restrict__ v,
                                                    double *__restrict__ w) {
    const uint32_t nu = (P.nx + 1) * P.ny * P.nz, nv = P.nx * (P.ny + 1) * P.nz,
                   nw = P.nx * P.ny * (P.nz + 1);
    const double scale = P.dt / P.rho;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nu + nv + nw;
         t += gridDim.x * blockDim.x) {
        if (t < nu) {
            const int i = t % (P.nx + 1), j = (t / (P.nx + 1)) % P.ny, k = t / ((P.nx + 1) * P.ny);
            if (i == 0 || i == (int)P.nx) continue;
            u[t] = us[t] - scale * (p[W3(i, j, k)] - p[W3(i - 1, j, k)]) / P.dx;
        } else if (t < nu + nv) {
            const uint32_t s = t - nu;
            const int i = s % P.nx, j = (s / P.nx) % (P.ny + 1), k = s / (P.nx * (P.ny + 1));
            if (j == 0 || j == (int)P.ny) continue;
            v[s] = vs[s] - scale * (p[W3(i, j, k)] - p[W3(i, j - 1, k)]) / P.dy;
        } else {
            const uint32_t s = t - nu - nv;
            const int i = s % P.nx, j = (s / P.nx) % P.ny, k = s / (P.nx * P.ny);
            if (k == 0 || k == (int)P.nz) continue;
            w[s] = ws[s] - scale * (p[W3(i, j, k)] - p[W3(i, j, k - 1)]) / P.dz;
        }
    }
}

struct DiagState {
    double div_l2sq, max_div, max_speed;
    unsigned long long first_bad[4];  // first non-finite flat index per field
    unsigned counter;
};

// max_divergence (:504-517), max_speed (:519-532) and sum(div^2).
__global__ void __launch_bounds__(kBlock) k_diag(StepParams P, const double *__restrict__ u,
                                                 const double *__restrict__ v,
                                                 const double *__restrict__ w,
                                                 double *partials, DiagState *dg) {
    const uint32_t n = P.nx * P.ny * P.nz;
    double s[1] = {0.0}, m[2] = {0.0, 0.0};
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        const int i = c % P.nx, j = (c / P.nx) % P.ny, k = c / (P.nx * P.ny);
        const double u0 = u[U3(i, j, k)], u1 = u[U3(i + 1, j, k)];
        const double v0 = v[V3(i, j, k)], v1 = v[V3(i, j + 1, k)];
        const double w0 = w[W3(i, j, k)], w1 = w[W3(i, j, k + 1)];
        const double div = (u1 - u0) / P.dx + (v1 - v0) / P.dy + (w1 - w0) / P.dz;
        s[0] += div * div;
        m[0] = fmax(m[0], fabs(div));
        const double a = 0.5 * (u0 + u1), b = 0.5 * (v0 + v1), cc = 0.5 * (w0 + w1);
        m[1] = fmax(m[1], sqrt(a * a + b * b + cc * cc));
    }
    if (grid_reduce<1, 2>(s, m, partials, &dg->counter) && threadIdx.x == 0) {
        dg->div_l2sq = s[0];
        dg->max_div = m[0];
        dg->max_speed = m[1];
    }
}

// ensure_finite (:534-554): first non-finite index of p, u, v, w.
__global__ void __launch_bounds__(kBlock) k_finite(const double *__restrict__ a0, uint32_t n0,
                                                   const double *__restrict__ a1, uint32_t n1,
                                                   const double *__restrict__ a2, uint32_t n2,
                                                   const double *__restrict__ a3, uint32_t n3,
                                                   DiagState *dg) {
    const double *arr[4] = {a0, a1, a2, a3};
    const uint32_t len[4] = {n0, n1, n2, n3};
    for (int f = 0; f < 4; ++f)
        for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < len[f];
             c += gridDim.x * blockDim.x)
            if (!isfinite(arr[f][c])) atomicMin(&dg->first_bad[f], (unsigned long long)c);
}

#undef U3
#undef V3
#undef W3

}  // namespace c3
}  // namespace fl

using namespace fl;
using namespace fl::c3;

struct fl_cavity3d {
    fl_ctx *ctx = nullptr;
    fl_cavity3d_config cfg{};
    Grid g{};
    StepParams P{};
    size_t nu = 0, nv = 0, nw = 0, n = 0;
    double *u = nullptr, *v = nullptr, *w = nullptr, *us = nullptr, *vs = nullptr,
           *ws = nullptr, *p = nullptr, *rhs = nullptr, *r = nullptr, *d = nullptr,
           *q = nullptr, *partials = nullptr;
    PcgState *st = nullptr, *st_host = nullptr;
    DiagState *dg = nullptr, *dg_host = nullptr;
    uint32_t blocks = 1, ew_blocks = 1;
    int64_t step = 0;
    // per-kernel timing (HIP events on the solver's stream)
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double t_apply = 0, t_update = 0, t_dir = 0;
    int64_t t_iters = 0;
};

namespace {

constexpr int kMaxBatch = 64;

fl_status alloc(fl_cavity3d *s, double **ptr, size_t count) {
    FL_HIP_TRY(s->ctx, hipMalloc(ptr, count * sizeof(double)));
    FL_HIP_TRY(s->ctx, hipMemsetAsync(*ptr, 0, count * sizeof(double), s->ctx->stream));
    return FL_OK;
}

void free_all(fl_cavity3d *s) {
    double *bufs[] = {s->u, s->v, s->w, s->us, s->vs, s->ws, s->p,
                      s->rhs, s->r, s->d, s->q, s->partials};
    for (double *b : bufs)
        if (b) (void)hipFree(b);
    if (s->st) (void)hipFree(s->st);
    if (s->dg) (void)hipFree(s->dg);
    if (s->st_host) (void)hipHostFree(s->st_host);
    if (s->dg_host) (void)hipHostFree(s->dg_host);
    for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
    s->ev.clear();
}

fl_status validate(fl_ctx *ctx, const fl_cavity3d_config *c) {
    // LidDrivenCavity3DConfig::validate (incompressible_3d.rs:49-84) and
    // UniformGrid3D::new (src/grid.rs:74-84).
    if (c->nx < 4 || c->ny < 4 || c->nz < 4)
        return set_error(ctx, FL_ERR_CONFIG, "nx, ny, and nz must all be at least 4");
    if (!(c->length > 0.0 && c->height > 0.0 && c->depth > 0.0))
        return set_error(ctx, FL_ERR_CONFIG, "Domain dimensions must be positive");
    if (!(std::isfinite(c->density) && c->density > 0.0))
        return set_error(ctx, FL_ERR_CONFIG, "density must be finite and positive");
    if (!(std::isfinite(c->lid_velocity) && c->lid_velocity > 0.0))
        return set_error(ctx, FL_ERR_CONFIG, "lid_velocity must be finite and positive");
    if (!(std::isfinite(c->reynolds) && c->reynolds > 0.0))
        return set_error(ctx, FL_ERR_CONFIG, "reynolds must be finite and positive");
    if (!(std::isfinite(c->dt) && c->dt > 0.0))
        return set_error(ctx, FL_ERR_CONFIG, "dt must be finite and positive");
    if (c->pressure_max_iters <= 0)
        return set_error(ctx, FL_ERR_CONFIG, "iteration limits must be positive");
    if (!(std::isfinite(c->pressure_tolerance) && c->pressure_tolerance > 0.0))
        return set_error(ctx, FL_ERR_CONFIG, "pressure_tolerance must be finite and positive");
    if (c->stop_rule != FL_STOP_ABS_INF && c->stop_rule != FL_STOP_REL_L2)
        return set_error(ctx, FL_ERR_CONFIG, "unknown stop rule %d", c->stop_rule);
    if (c->initial_guess != FL_GUESS_PREVIOUS && c->initial_guess != FL_GUESS_ZERO)
        return set_error(ctx, FL_ERR_CONFIG, "unknown initial guess %d", c->initial_guess);
    if (c->precision != FL_PRECISION_FP64)
        return set_error(ctx, FL_ERR_CONFIG, "unsupported precision %d", c->precision);
    const double cells = (double)c->nx * (double)c->ny * (double)c->nz;
    const double faces = (double)(c->nx + 1) * (double)(c->ny + 1) * (double)(c->nz + 1) * 3.0;
    if (cells + faces >= 4294967295.0)
        return set_error(ctx, FL_ERR_CONFIG, "grid too large for 32-bit cell indexing");
    return FL_OK;
}

const char *status_text(int status) {
    switch (status) {
        case ST_SINGULAR: return "PCG pressure solver encountered a singular search direction";
        case ST_ALPHA: return "PCG pressure solver produced a non-finite step length";
        case ST_RESIDUAL: return "PCG pressure solver produced a non-finite residual";
        default: return "";
    }
}

// Pinned Jacobi-PCG on the device.  p holds the guess on entry and the
// solution on exit; rhs is the device right-hand side.
fl_status solve_pressure(fl_cavity3d *s, fl_solve_report *rep) {
    fl_ctx *ctx = s->ctx;
    hipStream_t stm = ctx->stream;
    if (s->cfg.initial_guess == FL_GUESS_ZERO)
        FL_HIP_TRY(ctx, hipMemsetAsync(s->p, 0, s->n * sizeof(double), stm));
    else
        FL_HIP_TRY(ctx, hipMemsetAsync(s->p, 0, sizeof(double), stm));  // pin p(0,0,0) = 0
    PcgState init{};
    init.tol = s->cfg.pressure_tolerance;
    init.stop_rule = s->cfg.stop_rule;
    init.max_iter = s->cfg.pressure_max_iters;
    *s->st_host = init;
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->st, s->st_host, sizeof(PcgState), hipMemcpyHostToDevice, stm));
    k_pcg_init<<<s->blocks, kBlock, 0, stm>>>(s->g, s->rhs, s->p, s->r, s->d, s->partials, s->st);
    FL_HIP_TRY(ctx, hipGetLastError());
    int batch = 8;
    long long done_iters = 0;
    for (;;) {
        for (int b = 0; b < batch; ++b) {
            if (s->timing) FL_HIP_TRY(ctx, hipEventRecord(s->ev[4 * b], stm));
            k_pcg_apply<<<s->blocks, kBlock, 0, stm>>>(s->g, s->d, s->q, s->partials, s->st);
            if (s->timing) FL_HIP_TRY(ctx, hipEventRecord(s->ev[4 * b + 1], stm));
            k_pcg_update<<<s->blocks, kBlock, 0, stm>>>(s->g, s->p, s->r, s->d, s->q,
                                                        s->partials, s->st);
            if (s->timing) FL_HIP_TRY(ctx, hipEventRecord(s->ev[4 * b + 2], stm));
            k_pcg_direction<<<s->blocks, kBlock, 0, stm>>>(s->g, s->r, s->d, s->st);
            if (s->timing) FL_HIP_TRY(ctx, hipEventRecord(s->ev[4 * b + 3], stm));
        }
        FL_HIP_TRY(ctx, hipGetLastError());
        FL_HIP_TRY(ctx, hipMemcpyAsync(s->st_host, s->st, sizeof(PcgState),
                                       hipMemcpyDeviceToHost, stm));
        FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
        if (s->timing) {
            // Only iterations that did work count (later launches exit early).
            const long long active = std::min<long long>(batch, s->st_host->iter - done_iters);
            for (long long b = 0; b < active; ++b) {
                float a = 0, u = 0, dd = 0;
                FL_HIP_TRY(ctx, hipEventElapsedTime(&a, s->ev[4 * b], s->ev[4 * b + 1]));
                FL_HIP_TRY(ctx, hipEventElapsedTime(&u, s->ev[4 * b + 1], s->ev[4 * b + 2]));
                FL_HIP_TRY(ctx, hipEventElapsedTime(&dd, s->ev[4 * b + 2], s->ev[4 * b + 3]));
                s->t_apply += a;
                s->t_update += u;
                s->t_dir += dd;
                s->t_iters += 1;
            }
        }
        done_iters = s->st_host->iter;
        if (s->st_host->done) break;
        batch = std::min(batch * 2, kMaxBatch);
    }
    const PcgState &h = *s->st_host;
    if (rep) {
        rep->iterations = h.iter;
        rep->refinements = 0;
        rep->residual_abs_inf = h.rmax;
        rep->residual_l2 = std::sqrt(h.rr);
        rep->initial_l2 = std::sqrt(h.rr0);
        rep->converged = h.converged;
        rep->status = h.status == ST_OK ? FL_OK : FL_ERR_SOLVER;
    }
    if (h.status != ST_OK) return set_error(ctx, FL_ERR_SOLVER, "%s", status_text(h.status));
    return FL_OK;
}

fl_status run_diag(fl_cavity3d *s, fl_cavity3d_step_report *rep, bool check_finite) {
    fl_ctx *ctx = s->ctx;
    hipStream_t stm = ctx->stream;
    DiagState init{};
    for (int f = 0; f < 4; ++f) init.first_bad[f] = ~0ull;
    *s->dg_host = init;
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->dg, s->dg_host, sizeof(DiagState), hipMemcpyHostToDevice, stm));
    k_diag<<<s->ew_blocks, kBlock, 0, stm>>>(s->P, s->u, s->v, s->w, s->partials, s->dg);
    if (check_finite)
        k_finite<<<s->ew_blocks, kBlock, 0, stm>>>(s->p, (uint32_t)s->n, s->u, (uint32_t)s->nu,
                                                   s->v, (uint32_t)s->nv, s->w, (uint32_t)s->nw,
                                                   s->dg);
    FL_HIP_TRY(ctx, hipGetLastError());
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->dg_host, s->dg, sizeof(DiagState), hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
    const DiagState &h = *s->dg_host;
    if (rep) {
        rep->max_divergence = h.max_div;
        rep->divergence_l2 = std::sqrt(h.div_l2sq);
        rep->max_speed = h.max_speed;
    }
    if (check_finite) {
        const char *names[4] = {"pressure", "u", "v", "w"};
        double *arrs[4] = {s->p, s->u, s->v, s->w};
        for (int f = 0; f < 4; ++f)
            if (h.first_bad[f] != ~0ull) {
                double val = 0.0;
                FL_HIP_TRY(ctx, hipMemcpy(&val, arrs[f] + h.first_bad[f], sizeof(double),
                                          hipMemcpyDeviceToHost));
                return set_error(ctx, FL_ERR_SOLVER, "Non-finite 3D %s value at flat index %llu: %g",
                                 names[f], h.first_bad[f], val);
            }
    }
    return FL_OK;
}

fl_status do_step(fl_cavity3d *s, fl_cavity3d_step_report *rep) {
    fl_ctx *ctx = s->ctx;
    hipStream_t stm = ctx->stream;
    k_predict<<<s->ew_blocks, kBlock, 0, stm>>>(s->P, s->u, s->v, s->w, s->us, s->vs, s->ws);
    k_rhs<<<s->ew_blocks, kBlock, 0, stm>>>(s->P, s->us, s->vs, s->ws, s->rhs);
    FL_HIP_TRY(ctx, hipGetLastError());
    fl_solve_report sr{};
    fl_status st = solve_pressure(s, &sr);
    if (rep) {
        rep->pressure_iterations = sr.iterations;
        rep->pressure_refinements = sr.refinements;
        rep->pressure_residual = sr.residual_abs_inf;
        rep->pressure_residual_l2 = sr.residual_l2;
        rep->pressure_converged = sr.converged;
    }
    if (st != FL_OK) return st;
    k_correct<<<s->ew_blocks, kBlock, 0, stm>>>(s->P, s->p, s->us, s->vs, s->ws, s->u, s->v, s->w);
    FL_HIP_TRY(ctx, hipGetLastError());
    s->step += 1;
    st = run_diag(s, rep, true);
    if (rep) {
        rep->step = s->step;
        rep->time = (double)s->step * s->cfg.dt;
    }
    return st;
}

}  // namespace

extern "C" {

fl_status fl_cavity3d_create(fl_ctx *ctx, const fl_cavity3d_config *cfg, fl_cavity3d **out) {
    if (!ctx || !cfg || !out) return set_error(ctx, FL_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    fl_status st = validate(ctx, cfg);
    if (st != FL_OK) return st;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    fl_cavity3d *s = new (std::nothrow) fl_cavity3d();
    if (!s) return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "host allocation failed");
    s->ctx = ctx;
    s->cfg = *cfg;
    const uint32_t nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    const double dx = cfg->length / nx, dy = cfg->height / ny, dz = cfg->depth / nz;
    s->g = Grid{nx, ny, nz, nx * ny, nx * ny * nz, 1.0 / (dx * dx), 1.0 / (dy * dy),
                1.0 / (dz * dz)};
    s->P = StepParams{nx, ny, nz, dx, dy, dz, cfg->dt,
                      cfg->lid_velocity * cfg->length / cfg->reynolds,  // viscosity :86-88
                      cfg->lid_velocity, cfg->density};
    s->n = (size_t)nx * ny * nz;
    s->nu = (size_t)(nx + 1) * ny * nz;
    s->nv = (size_t)nx * (ny + 1) * nz;
    s->nw = (size_t)nx * ny * (nz + 1);
    const uint64_t ntiles = (s->n + kBlock - 1) / kBlock;
    s->blocks = grid_blocks(ctx, ntiles);
    s->ew_blocks = grid_blocks(ctx, (s->nu + s->nv + s->nw + kBlock - 1) / kBlock);
    double **fields[] = {&s->u, &s->v, &s->w, &s->us, &s->vs, &s->ws};
    size_t lens[] = {s->nu, s->nv, s->nw, s->nu, s->nv, s->nw};
    for (int a = 0; a < 6; ++a)
        if ((st = alloc(s, fields[a], lens[a])) != FL_OK) goto fail;
    {
        double **cells[] = {&s->p, &s->rhs, &s->r, &s->d, &s->q};
        for (double **c : cells)
            if ((st = alloc(s, c, s->n)) != FL_OK) goto fail;
    }
    if ((st = alloc(s, &s->partials, (size_t)std::max(s->blocks, s->ew_blocks) * 4)) != FL_OK)
        goto fail;
    if (hipMalloc(&s->st, sizeof(PcgState)) != hipSuccess ||
        hipMalloc(&s->dg, sizeof(DiagState)) != hipSuccess ||
        hipHostMalloc(&s->st_host, sizeof(PcgState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&s->dg_host, sizeof(DiagState), hipHostMallocDefault) != hipSuccess) {
        st = set_error(ctx, FL_ERR_OUT_OF_MEMORY, "device allocation failed");
        goto fail;
    }
    if (hipMemsetAsync(s->st, 0, sizeof(PcgState), ctx->stream) != hipSuccess ||
        hipMemsetAsync(s->dg, 0, sizeof(DiagState), ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        st = set_error(ctx, FL_ERR_HIP, "device initialisation failed");
        goto fail;
    }
    *out = s;
    return FL_OK;
fail:
    free_all(s);
    delete s;
    return st;
}

void fl_cavity3d_destroy(fl_cavity3d *s) {
    if (!s) return;
    (void)hipStreamSynchronize(s->ctx->stream);
    free_all(s);
    delete s;
}

fl_status fl_cavity3d_step(fl_cavity3d *s, fl_cavity3d_step_report *rep) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    FL_HIP_TRY(s->ctx, hipSetDevice(s->ctx->device));
    return do_step(s, rep);
}

fl_status fl_cavity3d_run(fl_cavity3d *s, int64_t steps, fl_cavity3d_step_report *rep) {
    if (!s || steps < 0) return FL_ERR_INVALID_ARGUMENT;
    FL_HIP_TRY(s->ctx, hipSetDevice(s->ctx->device));
    for (int64_t k = 0; k < steps; ++k) {
        fl_status st = do_step(s, rep);
        if (st != FL_OK) return st;
    }
    return FL_OK;
}

fl_status fl_cavity3d_pressure_solve(fl_cavity3d *s, const double *rhs, double *p_inout,
                                     fl_solve_report *rep) {
    if (!s || !rhs || !p_inout) return set_error(s ? s->ctx : nullptr, FL_ERR_INVALID_ARGUMENT, "null argument");
    fl_ctx *ctx = s->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    for (size_t c = 1; c < s->n; ++c)
        if (!std::isfinite(rhs[c]))
            return set_error(ctx, FL_ERR_NON_FINITE, "non-finite right-hand side value at index %zu", c);
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->rhs, rhs, s->n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->p, p_inout, s->n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    fl_status st = solve_pressure(s, rep);
    FL_HIP_TRY(ctx, hipMemcpyAsync(p_inout, s->p, s->n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return st;
}

fl_status fl_cavity3d_get_fields(fl_cavity3d *s, double *u, double *v, double *w, double *p) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = s->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t stm = ctx->stream;
    if (u) FL_HIP_TRY(ctx, hipMemcpyAsync(u, s->u, s->nu * 8, hipMemcpyDeviceToHost, stm));
    if (v) FL_HIP_TRY(ctx, hipMemcpyAsync(v, s->v, s->nv * 8, hipMemcpyDeviceToHost, stm));
    if (w) FL_HIP_TRY(ctx, hipMemcpyAsync(w, s->w, s->nw * 8, hipMemcpyDeviceToHost, stm));
    if (p) FL_HIP_TRY(ctx, hipMemcpyAsync(p, s->p, s->n * 8, hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
    return FL_OK;
}

fl_status fl_cavity3d_set_fields(fl_cavity3d *s, const double *u, const double *v, const double *w,
                                 const double *p) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = s->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t stm = ctx->stream;
    if (u) FL_HIP_TRY(ctx, hipMemcpyAsync(s->u, u, s->nu * 8, hipMemcpyHostToDevice, stm));
    if (v) FL_HIP_TRY(ctx, hipMemcpyAsync(s->v, v, s->nv * 8, hipMemcpyHostToDevice, stm));
    if (w) FL_HIP_TRY(ctx, hipMemcpyAsync(s->w, w, s->nw * 8, hipMemcpyHostToDevice, stm));
    if (p) FL_HIP_TRY(ctx, hipMemcpyAsync(s->p, p, s->n * 8, hipMemcpyHostToDevice, stm));
    FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
    return FL_OK;
}

fl_status fl_cavity3d_get_intermediate(fl_cavity3d *s, double *u_star, double *v_star,
                                       double *w_star, double *rhs) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = s->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t stm = ctx->stream;
    if (u_star) FL_HIP_TRY(ctx, hipMemcpyAsync(u_star, s->us, s->nu * 8, hipMemcpyDeviceToHost, stm));
    if (v_star) FL_HIP_TRY(ctx, hipMemcpyAsync(v_star, s->vs, s->nv * 8, hipMemcpyDeviceToHost, stm));
    if (w_star) FL_HIP_TRY(ctx, hipMemcpyAsync(w_star, s->ws, s->nw * 8, hipMemcpyDeviceToHost, stm));
    if (rhs) FL_HIP_TRY(ctx, hipMemcpyAsync(rhs, s->rhs, s->n * 8, hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
    return FL_OK;
}

fl_status fl_cavity3d_set_timing(fl_cavity3d *s, int32_t enable) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = s->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (enable && s->ev.empty()) {
        s->ev.resize(4 * kMaxBatch);
        for (auto &e : s->ev) FL_HIP_TRY(ctx, hipEventCreate(&e));
    }
    s->timing = enable != 0;
    s->t_apply = s->t_update = s->t_dir = 0.0;
    s->t_iters = 0;
    return FL_OK;
}

fl_status fl_cavity3d_kernel_times(fl_cavity3d *s, double *ms_per_iteration, double *ms_apply,
                                   double *ms_update, double *ms_direction,
                                   int64_t *timed_iterations) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    const double n = s->t_iters > 0 ? (double)s->t_iters : 1.0;
    if (ms_apply) *ms_apply = s->t_apply / n;
    if (ms_update) *ms_update = s->t_update / n;
    if (ms_direction) *ms_direction = s->t_dir / n;
    if (ms_per_iteration) *ms_per_iteration = (s->t_apply + s->t_update + s->t_dir) / n;
    if (timed_iterations) *timed_iterations = s->t_iters;
    return FL_OK;
}

}  // extern "C"
