// 2D structured projection step on gfx950, batched over independent grids.
//
// Replaces IncompressibleSolver with projection coupling
// (src/solver/incompressible.rs):
//   predict_momentum :748-922 (+ MomentumStencil :2089-2212) -> k_predict
//   apply_velocity_boundaries :1981-2044                      -> k_boundaries
//   build_pressure_rhs :924-970                               -> k_rhs
//   solve_pressure_pcg :1048-1214 (+ operator :1216-1255,
//     Jacobi :1257-1294, project_zero_mean :2338-2377)        -> k_pcg (persistent)
//   correct_velocity :1322-1379                               -> k_correct
//   compute_cell_fields :1381-1439, max_divergence :1569-1582,
//     ensure_finite :1707-1742, stability numbers :2294-2306  -> k_cells + k_fold_diag
//   advance_one_step :622-746, adaptive_time_step :2308-2336  -> host driver
//
// A single solver is a batch of one.  Cells of every grid are packed into
// 256-cell tiles (one grid per tile, tail padded); faces are packed per grid.
// Per-cell flag bytes carry the mask: which of the four neighbours is open,
// and whether the row is an identity row (solid or outlet cell).
//
// The pressure solve is one cooperative persistent kernel for all grids.
// Per iteration it runs two phases separated by grid barriers:
//   A: fold the update partials -> stop test, beta, the zero-mean shift of
//      z; d' = P(D^-1 r) + beta d (recomputed at the stencil points),
//      q = A d', partial d'.q
//   B: fold d'.q -> alpha; x += alpha d', r -= alpha q; partials r.z, sum z,
//      sum r, r.r, max|r|
// The reference's zero-mean projections (on x, b, r, z, d every iteration)
// are applied where they change anything: b once, z through its mean (folded
// into the d' formula and into r.z = sum r z - sum r sum z / count), x once
// at the end.  A d' and the CG recurrences keep r and d mean-free up to
// round-off because the Neumann operator's columns sum to zero.

#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "common.hpp"

namespace cgrp = cooperative_groups;

namespace fl {
namespace g2 {

constexpr int kT = 256;  // cells per tile (= threads per block)

enum : uint8_t { F_E = 1, F_W = 2, F_N = 4, F_S = 8, F_ID = 16, F_SOLID = 32, F_OUTLET = 64 };
enum { KV = FL_BC_VELOCITY, KS = FL_BC_SYMMETRY, KO = FL_BC_OUTLET };
enum { SL = FL_SIDE_LEFT, SR = FL_SIDE_RIGHT, SB = FL_SIDE_BOTTOM, ST = FL_SIDE_TOP };

struct Prob {
    uint32_t nx, ny, n;
    uint32_t tile0, ntiles;
    uint32_t uoff, voff;
    double dx, dy, cx, cy, nu, rho;
    int32_t scheme;
    int32_t kind[4];
    double bu[4], bv[4];
    double pout;
    int32_t zm;  // zero-mean gauge (no outlet cells)
    double inv_count;
    double tol;
    long long max_iter;
    double eta;  // mixed path: relative tolerance of each inner CG cycle
    int32_t stop_rule, guess;
    FastDiv fx;
    int32_t inlet;       // FL_INLET_* of a left Velocity side
    double step_h, height;  // inlet profile: step height and domain height
    double omega;        // SOR relaxation factor
    uint32_t ref;        // SOR gauge: the first fluid cell (find_reference_cell)
    int32_t simple;      // SIMPLE-style coupling (else projection)
    double au, ap;       // velocity relaxation (1 for projection), pressure relaxation
};

struct PcgA {  // a grid's PCG state after phase A
    double rz, beta, mz, rr0, rr, rmax;
    double rr0o, mb;  // r.r of the first true residual; mean of b
    int converged, status;
};
struct PcgB {  // a grid's PCG state after phase B (set up by the host)
    double rz_prev, alpha;
    long long iters;
    int first, status;
    int outer, pad;  // mixed path: the last pass was a true residual
};
enum { ST_OK = 0, ST_SINGULAR = 1, ST_ALPHA = 2, ST_RESIDUAL = 3, ST_BARRIER = 4 };

struct Diag {
    double max_div, max_speed, max_u, max_v, du, dv;
    unsigned long long bad[4];  // first non-finite index of p, u, v, speed
};

__device__ inline bool cell_of(const Prob &P, uint32_t t, uint32_t &c, uint32_t &i, uint32_t &j) {
    c = (t - P.tile0) * kT + threadIdx.x;
    if (c >= P.n) return false;
    j = P.fx.div(c);
    i = c - j * P.nx;
    return true;
}

// u_face_open / v_face_open (incompressible.rs:2046-2064) from the flags.
__device__ inline bool uopen(const Prob &P, const uint8_t *F, int i, int j) {
    if (i == 0) return !(F[j * P.nx] & F_SOLID);
    if (i == (int)P.nx) return !(F[P.nx - 1 + j * P.nx] & F_SOLID);
    const uint8_t f = F[i + j * P.nx];
    return !(f & F_SOLID) && (f & F_W);
}
__device__ inline bool vopen(const Prob &P, const uint8_t *F, int i, int j) {
    if (j == 0) return !(F[i] & F_SOLID);
    if (j == (int)P.ny) return !(F[i + (P.ny - 1) * P.nx] & F_SOLID);
    const uint8_t f = F[i + j * P.nx];
    return !(f & F_SOLID) && (f & F_S);
}

// MomentumStencil::u_neighbor / v_neighbor with the tangential wall ghosts.
__device__ inline double u_nb(const Prob &P, const uint8_t *F, const double *u, int i, int j,
                              int di, int dj, double c) {
    const int a = i + di, b = j + dj;
    if (a >= 0 && a <= (int)P.nx && b >= 0 && b < (int)P.ny)
        return uopen(P, F, a, b) ? u[a + (P.nx + 1) * b] : -c;
    const int side = b < 0 ? SB : (b >= (int)P.ny ? ST : -1);
    if (side < 0) return c;
    return P.kind[side] == KV ? 2.0 * P.bu[side] - c : c;
}
__device__ inline double v_nb(const Prob &P, const uint8_t *F, const double *v, int i, int j,
                              int di, int dj, double c) {
    const int a = i + di, b = j + dj;
    if (a >= 0 && a < (int)P.nx && b >= 0 && b <= (int)P.ny)
        return vopen(P, F, a, b) ? v[a + P.nx * b] : -c;
    const int side = a < 0 ? SL : (a >= (int)P.nx ? SR : -1);
    if (side < 0) return c;
    return P.kind[side] == KV ? 2.0 * P.bv[side] - c : c;
}
__device__ inline double transported(int scheme, double fv, double l, double r) {
    return scheme == FL_UPWIND ? (fv >= 0.0 ? l : r) : 0.5 * (l + r);
}

// ------------------------------------------------------------ SOR
// solve_pressure_sor (incompressible.rs:979-1046) in red-black order.  One
// sweep = the two colours (i + j) & 1, the gauge shift (the first fluid
// cell's value subtracted from every non-solid cell when no outlet pins the
// level) and the max residual of pressure_residual (:1532-1567); the verdict
// is kept per grid on the device.
struct SorState {
    unsigned long long res_bits;  // max |Lap p - rhs| of the sweep (non-negative double bits)
    double res, refval;
    long long iters;
    int done, converged, bad, pad;
};

__global__ void __launch_bounds__(kT) k_sor_sweep(const Prob *probs, const uint32_t *tile_prob,
                                                  uint32_t ntiles, const int *active, const uint8_t *flags,
                                                  const double *__restrict__ rhs, double *__restrict__ x,
                                                  uint32_t color, const SorState *st) {
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t pi = tile_prob[t];
        if (!active[pi] || st[pi].done) continue;
        const Prob &P = probs[pi];
        uint32_t c, i, j;
        if (!cell_of(P, t, c, i, j) || ((i + j) & 1) != color) continue;
        const size_t g = (size_t)P.tile0 * kT + c;
        const uint8_t f = flags[g];
        if (f & F_SOLID) {
            x[g] = 0.0;
            continue;
        }
        if (f & F_OUTLET) {
            x[g] = P.pout;
            continue;
        }
        double ap = 0.0, sum = 0.0;
        if (f & F_E) { ap += P.cx; sum += P.cx * x[g + 1]; }
        if (f & F_W) { ap += P.cx; sum += P.cx * x[g - 1]; }
        if (f & F_N) { ap += P.cy; sum += P.cy * x[g + P.nx]; }
        if (f & F_S) { ap += P.cy; sum += P.cy * x[g - P.nx]; }
        if (ap > 0.0) {
            const double candidate = (sum - rhs[g]) / ap;
            const double old = x[g];
            x[g] = old + P.omega * (candidate - old);
        }
    }
}
// The gauge value of each grid (one thread per grid).
__global__ void k_sor_ref(const Prob *probs, const int *active, int count, const double *__restrict__ x,
                          SorState *st) {
    const int pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= count || !active[pi] || st[pi].done) return;
    const Prob &P = probs[pi];
    st[pi].refval = P.zm ? x[(size_t)P.tile0 * kT + P.ref] : 0.0;
}
__global__ void __launch_bounds__(kT) k_sor_shift(const Prob *probs, const uint32_t *tile_prob,
                                                  uint32_t ntiles, const int *active, const uint8_t *flags,
                                                  double *__restrict__ x, const SorState *st) {
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t pi = tile_prob[t];
        if (!active[pi] || st[pi].done || !probs[pi].zm) continue;
        const Prob &P = probs[pi];
        uint32_t c, i, j;
        if (!cell_of(P, t, c, i, j)) continue;
        const size_t g = (size_t)P.tile0 * kT + c;
        if (!(flags[g] & F_SOLID)) x[g] -= st[pi].refval;
    }
}
// max |Lap p - rhs| over fluid, non-outlet cells (a closed neighbour reads the
// centre); one atomic max per tile on the non-negative double's bits.
__global__ void __launch_bounds__(kT) k_sor_residual(const Prob *probs, const uint32_t *tile_prob,
                                                     uint32_t ntiles, const int *active, const uint8_t *flags,
                                                     const double *__restrict__ rhs,
                                                     const double *__restrict__ x, SorState *st) {
    __shared__ double red[kT / 64];
    __shared__ int bad_any[1];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t pi = tile_prob[t];
        if (!active[pi] || st[pi].done) continue;  // uniform over the tile
        const Prob &P = probs[pi];
        if (threadIdx.x == 0) bad_any[0] = 0;
        __syncthreads();
        double r = 0.0;
        uint32_t c, i, j;
        if (cell_of(P, t, c, i, j)) {
            const size_t g = (size_t)P.tile0 * kT + c;
            const uint8_t f = flags[g];
            if (!(f & F_ID)) {
                const double p0 = x[g];
                const double pe = (f & F_E) ? x[g + 1] : p0, pw = (f & F_W) ? x[g - 1] : p0;
                const double pn = (f & F_N) ? x[g + P.nx] : p0, ps = (f & F_S) ? x[g - P.nx] : p0;
                const double lap = (pe - 2.0 * p0 + pw) * P.cx + (pn - 2.0 * p0 + ps) * P.cy;
                r = fabs(lap - rhs[g]);
                if (!isfinite(r)) {
                    bad_any[0] = 1;
                    r = 0.0;
                }
            }
        }
        for (int o = 32; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = red[0];
            for (int w = 1; w < kT / 64; ++w) m = fmax(m, red[w]);
            atomicMax(&st[pi].res_bits, (unsigned long long)__double_as_longlong(m));
            if (bad_any[0]) atomicOr(&st[pi].bad, 1);
        }
        __syncthreads();
    }
}
// The sweep's verdict per grid (one thread per grid).
__global__ void k_sor_check(const Prob *probs, const int *active, int count, SorState *st) {
    const int pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= count || !active[pi] || st[pi].done) return;
    SorState &S = st[pi];
    const double res = __longlong_as_double((long long)S.res_bits);
    S.res_bits = 0ull;
    S.iters += 1;
    S.res = S.bad ? (double)NAN : res;
    if (S.bad) S.done = 1;
    else if (res < probs[pi].tol) S.converged = S.done = 1;
    else if (S.iters >= probs[pi].max_iter) S.done = 1;
}

// ------------------------------------------------------------ step kernels
__global__ void __launch_bounds__(kT) k_predict(const Prob *probs, const uint32_t *tile_prob,
                                                uint32_t ntiles, const int *active,
                                                const double *dts, const uint8_t *flags,
                                                const double *__restrict__ Pp,
                                                const double *__restrict__ U,
                                                const double *__restrict__ V,
                                                double *__restrict__ US, double *__restrict__ VS) {
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t pi = tile_prob[t];
        if (!active[pi]) continue;
        const Prob &P = probs[pi];
        uint32_t c, ui, uj;
        if (!cell_of(P, t, c, ui, uj)) continue;
        const int i = ui, j = uj, nx = P.nx, ny = P.ny;
        const uint8_t *F = flags + (size_t)P.tile0 * kT;
        const double *u = U + P.uoff, *v = V + P.voff;
        double *us = US + P.uoff, *vs = VS + P.voff;
        const double *p = Pp + (size_t)P.tile0 * kT;
        const double dt = dts[pi], dx = P.dx, dy = P.dy;
        if (i >= 1) {
            double val = 0.0;
            if (uopen(P, F, i, j)) {
                const double up = u[i + (nx + 1) * j];
                const double ue_nb = u_nb(P, F, u, i, j, 1, 0, up), uw_nb = u_nb(P, F, u, i, j, -1, 0, up);
                const double un_nb = u_nb(P, F, u, i, j, 0, 1, up), us_nb = u_nb(P, F, u, i, j, 0, -1, up);
                const double ue = 0.5 * (up + ue_nb), uw = 0.5 * (uw_nb + up);
                const double vn = 0.5 * (v[i - 1 + nx * (j + 1)] + v[i + nx * (j + 1)]);
                const double vsv = 0.5 * (v[i - 1 + nx * j] + v[i + nx * j]);
                const double pe = transported(P.scheme, ue, up, ue_nb), pw = transported(P.scheme, uw, uw_nb, up);
                const double pn = transported(P.scheme, vn, up, un_nb), ps = transported(P.scheme, vsv, us_nb, up);
                const double conv = (ue * pe - uw * pw) / dx + (vn * pn - vsv * ps) / dy;
                const double diff = (ue_nb - 2.0 * up + uw_nb) / (dx * dx) + (un_nb - 2.0 * up + us_nb) / (dy * dy);
                // SIMPLE: the relaxed step with the current pressure's gradient
                val = P.simple ? up + P.au * dt * (-conv + P.nu * diff - (p[c] - p[c - 1]) / dx / P.rho)
                               : up + dt * (-conv + P.nu * diff);
            }
            us[i + (nx + 1) * j] = val;
        } else {
            us[(nx + 1) * j] = 0.0;
        }
        if (i == nx - 1) us[nx + (nx + 1) * j] = 0.0;
        if (j >= 1) {
            double val = 0.0;
            if (vopen(P, F, i, j)) {
                const double vp = v[i + nx * j];
                const double ve_nb = v_nb(P, F, v, i, j, 1, 0, vp), vw_nb = v_nb(P, F, v, i, j, -1, 0, vp);
                const double vn_nb = v_nb(P, F, v, i, j, 0, 1, vp), vs_nb = v_nb(P, F, v, i, j, 0, -1, vp);
                const double ue = 0.5 * (u[i + 1 + (nx + 1) * (j - 1)] + u[i + 1 + (nx + 1) * j]);
                const double uw = 0.5 * (u[i + (nx + 1) * (j - 1)] + u[i + (nx + 1) * j]);
                const double vn = 0.5 * (vp + vn_nb), vsv = 0.5 * (vs_nb + vp);
                const double pe = transported(P.scheme, ue, vp, ve_nb), pw = transported(P.scheme, uw, vw_nb, vp);
                const double pn = transported(P.scheme, vn, vp, vn_nb), ps = transported(P.scheme, vsv, vs_nb, vp);
                const double conv = (ue * pe - uw * pw) / dx + (vn * pn - vsv * ps) / dy;
                const double diff = (ve_nb - 2.0 * vp + vw_nb) / (dx * dx) + (vn_nb - 2.0 * vp + vs_nb) / (dy * dy);
                val = P.simple ? vp + P.au * dt * (-conv + P.nu * diff - (p[c] - p[c - nx]) / dy / P.rho)
                               : vp + dt * (-conv + P.nu * diff);
            }
            vs[i + nx * j] = val;
        } else {
            vs[i] = 0.0;
        }
        if (j == ny - 1) vs[i + nx * ny] = 0.0;
    }
}

// Boundary-normal faces (apply_velocity_boundaries :1981-2022); one block
// per grid.  Interior blocked faces are already zero.
__global__ void __launch_bounds__(kT) k_boundaries(const Prob *probs, const int *active,
                                                   double *__restrict__ U, double *__restrict__ V) {
    const uint32_t pi = blockIdx.x;
    if (!active[pi]) return;
    const Prob &P = probs[pi];
    double *u = U + P.uoff, *v = V + P.voff;
    const uint32_t nx = P.nx, ny = P.ny;
    for (uint32_t j = threadIdx.x; j < ny; j += blockDim.x) {
        double &l = u[(nx + 1) * j], &r = u[nx + (nx + 1) * j];
        double ul = P.bu[SL];
        if (P.inlet == FL_INLET_PARABOLIC) {
            // Case::boundary_velocity's Channel / BackwardFacingStep inlet at the
            // face centre y = (j + 1/2) dy (src/cases/mod.rs:141-170)
            const double y = ((double)j + 0.5) * P.dy;
            const double eta = fmin(fmax((y - P.step_h) / (P.height - P.step_h), 0.0), 1.0);
            ul = y <= P.step_h ? 0.0 : 6.0 * P.bu[SL] * eta * (1.0 - eta);
        }
        l = P.kind[SL] == KV ? ul : (P.kind[SL] == KS ? 0.0 : u[1 + (nx + 1) * j]);
        r = P.kind[SR] == KV ? P.bu[SR] : (P.kind[SR] == KS ? 0.0 : u[nx - 1 + (nx + 1) * j]);
    }
    for (uint32_t i = threadIdx.x; i < nx; i += blockDim.x) {
        double &b = v[i], &t = v[i + nx * ny];
        b = P.kind[SB] == KV ? P.bv[SB] : (P.kind[SB] == KS ? 0.0 : v[i + nx]);
        t = P.kind[ST] == KV ? P.bv[ST] : (P.kind[ST] == KS ? 0.0 : v[i + nx * (ny - 1)]);
    }
}

// build_pressure_rhs (:924-970): zero on solid and outlet cells; SIMPLE
// scales by the velocity relaxation (:936-941).
__global__ void __launch_bounds__(kT) k_rhs(const Prob *probs, const uint32_t *tile_prob,
                                            uint32_t ntiles, const int *active, const double *dts,
                                            const uint8_t *flags, const double *__restrict__ US,
                                            const double *__restrict__ VS, double *__restrict__ RHS) {
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t pi = tile_prob[t];
        if (!active[pi]) continue;
        const Prob &P = probs[pi];
        uint32_t c, i, j;
        if (!cell_of(P, t, c, i, j)) continue;
        const size_t base = (size_t)P.tile0 * kT;
        if (flags[base + c] & F_ID) {
            RHS[base + c] = 0.0;
            continue;
        }
        const double *us = US + P.uoff, *vs = VS + P.voff;
        const uint32_t nx = P.nx;
        const double div = (us[i + 1 + (nx + 1) * j] - us[i + (nx + 1) * j]) / P.dx +
                           (vs[i + nx * (j + 1)] - vs[i + nx * j]) / P.dy;
        RHS[base + c] = P.rho / (P.au * dts[pi]) * div;
    }
}

// correct_velocity (:1322-1379): u = u* - a dt/rho grad p on open interior
// faces (a = the velocity relaxation under SIMPLE, else 1), 0 on blocked
// ones, boundary faces copied from u*.
__global__ void __launch_bounds__(kT) k_correct(const Prob *probs, const uint32_t *tile_prob,
                                                uint32_t ntiles, const int *active,
                                                const double *dts, const uint8_t *flags,
                                                const double *__restrict__ Pp,
                                                const double *__restrict__ US,
                                                const double *__restrict__ VS,
                                                double *__restrict__ U, double *__restrict__ V) {
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t pi = tile_prob[t];
        if (!active[pi]) continue;
        const Prob &P = probs[pi];
        uint32_t c, ui, uj;
        if (!cell_of(P, t, c, ui, uj)) continue;
        const int i = ui, j = uj, nx = P.nx, ny = P.ny;
        const uint8_t *F = flags + (size_t)P.tile0 * kT;
        const double *p = Pp + (size_t)P.tile0 * kT;
        const double *us = US + P.uoff, *vs = VS + P.voff;
        double *u = U + P.uoff, *v = V + P.voff;
        const double dor = P.au * dts[pi] / P.rho;
        if (i >= 1)
            u[i + (nx + 1) * j] = uopen(P, F, i, j)
                                      ? us[i + (nx + 1) * j] - dor * ((p[c] - p[c - 1]) / P.dx)
                                      : 0.0;
        else
            u[(nx + 1) * j] = us[(nx + 1) * j];
        if (i == nx - 1) u[nx + (nx + 1) * j] = us[nx + (nx + 1) * j];
        if (j >= 1)
            v[i + nx * j] = vopen(P, F, i, j) ? vs[i + nx * j] - dor * ((p[c] - p[c - nx]) / P.dy)
                                              : 0.0;
        else
            v[i] = vs[i];
        if (j == ny - 1) v[i + nx * ny] = vs[i + nx * ny];
    }
}

// SIMPLE's pressure split (incompressible.rs:669-673, 683-694): before the
// solve the pressure moves to `pold` and the solve starts from p' = 0; after
// the correction p = p_old + alpha_p p'.  Projection grids are untouched.
__global__ void __launch_bounds__(kT) k_simple_save(const Prob *probs, const uint32_t *tile_prob,
                                                    uint32_t ntiles, const int *active,
                                                    double *__restrict__ X, double *__restrict__ PO) {
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t pi = tile_prob[t];
        if (!active[pi] || !probs[pi].simple) continue;
        const size_t k = (size_t)t * kT + threadIdx.x;
        PO[k] = X[k];
        X[k] = 0.0;
    }
}
__global__ void __launch_bounds__(kT) k_simple_combine(const Prob *probs, const uint32_t *tile_prob,
                                                       uint32_t ntiles, const int *active,
                                                       double *__restrict__ X,
                                                       const double *__restrict__ PO) {
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t pi = tile_prob[t];
        if (!active[pi] || !probs[pi].simple) continue;
        const size_t k = (size_t)t * kT + threadIdx.x;
        X[k] = PO[k] + probs[pi].ap * X[k];
    }
}

// Cell-centred fields (:1381-1439) and the per-tile diagnostic maxima: max
// divergence (:1569-1582), max speed, max|u|, max|v|, velocity change, and
// the first non-finite index of p, u, v, speed (:1707-1742).
__device__ inline double cell_u(const Prob &P, const double *u, int i, int j) {
    return 0.5 * (u[i + (P.nx + 1) * j] + u[i + 1 + (P.nx + 1) * j]);
}
__device__ inline double cell_v(const Prob &P, const double *v, int i, int j) {
    return 0.5 * (v[i + P.nx * j] + v[i + P.nx * (j + 1)]);
}
__global__ void __launch_bounds__(kT) k_cells(const Prob *probs, const uint32_t *tile_prob,
                                              uint32_t ntiles, const int *active,
                                              const uint8_t *flags, const double *__restrict__ Pp,
                                              const double *__restrict__ U,
                                              const double *__restrict__ V,
                                              const double *__restrict__ UO,
                                              const double *__restrict__ VO, double *UC,
                                              double *VC, double *SP, double *VORT,
                                              double *partials, Diag *diag) {
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const uint32_t pi = tile_prob[t];
        const Prob &P = probs[pi];
        double s[1] = {0.0}, m[6] = {0, 0, 0, 0, 0, 0};
        uint32_t c = 0, ui = 0, uj = 0;
        const bool live = active[pi] && cell_of(P, t, c, ui, uj);
        if (live) {
            const int i = ui, j = uj, nx = P.nx, ny = P.ny;
            const size_t base = (size_t)P.tile0 * kT;
            const uint8_t *F = flags + base;
            const double *u = U + P.uoff, *v = V + P.voff, *uo = UO + P.uoff, *vo = VO + P.voff;
            const bool solid = F[c] & F_SOLID;
            double uc = 0.0, vc = 0.0, sp = 0.0, vort = 0.0;
            if (!solid) {
                uc = cell_u(P, u, i, j);
                vc = cell_v(P, v, i, j);
                sp = sqrt(uc * uc + vc * vc);
                // derivative_x_slice (v_cell) - derivative_y_slice (u_cell) (:2228-2285)
                const bool eo = i + 1 < nx && !(F[c + 1] & F_SOLID), wo = i > 0 && !(F[c - 1] & F_SOLID);
                const bool no = j + 1 < ny && !(F[c + nx] & F_SOLID), so = j > 0 && !(F[c - nx] & F_SOLID);
                const double ve = eo ? cell_v(P, v, i + 1, j) : vc, vw = wo ? cell_v(P, v, i - 1, j) : vc;
                const double un = no ? cell_u(P, u, i, j + 1) : uc, us_ = so ? cell_u(P, u, i, j - 1) : uc;
                const double dvdx = (eo && wo) ? (ve - vw) / (2.0 * P.dx)
                                    : eo       ? (ve - vc) / P.dx
                                    : wo       ? (vc - vw) / P.dx
                                               : 0.0;
                const double dudy = (no && so) ? (un - us_) / (2.0 * P.dy)
                                    : no       ? (un - uc) / P.dy
                                    : so       ? (uc - us_) / P.dy
                                               : 0.0;
                vort = dvdx - dudy;
                const double div = (u[i + 1 + (nx + 1) * j] - u[i + (nx + 1) * j]) / P.dx +
                                   (v[i + nx * (j + 1)] - v[i + nx * j]) / P.dy;
                m[0] = fmax(m[0], fabs(div));
            }
            UC[base + c] = uc;
            VC[base + c] = vc;
            SP[base + c] = sp;
            VORT[base + c] = vort;
            m[1] = fmax(m[1], sp);
            // faces owned by this cell: u(i,j) [+ u(nx,j)], v(i,j) [+ v(i,ny)]
            const uint32_t fu = i + (nx + 1) * j, fv = i + nx * j;
            m[2] = fmax(m[2], fabs(u[fu]));
            m[3] = fmax(m[3], fabs(v[fv]));
            m[4] = fmax(m[4], fabs(u[fu] - uo[fu]));
            m[5] = fmax(m[5], fabs(v[fv] - vo[fv]));
            if (!isfinite(u[fu])) atomicMin(&diag[pi].bad[1], (unsigned long long)fu);
            if (!isfinite(v[fv])) atomicMin(&diag[pi].bad[2], (unsigned long long)fv);
            if (i == nx - 1) {
                const uint32_t f2 = nx + (nx + 1) * j;
                m[2] = fmax(m[2], fabs(u[f2]));
                m[4] = fmax(m[4], fabs(u[f2] - uo[f2]));
                if (!isfinite(u[f2])) atomicMin(&diag[pi].bad[1], (unsigned long long)f2);
            }
            if (j == ny - 1) {
                const uint32_t f2 = i + nx * ny;
                m[3] = fmax(m[3], fabs(v[f2]));
                m[5] = fmax(m[5], fabs(v[f2] - vo[f2]));
                if (!isfinite(v[f2])) atomicMin(&diag[pi].bad[2], (unsigned long long)f2);
            }
            if (!isfinite(Pp[base + c])) atomicMin(&diag[pi].bad[0], (unsigned long long)c);
            if (!isfinite(sp)) atomicMin(&diag[pi].bad[3], (unsigned long long)c);
        }
        block_reduce<1, 6>(s, m);
        if (threadIdx.x == 0)
            for (int a = 0; a < 6; ++a) partials[(size_t)t * 6 + a] = m[a];
        __syncthreads();
    }
}

// One block per grid folds its tiles' diagnostic maxima in tile order.
__global__ void __launch_bounds__(kT) k_fold_diag(const Prob *probs, const int *active,
                                                  const double *partials, Diag *diag) {
    const uint32_t pi = blockIdx.x;
    if (!active[pi]) return;
    const Prob &P = probs[pi];
    double s[1] = {0.0}, m[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t t = threadIdx.x; t < P.ntiles; t += blockDim.x)
        for (int a = 0; a < 6; ++a) m[a] = fmax(m[a], partials[(size_t)(P.tile0 + t) * 6 + a]);
    block_reduce<1, 6>(s, m);
    if (threadIdx.x == 0) {
        Diag &d = diag[pi];
        d.max_div = m[0];
        d.max_speed = m[1];
        d.max_u = m[2];
        d.max_v = m[3];
        d.du = m[4];
        d.dv = m[5];
    }
}

// ------------------------------------------------------ persistent PCG
// Work layout.  At the start of each iteration every block lists the
// 256-cell tiles of the grids still iterating, in grid order, and takes an
// equal contiguous share of that list, so the work stays balanced over all
// blocks as grids finish.  A wave takes whole tiles of its block's share
// (tile i goes to wave i mod NT/64); each lane holds four cells of the tile
// (lane, lane + 64, ...), so every lane has four cells' loads in flight,
// and the wave writes one partial row per tile: the lane sums its cells in
// order and the wave combines its lanes by a fixed butterfly.  Every block
// then folds the rows of the grids its share touches, with groups of 256
// threads: thread u sums rows u, u + 256, ... in order, each wave combines
// its lanes by the butterfly and the group adds its four wave totals in
// order.  So every block gets the same scalars, and results depend neither
// on the number of blocks, nor on their size, nor on the share a block
// gets.  The block whose share holds a grid's first tile writes the grid's
// state out (for the host and for the next iteration's blocks).
constexpr int kCells = kT / 64;   // cells per lane in a tile
constexpr int kMaxGrids = 256;    // grids per batch
constexpr uint32_t kFoldT = 256;  // threads folding one grid's rows

// T is the storage type of the CG vectors: double (the reference's FP64
// PCG) or float (FP32 CG cycles on the correction e inside FP64 iterative
// refinement: x and every true residual stay FP64).
template <typename T>
struct PcgArgs {
    const Prob *probs;
    const uint32_t *tile_prob;
    uint32_t ntiles, nprob;
    long long kmax;  // hard bound on the loop
    const uint8_t *flags;
    const double *rhs;
    double *x;
    T *r, *z, *d0, *d1, *q, *e;  // z = D^-1 r (identity rows: r); FP64: e is x
    double *pa, *pb, *p0;  // per-tile partial rows: 1, 5 and 2 columns
    PcgA *sa;
    PcgB *sb;
    const int *active;  // grids taking part in this solve
    int *fin;  // fin[parity * nprob + grid]: grid finished in that loop (the
               // host sets fin[1][g] = !active[g], the "loop -1" verdicts)
    unsigned *bar;  // grid barrier: arrival count [0], generation [32], abort [64]
};

// Grid-wide barrier for a co-resident (cooperative) grid.  The last block to
// arrive resets the arrival count and then advances the generation the
// others wait on, so the count is zero again after every barrier and between
// launches.  Thread 0's device-scope fences publish the block's writes
// before it arrives and make the other blocks' writes visible after.  The
// wait is bounded: a block that waits about four seconds (far beyond any
// phase of the solve) raises the abort word, every barrier after that
// returns false at once, and the kernel ends with the abort word set for
// the host to report.
constexpr unsigned kBarrierSpins = 1u << 22;
__device__ inline bool grid_barrier(unsigned *bar) {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned *count = bar, *gen = bar + 32, *abort = bar + 64;
        int ok = __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
        if (ok) {
            const unsigned g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __threadfence();
            if (atomicAdd(count, 1u) == gridDim.x - 1) {
                __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __threadfence();
                __hip_atomic_store(gen, g + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                unsigned spins = 0;
                while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                    if (++spins > kBarrierSpins ||
                        __hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                        __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            __threadfence();
        }
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// Jacobi inverse diagonal from the open-neighbour bits.
__device__ inline double inv_diag(double cx, double cy, uint8_t f) {
    const double dg = cx * (double)(((f & F_E) != 0) + ((f & F_W) != 0)) +
                      cy * (double)(((f & F_N) != 0) + ((f & F_S) != 0));
    return dg > 0.0 ? 1.0 / dg : 1.0;
}

// A grid as every PCG block sees it: geometry, its place in the iteration's
// sub-tile list, the scalars of the current phase and the solver state.
struct GridL {
    uint32_t tile0, ntiles, n, nx;
    uint32_t fxm, fxs;  // x = c mod nx via FastDiv's multiplier and shift
    uint32_t off;       // first list index (tiles of earlier iterating grids)
    int32_t act;        // iterating in this loop
    double cx, cy;
    double a, b;  // phase A: beta, mean of z; phase B: alpha; setup: mean of b / x
    double rz, rz_prev, rr0;
    double rr0o, mb;  // r.r of the first true residual; mean of b
    double tol, eta, inv_count;  // stop test (copied from Prob)
    long long iters, max_iter;
    int32_t stop_rule, zm;
    int32_t done, first;
    int32_t outer;    // mixed: the partials come from a true residual
    int32_t skip_a;   // no CG work in phase A
    int32_t apply;    // mixed: x += e, e = 0 in phase A
    int32_t outer_b;  // mixed: phase B is a true-residual pass
};

// Adds one cell's residual to the stop test's partial sums: r.z (r.r on
// identity rows), z, r (both on the gauge cells only) and r.r.  Returns the
// cell's Jacobi-preconditioned residual z = D^-1 r (identity rows: r).
__device__ inline double residual_partials(uint8_t f, double r, double inv_d, double (&v)[4]) {
    double z = r;
    if (!(f & F_ID)) {
        z = r * inv_d;
        v[0] += r * z;
        v[1] += z;
        v[2] += r;
    } else {
        v[0] += r * r;
    }
    v[3] += r * r;
    return z;
}

// Wave-uniform copies (scalar registers) of values read from LDS.
__device__ inline uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline double uni(double v) {
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// A grid's fields the cell loops use, in scalar registers.
struct GridU {
    uint32_t tile0, n, nx, fxm, fxs;
    double cx, cy, a, b;
    bool first, done, skip_a, apply;
};
__device__ inline GridU grid_uni(const GridL &G) {
    return {uni(G.tile0), uni(G.n), uni(G.nx), uni(G.fxm), uni(G.fxs), uni(G.cx), uni(G.cy),
            uni(G.a), uni(G.b), uni((uint32_t)G.first) != 0, uni((uint32_t)G.done) != 0,
            uni((uint32_t)G.skip_a) != 0, uni((uint32_t)G.apply) != 0};
}
__device__ inline uint32_t grid_col(const GridU &G, uint32_t c) {
    const uint32_t t = __umulhi(c, G.fxm);
    return c - ((t + ((c - t) >> 1)) >> G.fxs) * G.nx;
}

// Element i of a grid's array at a 32-bit byte offset from a scalar base
// (grids hold fewer than 2^28 cells).
template <typename T>
__device__ inline T ld(const T *base, uint32_t i) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ inline void st(T *base, uint32_t i, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + i * (uint32_t)sizeof(T)) = v;
}

// The iteration's tile list: offsets of the iterating grids (one wave
// scans gl[].act in grid order), this block's share [lo, hi) and the grids
// [g0, g1] that share touches.  Ends with a block barrier.
struct Share {
    uint32_t lo, hi, g0, g1, ng;
};
__device__ inline Share make_share(GridL *gl, uint32_t nprob) {
    __shared__ uint32_t s_total, s_g0, s_g1;
    if (threadIdx.x < 64) {
        uint32_t carry = 0;
        for (uint32_t c0 = 0; c0 < nprob; c0 += 64) {
            const uint32_t g = c0 + threadIdx.x;
            const uint32_t v = g < nprob && gl[g].act ? gl[g].ntiles : 0u;
            uint32_t incl = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t t = __shfl_up(incl, o, 64);
                if ((int)threadIdx.x >= o) incl += t;
            }
            if (g < nprob) gl[g].off = carry + incl - v;
            carry += __shfl(incl, 63, 64);
        }
        if (threadIdx.x == 0) {
            s_total = carry;
            s_g0 = 1u;
            s_g1 = 0u;
        }
    }
    __syncthreads();
    const uint64_t S = s_total;
    Share sh;
    sh.lo = (uint32_t)(S * blockIdx.x / gridDim.x);
    sh.hi = (uint32_t)(S * (blockIdx.x + 1) / gridDim.x);
    for (uint32_t g = threadIdx.x; g < nprob; g += blockDim.x) {
        const GridL &G = gl[g];
        if (!G.act || sh.lo >= sh.hi) continue;
        if (G.off <= sh.lo && sh.lo < G.off + G.ntiles) s_g0 = g;
        if (G.off < sh.hi && sh.hi <= G.off + G.ntiles) s_g1 = g;
    }
    __syncthreads();
    sh.g0 = s_g0;
    sh.g1 = s_g1;
    sh.ng = nprob;
    __syncthreads();
    return sh;
}

// Calls f(t, g, c0) for each of this wave's tiles in the block's share: t
// is the tile (its partial row), g its grid, c0 this lane's first cell (the
// others follow at +64, +128, +192).
template <uint32_t kW, typename F>
__device__ inline void for_share(const GridL *gl, const Share &sh, F &&f) {
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    uint32_t g = sh.g0;
    for (uint32_t i = sh.lo + wave; i < sh.hi; i += kW) {
        while (g + 1 < uni(sh.ng) && (!uni((uint32_t)gl[g].act) || i >= uni(gl[g].off) + uni(gl[g].ntiles))) ++g;
        const uint32_t lt = i - uni(gl[g].off);
        f(uni(gl[g].tile0) + lt, g, lt * kT + lane);
    }
}

// Folds the partial rows (NS sums, then NM maxima per row) of every grid g
// in [g0, g1] with want(g), then calls use(g, sums, maxima) in one thread.
// Ends with a block barrier, so what use() wrote to LDS is visible.
template <int NS, int NM, int NT, typename W, typename U>
__device__ inline void fold_grids(const double *__restrict__ part, const GridL *gl, uint32_t g0,
                                  uint32_t g1, double (*red)[8], W &&want, U &&use) {
    static_assert(NS + NM <= 8, "at most 8 columns");
    constexpr int K = NS + NM;
    constexpr uint32_t kG = NT / kFoldT;
    const uint32_t q = threadIdx.x / kFoldT, u = threadIdx.x % kFoldT;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (uint32_t j0 = g0; j0 <= g1; j0 += kG) {
        const uint32_t g = j0 + q;
        const bool on = g <= g1 && want(g);
        double v[K];
#pragma unroll
        for (int a = 0; a < K; ++a) v[a] = 0.0;
        if (on) {
            const uint32_t row0 = gl[g].tile0, rows = gl[g].ntiles;
#pragma unroll 2
            for (uint32_t i = u; i < rows; i += kFoldT) {
                const double *p = part + (size_t)(row0 + i) * K;
#pragma unroll
                for (int a = 0; a < NS; ++a) v[a] += p[a];
#pragma unroll
                for (int a = NS; a < K; ++a) v[a] = fmax(v[a], p[a]);
            }
        }
#pragma unroll
        for (int a = 0; a < NS; ++a) v[a] = wave_sum(v[a]);
#pragma unroll
        for (int a = NS; a < K; ++a) v[a] = wave_max(v[a]);
        if (lane == 0)
#pragma unroll
            for (int a = 0; a < K; ++a) red[wave][a] = v[a];
        __syncthreads();
        if (on && u == 0) {
            double sm[NS > 0 ? NS : 1], mx[NM > 0 ? NM : 1];
            const uint32_t w0 = q * (kFoldT / 64);
#pragma unroll
            for (int a = 0; a < NS; ++a) {
                double t = red[w0][a];
                for (uint32_t w = 1; w < kFoldT / 64; ++w) t += red[w0 + w][a];
                sm[a] = t;
            }
#pragma unroll
            for (int a = 0; a < NM; ++a) {
                double t = red[w0][NS + a];
                for (uint32_t w = 1; w < kFoldT / 64; ++w) t = fmax(t, red[w0 + w][NS + a]);
                mx[a] = t;
            }
            use(g, sm, mx);
        }
        __syncthreads();
    }
}

// Phase A over one tile (this lane's cells c0, c0 + 64, ...): the search
// direction d' = P(z) + beta d and q = A d'; returns the lane's sum of d'.q.
// kFirst: d' = P(z) (no previous direction is read).
template <bool kFirst, int kBatchA, typename T>
__device__ inline double search_dir(const PcgArgs<T> &A, const GridU &G, const T *dold, T *dnew,
                                    uint32_t c0) {
    const size_t base = (size_t)G.tile0 * kT;
    const uint8_t *F = A.flags + base;
    const T *z = A.z + base, *dv = dold + base;
    double v = 0.0;
#pragma unroll
    for (int u0 = 0; u0 < kCells; u0 += kBatchA) {
        // loads of kBatchA cells first; padding lanes read the grid's
        // last cell, and neighbours come from the geometry (a closed
        // side still reads a cell of the grid), so no load waits for
        // the flags
        uint8_t f[kBatchA];
        T zv[kBatchA][5], dd[kBatchA][kFirst ? 1 : 5];
#pragma unroll
        for (int u = 0; u < kBatchA; ++u) {
            const uint32_t c = min(c0 + 64 * (u0 + u), G.n - 1);
            const uint32_t i = grid_col(G, c);
            const uint32_t ce = i + 1 < G.nx ? c + 1 : c, cw = i > 0 ? c - 1 : c;
            const uint32_t cn = c + G.nx < G.n ? c + G.nx : c, cs = c >= G.nx ? c - G.nx : c;
            const uint32_t k5[5] = {c, ce, cw, cn, cs};
            f[u] = F[c];
#pragma unroll
            for (int a = 0; a < 5; ++a) {
                zv[u][a] = ld(z, k5[a]);
                if (!kFirst) dd[u][a] = ld(dv, k5[a]);
            }
        }
#pragma unroll
        for (int u = 0; u < kBatchA; ++u) {
            const uint32_t c = c0 + 64 * (u0 + u);
            if (c >= G.n) continue;
            // the projection P(z) = z - mean(z) on the gauge cells is
            // applied to the stored d' only: A annihilates the
            // constant on the gauge cells (their open neighbours are
            // gauge cells, identity rows only exist without a gauge)
            double dn[5];
#pragma unroll
            for (int a = 0; a < 5; ++a)
                dn[a] = kFirst ? (double)zv[u][a] : (double)zv[u][a] + G.a * (double)dd[u][kFirst ? 0 : a];
            double qc = dn[0], dc = dn[0];
            if (!(f[u] & F_ID)) {
                double dg = 0.0, sum = 0.0;
                if (f[u] & F_E) { dg += G.cx; sum += G.cx * dn[1]; }
                if (f[u] & F_W) { dg += G.cx; sum += G.cx * dn[2]; }
                if (f[u] & F_N) { dg += G.cy; sum += G.cy * dn[3]; }
                if (f[u] & F_S) { dg += G.cy; sum += G.cy * dn[4]; }
                qc = dg * dn[0] - sum;
                const double zs = (double)zv[u][0] - G.b;
                dc = kFirst ? zs : zs + G.a * (double)dd[u][0];
            }
            // d'.q from the stored (for the mixed path rounded) values
            const T dt = (T)dc, qt = (T)qc;
            st(dnew + base, c, dt);
            st(A.q + base, c, qt);
            v += (double)dt * (double)qt;
        }
    }
    return v;
}

template <int NT, typename T>
__global__ void __launch_bounds__(NT) k_pcg(PcgArgs<T> A) {
    constexpr bool kMixed = !std::is_same<T, double>::value;
    // cells per lane whose phase-A loads are in flight together
    constexpr int kBatchA = kMixed ? 4 : 2;
    constexpr uint32_t kW = NT / 64;
    __shared__ GridL gl[kMaxGrids];
    __shared__ double red[kW][8];
    const uint32_t lane = threadIdx.x & 63, nprob = A.nprob;
    for (uint32_t g = threadIdx.x; g < nprob; g += NT) {
        const Prob &P = A.probs[g];
        const PcgB sb = A.sb[g];
        GridL &G = gl[g];
        G.tile0 = P.tile0;
        G.ntiles = P.ntiles;
        G.n = P.n;
        G.nx = P.nx;
        G.fxm = P.fx.m;
        G.fxs = P.fx.s;
        G.act = A.active[g] != 0;
        G.cx = P.cx;
        G.cy = P.cy;
        G.tol = P.tol;
        G.eta = P.eta;
        G.inv_count = P.inv_count;
        G.max_iter = P.max_iter;
        G.stop_rule = P.stop_rule;
        G.zm = P.zm;
        G.rz_prev = sb.rz_prev;
        G.iters = sb.iters;
        G.first = sb.first;
        G.outer = kMixed ? 1 : 0;  // the set-up residual is a true residual
        G.rr0 = 0.0;
        G.rr0o = -1.0;  // not yet known
        G.rz = 0.0;
        G.done = G.skip_a = G.apply = G.outer_b = 0;
    }
    __syncthreads();
    // the set-up and final passes share out every active grid
    const Share all = make_share(gl, nprob);
    // ---- init 0: sums of x0 and b = -rhs over the gauge cells
    for_share<kW>(gl, all, [=](uint32_t t, uint32_t g, uint32_t c0) {
        const GridL &G = gl[g];
        const size_t base = (size_t)G.tile0 * kT;
        const bool zero = A.probs[g].guess == FL_GUESS_ZERO;
        double v0 = 0.0, v1 = 0.0;
#pragma unroll
        for (int u = 0; u < kCells; ++u) {
            const uint32_t c = c0 + 64 * u;
            if (c < G.n && !(A.flags[base + c] & F_ID)) {
                v0 += zero ? 0.0 : A.x[base + c];
                v1 += -A.rhs[base + c];
            }
        }
        v0 = wave_sum(v0);
        v1 = wave_sum(v1);
        if (lane == 0) {
            A.p0[(size_t)t * 2] = v0;
            A.p0[(size_t)t * 2 + 1] = v1;
        }
    });
    if (!grid_barrier(A.bar)) return;
    // ---- init 1: r = P(b) - A x0; identity rows x = value, r = 0 (the
    // mixed path stores r in T and clears the correction e)
    fold_grids<2, 0, NT>(A.p0, gl, all.g0, all.g1, red, [=](uint32_t g) { return gl[g].act != 0; },
                         [=](uint32_t g, const double *sm, const double *) {
                             const Prob &P = A.probs[g];
                             gl[g].mb = P.zm ? sm[1] * P.inv_count : 0.0;
                             if (gl[g].off >= all.lo && gl[g].off < all.hi) A.sa[g].mb = gl[g].mb;
                         });
    for_share<kW>(gl, all, [=](uint32_t t, uint32_t g, uint32_t c0) {
        const GridL &G = gl[g];
        const Prob &P = A.probs[g];
        const size_t base = (size_t)G.tile0 * kT;
        const uint8_t *F = A.flags + base;
        double *x = A.x + base;
        const bool zero = P.guess == FL_GUESS_ZERO;
        double v[4] = {0.0, 0.0, 0.0, 0.0}, mx = 0.0;
#pragma unroll
        for (int u = 0; u < kCells; ++u) {
            const uint32_t c = c0 + 64 * u;
            if (c >= G.n) continue;
            const uint8_t f = F[c];
            double rc = 0.0;
            if (f & F_ID) {
                x[c] = (f & F_OUTLET) ? P.pout : 0.0;
            } else {
                auto xv = [&](uint32_t k) -> double {
                    const uint8_t h = F[k];
                    if (h & F_ID) return (h & F_OUTLET) ? P.pout : 0.0;
                    return zero ? 0.0 : x[k];
                };
                const double xc = xv(c);
                double dg = 0.0, sum = 0.0;
                if (f & F_E) { dg += G.cx; sum += G.cx * xv(c + 1); }
                if (f & F_W) { dg += G.cx; sum += G.cx * xv(c - 1); }
                if (f & F_N) { dg += G.cy; sum += G.cy * xv(c + G.nx); }
                if (f & F_S) { dg += G.cy; sum += G.cy * xv(c - G.nx); }
                rc = (-A.rhs[base + c] - G.mb) - (dg * xc - sum);
            }
            const T rt = (T)rc;
            A.r[base + c] = rt;
            if (kMixed) A.e[base + c] = (T)0;
            A.z[base + c] = (T)residual_partials(f, (double)rt, inv_diag(G.cx, G.cy, f), v);
            mx = fmax(mx, fabs(rc));
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) v[a] = wave_sum(v[a]);
        mx = wave_max(mx);
        if (lane == 0) {
            double *pb = A.pb + (size_t)t * 5;
            pb[0] = v[0]; pb[1] = v[1]; pb[2] = v[2]; pb[3] = v[3]; pb[4] = mx;
        }
    });
    if (!grid_barrier(A.bar)) return;
    // zero-guess grids start from x = 0 on the fluid cells (read next in
    // phase B, after a barrier)
    for_share<kW>(gl, all, [=](uint32_t, uint32_t g, uint32_t c0) {
        const GridL &G = gl[g];
        const size_t base = (size_t)G.tile0 * kT;
        if (A.probs[g].guess != FL_GUESS_ZERO) return;
#pragma unroll
        for (int u = 0; u < kCells; ++u) {
            const uint32_t c = c0 + 64 * u;
            if (c < G.n && !(A.flags[base + c] & F_ID)) A.x[base + c] = 0.0;
        }
    });
    // ---- CG loop; the host's fin[1] verdicts say which grids iterate
    for (uint32_t g = threadIdx.x; g < nprob; g += NT) {
        gl[g].act = A.fin[A.nprob + g] == 0;
        gl[g].mb = A.sa[g].mb;
    }
    __syncthreads();
    Share sh = make_share(gl, nprob);
    for (long long k = 0; k < A.kmax; ++k) {
        const int par = (int)(k & 1);
        const T *dold = (k & 1) ? A.d1 : A.d0;
        T *dnew = (k & 1) ? A.d0 : A.d1;
        // phase A scalars: stop test, beta, mean of z; the mixed path also
        // decides between a CG step and a true-residual pass
        fold_grids<4, 1, NT>(A.pb, gl, sh.g0, sh.g1, red, [=](uint32_t g) { return gl[g].act != 0; },
                             [=](uint32_t g, const double *sm, const double *mx) {
            GridL &G = gl[g];
            const double rr = sm[3], rmax = mx[0];
            const double rz = G.zm ? sm[0] - sm[2] * sm[1] * G.inv_count : sm[0];
            const double mz = G.zm ? sm[1] * G.inv_count : 0.0;
            int status = ST_OK, conv = 0;
            bool done = false, skip_a = false, outer_b = false, first = G.first != 0;
            double beta = 0.0, rr0 = G.rr0, rr0o = G.rr0o;
            const bool owner = G.off >= sh.lo && G.off < sh.hi;
            if (!kMixed) {
                rr0 = first ? rr : G.rr0;
                if (!first && !isfinite(rmax)) {
                    status = ST_RESIDUAL;
                    done = true;
                } else if (G.stop_rule == FL_STOP_REL_L2 ? sqrt(rr) <= G.tol * sqrt(rr0)
                                                        : rmax < G.tol) {
                    conv = 1;
                    done = true;
                } else if (!first) {
                    if (fabs(G.rz_prev) < 1.0e-30) done = true;
                    else if (G.iters >= G.max_iter) done = true;
                    else beta = rz / G.rz_prev;
                }
            } else if (G.outer) {
                // a true residual: the reference's stop test, else a new
                // inner CG from it
                if (rr0o < 0.0) rr0o = rr;
                if (!isfinite(rmax)) {
                    status = ST_RESIDUAL;
                    done = true;
                } else if (G.stop_rule == FL_STOP_REL_L2 ? sqrt(rr) <= G.tol * sqrt(rr0o)
                                                        : rmax < G.tol) {
                    conv = 1;
                    done = true;
                } else if (G.iters >= G.max_iter) {
                    done = true;
                } else {
                    first = true;
                    rr0 = rr;
                }
            } else {
                // inner CG: end the cycle (true-residual pass next) when the
                // recurrence residual meets half the outer tolerance, has
                // dropped by the inner tolerance, or the budget is spent
                const bool met = G.stop_rule == FL_STOP_REL_L2 ? sqrt(rr) <= 0.5 * G.tol * sqrt(rr0o)
                                                               : rmax <= 0.5 * G.tol;
                if (!isfinite(rmax) || met || sqrt(rr) <= G.eta * sqrt(rr0) ||
                    G.iters >= G.max_iter || fabs(G.rz_prev) < 1.0e-30) {
                    skip_a = true;
                    outer_b = true;
                } else {
                    beta = rz / G.rz_prev;
                }
            }
            G.apply = kMixed && G.outer ? 1 : 0;  // x += e after a true-residual pass
            G.a = beta;
            G.b = mz;
            G.rz = rz;
            G.rr0 = rr0;
            G.rr0o = rr0o;
            G.first = first ? 1 : 0;
            G.done = done ? 1 : 0;
            G.skip_a = done || skip_a ? 1 : 0;
            G.outer_b = outer_b ? 1 : 0;
            if (owner) {
                PcgA &sa = A.sa[g];
                sa.rz = rz;
                sa.beta = beta;
                sa.mz = mz;
                sa.rr0 = rr0;
                sa.rr0o = rr0o;
                if (!kMixed || G.outer) {
                    sa.rr = rr;
                    sa.rmax = rmax;
                    sa.converged = conv;
                    sa.status = status;
                }
                A.fin[par * A.nprob + g] = done ? 1 : 0;
            }
        });
        // phase A: (mixed, after a true-residual pass) x += e, e = 0; then
        // d' = P(D^-1 r) + beta d at the stencil points, q = A d'
        for_share<kW>(gl, sh, [=](uint32_t t, uint32_t g, uint32_t c0) {
            const GridU G = grid_uni(gl[g]);
            const size_t base = (size_t)G.tile0 * kT;
            if (kMixed && G.apply) {
#pragma unroll
                for (int u = 0; u < kCells; ++u) {
                    const uint32_t c = c0 + 64 * u;
                    if (c >= G.n) continue;
                    const double ev = (double)ld(A.e + base, c);
                    st(A.x + base, c, ld(A.x + base, c) + ev);
                    st(A.e + base, c, (T)0);
                }
            }
            if (G.skip_a) return;
            const double v = G.first ? search_dir<true, kBatchA>(A, G, dold, dnew, c0)
                                     : search_dir<false, kBatchA>(A, G, dold, dnew, c0);
            const double vs = wave_sum(v);
            if (lane == 0) A.pa[t] = vs;
        });
        if (!grid_barrier(A.bar)) return;
        // phase B scalars: alpha
        fold_grids<1, 0, NT>(A.pa, gl, sh.g0, sh.g1, red,
                             [=](uint32_t g) { return gl[g].act != 0 && gl[g].skip_a == 0; },
                             [=](uint32_t g, const double *sm, const double *) {
            GridL &G = gl[g];
            const double den = sm[0];
            int status = ST_OK;
            double alpha = 0.0;
            if (fabs(den) < 1.0e-30 || !isfinite(den)) status = ST_SINGULAR;
            else {
                alpha = G.rz / den;
                if (!isfinite(alpha)) status = ST_ALPHA;
            }
            G.a = alpha;
            G.done = status != ST_OK ? 1 : 0;
            if (G.off >= sh.lo && G.off < sh.hi) {
                PcgB &sb = A.sb[g];
                sb.rz_prev = G.rz;
                sb.alpha = alpha;
                sb.first = 0;
                sb.outer = 0;
                sb.status = status;
                sb.iters = G.iters + (status == ST_OK ? 1 : 0);
                if (status != ST_OK) A.fin[par * A.nprob + g] = 1;
            }
        });
        // phase B: x (mixed: e) += alpha d', r -= alpha q, partials of the
        // next test; the mixed path's true-residual pass r = P(b) - A (x + e)
        for_share<kW>(gl, sh, [=](uint32_t t, uint32_t g, uint32_t c0) {
            const GridU G = grid_uni(gl[g]);
            const bool outer = kMixed && uni((uint32_t)gl[g].outer_b) != 0;
            if (G.done || (G.skip_a && !outer)) return;
            const size_t base = (size_t)G.tile0 * kT;
            const uint8_t *F = A.flags + base;
            double v[4] = {0.0, 0.0, 0.0, 0.0}, mx = 0.0;
            if (outer) {
                const double mb = uni(gl[g].mb);
#pragma unroll
                for (int u = 0; u < kCells; ++u) {
                    const uint32_t c = c0 + 64 * u;
                    if (c >= G.n) continue;
                    const uint8_t f = F[c];
                    double rc = 0.0;
                    if (!(f & F_ID)) {
                        const uint32_t i = grid_col(G, c);
                        const uint32_t ce = i + 1 < G.nx ? c + 1 : c, cw = i > 0 ? c - 1 : c;
                        const uint32_t cn = c + G.nx < G.n ? c + G.nx : c, cs = c >= G.nx ? c - G.nx : c;
                        const uint32_t k5[5] = {c, ce, cw, cn, cs};
                        double xs[5];
#pragma unroll
                        for (int a = 0; a < 5; ++a)
                            xs[a] = ld(A.x + base, k5[a]) + (double)ld(A.e + base, k5[a]);
                        // identity rows hold their value in x and 0 in e
                        double dg = 0.0, sum = 0.0;
                        if (f & F_E) { dg += G.cx; sum += G.cx * xs[1]; }
                        if (f & F_W) { dg += G.cx; sum += G.cx * xs[2]; }
                        if (f & F_N) { dg += G.cy; sum += G.cy * xs[3]; }
                        if (f & F_S) { dg += G.cy; sum += G.cy * xs[4]; }
                        rc = (-A.rhs[base + c] - mb) - (dg * xs[0] - sum);
                    }
                    const T rt = (T)rc;
                    st(A.r + base, c, rt);
                    st(A.z + base, c, (T)residual_partials(f, (double)rt, inv_diag(G.cx, G.cy, f), v));
                    mx = fmax(mx, fabs(rc));
                }
            } else {
                uint8_t f[kCells];
                T ev[kCells], dv[kCells], r0[kCells], qv[kCells];
#pragma unroll
                for (int u = 0; u < kCells; ++u) {
                    const uint32_t c = min(c0 + 64 * u, G.n - 1);
                    f[u] = F[c];
                    ev[u] = ld(A.e + base, c);
                    dv[u] = ld(dnew + base, c);
                    r0[u] = ld(A.r + base, c);
                    qv[u] = ld(A.q + base, c);
                }
#pragma unroll
                for (int u = 0; u < kCells; ++u) {
                    const uint32_t c = c0 + 64 * u;
                    if (c >= G.n) continue;
                    st(A.e + base, c, (T)((double)ev[u] + G.a * (double)dv[u]));
                    const T rt = (T)((double)r0[u] - G.a * (double)qv[u]);
                    st(A.r + base, c, rt);
                    st(A.z + base, c, (T)residual_partials(f[u], (double)rt, inv_diag(G.cx, G.cy, f[u]), v));
                    mx = fmax(mx, fabs((double)rt));
                }
            }
#pragma unroll
            for (int a = 0; a < 4; ++a) v[a] = wave_sum(v[a]);
            mx = wave_max(mx);
            if (lane == 0) {
                double *pb = A.pb + (size_t)t * 5;
                pb[0] = v[0]; pb[1] = v[1]; pb[2] = v[2]; pb[3] = v[3]; pb[4] = mx;
            }
        });
        if (kMixed)  // the true-residual pass's owner marks the next test
            for (uint32_t g = sh.g0 + threadIdx.x; g <= sh.g1; g += NT)
                if (gl[g].act && gl[g].outer_b && gl[g].off >= sh.lo && gl[g].off < sh.hi) A.sb[g].outer = 1;
        if (!grid_barrier(A.bar)) return;
        // next loop: which grids still iterate, and their state as the
        // blocks that held their first tiles wrote it
        int open = 0;
        for (uint32_t g = threadIdx.x; g < nprob; g += NT) {
            GridL &G = gl[g];
            if (!G.act) continue;
            const int f = A.fin[par * A.nprob + g];
            const PcgB sb = A.sb[g];
            const PcgA &sa = A.sa[g];
            G.act = f == 0;
            G.rz_prev = sb.rz_prev;
            G.iters = sb.iters;
            G.first = sb.first;
            G.outer = sb.outer;
            G.rr0 = sa.rr0;
            G.rr0o = sa.rr0o;
            G.done = G.skip_a = G.apply = G.outer_b = 0;
            open |= G.act;
        }
        if (!__syncthreads_or(open)) break;
        sh = make_share(gl, nprob);
    }
    // ---- final: project x to the zero-mean gauge (or write identity values)
    for (uint32_t g = threadIdx.x; g < nprob; g += NT) gl[g].act = A.active[g] != 0;
    __syncthreads();
    const Share fin_sh = make_share(gl, nprob);
    for_share<kW>(gl, fin_sh, [=](uint32_t t, uint32_t g, uint32_t c0) {
        const GridL &G = gl[g];
        const size_t base = (size_t)G.tile0 * kT;
        double v = 0.0;
#pragma unroll
        for (int u = 0; u < kCells; ++u) {
            const uint32_t c = c0 + 64 * u;
            if (c < G.n && !(A.flags[base + c] & F_ID)) v += A.x[base + c];
        }
        v = wave_sum(v);
        if (lane == 0) {
            A.p0[(size_t)t * 2] = v;
            A.p0[(size_t)t * 2 + 1] = 0.0;
        }
    });
    if (!grid_barrier(A.bar)) return;
    fold_grids<2, 0, NT>(A.p0, gl, fin_sh.g0, fin_sh.g1, red, [=](uint32_t g) { return gl[g].act != 0; },
                         [=](uint32_t g, const double *sm, const double *) {
                             const Prob &P = A.probs[g];
                             gl[g].b = P.zm ? sm[0] * P.inv_count : 0.0;
                         });
    for_share<kW>(gl, fin_sh, [=](uint32_t, uint32_t g, uint32_t c0) {
        const GridL &G = gl[g];
        const size_t base = (size_t)G.tile0 * kT;
#pragma unroll
        for (int u = 0; u < kCells; ++u) {
            const uint32_t c = c0 + 64 * u;
            if (c >= G.n) continue;
            const size_t i = base + c;
            const uint8_t f = A.flags[i];
            if (f & F_ID) A.x[i] = (f & F_OUTLET) ? A.probs[g].pout : 0.0;
            else A.x[i] -= G.b;
        }
    });
}

// ------------------------------------------------ device-side step control
// Runs of many steps (fl_batch2d_run) keep every decision on the device: the
// time-step rule, t_end, the step and time counters and the first error of
// each grid, so that a step is a fixed launch sequence (capturable in a HIP
// graph) and the host synchronises once per run.
struct StepCtl {
    // TimeStepSettings / t_end (incompressible.rs:35-90)
    double dt0, t_end, min_dt, max_dt, target_cfl, target_diff, growth;
    int32_t adaptive, simple;
    // state
    double time, current_dt, max_u, max_v;
    long long step, iters;
    double last_residual, max_div, max_speed;
    // first error: RE_MIN_DT (vals = stable, cfl, diffusion dt), RE_PCG
    // (which = PCG status), RE_NONFINITE (which = field, index)
    int32_t error, which;
    unsigned long long index;
    double vals[3];
};
enum { RE_NONE = 0, RE_MIN_DT = 1, RE_PCG = 2, RE_NONFINITE = 3 };

__host__ __device__ inline bool ctl_reached_end(double time, double t_end) {
    const double tol = 2.220446049250313e-16 * fmax(fabs(t_end), 1.0) * 16.0;
    return time >= t_end - tol;
}

// Step setup, one thread per grid: dt (fixed, or adaptive_time_step
// incompressible.rs:2308-2336) clamped to t_end, the activity flag, the
// PCG's per-solve state and the diagnostics reset.
__global__ void k_run_setup(const Prob *probs, StepCtl *ctl, int count, int *active, double *dts,
                            int *fin, PcgB *sb, Diag *diag) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    StepCtl &c = ctl[p];
    const Prob &P = probs[p];
    Diag d{};
    for (int f = 0; f < 4; ++f) d.bad[f] = ~0ull;
    diag[p] = d;
    int act = 0;
    double dt = c.dt0;
    if (c.error == RE_NONE && (c.simple || !ctl_reached_end(c.time, c.t_end))) {
        bool ok = true;
        if (c.adaptive) {
            const double rate = c.max_u / P.dx + c.max_v / P.dy;
            const double cfl_dt = rate > 2.220446049250313e-16 ? c.target_cfl / rate : c.max_dt;
            const double inv = 1.0 / (P.dx * P.dx) + 1.0 / (P.dy * P.dy);
            const double diff_dt = c.target_diff / (P.nu * inv);
            const double stable = fmin(fmin(c.max_dt, cfl_dt), diff_dt);
            if (stable < c.min_dt) {
                c.error = RE_MIN_DT;
                c.vals[0] = stable;
                c.vals[1] = cfl_dt;
                c.vals[2] = diff_dt;
                ok = false;
            } else {
                const double g = fmin(stable, c.current_dt * c.growth);
                dt = fmin(fmax(g, c.min_dt), c.max_dt);
            }
        }
        if (ok) {
            if (!c.simple) dt = fmin(dt, c.t_end - c.time);
            if (dt > 0.0) {
                act = 1;
                if (c.adaptive) c.current_dt = dt;
            }
        }
    }
    active[p] = act;
    dts[p] = dt;
    fin[p] = 0;
    fin[count + p] = act ? 0 : 1;
    PcgB b{};
    b.first = 1;
    sb[p] = b;
}

// Step finish, one thread per active grid: the PCG verdict, counters, the
// ensure_finite check (incompressible.rs:1707-1742; pressure, u, v, speed)
// and the maxima the next step's dt rule reads.
__global__ void k_run_finish(StepCtl *ctl, int count, const int *active, const double *dts,
                             const Diag *diag, const PcgA *sa, const PcgB *sb,
                             const unsigned *bar) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count || !active[p]) return;
    StepCtl &c = ctl[p];
    const int st = bar[64] != 0 ? ST_BARRIER : sb[p].status != ST_OK ? sb[p].status : sa[p].status;
    if (st != ST_OK) {
        c.error = RE_PCG;
        c.which = st;
        return;
    }
    c.step += 1;
    c.time = c.simple ? (double)c.step * c.dt0 : c.time + dts[p];
    const Diag &d = diag[p];
    for (int f = 0; f < 4; ++f)
        if (d.bad[f] != ~0ull) {
            c.error = RE_NONFINITE;
            c.which = f;
            c.index = d.bad[f];
            return;
        }
    c.max_u = d.max_u;
    c.max_v = d.max_v;
    c.iters += sb[p].iters;
    c.last_residual = sa[p].rmax;
    c.max_div = d.max_div;
    c.max_speed = d.max_speed;
}

// Autotuner input: a fixed pseudo-random right-hand side (splitmix64 of the
// cell index), zero on padding.
__global__ void k_fill_rhs(double *rhs, size_t n, uint64_t seed) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i + 1);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        rhs[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
}

}  // namespace g2
}  // namespace fl

// =========================================================== host driver
using namespace fl;
using namespace fl::g2;

namespace {

struct Batch2 {
    fl_ctx *ctx = nullptr;
    int count = 0;
    std::vector<fl_grid2d_config> cfg;
    std::vector<Prob> hp;
    uint32_t ntiles = 0;
    size_t ncell = 0, nuf = 0, nvf = 0;
    Prob *probs = nullptr;
    uint32_t *tile_prob = nullptr;
    uint8_t *flags = nullptr;
    double *x = nullptr, *rhs = nullptr, *r = nullptr, *d0 = nullptr, *d1 = nullptr, *q = nullptr;
    double *zv = nullptr;  // Jacobi-preconditioned residual
    double *uc = nullptr, *vc = nullptr, *sp = nullptr, *vort = nullptr;
    double *u = nullptr, *v = nullptr, *us = nullptr, *vs = nullptr, *uo = nullptr, *vo = nullptr;
    double *pa = nullptr, *pb = nullptr, *p0 = nullptr, *dpart = nullptr;
    PcgA *sa = nullptr;
    PcgB *sb = nullptr;
    int *fin = nullptr, *active = nullptr;
    double *dts = nullptr;
    Diag *diag = nullptr;
    // pinned host mirrors
    int *h_active = nullptr, *h_fin = nullptr;
    double *h_dts = nullptr;
    PcgA *h_sa = nullptr;
    PcgB *h_sb = nullptr;
    Diag *h_diag = nullptr;
    // host-side solver state per grid
    std::vector<int64_t> step;
    std::vector<double> time, current_dt, max_u, max_v;
    uint32_t pcg_blocks = 1, pcg_threads = 1024;
    bool mixed = false;       // FP32 CG cycles inside FP64 refinement
    float *e32 = nullptr;     // mixed: the FP32 correction
    unsigned *bar = nullptr;  // the persistent PCG's grid barrier
    long long kmax = 1;
    // device-driven runs: per-grid control state, the captured step (one
    // graph holding the cooperative PCG as a kernel node, or, where the
    // runtime cannot capture that, the two halves around it), per-step events
    StepCtl *ctl = nullptr, *h_ctl = nullptr;
    long long kmax_all = 1;
    int32_t graph_mode = 0;  // FL_GRAPH_* of the captured step
    hipGraph_t gr_step = nullptr, gr_pre = nullptr, gr_post = nullptr;
    hipGraphExec_t gx_step = nullptr, gx_pre = nullptr, gx_post = nullptr;
    std::vector<hipEvent_t> ev;
    // SOR (FL_SOLVER_SOR): per-grid iteration state
    bool sor = false;
    SorState *sor_st = nullptr, *h_sor = nullptr;
    // SIMPLE-style coupling on any grid: the pressure before the p' solve
    bool simple = false;
    double *pold = nullptr;
};

template <typename T>
fl_status dalloc(Batch2 *B, T **p, size_t n) {
    FL_HIP_TRY(B->ctx, hipMalloc(p, (n > 0 ? n : 1) * sizeof(T)));
    FL_HIP_TRY(B->ctx, hipMemsetAsync(*p, 0, (n > 0 ? n : 1) * sizeof(T), B->ctx->stream));
    return FL_OK;
}
template <typename T>
fl_status halloc(Batch2 *B, T **p, size_t n) {
    if (hipHostMalloc(p, (n > 0 ? n : 1) * sizeof(T), hipHostMallocDefault) != hipSuccess)
        return set_error(B->ctx, FL_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
    std::memset(*p, 0, (n > 0 ? n : 1) * sizeof(T));
    return FL_OK;
}

// Releases the captured step graphs (a launch-shape change re-captures).
void drop_graphs(Batch2 *B) {
    hipGraphExec_t *x[] = {&B->gx_step, &B->gx_pre, &B->gx_post};
    hipGraph_t *g[] = {&B->gr_step, &B->gr_pre, &B->gr_post};
    for (hipGraphExec_t *e : x)
        if (*e) {
            (void)hipGraphExecDestroy(*e);
            *e = nullptr;
        }
    for (hipGraph_t *e : g)
        if (*e) {
            (void)hipGraphDestroy(*e);
            *e = nullptr;
        }
    B->graph_mode = FL_GRAPH_NONE;
}

void free_batch(Batch2 *B) {
    // the graph execs first (the stream has drained): none refers to freed memory
    drop_graphs(B);
    void *d[] = {B->probs, B->tile_prob, B->flags, B->x,  B->rhs, B->r,  B->d0,    B->d1,
                 B->q,     B->uc,        B->vc,    B->sp, B->vort, B->u, B->v,     B->us,
                 B->vs,    B->uo,        B->vo,    B->pa, B->pb,   B->p0, B->dpart, B->sa,
                 B->sb,    B->fin,       B->active, B->dts, B->diag, B->pold};
    for (void *p : d)
        if (p) (void)hipFree(p);
    void *h[] = {B->h_active, B->h_fin, B->h_dts, B->h_sa, B->h_sb, B->h_diag, B->h_ctl};
    for (void *p : h)
        if (p) (void)hipHostFree(p);
    if (B->ctl) (void)hipFree(B->ctl);
    if (B->bar) (void)hipFree(B->bar);
    if (B->e32) (void)hipFree(B->e32);
    if (B->zv) (void)hipFree(B->zv);
    if (B->sor_st) (void)hipFree(B->sor_st);
    if (B->h_sor) (void)hipHostFree(B->h_sor);
    for (hipEvent_t e : B->ev) (void)hipEventDestroy(e);
}

// A relaxation factor of the config: 0 selects the reference's default.
double relaxation(double v, double dflt) { return v == 0.0 ? dflt : v; }

// SimulationConfig::validate (incompressible.rs:93-191), flow-only parts.
fl_status validate2(fl_ctx *ctx, const fl_grid2d_config &c) {
    if (c.nx < 4 || c.ny < 4) return set_error(ctx, FL_ERR_CONFIG, "nx and ny must be at least 4");
    if (!(c.length > 0.0 && c.height > 0.0))
        return set_error(ctx, FL_ERR_CONFIG, "Domain dimensions must be positive");
    if (!std::isfinite(c.dt) || c.dt <= 0.0)
        return set_error(ctx, FL_ERR_CONFIG, "dt must be finite and positive");
    if (!std::isfinite(c.t_end) || c.t_end < 0.0)
        return set_error(ctx, FL_ERR_CONFIG, "t_end must be finite and non-negative");
    if (c.coupling != FL_COUPLING_PROJECTION && c.coupling != FL_COUPLING_SIMPLE)
        return set_error(ctx, FL_ERR_CONFIG, "unknown coupling %d", c.coupling);
    if (c.coupling == FL_COUPLING_PROJECTION && c.t_end == 0.0)
        return set_error(ctx, FL_ERR_CONFIG, "t_end must be positive for projection coupling");
    if (c.pressure_max_iters <= 0)
        return set_error(ctx, FL_ERR_CONFIG, "pressure_max_iters must be positive");
    const double vr = relaxation(c.velocity_relaxation, 0.7), pr = relaxation(c.pressure_relaxation, 0.3);
    if (!std::isfinite(vr) || vr <= 0.0 || vr > 1.0)
        return set_error(ctx, FL_ERR_CONFIG, "velocity_relaxation must lie between 0 (exclusive) and 1");
    if (!std::isfinite(pr) || pr <= 0.0 || pr > 1.0)
        return set_error(ctx, FL_ERR_CONFIG, "pressure_relaxation must lie between 0 (exclusive) and 1");
    if (!(std::isfinite(c.pressure_tolerance) && c.pressure_tolerance > 0.0))
        return set_error(ctx, FL_ERR_CONFIG, "pressure_tolerance must be finite and positive");
    if (!(std::isfinite(c.density) && c.density > 0.0 && std::isfinite(c.kinematic_viscosity) &&
          c.kinematic_viscosity > 0.0))
        return set_error(ctx, FL_ERR_CONFIG,
                         "fluid density and kinematic viscosity must be finite and positive");
    if (c.convection != FL_UPWIND && c.convection != FL_CENTRAL)
        return set_error(ctx, FL_ERR_CONFIG, "unknown convection scheme %d", c.convection);
    if (c.precision != FL_PRECISION_FP64 && c.precision != FL_PRECISION_MIXED)
        return set_error(ctx, FL_ERR_CONFIG, "unknown precision %d", c.precision);
    if (!(std::isfinite(c.inner_tolerance) && c.inner_tolerance >= 0.0 && c.inner_tolerance < 1.0))
        return set_error(ctx, FL_ERR_CONFIG, "inner_tolerance must lie in [0, 1)");
    if (c.stop_rule != FL_STOP_ABS_INF && c.stop_rule != FL_STOP_REL_L2)
        return set_error(ctx, FL_ERR_CONFIG, "unknown stop rule %d", c.stop_rule);
    if (c.initial_guess != FL_GUESS_PREVIOUS && c.initial_guess != FL_GUESS_ZERO)
        return set_error(ctx, FL_ERR_CONFIG, "unknown initial guess %d", c.initial_guess);
    for (int s = 0; s < 4; ++s)
        if (c.bc_kind[s] < FL_BC_VELOCITY || c.bc_kind[s] > FL_BC_OUTLET)
            return set_error(ctx, FL_ERR_CONFIG, "unknown boundary kind %d", c.bc_kind[s]);
    if (c.pressure_solver != FL_SOLVER_PCG && c.pressure_solver != FL_SOLVER_SOR)
        return set_error(ctx, FL_ERR_CONFIG, "unknown pressure solver %d", c.pressure_solver);
    if (c.pressure_solver == FL_SOLVER_SOR) {
        if (!std::isfinite(c.pressure_omega) || c.pressure_omega <= 0.0 || c.pressure_omega >= 2.0)
            return set_error(ctx, FL_ERR_CONFIG, "pressure_omega must lie between 0 and 2");
        if (c.precision != FL_PRECISION_FP64 || c.stop_rule != FL_STOP_ABS_INF)
            return set_error(ctx, FL_ERR_CONFIG,
                             "the SOR pressure solver runs in FP64 with the reference's max-residual stop rule");
    }
    if (c.inlet_profile != FL_INLET_UNIFORM && c.inlet_profile != FL_INLET_PARABOLIC)
        return set_error(ctx, FL_ERR_CONFIG, "unknown inlet profile %d", c.inlet_profile);
    if (c.inlet_profile == FL_INLET_PARABOLIC &&
        !(std::isfinite(c.inlet_step_height) && c.inlet_step_height >= 0.0 && c.inlet_step_height < c.height))
        return set_error(ctx, FL_ERR_CONFIG, "inlet_step_height must lie in [0, height)");
    if (c.adaptive) {
        if (c.coupling != FL_COUPLING_PROJECTION)
            return set_error(ctx, FL_ERR_CONFIG, "adaptive time stepping requires transient projection coupling");
        if (!std::isfinite(c.min_dt) || c.min_dt <= 0.0)
            return set_error(ctx, FL_ERR_CONFIG, "adaptive min_dt must be finite and positive");
        if (!std::isfinite(c.max_dt) || c.max_dt <= 0.0)
            return set_error(ctx, FL_ERR_CONFIG, "adaptive max_dt must be finite and positive");
        if (c.min_dt > c.max_dt)
            return set_error(ctx, FL_ERR_CONFIG, "adaptive min_dt must not exceed max_dt");
        if (c.dt < c.min_dt || c.dt > c.max_dt)
            return set_error(ctx, FL_ERR_CONFIG, "adaptive initial dt must lie between min_dt and max_dt");
        if (!std::isfinite(c.target_cfl) || c.target_cfl <= 0.0 || c.target_cfl > 1.0)
            return set_error(ctx, FL_ERR_CONFIG, "adaptive target_cfl must lie between 0 (exclusive) and 1");
        if (!std::isfinite(c.target_diffusion_number) || c.target_diffusion_number <= 0.0 ||
            c.target_diffusion_number > 0.5)
            return set_error(ctx, FL_ERR_CONFIG,
                             "adaptive target_diffusion_number must lie between 0 (exclusive) and 0.5");
        if (!std::isfinite(c.max_growth_factor) || c.max_growth_factor < 1.0)
            return set_error(ctx, FL_ERR_CONFIG, "adaptive max_growth_factor must be finite and at least 1");
    }
    if ((double)c.nx * (double)c.ny > 268435456.0)
        return set_error(ctx, FL_ERR_CONFIG, "grid too large (at most 2^28 cells)");
    return FL_OK;
}

// adaptive_time_step (incompressible.rs:2308-2336), flow only.
fl_status adaptive_dt(fl_ctx *ctx, const fl_grid2d_config &c, double current, double umax,
                      double vmax, double dx, double dy, double *out) {
    const double rate = umax / dx + vmax / dy;
    const double cfl_dt = rate > 2.220446049250313e-16 ? c.target_cfl / rate : c.max_dt;
    const double inv = 1.0 / (dx * dx) + 1.0 / (dy * dy);
    const double diff_dt = c.target_diffusion_number / (c.kinematic_viscosity * inv);
    const double stable = std::min(std::min(c.max_dt, cfl_dt), diff_dt);
    if (stable < c.min_dt)
        return set_error(ctx, FL_ERR_SOLVER,
                         "required stable time step %.3e is below configured min_dt %.3e; CFL "
                         "constraint=%.3e, diffusion constraint=%.3e",
                         stable, c.min_dt, cfl_dt, diff_dt);
    const double g = std::min(stable, current * c.max_growth_factor);
    *out = std::min(std::max(g, c.min_dt), c.max_dt);
    return FL_OK;
}

// reached_end_time (incompressible.rs:557-561): SIMPLE never ends on time.
bool reached_end(const Batch2 *B, int p) {
    if (B->hp[p].simple) return false;
    const double t_end = B->cfg[p].t_end;
    const double tol = 2.220446049250313e-16 * std::max(std::fabs(t_end), 1.0) * 16.0;
    return B->time[p] >= t_end - tol;
}

// The persistent PCG's kernel for a block size.
const void *pcg_kernel(uint32_t threads, bool mixed) {
    switch (threads) {
        case 256: return mixed ? (const void *)k_pcg<256, float> : (const void *)k_pcg<256, double>;
        case 512: return mixed ? (const void *)k_pcg<512, float> : (const void *)k_pcg<512, double>;
        default: return mixed ? (const void *)k_pcg<1024, float> : (const void *)k_pcg<1024, double>;
    }
}
// Loop bound of a grid's solve: FP64 iterations, or (mixed) CG iterations
// plus at most as many true-residual passes.
long long pcg_kmax(const Prob &P, bool mixed) { return mixed ? 2 * P.max_iter + 4 : P.max_iter + 2; }
constexpr uint32_t kPcgThreads[3] = {256, 512, 1024};

// Blocks of `threads` threads that fit on the device at once.
fl_status pcg_resident(Batch2 *B, uint32_t threads, uint32_t *out) {
    int per_cu = 0;
    FL_HIP_TRY(B->ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pcg_kernel(threads, B->mixed),
                                                                     (int)threads, 0));
    *out = (uint32_t)std::max(per_cu, 0) * (uint32_t)B->ctx->num_cu;
    if (*out == 0)
        return set_error(B->ctx, FL_ERR_HIP, "the persistent PCG does not fit on the device");
    return FL_OK;
}

// The persistent PCG's launch shape: G co-resident blocks of `threads`
// threads (the blocks share the work out among themselves every iteration).
fl_status set_pcg_blocks(Batch2 *B, uint32_t G, uint32_t threads) {
    if (threads != 256 && threads != 512 && threads != 1024)
        return set_error(B->ctx, FL_ERR_INVALID_ARGUMENT, "PCG blocks hold 256, 512 or 1024 threads");
    uint32_t resident = 0;
    fl_status st = pcg_resident(B, threads, &resident);
    if (st != FL_OK) return st;
    G = std::max<uint32_t>(1, std::min(G, std::min(resident, B->ntiles)));
    if (G != B->pcg_blocks || threads != B->pcg_threads) drop_graphs(B);
    B->pcg_blocks = G;
    B->pcg_threads = threads;
    return FL_OK;
}

fl_status build(Batch2 *B, fl_ctx *ctx, int count, const fl_grid2d_config *cfgs,
                const uint8_t *const *solids) {
    B->ctx = ctx;
    B->count = count;
    B->cfg.assign(cfgs, cfgs + count);
    fl_status st;
    for (int p = 0; p < count; ++p)
        if ((st = validate2(ctx, cfgs[p])) != FL_OK) return st;
    // pack grids into tiles and faces
    uint32_t tile = 0;
    size_t uo = 0, vo = 0;
    B->hp.resize(count);
    std::vector<uint8_t> hflags;
    std::vector<uint32_t> htp;
    for (int p = 0; p < count; ++p) {
        const fl_grid2d_config &c = cfgs[p];
        Prob &P = B->hp[p];
        std::memset(&P, 0, sizeof(P));
        P.nx = (uint32_t)c.nx;
        P.ny = (uint32_t)c.ny;
        P.n = P.nx * P.ny;
        P.tile0 = tile;
        P.ntiles = (P.n + kT - 1) / kT;
        P.uoff = (uint32_t)uo;
        P.voff = (uint32_t)vo;
        P.dx = c.length / c.nx;
        P.dy = c.height / c.ny;
        P.cx = 1.0 / (P.dx * P.dx);
        P.cy = 1.0 / (P.dy * P.dy);
        P.nu = c.kinematic_viscosity;
        P.rho = c.density;
        P.scheme = c.convection;
        for (int s = 0; s < 4; ++s) {
            P.kind[s] = c.bc_kind[s];
            // SolverBoundaryOverrides::velocity returns (0, 0) for Symmetry and
            // PressureOutlet sides (src/preprocess.rs:1384-1391)
            P.bu[s] = c.bc_kind[s] == FL_BC_VELOCITY ? c.bc_u[s] : 0.0;
            P.bv[s] = c.bc_kind[s] == FL_BC_VELOCITY ? c.bc_v[s] : 0.0;
        }
        const bool outlet = c.bc_kind[FL_SIDE_RIGHT] == FL_BC_OUTLET;
        P.pout = outlet ? c.outlet_pressure[FL_SIDE_RIGHT] : 0.0;
        P.zm = outlet ? 0 : 1;
        P.tol = c.pressure_tolerance;
        P.max_iter = c.pressure_max_iters;
        P.eta = c.inner_tolerance > 0.0 ? c.inner_tolerance : 1.0e-5;
        P.stop_rule = c.stop_rule;
        P.guess = c.initial_guess;
        P.fx = FastDiv(P.nx);
        P.inlet = c.inlet_profile;
        P.step_h = c.inlet_step_height;
        P.height = c.height;
        P.omega = c.pressure_omega;
        P.simple = c.coupling == FL_COUPLING_SIMPLE;
        P.au = P.simple ? relaxation(c.velocity_relaxation, 0.7) : 1.0;
        P.ap = relaxation(c.pressure_relaxation, 0.3);
        if (P.simple) B->simple = true;
        const uint8_t *sol = solids ? solids[p] : nullptr;
        auto solid = [&](uint32_t i, uint32_t j) { return sol && sol[i + P.nx * j]; };
        hflags.resize((size_t)(tile + P.ntiles) * kT, 0);
        uint64_t mean_cells = 0;
        bool found_ref = false;  // find_reference_cell (:1614-1623): first non-solid cell
        for (uint32_t j = 0; j < P.ny; ++j)
            for (uint32_t i = 0; i < P.nx; ++i) {
                uint8_t f = 0;
                const bool s = solid(i, j);
                const bool out = outlet && i + 1 == P.nx && !s;  // pressure_outlet_cell
                if (i + 1 < P.nx && !solid(i + 1, j)) f |= F_E;
                if (i > 0 && !solid(i - 1, j)) f |= F_W;
                if (j + 1 < P.ny && !solid(i, j + 1)) f |= F_N;
                if (j > 0 && !solid(i, j - 1)) f |= F_S;
                if (s) f |= F_SOLID | F_ID;
                if (out) f |= F_OUTLET | F_ID;
                if (!(f & F_ID)) ++mean_cells;
                if (!s && !found_ref) {
                    P.ref = i + P.nx * j;
                    found_ref = true;
                }
                hflags[(size_t)tile * kT + i + P.nx * j] = f;
            }
        P.inv_count = mean_cells ? 1.0 / (double)mean_cells : 0.0;
        for (uint32_t t = 0; t < P.ntiles; ++t) htp.push_back(p);
        tile += P.ntiles;
        uo += (size_t)(P.nx + 1) * P.ny;
        vo += (size_t)P.nx * (P.ny + 1);
    }
    B->ntiles = tile;
    B->ncell = (size_t)tile * kT;
    B->nuf = uo;
    B->nvf = vo;
    // persistent-kernel grid: all blocks co-resident (cooperative launch)
    if (count > kMaxGrids)
        return set_error(ctx, FL_ERR_CONFIG, "a batch holds at most %d grids", kMaxGrids);
    B->mixed = cfgs[0].precision == FL_PRECISION_MIXED;
    B->sor = cfgs[0].pressure_solver == FL_SOLVER_SOR;
    for (int p = 1; p < count; ++p) {
        if (cfgs[p].precision != cfgs[0].precision)
            return set_error(ctx, FL_ERR_CONFIG, "all grids of a batch share one precision");
        if (cfgs[p].pressure_solver != cfgs[0].pressure_solver)
            return set_error(ctx, FL_ERR_CONFIG, "all grids of a batch share one pressure solver");
    }
    if ((st = dalloc(B, &B->probs, count)) || (st = dalloc(B, &B->tile_prob, B->ntiles)) ||
        (st = dalloc(B, &B->flags, B->ncell)) || (st = dalloc(B, &B->x, B->ncell)) ||
        (st = dalloc(B, &B->rhs, B->ncell)) || (st = dalloc(B, &B->r, B->ncell)) ||
        (st = dalloc(B, &B->d0, B->ncell)) || (st = dalloc(B, &B->d1, B->ncell)) ||
        (st = dalloc(B, &B->zv, B->ncell)) ||
        (st = dalloc(B, &B->q, B->ncell)) || (st = dalloc(B, &B->uc, B->ncell)) ||
        (st = dalloc(B, &B->vc, B->ncell)) || (st = dalloc(B, &B->sp, B->ncell)) ||
        (st = dalloc(B, &B->vort, B->ncell)) || (st = dalloc(B, &B->u, B->nuf)) ||
        (st = dalloc(B, &B->v, B->nvf)) || (st = dalloc(B, &B->us, B->nuf)) ||
        (st = dalloc(B, &B->vs, B->nvf)) || (st = dalloc(B, &B->uo, B->nuf)) ||
        (st = dalloc(B, &B->vo, B->nvf)) || (st = dalloc(B, &B->pa, (size_t)B->ntiles)) ||
        (st = dalloc(B, &B->pb, (size_t)B->ntiles * 5)) ||
        (st = dalloc(B, &B->p0, (size_t)B->ntiles * 2)) ||
        (st = dalloc(B, &B->dpart, (size_t)B->ntiles * 6)) || (st = dalloc(B, &B->sa, count)) ||
        (st = dalloc(B, &B->sb, count)) || (st = dalloc(B, &B->fin, 2 * count)) ||
        (st = dalloc(B, &B->active, count)) || (st = dalloc(B, &B->dts, count)) ||
        (st = dalloc(B, &B->diag, count)) || (st = dalloc(B, &B->ctl, count)) ||
        (st = dalloc(B, &B->bar, 128)) || (st = dalloc(B, &B->e32, B->mixed ? B->ncell : 1)) ||
        (st = dalloc(B, &B->sor_st, count)) || (st = dalloc(B, &B->pold, B->simple ? B->ncell : 1)))
        return st;
    if ((st = halloc(B, &B->h_sor, count)) != FL_OK) return st;
    if ((st = halloc(B, &B->h_active, count)) || (st = halloc(B, &B->h_fin, 2 * count)) ||
        (st = halloc(B, &B->h_dts, count)) || (st = halloc(B, &B->h_sa, count)) ||
        (st = halloc(B, &B->h_sb, count)) || (st = halloc(B, &B->h_diag, count)) ||
        (st = halloc(B, &B->h_ctl, count)))
        return st;
    for (int p = 0; p < count; ++p)
        B->kmax_all = std::max<long long>(B->kmax_all, pcg_kmax(B->hp[p], B->mixed));
    hipStream_t s = ctx->stream;
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->probs, B->hp.data(), count * sizeof(Prob), hipMemcpyHostToDevice, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->tile_prob, htp.data(), htp.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->flags, hflags.data(), hflags.size(), hipMemcpyHostToDevice, s));
    // Default launch shape: one 16-wave block per CU for batches that fill
    // the chip; small batches (fewer tiles than four per CU) take 4-wave
    // blocks of two tiles each, which keeps the barrier's participants few
    // and every wave busy (one 128^2 grid: 120 -> 143 steps/s).  Results do
    // not depend on the shape.
    const bool small = B->ntiles < 4u * (uint32_t)std::max(ctx->num_cu, 1);
    if ((st = small ? set_pcg_blocks(B, std::max<uint32_t>(1, B->ntiles / 2), 256)
                    : set_pcg_blocks(B, ctx->num_cu, 1024)) != FL_OK)
        return st;
    B->step.assign(count, 0);
    B->time.assign(count, 0.0);
    B->current_dt.resize(count);
    B->max_u.assign(count, 0.0);
    B->max_v.assign(count, 0.0);
    for (int p = 0; p < count; ++p) {
        B->current_dt[p] = cfgs[p].dt;
        B->h_active[p] = 1;
    }
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->active, B->h_active, count * sizeof(int), hipMemcpyHostToDevice, s));
    // initialize_velocity (:588-620): zero field, then boundary values at t = 0
    k_boundaries<<<count, kT, 0, s>>>(B->probs, B->active, B->u, B->v);
    FL_HIP_TRY(ctx, hipGetLastError());
    FL_HIP_TRY(ctx, hipStreamSynchronize(s));
    return FL_OK;
}

uint32_t ew_blocks(const Batch2 *B) { return std::min<uint32_t>(B->ntiles, B->ctx->num_cu * 16); }

// Cell fields and per-grid diagnostics into h_diag (after a synchronise).
fl_status run_cells(Batch2 *B) {
    fl_ctx *ctx = B->ctx;
    hipStream_t s = ctx->stream;
    for (int p = 0; p < B->count; ++p) {
        B->h_diag[p] = Diag{};
        for (int f = 0; f < 4; ++f) B->h_diag[p].bad[f] = ~0ull;
    }
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->diag, B->h_diag, B->count * sizeof(Diag), hipMemcpyHostToDevice, s));
    k_cells<<<ew_blocks(B), kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->flags, B->x,
                                        B->u, B->v, B->uo, B->vo, B->uc, B->vc, B->sp, B->vort,
                                        B->dpart, B->diag);
    k_fold_diag<<<B->count, kT, 0, s>>>(B->probs, B->active, B->dpart, B->diag);
    FL_HIP_TRY(ctx, hipGetLastError());
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->h_diag, B->diag, B->count * sizeof(Diag), hipMemcpyDeviceToHost, s));
    FL_HIP_TRY(ctx, hipStreamSynchronize(s));
    return FL_OK;
}

// One cooperative launch of the persistent PCG on rhs / x (the batch's own
// or scratch vectors).
fl_status launch_pcg_coop(Batch2 *B, hipStream_t s, const double *rhs, double *x, long long kmax,
                          bool capturable = false) {
    void *args[1];
    PcgArgs<double> a64{B->probs, B->tile_prob, B->ntiles, (uint32_t)B->count, kmax, B->flags, rhs,
                        x, B->r, B->zv, B->d0, B->d1, B->q, x, B->pa, B->pb, B->p0, B->sa, B->sb,
                        B->active, B->fin, B->bar};
    // the mixed path keeps its FP32 vectors in the FP64 vectors' memory
    PcgArgs<float> a32{B->probs, B->tile_prob, B->ntiles, (uint32_t)B->count, kmax, B->flags, rhs,
                       x, reinterpret_cast<float *>(B->r), reinterpret_cast<float *>(B->zv),
                       reinterpret_cast<float *>(B->d0),
                       reinterpret_cast<float *>(B->d1), reinterpret_cast<float *>(B->q), B->e32,
                       B->pa, B->pb, B->p0, B->sa, B->sb, B->active, B->fin, B->bar};
    args[0] = B->mixed ? (void *)&a32 : (void *)&a64;
    if (capturable) {
        // the extensible launch with the cooperative attribute, which a
        // stream capture records as a cooperative kernel node
        hipLaunchAttribute attr{};
        attr.id = hipLaunchAttributeCooperative;
        attr.val.cooperative = 1;
        hipLaunchConfig_t cfg{};
        cfg.gridDim = dim3(B->pcg_blocks);
        cfg.blockDim = dim3(B->pcg_threads);
        cfg.dynamicSmemBytes = 0;
        cfg.stream = s;
        cfg.attrs = &attr;
        cfg.numAttrs = 1;
        FL_HIP_TRY(B->ctx, hipLaunchKernelExC(&cfg, pcg_kernel(B->pcg_threads, B->mixed), args));
        return FL_OK;
    }
    FL_HIP_TRY(B->ctx, hipLaunchCooperativeKernel(pcg_kernel(B->pcg_threads, B->mixed), dim3(B->pcg_blocks),
                                                  dim3(B->pcg_threads), args, 0, s));
    return FL_OK;
}

// The persistent PCG over the active grids; states land in h_sa / h_sb.
fl_status run_pcg(Batch2 *B) {
    fl_ctx *ctx = B->ctx;
    hipStream_t s = ctx->stream;
    long long kmax = 1;
    for (int p = 0; p < B->count; ++p) {
        B->h_fin[p] = 0;
        B->h_fin[B->count + p] = B->h_active[p] ? 0 : 1;
        B->h_sb[p] = PcgB{};
        B->h_sb[p].first = 1;
        if (B->h_active[p]) kmax = std::max<long long>(kmax, pcg_kmax(B->hp[p], B->mixed));
    }
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->fin, B->h_fin, 2 * B->count * sizeof(int), hipMemcpyHostToDevice, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->sb, B->h_sb, B->count * sizeof(PcgB), hipMemcpyHostToDevice, s));
    fl_status st = launch_pcg_coop(B, s, B->rhs, B->x, kmax);
    if (st != FL_OK) return st;
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->h_sa, B->sa, B->count * sizeof(PcgA), hipMemcpyDeviceToHost, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->h_sb, B->sb, B->count * sizeof(PcgB), hipMemcpyDeviceToHost, s));
    FL_HIP_TRY(ctx, hipStreamSynchronize(s));
    return FL_OK;
}

// SOR for every active grid (FL_SOLVER_SOR batches); the results land in
// h_sa / h_sb like the PCG's for the step reports.
fl_status run_sor(Batch2 *B) {
    fl_ctx *ctx = B->ctx;
    hipStream_t s = ctx->stream;
    FL_HIP_TRY(ctx, hipMemsetAsync(B->sor_st, 0, B->count * sizeof(SorState), s));
    for (int p = 0; p < B->count; ++p)
        if (B->h_active[p] && B->hp[p].guess == FL_GUESS_ZERO)
            FL_HIP_TRY(ctx, hipMemsetAsync(B->x + (size_t)B->hp[p].tile0 * kT, 0, (size_t)B->hp[p].n * 8, s));
    const uint32_t G = ew_blocks(B);
    const unsigned gb = (unsigned)(B->count + 63) / 64;
    int batch = 4;
    for (;;) {
        for (int b = 0; b < batch; ++b) {
            k_sor_sweep<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->flags, B->rhs, B->x, 0u, B->sor_st);
            k_sor_sweep<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->flags, B->rhs, B->x, 1u, B->sor_st);
            k_sor_ref<<<gb, 64, 0, s>>>(B->probs, B->active, B->count, B->x, B->sor_st);
            k_sor_shift<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->flags, B->x, B->sor_st);
            k_sor_residual<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->flags, B->rhs, B->x, B->sor_st);
            k_sor_check<<<gb, 64, 0, s>>>(B->probs, B->active, B->count, B->sor_st);
        }
        FL_HIP_TRY(ctx, hipGetLastError());
        FL_HIP_TRY(ctx, hipMemcpyAsync(B->h_sor, B->sor_st, B->count * sizeof(SorState), hipMemcpyDeviceToHost, s));
        FL_HIP_TRY(ctx, hipStreamSynchronize(s));
        bool all = true;
        for (int p = 0; p < B->count; ++p)
            if (B->h_active[p] && !B->h_sor[p].done) all = false;
        if (all) break;
        batch = std::min(batch * 2, 256);
    }
    bool bad = false;
    for (int p = 0; p < B->count; ++p) {
        if (!B->h_active[p]) continue;
        const SorState &S = B->h_sor[p];
        B->h_sa[p] = PcgA{};
        B->h_sb[p] = PcgB{};
        B->h_sa[p].rmax = S.res;
        B->h_sa[p].converged = S.converged;
        B->h_sb[p].iters = S.iters;
        bad = bad || S.bad;
    }
    if (bad) return set_error(ctx, FL_ERR_SOLVER, "Pressure solver produced a non-finite residual");
    return FL_OK;
}

const char *pcg_text(int status) {
    switch (status) {
        case ST_SINGULAR: return "PCG pressure solver encountered a singular search direction";
        case ST_ALPHA: return "PCG pressure solver produced a non-finite step length";
        case ST_RESIDUAL: return "PCG pressure solver produced a non-finite residual";
        case ST_BARRIER: return "the persistent PCG's grid barrier timed out";
        default: return "";
    }
}

// After a synchronise: a persistent PCG whose barrier timed out leaves the
// abort word set; clear the barrier state and report it.
fl_status barrier_check(Batch2 *B) {
    unsigned word = 0;
    FL_HIP_TRY(B->ctx, hipMemcpy(&word, B->bar + 64, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (word == 0) return FL_OK;
    FL_HIP_TRY(B->ctx, hipMemset(B->bar, 0, 128 * sizeof(unsigned)));
    return set_error(B->ctx, FL_ERR_SOLVER, "%s", pcg_text(ST_BARRIER));
}

fl_status pcg_status(Batch2 *B, int p) {
    const int st = B->h_sb[p].status != ST_OK ? B->h_sb[p].status : B->h_sa[p].status;
    if (st != ST_OK) return set_error(B->ctx, FL_ERR_SOLVER, "%s", pcg_text(st));
    return FL_OK;
}

// advance_one_step (incompressible.rs:622-746) for every grid that has not
// reached t_end.  `strict` makes a grid at t_end an error (single solver).
fl_status step_batch(Batch2 *B, fl_grid2d_step_report *reps, bool strict) {
    fl_ctx *ctx = B->ctx;
    hipStream_t s = ctx->stream;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    int nact = 0;
    for (int p = 0; p < B->count; ++p) {
        const fl_grid2d_config &c = B->cfg[p];
        B->h_active[p] = 0;
        if (reached_end(B, p)) {
            if (strict) return set_error(ctx, FL_ERR_SOLVER, "projection simulation has already reached t_end");
            continue;
        }
        double dt = c.dt;
        if (c.adaptive) {
            fl_status st = adaptive_dt(ctx, c, B->current_dt[p], B->max_u[p], B->max_v[p],
                                       B->hp[p].dx, B->hp[p].dy, &dt);
            if (st != FL_OK) return st;
        }
        if (!B->hp[p].simple) {
            dt = std::min(dt, c.t_end - B->time[p]);
            if (dt <= 0.0) {
                if (strict) return set_error(ctx, FL_ERR_SOLVER, "projection simulation has already reached t_end");
                continue;
            }
        }
        if (c.adaptive) B->current_dt[p] = dt;
        B->h_dts[p] = dt;
        B->h_active[p] = 1;
        ++nact;
    }
    if (nact == 0) return FL_OK;
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->active, B->h_active, B->count * sizeof(int), hipMemcpyHostToDevice, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->dts, B->h_dts, B->count * sizeof(double), hipMemcpyHostToDevice, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->uo, B->u, B->nuf * sizeof(double), hipMemcpyDeviceToDevice, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->vo, B->v, B->nvf * sizeof(double), hipMemcpyDeviceToDevice, s));
    const uint32_t G = ew_blocks(B);
    k_predict<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->dts, B->flags,
                               B->x, B->u, B->v, B->us, B->vs);
    k_boundaries<<<B->count, kT, 0, s>>>(B->probs, B->active, B->us, B->vs);
    k_rhs<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->dts, B->flags, B->us,
                           B->vs, B->rhs);
    if (B->simple) k_simple_save<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->x, B->pold);
    FL_HIP_TRY(ctx, hipGetLastError());
    fl_status st = B->sor ? run_sor(B) : run_pcg(B);
    if (st != FL_OK || (st = barrier_check(B)) != FL_OK) return st;
    for (int p = 0; p < B->count; ++p)
        if (B->h_active[p] && (st = pcg_status(B, p)) != FL_OK) return st;
    k_correct<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->dts, B->flags, B->x,
                               B->us, B->vs, B->u, B->v);
    if (B->simple) k_simple_combine<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->x, B->pold);
    k_boundaries<<<B->count, kT, 0, s>>>(B->probs, B->active, B->u, B->v);
    FL_HIP_TRY(ctx, hipGetLastError());
    if ((st = run_cells(B)) != FL_OK) return st;
    for (int p = 0; p < B->count; ++p) {
        if (!B->h_active[p]) continue;
        const Prob &P = B->hp[p];
        const Diag &dg = B->h_diag[p];
        B->step[p] += 1;
        B->time[p] = P.simple ? (double)B->step[p] * B->cfg[p].dt : B->time[p] + B->h_dts[p];
        // ensure_finite (:1707-1742): pressure, u, v, speed in that order
        const char *names[4] = {"pressure", "u", "v", "speed"};
        double *arrs[4] = {B->x + (size_t)P.tile0 * kT, B->u + P.uoff, B->v + P.voff,
                           B->sp + (size_t)P.tile0 * kT};
        for (int f = 0; f < 4; ++f)
            if (dg.bad[f] != ~0ull) {
                double val = 0.0;
                FL_HIP_TRY(ctx, hipMemcpy(&val, arrs[f] + dg.bad[f], sizeof(double), hipMemcpyDeviceToHost));
                return set_error(ctx, FL_ERR_SOLVER, "Non-finite %s value at flat index %llu: %g (step %lld)",
                                 names[f], dg.bad[f], val, (long long)B->step[p]);
            }
        B->max_u[p] = dg.max_u;
        B->max_v[p] = dg.max_v;
        if (reps) {
            fl_grid2d_step_report &r = reps[p];
            const double dt = B->h_dts[p];
            r.iteration = B->step[p];
            r.time = B->time[p];
            r.time_step = dt;
            r.continuity_residual = dg.max_div;
            r.pressure_residual = B->h_sa[p].rmax;
            r.pressure_iterations = B->h_sb[p].iters;
            r.momentum_residual = std::max(dg.du, dg.dv);
            r.max_speed = dg.max_speed;
            r.momentum_cfl = dt * (dg.max_u / P.dx + dg.max_v / P.dy);
            r.viscous_diffusion_number = P.nu * dt * (1.0 / (P.dx * P.dx) + 1.0 / (P.dy * P.dy));
            r.pressure_converged = B->h_sa[p].converged;
            r.reached_end_time = reached_end(B, p);
        }
    }
    return FL_OK;
}

// ------------------------------------------------ device-driven runs
// The step without host decisions: setup, predictor, RHS (before the PCG)
// and correction, cell fields, diagnostics, finish (after it).
fl_status launch_pre(Batch2 *B, hipStream_t s) {
    fl_ctx *ctx = B->ctx;
    const uint32_t G = ew_blocks(B);
    k_run_setup<<<(B->count + 63) / 64, 64, 0, s>>>(B->probs, B->ctl, B->count, B->active, B->dts,
                                                    B->fin, B->sb, B->diag);
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->uo, B->u, B->nuf * sizeof(double), hipMemcpyDeviceToDevice, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->vo, B->v, B->nvf * sizeof(double), hipMemcpyDeviceToDevice, s));
    k_predict<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->dts, B->flags,
                               B->x, B->u, B->v, B->us, B->vs);
    k_boundaries<<<B->count, kT, 0, s>>>(B->probs, B->active, B->us, B->vs);
    k_rhs<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->dts, B->flags, B->us,
                           B->vs, B->rhs);
    if (B->simple) k_simple_save<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->x, B->pold);
    FL_HIP_TRY(ctx, hipGetLastError());
    return FL_OK;
}
fl_status launch_post(Batch2 *B, hipStream_t s) {
    fl_ctx *ctx = B->ctx;
    const uint32_t G = ew_blocks(B);
    k_correct<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->dts, B->flags, B->x,
                               B->us, B->vs, B->u, B->v);
    if (B->simple) k_simple_combine<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->x, B->pold);
    k_boundaries<<<B->count, kT, 0, s>>>(B->probs, B->active, B->u, B->v);
    k_cells<<<G, kT, 0, s>>>(B->probs, B->tile_prob, B->ntiles, B->active, B->flags, B->x, B->u,
                             B->v, B->uo, B->vo, B->uc, B->vc, B->sp, B->vort, B->dpart, B->diag);
    k_fold_diag<<<B->count, kT, 0, s>>>(B->probs, B->active, B->dpart, B->diag);
    k_run_finish<<<(B->count + 63) / 64, 64, 0, s>>>(B->ctl, B->count, B->active, B->dts, B->diag,
                                                     B->sa, B->sb, B->bar);
    FL_HIP_TRY(ctx, hipGetLastError());
    return FL_OK;
}

// Captures part of a step: 0 the part before the PCG, 1 the part after it,
// 2 the whole step with the PCG as a cooperative kernel node.
fl_status capture(Batch2 *B, int part, hipGraph_t *g, hipGraphExec_t *x) {
    fl_ctx *ctx = B->ctx;
    hipStream_t s = ctx->stream;
    FL_HIP_TRY(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    fl_status st = FL_OK;
    if (part != 1) st = launch_pre(B, s);
    if (st == FL_OK && part == 2) st = launch_pcg_coop(B, s, B->rhs, B->x, B->kmax_all, true);
    if (st == FL_OK && part != 0) st = launch_post(B, s);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (st == FL_OK && e != hipSuccess)
        st = set_error(ctx, FL_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
    if (st == FL_OK && hipGraphInstantiate(x, graph, nullptr, nullptr, 0) != hipSuccess)
        st = set_error(ctx, FL_ERR_HIP, "graph instantiation failed");
    if (st != FL_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return st;
    }
    *g = graph;
    return FL_OK;
}

// The step graph: one graph for the whole step where the runtime records the
// cooperative launch, else graphs of the two halves around it.
fl_status capture_step(Batch2 *B) {
    if (B->graph_mode != FL_GRAPH_NONE) return FL_OK;
    if (capture(B, 2, &B->gr_step, &B->gx_step) == FL_OK) {
        B->graph_mode = FL_GRAPH_WHOLE_STEP;
        return FL_OK;
    }
    fl_status st;
    if ((st = capture(B, 0, &B->gr_pre, &B->gx_pre)) != FL_OK) return st;
    if ((st = capture(B, 1, &B->gr_post, &B->gx_post)) != FL_OK) return st;
    B->graph_mode = FL_GRAPH_SPLIT;
    return FL_OK;
}

fl_status run_device(Batch2 *B, int64_t steps, bool graph, fl_batch2d_run_report *reps,
                     double *step_ms, double *pcg_ms) {
    fl_ctx *ctx = B->ctx;
    hipStream_t s = ctx->stream;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (B->sor)
        return set_error(ctx, FL_ERR_CONFIG, "device-driven batch runs use the PCG pressure solver");
    for (int p = 0; p < B->count; ++p) {
        const fl_grid2d_config &c = B->cfg[p];
        StepCtl &k = B->h_ctl[p];
        k = StepCtl{};
        k.dt0 = c.dt;
        k.t_end = c.t_end;
        k.min_dt = c.min_dt;
        k.max_dt = c.max_dt;
        k.target_cfl = c.target_cfl;
        k.target_diff = c.target_diffusion_number;
        k.growth = c.max_growth_factor;
        k.adaptive = c.adaptive;
        k.simple = B->hp[p].simple;
        k.time = B->time[p];
        k.current_dt = B->current_dt[p];
        k.max_u = B->max_u[p];
        k.max_v = B->max_v[p];
        k.step = B->step[p];
    }
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->ctl, B->h_ctl, B->count * sizeof(StepCtl), hipMemcpyHostToDevice, s));
    fl_status st;
    if (graph && (st = capture_step(B)) != FL_OK) return st;
    const bool whole = graph && B->graph_mode == FL_GRAPH_WHOLE_STEP;
    const bool timed = step_ms || (pcg_ms && !whole);
    if (timed)
        while ((int64_t)B->ev.size() < 4 * steps) {
            hipEvent_t e;
            FL_HIP_TRY(ctx, hipEventCreate(&e));
            B->ev.push_back(e);
        }
    for (int64_t k = 0; k < steps; ++k) {
        if (timed) FL_HIP_TRY(ctx, hipEventRecord(B->ev[4 * k], s));
        if (whole) {
            FL_HIP_TRY(ctx, hipGraphLaunch(B->gx_step, s));
        } else {
            if (graph) FL_HIP_TRY(ctx, hipGraphLaunch(B->gx_pre, s));
            else if ((st = launch_pre(B, s)) != FL_OK) return st;
            if (timed) FL_HIP_TRY(ctx, hipEventRecord(B->ev[4 * k + 1], s));
            if ((st = launch_pcg_coop(B, s, B->rhs, B->x, B->kmax_all)) != FL_OK) return st;
            if (timed) FL_HIP_TRY(ctx, hipEventRecord(B->ev[4 * k + 2], s));
            if (graph) FL_HIP_TRY(ctx, hipGraphLaunch(B->gx_post, s));
            else if ((st = launch_post(B, s)) != FL_OK) return st;
        }
        if (timed) FL_HIP_TRY(ctx, hipEventRecord(B->ev[4 * k + 3], s));
    }
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->h_ctl, B->ctl, B->count * sizeof(StepCtl), hipMemcpyDeviceToHost, s));
    FL_HIP_TRY(ctx, hipStreamSynchronize(s));
    for (int64_t k = 0; k < steps && timed; ++k) {
        float ms = 0.0f;
        if (step_ms) {
            FL_HIP_TRY(ctx, hipEventElapsedTime(&ms, B->ev[4 * k], B->ev[4 * k + 3]));
            step_ms[k] = ms;
        }
        if (pcg_ms && !whole) {
            FL_HIP_TRY(ctx, hipEventElapsedTime(&ms, B->ev[4 * k + 1], B->ev[4 * k + 2]));
            pcg_ms[k] = ms;
        }
    }
    if (pcg_ms && whole)
        for (int64_t k = 0; k < steps; ++k) pcg_ms[k] = -1.0;
    int first_bad = -1;
    const int64_t nsteps = steps;
    for (int p = 0; p < B->count; ++p) {
        const StepCtl &k = B->h_ctl[p];
        const int64_t taken = k.step - B->step[p];
        B->step[p] = k.step;
        B->time[p] = k.time;
        B->current_dt[p] = k.current_dt;
        B->max_u[p] = k.max_u;
        B->max_v[p] = k.max_v;
        if (k.error != RE_NONE && first_bad < 0) first_bad = p;
        if (reps) {
            fl_batch2d_run_report &r = reps[p];
            r.steps = taken;
            r.time = k.time;
            r.pressure_iterations = k.iters;
            r.last_pressure_residual = k.last_residual;
            r.max_divergence = k.max_div;
            r.max_speed = k.max_speed;
            r.reached_end_time = !k.simple && ctl_reached_end(k.time, k.t_end) ? 1 : 0;
            r.status = k.error == RE_NONE ? FL_OK : FL_ERR_SOLVER;
        }
    }
    (void)nsteps;
    if ((st = barrier_check(B)) != FL_OK) return st;
    if (first_bad < 0) return FL_OK;
    const StepCtl &k = B->h_ctl[first_bad];
    if (k.error == RE_MIN_DT)
        return set_error(ctx, FL_ERR_SOLVER,
                         "required stable time step %.3e is below configured min_dt %.3e; CFL "
                         "constraint=%.3e, diffusion constraint=%.3e",
                         k.vals[0], k.min_dt, k.vals[1], k.vals[2]);
    if (k.error == RE_PCG) return set_error(ctx, FL_ERR_SOLVER, "%s", pcg_text(k.which));
    const Prob &P = B->hp[first_bad];
    const char *names[4] = {"pressure", "u", "v", "speed"};
    const double *arrs[4] = {B->x + (size_t)P.tile0 * kT, B->u + P.uoff, B->v + P.voff,
                             B->sp + (size_t)P.tile0 * kT};
    double val = 0.0;
    FL_HIP_TRY(ctx, hipMemcpy(&val, arrs[k.which] + k.index, sizeof(double), hipMemcpyDeviceToHost));
    return set_error(ctx, FL_ERR_SOLVER, "Non-finite %s value at flat index %llu: %g (step %lld)",
                     names[k.which], k.index, val, (long long)k.step);
}

// Chooses the persistent PCG's launch shape: every candidate (threads per
// block, co-resident block count) solves
// the same pseudo-random right-hand sides from zero with each grid's own
// tolerance, capped at `iters` iterations, into scratch x and rhs (the
// solver's fields are untouched), and the fastest is kept.  Solver results
// do not depend on the choice.
fl_status autotune(Batch2 *B, int iters, fl_batch2d_tuning *chosen, int32_t cap,
                   fl_batch2d_tuning *table, int32_t *tlen) {
    fl_ctx *ctx = B->ctx;
    hipStream_t s = ctx->stream;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::vector<fl_batch2d_tuning> cand;
    fl_status st;
    for (uint32_t th : kPcgThreads) {
        uint32_t resident = 0;
        if ((st = pcg_resident(B, th, &resident)) != FL_OK) return st;
        std::vector<uint32_t> gs;
        for (uint32_t m : {1u, 2u, 4u, 8u})
            for (uint32_t G : {(uint32_t)ctx->num_cu * m / 2, (uint32_t)ctx->num_cu * m})
                if (G >= 1 && G <= resident && G <= B->ntiles &&
                    std::find(gs.begin(), gs.end(), G) == gs.end())
                    gs.push_back(G);
        if (gs.empty()) gs.push_back(std::min<uint32_t>(B->ntiles, resident));
        for (uint32_t G : gs) {
            fl_batch2d_tuning c{};
            c.threads = (int32_t)th;
            c.blocks = (int32_t)G;
            cand.push_back(c);
        }
    }
    const fl_batch2d_tuning keep{(int32_t)B->pcg_threads, (int32_t)B->pcg_blocks, {0, 0}, 0.0};
    double *rhs = nullptr, *x = nullptr;
    FL_HIP_TRY(ctx, hipMalloc(&rhs, B->ncell * sizeof(double)));
    if (hipMalloc(&x, B->ncell * sizeof(double)) != hipSuccess) {
        (void)hipFree(rhs);
        return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "autotune scratch allocation failed");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    st = FL_OK;
    if (tlen) *tlen = 0;
    fl_batch2d_tuning best = keep;
    best.ms = 1e300;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        st = set_error(ctx, FL_ERR_HIP, "event creation failed");
    if (st == FL_OK) {
        k_fill_rhs<<<256, 256, 0, s>>>(rhs, B->ncell, 12345);
        for (int p = 0; p < B->count; ++p) {
            B->h_active[p] = 1;
            B->h_fin[p] = 0;
            B->h_fin[B->count + p] = 0;
            B->h_sb[p] = PcgB{};
            B->h_sb[p].first = 1;
        }
        if (hipMemcpyAsync(B->active, B->h_active, B->count * sizeof(int), hipMemcpyHostToDevice, s) !=
            hipSuccess)
            st = set_error(ctx, FL_ERR_HIP, "autotune setup failed");
    }
    // an untimed solve with the first candidate, then one timed solve each
    for (size_t ci = 0; ci <= cand.size() && st == FL_OK; ++ci) {
        fl_batch2d_tuning &c = cand[ci == 0 ? 0 : ci - 1];
        if ((st = set_pcg_blocks(B, (uint32_t)c.blocks, (uint32_t)c.threads)) != FL_OK) break;
        c.blocks = (int32_t)B->pcg_blocks;
        if (hipMemsetAsync(x, 0, B->ncell * sizeof(double), s) != hipSuccess ||
            hipMemcpyAsync(B->fin, B->h_fin, 2 * B->count * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(B->sb, B->h_sb, B->count * sizeof(PcgB), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipEventRecord(e0, s) != hipSuccess) {
            st = set_error(ctx, FL_ERR_HIP, "autotune setup failed");
            break;
        }
        if ((st = launch_pcg_coop(B, s, rhs, x, iters)) != FL_OK) break;
        float ms = 0.0f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            st = set_error(ctx, FL_ERR_HIP, "autotune timing failed");
            break;
        }
        if (ci == 0) continue;
        c.ms = ms;
        if (table && tlen && *tlen < cap) table[(*tlen)++] = c;
        if (c.ms < best.ms) best = c;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    (void)hipFree(rhs);
    (void)hipFree(x);
    if (st == FL_OK) st = barrier_check(B);
    if (st != FL_OK) {
        (void)set_pcg_blocks(B, (uint32_t)keep.blocks, (uint32_t)keep.threads);
        return st;
    }
    if ((st = set_pcg_blocks(B, (uint32_t)best.blocks, (uint32_t)best.threads)) != FL_OK)
        return st;
    FL_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (chosen) *chosen = best;
    return FL_OK;
}

fl_status create_common(fl_ctx *ctx, int count, const fl_grid2d_config *cfgs,
                        const uint8_t *const *solids, Batch2 **out) {
    if (!ctx || !cfgs || !out || count < 1)
        return set_error(ctx, FL_ERR_INVALID_ARGUMENT, "null argument or empty batch");
    *out = nullptr;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    Batch2 *B = new (std::nothrow) Batch2();
    if (!B) return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "host allocation failed");
    fl_status st = build(B, ctx, count, cfgs, solids);
    if (st == FL_OK) st = run_cells(B);  // cell fields and |u|, |v| of the initial state
    if (st != FL_OK) {
        free_batch(B);
        delete B;
        return st;
    }
    ctx_retain(ctx);
    *out = B;
    return FL_OK;
}

void destroy_common(Batch2 *B) {
    if (!B) return;
    fl_ctx *ctx = B->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_batch(B);
    delete B;
    ctx_release(ctx);
}

fl_status get_fields(Batch2 *B, int p, double *u, double *v, double *pp) {
    fl_ctx *ctx = B->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const Prob &P = B->hp[p];
    hipStream_t s = ctx->stream;
    if (u) FL_HIP_TRY(ctx, hipMemcpyAsync(u, B->u + P.uoff, (size_t)(P.nx + 1) * P.ny * 8, hipMemcpyDeviceToHost, s));
    if (v) FL_HIP_TRY(ctx, hipMemcpyAsync(v, B->v + P.voff, (size_t)P.nx * (P.ny + 1) * 8, hipMemcpyDeviceToHost, s));
    if (pp) FL_HIP_TRY(ctx, hipMemcpyAsync(pp, B->x + (size_t)P.tile0 * kT, (size_t)P.n * 8, hipMemcpyDeviceToHost, s));
    FL_HIP_TRY(ctx, hipStreamSynchronize(s));
    return FL_OK;
}

}  // namespace

struct fl_grid2d {
    Batch2 *b;
};
struct fl_batch2d {
    Batch2 *b;
};

extern "C" {

fl_status fl_grid2d_create(fl_ctx *ctx, const fl_grid2d_config *cfg, const uint8_t *solid,
                           fl_grid2d **out) {
    if (!out) return set_error(ctx, FL_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    const uint8_t *solids[1] = {solid};
    Batch2 *B = nullptr;
    fl_status st = create_common(ctx, 1, cfg, solids, &B);
    if (st != FL_OK) return st;
    fl_grid2d *g = new (std::nothrow) fl_grid2d{B};
    if (!g) {
        destroy_common(B);
        return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    *out = g;
    return FL_OK;
}

void fl_grid2d_destroy(fl_grid2d *g) {
    if (!g) return;
    destroy_common(g->b);
    delete g;
}

fl_status fl_grid2d_step(fl_grid2d *g, fl_grid2d_step_report *rep) {
    if (!g) return FL_ERR_INVALID_ARGUMENT;
    return step_batch(g->b, rep, true);
}

fl_status fl_grid2d_pressure_solve(fl_grid2d *g, const double *rhs, double *p_inout,
                                   fl_solve_report *rep) {
    if (!g || !rhs || !p_inout) return FL_ERR_INVALID_ARGUMENT;
    Batch2 *B = g->b;
    fl_ctx *ctx = B->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const Prob &P = B->hp[0];
    for (uint32_t c = 0; c < P.n; ++c)
        if (!std::isfinite(rhs[c]))
            return set_error(ctx, FL_ERR_NON_FINITE, "non-finite right-hand side value at index %u", c);
    hipStream_t s = ctx->stream;
    B->h_active[0] = 1;
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->active, B->h_active, sizeof(int), hipMemcpyHostToDevice, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->rhs, rhs, (size_t)P.n * 8, hipMemcpyHostToDevice, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->x, p_inout, (size_t)P.n * 8, hipMemcpyHostToDevice, s));
    fl_status st = B->sor ? run_sor(B) : run_pcg(B);
    if (st != FL_OK || (st = barrier_check(B)) != FL_OK) return st;
    FL_HIP_TRY(ctx, hipMemcpyAsync(p_inout, B->x, (size_t)P.n * 8, hipMemcpyDeviceToHost, s));
    FL_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (rep) {
        rep->iterations = B->h_sb[0].iters;
        rep->refinements = 0;
        rep->fallback = 0;
        rep->warm_start = 0;
        rep->fallback_reason = FL_FALLBACK_NONE;
        rep->warm_start_ratio = 0.0;
        rep->fallback_iterations = 0;
        rep->residual_abs_inf = B->h_sa[0].rmax;
        rep->residual_l2 = std::sqrt(B->h_sa[0].rr);
        rep->initial_l2 = std::sqrt(B->h_sa[0].rr0);
        rep->converged = B->h_sa[0].converged;
        rep->status = FL_OK;
    }
    st = pcg_status(B, 0);
    if (rep && st != FL_OK) rep->status = st;
    return st;
}

fl_status fl_grid2d_get_fields(fl_grid2d *g, double *u, double *v, double *p) {
    if (!g) return FL_ERR_INVALID_ARGUMENT;
    return get_fields(g->b, 0, u, v, p);
}

fl_status fl_grid2d_set_fields(fl_grid2d *g, const double *u, const double *v, const double *p) {
    if (!g) return FL_ERR_INVALID_ARGUMENT;
    Batch2 *B = g->b;
    fl_ctx *ctx = B->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const Prob &P = B->hp[0];
    hipStream_t s = ctx->stream;
    if (u) FL_HIP_TRY(ctx, hipMemcpyAsync(B->u + P.uoff, u, (size_t)(P.nx + 1) * P.ny * 8, hipMemcpyHostToDevice, s));
    if (v) FL_HIP_TRY(ctx, hipMemcpyAsync(B->v + P.voff, v, (size_t)P.nx * (P.ny + 1) * 8, hipMemcpyHostToDevice, s));
    if (p) FL_HIP_TRY(ctx, hipMemcpyAsync(B->x + (size_t)P.tile0 * kT, p, (size_t)P.n * 8, hipMemcpyHostToDevice, s));
    B->h_active[0] = 1;
    FL_HIP_TRY(ctx, hipMemcpyAsync(B->active, B->h_active, sizeof(int), hipMemcpyHostToDevice, s));
    fl_status st = run_cells(B);
    if (st == FL_OK) {
        B->max_u[0] = B->h_diag[0].max_u;
        B->max_v[0] = B->h_diag[0].max_v;
    }
    return st;
}

fl_status fl_grid2d_cell_fields(fl_grid2d *g, double *u_cell, double *v_cell, double *speed,
                                double *vorticity) {
    if (!g) return FL_ERR_INVALID_ARGUMENT;
    Batch2 *B = g->b;
    fl_ctx *ctx = B->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const Prob &P = B->hp[0];
    hipStream_t s = ctx->stream;
    const size_t base = (size_t)P.tile0 * kT, bytes = (size_t)P.n * 8;
    if (u_cell) FL_HIP_TRY(ctx, hipMemcpyAsync(u_cell, B->uc + base, bytes, hipMemcpyDeviceToHost, s));
    if (v_cell) FL_HIP_TRY(ctx, hipMemcpyAsync(v_cell, B->vc + base, bytes, hipMemcpyDeviceToHost, s));
    if (speed) FL_HIP_TRY(ctx, hipMemcpyAsync(speed, B->sp + base, bytes, hipMemcpyDeviceToHost, s));
    if (vorticity) FL_HIP_TRY(ctx, hipMemcpyAsync(vorticity, B->vort + base, bytes, hipMemcpyDeviceToHost, s));
    FL_HIP_TRY(ctx, hipStreamSynchronize(s));
    return FL_OK;
}

fl_status fl_batch2d_create(fl_ctx *ctx, int32_t count, const fl_grid2d_config *cfgs,
                            fl_batch2d **out) {
    if (!out) return set_error(ctx, FL_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    Batch2 *B = nullptr;
    fl_status st = create_common(ctx, count, cfgs, nullptr, &B);
    if (st != FL_OK) return st;
    fl_batch2d *b = new (std::nothrow) fl_batch2d{B};
    if (!b) {
        destroy_common(B);
        return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    *out = b;
    return FL_OK;
}

void fl_batch2d_destroy(fl_batch2d *b) {
    if (!b) return;
    destroy_common(b->b);
    delete b;
}

fl_status fl_batch2d_step(fl_batch2d *b, fl_grid2d_step_report *reps) {
    if (!b) return FL_ERR_INVALID_ARGUMENT;
    return step_batch(b->b, reps, false);
}

fl_status fl_batch2d_get_fields(fl_batch2d *b, int32_t index, double *u, double *v, double *p) {
    if (!b || index < 0 || index >= b->b->count) return FL_ERR_INVALID_ARGUMENT;
    return get_fields(b->b, index, u, v, p);
}

fl_status fl_batch2d_run(fl_batch2d *b, int64_t steps, int32_t use_graphs,
                         fl_batch2d_run_report *reports, double *step_ms, double *pcg_ms) {
    if (!b) return FL_ERR_INVALID_ARGUMENT;
    if (steps < 0) return set_error(b->b->ctx, FL_ERR_INVALID_ARGUMENT, "negative step count");
    return run_device(b->b, steps, use_graphs != 0, reports, step_ms, pcg_ms);
}

fl_status fl_batch2d_autotune(fl_batch2d *b, int32_t iterations, fl_batch2d_tuning *chosen,
                              int32_t table_cap, fl_batch2d_tuning *table, int32_t *table_len) {
    if (!b) return FL_ERR_INVALID_ARGUMENT;
    if (iterations < 1)
        return set_error(b->b->ctx, FL_ERR_INVALID_ARGUMENT, "autotune needs at least one iteration");
    return autotune(b->b, iterations, chosen, table_cap, table, table_len);
}

fl_status fl_batch2d_set_launch(fl_batch2d *b, const fl_batch2d_tuning *launch) {
    if (!b || !launch) return FL_ERR_INVALID_ARGUMENT;
    Batch2 *B = b->b;
    FL_HIP_TRY(B->ctx, hipSetDevice(B->ctx->device));
    fl_status st = set_pcg_blocks(B, (uint32_t)std::max(launch->blocks, 1), (uint32_t)launch->threads);
    if (st != FL_OK) return st;
    FL_HIP_TRY(B->ctx, hipStreamSynchronize(B->ctx->stream));
    return FL_OK;
}

fl_status fl_batch2d_graph_mode(const fl_batch2d *b, int32_t *mode) {
    if (!b || !mode) return FL_ERR_INVALID_ARGUMENT;
    *mode = b->b->graph_mode;
    return FL_OK;
}

fl_status fl_batch2d_get_launch(const fl_batch2d *b, fl_batch2d_tuning *launch) {
    if (!b || !launch) return FL_ERR_INVALID_ARGUMENT;
    const Batch2 *B = b->b;
    *launch = fl_batch2d_tuning{(int32_t)B->pcg_threads, (int32_t)B->pcg_blocks, {0, 0}, 0.0};
    return FL_OK;
}

}  // extern "C"
