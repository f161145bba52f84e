// 3D lid-driven cavity projection step on gfx950.
//
// Replaces LidDrivenCavity3DSolver (src/solver/incompressible_3d.rs):
//   predict_momentum :174-273   -> k_predict3_t
//   build_pressure_rhs :275-291 -> k_rhs
//   solve_pressure_poisson :293-332 -> pinned Jacobi-PCG; the reference only
//       has SOR here, the PCG loop is the 2D solver's
//       (src/solver/incompressible.rs:1048-1214) on the 7-point operator with
//       cell (0,0,0) eliminated (src/poisson.rs:303-337 style)
//   correct_velocity :334-363 and
//   max_divergence / max_speed / ensure_finite :504-554 -> k_correct_diag_t (+ k_diag_fold)
//
// Data layout in HBM: the reference's own (src/field.rs:104-107), one flat
// array per field, x fastest.  The PCG vectors x (= p), r, q and two
// direction buffers d0/d1 (ping-pong) live beside the fields; z = D^-1 r is
// never stored.
//
// One PCG iteration is two bandwidth-bound kernels (10 vector passes):
//   k_cg_dirapply  d' = D^-1 r + beta d,  q = A d',  d'.q     (reads r d, writes d' q)
//   k_cg_update    x += alpha d', r -= alpha q,  r.z, r.r, max|r|  (reads x d' r q, writes x r)
// The direction update is fused into the operator apply: the halo values of
// d' are recomputed from r and d, so d' is written once and never re-read
// for the stencil.  Scalars never leave the device: each kernel folds the
// other kernel's per-block partials itself (common.hpp) and block 0 records
// the state the host polls.  Stencil loads use clamped indices (the
// reference's ghost = centre rule makes a missing neighbour contribute
// nothing), so all seven loads of a cell are in flight together.
//
// Mixed precision (FL_PRECISION_MIXED): the same two kernels run on FP32
// vectors (dots, alpha, beta and the stop test stay FP64) as the inner
// solver of FP64 iterative refinement: r = b - A x is recomputed in FP64
// (k_residual_z), the inner CG solves A e = r to a relative tolerance, and
// x += e in FP64, until the outer FP64 stop test holds.

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "kprof.hpp"
#include "unet.hpp"

namespace fl {
namespace c3 {

// The Jacobi diagonal of the Neumann operator is the sum of the coefficients
// of the neighbours that exist (ghost = centre, incompressible_3d.rs:450-474).
// Every axis contributes 1 or 2 neighbours, so there are eight diagonals;
// their reciprocals are tabulated (bit a of the index set = two neighbours
// along axis a) and z = D^-1 r is a multiply.
struct Grid {
    uint32_t nx, ny, nz, nxy, n;
    double ix2, iy2, iz2;
    FastDiv fx, fxy;
    double inv[8];
    float invf[8];
};

// Two-cell vector types (marches and streams with two cells per lane).
template <typename T> struct Pair;
template <> struct Pair<double> { typedef double2 type; };
template <> struct Pair<float> { typedef float2 type; };

__device__ inline void cell_ijk(const Grid &g, uint32_t c, uint32_t &i, uint32_t &j,
                                uint32_t &k) {
    k = g.fxy.div(c);
    const uint32_t rem = c - k * g.nxy;
    j = g.fx.div(rem);
    i = rem - j * g.nx;
}

__device__ inline uint32_t diag_index(const Grid &g, uint32_t i, uint32_t j, uint32_t k) {
    return (uint32_t)(i > 0 && i + 1 < g.nx) | ((uint32_t)(j > 0 && j + 1 < g.ny) << 1) |
           ((uint32_t)(k > 0 && k + 1 < g.nz) << 2);
}
template <typename T> __device__ inline T inv_diag(const Grid &g, uint32_t idx);
template <> __device__ inline double inv_diag<double>(const Grid &g, uint32_t idx) { return g.inv[idx]; }
template <> __device__ inline float inv_diag<float>(const Grid &g, uint32_t idx) { return g.invf[idx]; }

// The 8 reciprocal diagonals in LDS: indexed per lane, the kernel-argument
// copy would be a vector load from the argument segment per lookup.  Every
// thread of the block must call it (it ends with a block barrier).
template <typename T>
__device__ inline const T *inv_table(const Grid &g) {
    __shared__ T tab[8];
    if (threadIdx.x < 8) tab[threadIdx.x] = inv_diag<T>(g, threadIdx.x);
    __syncthreads();
    return tab;
}

// Neighbour offsets clamped at the walls (a missing neighbour reads the
// centre, contributing c * (x_c - x_c) = 0).
struct Nbr {
    uint32_t e, w, n, s, f, b;
};
__device__ inline Nbr neighbours(const Grid &g, uint32_t c, uint32_t i, uint32_t j, uint32_t k) {
    Nbr o;
    o.e = i + 1 < g.nx ? c + 1 : c;
    o.w = i > 0 ? c - 1 : c;
    o.n = j + 1 < g.ny ? c + g.nx : c;
    o.s = j > 0 ? c - g.nx : c;
    o.f = k + 1 < g.nz ? c + g.nxy : c;
    o.b = k > 0 ? c - g.nxy : c;
    return o;
}

// Host-polled state.  SA is written only by block 0 of k_cg_dirapply, SB
// only by block 0 of k_cg_update (and the host before a solve); a kernel
// never reads a field of its own struct that its block 0 may change, except
// `done`, which block 0 raises only when every block reaches the same
// verdict from the same folded partials.
struct CgStateA {
    double rz, beta, rr0, rr, rmax;
    int done, converged, status;
    int first;  // this iteration's direction is z itself (no beta d term)
};
struct CgStateB {
    double rz_prev, alpha;
    long long iters;
    int done, status, first, pad;
};
struct CgParams {
    double tol;
    long long max_iter;
    int stop_rule;
    double rr0_ref;  // > 0: the rel-L2 test is against this ||r0||^2
    double e_bound;  // multigrid FCG: bound of |A^-1 r| / max|r| (L_max^2 / pi^2)
    // Jacobi CG on the Neumann operator (row 0 a regular row) instead of the
    // pinned one; the right-hand side must be compatible (sum r = 0)
    int neumann;
};
enum { ST_OK = 0, ST_SINGULAR = 1, ST_ALPHA = 2, ST_RESIDUAL = 3 };

template <typename T, typename TX>
__global__ void __launch_bounds__(kBlock) k_cg_update(Grid g, TX *__restrict__ x,
                                                      T *__restrict__ r,
                                                      const T *__restrict__ d,
                                                      const T *__restrict__ q,
                                                      const double *__restrict__ pin,
                                                      uint32_t rows, double *__restrict__ pout,
                                                      const CgStateA *SA, CgStateB *SB, int neumann) {
    if (SA->done || SB->done) return;
    const T *inv_tab = inv_table<T>(g);
    double f[1], fm[1];
    fold_partials<1, 1>(pin, rows, f, fm);
    const double den = f[0];
    int status = ST_OK;
    double alpha = 0.0;
    if (fabs(den) < 1.0e-30 || !isfinite(den)) {
        status = ST_SINGULAR;
    } else {
        alpha = SA->rz / den;
        if (!isfinite(alpha)) status = ST_ALPHA;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        SB->rz_prev = SA->rz;
        SB->alpha = alpha;
        SB->status = status;
        SB->first = 0;
        if (status != ST_OK) SB->done = 1;
        else SB->iters += 1;
    }
    if (status != ST_OK) return;
    const T ta = (T)alpha;
    double s[3] = {0.0, 0.0, 0.0}, m[1] = {0.0};
    auto cell = [&](uint32_t c, TX xc, T dc, T rc0, T qc, TX &xo, T &ro) {
        xo = xc + (TX)(ta * dc);
        const T rc = rc0 - ta * qc;
        ro = rc;
        T z = rc;  // the same z = D^-1 r the direction kernel forms
        if (c != 0 || neumann) {
            uint32_t i, j, k;
            cell_ijk(g, c, i, j, k);
            z = rc * inv_tab[diag_index(g, i, j, k)];
        }
        const double rd = (double)rc;
        s[0] += rd * (double)z;
        s[1] += rd * rd;
        s[2] += rd;
        m[0] = fmax(m[0], fabs(rd));
    };
    if (sizeof(T) == 8 && sizeof(TX) == 8 && g.n % 2 == 0) {
        // FP64: two cells per lane (16-byte loads and stores), 512-cell tiles
        typedef typename Pair<T>::type T2;
        TileSched ts((g.n / 2 + kBlock - 1) / kBlock);
        for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
            const uint32_t c = 2 * (t * kBlock + threadIdx.x);
            if (c >= g.n) continue;
            const T2 xc = *(const T2 *)(x + c), dc = *(const T2 *)(d + c), rc = *(const T2 *)(r + c),
                     qc = *(const T2 *)(q + c);
            T2 xo, ro;
            cell(c, xc.x, dc.x, rc.x, qc.x, xo.x, ro.x);
            cell(c + 1, xc.y, dc.y, rc.y, qc.y, xo.y, ro.y);
            *(T2 *)(x + c) = xo;
            *(T2 *)(r + c) = ro;
        }
    } else {
        TileSched ts((g.n + kBlock - 1) / kBlock);
        for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
            const uint32_t c = t * kBlock + threadIdx.x;
            if (c >= g.n) continue;
            TX xo;
            T ro;
            cell(c, x[c], d[c], r[c], q[c], xo, ro);
            x[c] = xo;
            r[c] = ro;
        }
    }
    publish_partials<3, 1>(s, m, pout);
}

// x += e (FP64 += FP32): the refinement step of the mixed path.  A
// correction solved on the Neumann operator is defined up to a constant:
// with `shift` it is taken as e - e_0, the one that keeps the pin x_0 = 0.
// x += e - shift e_0 (overwrite: x = e - shift e_0, x is zero and not
// read).  The flexible-CG solution arrives in two parts: e (valid after two
// updates or more) and a deferred last step ta * pend (pend null: none).
__global__ void __launch_bounds__(kBlock) k_add_correction(uint32_t n, double *__restrict__ x,
                                                           const float *__restrict__ e, int shift, int e_valid,
                                                           const float *__restrict__ pend, float ta,
                                                           int overwrite) {
    const double e0 = shift ? (double)e[0] : 0.0;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) {
        double d = e_valid ? (double)e[c] - e0 : 0.0;
        if (pend) d += (double)(ta * pend[c]);
        x[c] = overwrite ? d : x[c] + d;
    }
}

// The compatible Neumann form of the pinned residual for the inner Jacobi
// CG: row 0 gets r_0 = -sum_{k!=0} r_k (outer[3], folded from the residual
// kernel's partials), stored in FP32, and its contribution {r.z, r.r, r,
// |r|} is written as one extra partial row after the residual's rows.
__global__ void k_row0_compat(Grid g, const double *__restrict__ outer, float *__restrict__ r,
                              double *__restrict__ row) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const float rf = (float)(-outer[3]);
    r[0] = rf;
    const float z = rf * g.invf[diag_index(g, 0, 0, 0)];
    const double rd = (double)rf;
    row[0] = rd * (double)z;
    row[1] = rd * rd;
    row[2] = rd;
    row[3] = fabs(rd);
}

// ---------------------------------------------------------------------------
// Multigrid-preconditioned flexible CG (FP32 vectors, FP64 dots and scalars).
//
// The preconditioner is one V-cycle of cell-centred geometric multigrid for
// the Neumann operator: 3 damped-Jacobi sweeps (omega = 6/7) before and after
// the coarse correction, restriction = mean of the 8 children, prolongation =
// piecewise constant scaled by 2.0, Galerkin-consistent coarse coefficients
// (each axis coefficient halves per level).  Levels down to <= 4096 cells run
// in one single-workgroup kernel out of LDS.  The pinned system is handed to
// the Neumann cycle in its exact equivalent form: r_0 is replaced by
// -sum_{k != 0} r_k (which makes the right-hand side compatible) and the
// result is shifted so that z_0 = 0.  With z = M(r) the FCG(1) recurrence is
//   p = z - (z.q_old / p_old.q_old) p_old,  q = A p,  alpha = r.p / p.q.
// ---------------------------------------------------------------------------
constexpr float kOmega = 6.0f / 7.0f;
constexpr float kCorr = 2.0f;
constexpr int kMaxLevels = 12;
constexpr uint32_t kTailCells = 4096;
// A level of at most kTailFast cells below the first one with <= kTailCells
// takes the tail instead: the one workgroup's sweeps over 4096 cells cost
// more (28 us per V-cycle at 16^3-4^3) than the fused passes over 8^3 tiles
// plus the smaller tail (10 us + 16 us).
constexpr uint32_t kTailFast = 512;

struct Level {
    uint32_t nx, ny, nz, nxy, n;
    float cx, cy, cz;
    FastDiv fx, fxy;
    float inv[8];
};

struct FcgA {  // written by block 0 of k_mg_check (the stop test)
    double rr0, rr, rmax, sumr;
    int done, converged, status;
    float escale;  // storage scale of the level-0 iterates this iteration
};
struct FcgB {  // written by block 0 of k_fcg_update (and the host)
    double pq;
    long long iters;
    int done, status, first;
    float alpha_even;  // alpha of the last even-numbered update (its x update is deferred)
};

__device__ inline void lvl_ijk(const Level &L, uint32_t c, uint32_t &i, uint32_t &j, uint32_t &k) {
    k = L.fxy.div(c);
    const uint32_t rem = c - k * L.nxy;
    j = L.fx.div(rem);
    i = rem - j * L.nx;
}
__device__ inline uint32_t lvl_didx(const Level &L, uint32_t i, uint32_t j, uint32_t k) {
    return (uint32_t)(i > 0 && i + 1 < L.nx) | ((uint32_t)(j > 0 && j + 1 < L.ny) << 1) |
           ((uint32_t)(k > 0 && k + 1 < L.nz) << 2);
}

// A level's 8 reciprocal diagonals in LDS (per-lane lookups of the
// kernel-argument copy are vector loads from the argument segment).  Every
// thread of the block must call it (it ends with a block barrier).
__device__ inline const float *lvl_inv_table(const Level &L) {
    __shared__ float tab[8];
    if (threadIdx.x < 8) tab[threadIdx.x] = L.inv[threadIdx.x];
    __syncthreads();
    return tab;
}

// f on level 0: the residual with the compatibility value at cell 0.
// The load is unconditional (then a select): a load under a branch would
// make the marches wait for every outstanding access where the paths join.
__device__ inline float f_at(const float *f, uint32_t c, bool level0, float f0) {
    const float v = f[c];
    return (level0 && c == 0) ? f0 : v;
}

// The first two damped-Jacobi sweeps from a zero iterate at cell c of a
// level: e1 = omega D^-1 f is pointwise, so e2 = e1 + omega D^-1 (f - A e1)
// needs only f and D^-1 on the 7-point stencil (a clamped neighbour is the
// cell itself and contributes nothing, as in the sweep).
__device__ inline float two_sweeps_from_zero(const Level &L, const float *inv, const float *f,
                                             uint32_t i, uint32_t j, uint32_t k, bool level0, float f0) {
    const uint32_t ii[7] = {i, i + 1 < L.nx ? i + 1 : i, i > 0 ? i - 1 : i, i, i, i, i};
    const uint32_t jj[7] = {j, j, j, j + 1 < L.ny ? j + 1 : j, j > 0 ? j - 1 : j, j, j};
    const uint32_t kk[7] = {k, k, k, k, k, k + 1 < L.nz ? k + 1 : k, k > 0 ? k - 1 : k};
    float fv[7], e1[7];
#pragma unroll
    for (int a = 0; a < 7; ++a) fv[a] = f_at(f, ii[a] + L.nx * (jj[a] + L.ny * kk[a]), level0, f0);
#pragma unroll
    for (int a = 0; a < 7; ++a) e1[a] = kOmega * inv[lvl_didx(L, ii[a], jj[a], kk[a])] * fv[a];
    const float ae = L.cx * ((e1[0] - e1[1]) + (e1[0] - e1[2])) + L.cy * ((e1[0] - e1[3]) + (e1[0] - e1[4])) +
                     L.cz * ((e1[0] - e1[5]) + (e1[0] - e1[6]));
    return e1[0] + kOmega * inv[lvl_didx(L, i, j, k)] * (fv[0] - ae);
}

// The learned coarse solve's residual on a level: f <- f - A x (in place;
// each cell reads only its own f), and the correction's sum a <- a + b.
__global__ void __launch_bounds__(kBlock) k_lvl_residual(Level L, float *__restrict__ f,
                                                         const float *__restrict__ x, const FcgA *SA,
                                                         const FcgB *SB) {
    if (SA->done || SB->done) return;
    TileSched ts((L.n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t c = t * kBlock + threadIdx.x;
        if (c >= L.n) continue;
        uint32_t i, j, k;
        lvl_ijk(L, c, i, j, k);
        const uint32_t ii[7] = {i, i + 1 < L.nx ? i + 1 : i, i > 0 ? i - 1 : i, i, i, i, i};
        const uint32_t jj[7] = {j, j, j, j + 1 < L.ny ? j + 1 : j, j > 0 ? j - 1 : j, j, j};
        const uint32_t kk[7] = {k, k, k, k, k, k + 1 < L.nz ? k + 1 : k, k > 0 ? k - 1 : k};
        float ev[7];
#pragma unroll
        for (int a = 0; a < 7; ++a) ev[a] = x[ii[a] + L.nx * (jj[a] + L.ny * kk[a])];
        const float ae = L.cx * ((ev[0] - ev[1]) + (ev[0] - ev[2])) +
                         L.cy * ((ev[0] - ev[3]) + (ev[0] - ev[4])) +
                         L.cz * ((ev[0] - ev[5]) + (ev[0] - ev[6]));
        f[c] = f[c] - ae;
    }
}
__global__ void __launch_bounds__(kBlock) k_lvl_add(uint32_t n, float *__restrict__ a,
                                                    const float *__restrict__ b, const FcgA *SA,
                                                    const FcgB *SB) {
    if (SA->done || SB->done) return;
    for (uint32_t c = blockIdx.x * kBlock + threadIdx.x; c < n; c += gridDim.x * kBlock) a[c] += b[c];
}

// One damped-Jacobi sweep on a level:
//   e_out = e' + omega D^-1 (f - A e'),  e' = e_in + corr P(e_c) (kProlong;
//   a null e_in is zero)
// or, kFirst, the first two sweeps from a zero iterate.  kPublish (level 0,
// last sweep) also publishes sum z.q_old and sum q_old for the FCG beta.
template <bool kFirst, bool kProlong, bool kLevel0, bool kPublish>
__global__ void __launch_bounds__(kBlock) k_mg_sweep(Level L, Level LC, const float *__restrict__ f,
                                                     const float *__restrict__ ein,
                                                     const float *__restrict__ ec,
                                                     float *__restrict__ eout,
                                                     const float *__restrict__ qold,
                                                     double *__restrict__ pout, const FcgA *SA,
                                                     const FcgB *SB, float corr) {
    if (SA->done || SB->done) return;
    const float *inv_tab = lvl_inv_table(L);
    const float f0 = kLevel0 ? (float)(-SA->sumr) : 0.0f;
    double s[2] = {0.0, 0.0}, m[1] = {0.0};
    TileSched ts((L.n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t c = t * kBlock + threadIdx.x;
        if (c >= L.n) continue;
        uint32_t i, j, k;
        lvl_ijk(L, c, i, j, k);
        const float fc = f_at(f, c, kLevel0, f0);
        const float inv = inv_tab[lvl_didx(L, i, j, k)];
        float out;
        if (kFirst) {
            out = two_sweeps_from_zero(L, inv_tab, f, i, j, k, kLevel0, f0);
        } else {
            const uint32_t ii[7] = {i, i + 1 < L.nx ? i + 1 : i, i > 0 ? i - 1 : i, i, i, i, i};
            const uint32_t jj[7] = {j, j, j, j + 1 < L.ny ? j + 1 : j, j > 0 ? j - 1 : j, j, j};
            const uint32_t kk[7] = {k, k, k, k, k, k + 1 < L.nz ? k + 1 : k, k > 0 ? k - 1 : k};
            float ev[7];
#pragma unroll
            for (int a = 0; a < 7; ++a) {
                const uint32_t idx = ii[a] + L.nx * (jj[a] + L.ny * kk[a]);
                ev[a] = (!kProlong || ein) ? ein[idx] : 0.0f;
                if (kProlong) ev[a] += corr * ec[(ii[a] >> 1) + LC.nx * ((jj[a] >> 1) + LC.ny * (kk[a] >> 1))];
            }
            const float ae = L.cx * ((ev[0] - ev[1]) + (ev[0] - ev[2])) +
                             L.cy * ((ev[0] - ev[3]) + (ev[0] - ev[4])) +
                             L.cz * ((ev[0] - ev[5]) + (ev[0] - ev[6]));
            out = ev[0] + kOmega * inv * (fc - ae);
        }
        eout[c] = out;
        if (kPublish && !SB->first) {
            const double qo = (double)qold[c];
            s[0] += (double)out * qo;
            s[1] += qo;
        }
    }
    if (kPublish) publish_partials<2, 1>(s, m, pout);
}

// Restriction of the residual, f_c(C) = mean over the 8 children of f - A e,
// one coarse cell per thread (the levels in global memory use the column
// march k_mg_restrict_z; this one serves the learned warm start's levels).
__global__ void __launch_bounds__(kBlock) k_mg_restrict(Level L, Level LC, const float *__restrict__ f,
                                                        const float *__restrict__ e,
                                                        float *__restrict__ fcoarse, const FcgA *SA,
                                                        const FcgB *SB) {
    if (SA->done || SB->done) return;
    for (uint32_t C = blockIdx.x * blockDim.x + threadIdx.x; C < LC.n; C += gridDim.x * blockDim.x) {
        uint32_t I, J, K;
        lvl_ijk(LC, C, I, J, K);
        float acc = 0.0f;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const uint32_t i = 2 * I + (a & 1), j = 2 * J + (a >> 1 & 1), k = 2 * K + (a >> 2);
            const uint32_t c = i + L.nx * (j + L.ny * k);
            const float ec = e[c];
            const float ee = e[i + 1 < L.nx ? c + 1 : c], ew = e[i > 0 ? c - 1 : c];
            const float en = e[j + 1 < L.ny ? c + L.nx : c], es = e[j > 0 ? c - L.nx : c];
            const float ef = e[k + 1 < L.nz ? c + L.nxy : c], eb = e[k > 0 ? c - L.nxy : c];
            const float ae = L.cx * ((ec - ee) + (ec - ew)) + L.cy * ((ec - en) + (ec - es)) +
                             L.cz * ((ec - ef) + (ec - eb));
            acc += f[c] - ae;
        }
        fcoarse[C] = 0.125f * acc;
    }
}

// The levels with <= 4096 cells: a whole V-cycle (the same sweeps and
// transfers) in one workgroup out of LDS.  `lv` holds levels t..nl-1 (at
// most three: a level of <= 4096 cells halves only from 16^3, to 8^3 and
// then 4^3); f is level t's right-hand side and e receives its solution.
// The coarsest level is solved exactly when `pinv` holds its dense
// pseudo-inverse (n x n, symmetric), else by 60 sweeps with the mean removed
// (the Neumann null space).
constexpr int kTailMax = 3;
constexpr int kFcgHints = 8;

// Level-0 multigrid iterates are stored as FP16 scaled by a power of two,
// halving the bytes of the level-0 sweeps' iterate streams: they only steer
// the flexible CG, whose vectors stay FP32 and whose dots and stop tests
// stay FP64.  Arithmetic is FP32; scaling by a power of two is exact.  The
// iterates of a V-cycle on r are bounded by |A^-1 r| <= max|r| L_max^2 / pi^2
// (the operator's smallest non-zero eigenvalue is at least pi^2 / L_max^2)
// and the smoothest ones a sweep makes are ~max|r| h^2 / 12; the scale
// S = 2^(10 - e), with max|r| L_max^2 / pi^2 in [2^(e-1), 2^e), maps the
// largest below 2^10 and leaves about n^2 of range above FP16's smallest
// normal (6e-5).  An overflow (a diverging solve) makes the inner solve
// stagnate, which hands over to the FP64 PCG.
typedef _Float16 e16;
__device__ inline float e_ld(const e16 *p, size_t i, float inv_scale) { return (float)p[i] * inv_scale; }
__device__ inline float e_ld(const float *p, size_t i, float) { return p[i]; }  // coarse levels: FP32
__device__ inline void e_st(e16 *p, size_t i, float v, float scale) { p[i] = (e16)(v * scale); }
__device__ inline float e_scale_for(double rmax, double e_bound) {
    int ex = 0;
    (void)frexp(rmax * e_bound, &ex);
    return ldexpf(1.0f, min(max(10 - ex, -120), 120));
}
constexpr uint32_t kDirectCells = 128;
struct TailLevels {
    Level lv[kTailMax];
    int count;
    const float *pinv;
};
// `itab` is the level's reciprocal-diagonal table in LDS (a per-lane index
// into the kernel-argument copy would put the whole argument in scratch).
__device__ inline void tail_sweep(const Level &L, const float *itab, const float *f, const float *ein,
                                  float *eout, const float *ec, const Level *LC, bool first) {
    for (uint32_t c = threadIdx.x; c < L.n; c += blockDim.x) {
        uint32_t i, j, k;
        lvl_ijk(L, c, i, j, k);
        const float inv = itab[lvl_didx(L, i, j, k)];
        if (first) {
            eout[c] = kOmega * inv * f[c];
            continue;
        }
        const uint32_t ii[7] = {i, i + 1 < L.nx ? i + 1 : i, i > 0 ? i - 1 : i, i, i, i, i};
        const uint32_t jj[7] = {j, j, j, j + 1 < L.ny ? j + 1 : j, j > 0 ? j - 1 : j, j, j};
        const uint32_t kk[7] = {k, k, k, k, k, k + 1 < L.nz ? k + 1 : k, k > 0 ? k - 1 : k};
        float ev[7];
#pragma unroll
        for (int a = 0; a < 7; ++a) {
            ev[a] = ein[ii[a] + L.nx * (jj[a] + L.ny * kk[a])];
            if (ec) ev[a] += kCorr * ec[(ii[a] >> 1) + LC->nx * ((jj[a] >> 1) + LC->ny * (kk[a] >> 1))];
        }
        const float ae = L.cx * ((ev[0] - ev[1]) + (ev[0] - ev[2])) +
                         L.cy * ((ev[0] - ev[3]) + (ev[0] - ev[4])) +
                         L.cz * ((ev[0] - ev[5]) + (ev[0] - ev[6]));
        eout[c] = ev[0] + kOmega * inv * (f[c] - ae);
    }
    __syncthreads();
}
// The V-cycle over C tail levels; every loop over levels is unrolled so the
// per-level LDS pointers stay in registers.
template <int C>
__device__ inline void tail_vcycle(const TailLevels &T, float *pool, float *red, float *itab,
                                   const float *__restrict__ fin, float *__restrict__ eout) {
    float *F[C], *E[C], *W[C], *I[C];
    uint32_t off = 0;
#pragma unroll
    for (int l = 0; l < C; ++l) {
        F[l] = pool + off; off += T.lv[l].n;
        E[l] = pool + off; off += T.lv[l].n;
        W[l] = pool + off; off += T.lv[l].n;
        I[l] = itab + 8 * l;
        if (threadIdx.x < 8) I[l][threadIdx.x] = T.lv[l].inv[threadIdx.x];
    }
    for (uint32_t c = threadIdx.x; c < T.lv[0].n; c += blockDim.x) F[0][c] = fin[c];
    constexpr int last = C - 1;
    // The exact coarsest solve splits each cell's column of the
    // pseudo-inverse over kPL lanes; the column is fetched into registers
    // here, so its global loads overlap the down pass instead of stalling
    // the solve.
    constexpr uint32_t kPL = 8, kPM = kDirectCells / kPL;
    static_assert(kPL * kDirectCells <= 1024 && kDirectCells % kPL == 0, "one block of 1024 covers the solve");
    const Level &LL = T.lv[last];
    const uint32_t pc = threadIdx.x / kPL, pl = threadIdx.x % kPL;
    float prow[kPM];
#pragma unroll
    for (uint32_t m = 0; m < kPM; ++m) {
        const uint32_t j = pl + kPL * m;
        prow[m] = (T.pinv && pc < LL.n && j < LL.n) ? T.pinv[(size_t)j * LL.n + pc] : 0.0f;
    }
    __syncthreads();
    // down: pre-smooth (result in E), restrict into F[l+1]
#pragma unroll
    for (int l = 0; l < last; ++l) {
        const Level &L = T.lv[l];
        tail_sweep(L, I[l], F[l], nullptr, E[l], nullptr, nullptr, true);
        tail_sweep(L, I[l], F[l], E[l], W[l], nullptr, nullptr, false);
        tail_sweep(L, I[l], F[l], W[l], E[l], nullptr, nullptr, false);
        const Level &LC = T.lv[l + 1];
        const float *e = E[l];
        for (uint32_t cc = threadIdx.x; cc < LC.n; cc += blockDim.x) {
            uint32_t I, J, K;
            lvl_ijk(LC, cc, I, J, K);
            float acc = 0.0f;
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const uint32_t i = 2 * I + (a & 1), j = 2 * J + (a >> 1 & 1), k = 2 * K + (a >> 2);
                const uint32_t c = i + L.nx * (j + L.ny * k);
                const float ec = e[c];
                const float ae = L.cx * ((ec - e[i + 1 < L.nx ? c + 1 : c]) + (ec - e[i > 0 ? c - 1 : c])) +
                                 L.cy * ((ec - e[j + 1 < L.ny ? c + L.nx : c]) + (ec - e[j > 0 ? c - L.nx : c])) +
                                 L.cz * ((ec - e[k + 1 < L.nz ? c + L.nxy : c]) + (ec - e[k > 0 ? c - L.nxy : c]));
                acc += F[l][c] - ae;
            }
            F[l + 1][cc] = 0.125f * acc;
        }
        __syncthreads();
    }
    // coarsest level, solution in E[last]
    if (T.pinv) {
        float acc = 0.0f;
#pragma unroll
        for (uint32_t m = 0; m < kPM; ++m) {
            const uint32_t j = pl + kPL * m;
            if (pc < LL.n && j < LL.n) acc += prow[m] * F[last][j];
        }
        acc += __shfl_xor(acc, 1);
        acc += __shfl_xor(acc, 2);
        acc += __shfl_xor(acc, 4);
        if (pc < LL.n && pl == 0) E[last][pc] = acc;
        __syncthreads();
    } else {
        tail_sweep(LL, I[last], F[last], nullptr, E[last], nullptr, nullptr, true);
        for (int it = 0; it < 59; it += 2) {
            tail_sweep(LL, I[last], F[last], E[last], W[last], nullptr, nullptr, false);
            tail_sweep(LL, I[last], F[last], W[last], E[last], nullptr, nullptr, false);
        }
        float acc = 0.0f;
        for (uint32_t c = threadIdx.x; c < LL.n; c += blockDim.x) acc += E[last][c];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        const float mean = red[0] / (float)LL.n;
        for (uint32_t c = threadIdx.x; c < LL.n; c += blockDim.x) E[last][c] -= mean;
        __syncthreads();
    }
    // up: prolongate-correct fused into the first post-sweep; a level's
    // result ends in W
    const float *sol = E[last];
#pragma unroll
    for (int l = last - 1; l >= 0; --l) {
        const Level &L = T.lv[l];
        tail_sweep(L, I[l], F[l], E[l], W[l], sol, &T.lv[l + 1], false);
        tail_sweep(L, I[l], F[l], W[l], E[l], nullptr, nullptr, false);
        tail_sweep(L, I[l], F[l], E[l], W[l], nullptr, nullptr, false);
        sol = W[l];
    }
    for (uint32_t c = threadIdx.x; c < T.lv[0].n; c += blockDim.x) eout[c] = sol[c];
}
__global__ void __launch_bounds__(1024) k_mg_tail(TailLevels T, const float *__restrict__ fin,
                                                  float *__restrict__ eout, const FcgA *SA,
                                                  const FcgB *SB) {
    if (SA->done || SB->done) return;
    // LDS: per level f, e, w (ping-pong); levels shrink by 8 so 3 * 1.15 * 4096 floats suffice
    __shared__ float pool[3 * 4800];
    __shared__ float red[1024];
    __shared__ float itab[8 * kTailMax];
    if (T.count == 1) tail_vcycle<1>(T, pool, red, itab, fin, eout);
    else if (T.count == 2) tail_vcycle<2>(T, pool, red, itab, fin, eout);
    else tail_vcycle<3>(T, pool, red, itab, fin, eout);
}

// ------------------------------------------------ fused coarse-level passes
// A level's down pass (two sweeps from zero, a third sweep, the residual
// restriction) and its up pass (prolongate-correct and three sweeps) as one
// launch each.  A block owns a kFT^3 tile of fine cells and recomputes, in
// LDS, each intermediate iterate over the halo the later stages read: the
// down pass stages f over a halo of 3, the second iterate over 2, the third
// over 1; the up pass likewise.  Every array shares one (kFT + 6)^3 frame, so
// a neighbour is +-1, +-kFH, +-kFH^2 away.  Neighbours outside the domain
// are the cell itself (the Neumann clamp), so they drop out of A e.
constexpr int kFT = 8, kFH = kFT + 6, kFF = kFH * kFH * kFH;
constexpr uint32_t kFB = 1024;

struct FTile {
    int x0, y0, z0;  // global coordinates of frame cell (0, 0, 0)
};
__device__ inline FTile ftile(const Level &L) {
    const uint32_t tx = (L.nx + kFT - 1) / kFT, ty = (L.ny + kFT - 1) / kFT;
    // XCD-contiguous tile order (the dispatcher deals blocks round-robin)
    const uint32_t nb = gridDim.x, q = nb / kXcds, r = nb % kXcds, x = blockIdx.x % kXcds,
                   slot = blockIdx.x / kXcds;
    const uint32_t t = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + slot;
    return FTile{(int)(t % tx) * kFT - 3, (int)((t / tx) % ty) * kFT - 3, (int)(t / (tx * ty)) * kFT - 3};
}
// Visits the frame cells of the tile grown by halo h that lie in the domain:
// fn(frame index, x, y, z).
template <int H, typename Fn>
__device__ inline void fvisit(const Level &L, const FTile &T, Fn &&fn) {
    constexpr int E = kFT + 2 * H, N = E * E * E, o = 3 - H;
    for (int t = threadIdx.x; t < N; t += kFB) {
        const int a = t % E + o, b = (t / E) % E + o, c = t / (E * E) + o;
        const int x = T.x0 + a, y = T.y0 + b, z = T.z0 + c;
        if (x < 0 || y < 0 || z < 0 || x >= (int)L.nx || y >= (int)L.ny || z >= (int)L.nz) continue;
        fn(a + kFH * (b + kFH * c), x, y, z);
    }
}
// Staging form of fvisit over the full halo of 3: ld(global index, x, y, z) runs for
// every cell of this thread first (coordinates clamped into the domain, so
// all of them are issued together), then st(frame index, x, y, z, value)
// for the cells that lie in the domain.
template <typename Ld, typename St>
__device__ inline void fstage(const Level &L, const FTile &T, Ld &&ld, St &&st) {
    constexpr int E = kFH, N = kFF, R = (N + kFB - 1) / kFB;
    decltype(ld(0u, 0, 0, 0)) v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int t = min((int)threadIdx.x + q * (int)kFB, N - 1);
        const int x = min(max(T.x0 + t % E, 0), (int)L.nx - 1);
        const int y = min(max(T.y0 + (t / E) % E, 0), (int)L.ny - 1);
        const int z = min(max(T.z0 + t / (E * E), 0), (int)L.nz - 1);
        v[q] = ld((uint32_t)x + L.nx * ((uint32_t)y + L.ny * (uint32_t)z), x, y, z);
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int t = (int)threadIdx.x + q * (int)kFB;
        if (t >= N) break;
        const int x = T.x0 + t % E, y = T.y0 + (t / E) % E, z = T.z0 + t / (E * E);
        if (x < 0 || y < 0 || z < 0 || x >= (int)L.nx || y >= (int)L.ny || z >= (int)L.nz) continue;
        st(t, x, y, z, v[q]);
    }
}
struct FNbr {
    bool e, w, n, s, f, b;
};
__device__ inline FNbr fnbr(const Level &L, int x, int y, int z) {
    return FNbr{x + 1 < (int)L.nx, x > 0, y + 1 < (int)L.ny, y > 0, z + 1 < (int)L.nz, z > 0};
}
__device__ inline uint32_t fdidx(const FNbr &m) {
    return (uint32_t)(m.e && m.w) | ((uint32_t)(m.n && m.s) << 1) | ((uint32_t)(m.f && m.b) << 2);
}
// (A e)(idx) over the frame array e, clamped neighbours dropped
__device__ inline float flap(const Level &L, const float *e, int idx, const FNbr &m) {
    const float c = e[idx];
    return L.cx * ((m.e ? c - e[idx + 1] : 0.0f) + (m.w ? c - e[idx - 1] : 0.0f)) +
           L.cy * ((m.n ? c - e[idx + kFH] : 0.0f) + (m.s ? c - e[idx - kFH] : 0.0f)) +
           L.cz * ((m.f ? c - e[idx + kFH * kFH] : 0.0f) + (m.b ? c - e[idx - kFH * kFH] : 0.0f));
}
__device__ inline uint32_t fglobal(const Level &L, int x, int y, int z) {
    return (uint32_t)x + L.nx * ((uint32_t)y + L.ny * (uint32_t)z);
}

// Down pass of level L: e3 (three damped-Jacobi sweeps from zero on f) into
// e_out, the restricted residual mean_8(f - A e3) into the next level's f.
__global__ void __launch_bounds__(kFB) k_mg_down(Level L, Level LC, const float *__restrict__ f,
                                                 float *__restrict__ eout, float *__restrict__ fcoarse,
                                                 const FcgA *SA, const FcgB *SB) {
    if (SA->done || SB->done) return;
    __shared__ float sf[kFF], s2[kFF], s3[kFF];
    const float *inv = lvl_inv_table(L);
    const FTile T = ftile(L);
    // f, and e1 = omega D^-1 f (pointwise), over the halo of 3
    fstage(L, T, [&](uint32_t g, int, int, int) { return f[g]; },
           [&](int id, int x, int y, int z, float v) {
               sf[id] = v;
               s3[id] = kOmega * inv[fdidx(fnbr(L, x, y, z))] * v;
           });
    __syncthreads();
    // e2 = e1 + omega D^-1 (f - A e1) over 2, e3 = e2 + omega D^-1 (f - A e2) over 1
    fvisit<2>(L, T, [&](int id, int x, int y, int z) {
        const FNbr m = fnbr(L, x, y, z);
        s2[id] = s3[id] + kOmega * inv[fdidx(m)] * (sf[id] - flap(L, s3, id, m));
    });
    __syncthreads();
    fvisit<1>(L, T, [&](int id, int x, int y, int z) {
        const FNbr m = fnbr(L, x, y, z);
        s3[id] = s2[id] + kOmega * inv[fdidx(m)] * (sf[id] - flap(L, s2, id, m));
    });
    __syncthreads();
    // residual of e3 on the tile (into s2), e3 out
    fvisit<0>(L, T, [&](int id, int x, int y, int z) {
        const FNbr m = fnbr(L, x, y, z);
        eout[fglobal(L, x, y, z)] = s3[id];
        s2[id] = sf[id] - flap(L, s3, id, m);
    });
    __syncthreads();
    if (threadIdx.x < (kFT / 2) * (kFT / 2) * (kFT / 2)) {
        constexpr int h = kFT / 2;
        const int a = 3 + 2 * (threadIdx.x % h), b = 3 + 2 * ((threadIdx.x / h) % h), c = 3 + 2 * (threadIdx.x / (h * h));
        const int x = T.x0 + a, y = T.y0 + b, z = T.z0 + c;
        if (x < (int)L.nx && y < (int)L.ny && z < (int)L.nz) {
            const int id = a + kFH * (b + kFH * c);
            float acc = 0.0f;
#pragma unroll
            for (int q = 0; q < 8; ++q) acc += s2[id + (q & 1) + kFH * ((q >> 1) & 1) + kFH * kFH * (q >> 2)];
            fcoarse[(uint32_t)(x >> 1) + LC.nx * ((uint32_t)(y >> 1) + LC.ny * (uint32_t)(z >> 1))] = 0.125f * acc;
        }
    }
}

// Up pass of level L: e' = e3 + corr P(e_c), then three sweeps; the result
// into e_out.
__global__ void __launch_bounds__(kFB) k_mg_up(Level L, Level LC, const float *__restrict__ f,
                                               const float *__restrict__ e3, const float *__restrict__ ec,
                                               float *__restrict__ eout, const FcgA *SA, const FcgB *SB,
                                               float corr) {
    if (SA->done || SB->done) return;
    __shared__ float sf[kFF], sa[kFF], sb[kFF];
    const float *inv = lvl_inv_table(L);
    const FTile T = ftile(L);
    struct Raw {
        float e, c, f;
    };
    fstage(L, T,
           [&](uint32_t g, int x, int y, int z) {
               return Raw{e3[g], ec[(uint32_t)(x >> 1) + LC.nx * ((uint32_t)(y >> 1) + LC.ny * (uint32_t)(z >> 1))],
                          f[g]};
           },
           [&](int id, int, int, int, const Raw &v) {
               sa[id] = v.e + corr * v.c;
               sf[id] = v.f;
           });
    __syncthreads();
    fvisit<2>(L, T, [&](int id, int x, int y, int z) {
        const FNbr m = fnbr(L, x, y, z);
        sb[id] = sa[id] + kOmega * inv[fdidx(m)] * (sf[id] - flap(L, sa, id, m));
    });
    __syncthreads();
    fvisit<1>(L, T, [&](int id, int x, int y, int z) {
        const FNbr m = fnbr(L, x, y, z);
        sa[id] = sb[id] + kOmega * inv[fdidx(m)] * (sf[id] - flap(L, sb, id, m));
    });
    __syncthreads();
    fvisit<0>(L, T, [&](int id, int x, int y, int z) {
        const FNbr m = fnbr(L, x, y, z);
        eout[fglobal(L, x, y, z)] = sa[id] + kOmega * inv[fdidx(m)] * (sf[id] - flap(L, sa, id, m));
    });
}
inline uint32_t ftiles(const Level &L) {
    return ((L.nx + kFT - 1) / kFT) * ((L.ny + kFT - 1) / kFT) * ((L.nz + kFT - 1) / kFT);
}

// ------------------------------------------------ column-marching kernels
// The level-0 stencil passes of the solvers as column marches: a block of
// kZX x kZY threads owns as many (x, y) columns over nz / gridDim.z planes
// (rounded up); a thread walks its column up in z keeping the operand at
// k-1, k and k+1 in registers, so each cell loads only the next plane's
// centre and its four in-plane neighbours (the per-cell form loads all
// seven, the two z neighbours a plane apart).  Each block publishes one
// partial row (blocks in x, y, z order); consumers fold ZGrid::count rows.
// Marches of kZK planes for the one-operand sweeps (twice the blocks of
// kZKD: more waves in flight), kZKD for the two-operand search directions
// (fewer chunk-edge planes read twice, fewer partial rows to fold).
constexpr uint32_t kZX = 128, kZY = 2, kZK = 32, kZKD = 64, kZT = kZX * kZY;

struct ZGrid {
    uint32_t bx, by, bz, count;
};

// The dispatcher deals workgroups to the XCDs round-robin by linear id, and
// each XCD has its own L2.  This maps the hardware id to a logical block so
// that each XCD gets one contiguous run of logical blocks: the y-neighbours
// that read each other's edge rows then share an L2 instead of fetching
// those rows from HBM twice.
__device__ inline uint3 zblock() {
    const uint32_t nb = gridDim.x * gridDim.y * gridDim.z;
    const uint32_t hw = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const uint32_t q = nb / kXcds, r = nb % kXcds, x = hw % kXcds, slot = hw / kXcds;
    const uint32_t lid = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + slot;
    return make_uint3(lid % gridDim.x, (lid / gridDim.x) % gridDim.y, lid / (gridDim.x * gridDim.y));
}
inline ZGrid zgrid(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t kz) {
    ZGrid z{(nx + kZX - 1) / kZX, (ny + kZY - 1) / kZY, (nz + kz - 1) / kz, 0};
    z.count = z.bx * z.by * z.bz;
    return z;
}

// Walks this thread's column.  load(c, i, j, k) fetches the raw values of a cell,
// make(raw, c, i, j, k) turns them into the stencil operand; side(c) fetches
// centre-only values of the current cell (no neighbours); every load of a
// plane step is issued before any operand is formed.  body(i, j, k, c, v, sd)
// gets v = {centre, +x, -x, +y, -y, +z, -z} (neighbours clamped at the walls
// to the cell itself) and side's value.  Threads past the grid return.
template <typename Ld, typename Mk, typename Sd, typename Body>
__device__ inline void zmarch_side(uint32_t nx, uint32_t ny, uint32_t nz, Ld &&load, Mk &&make, Sd &&side,
                                   Body &&body) {
    const uint3 b = zblock();
    const uint32_t i = b.x * kZX + threadIdx.x % kZX, j = b.y * kZY + threadIdx.x / kZX;
    if (i >= nx || j >= ny) return;
    const uint32_t kz = (nz + gridDim.z - 1) / gridDim.z, k0 = b.z * kz;
    if (k0 >= nz) return;
    const uint32_t k1 = min(k0 + kz, nz), nxy = nx * ny;
    const uint32_t ie = i + 1 < nx ? i + 1 : i, iw = i > 0 ? i - 1 : i;
    const uint32_t jn = j + 1 < ny ? j + 1 : j, js = j > 0 ? j - 1 : j;
    const uint32_t de = ie - i, dw = i - iw, dn = (jn - j) * nx, ds = (j - js) * nx;
    uint32_t c = i + nx * (j + ny * k0);
    const uint32_t cb = k0 > 0 ? c - nxy : c;
    auto em = make(load(cb, i, j, k0 > 0 ? k0 - 1 : k0), cb, i, j, k0 > 0 ? k0 - 1 : k0);
    auto ec = make(load(c, i, j, k0), c, i, j, k0);
    for (uint32_t k = k0; k < k1; ++k, c += nxy) {
        const uint32_t kp = k + 1 < nz ? k + 1 : k, cp = c + (kp - k) * nxy;
        const auto rp = load(cp, i, j, kp), re = load(c + de, ie, j, k), rw = load(c - dw, iw, j, k),
                   rn = load(c + dn, i, jn, k), rs = load(c - ds, i, js, k);
        const auto sd = side(c);
        const auto ep = make(rp, cp, i, j, kp);
        const decltype(ec) v[7] = {ec, make(re, c + de, ie, j, k), make(rw, c - dw, iw, j, k),
                                   make(rn, c + dn, i, jn, k), make(rs, c - ds, i, js, k), ep, em};
        body(i, j, k, c, v, sd);
        em = ec;
        ec = ep;
    }
}

template <typename Ld, typename Mk, typename Body>
__device__ inline void zmarch(uint32_t nx, uint32_t ny, uint32_t nz, Ld &&load, Mk &&make, Body &&body) {
    zmarch_side(
        nx, ny, nz, load, make, [](uint32_t) { return 0; },
        [&](uint32_t i, uint32_t j, uint32_t k, uint32_t c, const auto &v, int) { body(i, j, k, c, v); });
}

// Residual restriction, f_c = mean over the 8 children of f - A e, as a
// column march: a thread owns the column pair (i, j), (i, j + 1) with j even
// and walks an even number of planes, so the children's y pair sums in
// registers, the x pair across neighbouring lanes and the z pair across two
// plane steps.  Reads f and e once, writes f_c.  E = e16: level 0 (scaled
// FP16 iterate, f pinned at cell 0); E = float: a coarser level.
constexpr uint32_t kRX = 128;
template <typename E>
__global__ void __launch_bounds__(kRX) k_mg_restrict_z(Level L, Level LC, const float *__restrict__ f,
                                                       const E *__restrict__ e,
                                                       float *__restrict__ fcoarse, const FcgA *SA,
                                                       const FcgB *SB) {
    if (SA->done || SB->done) return;
    constexpr bool kL0 = sizeof(E) == 2;
    const float iS = kL0 ? 1.0f / SA->escale : 1.0f;
    const float f0 = kL0 ? (float)(-SA->sumr) : 0.0f;
    const uint3 b = zblock();
    const uint32_t nx = L.nx, ny = L.ny, nz = L.nz, nxy = L.nxy;
    const uint32_t kz = (((nz + gridDim.z - 1) / gridDim.z) + 1) & ~1u, k0 = b.z * kz;
    if (k0 >= nz) return;  // uniform over the block
    const uint32_t k1 = min(k0 + kz, nz);
    const uint32_t iraw = b.x * kRX + threadIdx.x, j = 2 * b.y;
    const bool live = iraw < nx;  // nx is even, so a live lane's x partner is live
    const uint32_t i = live ? iraw : nx - 1;
    const uint32_t de = i + 1 < nx ? 1 : 0, dw = i > 0 ? 1 : 0;
    const uint32_t ds = j > 0 ? nx : 0, dn = j + 2 < ny ? nx : 0;  // row j+1 always exists
    uint32_t c = i + nx * (j + ny * k0);  // row j; row j+1 is c + nx
    const uint32_t cb = k0 > 0 ? c - nxy : c;
    float am = e_ld(e, cb, iS), bm = e_ld(e, cb + nx, iS), ac = e_ld(e, c, iS), bc = e_ld(e, c + nx, iS);
    const float cx = L.cx, cy = L.cy, cz = L.cz;
    float acc = 0.0f;
    for (uint32_t k = k0; k < k1; ++k, c += nxy) {
        const uint32_t cp = k + 1 < nz ? c + nxy : c;
        const float ap = e_ld(e, cp, iS), bp = e_ld(e, cp + nx, iS);
        const float ae = e_ld(e, c + de, iS), aw = e_ld(e, c - dw, iS), be = e_ld(e, c + nx + de, iS), bw = e_ld(e, c + nx - dw, iS);
        const float as = e_ld(e, c - ds, iS), bn = e_ld(e, c + nx + dn, iS);
        const float fr = f[c], fb = f[c + nx];
        const float fa = (kL0 && c == 0) ? f0 : fr;
        const float ra = fa - (cx * ((ac - ae) + (ac - aw)) + cy * ((ac - bc) + (ac - as)) +
                               cz * ((ac - ap) + (ac - am)));
        const float rb = fb - (cx * ((bc - be) + (bc - bw)) + cy * ((bc - bn) + (bc - ac)) +
                               cz * ((bc - bp) + (bc - bm)));
        float sum = ra + rb;
        sum += __shfl_xor(sum, 1);
        if (((k - k0) & 1) == 0) {
            acc = sum;
        } else if (live && (i & 1) == 0) {
            fcoarse[(i >> 1) + LC.nx * ((j >> 1) + LC.ny * (k >> 1))] = 0.125f * (acc + sum);
        }
        am = ac; ac = ap; bm = bc; bc = bp;
    }
}

// Block reduction of NS sums and NM maxima into this block's partial row.
template <int NS, int NM>
__device__ inline void publish_row(double (&s)[NS], double (&m)[NM], double *pout) {
    block_reduce<NS, NM>(s, m);
    if (threadIdx.x == 0) {
        const size_t row = blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z);
#pragma unroll
        for (int a = 0; a < NS; ++a) pout[row * (NS + NM) + a] = s[a];
#pragma unroll
        for (int a = 0; a < NM; ++a) pout[row * (NS + NM) + NS + a] = m[a];
    }
}

// r = b - A x with b = -rhs (b_0 = 0, row 0 the identity of the pin) as a
// column march over x; each block publishes one row {r.z, r.r, sum r |
// max|r|} (z = D^-1 r), the "update" partials the first iteration of an
// inner solve folds.  For the mixed path the residual is computed in FP64
// and stored as FP32, with the partials taken from the stored values.
//
// kZeroX: x is known to be zero (a zero initial guess), so no x is read;
// A 0 is formed from zeros in the same order, and r = -rhs bit for bit.
template <typename TR, bool kZeroX>
__global__ void __launch_bounds__(kZT) k_residual_z(Grid g, const double *__restrict__ rhs,
                                                    const double *__restrict__ x, TR *__restrict__ r,
                                                    double *__restrict__ pout) {
    const TR *inv_tab = inv_table<TR>(g);
    double s[3] = {0.0, 0.0, 0.0}, m[1] = {0.0};
    zmarch_side(
        g.nx, g.ny, g.nz, [&](uint32_t c, uint32_t, uint32_t, uint32_t) -> double { return kZeroX ? 0.0 : x[c]; },
        [&](double v, uint32_t, uint32_t, uint32_t, uint32_t) -> double { return v; },
        [&](uint32_t c) -> double { return rhs[c]; },
        [&](uint32_t i, uint32_t j, uint32_t k, uint32_t c, const double (&v)[7], double rh) {
            TR rc = 0;
            double inv = 1.0;
            if (c != 0) {
                const double ax = g.ix2 * ((v[0] - v[1]) + (v[0] - v[2])) + g.iy2 * ((v[0] - v[3]) + (v[0] - v[4])) +
                                  g.iz2 * ((v[0] - v[5]) + (v[0] - v[6]));
                rc = (TR)(-rh - ax);
                inv = (double)inv_tab[diag_index(g, i, j, k)];
            }
            r[c] = rc;
            const double rd = (double)rc;
            s[0] += rd * (rd * inv);
            s[1] += rd * rd;
            s[2] += rd;
            m[0] = fmax(m[0], fabs(rd));
        });
    publish_row<3, 1>(s, m, pout);
}

// The same residual with two cells per lane (x even): the centre, the +-y
// and the +-z operands are 16-byte loads of the pair, so a cell costs half
// the vector-memory instructions (the single-cell march is bound by the
// texture-address unit); only the pair's outer x neighbours are 8-byte
// loads.  Arithmetic, in the same order as k_residual_z, gives the same r.
template <typename TR, bool kZeroX>
__global__ void __launch_bounds__(kZT) k_residual_z2(Grid g, const double *__restrict__ rhs,
                                                     const double *__restrict__ x, TR *__restrict__ r,
                                                     double *__restrict__ pout) {
    typedef typename Pair<TR>::type TR2;
    const TR *inv_tab = inv_table<TR>(g);
    double s[3] = {0.0, 0.0, 0.0}, m[1] = {0.0};
    const uint3 b = zblock();
    const uint32_t nx = g.nx, ny = g.ny, nz = g.nz, nxy = nx * ny;
    const uint32_t i0 = 2 * (b.x * kZX + threadIdx.x % kZX), j = b.y * kZY + threadIdx.x / kZX;
    const uint32_t kz = (nz + gridDim.z - 1) / gridDim.z, k0 = b.z * kz;
    if (i0 < nx && j < ny && k0 < nz) {  // nx is even, so i0 + 1 < nx
        const uint32_t k1 = min(k0 + kz, nz);
        const bool xw = i0 > 0, xe = i0 + 2 < nx, ys = j > 0, yn = j + 1 < ny;
        const double ix2 = g.ix2, iy2 = g.iy2, iz2 = g.iz2;
        const double2 z2 = make_double2(0.0, 0.0);
        auto ld2 = [&](uint32_t a) -> double2 { return kZeroX ? z2 : *(const double2 *)(x + a); };
        auto ld1 = [&](uint32_t a) -> double { return kZeroX ? 0.0 : x[a]; };
        uint32_t c = i0 + nx * (j + ny * k0);
        double2 xc = ld2(c);
        double2 xm = k0 > 0 ? ld2(c - nxy) : xc;
        for (uint32_t k = k0; k < k1; ++k, c += nxy) {
            const double2 xp = k + 1 < nz ? ld2(c + nxy) : xc;
            const double2 xn = yn ? ld2(c + nx) : xc;
            const double2 xs = ys ? ld2(c - nx) : xc;
            const double xe1 = xe ? ld1(c + 2) : xc.y, xw0 = xw ? ld1(c - 1) : xc.x;
            const double2 rh = *(const double2 *)(rhs + c);
            const double ax0 = ix2 * ((xc.x - xc.y) + (xc.x - xw0)) + iy2 * ((xc.x - xn.x) + (xc.x - xs.x)) +
                               iz2 * ((xc.x - xp.x) + (xc.x - xm.x));
            const double ax1 = ix2 * ((xc.y - xe1) + (xc.y - xc.x)) + iy2 * ((xc.y - xn.y) + (xc.y - xs.y)) +
                               iz2 * ((xc.y - xp.y) + (xc.y - xm.y));
            TR r0 = 0, r1 = (TR)(-rh.y - ax1);
            double inv0 = 1.0;
            if (c != 0) {
                r0 = (TR)(-rh.x - ax0);
                inv0 = (double)inv_tab[diag_index(g, i0, j, k)];
            }
            const double inv1 = (double)inv_tab[diag_index(g, i0 + 1, j, k)];
            TR2 rv;
            rv.x = r0;
            rv.y = r1;
            *(TR2 *)(r + c) = rv;
            const double d0 = (double)r0, d1 = (double)r1;
            s[0] += d0 * (d0 * inv0) + d1 * (d1 * inv1);
            s[1] += d0 * d0 + d1 * d1;
            s[2] += d0 + d1;
            m[0] = fmax(m[0], fmax(fabs(d0), fabs(d1)));
            xm = xc;
            xc = xp;
        }
    }
    publish_row<3, 1>(s, m, pout);
}



// Level-0 damped-Jacobi sweep, e_out = e_in + omega D^-1 (f - A e_in)
// (the prolongating one is k_mg_prolong0_z); kPublish also publishes
// sum e_out.q_old and sum q_old for the FCG beta.
template <bool kPublish>
__global__ void __launch_bounds__(kZT) k_mg_sweep0_z(Level L, const float *__restrict__ f,
                                                     const e16 *__restrict__ ein, e16 *__restrict__ eout,
                                                     const float *__restrict__ qold,
                                                     double *__restrict__ pout, const FcgA *SA,
                                                     const FcgB *SB) {
    if (SA->done || SB->done) return;
    const float S = SA->escale, iS = 1.0f / S;
    const float *inv = lvl_inv_table(L);
    const float f0 = (float)(-SA->sumr);
    const bool publish = kPublish && !SB->first;
    double s[2] = {0.0, 0.0}, m[1] = {0.0};
    struct Side {
        float f, q;
    };
    // f (and q_old) of the cell are fetched with the plane's stencil loads
    zmarch_side(
        L.nx, L.ny, L.nz,
        [&](uint32_t c, uint32_t, uint32_t, uint32_t) -> float { return e_ld(ein, c, iS); },
        [&](float e, uint32_t, uint32_t, uint32_t, uint32_t) -> float { return e; },
        [&](uint32_t c) -> Side { return Side{f_at(f, c, true, f0), publish ? qold[c] : 0.0f}; },
        [&](uint32_t i, uint32_t j, uint32_t k, uint32_t c, const float (&v)[7], const Side &sd) {
            const float ae = L.cx * ((v[0] - v[1]) + (v[0] - v[2])) + L.cy * ((v[0] - v[3]) + (v[0] - v[4])) +
                             L.cz * ((v[0] - v[5]) + (v[0] - v[6]));
            const float out = v[0] + kOmega * inv[lvl_didx(L, i, j, k)] * (sd.f - ae);
            e_st(eout, c, out, S);
            if (publish) {
                const double qo = (double)sd.q;
                s[0] += (double)out * qo;
                s[1] += qo;
            }
        });
    if (kPublish) publish_row<2, 1>(s, m, pout);
}

// The level-0 prolongate-correct sweep, e_out = e' + omega D^-1 (f - A e')
// with e' = e_in + corr P(e_c), as an explicit march.  P is constant on the
// 2x2x2 children of a coarse cell, so a plane step needs the coarse values of
// the column's own coarse cell and of the one x and one y neighbour outside
// it; they sit in registers and are refetched every second plane, instead of
// one coarse load per stencil operand.
__global__ void __launch_bounds__(kZT) k_mg_prolong0_z(Level L, Level LC, const float *__restrict__ f,
                                                       const e16 *__restrict__ ein,
                                                       const float *__restrict__ ecoarse,
                                                       e16 *__restrict__ eout, const FcgA *SA,
                                                       const FcgB *SB, float corr) {
    if (SA->done || SB->done) return;
    const float S = SA->escale, iS = 1.0f / S;
    const float *inv = lvl_inv_table(L);
    const float f0 = (float)(-SA->sumr);
    const uint3 b = zblock();
    const uint32_t nx = L.nx, ny = L.ny, nz = L.nz, nxy = L.nxy, ncxy = LC.nxy;
    const uint32_t i = b.x * kZX + threadIdx.x % kZX, j = b.y * kZY + threadIdx.x / kZX;
    if (i >= nx || j >= ny) return;
    const uint32_t kz = (nz + gridDim.z - 1) / gridDim.z, k0 = b.z * kz;
    if (k0 >= nz) return;
    const uint32_t k1 = min(k0 + kz, nz);
    const uint32_t ie = i + 1 < nx ? i + 1 : i, iw = i > 0 ? i - 1 : i;
    const uint32_t jn = j + 1 < ny ? j + 1 : j, js = j > 0 ? j - 1 : j;
    const uint32_t de = ie - i, dw = i - iw, dn = (jn - j) * nx, ds = (j - js) * nx;
    // coarse columns: own, and of the x / y neighbour outside the own cell
    // (the own column again where that neighbour is clamped at a wall)
    const bool xodd = (i & 1) != 0, yodd = (j & 1) != 0;
    const uint32_t cc = (i >> 1) + LC.nx * (j >> 1);
    const uint32_t cx = ((xodd ? ie : iw) >> 1) + LC.nx * (j >> 1);
    const uint32_t cy = (i >> 1) + LC.nx * ((yodd ? jn : js) >> 1);
    uint32_t c = i + nx * (j + ny * k0);
    const uint32_t km = k0 > 0 ? k0 - 1 : k0;
    float Cc = ecoarse[cc + ncxy * (k0 >> 1)];
    float Cx = ecoarse[cx + ncxy * (k0 >> 1)], Cy = ecoarse[cy + ncxy * (k0 >> 1)];
    float em = e_ld(ein, c - (k0 - km) * nxy, iS) + corr * ecoarse[cc + ncxy * (km >> 1)];
    float ev = e_ld(ein, c, iS) + corr * Cc;
    for (uint32_t k = k0; k < k1; ++k, c += nxy) {
        const uint32_t kp = k + 1 < nz ? k + 1 : k, cp = c + (kp - k) * nxy;
        if ((k & 1) == 0 && k != k0) {  // a new coarse plane (Cc came with the last +z operand)
            Cx = ecoarse[cx + ncxy * (k >> 1)];
            Cy = ecoarse[cy + ncxy * (k >> 1)];
        }
        const float Cn = (kp >> 1) == (k >> 1) ? Cc : ecoarse[cc + ncxy * (kp >> 1)];
        const float rp = e_ld(ein, cp, iS), re = e_ld(ein, c + de, iS), rw = e_ld(ein, c - dw, iS),
                    rn = e_ld(ein, c + dn, iS), rs = e_ld(ein, c - ds, iS);
        const float fc = f_at(f, c, true, f0);
        const float ep = rp + corr * Cn;
        const float v1 = re + corr * (xodd ? Cx : Cc), v2 = rw + corr * (xodd ? Cc : Cx);
        const float v3 = rn + corr * (yodd ? Cy : Cc), v4 = rs + corr * (yodd ? Cc : Cy);
        const float ae = L.cx * ((ev - v1) + (ev - v2)) + L.cy * ((ev - v3) + (ev - v4)) +
                         L.cz * ((ev - ep) + (ev - em));
        e_st(eout, c, ev + kOmega * inv[lvl_didx(L, i, j, k)] * (fc - ae), S);
        em = ev;
        ev = ep;
        Cc = Cn;
    }
}

// Damped-Jacobi sweep on a coarse level (FP32 iterates) as a column march,
// k_mg_sweep's arithmetic: kFirst gives the first two sweeps from a zero
// iterate (e1 = omega D^-1 f marched, e2 = e1 + omega D^-1 (f - A e1)),
// otherwise e_out = e' + omega D^-1 (f - A e') with e' = e_in
// (+ corr P(e_c), kProlong; a null e_in is zero).
template <bool kFirst, bool kProlong>
__global__ void __launch_bounds__(kZT) k_mg_sweep_z(Level L, Level LC, const float *__restrict__ f,
                                                    const float *__restrict__ ein,
                                                    const float *__restrict__ ec,
                                                    float *__restrict__ eout, const FcgA *SA,
                                                    const FcgB *SB, float corr) {
    if (SA->done || SB->done) return;
    const float *inv = lvl_inv_table(L);
    struct Raw {
        float e, c;
    };
    zmarch_side(
        L.nx, L.ny, L.nz,
        [&](uint32_t c, uint32_t i, uint32_t j, uint32_t k) -> Raw {
            if (kFirst) return Raw{f[c], 0.0f};
            return Raw{ein ? ein[c] : 0.0f,
                       kProlong ? ec[(i >> 1) + LC.nx * ((j >> 1) + LC.ny * (k >> 1))] : 0.0f};
        },
        [&](const Raw &r, uint32_t, uint32_t i, uint32_t j, uint32_t k) -> float {
            if (kFirst) return kOmega * inv[lvl_didx(L, i, j, k)] * r.e;
            return kProlong ? r.e + corr * r.c : r.e;
        },
        [&](uint32_t c) -> float { return f[c]; },
        [&](uint32_t i, uint32_t j, uint32_t k, uint32_t c, const float (&v)[7], float fc) {
            const float ae = L.cx * ((v[0] - v[1]) + (v[0] - v[2])) + L.cy * ((v[0] - v[3]) + (v[0] - v[4])) +
                             L.cz * ((v[0] - v[5]) + (v[0] - v[6]));
            eout[c] = v[0] + kOmega * inv[lvl_didx(L, i, j, k)] * (fc - ae);
        });
}

// Stop test of the inner FCG: folds the `rows` update partials (r.r, sum r,
// max|r| after the last update), and the lead thread records the verdict.
// Returns done; *sumr gets sum r.
__device__ inline int fcg_stop_test(const CgParams &prm, const double *__restrict__ pin, uint32_t rows,
                                    FcgA *SA, const FcgB *SB, bool lead, double *sumr, double *rmax_out) {
    double fs[3], fm[1];
    fold_partials<3, 1>(pin, rows, fs, fm);
    const double rr = fs[1], rmax = fm[0];
    *sumr = fs[2];
    *rmax_out = rmax;
    const bool first = SB->first != 0;
    const double rr0 = first ? rr : SA->rr0;
    int done = 0, conv = 0, status = ST_OK;
    if (!isfinite(rmax) || !isfinite(rr)) {
        status = ST_RESIDUAL;
        done = 1;
    } else if (prm.stop_rule == FL_STOP_REL_L2 ? sqrt(rr) <= prm.tol * sqrt(rr0) : rmax < prm.tol) {
        conv = done = 1;
    } else if (SB->iters >= prm.max_iter) {
        done = 1;
    }
    if (lead && threadIdx.x == 0) {
        SA->rr0 = rr0;
        SA->rr = rr;
        SA->rmax = rmax;
        SA->sumr = *sumr;
        SA->converged = conv;
        SA->status = status;
        SA->done = done;
    }
    return done;
}

// The state of a new inner flexible-CG solve: all zero, first iteration.
__global__ void k_fcg_reset(FcgA *SA, FcgB *SB) {
    *SA = FcgA{};
    FcgB b{};
    b.first = 1;
    *SB = b;
}

// The stop test alone (one block), launched after a batch of iterations so
// the host learns the verdict without launching a whole further iteration.
__global__ void __launch_bounds__(kBlock) k_fcg_check(CgParams prm, const double *__restrict__ pin,
                                                      uint32_t rows, FcgA *SA, const FcgB *SB) {
    if (SA->done || SB->done) return;
    double sumr, rmax;
    (void)fcg_stop_test(prm, pin, rows, SA, SB, true, &sumr, &rmax);
}

// Stop test of the inner FCG, then the first two pre-sweeps of the cycle on
// level 0 from a zero iterate: e1 = omega D^-1 f is pointwise and marched,
// e2 = e1 + omega D^-1 (f - A e1).
__global__ void __launch_bounds__(kZT) k_mg_check_z(Level L, CgParams prm, const float *__restrict__ r,
                                                    e16 *__restrict__ eout,
                                                    const double *__restrict__ pin, uint32_t rows,
                                                    FcgA *SA, const FcgB *SB) {
    if (SA->done || SB->done) return;
    double sumr, rmax;
    const bool lead = blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0;
    const int done = fcg_stop_test(prm, pin, rows, SA, SB, lead, &sumr, &rmax);
    if (done) return;
    const float S = e_scale_for(rmax, prm.e_bound);
    if (lead && threadIdx.x == 0) SA->escale = S;
    const float *inv = lvl_inv_table(L);
    const float f0 = (float)(-sumr);
    // The march written out: the x and y wall bits of the column and of its
    // four side neighbours are fixed, so each one's omega D^-1 for a z-interior
    // and a z-wall plane sits in registers; the raw f of the +z plane is kept
    // as the next plane's centre.
    const uint3 b = zblock();
    const uint32_t nx = L.nx, ny = L.ny, nz = L.nz, nxy = L.nxy;
    const uint32_t i = b.x * kZX + threadIdx.x % kZX, j = b.y * kZY + threadIdx.x / kZX;
    if (i >= nx || j >= ny) return;
    const uint32_t kz = (nz + gridDim.z - 1) / gridDim.z, k0 = b.z * kz;
    if (k0 >= nz) return;
    const uint32_t k1 = min(k0 + kz, nz);
    const uint32_t ie = i + 1 < nx ? i + 1 : i, iw = i > 0 ? i - 1 : i;
    const uint32_t jn = j + 1 < ny ? j + 1 : j, js = j > 0 ? j - 1 : j;
    const uint32_t de = ie - i, dw = i - iw, dn = (jn - j) * nx, ds = (j - js) * nx;
    auto xyb = [&](uint32_t ii, uint32_t jj) -> uint32_t {
        return (uint32_t)(ii > 0 && ii + 1 < nx) | ((uint32_t)(jj > 0 && jj + 1 < ny) << 1);
    };
    const uint32_t bc = xyb(i, j), be = xyb(ie, j), bw = xyb(iw, j), bn = xyb(i, jn), bs = xyb(i, js);
    const float wc0 = kOmega * inv[bc], wc1 = kOmega * inv[bc | 4], we0 = kOmega * inv[be],
                we1 = kOmega * inv[be | 4], ww0 = kOmega * inv[bw], ww1 = kOmega * inv[bw | 4],
                wn0 = kOmega * inv[bn], wn1 = kOmega * inv[bn | 4], ws0 = kOmega * inv[bs],
                ws1 = kOmega * inv[bs | 4];
    auto zin = [&](uint32_t k) { return k > 0 && k + 1 < nz; };
    uint32_t c = i + nx * (j + ny * k0);
    const uint32_t km = k0 > 0 ? k0 - 1 : k0;
    float fc = f_at(r, c, true, f0);
    float em = (zin(km) ? wc1 : wc0) * f_at(r, c - (k0 - km) * nxy, true, f0);
    float ec = (zin(k0) ? wc1 : wc0) * fc;
    for (uint32_t k = k0; k < k1; ++k, c += nxy) {
        const uint32_t kp = k + 1 < nz ? k + 1 : k, cp = c + (kp - k) * nxy;
        const float fp = f_at(r, cp, true, f0), fe = f_at(r, c + de, true, f0), fw = f_at(r, c - dw, true, f0),
                    fn = f_at(r, c + dn, true, f0), fs = f_at(r, c - ds, true, f0);
        const bool zk = zin(k);
        const float ep = (zin(kp) ? wc1 : wc0) * fp;
        const float v1 = (zk ? we1 : we0) * fe, v2 = (zk ? ww1 : ww0) * fw, v3 = (zk ? wn1 : wn0) * fn,
                    v4 = (zk ? ws1 : ws0) * fs;
        const float ae = L.cx * ((ec - v1) + (ec - v2)) + L.cy * ((ec - v3) + (ec - v4)) +
                         L.cz * ((ec - ep) + (ec - em));
        e_st(eout, c, ec + (zk ? wc1 : wc0) * (fc - ae), S);
        em = ec;
        ec = ep;
        fc = fp;
    }
}

// ---------------------------------------------------------------------------
// SOR (PressureSolverKind::Sor, the reference's 3D solver,
// incompressible_3d.rs:293-332) in red-black order.  Each half sweep updates
// the cells of one colour (i + j + k) & 1 from the other colour's values:
// candidate = (sum_nb c_nb p_nb - rhs) / (2 (ix2 + iy2 + iz2)) with a missing
// neighbour reading the centre (pressure_neighbor :450-474), p += omega
// (candidate - p); p(0,0,0) stays 0.  After every sweep the max residual
// |Lap p - rhs| over the cells other than (0,0,0) (pressure_residual
// :476-502, k_residual_z with the pinned row) decides the stop.
struct SorState {
    double res;
    long long iters;
    int done, converged, status, pad;
};

__global__ void __launch_bounds__(kBlock) k_sor3(Grid g, const double *__restrict__ rhs,
                                                 double *__restrict__ p, double omega, uint32_t color,
                                                 const SorState *st) {
    if (st->done) return;
    const uint32_t nx = g.nx, ny = g.ny, nxy = g.nxy;
    const uint32_t nh = (nx + 1) / 2, total = nh * ny * g.nz;
    const double diagonal = 2.0 * (g.ix2 + g.iy2 + g.iz2);
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const uint32_t row = t / nh, i2 = t - row * nh;
        const uint32_t k = row / ny, j = row - k * ny;
        const uint32_t i = 2 * i2 + ((color + j + k) & 1);
        if (i >= nx) continue;
        const uint32_t c = i + nx * (j + ny * k);
        if (c == 0) {
            p[0] = 0.0;
            continue;
        }
        const double ctr = p[c];
        const double e = i + 1 < nx ? p[c + 1] : ctr, w = i > 0 ? p[c - 1] : ctr;
        const double n = j + 1 < ny ? p[c + nx] : ctr, s = j > 0 ? p[c - nx] : ctr;
        const double f = k + 1 < g.nz ? p[c + nxy] : ctr, b = k > 0 ? p[c - nxy] : ctr;
        const double sum = g.ix2 * (e + w) + g.iy2 * (n + s) + g.iz2 * (f + b);
        const double candidate = (sum - rhs[c]) / diagonal;
        p[c] = ctr + omega * (candidate - ctr);
    }
}

// The sweep's verdict from the folded max residual (outer[2]).
__global__ void k_sor_check(const double *__restrict__ outer, double tol, long long max_iter, SorState *st) {
    if (blockIdx.x != 0 || threadIdx.x != 0 || st->done) return;
    const double res = outer[2];
    st->res = res;
    st->iters += 1;
    if (!isfinite(res)) {
        st->status = ST_RESIDUAL;
        st->done = 1;
    } else if (res < tol) {
        st->converged = 1;
        st->done = 1;
    } else if (st->iters >= max_iter) {
        st->done = 1;
    }
}

// Jacobi-PCG direction: d' = D^-1 r + beta d (the pinned row: r_0 + beta d_0),
// q = A d' (q_0 = d'_0), partial d'.q.  Folds the `rows` update partials.
// With `neu` row 0 is a regular row of the Neumann operator.
template <bool kFirst, typename T>
__device__ inline void cg_direction(const Grid &g, const T *inv, const T *r, const T *dold, T *dnew,
                                    T *q, T tb, double (&s)[1], bool neu) {
    const T cx = (T)g.ix2, cy = (T)g.iy2, cz = (T)g.iz2;
    struct Raw {
        T r, d;
    };
    zmarch(
        g.nx, g.ny, g.nz,
        [&](uint32_t c, uint32_t, uint32_t, uint32_t) -> Raw { return Raw{r[c], kFirst ? (T)0 : dold[c]}; },
        [&](const Raw &v, uint32_t c, uint32_t i, uint32_t j, uint32_t k) -> T {
            const T z = (c == 0 && !neu) ? v.r : v.r * inv[diag_index(g, i, j, k)];
            return kFirst ? z : z + tb * v.d;
        },
        [&](uint32_t, uint32_t, uint32_t, uint32_t c, const T (&v)[7]) {
            const T qc = (c == 0 && !neu) ? v[0]
                                : cx * ((v[0] - v[1]) + (v[0] - v[2])) + cy * ((v[0] - v[3]) + (v[0] - v[4])) +
                                      cz * ((v[0] - v[5]) + (v[0] - v[6]));
            dnew[c] = v[0];
            q[c] = qc;
            s[0] += (double)v[0] * (double)qc;
        });
}
// cg_direction with two cells per lane (x even; launched on ZGrid zp): the
// centre, +-y and +-z operands are 16-byte loads of the pair, only the pair's
// outer x neighbours are single loads.  Same arithmetic, same d' and q.
// Used for FP64 (185 -> 168 us at 256^3); the FP32 direction was slower so.
template <bool kFirst, typename T>
__device__ inline void cg_direction2(const Grid &g, const T *inv, const T *r, const T *dold, T *dnew, T *q,
                                     T tb, double (&s)[1]) {
    typedef typename Pair<T>::type T2;
    const T cx = (T)g.ix2, cy = (T)g.iy2, cz = (T)g.iz2;
    const uint3 b = zblock();
    const uint32_t nx = g.nx, ny = g.ny, nz = g.nz, nxy = nx * ny;
    const uint32_t i0 = 2 * (b.x * kZX + threadIdx.x % kZX), j = b.y * kZY + threadIdx.x / kZX;
    const uint32_t kz = (nz + gridDim.z - 1) / gridDim.z, k0 = b.z * kz;
    if (i0 >= nx || j >= ny || k0 >= nz) return;
    const uint32_t k1 = min(k0 + kz, nz);
    const bool xw = i0 > 0, xe = i0 + 2 < nx, ys = j > 0, yn = j + 1 < ny;
    // d' of one cell from its raw r, d
    auto mk = [&](T rv, T dv, uint32_t c, uint32_t i, uint32_t jj, uint32_t k) -> T {
        const T z = c == 0 ? rv : rv * inv[diag_index(g, i, jj, k)];
        return kFirst ? z : z + tb * dv;
    };
    auto mk2 = [&](uint32_t c, uint32_t jj, uint32_t k) -> T2 {
        const T2 rv = *(const T2 *)(r + c);
        T2 dv;
        dv.x = 0;
        dv.y = 0;
        if (!kFirst) dv = *(const T2 *)(dold + c);
        T2 o;
        o.x = mk(rv.x, dv.x, c, i0, jj, k);
        o.y = mk(rv.y, dv.y, c + 1, i0 + 1, jj, k);
        return o;
    };
    auto mk1 = [&](uint32_t c, uint32_t i, uint32_t k) -> T {
        return mk(r[c], kFirst ? (T)0 : dold[c], c, i, j, k);
    };
    uint32_t c = i0 + nx * (j + ny * k0);
    T2 ec = mk2(c, j, k0);
    T2 em = k0 > 0 ? mk2(c - nxy, j, k0 - 1) : ec;
    for (uint32_t k = k0; k < k1; ++k, c += nxy) {
        const T2 ep = k + 1 < nz ? mk2(c + nxy, j, k + 1) : ec;
        const T2 en = yn ? mk2(c + nx, j + 1, k) : ec;
        const T2 es = ys ? mk2(c - nx, j - 1, k) : ec;
        const T we = xw ? mk1(c - 1, i0 - 1, k) : ec.x;
        const T ee = xe ? mk1(c + 2, i0 + 2, k) : ec.y;
        const T qa = c == 0 ? ec.x
                            : cx * ((ec.x - ec.y) + (ec.x - we)) + cy * ((ec.x - en.x) + (ec.x - es.x)) +
                                  cz * ((ec.x - ep.x) + (ec.x - em.x));
        const T qb = cx * ((ec.y - ee) + (ec.y - ec.x)) + cy * ((ec.y - en.y) + (ec.y - es.y)) +
                     cz * ((ec.y - ep.y) + (ec.y - em.y));
        *(T2 *)(dnew + c) = ec;
        T2 qv;
        qv.x = qa;
        qv.y = qb;
        *(T2 *)(q + c) = qv;
        s[0] += (double)ec.x * (double)qa + (double)ec.y * (double)qb;
        em = ec;
        ec = ep;
    }
}
template <typename T>
__global__ void __launch_bounds__(kZT) k_cg_dirapply_z(Grid g, CgParams prm, const T *__restrict__ r,
                                                       const T *__restrict__ dold, T *__restrict__ dnew,
                                                       T *__restrict__ q, const double *__restrict__ pin,
                                                       uint32_t rows, double *__restrict__ pout,
                                                       CgStateA *SA, const CgStateB *SB) {
    if (SA->done || SB->done) return;
    const T *inv = inv_table<T>(g);
    double f[3], fm[1];
    fold_partials<3, 1>(pin, rows, f, fm);
    const double rz = f[0], rr = f[1], rmax = fm[0];
    const bool first = SB->first != 0;
    const double rr0 = prm.rr0_ref > 0.0 ? prm.rr0_ref : (first ? rr : SA->rr0);
    const long long iters = SB->iters;
    int done = 0, converged = 0, status = ST_OK;
    double beta = 0.0;
    if (!first && !isfinite(rmax)) {
        status = ST_RESIDUAL;
        done = 1;
    } else {
        const bool conv = prm.stop_rule == FL_STOP_REL_L2 ? (sqrt(rr) <= prm.tol * sqrt(rr0))
                                                          : (rmax < prm.tol);
        if (conv) {
            converged = done = 1;
        } else if (!first) {
            if (fabs(SB->rz_prev) < 1.0e-30) done = 1;  // the silent break (:1171-1173)
            else if (iters >= prm.max_iter) done = 1;   // MaxIterations
            else beta = rz / SB->rz_prev;
        }
    }
    const bool lead = blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0;
    if (lead && threadIdx.x == 0) {
        SA->rz = rz;
        SA->beta = beta;
        SA->rr0 = rr0;
        SA->rr = rr;
        SA->rmax = rmax;
        SA->converged = converged;
        SA->status = status;
        SA->done = done;
    }
    if (done) return;
    double s[1] = {0.0}, m[1] = {0.0};
    if (sizeof(T) == 8 && g.nx % 2 == 0) {  // FP64 on even x: launched on ZGrid zp
        if (first) cg_direction2<true, T>(g, inv, r, dold, dnew, q, (T)0, s);
        else cg_direction2<false, T>(g, inv, r, dold, dnew, q, (T)beta, s);
    } else {
        if (first) cg_direction<true, T>(g, inv, r, dold, dnew, q, (T)0, s, prm.neumann != 0);
        else cg_direction<false, T>(g, inv, r, dold, dnew, q, (T)beta, s, prm.neumann != 0);
    }
    publish_row<1, 1>(s, m, pout);
}

// ------------------------------------------------ LDS plane-tile Jacobi-PCG
// One CG iteration as two passes over (r, d), each a z-march of plane tiles:
//   dot:  d' = D^-1 r + beta d, q = A d' (never stored), partial d'.q;
//   step: d' and q again from the same r, d; x += alpha d', r' = r - alpha q
//         into the other residual buffer, d' stored; partials of r'.
// Per cell and iteration that is 16 + 48 bytes instead of the 32 + 48 of a
// direction kernel that stores q and an update that reloads it; the
// arithmetic of d', q, x and r' is the same expression for expression.
//
// A block of 4 waves owns a tile of W = 64 V cells in x (V cells, 16 bytes,
// per lane) by 4 rows in y (one per wave) and marches it over its z-chunk.
// Each plane step loads the next plane's r, d once per cell (16-byte loads),
// forms d' into an LDS plane buffer with a one-cell halo, and takes the
// in-plane neighbours of the current plane from LDS and the z neighbours
// from registers; two LDS buffers alternate, one barrier per plane.  Halo
// cells inside the domain are loaded and formed by the wave or lane that
// sits next to them; at a wall the edge cell writes its own d' into the halo
// slot, the clamped neighbour (incompressible_3d.rs:450-474).  Needs nx a
// multiple of V (a lane's cells all exist or none does) and 16-byte aligned
// vectors.
template <typename T> struct Vec16;
template <> struct Vec16<double> { typedef double2 type; };
template <> struct Vec16<float> { typedef float4 type; };
template <typename T> struct Lanes { T a[16 / sizeof(T)]; };
constexpr uint32_t kTileRows = 4;  // rows of a tile: one per wave
template <typename T> constexpr uint32_t tile_width() { return 64u * (16u / (uint32_t)sizeof(T)); }
template <typename T> inline ZGrid tile_grid(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t kz) {
    ZGrid z{(nx + tile_width<T>() - 1) / tile_width<T>(), (ny + kTileRows - 1) / kTileRows, (nz + kz - 1) / kz, 0};
    z.count = z.bx * z.by * z.bz;
    return z;
}
// LDS of a tile pass: two plane buffers of kTileRows + 2 rows of W + 2 V values
template <typename T> constexpr uint32_t tile_lds_len() {
    return 2u * (kTileRows + 2u) * (tile_width<T>() + 2u * (16u / (uint32_t)sizeof(T)));
}

// kStep false: partial d'.q into s[0].  kStep true: the step, partials
// {r'.z', r'.r', sum r' | max|r'|} (z' = D^-1 r', the pinned row z'_0 = r'_0).
template <typename T, bool kFirst, bool kStep>
__device__ inline void cg_tile_pass(const Grid &g, T *lds, const T *__restrict__ r, const T *__restrict__ dold,
                                    T tb, T ta, bool neu, T *__restrict__ x, T *__restrict__ rnew,
                                    T *__restrict__ dnew, double (&s)[3], double (&m)[1]) {
    constexpr int V = 16 / sizeof(T);
    constexpr int W = 64 * V, TY = kTileRows, OFF = V, S = W + 2 * V, BUF = (TY + 2) * S;
    typedef typename Vec16<T>::type TV;
    typedef Lanes<T> RV;
    const T cx = (T)g.ix2, cy = (T)g.iy2, cz = (T)g.iz2;
    const uint3 b = zblock();
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t nx = g.nx, ny = g.ny, nz = g.nz, nxy = nx * ny;
    const uint32_t tx0 = b.x * W, ty0 = b.y * TY, i0 = tx0 + lane * V, j = ty0 + wv;
    const uint32_t kz = (nz + gridDim.z - 1) / gridDim.z, k0 = b.z * kz;
    if (k0 >= nz) return;  // block-uniform; the caller's reduction follows
    const uint32_t k1 = min(k0 + kz, nz);
    const bool own = i0 < nx && j < ny;
    // halo rows: wave 0 forms the row below the tile, wave 1 the row above
    const uint32_t jh = wv == 0 ? ty0 - 1 : ty0 + TY;
    const bool hrow = wv < 2 && i0 < nx && (wv == 0 ? ty0 > 0 : ty0 + TY < ny);
    // halo columns: thread t < TY the cell west of row t, TY <= t < 2 TY east of row t - TY
    const uint32_t t = threadIdx.x;
    const bool east = t >= TY;
    const uint32_t jx = ty0 + (east ? t - TY : t), ix = east ? tx0 + W : tx0 - 1;
    const bool xh = t < 2 * TY && jx < ny && (east ? tx0 + W < nx : tx0 > 0);
    // Per-lane constants of the march.  Plane-0 offsets of the three loads:
    // a lane with nothing to load reads its own (or cell 0's) values instead,
    // so every lane issues the same loads and the compiler's wait counts on
    // them are static.
    const uint32_t co = own ? i0 + nx * j : 0u, ch = hrow ? i0 + nx * jh : co, cxh = xh ? ix + nx * jx : co;
    // reciprocal diagonals: [0] on an interior plane, [1] on a wall plane
    auto inv_of = [&](uint32_t i, uint32_t jj, int wall) -> T {
        const uint32_t a = (uint32_t)(i > 0 && i + 1 < nx) | ((uint32_t)(jj > 0 && jj + 1 < ny) << 1) |
                           ((uint32_t)(!wall) << 2);
        return inv_diag<T>(g, a);
    };
    T io[2][V], ih[2][V], ixv[2];
#pragma unroll
    for (int zw = 0; zw < 2; ++zw) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            io[zw][v] = inv_of(i0 + v, j, zw);
            ih[zw][v] = inv_of(i0 + v, jh, zw);
        }
        ixv[zw] = inv_of(ix, jx, zw);
    }
    auto wall_plane = [&](uint32_t k) { return (k == 0 || k + 1 == nz) ? 1 : 0; };
    const bool pinlane = !neu && i0 == 0 && j == 0;  // owns cell 0 (plane 0, v = 0)
    T *const own_slot = lds + (1 + wv) * S + OFF + lane * V;
    T *const h_slot = lds + (wv == 0 ? 0 : TY + 1) * S + OFF + lane * V;
    T *const x_slot = lds + (1 + (east ? t - TY : t)) * S + (east ? OFF + W : OFF - 1);
    const bool wl = i0 == 0, el = i0 + V == nx, sl = j == 0, nl = j + 1 == ny;
    // the raw loads of one plane (those of plane k + 1 are in flight while
    // plane k is computed)
    struct Ld {
        TV rr, dd, xx, hr, hd;
        T xr, xd;
    };
    Ld LA{};
    auto load = [&](Ld &L, uint32_t k) {
        const uint32_t pk = nxy * k;
        const uint32_t o = own ? pk : 0u, h = hrow ? pk : o, q = xh ? pk : o;
        L.rr = *(const TV *)(r + co + o);
        if (!kFirst) L.dd = *(const TV *)(dold + co + o);
        if (kStep) L.xx = *(const TV *)(x + co + o);
        L.hr = *(const TV *)(r + ch + h);
        if (!kFirst) L.hd = *(const TV *)(dold + ch + h);
        L.xr = r[cxh + q];
        if (!kFirst) L.xd = dold[cxh + q];
    };
    // d' of the own cells of plane k from the loaded r, d
    auto own_dprime = [&](const Ld &L, uint32_t k) -> RV {
        const int zw = wall_plane(k);
        const bool pin = pinlane && k == 0;
        const T *rv = (const T *)&L.rr, *dv = (const T *)&L.dd;
        RV o;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const T z = (v == 0 && pin) ? rv[v] : rv[v] * io[zw][v];
            o.a[v] = kFirst ? z : z + tb * dv[v];
        }
        return o;
    };
    RV em{}, ec{}, ep{};
    TV rc{}, rp{}, xc{}, xp{};
    // d' of plane k into `out` (own cells) and LDS buffer bf (own cells, halos,
    // and at a wall the edge cell's own value as its clamped neighbour)
    auto form = [&](const Ld &L, uint32_t k, int bf, RV &out, TV &rs, TV &xs) {
        const int zw = wall_plane(k);
        if (own) {
            out = own_dprime(L, k);
            rs = L.rr;
            if (kStep) xs = L.xx;
            T *const ob = own_slot + bf * BUF;
            *(TV *)ob = *(const TV *)&out;
            if (wl) ob[-1] = out.a[0];
            if (el) ob[V] = out.a[V - 1];
            if (sl) *(TV *)(ob - S) = *(const TV *)&out;
            if (nl) *(TV *)(ob + S) = *(const TV *)&out;
        }
        if (hrow) {
            const T *rv = (const T *)&L.hr, *dv = (const T *)&L.hd;
            RV h;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const T z = rv[v] * ih[zw][v];
                h.a[v] = kFirst ? z : z + tb * dv[v];
            }
            *(TV *)(h_slot + bf * BUF) = *(const TV *)&h;
        }
        if (xh) {
            const T z = L.xr * ixv[zw];
            x_slot[bf * BUF] = kFirst ? z : z + tb * L.xd;
        }
    };
    // plane k from LDS buffer bf and the registers (em, ec, ep); the step
    // pass stores x, r', d' of its cells
    uint32_t c = co + nxy * k0;
    auto compute = [&](uint32_t k, int bf) {
        if (own) {
            const T *row = own_slot + bf * BUF;
            const bool pin = pinlane && k == 0;
            const int zw = wall_plane(k);
            T qv[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const T v0 = ec.a[v];
                const T ve = v + 1 < V ? ec.a[v + 1] : row[V];
                const T vw = v > 0 ? ec.a[v - 1] : row[-1];
                const T vn = row[S + v], vs = row[-S + v];
                const T qf = cx * ((v0 - ve) + (v0 - vw)) + cy * ((v0 - vn) + (v0 - vs)) +
                             cz * ((v0 - ep.a[v]) + (v0 - em.a[v]));
                qv[v] = (v == 0 && pin) ? v0 : qf;
            }
            if (!kStep) {
#pragma unroll
                for (int v = 0; v < V; ++v) s[0] += (double)ec.a[v] * (double)qv[v];
            } else {
                T xo[V], ro[V];
                const T *xi = (const T *)&xc, *ri = (const T *)&rc;
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    xo[v] = xi[v] + (T)(ta * ec.a[v]);
                    const T rn = ri[v] - ta * qv[v];
                    ro[v] = rn;
                    const T z = (v == 0 && pin) ? rn : rn * io[zw][v];
                    const double rd = (double)rn;
                    s[0] += rd * (double)z;
                    s[1] += rd * rd;
                    s[2] += rd;
                    m[0] = fmax(m[0], fabs(rd));
                }
                *(TV *)(x + c) = *(const TV *)xo;
                *(TV *)(rnew + c) = *(const TV *)ro;
                *(TV *)(dnew + c) = *(const TV *)&ec;
            }
        }
    };
    auto advance = [&]() {
        em = ec;
        ec = ep;
        rc = rp;
        if (kStep) xc = xp;
        c += nxy;
    };
    if (k0 > 0) {
        load(LA, k0 - 1);
        em = own_dprime(LA, k0 - 1);
    }
    load(LA, k0);
    form(LA, k0, 0, ec, rc, xc);
    if (k0 == 0) em = ec;
    if (k0 + 1 < nz) load(LA, k0 + 1);
    __syncthreads();
    for (uint32_t k = k0; k < k1; ++k) {
        const int bf = (k - k0) & 1;
        if (k + 1 < nz) {
            form(LA, k + 1, bf ^ 1, ep, rp, xp);
            if (k + 2 < nz && k + 1 < k1) load(LA, k + 2);
        } else {
            ep = ec;
        }
        compute(k, bf);
        __syncthreads();
        advance();
    }
}

// The dot pass with the stop test and beta of k_cg_dirapply_z; publishes
// {d'.q | 0} per block.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_dot_t(Grid g, CgParams prm, const T *__restrict__ r,
                                                     const T *__restrict__ dold, const double *__restrict__ pin,
                                                     uint32_t rows, double *__restrict__ pout, CgStateA *SA,
                                                     const CgStateB *SB) {
    if (SA->done || SB->done) return;
    double f[3], fm[1];
    fold_partials<3, 1>(pin, rows, f, fm);
    const double rz = f[0], rr = f[1], rmax = fm[0];
    const bool first = SB->first != 0;
    const double rr0 = prm.rr0_ref > 0.0 ? prm.rr0_ref : (first ? rr : SA->rr0);
    int done = 0, converged = 0, status = ST_OK;
    double beta = 0.0;
    if (!first && !isfinite(rmax)) {
        status = ST_RESIDUAL;
        done = 1;
    } else {
        const bool conv = prm.stop_rule == FL_STOP_REL_L2 ? (sqrt(rr) <= prm.tol * sqrt(rr0))
                                                          : (rmax < prm.tol);
        if (conv) {
            converged = done = 1;
        } else if (!first) {
            if (fabs(SB->rz_prev) < 1.0e-30) done = 1;   // the silent break (:1171-1173)
            else if (SB->iters >= prm.max_iter) done = 1;  // MaxIterations
            else beta = rz / SB->rz_prev;
        }
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0) {
        SA->rz = rz;
        SA->beta = beta;
        SA->rr0 = rr0;
        SA->rr = rr;
        SA->rmax = rmax;
        SA->converged = converged;
        SA->status = status;
        SA->first = first;
        SA->done = done;
    }
    if (done) return;
    __shared__ __align__(16) T lds[tile_lds_len<T>()];
    double s[3] = {0.0, 0.0, 0.0}, m[1] = {0.0};
    const bool neu = prm.neumann != 0;
    if (first) cg_tile_pass<T, true, false>(g, lds, r, dold, (T)0, (T)0, neu, nullptr, nullptr, nullptr, s, m);
    else cg_tile_pass<T, false, false>(g, lds, r, dold, (T)beta, (T)0, neu, nullptr, nullptr, nullptr, s, m);
    double s1[1] = {s[0]};
    publish_row<1, 1>(s1, m, pout);
}

// The step pass with alpha from the folded d'.q (k_cg_update's scalars and
// state); reads r, d of this iteration, writes x, r' and d'.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_step_t(Grid g, T *__restrict__ x, const T *__restrict__ r,
                                                      const T *__restrict__ dold, T *__restrict__ rnew,
                                                      T *__restrict__ dnew, const double *__restrict__ pin,
                                                      uint32_t rows, double *__restrict__ pout,
                                                      const CgStateA *SA, CgStateB *SB, int neumann) {
    if (SA->done || SB->done) return;
    double f[1], fm[1];
    fold_partials<1, 1>(pin, rows, f, fm);
    const double den = f[0];
    int status = ST_OK;
    double alpha = 0.0;
    if (fabs(den) < 1.0e-30 || !isfinite(den)) {
        status = ST_SINGULAR;
    } else {
        alpha = SA->rz / den;
        if (!isfinite(alpha)) status = ST_ALPHA;
    }
    if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0) {
        SB->rz_prev = SA->rz;
        SB->alpha = alpha;
        SB->status = status;
        SB->first = 0;
        if (status != ST_OK) SB->done = 1;
        else SB->iters += 1;
    }
    if (status != ST_OK) return;
    __shared__ __align__(16) T lds[tile_lds_len<T>()];
    double s[3] = {0.0, 0.0, 0.0}, m[1] = {0.0};
    const bool neu = neumann != 0;
    if (SA->first) cg_tile_pass<T, true, true>(g, lds, r, dold, (T)0, (T)alpha, neu, x, rnew, dnew, s, m);
    else cg_tile_pass<T, false, true>(g, lds, r, dold, (T)SA->beta, (T)alpha, neu, x, rnew, dnew, s, m);
    publish_row<3, 1>(s, m, pout);
}

// p = (z~ - z~_0) - beta p_old, q = A p (pinned), partials p.q and r.p;
// folds the `rows` partials of the publishing sweep.  A column march: each
// block publishes one row, ZGrid::count rows in all.
__global__ void __launch_bounds__(kZT) k_fcg_dirapply_z(Grid g, const e16 *__restrict__ zt,
                                                        const float *__restrict__ pold,
                                                        float *__restrict__ pnew, float *__restrict__ q,
                                                        const float *__restrict__ r,
                                                        const double *__restrict__ pz, uint32_t rows,
                                                        double *__restrict__ pout, const FcgA *SA,
                                                        const FcgB *SB) {
    if (SA->done || SB->done) return;
    const float iS = 1.0f / SA->escale;
    const bool first = SB->first != 0;
    double fz[2], fm[1];
    fold_partials<2, 1>(pz, rows, fz, fm);
    const float z0 = e_ld(zt, 0, iS);
    const double beta = first ? 0.0 : (fz[0] - (double)z0 * fz[1]) / SB->pq;
    const float tb = (float)beta;
    const float cx = (float)g.ix2, cy = (float)g.iy2, cz = (float)g.iz2;
    double s[2] = {0.0, 0.0}, m[1] = {0.0};
    struct Raw {
        float z, p;
    };
    zmarch_side(
        g.nx, g.ny, g.nz,
        [&](uint32_t c, uint32_t, uint32_t, uint32_t) -> Raw { return Raw{e_ld(zt, c, iS), first ? 0.0f : pold[c]}; },
        [&](const Raw &v, uint32_t c, uint32_t, uint32_t, uint32_t) -> float {
            return c == 0 ? 0.0f : (v.z - z0) - tb * v.p;  // the pinned cell's p stays 0
        },
        [&](uint32_t c) -> float { return r[c]; },
        [&](uint32_t, uint32_t, uint32_t, uint32_t c, const float (&v)[7], float rc) {
            const float qc = c == 0 ? 0.0f
                                    : cx * ((v[0] - v[1]) + (v[0] - v[2])) + cy * ((v[0] - v[3]) + (v[0] - v[4])) +
                                          cz * ((v[0] - v[5]) + (v[0] - v[6]));
            pnew[c] = v[0];
            q[c] = qc;
            s[0] += (double)v[0] * (double)qc;
            s[1] += (double)rc * (double)v[0];
        });
    publish_row<2, 1>(s, m, pout);
}

// The same direction as a plane-tile pass (FP32, four cells per lane; nx a
// multiple of 4): a block owns a 256 x 4 (x, y) tile over a chunk of planes,
// forms p of the next plane into an LDS plane buffer with a one-cell halo
// (rows above and below by waves 0 and 1, the cells west and east of each
// row by threads 0-7; at a wall the edge cell's own p, the clamped
// neighbour), takes the in-plane neighbours from LDS and the z neighbours
// from registers.  Two LDS buffers alternate, one barrier per plane.  The
// stencil arithmetic is k_fcg_dirapply_z's, in the same order.
__global__ void __launch_bounds__(kBlock) k_fcg_dirapply_t(Grid g, const e16 *__restrict__ zt,
                                                        const float *__restrict__ pold,
                                                        float *__restrict__ pnew, float *__restrict__ q,
                                                        const float *__restrict__ r,
                                                        const double *__restrict__ pz, uint32_t rows,
                                                        double *__restrict__ pout, const FcgA *SA,
                                                        const FcgB *SB) {
    if (SA->done || SB->done) return;
    const float iS = 1.0f / SA->escale;
    const bool first = SB->first != 0;
    double fz[2], fm[1];
    fold_partials<2, 1>(pz, rows, fz, fm);
    const float z0 = e_ld(zt, 0, iS);
    const double beta = first ? 0.0 : (fz[0] - (double)z0 * fz[1]) / SB->pq;
    const float tb = (float)beta;
    const float cx = (float)g.ix2, cy = (float)g.iy2, cz = (float)g.iz2;
    double s[2] = {0.0, 0.0}, m[1] = {0.0};
    __shared__ __align__(16) float lds[tile_lds_len<float>()];
    constexpr int V = 4;
    constexpr int W = 64 * V, TY = kTileRows, OFF = V, S = W + 2 * V, BUF = (TY + 2) * S;
    typedef e16 Z4 __attribute__((ext_vector_type(4)));  // 8-byte loads of 4 iterates
    struct F4 {
        float a[4];
    };
    const uint3 b = zblock();
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t nx = g.nx, ny = g.ny, nz = g.nz, nxy = nx * ny;
    const uint32_t tx0 = b.x * W, ty0 = b.y * TY, i0 = tx0 + lane * V, j = ty0 + wv;
    const uint32_t kz = (nz + gridDim.z - 1) / gridDim.z, k0 = b.z * kz;
    if (k0 < nz) {
        const uint32_t k1 = min(k0 + kz, nz);
        const bool own = i0 < nx && j < ny;
        const uint32_t jh = wv == 0 ? ty0 - 1 : ty0 + TY;
        const bool hrow = wv < 2 && i0 < nx && (wv == 0 ? ty0 > 0 : ty0 + TY < ny);
        const uint32_t t = threadIdx.x;
        const bool east = t >= TY;
        const uint32_t jx = ty0 + (east ? t - TY : t), ix = east ? tx0 + W : tx0 - 1;
        const bool xh = t < 2 * TY && jx < ny && (east ? tx0 + W < nx : tx0 > 0);
        // every lane issues the same loads (a lane with nothing to load reads
        // its own or cell 0's values), so the loads' wait counts are static
        const uint32_t co = own ? i0 + nx * j : 0u, ch = hrow ? i0 + nx * jh : co, cxh = xh ? ix + nx * jx : co;
        const bool pin0 = i0 == 0 && j == 0;  // owns cell 0 (plane 0, v = 0): p_0 = q_0 = 0
        float *const own_slot = lds + (1 + wv) * S + OFF + lane * V;
        float *const h_slot = lds + (wv == 0 ? 0 : TY + 1) * S + OFF + lane * V;
        float *const x_slot = lds + (1 + (east ? t - TY : t)) * S + (east ? OFF + W : OFF - 1);
        const bool wl = i0 == 0, el = i0 + V == nx, sl = j == 0, nl = j + 1 == ny;
        Z4 oz{}, hz{};
        float4 op{}, hp{};
        e16 xz = 0;
        float xp = 0.0f;
        auto load = [&](uint32_t k) {
            const uint32_t pk = nxy * k;
            const uint32_t o = own ? pk : 0u, h = hrow ? pk : o, qq = xh ? pk : o;
            oz = *(const Z4 *)(zt + co + o);
            hz = *(const Z4 *)(zt + ch + h);
            xz = zt[cxh + qq];
            if (!first) {
                op = *(const float4 *)(pold + co + o);
                hp = *(const float4 *)(pold + ch + h);
                xp = pold[cxh + qq];
            }
        };
        auto pval = [&](e16 z, float po) { return ((float)z * iS - z0) - tb * po; };
        F4 pm{}, pc{}, pn{};
        // p of plane k: own cells into `out`, and into LDS buffer bf
        auto form = [&](uint32_t k, int bf, F4 &out) {
            if (own) {
#pragma unroll
                for (int v = 0; v < V; ++v) out.a[v] = pval(oz[v], first ? 0.0f : op[v]);
                if (pin0 && k == 0) out.a[0] = 0.0f;
                float *const ob = own_slot + bf * BUF;
                *(float4 *)ob = *(const float4 *)&out;
                if (wl) ob[-1] = out.a[0];
                if (el) ob[V] = out.a[V - 1];
                if (sl) *(float4 *)(ob - S) = *(const float4 *)&out;
                if (nl) *(float4 *)(ob + S) = *(const float4 *)&out;
            }
            if (hrow) {
                F4 h;
#pragma unroll
                for (int v = 0; v < V; ++v) h.a[v] = pval(hz[v], first ? 0.0f : hp[v]);
                *(float4 *)(h_slot + bf * BUF) = *(const float4 *)&h;
            }
            if (xh) x_slot[bf * BUF] = pval(xz, first ? 0.0f : xp);
        };
        // the plane k0 - 1 below the chunk (or plane k0 itself at the bottom wall)
        if (k0 > 0) {
            load(k0 - 1);
            if (own) {
#pragma unroll
                for (int v = 0; v < V; ++v) pm.a[v] = pval(oz[v], first ? 0.0f : op[v]);
                if (pin0 && k0 == 1) pm.a[0] = 0.0f;
            }
        }
        load(k0);
        form(k0, 0, pc);
        if (k0 == 0) pm = pc;
        if (k0 + 1 < nz) load(k0 + 1);
        __syncthreads();
        uint32_t c = co + nxy * k0;
        for (uint32_t k = k0; k < k1; ++k, c += nxy) {
            const int bf = (k - k0) & 1;
            const float4 rc = *(const float4 *)(r + c);  // (cell (0, 0, k)'s for a lane without cells)
            if (k + 1 < nz) {
                form(k + 1, bf ^ 1, pn);
                if (k + 2 < nz && k + 1 < k1) load(k + 2);
            } else {
                pn = pc;
            }
            if (own) {
                const float *row = own_slot + bf * BUF;
                F4 qo;
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const float v0 = pc.a[v];
                    const float ve = v + 1 < V ? pc.a[v + 1] : row[V];
                    const float vw = v > 0 ? pc.a[v - 1] : row[-1];
                    const float vn = row[S + v], vs = row[-S + v];
                    const float qf = cx * ((v0 - ve) + (v0 - vw)) + cy * ((v0 - vn) + (v0 - vs)) +
                                     cz * ((v0 - pn.a[v]) + (v0 - pm.a[v]));
                    qo.a[v] = (v == 0 && pin0 && k == 0) ? 0.0f : qf;
                    s[0] += (double)v0 * (double)qo.a[v];
                    s[1] += (double)rc[v] * (double)v0;
                }
                *(float4 *)(pnew + c) = *(const float4 *)&pc;
                *(float4 *)(q + c) = *(const float4 *)&qo;
            }
            __syncthreads();
            pm = pc;
            pc = pn;
        }
    }
    publish_row<2, 1>(s, m, pout);
}

// ------------------------------------------------ level-0 plane-tile marches
// The level-0 stencil passes of the multigrid cycle on plane tiles (FP32
// operand, four cells per lane, nx a multiple of 4), k_fcg_dirapply_t's
// layout as a helper: a block of 4 waves owns a 256 x 4 (x, y) tile over a
// chunk of planes; each plane's operand goes into an LDS buffer with a
// one-cell halo (rows below and above by waves 0 and 1, the cells west and
// east of each row by threads 0-7; at a wall the edge cell's own value, the
// clamped neighbour), the in-plane neighbours come from LDS and the z
// neighbours from registers.  Two buffers alternate, one barrier per plane;
// the loads of plane k + 1 are in flight while plane k is computed.  A cell
// then costs one load of each operand array instead of five.
struct Quad {
    float a[4];
};
struct Tile4 {
    uint32_t i0, j;     // own quad: cells (i0 .. i0 + 3, j)
    uint32_t jh;        // halo row (waves 0 and 1)
    uint32_t ix, jx;    // halo cell (threads 0 .. 2 kTileRows - 1)
    bool own, hrow, xh;
    uint32_t co, ch, cxh;  // their plane-0 offsets (a lane without one reads its own, or cell 0)
    uint32_t k0, k1;       // the chunk of planes; k0 >= nz: nothing to do
};
__device__ inline Tile4 tile4_geometry(uint32_t nx, uint32_t ny, uint32_t nz) {
    constexpr uint32_t V = 4, W = 64 * V, TY = kTileRows;
    const uint3 b = zblock();
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6, t = threadIdx.x;
    const uint32_t tx0 = b.x * W, ty0 = b.y * TY;
    Tile4 T;
    T.i0 = tx0 + lane * V;
    T.j = ty0 + wv;
    const uint32_t kz = (nz + gridDim.z - 1) / gridDim.z;
    T.k0 = b.z * kz;
    T.k1 = min(T.k0 + kz, nz);
    T.own = T.i0 < nx && T.j < ny;
    T.jh = wv == 0 ? ty0 - 1 : ty0 + TY;
    T.hrow = wv < 2 && T.i0 < nx && (wv == 0 ? ty0 > 0 : ty0 + TY < ny);
    const bool east = t >= TY;
    T.jx = ty0 + (east ? t - TY : t);
    T.ix = east ? tx0 + W : tx0 - 1;
    T.xh = t < 2 * TY && T.jx < ny && (east ? tx0 + W < nx : tx0 > 0);
    T.co = T.own ? T.i0 + nx * T.j : 0u;
    T.ch = T.hrow ? T.i0 + nx * T.jh : T.co;
    T.cxh = T.xh ? T.ix + nx * T.jx : T.co;
    return T;
}
// omega D^-1 of level 0 for the tile's cells, [0] on a wall plane, [1] on an
// interior one (tab: the level's reciprocal diagonals in LDS)
struct TileW {
    float o[2][4], h[2][4], x[2];
};
__device__ inline TileW tile4_weights(const Level &L, const Tile4 &T, const float *tab) {
    auto wgt = [&](uint32_t i, uint32_t j, int zin) -> float {
        const uint32_t a = (uint32_t)(i > 0 && i + 1 < L.nx) | ((uint32_t)(j > 0 && j + 1 < L.ny) << 1) |
                           ((uint32_t)zin << 2);
        return kOmega * tab[a];
    };
    TileW w;
#pragma unroll
    for (int z = 0; z < 2; ++z) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            w.o[z][v] = wgt(T.i0 + v, T.j, z);
            w.h[z][v] = wgt(T.i0 + v, T.jh, z);
        }
        w.x[z] = wgt(T.ix, T.jx, z);
    }
    return w;
}
// load(Ld &, k, own offset, halo-row offset, halo-cell offset) issues plane
// k's loads; form(const Ld &, k, own Quad &, halo Quad &, halo float &, Sd &)
// makes the operand of plane k (and side values of the own cells, kept for
// the body); body(k, c, below, centre, above, row, side) computes plane k:
// c is the own quad's first cell, `row` its LDS slot (row[-1], row[4] the x
// neighbours outside the quad, row[-S + v], row[S + v] the y neighbours).
template <typename Ld, typename Sd, typename LoadF, typename FormF, typename BodyF>
__device__ inline void tile4_march(const Tile4 &T, uint32_t nx, uint32_t ny, uint32_t nz, float *lds, LoadF &&load,
                                   FormF &&form, BodyF &&body) {
    constexpr int V = 4, W = 64 * V, TY = kTileRows, OFF = V, S = W + 2 * V, BUF = (TY + 2) * S;
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6, t = threadIdx.x, nxy = nx * ny;
    const uint32_t k0 = T.k0, k1 = T.k1;
    const bool east = t >= TY;
    float *const own_slot = lds + (1 + wv) * S + OFF + lane * V;
    float *const h_slot = lds + (wv == 0 ? 0 : TY + 1) * S + OFF + lane * V;
    float *const x_slot = lds + (1 + (east ? t - TY : t)) * S + (east ? OFF + W : OFF - 1);
    const bool wl = T.i0 == 0, el = T.i0 + V == nx, sl = T.j == 0, nl = T.j + 1 == ny;
    Ld L{};
    auto ld = [&](uint32_t k) {
        const uint32_t pk = nxy * k;
        const uint32_t o = T.own ? pk : 0u, h = T.hrow ? pk : o, q = T.xh ? pk : o;
        load(L, k, T.co + o, T.ch + h, T.cxh + q);
    };
    auto put = [&](uint32_t k, int bf, Quad &out, Sd &sd) {
        Quad hv;
        float xv;
        form(L, k, out, hv, xv, sd);
        if (T.own) {
            float *const ob = own_slot + bf * BUF;
            *(float4 *)ob = *(const float4 *)&out;
            if (wl) ob[-1] = out.a[0];
            if (el) ob[V] = out.a[V - 1];
            if (sl) *(float4 *)(ob - S) = *(const float4 *)&out;
            if (nl) *(float4 *)(ob + S) = *(const float4 *)&out;
        }
        if (T.hrow) *(float4 *)(h_slot + bf * BUF) = *(const float4 *)&hv;
        if (T.xh) x_slot[bf * BUF] = xv;
    };
    Quad pm{}, pc{}, pn{};
    Sd sc{}, sn{};
    if (k0 > 0) {  // the plane below the chunk: own values only
        Quad hv;
        float xv;
        Sd sd;
        ld(k0 - 1);
        form(L, k0 - 1, pm, hv, xv, sd);
    }
    ld(k0);
    put(k0, 0, pc, sc);
    if (k0 == 0) pm = pc;
    if (k0 + 1 < nz) ld(k0 + 1);
    __syncthreads();
    uint32_t c = T.co + nxy * k0;
    for (uint32_t k = k0; k < k1; ++k, c += nxy) {
        const int bf = (k - k0) & 1;
        if (k + 1 < nz) {
            put(k + 1, bf ^ 1, pn, sn);
            if (k + 2 < nz && k + 1 < k1) ld(k + 2);
        } else {
            pn = pc;
        }
        if (T.own) body(k, c, pm, pc, pn, (const float *)(own_slot + bf * BUF), sc);
        __syncthreads();
        pm = pc;
        pc = pn;
        sc = sn;
    }
}
constexpr int kTileS = 64 * 4 + 2 * 4;  // LDS row stride of the float tiles

// k_mg_check_z as a plane-tile march: the stop test, then e1 = omega D^-1 f
// formed per plane (f with the compatibility value at cell 0) and
// e2 = e1 + omega D^-1 (f - A e1) from it, k_mg_check_z's arithmetic.
__global__ void __launch_bounds__(kBlock) k_mg_check_t(Level L, CgParams prm, const float *__restrict__ r,
                                                       e16 *__restrict__ eout, const double *__restrict__ pin,
                                                       uint32_t rows, FcgA *SA, const FcgB *SB) {
    if (SA->done || SB->done) return;
    double sumr, rmax;
    const bool lead = blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0;
    const int done = fcg_stop_test(prm, pin, rows, SA, SB, lead, &sumr, &rmax);
    if (done) return;
    const float S = e_scale_for(rmax, prm.e_bound);
    if (lead && threadIdx.x == 0) SA->escale = S;
    const float f0 = (float)(-sumr);
    const float *tab = lvl_inv_table(L);
    const uint32_t nx = L.nx, ny = L.ny, nz = L.nz;
    const Tile4 T = tile4_geometry(nx, ny, nz);
    if (T.k0 >= nz) return;
    const TileW w = tile4_weights(L, T, tab);
    __shared__ __align__(16) float lds[tile_lds_len<float>()];
    const float cx = L.cx, cy = L.cy, cz = L.cz;
    const bool pin0 = T.i0 == 0 && T.j == 0;
    auto zin = [&](uint32_t k) { return (k > 0 && k + 1 < nz) ? 1 : 0; };
    struct Ld {
        float4 fo, fh;
        float fx;
    };
    tile4_march<Ld, Quad>(
        T, nx, ny, nz, lds,
        [&](Ld &Q, uint32_t, uint32_t o, uint32_t h, uint32_t x) {
            Q.fo = *(const float4 *)(r + o);
            Q.fh = *(const float4 *)(r + h);
            Q.fx = r[x];
        },
        [&](const Ld &Q, uint32_t k, Quad &out, Quad &hv, float &xv, Quad &f) {
            const int z = zin(k);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                f.a[v] = (v == 0 && pin0 && k == 0) ? f0 : Q.fo[v];
                out.a[v] = w.o[z][v] * f.a[v];
                hv.a[v] = w.h[z][v] * Q.fh[v];
            }
            xv = w.x[z] * Q.fx;
        },
        [&](uint32_t k, uint32_t c, const Quad &em, const Quad &ec, const Quad &ep, const float *row,
            const Quad &f) {
            const int z = zin(k);
            typedef e16 H4 __attribute__((ext_vector_type(4)));
            H4 o;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float v0 = ec.a[v];
                const float ve = v + 1 < 4 ? ec.a[v + 1] : row[4];
                const float vw = v > 0 ? ec.a[v - 1] : row[-1];
                const float vn = row[kTileS + v], vs = row[-kTileS + v];
                const float ae = cx * ((v0 - ve) + (v0 - vw)) + cy * ((v0 - vn) + (v0 - vs)) +
                                 cz * ((v0 - ep.a[v]) + (v0 - em.a[v]));
                o[v] = (e16)((v0 + w.o[z][v] * (f.a[v] - ae)) * S);
            }
            *(H4 *)(eout + c) = o;
        });
}

// k_mg_sweep0_z as a plane-tile march: e_out = e_in + omega D^-1 (f - A e_in)
// on the FP16 iterates; kPublish also publishes sum e_out.q_old and
// sum q_old per block.
template <bool kPublish>
__global__ void __launch_bounds__(kBlock) k_mg_sweep0_t(Level L, const float *__restrict__ f,
                                                        const e16 *__restrict__ ein, e16 *__restrict__ eout,
                                                        const float *__restrict__ qold, double *__restrict__ pout,
                                                        const FcgA *SA, const FcgB *SB) {
    if (SA->done || SB->done) return;
    const float S = SA->escale, iS = 1.0f / S;
    const float f0 = (float)(-SA->sumr);
    const bool publish = kPublish && !SB->first;
    const float *tab = lvl_inv_table(L);
    double s[2] = {0.0, 0.0}, m[1] = {0.0};
    const uint32_t nx = L.nx, ny = L.ny, nz = L.nz;
    const Tile4 T = tile4_geometry(nx, ny, nz);
    if (T.k0 < nz) {
        const TileW w = tile4_weights(L, T, tab);
        __shared__ __align__(16) float lds[tile_lds_len<float>()];
        const float cx = L.cx, cy = L.cy, cz = L.cz;
        const bool pin0 = T.i0 == 0 && T.j == 0;
        auto zin = [&](uint32_t k) { return (k > 0 && k + 1 < nz) ? 1 : 0; };
        typedef e16 H4 __attribute__((ext_vector_type(4)));
        struct Ld {
            H4 eo, eh;
            e16 ex;
            float4 fo, qo;
        };
        struct Sd {
            float f[4], q[4];
        };
        tile4_march<Ld, Sd>(
            T, nx, ny, nz, lds,
            [&](Ld &Q, uint32_t, uint32_t o, uint32_t h, uint32_t x) {
                Q.eo = *(const H4 *)(ein + o);
                Q.eh = *(const H4 *)(ein + h);
                Q.ex = ein[x];
                Q.fo = *(const float4 *)(f + o);
                if (kPublish) Q.qo = *(const float4 *)(qold + o);  // (unused on the first iteration)
            },
            [&](const Ld &Q, uint32_t k, Quad &out, Quad &hv, float &xv, Sd &sd) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    out.a[v] = (float)Q.eo[v] * iS;
                    hv.a[v] = (float)Q.eh[v] * iS;
                    sd.f[v] = (v == 0 && pin0 && k == 0) ? f0 : Q.fo[v];
                    sd.q[v] = kPublish ? Q.qo[v] : 0.0f;
                }
                xv = (float)Q.ex * iS;
            },
            [&](uint32_t k, uint32_t c, const Quad &em, const Quad &ec, const Quad &ep, const float *row,
                const Sd &sd) {
                const int z = zin(k);
                H4 o;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const float v0 = ec.a[v];
                    const float ve = v + 1 < 4 ? ec.a[v + 1] : row[4];
                    const float vw = v > 0 ? ec.a[v - 1] : row[-1];
                    const float vn = row[kTileS + v], vs = row[-kTileS + v];
                    const float ae = cx * ((v0 - ve) + (v0 - vw)) + cy * ((v0 - vn) + (v0 - vs)) +
                                     cz * ((v0 - ep.a[v]) + (v0 - em.a[v]));
                    const float out = v0 + w.o[z][v] * (sd.f[v] - ae);
                    o[v] = (e16)(out * S);
                    if (publish) {
                        const double qo = (double)sd.q[v];
                        s[0] += (double)out * qo;
                        s[1] += qo;
                    }
                }
                *(H4 *)(eout + c) = o;
            });
    }
    if (kPublish) publish_row<2, 1>(s, m, pout);
}

// k_mg_prolong0_z as a plane-tile march: e' = e_in + corr P(e_c) is formed
// per plane (P constant on the 2 x 2 x 2 children of a coarse cell, so a
// quad's four cells take two coarse values), then
// e_out = e' + omega D^-1 (f - A e').
__global__ void __launch_bounds__(kBlock) k_mg_prolong0_t(Level L, Level LC, const float *__restrict__ f,
                                                          const e16 *__restrict__ ein,
                                                          const float *__restrict__ ecoarse,
                                                          e16 *__restrict__ eout, const FcgA *SA,
                                                          const FcgB *SB, float corr) {
    if (SA->done || SB->done) return;
    const float S = SA->escale, iS = 1.0f / S;
    const float f0 = (float)(-SA->sumr);
    const float *tab = lvl_inv_table(L);
    const uint32_t nx = L.nx, ny = L.ny, nz = L.nz;
    const Tile4 T = tile4_geometry(nx, ny, nz);
    if (T.k0 >= nz) return;
    const TileW w = tile4_weights(L, T, tab);
    __shared__ __align__(16) float lds[tile_lds_len<float>()];
    const float cx = L.cx, cy = L.cy, cz = L.cz;
    const bool pin0 = T.i0 == 0 && T.j == 0;
    auto zin = [&](uint32_t k) { return (k > 0 && k + 1 < nz) ? 1 : 0; };
    // coarse columns of the own quad, the halo-row quad and the halo cell
    // (a lane without one reads the own column, or column 0)
    const uint32_t qo = T.own ? (T.i0 >> 1) + LC.nx * (T.j >> 1) : 0u;
    const uint32_t qh = T.hrow ? (T.i0 >> 1) + LC.nx * (T.jh >> 1) : qo;
    const uint32_t qx = T.xh ? (T.ix >> 1) + LC.nx * (T.jx >> 1) : qo;
    typedef e16 H4 __attribute__((ext_vector_type(4)));
    struct Ld {
        H4 eo, eh;
        e16 ex;
        float2 co, ch;
        float cxv;
        float4 fo;
    };
    tile4_march<Ld, Quad>(
        T, nx, ny, nz, lds,
        [&](Ld &Q, uint32_t k, uint32_t o, uint32_t h, uint32_t x) {
            const uint32_t pc = LC.nxy * (k >> 1);
            Q.eo = *(const H4 *)(ein + o);
            Q.eh = *(const H4 *)(ein + h);
            Q.ex = ein[x];
            Q.co = *(const float2 *)(ecoarse + qo + pc);
            Q.ch = *(const float2 *)(ecoarse + qh + pc);
            Q.cxv = ecoarse[qx + pc];
            Q.fo = *(const float4 *)(f + o);
        },
        [&](const Ld &Q, uint32_t k, Quad &out, Quad &hv, float &xv, Quad &fq) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                out.a[v] = (float)Q.eo[v] * iS + corr * (v < 2 ? Q.co.x : Q.co.y);
                hv.a[v] = (float)Q.eh[v] * iS + corr * (v < 2 ? Q.ch.x : Q.ch.y);
                fq.a[v] = (v == 0 && pin0 && k == 0) ? f0 : Q.fo[v];
            }
            xv = (float)Q.ex * iS + corr * Q.cxv;
        },
        [&](uint32_t k, uint32_t c, const Quad &em, const Quad &ec, const Quad &ep, const float *row,
            const Quad &fq) {
            const int z = zin(k);
            H4 o;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float v0 = ec.a[v];
                const float ve = v + 1 < 4 ? ec.a[v + 1] : row[4];
                const float vw = v > 0 ? ec.a[v - 1] : row[-1];
                const float vn = row[kTileS + v], vs = row[-kTileS + v];
                const float ae = cx * ((v0 - ve) + (v0 - vw)) + cy * ((v0 - vn) + (v0 - vs)) +
                                 cz * ((v0 - ep.a[v]) + (v0 - em.a[v]));
                o[v] = (e16)((v0 + w.o[z][v] * (fq.a[v] - ae)) * S);
            }
            *(H4 *)(eout + c) = o;
        });
}

// alpha = r.p / p.q; r -= alpha q; partials r.r, sum r, max|r|.  x is
// updated every second iteration: update k even only keeps its alpha (in
// SB->alpha_even, read by update k + 1, so no block reads what this launch
// writes), update k odd adds both, x += alpha_{k-1} p_{k-1} + alpha_k p_k
// (p_{k-1} is still in the other direction buffer), and update 1 writes x
// outright, so x needs no zero fill.  This saves reading and writing x in
// half of the iterations; a solve that ends after an even-numbered update
// leaves alpha_even p_{k} for k_add_correction.
__global__ void __launch_bounds__(kBlock) k_fcg_update(uint32_t n, float *__restrict__ x,
                                                       float *__restrict__ r,
                                                       const float *__restrict__ p,
                                                       const float *__restrict__ pprev,
                                                       const float *__restrict__ q,
                                                       const double *__restrict__ pin, uint32_t rows,
                                                       double *__restrict__ pout, const FcgA *SA,
                                                       FcgB *SB, long long k) {
    if (SA->done || SB->done) return;
    double f[2], fm[1];
    fold_partials<2, 1>(pin, rows, f, fm);
    const double pq = f[0], rp = f[1];
    int status = ST_OK;
    double alpha = 0.0;
    if (fabs(pq) < 1.0e-30 || !isfinite(pq)) status = ST_SINGULAR;
    else {
        alpha = rp / pq;
        if (!isfinite(alpha)) status = ST_ALPHA;
    }
    const float ta = (float)alpha;
    const bool odd = (k & 1) != 0;
    float tprev = 0.0f;
    if (odd) tprev = SB->alpha_even;  // written by update k - 1 (an even k writes, never reads it)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        SB->pq = pq;
        SB->first = 0;
        SB->status = status;
        if (!odd) SB->alpha_even = ta;
        if (status != ST_OK) SB->done = 1;
        else SB->iters += 1;
    }
    if (status != ST_OK) return;
    double s[3] = {0.0, 0.0, 0.0}, m[1] = {0.0};
    auto acc = [&](float rc) {
        s[1] += (double)rc * (double)rc;
        s[2] += (double)rc;
        m[0] = fmax(m[0], fabs((double)rc));
    };
    // x: 0 untouched (k even), 1 written (k == 1), 2 accumulated
    const int xmode = !odd ? 0 : (k == 1 ? 1 : 2);
    if (n % 2 == 0) {  // two cells per lane, 8-byte loads and stores
        TileSched ts((n / 2 + kBlock - 1) / kBlock);
        for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
            const uint32_t c = 2 * (t * kBlock + threadIdx.x);
            if (c >= n) continue;
            float2 rc = *(const float2 *)(r + c);
            const float2 qc = *(const float2 *)(q + c);
            if (xmode != 0) {
                const float2 pc = *(const float2 *)(p + c), pm = *(const float2 *)(pprev + c);
                float2 xc = make_float2(0.0f, 0.0f);
                if (xmode == 2) xc = *(const float2 *)(x + c);
                xc.x += tprev * pm.x + ta * pc.x;
                xc.y += tprev * pm.y + ta * pc.y;
                *(float2 *)(x + c) = xc;
            }
            rc.x -= ta * qc.x;
            rc.y -= ta * qc.y;
            *(float2 *)(r + c) = rc;
            acc(rc.x);
            acc(rc.y);
        }
    } else {
        TileSched ts((n + kBlock - 1) / kBlock);
        for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
            const uint32_t c = t * kBlock + threadIdx.x;
            if (c >= n) continue;
            if (xmode != 0) x[c] = (xmode == 2 ? x[c] : 0.0f) + (tprev * pprev[c] + ta * p[c]);
            const float rc = r[c] - ta * q[c];
            r[c] = rc;
            acc(rc);
        }
    }
    publish_partials<3, 1>(s, m, pout);
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

// Cell-centred velocity for output: the mean of each cell's two faces per
// component, 0.5 (u_i + u_{i+1}) etc. (write_legacy_vtk_3d, output.rs:170-180).
__global__ void __launch_bounds__(kBlock) k_cell_velocity(StepParams P, const double *__restrict__ u,
                                                          const double *__restrict__ v,
                                                          const double *__restrict__ w,
                                                          double *__restrict__ out) {
    const size_t n = (size_t)P.nx * P.ny * P.nz;
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % P.nx), j = (int)((c / P.nx) % P.ny), k = (int)(c / ((size_t)P.nx * P.ny));
        out[c] = 0.5 * (u[U3(i, j, k)] + u[U3(i + 1, j, k)]);
        out[n + c] = 0.5 * (v[V3(i, j, k)] + v[V3(i, j + 1, k)]);
        out[2 * n + c] = 0.5 * (w[W3(i, j, k)] + w[W3(i, j, k + 1)]);
    }
}

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

// predict_momentum (incompressible_3d.rs:174-273): central, advective form,
// all three face families in one march over the cells: the thread of cell
// (i, j) walks k and writes u(i, j, k), v(i, j, k) and w(i, j, k) (the last
// column, row and plane also write the boundary faces u(nx, .), v(., ny),
// w(., ., nz)).  Boundary-normal faces are written 0 (the reference
// zero-fills u*, v*, w*); neighbours outside the domain follow
// u/v/w_neighbor (:366-447): the mirror -c, or 2 lid - c above the lid for
// u.  The divisions by 2 dx and dx^2 are multiplications by reciprocals
// computed once (within an ulp of the quotient).
//
// The march runs on plane tiles.  A block of 4 waves owns a 64 x 4 (x, y) tile of cells over a
// chunk of planes; per plane it stages u, v and w over the tile and a
// one-cell frame (66 x 6 values each, frame cells outside an array read a
// clamped in-bounds cell that the selects never use) into one of three LDS
// buffers.  Computing plane k reads planes k and k + 1 from LDS and plane
// k - 1's values from registers; plane k + 2 is written to the third buffer
// after plane k is computed, from loads issued one plane earlier, so a
// buffer is never written while a slower wave still reads it and one
// barrier per plane suffices.  A cell costs about 4.6 global loads (its
// share of the three frames) instead of 17.
constexpr int kPTX = 64, kPTY = 4, kPFX = kPTX + 2, kPFY = kPTY + 2, kPF = kPFX * kPFY;
__global__ void __launch_bounds__(kBlock) k_predict3_t(StepParams P, const double *__restrict__ u,
                                                       const double *__restrict__ v,
                                                       const double *__restrict__ w,
                                                       double *__restrict__ us, double *__restrict__ vs,
                                                       double *__restrict__ ws) {
    __shared__ double tile[3][3][kPF];  // [buffer][u, v, w][frame cell]
    const int nx = (int)P.nx, ny = (int)P.ny, nz = (int)P.nz;
    const uint3 b = zblock();
    const int kz = (nz + (int)gridDim.z - 1) / (int)gridDim.z, k0 = (int)b.z * kz;
    if (k0 >= nz) return;  // block-uniform
    const int k1 = min(k0 + kz, nz);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x0 = (int)b.x * kPTX, y0 = (int)b.y * kPTY;
    const int i = x0 + lane, j = y0 + wave;
    const bool own = i < nx && j < ny;
    const uint32_t uy = (uint32_t)nx + 1, uz = (uint32_t)(nx + 1) * ny;  // u strides
    const uint32_t vy = (uint32_t)nx, vz = (uint32_t)nx * (ny + 1);      // v strides
    const uint32_t wy = (uint32_t)nx, wz = (uint32_t)nx * ny;            // w strides
    auto ld = [](const double *base, uint32_t e) -> double {
        return *(const double *)((const char *)base + (e << 3));
    };
    // this thread's frame cells (f = t and t + 256) and their clamped
    // in-plane offsets in u, v and w
    constexpr int kQ = (kPF + kBlock - 1) / kBlock;
    uint32_t fu[kQ], fv[kQ], fw[kQ];
#pragma unroll
    for (int q = 0; q < kQ; ++q) {
        const int f = min((int)threadIdx.x + q * (int)kBlock, kPF - 1);
        const int gx = x0 - 1 + f % kPFX, gy = y0 - 1 + f / kPFX;
        const int cxu = min(max(gx, 0), nx), cxc = min(max(gx, 0), nx - 1);
        const int cyc = min(max(gy, 0), ny - 1), cyv = min(max(gy, 0), ny);
        fu[q] = (uint32_t)cxu + uy * (uint32_t)cyc;
        fv[q] = (uint32_t)cxc + vy * (uint32_t)cyv;
        fw[q] = (uint32_t)cxc + wy * (uint32_t)cyc;
    }
    double su[kQ], sv[kQ], sw[kQ];
    auto stage_load = [&](int k) {  // plane k of u, v (clamped to nz - 1) and w (k <= nz)
        const uint32_t pu = uz * (uint32_t)min(k, nz - 1), pv = vz * (uint32_t)min(k, nz - 1),
                       pw = wz * (uint32_t)min(k, nz);
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            su[q] = ld(u, fu[q] + pu);
            sv[q] = ld(v, fv[q] + pv);
            sw[q] = ld(w, fw[q] + pw);
        }
    };
    auto stage_store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int f = (int)threadIdx.x + q * (int)kBlock;
            if (f < kPF) {
                tile[buf][0][f] = su[q];
                tile[buf][1][f] = sv[q];
                tile[buf][2][f] = sw[q];
            }
        }
    };
    const double hx = 1.0 / (2.0 * P.dx), hy = 1.0 / (2.0 * P.dy), hz = 1.0 / (2.0 * P.dz);
    const double ix2 = 1.0 / (P.dx * P.dx), iy2 = 1.0 / (P.dy * P.dy), iz2 = 1.0 / (P.dz * P.dz);
    const bool xw = i > 0, xe = i + 1 < nx, ys = j > 0, yn = j + 1 < ny;
    const int fc = (wave + 1) * kPFX + lane + 1;  // own frame cell
    // plane k0 - 1 (own values; the bottom wall's are never used)
    const uint32_t iu = (uint32_t)min(i, nx - 1) + uy * (uint32_t)min(j, ny - 1),
                   iv = (uint32_t)min(i, nx - 1) + vy * (uint32_t)min(j, ny - 1),
                   iw = (uint32_t)min(i, nx - 1) + wy * (uint32_t)min(j, ny - 1);
    const int km = k0 > 0 ? k0 - 1 : k0;
    double um = ld(u, iu + uz * km), vm = ld(v, iv + vz * km), wm = ld(w, iw + wz * km),
           uem = ld(u, iu + uz * km + 1), vnm = ld(v, iv + vz * km + vy);
    stage_load(k0);
    stage_store(k0 % 3);
    stage_load(k0 + 1);
    stage_store((k0 + 1) % 3);
    if (k0 + 1 < k1) stage_load(k0 + 2);
    __syncthreads();
    for (int k = k0; k < k1; ++k) {
        const double *U = tile[k % 3][0], *V = tile[k % 3][1], *W = tile[k % 3][2];
        const double *U1 = tile[(k + 1) % 3][0], *V1 = tile[(k + 1) % 3][1], *W1 = tile[(k + 1) % 3][2];
        const double u0 = U[fc], ue = U[fc + 1], uw = U[fc - 1], un = U[fc + kPFX], usn = U[fc - kPFX],
                     use = U[fc - kPFX + 1];
        const double v0 = V[fc], ve = V[fc + 1], vw = V[fc - 1], vn = V[fc + kPFX], vsn = V[fc - kPFX],
                     vnw = V[fc + kPFX - 1];
        const double w0 = W[fc], we = W[fc + 1], wn = W[fc + kPFX], ww0 = W[fc - 1], ws0 = W[fc - kPFX];
        const double up = U1[fc], vp = V1[fc], wp = W1[fc], wwp = W1[fc - 1], wsp = W1[fc - kPFX];
        if (own) {
            const uint32_t cu = (uint32_t)i + uy * j + uz * k, cv = (uint32_t)i + vy * j + vz * k,
                           cw = (uint32_t)i + wy * j + wz * k;
            const bool zt = k + 1 < nz;
            {
                const double c = u0, e = ue, wv = uw;
                const double n = yn ? un : 2.0 * P.lid - c, so = ys ? usn : -c;
                const double f = zt ? up : -c, bk = k > 0 ? um : -c;
                const double vv = 0.25 * (vw + v0 + vnw + vn);
                const double wwv = 0.25 * (ww0 + w0 + wwp + wp);
                const double val = c + P.dt * (-c * (e - wv) * hx - vv * (n - so) * hy - wwv * (f - bk) * hz +
                                               P.nu * ((e - 2.0 * c + wv) * ix2 + (n - 2.0 * c + so) * iy2 +
                                                       (f - 2.0 * c + bk) * iz2));
                us[cu] = xw ? val : 0.0;
            }
            if (!xe) us[cu + 1] = 0.0;
            {
                const double c = v0;
                const double e = xe ? ve : -c, wv = xw ? vw : -c, n = vn, so = vsn;
                const double f = zt ? vp : -c, bk = k > 0 ? vm : -c;
                const double uu = 0.25 * (usn + use + u0 + ue);
                const double wwv = 0.25 * (ws0 + w0 + wsp + wp);
                const double val = c + P.dt * (-uu * (e - wv) * hx - c * (n - so) * hy - wwv * (f - bk) * hz +
                                               P.nu * ((e - 2.0 * c + wv) * ix2 + (n - 2.0 * c + so) * iy2 +
                                                       (f - 2.0 * c + bk) * iz2));
                vs[cv] = ys ? val : 0.0;
            }
            if (!yn) vs[cv + vy] = 0.0;
            {
                const double c = w0;
                const double e = xe ? we : -c, wv = xw ? ww0 : -c, n = yn ? wn : -c, so = ys ? ws0 : -c;
                const double f = wp, bk = wm;
                const double uu = 0.25 * (um + uem + u0 + ue);
                const double vv = 0.25 * (vm + vnm + v0 + vn);
                const double val = c + P.dt * (-uu * (e - wv) * hx - vv * (n - so) * hy - c * (f - bk) * hz +
                                               P.nu * ((e - 2.0 * c + wv) * ix2 + (n - 2.0 * c + so) * iy2 +
                                                       (f - 2.0 * c + bk) * iz2));
                ws[cw] = k > 0 ? val : 0.0;
            }
            if (!zt) ws[cw + wz] = 0.0;
        }
        um = u0; vm = v0; wm = w0; uem = ue; vnm = vn;
        if (k + 1 < k1) {
            stage_store((k + 2) % 3);
            if (k + 2 < k1) stage_load(k + 3);
            __syncthreads();
        }
    }
}

// build_pressure_rhs (incompressible_3d.rs:275-291), rhs(0,0,0) = 0.
//
// kRes: the step's pressure solve starts from a zero guess in FP32-in-FP64
// refinement, whose first residual is r = b = -rhs (b_0 = 0): it is written
// here as FP32, with the residual kernel's partial row {r.z, r.r, sum r |
// max|r|} per block, so that kernel (a second pass over rhs) is skipped.
// The values are the residual kernel's bit for bit; only the summation
// order of the partial sums differs.
template <bool kRes>
__global__ void __launch_bounds__(kBlock) k_rhs(StepParams P, Grid g, const double *__restrict__ us,
                                                const double *__restrict__ vs,
                                                const double *__restrict__ ws,
                                                double *__restrict__ rhs, float *__restrict__ r32,
                                                double *__restrict__ pout) {
    const uint32_t n = P.nx * P.ny * P.nz;
    const double scale = P.rho / P.dt;
    const float *inv_tab = kRes ? inv_table<float>(g) : nullptr;
    double sa[3] = {0.0, 0.0, 0.0}, ma[1] = {0.0};
    const XcdRange r(n);
    for (uint32_t c = r.first; c < r.end; c += r.step) {
        uint32_t ui, uj, uk;
        cell_ijk(g, c, ui, uj, uk);  // multiply-shift dividers, no runtime divides
        const int i = (int)ui, j = (int)uj, k = (int)uk;
        const double rv = c == 0 ? 0.0
                                 : scale * (us[U3(i + 1, j, k)] - us[U3(i, j, k)]) / P.dx +
                                       scale * (vs[V3(i, j + 1, k)] - vs[V3(i, j, k)]) / P.dy +
                                       scale * (ws[W3(i, j, k + 1)] - ws[W3(i, j, k)]) / P.dz;
        rhs[c] = rv;
        if (kRes) {
            float rf = 0.0f;
            double inv = 1.0;
            if (c != 0) {
                rf = (float)(-rv);
                inv = (double)inv_tab[diag_index(g, (uint32_t)i, (uint32_t)j, (uint32_t)k)];
            }
            r32[c] = rf;
            const double rd = (double)rf;
            sa[0] += rd * (rd * inv);
            sa[1] += rd * rd;
            sa[2] += rd;
            ma[0] = fmax(ma[0], fabs(rd));
        }
    }
    if (kRes) publish_row<3, 1>(sa, ma, pout);
}

struct DiagState {
    double div_l2sq, max_div, max_speed;
    unsigned long long first_bad[4];  // first non-finite flat index per field
    unsigned counter;
};

// correct_velocity (:334-363) and, on the corrected field, max_divergence
// (:504-517), max_speed (:519-532), sum(div^2) and ensure_finite's scan
// (:534-554) of p, u, v, w, as one march over the cells.  The thread of cell
// (i, j) at plane k corrects its low faces u(i, j, k), v(i, j, k) (interior
// faces only, as the reference; a boundary face is stored back unchanged)
// and the top face w(i, j, k+1), carried up as the next cell's bottom face
// (a chunk's first bottom face is recomputed, not read, since the chunk
// below writes it).  The cell's high faces u(i+1, .), v(., j+1) are
// recomputed from us, vs and p with the same arithmetic their owners use.
// A face is checked for finiteness by the cell on its low side, the top
// boundary faces by the last cell of their row; each thread keeps the
// lowest non-finite index per field and posts it once.  Each block
// publishes one row {sum div^2 | max |div|, max speed}; k_diag_fold folds
// them.
//
// The march runs on plane tiles.  A block of 4 waves owns a 64 x 4 (x, y) tile of cells over a
// chunk of planes.  Per plane it stages p over the tile and a one-cell
// frame (66 x 6), the x faces us over the tile's 65 faces per row and the
// y faces vs over its 5 face rows; a boundary face is staged from the
// corrected field itself (u at x = 0 and x = nx, v at y = 0 and y = ny),
// which is the value the correction keeps there.  Cell (i, j, k) then reads
// p at its four in-plane neighbours and both of its x and y faces from LDS,
// p at k + 1 from the next plane's tile and its top face ws (w on the top
// wall) from a per-lane load; its bottom face is the previous plane's top.
// Three LDS buffers rotate as in k_predict3_t (one barrier per plane).
constexpr int kCFU = kPTX + 1, kCFV = kPTY + 1;  // face-tile widths
constexpr int kCU = kCFU * kPTY, kCV = kPTX * kCFV;
__global__ void __launch_bounds__(kBlock) k_correct_diag_t(StepParams P, const double *__restrict__ p,
                                                           const double *__restrict__ us,
                                                           const double *__restrict__ vs,
                                                           const double *__restrict__ ws,
                                                           double *__restrict__ u, double *__restrict__ v,
                                                           double *__restrict__ w,
                                                           double *__restrict__ pout, DiagState *dg) {
    __shared__ double tp[3][kPF], tu[3][kCU], tv[3][kCV];
    const int nx = (int)P.nx, ny = (int)P.ny, nz = (int)P.nz;
    double s[1] = {0.0}, m[2] = {0.0, 0.0};
    const uint3 b = zblock();
    const int kz = (nz + (int)gridDim.z - 1) / (int)gridDim.z, k0 = (int)b.z * kz;
    const int k1 = min(k0 + kz, nz);
    if (k0 < nz) {  // block-uniform
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int x0 = (int)b.x * kPTX, y0 = (int)b.y * kPTY;
        const int i = x0 + lane, j = y0 + wave;
        const bool own = i < nx && j < ny;
        const double scale = P.dt / P.rho;
        const uint32_t uy = (uint32_t)nx + 1, uz = (uint32_t)(nx + 1) * ny, vy = nx, vz = (uint32_t)nx * (ny + 1),
                       wy = nx, wz = (uint32_t)nx * ny;
        auto ld = [](const double *base, uint32_t e) -> double {
            return *(const double *)((const char *)base + (e << 3));
        };
        // staging duties: p frame cells t, t + 256; x faces t, t + 256 (of 260);
        // y faces t, t + 256 (of 320).  Offsets are clamped in-plane; a face on
        // the domain boundary is read from the corrected field.
        constexpr int kQ = 2;
        uint32_t op[kQ], ou[kQ], ov[kQ];
        bool bu[kQ], bv[kQ];
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int t = (int)threadIdx.x + q * (int)kBlock;
            const int fp = min(t, kPF - 1), fu = min(t, kCU - 1), fv = min(t, kCV - 1);
            const int px = min(max(x0 - 1 + fp % kPFX, 0), nx - 1), py = min(max(y0 - 1 + fp / kPFX, 0), ny - 1);
            op[q] = (uint32_t)px + wy * (uint32_t)py;
            const int ux = min(x0 + fu % kCFU, nx), uyy = min(y0 + fu / kCFU, ny - 1);
            ou[q] = (uint32_t)ux + uy * (uint32_t)uyy;
            bu[q] = ux == 0 || ux == nx;
            const int vx = min(x0 + fv % kPTX, nx - 1), vyy = min(y0 + fv / kPTX, ny);
            ov[q] = (uint32_t)vx + vy * (uint32_t)vyy;
            bv[q] = vyy == 0 || vyy == ny;
        }
        double sp[kQ], su[kQ], sv[kQ];
        auto stage_load = [&](int k) {  // plane k (p clamped to nz - 1)
            const uint32_t kk = (uint32_t)min(k, nz - 1);
            const uint32_t pp = wz * kk, pu = uz * kk, pv = vz * kk;
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                sp[q] = ld(p, op[q] + pp);
                su[q] = ld(bu[q] ? u : us, ou[q] + pu);
                sv[q] = ld(bv[q] ? v : vs, ov[q] + pv);
            }
        };
        auto stage_store = [&](int buf) {
#pragma unroll
            for (int q = 0; q < kQ; ++q) {
                const int t = (int)threadIdx.x + q * (int)kBlock;
                if (t < kPF) tp[buf][t] = sp[q];
                if (t < kCU) tu[buf][t] = su[q];
                if (t < kCV) tv[buf][t] = sv[q];
            }
        };
        const bool xw = i > 0, xe = i + 1 < nx, ys = j > 0, yn = j + 1 < ny;
        const int fcp = (wave + 1) * kPFX + lane + 1, fcu = wave * kCFU + lane, fcv = wave * kPTX + lane;
        const uint32_t iw = (uint32_t)min(i, nx - 1) + wy * (uint32_t)min(j, ny - 1);
        // the top face of plane k: ws, or w on the top wall
        auto top_face = [&](int k) -> double { return ld(k + 1 < nz ? ws : w, iw + wz * (uint32_t)(k + 1)); };
        unsigned long long bad[4] = {~0ull, ~0ull, ~0ull, ~0ull};
        auto flag = [&](int f, bool finite, uint32_t idx) {
            if (!finite) bad[f] = min(bad[f], (unsigned long long)idx);
        };
        // bottom face of the first cell (the chunk below writes it)
        double w0 = k0 == 0 ? ld(w, iw) : ld(ws, iw + wz * k0) - scale * (ld(p, iw + wz * k0) - ld(p, iw + wz * (k0 - 1))) / P.dz;
        double wt = top_face(k0);
        stage_load(k0);
        stage_store(k0 % 3);
        stage_load(k0 + 1);
        stage_store((k0 + 1) % 3);
        if (k0 + 1 < k1) stage_load(k0 + 2);
        __syncthreads();
        for (int k = k0; k < k1; ++k) {
            const double wtn = top_face(min(k + 1, nz - 1));  // the next plane's, in flight meanwhile
            const double *TP = tp[k % 3], *TU = tu[k % 3], *TV = tv[k % 3], *TP1 = tp[(k + 1) % 3];
            const double pc = TP[fcp], pw = TP[fcp - 1], pe = TP[fcp + 1], psn = TP[fcp - kPFX],
                         pn = TP[fcp + kPFX], pp = TP1[fcp];
            const double us0 = TU[fcu], us1 = TU[fcu + 1], vs0 = TV[fcv], vs1 = TV[fcv + kPTX];
            if (own) {
                const uint32_t cu = (uint32_t)i + uy * j + uz * k, cv = (uint32_t)i + vy * j + vz * k,
                               cw = (uint32_t)i + wy * j + wz * k;
                const bool zt = k + 1 < nz;
                const double u0 = xw ? us0 - scale * (pc - pw) / P.dx : us0;
                const double u1 = xe ? us1 - scale * (pe - pc) / P.dx : us1;
                const double v0 = ys ? vs0 - scale * (pc - psn) / P.dy : vs0;
                const double v1 = yn ? vs1 - scale * (pn - pc) / P.dy : vs1;
                const double w1 = zt ? wt - scale * (pp - pc) / P.dz : wt;
                *(double *)((char *)u + (cu << 3)) = u0;
                *(double *)((char *)v + (cv << 3)) = v0;
                *(double *)((char *)w + ((cw + wz) << 3)) = w1;
                const double div = (u1 - u0) / P.dx + (v1 - v0) / P.dy + (w1 - w0) / P.dz;
                s[0] += div * div;
                m[0] = fmax(m[0], fabs(div));
                const double a = 0.5 * (u0 + u1), bb = 0.5 * (v0 + v1), cc = 0.5 * (w0 + w1);
                m[1] = fmax(m[1], sqrt(a * a + bb * bb + cc * cc));
                flag(0, isfinite(pc), cw);
                flag(1, isfinite(u0), cu);
                flag(1, xe || isfinite(u1), cu + 1);
                flag(2, isfinite(v0), cv);
                flag(2, yn || isfinite(v1), cv + vy);
                flag(3, isfinite(w0), cw);
                flag(3, zt || isfinite(w1), cw + wz);
                w0 = w1;
            }
            wt = wtn;
            if (k + 1 < k1) {
                stage_store((k + 2) % 3);
                if (k + 2 < k1) stage_load(k + 3);
                __syncthreads();
            }
        }
        if (own) {
#pragma unroll
            for (int f = 0; f < 4; ++f)
                if (bad[f] != ~0ull) atomicMin(&dg->first_bad[f], bad[f]);
        }
    }
    publish_row<1, 2>(s, m, pout);
}
// The diagnostics state of a step: sums and maxima zero, no non-finite value.
__global__ void k_diag_reset(DiagState *dg) {
    DiagState d{};
    for (int f = 0; f < 4; ++f) d.first_bad[f] = ~0ull;
    *dg = d;
}
__global__ void __launch_bounds__(kBlock) k_diag_fold(const double *__restrict__ pin, uint32_t rows,
                                                      DiagState *dg) {
    double s[1], m[2];
    fold_partials<1, 2>(pin, rows, s, m);
    if (threadIdx.x == 0) {
        dg->div_l2sq = s[0];
        dg->max_div = m[0];
        dg->max_speed = m[1];
    }
}

#undef U3
#undef V3
#undef W3

// ---------------------------------------------------------------------------
// Learned warm start (see warm_start): the U-Net sees the initial residual
// mean-restricted to its resolution and predicts the correction there.
// ---------------------------------------------------------------------------

// Mean of the 8 children: level L (fine) -> L+1.
__global__ void __launch_bounds__(kBlock) k_restrict_mean(Level L, Level LC, const float *__restrict__ f,
                                                          float *__restrict__ fc) {
    for (uint32_t C = blockIdx.x * blockDim.x + threadIdx.x; C < LC.n; C += gridDim.x * blockDim.x) {
        uint32_t I, J, K;
        lvl_ijk(LC, C, I, J, K);
        float acc = 0.0f;
#pragma unroll
        for (int a = 0; a < 8; ++a)
            acc += f[(2 * I + (a & 1)) + L.nx * ((2 * J + (a >> 1 & 1)) + L.ny * (2 * K + (a >> 2)))];
        fc[C] = 0.125f * acc;
    }
}

// x += sign (e - e_0) (FP64 += FP32), keeping the pinned cell's value;
// sign -1 undoes sign +1.
__global__ void __launch_bounds__(kBlock) k_add_shifted(uint32_t n, double *__restrict__ x,
                                                        const float *__restrict__ e, double sign) {
    const double e0 = (double)e[0];
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        x[c] += sign * ((double)e[c] - e0);
}

// Partials of r.e~ and e~.A e~ with e~ = e - e_0 on the pinned operator
// (published as <3,1> rows {r.e~, e~.Ae~, 0 | 0}).
__global__ void __launch_bounds__(kBlock) k_warm_dots(Grid g, const float *__restrict__ r,
                                                      const float *__restrict__ e,
                                                      double *__restrict__ partials) {
    double s[3] = {0.0, 0.0, 0.0}, m[1] = {0.0};
    const float e0 = e[0];
    const float cx = (float)g.ix2, cy = (float)g.iy2, cz = (float)g.iz2;
    TileSched ts((g.n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t c = t * kBlock + threadIdx.x;
        if (c >= g.n || c == 0) continue;
        uint32_t i, j, k;
        cell_ijk(g, c, i, j, k);
        const Nbr o = neighbours(g, c, i, j, k);
        const uint32_t idx[7] = {c, o.e, o.w, o.n, o.s, o.f, o.b};
        float ev[7];
#pragma unroll
        for (int a = 0; a < 7; ++a) ev[a] = e[idx[a]] - e0;  // the pinned cell's e~ is 0
        const float ae = cx * ((ev[0] - ev[1]) + (ev[0] - ev[2])) + cy * ((ev[0] - ev[3]) + (ev[0] - ev[4])) +
                         cz * ((ev[0] - ev[5]) + (ev[0] - ev[6]));
        s[0] += (double)r[c] * (double)ev[0];
        s[1] += (double)ev[0] * (double)ae;
    }
    publish_partials<3, 1>(s, m, partials);
}

// out[C] = sign * mean of the f^3 fine cells of coarse cell C (nc^3 cells).
__global__ void __launch_bounds__(kBlock) k_block_mean(Grid g, const double *__restrict__ src, double sign,
                                                       uint32_t nc, uint32_t f, float *__restrict__ out) {
    const uint32_t ncell = nc * nc * nc;
    const double w = sign / ((double)f * f * f);
    for (uint32_t C = blockIdx.x * blockDim.x + threadIdx.x; C < ncell; C += gridDim.x * blockDim.x) {
        const uint32_t I = C % nc, J = (C / nc) % nc, K = C / (nc * nc);
        double acc = 0.0;
        for (uint32_t dk = 0; dk < f; ++dk)
            for (uint32_t dj = 0; dj < f; ++dj)
                for (uint32_t di = 0; di < f; ++di)
                    acc += src[(I * f + di) + g.nx * ((J * f + dj) + g.ny * (K * f + dk))];
        out[C] = (float)(w * acc);
    }
}

// Kernel kinds of the per-kernel profiler, named as the kernels are (with
// their template arguments).
enum Kind {
    K_PREDICT, K_RHS, K_CORRECT,
    K_RESIDUAL_D, K_RESIDUAL_F, K_FOLD_OUTER, K_ADD_CORRECTION,
    K_DIRAPPLY_D, K_UPDATE_D, K_DIRAPPLY_F, K_UPDATE_F,
    K_MG_CHECK, K_MG_SWEEP_L0, K_MG_SWEEP_L0_PROLONG, K_MG_SWEEP_L0_PUBLISH,
    K_MG_RESTRICT_L0, K_MG_DOWN, K_MG_TAIL, K_MG_UP, K_FCG_DIRAPPLY, K_FCG_UPDATE, K_WARM_START,
    K_UNET_COARSE,
    K_CGDOT_D, K_CGSTEP_D, K_CGDOT_F, K_CGSTEP_F, K_FCG_DIRAPPLY_T, K_MG_CHECK_T,
    K_MG_SWEEP_L0_T, K_MG_PROLONG_L0_T, K_MG_PUBLISH_L0_T,
    K_COUNT
};
const char *const kKindNames[K_COUNT] = {
    "k_predict3_t", "k_rhs", "k_correct_diag_t",
    "k_residual_z<double>", "k_residual_z<float>", "k_fold_outer", "k_add_correction",
    "k_cg_dirapply_z<double>", "k_cg_update<double, double>",
    "k_cg_dirapply_z<float>", "k_cg_update<float, float>",
    "k_mg_check_z", "k_mg_sweep0_z<false>",
    "k_mg_prolong0_z", "k_mg_sweep0_z<true>",
    "k_mg_restrict_z<e16>", "k_mg_down", "k_mg_tail", "k_mg_up", "k_fcg_dirapply_z", "k_fcg_update",
    "unet_warm_start",  // restriction + U-Net inference + interpolation (many kernels)
    "unet_coarse_solve",  // U-Net inference + the residual on its level (many kernels)
    "k_cg_dot_t<double>", "k_cg_step_t<double>", "k_cg_dot_t<float>", "k_cg_step_t<float>",
    "k_fcg_dirapply_t", "k_mg_check_t",
    "k_mg_sweep0_t<false>", "k_mg_prolong0_t", "k_mg_sweep0_t<true>",
};

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
           *ws = nullptr, *p = nullptr, *rhs = nullptr, *partials = nullptr;
    // FP64 PCG vectors
    double *r = nullptr, *d[2] = {nullptr, nullptr}, *q = nullptr;
    // FP32 inner-solver vectors of the mixed path
    float *e32 = nullptr, *r32 = nullptr, *d32[2] = {nullptr, nullptr}, *q32 = nullptr;
    double *pa = nullptr, *pb = nullptr;  // partials of dirapply / update
    CgStateA *sa = nullptr;
    CgStateB *sb = nullptr;
    DiagState *dg = nullptr;
    // multigrid hierarchy (FL_PRECOND_MULTIGRID): levels 0..tail-1 in global
    // memory, levels tail..nlev-1 inside k_mg_tail
    int nlev = 0, tail = 0;
    Level lv[kMaxLevels]{};
    float *mf[kMaxLevels] = {}, *mea[kMaxLevels] = {}, *meb[kMaxLevels] = {};  // levels >= 1
    e16 *m0a = nullptr, *m0b = nullptr;  // level 0's iterates
    double e_bound = 1.0;                 // L_max^2 / pi^2 (see e_scale_for)
    uint32_t lblocks[kMaxLevels] = {};
    ZGrid zl[kMaxLevels] = {};  // column-marching launch of levels 1..kFusedFrom-1
    float *pinv = nullptr;  // coarsest level's pseudo-inverse (<= kDirectCells cells)
    int fcg_hint[kFcgHints] = {};  // inner FCG updates of the last step, per refinement round
    // learned warm start (FL_PRECOND_LEARNED): the network works on level
    // wl_level; wl_in / wl_out hold its input and output there
    fl_unet *unet = nullptr;
    int wl_level = 0;
    int learned_mode = FL_LEARNED_CORRECTED;  // FL_LEARNED_*
    double wl_alpha = 0.0;   // step of the last warm start along wl_e
    const float *wl_e = nullptr;
    float *wl_in = nullptr, *wl_out = nullptr;
    float *pair = nullptr;  // fl_cavity3d_coarse_pair staging (2 n^3)
    size_t pair_cells = 0;
    // training capture: the right-hand side of V-cycle cap_iter of the first
    // inner FCG solve of each pressure solve, on level cap_level (0: off)
    int cap_level = 0, cap_iter = 0;
    bool cap_taken = false;
    // p is all zeros (a zero initial guess, nothing written since): the next
    // residual is r = b and reads no p
    bool p_zero = false;
    // p's contents stand for zero and were not cleared (a zero guess solved
    // by the mixed multigrid path): the first correction writes p = e, and a
    // solve that ends before any correction clears p then
    bool p_stale = false;
    // partial rows of the latest residual in pb; r_ready: the zero guess's
    // FP32 residual is already in r32 (written with the right-hand side)
    uint32_t rrows = 0;
    bool r_ready = false;
    float *cap = nullptr;
    double *pz = nullptr;  // partials z.q_old, sum q_old
    // the flexible-CG state: one allocation, fb directly after fa (as in
    // Host), so one copy brings both to the host
    FcgA *fa = nullptr;
    FcgB *fb = nullptr;
    struct Host {
        CgStateA a;
        CgStateB b;
        FcgA fa;
        FcgB fb;
        DiagState dg;
        double outer[4];
    } *host = nullptr;
    static_assert(offsetof(Host, fb) == offsetof(Host, fa) + sizeof(FcgA), "fa, fb adjacent in Host");
    static_assert(sizeof(FcgA) % alignof(FcgB) == 0, "fb after fa is aligned");
    double *outer = nullptr;  // folded outer residual: r.z, r.r, max|r|
    SorState *sor = nullptr;  // SOR iteration state (FL_SOLVER_SOR)
    uint32_t blocks = 1, ew_blocks = 1;
    ZGrid zg{1, 1, 1, 1};   // column-marching launch of the level-0 sweeps
    ZGrid zgd{1, 1, 1, 1};  // and of the search directions
    ZGrid zr{1, 1, 1, 1};   // of the residual (two cells per lane when nx is even)
    ZGrid zp{1, 1, 1, 1};   // of the FP64 search direction (likewise, 32-plane marches)
    // plane-tile Jacobi-PCG passes (FP64, FP32): the dot pass's and the step
    // pass's launches; count 0 where nx is not a multiple of the lane width
    // (2 FP64, 4 FP32 cells)
    ZGrid zt{0, 0, 0, 0}, ztf{0, 0, 0, 0}, zts{0, 0, 0, 0}, ztfs{0, 0, 0, 0};
    // plane-tile FCG search direction (count 0: nx not a multiple of 4, the
    // column march runs instead)
    ZGrid ztd{0, 0, 0, 0};
    // plane-tile level-0 cycle passes: stop test and first sweeps, sweep,
    // prolongation, publishing sweep (count 0: the column marches run)
    ZGrid ztc{0, 0, 0, 0}, zts0{0, 0, 0, 0}, ztp0{0, 0, 0, 0}, zts1{0, 0, 0, 0};
    ZGrid zpt{1, 1, 1, 1}, zct{1, 1, 1, 1};  // plane-tile predictor and correction
    int64_t step = 0;
    KProf prof;  // per-kernel timing (HIP events on the solver's stream)
    double *cell_out = nullptr;  // output staging: cell-centred u, v, w (allocated on first use)
};

namespace {

constexpr int kMaxBatch = 64;

// out = {s0, s1, max, s2} of <3,1> partials (for k_residual_z: r.z, r.r,
// max|r|, sum r).
__global__ void k_fold_outer(const double *__restrict__ partials, uint32_t rows, double *out) {
    double s[3], m[1];
    fold_partials<3, 1>(partials, rows, s, m);
    if (threadIdx.x == 0) {
        out[0] = s[0];
        out[1] = s[1];
        out[2] = m[0];
        out[3] = s[2];
    }
}

template <typename T>
fl_status alloc(fl_cavity3d *s, T **ptr, size_t count) {
    FL_HIP_TRY(s->ctx, hipMalloc(ptr, count * sizeof(T)));
    FL_HIP_TRY(s->ctx, hipMemsetAsync(*ptr, 0, count * sizeof(T), s->ctx->stream));
    return FL_OK;
}

void free_all(fl_cavity3d *s) {
    void *bufs[] = {s->u,   s->v,      s->w,      s->us,  s->vs,     s->ws,     s->p,
                    s->rhs, s->r,      s->d[0],   s->d[1], s->q,     s->e32,    s->r32,
                    s->d32[0], s->d32[1], s->q32, s->pa,   s->pb,     s->partials, s->sa,
                    s->sb,  s->dg,     s->outer};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    void *more[] = {s->pz, s->fa, s->wl_in, s->wl_out, s->pair, s->pinv, s->m0a, s->m0b, s->cap,
                    s->sor, s->cell_out};
    for (void *b : more)
        if (b) (void)hipFree(b);
    for (int l = 0; l < kMaxLevels; ++l) {
        if (s->mf[l]) (void)hipFree(s->mf[l]);
        if (s->mea[l]) (void)hipFree(s->mea[l]);
        if (s->meb[l]) (void)hipFree(s->meb[l]);
    }
    if (s->host) (void)hipHostFree(s->host);
    s->prof.destroy();
}

// Dense pseudo-inverse of a level's operator (the 7-point Neumann Laplacian
// with weights cx, cy, cz; null space the constants), for the exact solve of
// the coarsest level:  P = M^-1 - 11^T / (s n) with M = A + (s / n) 11^T,
// s = mean diagonal.  M is SPD, M^-1 maps the mean-free vectors to
// themselves, so P b is the mean-free solution of A e = b - mean(b).
// Gauss-Jordan elimination in double; n <= kDirectCells.
std::vector<float> coarse_pseudo_inverse(const Level &L) {
    const uint32_t n = L.n;
    std::vector<double> M((size_t)n * n, 0.0), I((size_t)n * n, 0.0);
    const double w[3] = {L.cx, L.cy, L.cz};
    double diag = 0.0;
    for (uint32_t c = 0; c < n; ++c) {
        const uint32_t i = c % L.nx, j = (c / L.nx) % L.ny, k = c / L.nxy;
        const int64_t cc = c;
        const int64_t nb[6] = {i + 1 < L.nx ? cc + 1 : -1,     i > 0 ? cc - 1 : -1,
                               j + 1 < L.ny ? cc + L.nx : -1,  j > 0 ? cc - L.nx : -1,
                               k + 1 < L.nz ? cc + L.nxy : -1, k > 0 ? cc - L.nxy : -1};
        for (int a = 0; a < 6; ++a) {
            if (nb[a] < 0) continue;
            M[(size_t)c * n + c] += w[a / 2];
            M[(size_t)c * n + nb[a]] -= w[a / 2];
        }
        diag += M[(size_t)c * n + c];
        I[(size_t)c * n + c] = 1.0;
    }
    const double sc = diag / n;
    for (size_t a = 0; a < (size_t)n * n; ++a) M[a] += sc / n;
    for (uint32_t col = 0; col < n; ++col) {
        uint32_t piv = col;
        for (uint32_t r = col + 1; r < n; ++r)
            if (std::fabs(M[(size_t)r * n + col]) > std::fabs(M[(size_t)piv * n + col])) piv = r;
        if (piv != col)
            for (uint32_t a = 0; a < n; ++a) {
                std::swap(M[(size_t)piv * n + a], M[(size_t)col * n + a]);
                std::swap(I[(size_t)piv * n + a], I[(size_t)col * n + a]);
            }
        const double d = 1.0 / M[(size_t)col * n + col];
        for (uint32_t a = 0; a < n; ++a) {
            M[(size_t)col * n + a] *= d;
            I[(size_t)col * n + a] *= d;
        }
        for (uint32_t r = 0; r < n; ++r) {
            if (r == col) continue;
            const double f = M[(size_t)r * n + col];
            if (f == 0.0) continue;
            for (uint32_t a = 0; a < n; ++a) {
                M[(size_t)r * n + a] -= f * M[(size_t)col * n + a];
                I[(size_t)r * n + a] -= f * I[(size_t)col * n + a];
            }
        }
    }
    std::vector<float> P((size_t)n * n);
    for (size_t a = 0; a < (size_t)n * n; ++a) P[a] = (float)(I[a] - 1.0 / (sc * n));
    return P;
}

// Multigrid levels: halve every axis while all are even and at least 8;
// the coarse coefficients halve per level (Galerkin-consistent with the
// mean restriction and piecewise-constant prolongation).  Levels from the
// first one with <= kTailCells cells (or the next, when it has <= kTailFast)
// run inside k_mg_tail.
fl_status build_hierarchy(fl_cavity3d *s) {
    uint32_t nx = s->g.nx, ny = s->g.ny, nz = s->g.nz;
    double cx = s->g.ix2, cy = s->g.iy2, cz = s->g.iz2;
    int nl = 0;
    for (;;) {
        Level &L = s->lv[nl];
        L.nx = nx; L.ny = ny; L.nz = nz; L.nxy = nx * ny; L.n = nx * ny * nz;
        L.cx = (float)cx; L.cy = (float)cy; L.cz = (float)cz;
        L.fx = FastDiv(nx);
        L.fxy = FastDiv(nx * ny);
        for (uint32_t b = 0; b < 8; ++b) {
            const double dg = cx * (1 + (b & 1)) + cy * (1 + (b >> 1 & 1)) + cz * (1 + (b >> 2 & 1));
            L.inv[b] = (float)(1.0 / dg);
        }
        ++nl;
        const bool can = nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0 && nx >= 8 && ny >= 8 &&
                         nz >= 8 && nl < kMaxLevels;
        if (!can) break;
        nx /= 2; ny /= 2; nz /= 2;
        cx *= 0.5; cy *= 0.5; cz *= 0.5;
    }
    s->nlev = nl;
    if (nl < 2 || s->lv[nl - 1].n > kTailCells)
        return set_error(s->ctx, FL_ERR_CONFIG,
                         "multigrid needs grid dimensions that halve (all even, >= 8) down to at "
                         "most %u cells", kTailCells);
    int t = 1;
    while (s->lv[t].n > kTailCells) ++t;
    if (t + 1 < nl && s->lv[t + 1].n <= kTailFast) ++t;
    s->tail = t;
    // the longest side from the level-0 coefficients (cx = 1 / hx^2)
    const double lx = s->g.nx / std::sqrt(s->g.ix2), ly = s->g.ny / std::sqrt(s->g.iy2),
                 lz = s->g.nz / std::sqrt(s->g.iz2);
    const double lmax = std::max(lx, std::max(ly, lz));
    s->e_bound = lmax * lmax / (M_PI * M_PI);
    if (nl - t > kTailMax)
        return set_error(s->ctx, FL_ERR_CONFIG, "multigrid tail of %d levels (at most %d)", nl - t,
                         kTailMax);
    for (int l = 0; l < nl; ++l) {
        const Level &L = s->lv[l];
        s->lblocks[l] = grid_blocks(s->ctx, (L.n + kBlock - 1) / kBlock, 4);
        // an even march length (the restriction pairs planes) short enough
        // for about 1024 workgroups
        uint32_t kz = 4;
        while (kz < kZK && L.n / ((uint64_t)kZT * kz * 2) >= 1024) kz *= 2;
        s->zl[l] = zgrid(L.nx, L.ny, L.nz, kz);
    }
    return FL_OK;
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
    if (c->precision != FL_PRECISION_FP64 && c->precision != FL_PRECISION_MIXED)
        return set_error(ctx, FL_ERR_CONFIG, "unknown precision %d", c->precision);
    if (!(c->inner_tolerance >= 0.0 && c->inner_tolerance < 1.0))
        return set_error(ctx, FL_ERR_CONFIG, "inner_tolerance must lie in [0, 1)");
    if (c->preconditioner != FL_PRECOND_JACOBI && c->preconditioner != FL_PRECOND_MULTIGRID &&
        c->preconditioner != FL_PRECOND_LEARNED)
        return set_error(ctx, FL_ERR_CONFIG, "unknown preconditioner %d", c->preconditioner);
    if (c->preconditioner != FL_PRECOND_JACOBI && c->precision != FL_PRECISION_MIXED)
        return set_error(ctx, FL_ERR_CONFIG, "the multigrid preconditioner runs in the mixed-precision solve");
    if (c->pressure_solver != FL_SOLVER_PCG && c->pressure_solver != FL_SOLVER_SOR)
        return set_error(ctx, FL_ERR_CONFIG, "unknown pressure solver %d", c->pressure_solver);
    if (c->pressure_solver == FL_SOLVER_SOR) {
        if (!(std::isfinite(c->pressure_omega) && c->pressure_omega > 0.0 && c->pressure_omega < 2.0))
            return set_error(ctx, FL_ERR_CONFIG, "pressure_omega must lie between 0 and 2");
        if (c->precision != FL_PRECISION_FP64 || c->preconditioner != FL_PRECOND_JACOBI ||
            c->stop_rule != FL_STOP_ABS_INF)
            return set_error(ctx, FL_ERR_CONFIG,
                             "the SOR pressure solver runs in FP64 with the reference's max-residual stop rule");
    }
    const double cells = (double)c->nx * (double)c->ny * (double)c->nz;
    const double faces = (double)(c->nx + 1) * (double)(c->ny + 1) * (double)(c->nz + 1) * 3.0;
    if (cells + faces >= 2147483647.0)
        return set_error(ctx, FL_ERR_CONFIG, "grid too large for 32-bit cell indexing");
    // the step kernels address each FP64 field by a 32-bit byte offset
    const double face_array = std::max((double)(c->nx + 1) * c->ny * c->nz,
                                       std::max((double)c->nx * (c->ny + 1) * c->nz, (double)c->nx * c->ny * (c->nz + 1)));
    if (face_array * 8.0 >= 4294967296.0)
        return set_error(ctx, FL_ERR_CONFIG, "grid too large for 32-bit byte offsets (a field array reaches 4 GiB)");
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

// Runs CG iterations (dirapply + update pairs) on vectors of type T until
// the device state says done.  The "update" partials of iteration 0 must
// already be in s->pb (from k_residual_z: `rows0` rows, by default ZGrid
// zr's count).
template <typename T, typename TX>
fl_status run_cg(fl_cavity3d *s, const CgParams &prm, TX *x, T *r, T *const dbuf[2], T *q,
                 uint32_t rows0 = 0) {
    if (rows0 == 0) rows0 = s->zr.count;
    fl_ctx *ctx = s->ctx;
    hipStream_t stm = ctx->stream;
    s->host->a = CgStateA{};
    s->host->b = CgStateB{};
    s->host->b.first = 1;
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->sa, &s->host->a, sizeof(CgStateA), hipMemcpyHostToDevice, stm));
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->sb, &s->host->b, sizeof(CgStateB), hipMemcpyHostToDevice, stm));
    const uint32_t G = s->blocks;
    const bool fp64 = sizeof(T) == 8;
    const ZGrid &zt = fp64 ? s->zt : s->ztf, &zts = fp64 ? s->zts : s->ztfs;
    auto aligned = [](const void *p) { return ((uintptr_t)p & 15u) == 0; };
    if (sizeof(TX) == sizeof(T) && zt.count > 0 && aligned(x) && aligned(r) && aligned(q) &&
        aligned(dbuf[0]) && aligned(dbuf[1])) {
        // plane-tile passes; q is free (A d' is recomputed), so it is the
        // second residual buffer: iteration i reads r, d from slot i % 2 and
        // writes r', d' to the other.  The final residual is scratch.
        T *const rbuf[2] = {r, q};
        const dim3 tg(zt.bx, zt.by, zt.bz), tgs(zts.bx, zts.by, zts.bz);
        const int kd = fp64 ? K_CGDOT_D : K_CGDOT_F, ks = fp64 ? K_CGSTEP_D : K_CGSTEP_F;
        // algorithmic bytes: the dot pass reads r, d; the step reads x, r, d
        // and writes x, r', d'
        const double bd = 2.0 * sizeof(T) * s->n, bs = 6.0 * sizeof(T) * s->n;
        int batch = 8;
        long long launched = 0;
        for (;;) {
            for (int b = 0; b < batch; ++b) {
                const int a = (int)(launched & 1);
                FL_TIMED(s, kd, launched, bd,
                         k_cg_dot_t<T><<<tg, kBlock, 0, stm>>>(s->g, prm, rbuf[a], dbuf[a], s->pb,
                                                               launched == 0 ? rows0 : zts.count, s->pa,
                                                               s->sa, s->sb));
                FL_TIMED(s, ks, launched, bs,
                         k_cg_step_t<T><<<tgs, kBlock, 0, stm>>>(s->g, (T *)x, rbuf[a], dbuf[a], rbuf[a ^ 1],
                                                                 dbuf[a ^ 1], s->pa, zt.count, s->pb, s->sa,
                                                                 s->sb, prm.neumann));
                ++launched;
            }
            FL_HIP_TRY(ctx, hipGetLastError());
            FL_HIP_TRY(ctx, hipMemcpyAsync(&s->host->a, s->sa, sizeof(CgStateA), hipMemcpyDeviceToHost, stm));
            FL_HIP_TRY(ctx, hipMemcpyAsync(&s->host->b, s->sb, sizeof(CgStateB), hipMemcpyDeviceToHost, stm));
            FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
            FL_HIP_TRY(ctx, s->prof.flush(s->host->b.iters));
            if (s->host->a.done || s->host->b.done) break;
            batch = std::min(batch * 2, kMaxBatch);
        }
        return FL_OK;
    }
    // the FP64 direction marches two cells per lane on even-x grids (zp); the
    // FP32 one is faster one cell per lane (zgd)
    const ZGrid &zd = fp64 ? s->zp : s->zgd;
    const dim3 zg(zd.bx, zd.by, zd.bz);
    const int kd = fp64 ? K_DIRAPPLY_D : K_DIRAPPLY_F, ku = fp64 ? K_UPDATE_D : K_UPDATE_F;
    // algorithmic bytes: dirapply reads r, d and writes d', q; update reads
    // x, d', r, q and writes x, r
    const double bd = 4.0 * sizeof(T) * s->n, bu = (2.0 * sizeof(TX) + 4.0 * sizeof(T)) * s->n;
    int batch = 8;
    long long launched = 0;
    for (;;) {
        for (int b = 0; b < batch; ++b) {
            // iteration `launched` reads d[launched % 2] and writes the other
            T *dold = dbuf[launched & 1], *dnew = dbuf[(launched + 1) & 1];
            FL_TIMED(s, kd, launched, bd,
                     k_cg_dirapply_z<T><<<zg, kZT, 0, stm>>>(s->g, prm, r, dold, dnew, q, s->pb,
                                                              launched == 0 ? rows0 : G, s->pa,
                                                              s->sa, s->sb));
            FL_TIMED(s, ku, launched, bu,
                     k_cg_update<T, TX><<<G, kBlock, 0, stm>>>(s->g, x, r, dnew, q, s->pa, zd.count,
                                                               s->pb, s->sa, s->sb, prm.neumann));
            ++launched;
        }
        FL_HIP_TRY(ctx, hipGetLastError());
        FL_HIP_TRY(ctx, hipMemcpyAsync(&s->host->a, s->sa, sizeof(CgStateA), hipMemcpyDeviceToHost, stm));
        FL_HIP_TRY(ctx, hipMemcpyAsync(&s->host->b, s->sb, sizeof(CgStateB), hipMemcpyDeviceToHost, stm));
        FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
        // only iterations whose update ran count (later launches exit early)
        FL_HIP_TRY(ctx, s->prof.flush(s->host->b.iters));
        if (s->host->a.done || s->host->b.done) break;
        batch = std::min(batch * 2, kMaxBatch);
    }
    return FL_OK;
}

// r = b - A p into r, partials (s->zr.count rows) into pb.
template <typename TR>
void launch_residual(fl_cavity3d *s, TR *r) {
    const dim3 grid(s->zr.bx, s->zr.by, s->zr.bz);
    s->rrows = s->zr.count;
    const bool z = s->p_zero;
    s->p_zero = false;
    hipStream_t stm = s->ctx->stream;
    if (s->g.nx % 2 == 0) {
        if (z) k_residual_z2<TR, true><<<grid, kZT, 0, stm>>>(s->g, s->rhs, s->p, r, s->pb);
        else k_residual_z2<TR, false><<<grid, kZT, 0, stm>>>(s->g, s->rhs, s->p, r, s->pb);
    } else {
        if (z) k_residual_z<TR, true><<<grid, kZT, 0, stm>>>(s->g, s->rhs, s->p, r, s->pb);
        else k_residual_z<TR, false><<<grid, kZT, 0, stm>>>(s->g, s->rhs, s->p, r, s->pb);
    }
}

// Bytes of the next launch_residual: rhs, r and x unless x is known zero.
template <typename TR>
double residual_bytes(const fl_cavity3d *s) {
    return (8.0 + sizeof(TR) + (s->p_zero ? 0.0 : 8.0)) * s->n;
}

fl_status solve_fp64(fl_cavity3d *s, fl_solve_report *rep, double rr0_ref = 0.0) {
    fl_ctx *ctx = s->ctx;
    FL_TIMED(s, K_RESIDUAL_D, -1, residual_bytes<double>(s),
             launch_residual<double>(s, s->r));
    FL_HIP_TRY(ctx, hipGetLastError());
    const CgParams prm{s->cfg.pressure_tolerance, s->cfg.pressure_max_iters, s->cfg.stop_rule,
                       rr0_ref, 0.0};
    fl_status st = run_cg<double, double>(s, prm, s->p, s->r, s->d, s->q);
    if (st != FL_OK) return st;
    const CgStateA &a = s->host->a;
    const CgStateB &b = s->host->b;
    const int status = b.status != ST_OK ? b.status : a.status;
    if (rep) {
        rep->iterations = b.iters;
        rep->refinements = 0;
        rep->fallback = 0;
        rep->warm_start = 0;
        rep->warm_start_ratio = 0.0;
        rep->fallback_reason = FL_FALLBACK_NONE;
        rep->fallback_iterations = 0;
        rep->residual_abs_inf = a.rmax;
        rep->residual_l2 = std::sqrt(a.rr);
        rep->initial_l2 = std::sqrt(a.rr0);
        rep->converged = a.converged;
        rep->status = status == ST_OK ? FL_OK : FL_ERR_SOLVER;
    }
    if (status != ST_OK) return set_error(ctx, FL_ERR_SOLVER, "%s", status_text(status));
    return FL_OK;
}

// One multigrid-preconditioned FCG iteration: the V-cycle on r (levels
// 0..tail-1 here, the rest in k_mg_tail), then the direction and update.
// Algorithmic HBM bytes per launch (FP32 words, level 0's iterates FP16): a
// sweep reads f and e_in and writes e_out (the first sweep of a level reads
// no e_in); prolongation adds the coarse level once, the publishing sweep
// q_old; a restriction reads f and e of the fine level and writes the
// coarse f.
// Levels from kFusedFrom on run their down and up passes as one fused launch
// each (k_mg_down / k_mg_up); finer ones as separate sweep and restriction
// launches.
constexpr int kFusedFrom = 2;
// Damped-Jacobi sweeps on level 0 before and after the coarse correction.
constexpr int kL0Smooth = 3;
// The multigrid level whose coarse solve is the U-Net (FL_PRECOND_LEARNED
// with a network attached at a level below the finest), or 0 for none.
int learned_level(const fl_cavity3d *s) {
    if (s->cfg.preconditioner != FL_PRECOND_LEARNED || !s->unet) return 0;
    return s->wl_level >= 1 && s->wl_level <= s->tail ? s->wl_level : 0;
}

fl_status launch_fcg_iteration(fl_cavity3d *s, const CgParams &prm, long long k) {
    fl_ctx *ctx = s->ctx;
    hipStream_t stm = ctx->stream;
    const int t = s->tail;
    const Level *L = s->lv;
    float *pold = s->d32[k & 1], *pnew = s->d32[(k + 1) & 1];
    const uint32_t G = s->blocks;
    const dim3 zg(s->zg.bx, s->zg.by, s->zg.bz);
    const double n0 = L[0].n, n1 = L[1].n;
    // down: the first two pre-sweeps in one pass (into meb), the third into
    // mea, which the restriction and the up pass read
    const uint32_t crows = k == 0 ? s->rrows : G;
    if (s->ztc.count > 0)
        FL_TIMED(s, K_MG_CHECK_T, k, 6.0 * n0,
                 k_mg_check_t<<<dim3(s->ztc.bx, s->ztc.by, s->ztc.bz), kBlock, 0, stm>>>(L[0], prm, s->r32, s->m0b, s->pb, crows, s->fa, s->fb));
    else
        FL_TIMED(s, K_MG_CHECK, k, 6.0 * n0,
                 k_mg_check_z<<<zg, kZT, 0, stm>>>(L[0], prm, s->r32, s->m0b, s->pb, crows, s->fa, s->fb));
    e16 *pre0 = s->m0b, *post0 = s->m0a;  // level 0's pre-smoothed iterate, and the other buffer
    const bool tiles = s->ztc.count > 0;
    auto sweep0 = [&](const e16 *ein, e16 *eout) -> fl_status {
        if (tiles)
            FL_TIMED(s, K_MG_SWEEP_L0_T, k, 8.0 * n0,
                     k_mg_sweep0_t<false><<<dim3(s->zts0.bx, s->zts0.by, s->zts0.bz), kBlock, 0, stm>>>(L[0], s->r32, ein, eout, nullptr, nullptr, s->fa, s->fb));
        else
            FL_TIMED(s, K_MG_SWEEP_L0, k, 8.0 * n0,
                     k_mg_sweep0_z<false><<<zg, kZT, 0, stm>>>(L[0], s->r32, ein, eout, nullptr, nullptr, s->fa, s->fb));
        return FL_OK;
    };
    if (kL0Smooth == 3) {
        if (fl_status st = sweep0(s->m0b, s->m0a)) return st;
        pre0 = s->m0a;
        post0 = s->m0b;
    }
    FL_TIMED(s, K_MG_RESTRICT_L0, k, 6.0 * n0 + 4.0 * n1,
             k_mg_restrict_z<e16><<<dim3((L[0].nx + kRX - 1) / kRX, L[0].ny / 2, s->zg.bz), kRX, 0, stm>>>(L[0], L[1], s->r32, pre0, s->mf[1], s->fa, s->fb));
    // The learned coarse solve: on the network's level Lw the U-Net predicts
    // x = A_Lw^-1 f_Lw (the level's Galerkin diagonal sets its scale); the
    // cycle goes on below Lw with the residual f_Lw - A x in place of f_Lw,
    // and x is added back to level Lw's result on the way up, so the coarse
    // solve is the multigrid correction of the network's prediction (its
    // error is the cycle's contraction of the network's error).
    const int Lw = learned_level(s);
    auto capture = [&](int l) -> fl_status {
        if (s->cap_level == l && !s->cap_taken && k == s->cap_iter) {
            FL_HIP_TRY(ctx, hipMemcpyAsync(s->cap, s->mf[l], L[l].n * sizeof(float), hipMemcpyDeviceToDevice, stm));
            s->cap_taken = true;
        }
        return FL_OK;
    };
    auto learned = [&]() -> fl_status {
        const Level &LL = L[Lw];
        const double diag = 2.0 * ((double)LL.cx + (double)LL.cy + (double)LL.cz);
        const uint32_t gl = s->lblocks[Lw];
        fl_status ust = FL_OK;
        FL_TIMED(s, K_UNET_COARSE, k, 0.0,
                 ust = unet_infer_device(s->unet, stm, s->mf[Lw], diag, s->wl_out);
                 k_lvl_residual<<<gl, kBlock, 0, stm>>>(LL, s->mf[Lw], s->wl_out, s->fa, s->fb));
        if (ust != FL_OK)
            return set_error(ctx, FL_ERR_HIP, "U-Net inference failed: %s", fl_last_error(unet_context(s->unet)));
        return FL_OK;
    };
    auto learned_add = [&]() {
        k_lvl_add<<<s->lblocks[Lw], kBlock, 0, stm>>>(L[Lw].n, s->meb[Lw], s->wl_out, s->fa, s->fb);
    };
    // FL_LEARNED_REPLACE: levels Lw and below are replaced by the prediction
    // and two damped-Jacobi sweeps on level Lw
    const bool replace = Lw > 0 && s->learned_mode == FL_LEARNED_REPLACE;
    const int down_end = replace ? Lw : t;
    for (int l = 1; l < down_end; ++l) {
        fl_status lst = capture(l);
        if (lst == FL_OK && l == Lw) lst = learned();
        if (lst != FL_OK) return lst;
        const double nl = L[l].n, nc = L[l + 1].n;
        if (l >= kFusedFrom) {
            FL_TIMED(s, K_MG_DOWN, k, 8.0 * nl + 4.0 * nc,
                     k_mg_down<<<ftiles(L[l]), kFB, 0, stm>>>(L[l], L[l + 1], s->mf[l], s->mea[l], s->mf[l + 1], s->fa, s->fb));
        } else {
            const dim3 g(s->zl[l].bx, s->zl[l].by, s->zl[l].bz);
            const dim3 gr((L[l].nx + kRX - 1) / kRX, L[l].ny / 2, s->zl[l].bz);
            FL_TIMED(s, K_MG_DOWN, k, 8.0 * nl + 4.0 * nc,
                     k_mg_sweep_z<true, false><<<g, kZT, 0, stm>>>(L[l], L[l], s->mf[l], nullptr, nullptr, s->meb[l], s->fa, s->fb, kCorr);
                     k_mg_sweep_z<false, false><<<g, kZT, 0, stm>>>(L[l], L[l], s->mf[l], s->meb[l], nullptr, s->mea[l], s->fa, s->fb, kCorr);
                     k_mg_restrict_z<float><<<gr, kRX, 0, stm>>>(L[l], L[l + 1], s->mf[l], s->mea[l], s->mf[l + 1], s->fa, s->fb));
        }
    }
    if (replace) {
        fl_status lst = capture(Lw);
        if (lst != FL_OK) return lst;
        const Level &LL = L[Lw];
        const double diag = 2.0 * ((double)LL.cx + (double)LL.cy + (double)LL.cz);
        const uint32_t gl = s->lblocks[Lw];
        fl_status ust = FL_OK;
        FL_TIMED(s, K_UNET_COARSE, k, 0.0,
                 ust = unet_infer_device(s->unet, stm, s->mf[Lw], diag, s->wl_out);
                 k_mg_sweep<false, false, false, false><<<gl, kBlock, 0, stm>>>(LL, LL, s->mf[Lw], s->wl_out, nullptr, s->wl_in, nullptr, nullptr, s->fa, s->fb, 1.0f);
                 k_mg_sweep<false, false, false, false><<<gl, kBlock, 0, stm>>>(LL, LL, s->mf[Lw], s->wl_in, nullptr, s->meb[Lw], nullptr, nullptr, s->fa, s->fb, 1.0f));
        if (ust != FL_OK)
            return set_error(ctx, FL_ERR_HIP, "U-Net inference failed: %s", fl_last_error(unet_context(s->unet)));
    } else {
        fl_status lst = capture(t);
        if (lst == FL_OK && Lw == t) lst = learned();
        if (lst != FL_OK) return lst;
        TailLevels T{};
        T.count = s->nlev - t;
        for (int l = 0; l < T.count; ++l) T.lv[l] = L[t + l];
        T.pinv = s->pinv;
        FL_TIMED(s, K_MG_TAIL, k, 8.0 * L[t].n,
                 k_mg_tail<<<1, 1024, 0, stm>>>(T, s->mf[t], s->meb[t], s->fa, s->fb));
        if (Lw == t) learned_add();
    }
    // up: the solution of level l ends in meb[l]
    for (int l = down_end - 1; l >= 1; --l) {
        const double nl = L[l].n, nc = L[l + 1].n;
        if (l >= kFusedFrom) {
            FL_TIMED(s, K_MG_UP, k, 12.0 * nl + 4.0 * nc,
                     k_mg_up<<<ftiles(L[l]), kFB, 0, stm>>>(L[l], L[l + 1], s->mf[l], s->mea[l], s->meb[l + 1], s->meb[l], s->fa, s->fb, kCorr));
        } else {
            const dim3 g(s->zl[l].bx, s->zl[l].by, s->zl[l].bz);
            FL_TIMED(s, K_MG_UP, k, 12.0 * nl + 4.0 * nc,
                     k_mg_sweep_z<false, true><<<g, kZT, 0, stm>>>(L[l], L[l + 1], s->mf[l], s->mea[l], s->meb[l + 1], s->meb[l], s->fa, s->fb, kCorr);
                     k_mg_sweep_z<false, false><<<g, kZT, 0, stm>>>(L[l], L[l], s->mf[l], s->meb[l], nullptr, s->mea[l], s->fa, s->fb, kCorr);
                     k_mg_sweep_z<false, false><<<g, kZT, 0, stm>>>(L[l], L[l], s->mf[l], s->mea[l], nullptr, s->meb[l], s->fa, s->fb, kCorr));
        }
        if (!replace && l == Lw) learned_add();
    }
    // up on level 0: prolongate-correct sweep into post0 (kL0Smooth == 2:
    // mea) or into meb (3, then one more sweep into mea); the publishing
    // sweep leaves z~ in meb
    if (tiles)
        FL_TIMED(s, K_MG_PROLONG_L0_T, k, 8.0 * n0 + 4.0 * n1,
                 k_mg_prolong0_t<<<dim3(s->ztp0.bx, s->ztp0.by, s->ztp0.bz), kBlock, 0, stm>>>(L[0], L[1], s->r32, pre0, s->meb[1], post0, s->fa, s->fb, kCorr));
    else
        FL_TIMED(s, K_MG_SWEEP_L0_PROLONG, k, 8.0 * n0 + 4.0 * n1,
                 k_mg_prolong0_z<<<zg, kZT, 0, stm>>>(L[0], L[1], s->r32, pre0, s->meb[1], post0, s->fa, s->fb, kCorr));
    if (kL0Smooth == 3)
        if (fl_status st = sweep0(s->m0b, s->m0a)) return st;
    // the publishing sweep's partial rows, folded by the direction pass
    uint32_t zrows = s->zg.count;
    if (tiles) {
        zrows = s->zts1.count;
        FL_TIMED(s, K_MG_PUBLISH_L0_T, k, 12.0 * n0,
                 k_mg_sweep0_t<true><<<dim3(s->zts1.bx, s->zts1.by, s->zts1.bz), kBlock, 0, stm>>>(L[0], s->r32, s->m0a, s->m0b, s->q32, s->pz, s->fa, s->fb));
    } else {
        FL_TIMED(s, K_MG_SWEEP_L0_PUBLISH, k, 12.0 * n0,
                 k_mg_sweep0_z<true><<<zg, kZT, 0, stm>>>(L[0], s->r32, s->m0a, s->m0b, s->q32, s->pz, s->fa, s->fb));
    }
    // FCG direction (reads z~, p_old, r; writes p, q) and update (x, r
    // read and written, p and q read)
    uint32_t drows = s->zgd.count;
    if (s->ztd.count > 0) {
        drows = s->ztd.count;
        FL_TIMED(s, K_FCG_DIRAPPLY_T, k, 18.0 * n0,
                 k_fcg_dirapply_t<<<dim3(s->ztd.bx, s->ztd.by, s->ztd.bz), kBlock, 0, stm>>>(s->g, s->m0b, pold, pnew, s->q32, s->r32, s->pz, zrows, s->pa, s->fa, s->fb));
    } else {
        FL_TIMED(s, K_FCG_DIRAPPLY, k, 18.0 * n0,
                 k_fcg_dirapply_z<<<dim3(s->zgd.bx, s->zgd.by, s->zgd.bz), kZT, 0, stm>>>(s->g, s->m0b, pold, pnew, s->q32, s->r32, s->pz, zrows, s->pa, s->fa, s->fb));
    }
    // bytes per cell: r, q read and r written; with the x update (k odd)
    // also p and p_prev read and x written (and read, after the first pair)
    const double ubytes = (k & 1) == 0 ? 12.0 : (k == 1 ? 24.0 : 28.0);
    FL_TIMED(s, K_FCG_UPDATE, k, ubytes * n0,
             k_fcg_update<<<G, kBlock, 0, stm>>>((uint32_t)s->n, s->e32, s->r32, pnew, pold, s->q32, s->pa, drows, s->pb, s->fa, s->fb, k));
    return FL_OK;
}

// Inner FCG on A e = r32 (e32 = 0 on entry, r32 and its partials in pb from
// k_residual_z, zgd.count rows).  Returns with the device state in host->fa / host->fb;
// *stagnated is set when the residual fails to halve over a whole batch.
// `hint` is the update count this solve needed last time (0: unknown).
fl_status run_fcg(fl_cavity3d *s, const CgParams &prm, int hint, bool *stagnated) {
    fl_ctx *ctx = s->ctx;
    hipStream_t stm = ctx->stream;
    s->host->fa = FcgA{};
    s->host->fb = FcgB{};
    s->host->fb.first = 1;
    k_fcg_reset<<<1, 1, 0, stm>>>(s->fa, s->fb);  // the same state, without a copy
    *stagnated = false;
    // A batch of iterations is followed by the stop test alone (k_fcg_check),
    // so after launching iterations 0..L-1 the host knows rr_L, the residual
    // after L updates.  The first batch is the hint (a solve in a time loop
    // needs about what it needed last step), later ones are sized from the
    // observed contraction, so few iterations land after the device-side
    // stop test.
    long long launched = 0, at_prev = 0;
    int batch = hint > 0 ? std::min(hint, 16) : 4;
    double rr_prev = -1.0;
    for (;;) {
        for (int b = 0; b < batch; ++b) {
            fl_status st = launch_fcg_iteration(s, prm, launched++);
            if (st != FL_OK) return st;
        }
        k_fcg_check<<<1, kBlock, 0, stm>>>(prm, s->pb, s->blocks, s->fa, s->fb);
        FL_HIP_TRY(ctx, hipGetLastError());
        FL_HIP_TRY(ctx, hipMemcpyAsync(&s->host->fa, s->fa, sizeof(FcgA) + sizeof(FcgB), hipMemcpyDeviceToHost, stm));
        FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
        FL_HIP_TRY(ctx, s->prof.flush(s->host->fb.iters));
        if (s->host->fa.done || s->host->fb.done) break;
        const double rr = s->host->fa.rr, rr0 = s->host->fa.rr0;
        const long long at = launched;
        if (rr_prev < 0.0) rr_prev = rr0;
        const long long span = at - at_prev;
        const double f = (span > 0 && rr_prev > 0.0) ? std::pow(rr / rr_prev, 1.0 / (double)span) : 0.5;
        // stagnation: rr fails to halve over 4 iterations (0.84^4 = 0.5)
        if (span >= 4 && !(f < 0.84)) {
            *stagnated = true;
            break;
        }
        long long need = 16;
        const double target = prm.tol * prm.tol * rr0;
        if (f > 0.0 && f < 1.0 && rr > 0.0 && target > 0.0)
            need = (long long)std::ceil(std::log(target / rr) / std::log(f));
        batch = (int)std::min<long long>(std::max<long long>(need, 1), 16);
        rr_prev = rr;
        at_prev = at;
    }
    return FL_OK;
}

// Learned warm start from the FP32 residual r0 in r32 (sum_r = sum of r0):
// the upward half of a V-cycle whose coarse solve is the U-Net.  r0 is
// mean-restricted to the network's level L, the network predicts the
// correction there, and levels L-1 .. 0 each prolongate the coarser
// correction and smooth it (damped Jacobi against the restricted residual;
// level 0 with the pinned row replaced as in the multigrid preconditioner).
// The correction e is then added to p with the energy-minimising step
// alpha = r0.e~ / e~.A e~ (e~ = e - e_0, the pinned operator), so that the
// A-norm of the error cannot grow.  The diagonal handed to the network is
// the rediscretised Laplacian's at its resolution, as in training.
fl_status warm_start(fl_cavity3d *s, double sum_r) {
    fl_ctx *ctx = s->ctx;
    hipStream_t stm = ctx->stream;
    const int L = s->wl_level;
    const Level *lv = s->lv;
    const float *src = s->r32;
    for (int l = 0; l < L; ++l) {
        float *dst = l + 1 == L ? s->wl_in : s->mf[l + 1];
        k_restrict_mean<<<s->lblocks[l + 1], kBlock, 0, stm>>>(lv[l], lv[l + 1], src, dst);
        src = dst;
    }
    FL_HIP_TRY(ctx, hipGetLastError());
    const double f = (double)(1u << L);
    const double diag = 2.0 * (s->g.ix2 + s->g.iy2 + s->g.iz2) / (f * f);
    if (unet_infer_device(s->unet, stm, src, diag, s->wl_out) != FL_OK)
        return set_error(ctx, FL_ERR_HIP, "U-Net inference failed: %s",
                         fl_last_error(unet_context(s->unet)));
    // multigrid state for the level-0 sweeps: running, r_0 -> -sum_{k!=0} r_k
    s->host->fa = FcgA{};
    s->host->fa.sumr = sum_r;
    s->host->fb = FcgB{};
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->fa, &s->host->fa, sizeof(FcgA), hipMemcpyHostToDevice, stm));
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->fb, &s->host->fb, sizeof(FcgB), hipMemcpyHostToDevice, stm));
    const float *ec = s->wl_out;
    for (int l = L - 1; l >= 1; --l) {
        const uint32_t g = s->lblocks[l];
        k_mg_sweep<false, true, false, false><<<g, kBlock, 0, stm>>>(lv[l], lv[l + 1], s->mf[l], nullptr, ec, s->mea[l], nullptr, nullptr, s->fa, s->fb, 1.0f);
        k_mg_sweep<false, false, false, false><<<g, kBlock, 0, stm>>>(lv[l], lv[l], s->mf[l], s->mea[l], nullptr, s->meb[l], nullptr, nullptr, s->fa, s->fb, 1.0f);
        ec = s->meb[l];
    }
    // level 0 in FP32 scratch: the inner solve's direction buffers, free until
    // it starts (e0 must outlive the residual check that may undo it)
    float *e0 = s->d32[1], *e1 = s->d32[0];
    if (L >= 1)
        k_mg_sweep<false, true, true, false><<<s->blocks, kBlock, 0, stm>>>(lv[0], lv[1], s->r32, nullptr, ec, e1, nullptr, nullptr, s->fa, s->fb, 1.0f);
    else
        k_mg_sweep<false, false, true, false><<<s->blocks, kBlock, 0, stm>>>(lv[0], lv[0], s->r32, ec, nullptr, e1, nullptr, nullptr, s->fa, s->fb, 1.0f);
    k_mg_sweep<false, false, true, false><<<s->blocks, kBlock, 0, stm>>>(lv[0], lv[0], s->r32, e1, nullptr, e0, nullptr, nullptr, s->fa, s->fb, 1.0f);
    // alpha = r0.e~ / e~.A e~
    k_warm_dots<<<s->blocks, kBlock, 0, stm>>>(s->g, s->r32, e0, s->pb);
    k_fold_outer<<<1, kBlock, 0, stm>>>(s->pb, s->blocks, s->outer);
    FL_HIP_TRY(ctx, hipGetLastError());
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->host->outer, s->outer, 4 * sizeof(double), hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
    const double re = s->host->outer[0], eae = s->host->outer[1];
    const double alpha = (eae > 0.0 && std::isfinite(re / eae)) ? re / eae : 0.0;
    s->wl_alpha = alpha;
    s->wl_e = e0;
    k_add_shifted<<<s->ew_blocks, kBlock, 0, stm>>>((uint32_t)s->n, s->p, e0, alpha);
    FL_HIP_TRY(ctx, hipGetLastError());
    return FL_OK;
}

// FP32 CG (Jacobi or multigrid-preconditioned flexible) inside FP64
// iterative refinement; multigrid falls back to plain FP64 Jacobi-PCG from
// the current iterate when an inner solve stagnates or exhausts its budget.
fl_status solve_mixed(fl_cavity3d *s, fl_solve_report *rep) {
    fl_ctx *ctx = s->ctx;
    hipStream_t stm = ctx->stream;
    const bool mg = s->cfg.preconditioner != FL_PRECOND_JACOBI;
    const double inner_tol =
        s->cfg.inner_tolerance > 0.0 ? s->cfg.inner_tolerance : (mg ? 1.0e-4 : 1.0e-5);
    const bool rel = s->cfg.stop_rule == FL_STOP_REL_L2;
    double rr0 = 0.0;
    long long total = 0, sweeps = 0;
    bool converged = false, fallback = false, first_pass = true;
    int reason = FL_FALLBACK_NONE;
    double rr = 0.0, rmax = 0.0;
    // learned warm start: 1 pending, 2 applied and awaiting its residual;
    // warm_used ends 1 (kept) or -1 (rejected: it did not lower the residual).
    // With the network inside the V-cycle (learned_level > 0) the first FCG
    // step, alpha z_0 with z_0 from the learned cycle and the energy-
    // minimising alpha, is that warm start, so none is taken separately.
    int warm = (s->cfg.preconditioner == FL_PRECOND_LEARNED && s->unet && learned_level(s) == 0) ? 1 : 0,
        warm_used = 0;
    double warm_ratio = 0.0;
    for (;;) {
        if (s->r_ready) {  // r = b in r32 and its rows in pb, from the right-hand side kernel
            s->r_ready = false;
            s->p_zero = false;
            s->rrows = s->ew_blocks;
        } else {
            FL_TIMED(s, K_RESIDUAL_F, -1, residual_bytes<float>(s),
                     launch_residual<float>(s, s->r32));
        }
        FL_TIMED(s, K_FOLD_OUTER, -1, 32.0 * s->rrows,
                 k_fold_outer<<<1, kBlock, 0, stm>>>(s->pb, s->rrows, s->outer));
        FL_HIP_TRY(ctx, hipGetLastError());
        FL_HIP_TRY(ctx, hipMemcpyAsync(s->host->outer, s->outer, 4 * sizeof(double), hipMemcpyDeviceToHost, stm));
        FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
        FL_HIP_TRY(ctx, s->prof.flush());
        rr = s->host->outer[1];
        rmax = s->host->outer[2];
        if (!mg) {
            // The eliminated row 0 of the Neumann system has the residual
            // -sum_{k!=0} r_k (the columns of A sum to zero and the closed
            // cavity's right-hand side telescopes).  An FP32 Jacobi-CG leaves
            // error along the pinned operator's lowest mode, whose residual is
            // spread almost evenly over all cells: small in ||r||_2, yet
            // sqrt(n) times larger in that sum, which is the divergence left
            // at cell (0,0,0).  The stop test is therefore taken on the full
            // Neumann residual (rows k != 0 and row 0).  The multigrid path
            // hands row 0 to its Neumann cycle in compatible form instead.
            const double r0 = s->host->outer[3];
            rr += r0 * r0;
            rmax = std::max(rmax, std::fabs(r0));
        }
        if (first_pass) rr0 = rr;
        first_pass = false;
        if (!std::isfinite(rmax))
            return set_error(ctx, FL_ERR_SOLVER, "%s", status_text(ST_RESIDUAL));
        if (warm == 2) {
            warm = 0;
            warm_ratio = rr0 > 0.0 ? std::sqrt(rr / rr0) : 0.0;
            if (!(rr < rr0)) {  // undo the prediction and restart from the guess
                k_add_shifted<<<s->ew_blocks, kBlock, 0, stm>>>((uint32_t)s->n, s->p, s->wl_e, -s->wl_alpha);
                FL_HIP_TRY(ctx, hipGetLastError());
                warm_used = -1;
                continue;
            }
            warm_used = 1;
        }
        converged = rel ? std::sqrt(rr) <= s->cfg.pressure_tolerance * std::sqrt(rr0)
                        : rmax < s->cfg.pressure_tolerance;
        if (converged || total >= s->cfg.pressure_max_iters || sweeps >= 64) break;
        if (warm == 1) {
            fl_status wst = FL_OK;
            FL_TIMED(s, K_WARM_START, -1, 28.0 * s->n, wst = warm_start(s, s->host->outer[3]));
            if (wst != FL_OK) return wst;
            warm = 2;
            continue;
        }
        // the flexible CG writes its solution at its second update (no fill)
        if (!mg) FL_HIP_TRY(ctx, hipMemsetAsync(s->e32, 0, s->n * sizeof(float), stm));
        long long it = 0;
        int status = ST_OK;
        bool budget_hit = false, stagnated = false;
        // With the default inner tolerance under the rel-L2 rule an inner
        // solve need not go below half of the reduction still missing: the
        // last sweep stops near the target instead of overshooting it by up
        // to the inner tolerance.  An explicit inner tolerance is kept as is.
        double itol = inner_tol;
        if (rel && rr > 0.0 && !(s->cfg.inner_tolerance > 0.0))
            itol = std::max(inner_tol, 0.5 * s->cfg.pressure_tolerance * std::sqrt(rr0 / rr));
        if (mg) {
            const long long budget = std::min<long long>(s->cfg.pressure_max_iters - total, 100);
            const CgParams prm{itol, budget, FL_STOP_REL_L2, 0.0, s->e_bound};
            const int slot = (int)std::min<long long>(sweeps, kFcgHints - 1);
            fl_status st = run_fcg(s, prm, s->fcg_hint[slot], &stagnated);
            if (st != FL_OK) return st;
            const FcgB &b = s->host->fb;
            status = b.status != ST_OK ? b.status : s->host->fa.status;
            it = b.iters;
            s->fcg_hint[slot] = (int)std::min<long long>(it, 16);
            budget_hit = !s->host->fa.converged && !stagnated && status == ST_OK;
        } else {
            // the inner Jacobi-CG solves the compatible Neumann form of the
            // pinned correction equation (row 0 with r_0 = -sum_{k!=0} r_k);
            // its solution, shifted to e_0 = 0, is the pinned correction, and
            // the Neumann operator lacks the pinned operator's near-null mode
            k_row0_compat<<<1, 64, 0, stm>>>(s->g, s->outer, s->r32, s->pb + 4 * (size_t)s->rrows);
            const CgParams prm{itol, s->cfg.pressure_max_iters - total, FL_STOP_REL_L2, 0.0, 0.0, 1};
            fl_status st = run_cg<float, float>(s, prm, s->e32, s->r32, s->d32, s->q32, s->rrows + 1);
            if (st != FL_OK) return st;
            const CgStateB &b = s->host->b;
            status = b.status != ST_OK ? b.status : s->host->a.status;
            it = b.iters;
        }
        if (status != ST_OK && !mg) return set_error(ctx, FL_ERR_SOLVER, "%s", status_text(status));
        total += it;
        sweeps += 1;
        if (status == ST_OK && it > 0) {
            // flexible CG: e holds updates 0 .. 2 floor(it / 2) - 1; an odd
            // count leaves alpha_even p_{it-1} (in d32[it & 1]) pending
            const bool pend = mg && (it & 1);
            const int e_valid = mg ? (it >= 2) : 1;
            FL_TIMED(s, K_ADD_CORRECTION, -1, ((pend ? 16.0 : 12.0) + (s->p_stale ? 0.0 : 8.0)) * s->n,
                     k_add_correction<<<s->ew_blocks, kBlock, 0, stm>>>((uint32_t)s->n, s->p, s->e32, mg ? 0 : 1, e_valid,
                                                                        pend ? s->d32[it & 1] : nullptr,
                                                                        s->host->fb.alpha_even, s->p_stale ? 1 : 0));
            s->p_stale = false;
        }
        if (mg && (status != ST_OK || stagnated || budget_hit)) {
            fallback = true;
            reason = status != ST_OK ? FL_FALLBACK_BREAKDOWN
                                     : (stagnated ? FL_FALLBACK_STAGNATION : FL_FALLBACK_BUDGET);
            break;
        }
        if (it == 0) break;  // no progress possible in FP32
    }
    if (s->p_stale) {  // no correction was added: the solution is the zero guess
        FL_HIP_TRY(ctx, hipMemsetAsync(s->p, 0, s->n * sizeof(double), stm));
        s->p_stale = false;
    }
    fl_solve_report fb{};
    const bool fallback_runs = fallback && total < s->cfg.pressure_max_iters;
    if (fallback_runs) {
        // the judge: plain FP64 Jacobi-PCG from the current iterate
        fl_status st = solve_fp64(s, &fb, rr0);
        if (st != FL_OK) return st;
        converged = fb.converged;
        rmax = fb.residual_abs_inf;
        rr = fb.residual_l2 * fb.residual_l2;
    }
    if (rep) {
        rep->iterations = total;
        rep->refinements = sweeps;
        rep->fallback = fallback_runs ? 1 : 0;
        rep->fallback_reason = fallback ? reason : FL_FALLBACK_NONE;
        rep->warm_start = warm_used;
        rep->warm_start_ratio = warm_ratio;
        rep->fallback_iterations = fb.iterations;
        rep->residual_abs_inf = rmax;
        rep->residual_l2 = std::sqrt(rr);
        rep->initial_l2 = std::sqrt(rr0);
        rep->converged = converged;
        rep->status = FL_OK;
    }
    return FL_OK;
}

// SOR to max|Lap p - rhs| < tol (the reference's solver and stop rule),
// batches of sweeps with the verdict kept on the device.
fl_status solve_sor(fl_cavity3d *s, fl_solve_report *rep) {
    fl_ctx *ctx = s->ctx;
    hipStream_t stm = ctx->stream;
    s->p_zero = false;  // the sweeps write p before each residual
    const SorState zero{};
    FL_HIP_TRY(ctx, hipMemcpyAsync(s->sor, &zero, sizeof(SorState), hipMemcpyHostToDevice, stm));
    const uint32_t half = ((s->g.nx + 1) / 2) * s->g.ny * s->g.nz;
    const uint32_t G = grid_blocks(ctx, (half + kBlock - 1) / kBlock, 8);
    SorState h{};
    int batch = 4;
    for (;;) {
        for (int b = 0; b < batch; ++b) {
            k_sor3<<<G, kBlock, 0, stm>>>(s->g, s->rhs, s->p, s->cfg.pressure_omega, 0u, s->sor);
            k_sor3<<<G, kBlock, 0, stm>>>(s->g, s->rhs, s->p, s->cfg.pressure_omega, 1u, s->sor);
            launch_residual<double>(s, s->r);
            k_fold_outer<<<1, kBlock, 0, stm>>>(s->pb, s->zr.count, s->outer);
            k_sor_check<<<1, 64, 0, stm>>>(s->outer, s->cfg.pressure_tolerance, s->cfg.pressure_max_iters, s->sor);
        }
        FL_HIP_TRY(ctx, hipGetLastError());
        FL_HIP_TRY(ctx, hipMemcpyAsync(&h, s->sor, sizeof(SorState), hipMemcpyDeviceToHost, stm));
        FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
        if (h.done) break;
        batch = std::min(batch * 2, kMaxBatch);
    }
    if (h.status != ST_OK)
        return set_error(ctx, FL_ERR_SOLVER, "3D pressure solver produced a non-finite residual");
    if (rep) {
        *rep = fl_solve_report{};
        rep->iterations = h.iters;
        rep->residual_abs_inf = h.res;
        rep->converged = h.converged;
        rep->status = FL_OK;
    }
    return FL_OK;
}

// Pressure solve: p holds the guess on entry and the solution on exit; rhs
// is the device right-hand side.
fl_status solve_pressure(fl_cavity3d *s, fl_solve_report *rep) {
    fl_ctx *ctx = s->ctx;
    hipStream_t stm = ctx->stream;
    const bool zero = s->cfg.initial_guess == FL_GUESS_ZERO;
    // the mixed multigrid path reads p only after its first correction
    // (which writes it), unless a learned warm start adds to p first
    const bool learned_warm = s->cfg.preconditioner == FL_PRECOND_LEARNED && s->unet && learned_level(s) == 0;
    s->p_stale = zero && s->cfg.pressure_solver != FL_SOLVER_SOR && s->cfg.precision == FL_PRECISION_MIXED &&
                 s->cfg.preconditioner != FL_PRECOND_JACOBI && !learned_warm;
    if (zero && !s->p_stale)
        FL_HIP_TRY(ctx, hipMemsetAsync(s->p, 0, s->n * sizeof(double), stm));
    else if (!zero)
        FL_HIP_TRY(ctx, hipMemsetAsync(s->p, 0, sizeof(double), stm));  // pin p(0,0,0) = 0
    s->cap_taken = false;
    s->p_zero = s->cfg.initial_guess == FL_GUESS_ZERO;
    if (s->cfg.pressure_solver == FL_SOLVER_SOR) return solve_sor(s, rep);
    return s->cfg.precision == FL_PRECISION_MIXED ? solve_mixed(s, rep) : solve_fp64(s, rep);
}

// The correction and the step's diagnostics (one fused launch + the fold),
// then the finite check's verdict.
fl_status correct_and_diag(fl_cavity3d *s, fl_cavity3d_step_report *rep) {
    fl_ctx *ctx = s->ctx;
    hipStream_t stm = ctx->stream;
    k_diag_reset<<<1, 1, 0, stm>>>(s->dg);  // (a kernel, not a host-to-device copy)
    // algorithmic bytes: us, vs, ws read, u, v, w written, p read
    const double faces = 8.0 * (double)(s->nu + s->nv + s->nw), cells = 8.0 * s->n;
    const dim3 zg(s->zg.bx, s->zg.by, s->zg.bz);
    FL_TIMED(s, K_CORRECT, -1, 2.0 * faces + cells,
             k_correct_diag_t<<<dim3(s->zct.bx, s->zct.by, s->zct.bz), kBlock, 0, stm>>>(s->P, s->p, s->us, s->vs, s->ws, s->u, s->v, s->w, s->partials, s->dg);
             k_diag_fold<<<1, kBlock, 0, stm>>>(s->partials, s->zct.count, s->dg));
    FL_HIP_TRY(ctx, hipGetLastError());
    FL_HIP_TRY(ctx, hipMemcpyAsync(&s->host->dg, s->dg, sizeof(DiagState), hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
    FL_HIP_TRY(ctx, s->prof.flush());
    const DiagState &h = s->host->dg;
    if (rep) {
        rep->max_divergence = h.max_div;
        rep->divergence_l2 = std::sqrt(h.div_l2sq);
        rep->max_speed = h.max_speed;
    }
    {
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
    // algorithmic bytes: every face array read or written once, cells once
    const double faces = 8.0 * (double)(s->nu + s->nv + s->nw), cells = 8.0 * s->n;
    FL_TIMED(s, K_PREDICT, -1, 2.0 * faces,
             k_predict3_t<<<dim3(s->zpt.bx, s->zpt.by, s->zpt.bz), kBlock, 0, stm>>>(s->P, s->u, s->v, s->w, s->us, s->vs, s->ws));
    // the zero guess's FP32 residual with the right-hand side (see k_rhs)
    s->r_ready = s->cfg.initial_guess == FL_GUESS_ZERO && s->cfg.precision == FL_PRECISION_MIXED &&
                 s->cfg.pressure_solver != FL_SOLVER_SOR;
    if (s->r_ready)
        FL_TIMED(s, K_RHS, -1, faces + 1.5 * cells,
                 k_rhs<true><<<s->ew_blocks, kBlock, 0, stm>>>(s->P, s->g, s->us, s->vs, s->ws, s->rhs, s->r32, s->pb));
    else
        FL_TIMED(s, K_RHS, -1, faces + cells,
                 k_rhs<false><<<s->ew_blocks, kBlock, 0, stm>>>(s->P, s->g, s->us, s->vs, s->ws, s->rhs, nullptr, nullptr));
    FL_HIP_TRY(ctx, hipGetLastError());
    fl_solve_report sr{};
    fl_status st = solve_pressure(s, &sr);
    if (rep) {
        rep->pressure_iterations = sr.iterations;
        rep->pressure_refinements = sr.refinements;
        rep->pressure_residual = sr.residual_abs_inf;
        rep->pressure_residual_l2 = sr.residual_l2;
        rep->pressure_converged = sr.converged;
        rep->pressure_fallback = sr.fallback;
        rep->pressure_fallback_reason = sr.fallback_reason;
        rep->pressure_warm_start = sr.warm_start;
        rep->pressure_warm_ratio = sr.warm_start_ratio;
    }
    if (st != FL_OK) return st;
    s->step += 1;
    st = correct_and_diag(s, rep);
    if (rep) {
        rep->step = s->step;
        rep->time = (double)s->step * s->cfg.dt;
    }
    return st;
}

// Launch of a plane-tile pass with at most `per_cu` blocks per CU (0: as
// many as the kernel's registers and LDS let one CU hold), all resident at
// once: the z chunks are as long as needed, at least 8 planes.  A pass's
// block marches its chunk bound by the latency of the next plane's loads, so
// one round of resident blocks beats a second, partly idle one (measured at
// 256^3: FP64 dot pass 87 -> 81 us against 16-plane chunks).  The FP64 step
// pass is fastest at 2 blocks per CU (64-plane chunks at 256^3: 146 -> 140
// us), the FP32 passes at full occupancy.
template <typename T>
ZGrid tile_grid_resident(fl_ctx *ctx, const void *kernel, uint32_t nx, uint32_t ny, uint32_t nz,
                         int per_cu = 0) {
    ZGrid z = tile_grid<T>(nx, ny, nz, nz);
    int fit = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, kernel, kBlock, 0) != hipSuccess || fit < 1) fit = 1;
    if (per_cu > 0) fit = std::min(fit, per_cu);
    const uint64_t slots = (uint64_t)fit * (uint64_t)std::max(ctx->num_cu, 1);
    const uint64_t layer = (uint64_t)z.bx * z.by;
    uint32_t bz = (uint32_t)std::max<uint64_t>(1, slots / layer);
    bz = std::min(bz, std::max(1u, nz / 8));
    return tile_grid<T>(nx, ny, nz, (nz + bz - 1) / bz);
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
    s->prof.init(kKindNames, K_COUNT);
    const uint32_t nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    const double dx = cfg->length / nx, dy = cfg->height / ny, dz = cfg->depth / nz;
    s->g = Grid{nx, ny, nz, nx * ny, nx * ny * nz, 1.0 / (dx * dx), 1.0 / (dy * dy),
                1.0 / (dz * dz), FastDiv(nx), FastDiv(nx * ny), {}, {}};
    for (uint32_t b = 0; b < 8; ++b) {
        // neighbours per axis: 1 + bit; summed in x, y, z order
        const double diag = s->g.ix2 * (double)(1 + (b & 1)) + s->g.iy2 * (double)(1 + (b >> 1 & 1)) +
                            s->g.iz2 * (double)(1 + (b >> 2 & 1));
        s->g.inv[b] = 1.0 / diag;
        s->g.invf[b] = (float)(1.0 / diag);
    }
    s->P = StepParams{nx, ny, nz, dx, dy, dz, cfg->dt,
                      cfg->lid_velocity * cfg->length / cfg->reynolds,  // viscosity :86-88
                      cfg->lid_velocity, cfg->density};
    s->n = (size_t)nx * ny * nz;
    s->nu = (size_t)(nx + 1) * ny * nz;
    s->nv = (size_t)nx * (ny + 1) * nz;
    s->nw = (size_t)nx * ny * (nz + 1);
    const uint64_t ntiles = (s->n + kBlock - 1) / kBlock;
    s->blocks = grid_blocks(ctx, ntiles, 4);
    s->ew_blocks = grid_blocks(ctx, (s->nu + s->nv + s->nw + kBlock - 1) / kBlock);
    const bool mixed = cfg->precision == FL_PRECISION_MIXED;
    s->zg = zgrid(nx, ny, nz, kZK);
    s->zgd = zgrid(nx, ny, nz, kZKD);
    s->zr = s->zgd;
    s->zp = s->zgd;
    if (nx % 2 == 0) {
        s->zr.bx = (nx / 2 + kZX - 1) / kZX;
        s->zr.count = s->zr.bx * s->zr.by * s->zr.bz;
        s->zp = s->zg;
        s->zp.bx = s->zr.bx;
        s->zp.count = s->zp.bx * s->zp.by * s->zp.bz;
    }
    if (nx % 2 == 0) {
        s->zt = tile_grid_resident<double>(ctx, (const void *)k_cg_dot_t<double>, nx, ny, nz);
        s->zts = tile_grid_resident<double>(ctx, (const void *)k_cg_step_t<double>, nx, ny, nz, 2);
    }
    if (nx % 4 == 0) {
        s->ztf = tile_grid_resident<float>(ctx, (const void *)k_cg_dot_t<float>, nx, ny, nz);
        s->ztfs = tile_grid_resident<float>(ctx, (const void *)k_cg_step_t<float>, nx, ny, nz);
    }
    // the FCG direction's tile pass (16-byte float and 8-byte FP16 lane loads)
    {  // plane-tile predictor: 64 x 4 tiles, one round of resident blocks
        int fit = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, (const void *)k_predict3_t, kBlock, 0) != hipSuccess ||
            fit < 1)
            fit = 1;
        const uint32_t bx = (nx + kPTX - 1) / kPTX, by = (ny + kPTY - 1) / kPTY;
        const uint64_t slots = (uint64_t)fit * (uint64_t)std::max(ctx->num_cu, 1);
        uint32_t bz = (uint32_t)std::max<uint64_t>(1, slots / ((uint64_t)bx * by));
        bz = std::min(bz, std::max(1u, nz / 4));
        const uint32_t kz = (nz + bz - 1) / bz;
        bz = (nz + kz - 1) / kz;
        s->zpt = ZGrid{bx, by, bz, bx * by * bz};
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, (const void *)k_correct_diag_t, kBlock, 0) != hipSuccess ||
            fit < 1)
            fit = 1;
        const uint64_t cslots = (uint64_t)fit * (uint64_t)std::max(ctx->num_cu, 1);
        uint32_t cz = (uint32_t)std::max<uint64_t>(1, cslots / ((uint64_t)bx * by));
        cz = std::min(cz, std::max(1u, nz / 4));
        const uint32_t ckz = (nz + cz - 1) / cz;
        cz = (nz + ckz - 1) / ckz;
        s->zct = ZGrid{bx, by, cz, bx * by * cz};
    }
    if (nx % 4 == 0 && cfg->preconditioner != FL_PRECOND_JACOBI) {
        s->ztd = tile_grid_resident<float>(ctx, (const void *)k_fcg_dirapply_t, nx, ny, nz);
        s->ztc = tile_grid_resident<float>(ctx, (const void *)k_mg_check_t, nx, ny, nz);
        s->zts0 = tile_grid_resident<float>(ctx, (const void *)k_mg_sweep0_t<false>, nx, ny, nz);
        s->ztp0 = tile_grid_resident<float>(ctx, (const void *)k_mg_prolong0_t, nx, ny, nz);
        s->zts1 = tile_grid_resident<float>(ctx, (const void *)k_mg_sweep0_t<true>, nx, ny, nz);
    }
    const size_t prow =
        (size_t)std::max(std::max(std::max(std::max(s->blocks, s->ew_blocks), std::max(s->zg.count, s->zgd.count)),
                                  std::max(s->zpt.count, s->zct.count)),
                         std::max(std::max(std::max(s->zt.count, s->ztf.count), std::max(s->zts.count, s->ztfs.count)),
                                  std::max(s->ztd.count, s->zts1.count))) * 4;
    if ((st = alloc(s, &s->u, s->nu)) || (st = alloc(s, &s->v, s->nv)) ||
        (st = alloc(s, &s->w, s->nw)) || (st = alloc(s, &s->us, s->nu)) ||
        (st = alloc(s, &s->vs, s->nv)) || (st = alloc(s, &s->ws, s->nw)) ||
        (st = alloc(s, &s->p, s->n)) || (st = alloc(s, &s->rhs, s->n)) ||
        (st = alloc(s, &s->partials, prow)) || (st = alloc(s, &s->pa, prow)) ||
        (st = alloc(s, &s->pb, prow + 4)) || (st = alloc(s, &s->outer, 4)) ||
        (st = alloc(s, &s->sa, 1)) || (st = alloc(s, &s->sb, 1)) || (st = alloc(s, &s->dg, 1)) ||
        (st = alloc(s, &s->sor, 1)))
        goto fail;
    if (mixed) {
        if ((st = alloc(s, &s->e32, s->n)) || (st = alloc(s, &s->r32, s->n)) ||
            (st = alloc(s, &s->d32[0], s->n)) || (st = alloc(s, &s->d32[1], s->n)) ||
            (st = alloc(s, &s->q32, s->n)))
            goto fail;
    }
    if (!mixed || cfg->preconditioner != FL_PRECOND_JACOBI) {
        // FP64 PCG vectors (the multigrid path keeps them for its fallback)
        if ((st = alloc(s, &s->r, s->n)) || (st = alloc(s, &s->d[0], s->n)) ||
            (st = alloc(s, &s->d[1], s->n)) || (st = alloc(s, &s->q, s->n)))
            goto fail;
    }
    if (cfg->preconditioner != FL_PRECOND_JACOBI) {
        if ((st = build_hierarchy(s)) != FL_OK) goto fail;
        if ((st = alloc(s, &s->pz, prow)) || (st = alloc(s, &s->fa, 2)))
            goto fail;
        s->fb = reinterpret_cast<FcgB *>(s->fa + 1);  // sizeof(FcgA) keeps FcgB's alignment
        for (int l = 0; l <= s->tail; ++l) {
            const size_t n = s->lv[l].n;
            if (l == 0) {
                if ((st = alloc(s, &s->m0a, n)) || (st = alloc(s, &s->m0b, n))) goto fail;
                continue;
            }
            if ((st = alloc(s, &s->mf[l], n))) goto fail;
            if (l < s->tail && (st = alloc(s, &s->mea[l], n))) goto fail;
            if ((st = alloc(s, &s->meb[l], n))) goto fail;
        }
        if (s->lv[s->nlev - 1].n <= kDirectCells) {
            const std::vector<float> P = coarse_pseudo_inverse(s->lv[s->nlev - 1]);
            if ((st = alloc(s, &s->pinv, P.size()))) goto fail;
            const hipError_t e = hipMemcpyAsync(s->pinv, P.data(), P.size() * sizeof(float),
                                                hipMemcpyHostToDevice, ctx->stream);
            // the host vector P must stay alive until the copy has finished
            const hipError_t e2 = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess || e2 != hipSuccess) {
                st = set_error(ctx, FL_ERR_HIP, "pseudo-inverse upload failed: %s",
                               hipGetErrorString(e != hipSuccess ? e : e2));
                goto fail;
            }
        }
    }
    if (hipHostMalloc(&s->host, sizeof(fl_cavity3d::Host), hipHostMallocDefault) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        st = set_error(ctx, FL_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
        goto fail;
    }
    ctx_retain(ctx);
    *out = s;
    return FL_OK;
fail:
    free_all(s);
    delete s;
    return st;
}

void fl_cavity3d_destroy(fl_cavity3d *s) {
    if (!s) return;
    fl_ctx *ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_all(s);
    delete s;
    ctx_release(ctx);
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
    FL_HIP_TRY(ctx, s->prof.flush());
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

fl_status fl_cavity3d_cell_fields(fl_cavity3d *s, double *p, double *uc, double *vc, double *wc) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = s->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t stm = ctx->stream;
    if (uc || vc || wc) {
        if (!s->cell_out) {  // staging for the three components, kept for later snapshots
            if (hipMalloc(&s->cell_out, 3 * s->n * sizeof(double)) != hipSuccess) {
                s->cell_out = nullptr;
                return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "cannot allocate the output staging buffer");
            }
        }
        k_cell_velocity<<<s->blocks, kBlock, 0, stm>>>(s->P, s->u, s->v, s->w, s->cell_out);
        FL_HIP_TRY(ctx, hipGetLastError());
        if (uc) FL_HIP_TRY(ctx, hipMemcpyAsync(uc, s->cell_out, s->n * 8, hipMemcpyDeviceToHost, stm));
        if (vc) FL_HIP_TRY(ctx, hipMemcpyAsync(vc, s->cell_out + s->n, s->n * 8, hipMemcpyDeviceToHost, stm));
        if (wc) FL_HIP_TRY(ctx, hipMemcpyAsync(wc, s->cell_out + 2 * s->n, s->n * 8, hipMemcpyDeviceToHost, stm));
    }
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

fl_status fl_cavity3d_set_timing(fl_cavity3d *s, int32_t enable, const char *kernel) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = s->ctx;
    int only = -1;
    if (kernel && kernel[0]) {
        only = s->prof.find(kernel);
        if (only < 0) return set_error(ctx, FL_ERR_INVALID_ARGUMENT, "unknown kernel '%s'", kernel);
    }
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    FL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    s->prof.on = enable != 0;
    s->prof.only = only;
    s->prof.clear();
    return FL_OK;
}

fl_status fl_cavity3d_set_learned_mode(fl_cavity3d *s, int32_t mode) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    if (mode != FL_LEARNED_CORRECTED && mode != FL_LEARNED_REPLACE)
        return set_error(s->ctx, FL_ERR_CONFIG, "unknown learned coarse-solve mode %d", mode);
    s->learned_mode = mode;
    return FL_OK;
}

fl_status fl_cavity3d_set_unet(fl_cavity3d *s, fl_unet *net) {
    if (!s) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = s->ctx;
    if (!net) {
        s->unet = nullptr;
        return FL_OK;
    }
    if (s->cfg.preconditioner != FL_PRECOND_LEARNED)
        return set_error(ctx, FL_ERR_CONFIG, "the solver was not created with the learned preconditioner");
    if (unet_context(net)->device != ctx->device)
        return set_error(ctx, FL_ERR_CONFIG, "the U-Net lives on device %d, the solver on %d",
                         unet_context(net)->device, ctx->device);
    if (s->g.nx != s->g.ny || s->g.ny != s->g.nz)
        return set_error(ctx, FL_ERR_CONFIG, "the learned warm start needs a cubic grid");
    const uint32_t nn = (uint32_t)unet_resolution(net);
    int L = -1;
    for (int l = 0; l <= s->tail && l < s->nlev; ++l)
        if (s->lv[l].nx == nn) L = l;
    if (L < 0)
        return set_error(ctx, FL_ERR_CONFIG,
                         "no multigrid level of the %u^3 grid has the U-Net's resolution %u^3", s->g.nx, nn);
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const size_t cells = (size_t)nn * nn * nn;
    if (!s->wl_in || s->lv[s->wl_level].n != cells) {
        if (s->wl_in) (void)hipFree(s->wl_in);
        if (s->wl_out) (void)hipFree(s->wl_out);
        s->wl_in = s->wl_out = nullptr;
        fl_status st = alloc(s, &s->wl_in, cells);
        if (st == FL_OK) st = alloc(s, &s->wl_out, cells);
        if (st != FL_OK) return st;
    }
    s->unet = net;
    s->wl_level = L;
    return FL_OK;
}

fl_status fl_cavity3d_capture_coarse(fl_cavity3d *s, int32_t n, int32_t iteration) {
    if (!s || iteration < 0) return FL_ERR_INVALID_ARGUMENT;
    s->cap_iter = iteration;
    fl_ctx *ctx = s->ctx;
    if (n == 0) {
        s->cap_level = 0;
        return FL_OK;
    }
    if (s->cfg.preconditioner == FL_PRECOND_JACOBI)
        return set_error(ctx, FL_ERR_CONFIG, "coarse capture needs a multigrid preconditioner");
    int L = -1;
    for (int l = 1; l <= s->tail && l < s->nlev; ++l)
        if (s->lv[l].nx == (uint32_t)n && s->lv[l].ny == (uint32_t)n && s->lv[l].nz == (uint32_t)n) L = l;
    if (L < 0)
        return set_error(ctx, FL_ERR_CONFIG, "no multigrid level below the finest has %d^3 cells", n);
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (s->cap) (void)hipFree(s->cap);
    s->cap = nullptr;
    fl_status st = alloc(s, &s->cap, s->lv[L].n);
    if (st != FL_OK) return st;
    s->cap_level = L;
    return FL_OK;
}

fl_status fl_cavity3d_captured_rhs(fl_cavity3d *s, float *rhs_c, double *diag) {
    if (!s || !rhs_c) return set_error(s ? s->ctx : nullptr, FL_ERR_INVALID_ARGUMENT, "null argument");
    fl_ctx *ctx = s->ctx;
    if (s->cap_level <= 0) return set_error(ctx, FL_ERR_CONFIG, "coarse capture is not armed");
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    const Level &L = s->lv[s->cap_level];
    FL_HIP_TRY(ctx, hipMemcpyAsync(rhs_c, s->cap, L.n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    FL_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (diag) {
        const double f = (double)(1u << s->cap_level);
        *diag = 2.0 * (s->g.ix2 + s->g.iy2 + s->g.iz2) / (f * f);
    }
    return FL_OK;
}

fl_status fl_cavity3d_coarse_pair(fl_cavity3d *s, int32_t n, float *rhs_c, float *p_c, double *diag) {
    if (!s || !rhs_c || !p_c) return set_error(s ? s->ctx : nullptr, FL_ERR_INVALID_ARGUMENT, "null argument");
    fl_ctx *ctx = s->ctx;
    if (s->g.nx != s->g.ny || s->g.ny != s->g.nz || n < 1 || s->g.nx % (uint32_t)n != 0)
        return set_error(ctx, FL_ERR_INVALID_ARGUMENT,
                         "coarse pairs need a cubic grid whose size %u is a multiple of n = %d", s->g.nx, n);
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t stm = ctx->stream;
    const size_t cells = (size_t)n * n * n;
    if (s->pair_cells < cells) {
        if (s->pair) (void)hipFree(s->pair);
        s->pair = nullptr;
        s->pair_cells = 0;
        fl_status st = alloc(s, &s->pair, 2 * cells);
        if (st != FL_OK) return st;
        s->pair_cells = cells;
    }
    const uint32_t f = s->g.nx / (uint32_t)n;
    const unsigned grid = (unsigned)std::min<size_t>((cells + kBlock - 1) / kBlock, 4096);
    // b = -rhs, the right-hand side of the solver's A p = b
    k_block_mean<<<grid, kBlock, 0, stm>>>(s->g, s->rhs, -1.0, (uint32_t)n, f, s->pair);
    k_block_mean<<<grid, kBlock, 0, stm>>>(s->g, s->p, 1.0, (uint32_t)n, f, s->pair + cells);
    FL_HIP_TRY(ctx, hipGetLastError());
    FL_HIP_TRY(ctx, hipMemcpyAsync(rhs_c, s->pair, cells * sizeof(float), hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipMemcpyAsync(p_c, s->pair + cells, cells * sizeof(float), hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
    if (diag) *diag = 2.0 * (s->g.ix2 + s->g.iy2 + s->g.iz2) / ((double)f * f);
    return FL_OK;
}

fl_status fl_cavity3d_kernel_stats(fl_cavity3d *s, int32_t index, fl_kernel_stat *out) {
    if (!s || !out) return FL_ERR_INVALID_ARGUMENT;
    if (index < 0 || index >= s->prof.kinds)
        return set_error(s->ctx, FL_ERR_INVALID_ARGUMENT, "kernel index %d out of range [0, %d)",
                         index, s->prof.kinds);
    std::memset(out, 0, sizeof *out);
    std::snprintf(out->name, sizeof out->name, "%s", s->prof.names[index]);
    out->launches = s->prof.launches[index];
    out->total_ms = s->prof.ms[index];
    out->bytes = s->prof.bytes[index];
    return FL_OK;
}

}  // extern "C"
