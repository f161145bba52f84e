// CSR linear algebra on gfx950: SpMV, CG / Jacobi-PCG and BiCGSTAB.
//
// Replaces src/linear/ (csr.rs:89-106 spmv_into, csr.rs:195-251 validate_csr,
// preconditioner.rs:9-70, solver.rs:62-256).  Errors carry the reference's
// LinearAlgebraError kinds and &'static str texts.
//
// Device layout: int32 row offsets and columns, FP64 values, FP64 vectors.
// Rows are cut into tiles handed out XCD-slab-wise (common.hpp); an SpMV
// works each row with four adjacent lanes (the unstructured Poisson rows
// hold ~4 entries, so a row's entries are one coalesced load per operand),
// the vector kernels one element per lane.
//
// Per CG iteration (3 kernels, scalars never leave the device):
//   k_cg_spmv    q = A d, d.q
//   k_cg_update  alpha; x += a d, r -= a q, r.r [, z = r/diag], r.z
//   k_cg_dir     stop test, beta; d = z + beta d
// Per BiCGSTAB iteration (4 kernels): s = r - alpha v is never stored; the
// second SpMV gathers r and v and forms s on the fly, and the final update
// recomputes it:
//   k_bi_p       stop test of the previous iteration, rho, beta;
//                p = r + beta (p - omega v)
//   k_bi_spmv_v  v = A p, rhat.v
//   k_bi_spmv_t  alpha; t = A s, s.s, t.t, t.s
//   k_bi_update  s-test, omega; x += a p + w s, r = s - w t, r.r, rhat.r
//
// Reductions without fences: a kernel publishes one partial row per block
// and the next kernel folds all rows in every block in a fixed order, so
// every block derives the same scalars and verdicts; block 0 records them in
// the state for the host and later kernels.  Scalars a kernel reads while
// its block 0 writes the next value (rho, and BiCGSTAB's alpha and omega)
// are kept per iteration parity.  Every launch carries a sequence number:
// the launch that ends the solve stores its own in done_seq, and launches
// with a larger number return at once, while the deciding launch's blocks
// all finish their share (the BiCGSTAB early exit updates x there).

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "common.hpp"

namespace fl {
namespace csr {

constexpr double kBreakdown = 1.0e-30;  // BREAKDOWN_EPSILON, solver.rs:3

// error contexts (the reference's &'static str values)
enum Ctx : int {
    C_NONE = 0,
    C_SPMV_OUTPUT,
    C_DOT,
    C_L2,
    C_PRECOND,
    C_RESIDUAL,
    C_CG_DEN,
    C_CG_STEP,
    C_CG_REC_DEN,
    C_CG_REC,
    C_BI_RHO,
    C_BI_BETA,
    C_BI_ALPHA_DEN,
    C_BI_ALPHA,
    C_BI_OMEGA_DEN,
    C_BI_OMEGA,
    C_DIVERGED,
};
const char *ctx_text(int c) {
    switch (c) {
        case C_SPMV_OUTPUT: return "SpMV output";
        case C_DOT: return "dot product";
        case C_L2: return "L2 norm";
        case C_PRECOND: return "preconditioned residual";
        case C_RESIDUAL: return "residual";
        case C_CG_DEN: return "zero CG search denominator";
        case C_CG_STEP: return "non-finite CG step";
        case C_CG_REC_DEN: return "zero CG recurrence denominator";
        case C_CG_REC: return "non-finite CG recurrence";
        case C_BI_RHO: return "zero BiCGSTAB rho";
        case C_BI_BETA: return "non-finite BiCGSTAB beta";
        case C_BI_ALPHA_DEN: return "zero BiCGSTAB alpha denominator";
        case C_BI_ALPHA: return "non-finite BiCGSTAB alpha";
        case C_BI_OMEGA_DEN: return "zero BiCGSTAB omega denominator";
        case C_BI_OMEGA: return "zero BiCGSTAB omega";
        case C_DIVERGED: return "diverging inner solve";  // internal to the mixed path
        default: return "";
    }
}

template <typename T> struct MatT {
    const int *off;
    const int *col;
    const T *val;
    uint32_t n;
};
typedef MatT<double> Mat;

struct LinState {
    double rho[2];              // r.z (CG) or rhat.r (BiCGSTAB) by iteration parity
    double alpha[2], omega[2];  // BiCGSTAB, by iteration parity
    double atol, rtol;          // the options (threshold = atol + rtol ||r0||)
    double in_scale;            // 1, or (mixed inner solve) 1 / ||r|| of the round's residual
    double diverge;             // mixed inner solve: end it when ||r|| exceeds this (0: never)
    double threshold, initial, rnorm;
    long long iter, max_iter;
    long long done_seq;         // sequence number of the launch that ended the solve
    int converged, status, ctx, pad;
    long long err_index;
    unsigned long long bad_row;   // first non-finite row of the (first) SpMV of an iteration
    unsigned long long bad_row2;  // of BiCGSTAB's second SpMV
    unsigned long long bad_r;     // first non-finite initial residual entry
    unsigned long long bad_z;     // first non-finite preconditioned entry
};
constexpr long long kRunning = 0x7fffffffffffffffll;  // done_seq of a solve still running

// SpMV rows are worked by groups of kRowLanes adjacent lanes: per chunk of
// four entries lane q loads entry q (the ~4 entries of a Poisson row are one
// coalesced load per operand across the group) and forms its product; the
// group then adds the chunk's products in entry order, so the row value is
// the reference's left-to-right sum of rounded products (spmv_into,
// csr.rs:89-106) and every lane of the group holds it.  A 256-thread block
// covers a tile of 64 rows.
constexpr uint32_t kRowLanes = 4, kRowsPerTile = kBlock / kRowLanes;
struct RowSlot {
    uint32_t row, q;
    bool valid;
};
__device__ inline RowSlot row_slot(uint32_t tile, uint32_t n) {
    RowSlot r;
    r.row = tile * kRowsPerTile + threadIdx.x / kRowLanes;
    r.q = threadIdx.x % kRowLanes;
    r.valid = r.row < n;
    return r;
}
template <typename T, typename F>
__device__ inline T row_dot_lanes(const MatT<T> &A, const RowSlot &r, F &&xval) {
    T acc = 0;
    const int b = r.valid ? A.off[r.row] : 0, e = r.valid ? A.off[r.row + 1] : 0;
    const int lead = (int)(threadIdx.x & 63u) & ~(int)(kRowLanes - 1);
    for (int k0 = b; k0 < e; k0 += kRowLanes) {  // uniform over the group
        const int k = k0 + (int)r.q;
        const T prod = k < e ? A.val[k] * xval(A.col[k]) : (T)0;
#pragma unroll
        for (int t = 0; t < (int)kRowLanes; ++t) acc += __shfl(prod, lead + t);
    }
    return acc;
}

__device__ inline bool lead() { return blockIdx.x == 0 && threadIdx.x == 0; }

// The solve ends in this launch with an error (every block returns; block 0
// records it).
__device__ inline void fail(LinState *st, long long seq, int status, int ctx, long long index) {
    if (!lead()) return;
    st->status = status;
    st->ctx = ctx;
    st->err_index = index;
    st->done_seq = seq;
}
// The solve ends in this launch without an error (converged or
// MaxIterations, recorded by block 0).
__device__ inline void finish(LinState *st, long long seq, int converged, long long iter) {
    if (!lead()) return;
    st->converged = converged;
    st->iter = iter;
    st->done_seq = seq;
}

// ------------------------------------------------------------------ SpMV
__global__ void __launch_bounds__(kBlock) k_spmv(Mat A, const double *__restrict__ x,
                                                 double *__restrict__ y, LinState *st) {
    TileSched ts((A.n + kRowsPerTile - 1) / kRowsPerTile);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const RowSlot r = row_slot(t, A.n);
        const double v = row_dot_lanes(A, r, [&](int c) { return x[c]; });
        if (!r.valid || r.q != 0) continue;
        if (!isfinite(v)) atomicMin(&st->bad_row, (unsigned long long)r.row);
        y[r.row] = v;
    }
}

// --------------------------------------------------------- CG / PCG init
// r = b - q (q = A x from k_spmv), partial r.r.
__global__ void __launch_bounds__(kBlock) k_cg_init_r(uint32_t n, const double *__restrict__ b,
                                                      const double *__restrict__ q,
                                                      double *__restrict__ r, double *__restrict__ pout,
                                                      LinState *st) {
    double s[1] = {0.0}, m[1] = {0.0};
    TileSched ts((n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t i = t * kBlock + threadIdx.x;
        if (i >= n) continue;
        const double v = b[i] - q[i];
        if (!isfinite(v)) atomicMin(&st->bad_r, (unsigned long long)i);
        r[i] = v;
        s[0] += v * v;
    }
    publish_partials<1, 1>(s, m, pout);
}

// The checks of residual_into and norm_l2 on the initial residual and the
// stop test at iteration 0 (solver.rs:82-88), from k_cg_init_r's rows.
// Returns false when the solve ends here.
__device__ inline bool initial_tests(const double *__restrict__ pin, uint32_t rows, LinState *st,
                                     long long seq) {
    double f[1], fm[1];
    fold_partials<1, 1>(pin, rows, f, fm);
    if (st->bad_row != ~0ull) return fail(st, seq, FL_ERR_NON_FINITE, C_SPMV_OUTPUT, st->bad_row), false;
    if (st->bad_r != ~0ull) return fail(st, seq, FL_ERR_NON_FINITE, C_RESIDUAL, st->bad_r), false;
    if (!isfinite(f[0])) return fail(st, seq, FL_ERR_NON_FINITE, C_DOT, 0), false;
    const double init = sqrt(f[0]) * st->in_scale;
    const double threshold = st->atol + st->rtol * init;
    if (lead()) {
        st->initial = init;
        st->rnorm = init;
        st->threshold = threshold;
    }
    if (init <= threshold) return finish(st, seq, 1, 0), false;
    return true;
}

// Initial tests; z = D^-1 r (or r), d = z, partial r.z.
// (T = float: the inner solve of the mixed path, which starts from e = 0 on
// the round's residual scaled to unit norm: r is scaled here by in_scale.)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_init_d(uint32_t n, T *__restrict__ r,
                                                      const T *__restrict__ inv,
                                                      T *__restrict__ d, const double *__restrict__ pin,
                                                      uint32_t rows, double *__restrict__ pout, LinState *st,
                                                      long long seq, T *__restrict__ x0) {
    if (st->done_seq < seq || !initial_tests(pin, rows, st, seq)) return;
    const double sc = st->in_scale;
    double s[1] = {0.0}, m[1] = {0.0};
    TileSched ts((n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t i = t * kBlock + threadIdx.x;
        if (i >= n) continue;
        if (sc != 1.0) r[i] *= (T)sc;
        const T z = inv ? r[i] * inv[i] : r[i];
        if (!isfinite(z)) atomicMin(&st->bad_z, (unsigned long long)i);
        d[i] = z;
        if (x0) x0[i] = 0;
        s[0] += (double)r[i] * (double)z;
    }
    publish_partials<1, 1>(s, m, pout);
}

// rho = r.z for iteration 1 (one workgroup).
__global__ void __launch_bounds__(kBlock) k_cg_init_rho(const double *__restrict__ pin, uint32_t rows,
                                                        LinState *st, long long seq) {
    if (st->done_seq < seq) return;
    double f[1], fm[1];
    fold_partials<1, 1>(pin, rows, f, fm);
    if (st->bad_z != ~0ull) return fail(st, seq, FL_ERR_NON_FINITE, C_PRECOND, st->bad_z);
    if (!isfinite(f[0])) return fail(st, seq, FL_ERR_NON_FINITE, C_DOT, 0);
    if (lead()) st->rho[1] = f[0];
    if (st->max_iter < 1) return finish(st, seq, 0, 0);  // the loop 1..=0 never runs
}

// ------------------------------------------------------------ CG loop
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_spmv(MatT<T> A, const T *__restrict__ d,
                                                    T *__restrict__ q, double *__restrict__ pout,
                                                    LinState *st, long long seq) {
    if (st->done_seq < seq) return;
    double s[1] = {0.0}, m[1] = {0.0};
    TileSched ts((A.n + kRowsPerTile - 1) / kRowsPerTile);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const RowSlot r = row_slot(t, A.n);
        const T v = row_dot_lanes(A, r, [&](int c) { return d[c]; });
        if (!r.valid || r.q != 0) continue;
        if (!isfinite(v)) atomicMin(&st->bad_row, (unsigned long long)r.row);
        q[r.row] = v;
        s[0] += (double)d[r.row] * (double)v;
    }
    publish_partials<1, 1>(s, m, pout);
}

// alpha = rho / d.q (solver.rs:98-110); x += alpha d, r -= alpha q, partials
// r.r and r.z.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_update(uint32_t n, T *__restrict__ x,
                                                      T *__restrict__ r,
                                                      const T *__restrict__ d,
                                                      const T *__restrict__ q,
                                                      const T *__restrict__ inv,
                                                      const double *__restrict__ pin, uint32_t rows,
                                                      double *__restrict__ pout, LinState *st,
                                                      long long it, long long seq) {
    if (st->done_seq < seq) return;
    double f[1], fm[1];
    fold_partials<1, 1>(pin, rows, f, fm);
    if (st->bad_row != ~0ull) return fail(st, seq, FL_ERR_NON_FINITE, C_SPMV_OUTPUT, st->bad_row);
    const double den = f[0];
    if (!isfinite(den)) return fail(st, seq, FL_ERR_NON_FINITE, C_DOT, 0);
    if (fabs(den) <= kBreakdown) return fail(st, seq, FL_ERR_BREAKDOWN, C_CG_DEN, 0);
    const double alpha = st->rho[it & 1] / den;
    if (!isfinite(alpha)) return fail(st, seq, FL_ERR_BREAKDOWN, C_CG_STEP, 0);
    double s[2] = {0.0, 0.0}, m[1] = {0.0};
    TileSched ts((n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t i = t * kBlock + threadIdx.x;
        if (i >= n) continue;
        x[i] += (T)alpha * d[i];
        const T ri = r[i] - (T)alpha * q[i];
        r[i] = ri;
        s[0] += (double)ri * (double)ri;
        const T z = inv ? ri * inv[i] : ri;
        if (!isfinite(z)) atomicMin(&st->bad_z, (unsigned long long)i);
        s[1] += (double)ri * (double)z;
    }
    publish_partials<2, 1>(s, m, pout);
}

// The stop test (norm_l2, solver.rs:120-133), rho_new = r.z and beta
// (:134-148), MaxIterations after the last iteration; d = z + beta d.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cg_dir(uint32_t n, const T *__restrict__ r,
                                                   const T *__restrict__ inv,
                                                   T *__restrict__ d, const double *__restrict__ pin,
                                                   uint32_t rows, LinState *st, long long it, long long seq) {
    if (st->done_seq < seq) return;
    double f[2], fm[1];
    fold_partials<2, 1>(pin, rows, f, fm);
    if (!isfinite(f[0])) return fail(st, seq, FL_ERR_NON_FINITE, C_DOT, 0);
    const double rn = sqrt(f[0]);
    if (lead()) st->rnorm = rn;
    if (rn <= st->threshold) return finish(st, seq, 1, it);
    if (st->diverge > 0.0 && rn > st->diverge) return fail(st, seq, FL_ERR_BREAKDOWN, C_DIVERGED, 0);
    if (st->bad_z != ~0ull) return fail(st, seq, FL_ERR_NON_FINITE, C_PRECOND, st->bad_z);
    const double rho_new = f[1], rho = st->rho[it & 1];
    if (!isfinite(rho_new)) return fail(st, seq, FL_ERR_NON_FINITE, C_DOT, 0);
    if (fabs(rho) <= kBreakdown) return fail(st, seq, FL_ERR_BREAKDOWN, C_CG_REC_DEN, 0);
    const double beta = rho_new / rho;
    if (!isfinite(beta)) return fail(st, seq, FL_ERR_BREAKDOWN, C_CG_REC, 0);
    if (lead()) {
        st->rho[(it + 1) & 1] = rho_new;
        st->iter = it;
    }
    if (it >= st->max_iter) return finish(st, seq, 0, it);  // MaxIterations: a status, not an error
    TileSched ts((n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t i = t * kBlock + threadIdx.x;
        if (i >= n) continue;
        const T z = inv ? r[i] * inv[i] : r[i];
        d[i] = z + (T)beta * d[i];
    }
}

// ------------------------------------------------------------ BiCGSTAB
// Initial r = b - A x as for CG (k_spmv + k_cg_init_r); then the initial
// tests, rhat = r, p = v = 0, the iteration-0 scalars rho = alpha = omega
// = 1 (solver.rs:175-177) and partials {r.r, rhat.r}.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bi_init(uint32_t n, T *__restrict__ r,
                                                    T *__restrict__ rhat, T *__restrict__ p,
                                                    T *__restrict__ v, const double *__restrict__ pin,
                                                    uint32_t rows, double *__restrict__ pout, LinState *st,
                                                    long long seq, T *__restrict__ x0) {
    if (st->done_seq < seq || !initial_tests(pin, rows, st, seq)) return;
    if (lead()) {
        st->rho[0] = 1.0;
        st->alpha[0] = 1.0;
        st->omega[0] = 1.0;
    }
    if (st->max_iter < 1) return finish(st, seq, 0, 0);
    const double sc = st->in_scale;
    double s[2] = {0.0, 0.0}, m[1] = {0.0};
    TileSched ts((n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t i = t * kBlock + threadIdx.x;
        if (i >= n) continue;
        if (sc != 1.0) r[i] *= (T)sc;
        const T ri = r[i];
        rhat[i] = ri;
        p[i] = 0;
        v[i] = 0;
        if (x0) x0[i] = 0;
        s[0] += (double)ri * (double)ri;
        s[1] += (double)ri * (double)ri;
    }
    publish_partials<2, 1>(s, m, pout);
}

// The end of iteration it - 1 (its residual norm test, solver.rs:236-245,
// and MaxIterations), then iteration it's rho = rhat.r and beta (:179-190);
// p = r + beta (p - omega v).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bi_p(uint32_t n, const T *__restrict__ r,
                                                 T *__restrict__ p, const T *__restrict__ v,
                                                 const double *__restrict__ pin, uint32_t rows, LinState *st,
                                                 long long it, long long seq) {
    if (st->done_seq < seq) return;
    double f[2], fm[1];
    fold_partials<2, 1>(pin, rows, f, fm);
    if (it > 1) {
        if (!isfinite(f[0])) return fail(st, seq, FL_ERR_NON_FINITE, C_DOT, 0);
        const double rn = sqrt(f[0]);
        if (lead()) st->rnorm = rn;
        if (rn <= st->threshold) return finish(st, seq, 1, it - 1);
        if (st->diverge > 0.0 && rn > st->diverge) return fail(st, seq, FL_ERR_BREAKDOWN, C_DIVERGED, 0);
        if (it - 1 >= st->max_iter) return finish(st, seq, 0, it - 1);
    }
    const int c = (int)(it & 1), pv = c ^ 1;
    const double rho = f[1];
    if (!isfinite(rho)) return fail(st, seq, FL_ERR_NON_FINITE, C_DOT, 0);
    if (fabs(rho) <= kBreakdown) return fail(st, seq, FL_ERR_BREAKDOWN, C_BI_RHO, 0);
    const double omega = st->omega[pv];
    const double beta = (rho / st->rho[pv]) * (st->alpha[pv] / omega);
    if (!isfinite(beta)) return fail(st, seq, FL_ERR_BREAKDOWN, C_BI_BETA, 0);
    if (lead()) {
        st->rho[c] = rho;
        st->iter = it;
    }
    TileSched ts((n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t i = t * kBlock + threadIdx.x;
        if (i >= n) continue;
        p[i] = r[i] + (T)beta * (p[i] - (T)omega * v[i]);
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_bi_spmv_v(MatT<T> A, const T *__restrict__ p,
                                                      const T *__restrict__ rhat,
                                                      T *__restrict__ v, double *__restrict__ pout,
                                                      LinState *st, long long seq) {
    if (st->done_seq < seq) return;
    double s[1] = {0.0}, m[1] = {0.0};
    TileSched ts((A.n + kRowsPerTile - 1) / kRowsPerTile);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const RowSlot r = row_slot(t, A.n);
        const T vv = row_dot_lanes(A, r, [&](int c) { return p[c]; });
        if (!r.valid || r.q != 0) continue;
        if (!isfinite(vv)) atomicMin(&st->bad_row, (unsigned long long)r.row);
        v[r.row] = vv;
        s[0] += (double)rhat[r.row] * (double)vv;
    }
    publish_partials<1, 1>(s, m, pout);
}

// alpha = rho / rhat.v (solver.rs:192-203); t = A s with s = r - alpha v
// formed at each gathered column; partials s.s, t.t, t.s.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bi_spmv_t(MatT<T> A, const T *__restrict__ r,
                                                      const T *__restrict__ v,
                                                      T *__restrict__ tvec,
                                                      const double *__restrict__ pin, uint32_t rows,
                                                      double *__restrict__ pout, LinState *st, long long it,
                                                      long long seq) {
    if (st->done_seq < seq) return;
    double f[1], fm[1];
    fold_partials<1, 1>(pin, rows, f, fm);
    if (st->bad_row != ~0ull) return fail(st, seq, FL_ERR_NON_FINITE, C_SPMV_OUTPUT, st->bad_row);
    const double den = f[0];
    if (!isfinite(den)) return fail(st, seq, FL_ERR_NON_FINITE, C_DOT, 0);
    if (fabs(den) <= kBreakdown) return fail(st, seq, FL_ERR_BREAKDOWN, C_BI_ALPHA_DEN, 0);
    const double alpha = st->rho[it & 1] / den;
    if (!isfinite(alpha)) return fail(st, seq, FL_ERR_BREAKDOWN, C_BI_ALPHA, 0);
    if (lead()) st->alpha[it & 1] = alpha;
    double s[3] = {0.0, 0.0, 0.0}, m[1] = {0.0};
    TileSched ts((A.n + kRowsPerTile - 1) / kRowsPerTile);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const RowSlot rs = row_slot(t, A.n);
        const T ta = (T)alpha;
        const T acc = row_dot_lanes(A, rs, [&](int c) { return r[c] - ta * v[c]; });
        if (!rs.valid || rs.q != 0) continue;
        const uint32_t row = rs.row;
        if (!isfinite(acc)) atomicMin(&st->bad_row2, (unsigned long long)row);
        tvec[row] = acc;
        const T si = r[row] - ta * v[row];
        s[0] += (double)si * (double)si;
        s[1] += (double)acc * (double)acc;
        s[2] += (double)acc * (double)si;
    }
    publish_partials<3, 1>(s, m, pout);
}

// The s-test (early exit with x += alpha p, solver.rs:206-218), omega =
// t.s / t.t (:220-231); x += alpha p + omega s, r = s - omega t, partials
// r.r and rhat.r.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bi_update(uint32_t n, T *__restrict__ x,
                                                      T *__restrict__ r,
                                                      const T *__restrict__ p,
                                                      const T *__restrict__ v,
                                                      const T *__restrict__ tvec,
                                                      const T *__restrict__ rhat,
                                                      const double *__restrict__ pin, uint32_t rows,
                                                      double *__restrict__ pout, LinState *st, long long it,
                                                      long long seq) {
    if (st->done_seq < seq) return;
    double f[3], fm[1];
    fold_partials<3, 1>(pin, rows, f, fm);
    const double alpha = st->alpha[it & 1];
    if (!isfinite(f[0])) return fail(st, seq, FL_ERR_NON_FINITE, C_DOT, 0);
    const double sn = sqrt(f[0]);
    if (sn <= st->threshold) {  // solution += alpha p and return Converged
        TileSched ts((n + kBlock - 1) / kBlock);
        for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
            const uint32_t i = t * kBlock + threadIdx.x;
            if (i < n) x[i] += (T)alpha * p[i];
        }
        if (lead()) st->rnorm = sn;
        return finish(st, seq, 1, it);
    }
    if (st->bad_row2 != ~0ull) return fail(st, seq, FL_ERR_NON_FINITE, C_SPMV_OUTPUT, st->bad_row2);
    const double tt = f[1];
    if (!isfinite(tt)) return fail(st, seq, FL_ERR_NON_FINITE, C_DOT, 0);
    if (fabs(tt) <= kBreakdown) return fail(st, seq, FL_ERR_BREAKDOWN, C_BI_OMEGA_DEN, 0);
    if (!isfinite(f[2])) return fail(st, seq, FL_ERR_NON_FINITE, C_DOT, 0);
    const double omega = f[2] / tt;
    if (!isfinite(omega) || fabs(omega) <= kBreakdown) return fail(st, seq, FL_ERR_BREAKDOWN, C_BI_OMEGA, 0);
    if (lead()) st->omega[it & 1] = omega;
    double s[2] = {0.0, 0.0}, m[1] = {0.0};
    TileSched ts((n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t i = t * kBlock + threadIdx.x;
        if (i >= n) continue;
        const T ta = (T)alpha, tw = (T)omega;
        const T si = r[i] - ta * v[i];
        x[i] += ta * p[i] + tw * si;
        const T ri = si - tw * tvec[i];
        r[i] = ri;
        s[0] += (double)ri * (double)ri;
        s[1] += (double)rhat[i] * (double)ri;
    }
    publish_partials<2, 1>(s, m, pout);
}

// ------------------------------------------------- FP32 inner, FP64 refinement
// The mixed path (fl_csr_set_precision, FL_PRECISION_MIXED): the solver
// runs in FP32 (values, vectors, SpMV, axpys; dots, alpha/beta and tests in
// FP64) on the correction e of A e = r inside FP64 iterative refinement.
// Each round forms the true FP64 residual r = b - A x (k_spmv with the FP64
// values, k_res_mixed), makes the reference's stop test on it against the
// threshold fixed by the initial residual (k_outer_test), solves for e from
// e = 0 and adds it, x += e (k_add_e).
struct OuterState {
    double atol, rtol;        // the caller's options
    double threshold, initial, rnorm, prev;
    long long round;
    int converged, stalled, exhausted, pad;
};

// r = b - q in FP64 (q = A x), stored FP32; partial r.r of the FP64 values;
// the round's correction e = 0.
__global__ void __launch_bounds__(kBlock) k_res_mixed(uint32_t n, const double *__restrict__ b,
                                                      const double *__restrict__ q, float *__restrict__ r32,
                                                      float *__restrict__ e, double *__restrict__ pout,
                                                      LinState *st) {
    double s[1] = {0.0}, m[1] = {0.0};
    TileSched ts((n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t i = t * kBlock + threadIdx.x;
        if (i >= n) continue;
        const double v = b[i] - q[i];
        if (!isfinite(v)) atomicMin(&st->bad_r, (unsigned long long)i);
        r32[i] = (float)v;
        e[i] = 0.0f;  // the round's correction starts from zero
        s[0] += v * v;
    }
    publish_partials<1, 1>(s, m, pout);
}

// The outer stop test of a round (one workgroup): ||r||_2 of the FP64
// residual against atol + rtol ||r_0||_2 (solver.rs:82-88, 120-133); the
// round's inner solve targets ||r_e|| <= threshold / 2 + inner_rel ||r||
// and is skipped (done_seq 0) when the test passes, when the last round did
// not halve the residual, or when the iteration budget is spent.
__global__ void __launch_bounds__(kBlock) k_outer_test(const double *__restrict__ pin, uint32_t rows,
                                                       OuterState *os, LinState *st, double inner_rel,
                                                       int exhausted) {
    double f[1], fm[1];
    fold_partials<1, 1>(pin, rows, f, fm);
    if (threadIdx.x != 0 || !isfinite(f[0])) return;  // (a non-finite residual: the inner tests report it)
    const double rn = sqrt(f[0]);
    if (os->round == 0) {
        os->initial = rn;
        os->threshold = os->atol + os->rtol * rn;
    } else if (rn > 0.5 * os->prev) {
        os->stalled = 1;  // a round must at least halve the residual
    }
    os->rnorm = rn;
    os->prev = rn;
    os->round += 1;
    if (rn <= os->threshold) os->converged = 1;
    if (exhausted) os->exhausted = 1;
    if (os->converged || os->stalled || os->exhausted) {
        st->done_seq = 0;
        return;
    }
    // the inner solve works on r / ||r|| (unit norm): its absolute tests act
    // as on a problem of ordinary scale however small the residual has become
    st->in_scale = 1.0 / rn;
    st->atol = 0.5 * os->threshold / rn;
    st->rtol = inner_rel;
    st->diverge = 1.0e4;  // (unit initial residual) an FP32 BiCGSTAB that blows up ends early
}

// x += ||r|| e (the inner solve's e solves the unit-norm system)
__global__ void __launch_bounds__(kBlock) k_add_e(uint32_t n, double *__restrict__ x,
                                                  const float *__restrict__ e, const OuterState *os) {
    const double sc = os->rnorm;
    TileSched ts((n + kBlock - 1) / kBlock);
    for (uint32_t t = ts.first; t < ts.count; t += ts.step) {
        const uint32_t i = t * kBlock + threadIdx.x;
        if (i < n) x[i] += (double)e[i] * sc;
    }
}

__global__ void __launch_bounds__(kBlock) k_to_f32(long long n, const double *__restrict__ a,
                                                   float *__restrict__ b) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock)
        b[i] = (float)a[i];
}

// ------------------------------------------- allocation-free solve buffers
// fl_pcg_jacobi_noalloc keeps every device buffer of a solve -- the five
// work vectors x, b, r, d, q, the per-block partial rows and the solver
// state -- in this file-scope array instead of library allocations.  Each
// GPU the library runs on holds its own copy (42 MiB) until the process
// ends.  Host code never passes `scratch` itself to a kernel or copy (in host
// code the name is not a device address); it uses the per-device address
// from hipGetSymbolAddress, looked up once per device.
constexpr size_t kNoAllocMaxRows = (size_t)1 << 20;  // largest matrix dimension
constexpr size_t kNoAllocStride = kNoAllocMaxRows + 32;  // one vector, 256-byte aligned
constexpr uint32_t kNoAllocMaxBlocks = 4096;
constexpr size_t kNoAllocPartials = (size_t)kNoAllocMaxBlocks * 8;  // two buffers of 4-double rows
constexpr size_t kNoAllocState = 64;  // doubles reserved for LinState
constexpr size_t kScratchDoubles = 5 * kNoAllocStride + kNoAllocPartials + kNoAllocState;
static_assert(sizeof(LinState) <= kNoAllocState * sizeof(double), "LinState outgrew its slot");
static_assert(kNoAllocStride % 32 == 0, "vectors must stay 256-byte aligned");
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (the current device), looked
// up on the first call for that device.
hipError_t scratch_address(int device, double **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<double *>(ptr);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

}  // namespace csr

// Called by fl_create with the context's device current.
void noalloc_init(int device) {
    double *p = nullptr;
    (void)csr::scratch_address(device, &p);
}

}  // namespace fl

using namespace fl;
using namespace fl::csr;

struct fl_csr {
    fl_ctx *ctx = nullptr;
    uint32_t n = 0, ncols = 0;
    int64_t nnz = 0;
    int *off = nullptr, *col = nullptr;
    double *val = nullptr;
    double *inv = nullptr;     // Jacobi inverse diagonal (when valid)
    int inv_status = FL_OK;    // JacobiPreconditioner::new outcome
    int64_t inv_row = 0;
    const char *inv_ctx = "";
    double *vec[8] = {};       // x, b, r, d/p, q/v, t, rhat, spare
    int precision = FL_PRECISION_FP64;  // fl_csr_set_precision
    double inner_rel = 1e-5;   // mixed: the inner solve's relative target
    float *val32 = nullptr, *inv32 = nullptr;  // mixed: FP32 copies (made on first use)
    float *vec32[6] = {};      // mixed: r, d/p, q/v, e, t, rhat
    OuterState *os = nullptr, *os_host = nullptr;
    double *partials = nullptr;
    LinState *st = nullptr, *st_host = nullptr;
    uint32_t blocks = 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    unsigned long long *first = nullptr;  // Jacobi refresh: the first failing row
};

namespace {

void set_detail(fl_err_detail *d, int kind, const char *ctx, int64_t index = 0, int64_t expected = 0,
                int64_t actual = 0, int64_t row = 0, int64_t column = 0) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->kind = kind;
    d->index = index;
    d->expected = expected;
    d->actual = actual;
    d->row = row;
    d->column = column;
    std::snprintf(d->context, sizeof d->context, "%s", ctx ? ctx : "");
}

fl_status err(fl_ctx *ctx, fl_err_detail *d, fl_status st, const char *text, int64_t index = 0,
              int64_t row = 0) {
    set_detail(d, st, text, index, 0, 0, row);
    return set_error(ctx, st, "%s (index %lld)", text, (long long)index);
}

void free_csr(fl_csr *m) {
    void *bufs[] = {m->off, m->col, m->val, m->inv, m->partials, m->st, m->first, m->val32, m->inv32, m->os};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (float *v : m->vec32)
        if (v) (void)hipFree(v);
    if (m->os_host) (void)hipHostFree(m->os_host);
    for (double *v : m->vec)
        if (v) (void)hipFree(v);
    if (m->st_host) (void)hipHostFree(m->st_host);
    if (m->e0) (void)hipEventDestroy(m->e0);
    if (m->e1) (void)hipEventDestroy(m->e1);
}

fl_status ensure_vectors(fl_csr *m, int count) {
    for (int k = 0; k < count; ++k)
        if (!m->vec[k]) FL_HIP_TRY(m->ctx, hipMalloc(&m->vec[k], (size_t)m->n * sizeof(double) + 8));
    return FL_OK;
}

// validate_solve_inputs (solver.rs:258-300) on the host arrays.
fl_status validate_inputs(fl_csr *m, const double *rhs, const double *x, const fl_lin_opts *o,
                          fl_err_detail *d) {
    if (m->n != m->ncols) {
        set_detail(d, FL_ERR_DIMENSION_MISMATCH, "", 0, m->n, m->ncols);
        return set_error(m->ctx, FL_ERR_DIMENSION_MISMATCH, "dimension mismatch: expected %u, got %u",
                         m->n, m->ncols);
    }
    if (!std::isfinite(o->absolute_tolerance) || o->absolute_tolerance < 0.0 ||
        !std::isfinite(o->relative_tolerance) || o->relative_tolerance < 0.0)
        return err(m->ctx, d, FL_ERR_BREAKDOWN, "invalid solver tolerance");
    for (uint32_t i = 0; i < m->n; ++i)
        if (!std::isfinite(rhs[i])) return err(m->ctx, d, FL_ERR_NON_FINITE, "solver vector", i);
    for (uint32_t i = 0; i < m->n; ++i)
        if (!std::isfinite(x[i]))
            return err(m->ctx, d, FL_ERR_NON_FINITE, "solver vector", (int64_t)m->n + i);
    return FL_OK;
}

fl_status state_error(fl_csr *m, const LinState &h, fl_err_detail *d) {
    const char *text = ctx_text(h.ctx);
    return err(m->ctx, d, (fl_status)h.status, text, h.err_index);
}

enum Kind { KIND_CG = 0, KIND_PCG = 1, KIND_BICG = 2 };

// Device buffers of one solve: vectors x, b, r, d/p, q/v (+ t, rhat for
// BiCGSTAB), two buffers of per-block partial rows (blocks x 4 doubles
// each, the second right after the first) and the solver state.
struct Work {
    double *vec[7];
    double *partials;
    LinState *st;
};

// The checks every solve entry point makes before touching the device.
fl_status check_solve(fl_csr *m, int kind, const double *rhs, const double *x_inout,
                      const fl_lin_opts *opts, fl_err_detail *d) {
    fl_ctx *ctx = m->ctx;
    if (!rhs || !x_inout || !opts) return err(ctx, d, FL_ERR_INVALID_ARGUMENT, "null argument");
    fl_status st = validate_inputs(m, rhs, x_inout, opts, d);
    if (st != FL_OK) return st;
    if (kind == KIND_PCG && m->inv_status != FL_OK) {
        set_detail(d, m->inv_status, m->inv_ctx, m->inv_row, 0, 0, m->inv_row);
        return set_error(ctx, (fl_status)m->inv_status, "Jacobi preconditioner: %s (row %lld)",
                         m->inv_ctx, (long long)m->inv_row);
    }
    return FL_OK;
}

// The solve proper on caller-provided work buffers (check_solve passed and
// the context's device is current).
fl_status solve_on(fl_csr *m, int kind, const double *rhs, double *x_inout, const fl_lin_opts *opts,
                   fl_lin_report *rep, fl_err_detail *d, const Work &w, bool device_io = false) {
    fl_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    const size_t bytes = (size_t)m->n * sizeof(double);
    double *x = w.vec[0], *b = w.vec[1], *r = w.vec[2], *dp = w.vec[3], *qv = w.vec[4];
    double *tv = w.vec[5], *rhat = w.vec[6];
    const hipMemcpyKind in = device_io ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const hipMemcpyKind out = device_io ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    FL_HIP_TRY(ctx, hipMemcpyAsync(x, x_inout, bytes, in, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(b, rhs, bytes, in, s));
    LinState init{};
    init.atol = opts->absolute_tolerance;
    init.rtol = opts->relative_tolerance;
    init.max_iter = opts->max_iterations;
    init.in_scale = 1.0;
    init.done_seq = kRunning;
    init.bad_row = init.bad_row2 = init.bad_r = init.bad_z = ~0ull;
    *m->st_host = init;
    LinState *const dst = w.st;
    FL_HIP_TRY(ctx, hipMemcpyAsync(dst, m->st_host, sizeof(LinState), hipMemcpyHostToDevice, s));
    const Mat A{m->off, m->col, m->val, m->n};
    const double *inv = kind == KIND_PCG ? m->inv : nullptr;
    const uint32_t G = m->blocks;
    // two partial-row buffers: a kernel folds one and publishes into the other
    double *const pa = w.partials, *const pb = w.partials + (size_t)G * 4;
    long long seq = 0;  // launch sequence number (k_spmv and k_cg_init_r: 0, 1)
    FL_HIP_TRY(ctx, hipEventRecord(m->e0, s));
    k_spmv<<<G, kBlock, 0, s>>>(A, x, qv, dst);
    k_cg_init_r<<<G, kBlock, 0, s>>>(m->n, b, qv, r, pb, dst);
    seq = 2;
    if (kind == KIND_BICG) {
        k_bi_init<<<G, kBlock, 0, s>>>(m->n, r, rhat, dp, qv, pb, G, pa, dst, seq++, (double *)nullptr);
    } else {
        k_cg_init_d<<<G, kBlock, 0, s>>>(m->n, r, inv, dp, pb, G, pa, dst, seq++, (double *)nullptr);
        k_cg_init_rho<<<1, kBlock, 0, s>>>(pa, G, dst, seq++);
    }
    FL_HIP_TRY(ctx, hipGetLastError());
    int batch = 4;
    long long it = 1;
    for (;;) {
        for (int k = 0; k < batch; ++k, ++it) {
            if (kind == KIND_BICG) {
                k_bi_p<<<G, kBlock, 0, s>>>(m->n, r, dp, qv, pa, G, dst, it, seq++);
                k_bi_spmv_v<<<G, kBlock, 0, s>>>(A, dp, rhat, qv, pa, dst, seq++);
                k_bi_spmv_t<<<G, kBlock, 0, s>>>(A, r, qv, tv, pa, G, pb, dst, it, seq++);
                k_bi_update<<<G, kBlock, 0, s>>>(m->n, x, r, dp, qv, tv, rhat, pb, G, pa, dst, it, seq++);
            } else {
                k_cg_spmv<<<G, kBlock, 0, s>>>(A, dp, qv, pa, dst, seq++);
                k_cg_update<<<G, kBlock, 0, s>>>(m->n, x, r, dp, qv, inv, pa, G, pb, dst, it, seq++);
                k_cg_dir<<<G, kBlock, 0, s>>>(m->n, r, inv, dp, pb, G, dst, it, seq++);
            }
        }
        FL_HIP_TRY(ctx, hipGetLastError());
        FL_HIP_TRY(ctx, hipMemcpyAsync(m->st_host, dst, sizeof(LinState), hipMemcpyDeviceToHost, s));
        FL_HIP_TRY(ctx, hipStreamSynchronize(s));
        if (m->st_host->done_seq != kRunning) break;
        batch = std::min(batch * 2, 64);
    }
    FL_HIP_TRY(ctx, hipEventRecord(m->e1, s));
    const LinState h = *m->st_host;
    if (h.status != FL_OK) {
        FL_HIP_TRY(ctx, hipStreamSynchronize(s));
        return state_error(m, h, d);
    }
    FL_HIP_TRY(ctx, hipMemcpyAsync(x_inout, x, bytes, out, s));
    FL_HIP_TRY(ctx, hipStreamSynchronize(s));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, m->e0, m->e1);
    if (rep) {
        rep->converged = h.converged;
        rep->iterations = h.converged ? (h.initial <= h.threshold ? 0 : h.iter) : h.max_iter;
        rep->initial_residual = h.initial;
        rep->final_residual = h.rnorm;
        rep->relative_residual = h.initial == 0.0 ? 0.0 : h.rnorm / h.initial;
        rep->device_ms = ms;
        rep->refinements = 0;
        rep->fallback = 0;
    }
    if (d) set_detail(d, FL_OK, "");
    return FL_OK;
}

fl_status solve_mixed(fl_csr *m, int kind, const double *rhs, double *x_inout, const fl_lin_opts *opts,
                      fl_lin_report *rep, fl_err_detail *d);

// The library-allocated path: work vectors allocated on first use and kept
// with the matrix.
fl_status solve(fl_csr *m, int kind, const double *rhs, double *x_inout, const fl_lin_opts *opts,
                fl_lin_report *rep, fl_err_detail *d) {
    fl_ctx *ctx = m->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    fl_status st = check_solve(m, kind, rhs, x_inout, opts, d);
    if (st != FL_OK) return st;
    if (m->precision == FL_PRECISION_MIXED) return solve_mixed(m, kind, rhs, x_inout, opts, rep, d);
    if ((st = ensure_vectors(m, kind == KIND_BICG ? 7 : 5)) != FL_OK) return st;
    Work w{};
    for (int k = 0; k < 7; ++k) w.vec[k] = m->vec[k];
    w.partials = m->partials;
    w.st = m->st;
    return solve_on(m, kind, rhs, x_inout, opts, rep, d, w);
}

// FP32 copies of the values and the Jacobi inverse, and the FP32 vectors.
fl_status ensure_mixed(fl_csr *m) {
    fl_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    const unsigned G = m->blocks;
    if (!m->val32) {
        FL_HIP_TRY(ctx, hipMalloc(&m->val32, (size_t)std::max<int64_t>(m->nnz, 1) * sizeof(float)));
        k_to_f32<<<G, kBlock, 0, s>>>(m->nnz, m->val, m->val32);
        FL_HIP_TRY(ctx, hipGetLastError());
    }
    if (!m->inv32) {
        FL_HIP_TRY(ctx, hipMalloc(&m->inv32, (size_t)std::max<uint32_t>(m->n, 1) * sizeof(float)));
        k_to_f32<<<G, kBlock, 0, s>>>(m->n, m->inv, m->inv32);
        FL_HIP_TRY(ctx, hipGetLastError());
    }
    for (float *&v : m->vec32)
        if (!v) FL_HIP_TRY(ctx, hipMalloc(&v, (size_t)m->n * sizeof(float) + 16));
    if (!m->os) FL_HIP_TRY(ctx, hipMalloc(&m->os, sizeof(OuterState)));
    if (!m->os_host && hipHostMalloc(&m->os_host, sizeof(OuterState), hipHostMallocDefault) != hipSuccess)
        return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
    return FL_OK;
}

// The FP32 solver inside FP64 iterative refinement (see OuterState).  The
// report counts the inner iterations of every round; an inner solve that
// breaks down, or a round that does not lower the FP64 residual, hands the
// solve over to the FP64 solver from the current x (fallback = 1) with the
// same threshold and the rest of the iteration budget.
fl_status solve_mixed(fl_csr *m, int kind, const double *rhs, double *x_inout, const fl_lin_opts *opts,
                      fl_lin_report *rep, fl_err_detail *d) {
    fl_ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    fl_status st = ensure_vectors(m, 3);
    if (st == FL_OK) st = ensure_mixed(m);
    if (st != FL_OK) return st;
    const size_t bytes = (size_t)m->n * sizeof(double);
    double *x = m->vec[0], *b = m->vec[1], *q64 = m->vec[2];
    float *r = m->vec32[0], *dp = m->vec32[1], *qv = m->vec32[2], *e = m->vec32[3], *tv = m->vec32[4],
          *rhat = m->vec32[5];
    FL_HIP_TRY(ctx, hipMemcpyAsync(x, x_inout, bytes, hipMemcpyHostToDevice, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(b, rhs, bytes, hipMemcpyHostToDevice, s));
    OuterState o{};
    o.atol = opts->absolute_tolerance;
    o.rtol = opts->relative_tolerance;
    *m->os_host = o;
    FL_HIP_TRY(ctx, hipMemcpyAsync(m->os, m->os_host, sizeof(OuterState), hipMemcpyHostToDevice, s));
    const Mat A{m->off, m->col, m->val, m->n};
    const MatT<float> A32{m->off, m->col, m->val32, m->n};
    const float *inv = kind == KIND_PCG ? m->inv32 : nullptr;
    const unsigned G = m->blocks;
    double *const pa = m->partials, *const pb = m->partials + (size_t)G * 4;
    LinState *const dst = m->st;
    const long long budget = opts->max_iterations;
    long long used = 0;
    bool fallback = false;
    FL_HIP_TRY(ctx, hipEventRecord(m->e0, s));
    for (;;) {
        LinState init{};
        init.in_scale = 1.0;
        init.max_iter = std::max<long long>(budget - used, 0);
        init.done_seq = kRunning;
        init.bad_row = init.bad_row2 = init.bad_r = init.bad_z = ~0ull;
        *m->st_host = init;
        FL_HIP_TRY(ctx, hipMemcpyAsync(dst, m->st_host, sizeof(LinState), hipMemcpyHostToDevice, s));
        k_spmv<<<G, kBlock, 0, s>>>(A, x, q64, dst);
        k_res_mixed<<<G, kBlock, 0, s>>>(m->n, b, q64, r, e, pb, dst);
        k_outer_test<<<1, kBlock, 0, s>>>(pb, G, m->os, dst, m->inner_rel, used >= budget ? 1 : 0);
        long long seq = 2;
        if (kind == KIND_BICG) {
            k_bi_init<<<G, kBlock, 0, s>>>(m->n, r, rhat, dp, qv, pb, G, pa, dst, seq++,
                                           (float *)nullptr);
        } else {
            k_cg_init_d<<<G, kBlock, 0, s>>>(m->n, r, inv, dp, pb, G, pa, dst, seq++,
                                             (float *)nullptr);
            k_cg_init_rho<<<1, kBlock, 0, s>>>(pa, G, dst, seq++);
        }
        FL_HIP_TRY(ctx, hipGetLastError());
        int batch = 4;
        long long it = 1;
        for (;;) {
            FL_HIP_TRY(ctx, hipMemcpyAsync(m->st_host, dst, sizeof(LinState), hipMemcpyDeviceToHost, s));
            FL_HIP_TRY(ctx, hipStreamSynchronize(s));
            if (m->st_host->done_seq != kRunning) break;
            for (int k = 0; k < batch; ++k, ++it) {
                if (kind == KIND_BICG) {
                    k_bi_p<<<G, kBlock, 0, s>>>(m->n, (const float *)r, dp, qv, pa, G, dst, it, seq++);
                    k_bi_spmv_v<<<G, kBlock, 0, s>>>(A32, dp, rhat, qv, pa, dst, seq++);
                    k_bi_spmv_t<<<G, kBlock, 0, s>>>(A32, r, qv, tv, pa, G, pb, dst, it, seq++);
                    k_bi_update<<<G, kBlock, 0, s>>>(m->n, e, r, dp, qv, tv, rhat, pb, G, pa, dst, it, seq++);
                } else {
                    k_cg_spmv<<<G, kBlock, 0, s>>>(A32, dp, qv, pa, dst, seq++);
                    k_cg_update<<<G, kBlock, 0, s>>>(m->n, e, r, dp, qv, inv, pa, G, pb, dst, it, seq++);
                    k_cg_dir<<<G, kBlock, 0, s>>>(m->n, (const float *)r, inv, dp, pb, G, dst, it, seq++);
                }
            }
            FL_HIP_TRY(ctx, hipGetLastError());
            batch = std::min(batch * 2, 64);
        }
        FL_HIP_TRY(ctx, hipMemcpyAsync(m->os_host, m->os, sizeof(OuterState), hipMemcpyDeviceToHost, s));
        FL_HIP_TRY(ctx, hipStreamSynchronize(s));
        const LinState h = *m->st_host;
        const OuterState oh = *m->os_host;
        if (oh.converged || oh.exhausted) break;
        if (oh.stalled) {
            fallback = true;
            break;
        }
        if (h.status != FL_OK) {
            // the FP64 residual's own faults are the caller's errors (the
            // initial tests, launch 2); an FP32 breakdown after the inner
            // residual has fallen keeps its correction and restarts with a
            // new round; anything else hands over to FP64
            if (h.done_seq == 2) return state_error(m, h, d);
            if (!(h.status == FL_ERR_BREAKDOWN && h.rnorm < h.initial)) {
                fallback = true;
                break;
            }
            used += std::max<long long>(h.iter, 1);
        } else {
            used += h.converged ? h.iter : h.max_iter;
        }
        k_add_e<<<G, kBlock, 0, s>>>(m->n, x, e, m->os);
        FL_HIP_TRY(ctx, hipGetLastError());
    }
    FL_HIP_TRY(ctx, hipEventRecord(m->e1, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(x_inout, x, bytes, hipMemcpyDeviceToHost, s));
    FL_HIP_TRY(ctx, hipStreamSynchronize(s));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, m->e0, m->e1);
    const OuterState oh = *m->os_host;
    if (fallback) {
        // FP64 from the current x, to the original threshold
        fl_lin_opts o64{oh.threshold, 0.0, std::max<long long>(budget - used, 0)};
        fl_lin_report r64{};
        if ((st = ensure_vectors(m, kind == KIND_BICG ? 7 : 5)) != FL_OK) return st;
        Work w{};
        for (int k = 0; k < 7; ++k) w.vec[k] = m->vec[k];
        w.partials = m->partials;
        w.st = m->st;
        if ((st = solve_on(m, kind, rhs, x_inout, &o64, &r64, d, w)) != FL_OK) return st;
        if (rep) {
            rep->converged = r64.converged;
            rep->iterations = used + r64.iterations;
            rep->initial_residual = oh.initial;
            rep->final_residual = r64.final_residual;
            rep->relative_residual = oh.initial == 0.0 ? 0.0 : r64.final_residual / oh.initial;
            rep->device_ms = ms + r64.device_ms;
            rep->refinements = (int32_t)oh.round;
            rep->fallback = 1;
        }
        return FL_OK;
    }
    if (rep) {
        rep->converged = oh.converged;
        rep->iterations = oh.converged ? used : budget;
        rep->initial_residual = oh.initial;
        rep->final_residual = oh.rnorm;
        rep->relative_residual = oh.initial == 0.0 ? 0.0 : oh.rnorm / oh.initial;
        rep->device_ms = ms;
        rep->refinements = (int32_t)oh.round;
        rep->fallback = 0;
    }
    if (d) set_detail(d, FL_OK, "");
    return FL_OK;
}

}  // namespace

namespace {
// JacobiPreconditioner::new (preconditioner.rs:9-28) on the device: the
// inverse diagonal of every row and, in `first`, the smallest failing row
// packed with its kind (row << 2 | 1 missing, 2 zero, 3 non-finite inverse).
__global__ void k_jacobi_inverse(uint32_t n, const int *__restrict__ off, const int *__restrict__ col,
                                 const double *__restrict__ val, double *__restrict__ inv,
                                 unsigned long long *first) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        int found = -1;
        for (int e = off[r]; e < off[r + 1]; ++e)
            if (col[e] == (int)r) found = e;
        unsigned code = 0;
        if (found < 0) {
            code = 1;
        } else if (val[found] == 0.0) {
            code = 2;
        } else {
            inv[r] = 1.0 / val[found];
            if (!isfinite(inv[r])) code = 3;
        }
        if (code) atomicMin(first, ((unsigned long long)r << 2) | code);
    }
}

}  // namespace

namespace {
// JacobiPreconditioner::new on the matrix's current values: the inverse
// diagonal and the outcome (missing, zero or non-finite row) in inv_status.
fl_status refresh_jacobi(fl_csr *m, unsigned long long *first) {
    hipStream_t s = m->ctx->stream;
    unsigned long long h = ~0ull;
    FL_HIP_TRY(m->ctx, hipMemcpyAsync(first, &h, sizeof h, hipMemcpyHostToDevice, s));
    if (m->n > 0) k_jacobi_inverse<<<m->blocks, kBlock, 0, s>>>(m->n, m->off, m->col, m->val, m->inv, first);
    FL_HIP_TRY(m->ctx, hipGetLastError());
    FL_HIP_TRY(m->ctx, hipMemcpyAsync(&h, first, sizeof h, hipMemcpyDeviceToHost, s));
    FL_HIP_TRY(m->ctx, hipStreamSynchronize(s));
    m->inv_status = FL_OK;
    m->inv_row = 0;
    m->inv_ctx = "";
    if (h != ~0ull) {
        m->inv_row = (int64_t)(h >> 2);
        switch (h & 3) {
            case 1: m->inv_status = FL_ERR_MISSING_DIAGONAL; m->inv_ctx = "missing diagonal entry"; break;
            case 2: m->inv_status = FL_ERR_ZERO_DIAGONAL; m->inv_ctx = "zero diagonal entry"; break;
            default: m->inv_status = FL_ERR_NON_FINITE; m->inv_ctx = "inverse diagonal"; break;
        }
    }
    return FL_OK;
}
}  // namespace

namespace fl {
double *csr_values(fl_csr *m) { return m->val; }

fl_status csr_refresh_jacobi(fl_csr *m) {
    if (!m->first) FL_HIP_TRY(m->ctx, hipMalloc(&m->first, sizeof(unsigned long long)));
    return refresh_jacobi(m, m->first);
}

fl_status csr_solve_device(fl_csr *m, int kind, const double *rhs, double *x_inout, const fl_lin_opts *opts,
                           fl_lin_report *rep, fl_err_detail *d) {
    fl_ctx *ctx = m->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (!rhs || !x_inout || !opts) return err(ctx, d, FL_ERR_INVALID_ARGUMENT, "null argument");
    if (!std::isfinite(opts->absolute_tolerance) || opts->absolute_tolerance < 0.0 ||
        !std::isfinite(opts->relative_tolerance) || opts->relative_tolerance < 0.0)
        return err(ctx, d, FL_ERR_BREAKDOWN, "invalid solver tolerance");
    if (kind == KIND_PCG && m->inv_status != FL_OK) {
        set_detail(d, m->inv_status, m->inv_ctx, m->inv_row, 0, 0, m->inv_row);
        return set_error(ctx, (fl_status)m->inv_status, "Jacobi preconditioner: %s (row %lld)", m->inv_ctx,
                         (long long)m->inv_row);
    }
    fl_status st = ensure_vectors(m, kind == KIND_BICG ? 7 : 5);
    if (st != FL_OK) return st;
    Work w{};
    for (int k = 0; k < 7; ++k) w.vec[k] = m->vec[k];
    w.partials = m->partials;
    w.st = m->st;
    return solve_on(m, kind, rhs, x_inout, opts, rep, d, w, true);
}

fl_status csr_adopt(fl_ctx *ctx, uint32_t n, uint32_t ncols, int64_t nnz, int *off, int *col, double *val,
                    fl_csr **out) {
    *out = nullptr;
    fl_csr *m = new (std::nothrow) fl_csr();
    if (!m) {
        (void)hipFree(off);
        (void)hipFree(col);
        (void)hipFree(val);
        return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    m->ctx = ctx;
    m->n = n;
    m->ncols = ncols;
    m->nnz = nnz;
    m->off = off;
    m->col = col;
    m->val = val;
    const int64_t tiles = (n + kBlock - 1) / kBlock;
    m->blocks = grid_blocks(ctx, tiles > 0 ? tiles : 1);
    hipStream_t s = ctx->stream;
    unsigned long long *first = nullptr;
    bool okay = hipMalloc(&m->inv, (n > 0 ? n : 1) * sizeof(double)) == hipSuccess &&
                hipMalloc(&m->partials, (size_t)m->blocks * 8 * sizeof(double)) == hipSuccess &&
                hipMalloc(&m->st, sizeof(LinState)) == hipSuccess &&
                hipHostMalloc(&m->st_host, sizeof(LinState), hipHostMallocDefault) == hipSuccess &&
                hipEventCreate(&m->e0) == hipSuccess && hipEventCreate(&m->e1) == hipSuccess &&
                hipMalloc(&first, sizeof(unsigned long long)) == hipSuccess;
    okay = okay && hipMemsetAsync(m->st, 0, sizeof(LinState), s) == hipSuccess;
    okay = okay && refresh_jacobi(m, first) == FL_OK;
    if (first) (void)hipFree(first);
    if (!okay) {
        free_csr(m);
        delete m;
        return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "device allocation failed");
    }
    ctx_retain(ctx);
    *out = m;
    return FL_OK;
}
}  // namespace fl

extern "C" {

fl_status fl_csr_upload(fl_ctx *ctx, int64_t nrows, int64_t ncols, const int64_t *off,
                        const int64_t *cols, const double *vals, int64_t nnz, fl_csr **out,
                        fl_err_detail *d) {
    if (!ctx || !out || !off || (nnz > 0 && (!cols || !vals)) || nrows < 0 || ncols < 0 || nnz < 0)
        return err(ctx, d, FL_ERR_INVALID_ARGUMENT, "null or negative argument");
    *out = nullptr;
    // validate_csr (csr.rs:195-251), in the reference's order
    if (off[0] != 0) {
        set_detail(d, FL_ERR_INVALID_CSR, "first row offset must be zero");
        return set_error(ctx, FL_ERR_INVALID_CSR, "invalid CSR matrix: first row offset must be zero");
    }
    if (off[nrows] != nnz) {
        set_detail(d, FL_ERR_INVALID_CSR, "last row offset must equal value count");
        return set_error(ctx, FL_ERR_INVALID_CSR,
                         "invalid CSR matrix: last row offset must equal value count");
    }
    for (int64_t r = 0; r < nrows; ++r) {
        if (off[r] > off[r + 1]) {
            set_detail(d, FL_ERR_INVALID_CSR, "row offsets must be monotonic");
            return set_error(ctx, FL_ERR_INVALID_CSR, "invalid CSR matrix: row offsets must be monotonic");
        }
        for (int64_t e = off[r] + 1; e < off[r + 1]; ++e)
            if (cols[e - 1] >= cols[e]) {
                set_detail(d, FL_ERR_INVALID_CSR, "row columns must be strictly ascending");
                return set_error(ctx, FL_ERR_INVALID_CSR,
                                 "invalid CSR matrix: row columns must be strictly ascending");
            }
    }
    for (int64_t e = 0; e < nnz; ++e)
        if (cols[e] < 0 || cols[e] >= ncols) {
            set_detail(d, FL_ERR_INVALID_CSR, "column exceeds ncols", e);
            return set_error(ctx, FL_ERR_INVALID_CSR,
                             "invalid CSR matrix: column %lld at entry %lld exceeds ncols",
                             (long long)cols[e], (long long)e);
        }
    for (int64_t e = 0; e < nnz; ++e)
        if (!std::isfinite(vals[e])) return err(ctx, d, FL_ERR_NON_FINITE, "matrix", e);
    if (nnz >= (int64_t)1 << 31 || nrows >= (int64_t)1 << 31 || ncols >= (int64_t)1 << 31)
        return err(ctx, d, FL_ERR_INVALID_ARGUMENT, "matrix too large for 32-bit device indices");
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    fl_csr *m = new (std::nothrow) fl_csr();
    if (!m) return err(ctx, d, FL_ERR_OUT_OF_MEMORY, "host allocation failed");
    m->ctx = ctx;
    m->n = (uint32_t)nrows;
    m->ncols = (uint32_t)ncols;
    m->nnz = nnz;
    std::vector<int> ho(nrows + 1), hc(nnz > 0 ? nnz : 1);
    for (int64_t r = 0; r <= nrows; ++r) ho[r] = (int)off[r];
    for (int64_t e = 0; e < nnz; ++e) hc[e] = (int)cols[e];
    // JacobiPreconditioner::new (preconditioner.rs:9-28): the first failing row wins.
    std::vector<double> inv(nrows > 0 ? nrows : 1);
    for (int64_t r = 0; r < nrows && m->inv_status == FL_OK; ++r) {
        int64_t found = -1;
        for (int64_t e = off[r]; e < off[r + 1]; ++e)
            if (cols[e] == r) found = e;
        if (found < 0) {
            m->inv_status = FL_ERR_MISSING_DIAGONAL;
            m->inv_ctx = "missing diagonal entry";
            m->inv_row = r;
        } else if (vals[found] == 0.0) {
            m->inv_status = FL_ERR_ZERO_DIAGONAL;
            m->inv_ctx = "zero diagonal entry";
            m->inv_row = r;
        } else {
            inv[r] = 1.0 / vals[found];
            if (!std::isfinite(inv[r])) {
                m->inv_status = FL_ERR_NON_FINITE;
                m->inv_ctx = "inverse diagonal";
                m->inv_row = r;
            }
        }
    }
    const int64_t tiles = (nrows + kBlock - 1) / kBlock;
    m->blocks = grid_blocks(ctx, tiles > 0 ? tiles : 1);
    hipStream_t s = ctx->stream;
    bool okay = hipMalloc(&m->off, (nrows + 1) * sizeof(int)) == hipSuccess &&
                hipMalloc(&m->col, (nnz > 0 ? nnz : 1) * sizeof(int)) == hipSuccess &&
                hipMalloc(&m->val, (nnz > 0 ? nnz : 1) * sizeof(double)) == hipSuccess &&
                hipMalloc(&m->inv, (nrows > 0 ? nrows : 1) * sizeof(double)) == hipSuccess &&
                hipMalloc(&m->partials, (size_t)m->blocks * 8 * sizeof(double)) == hipSuccess &&
                hipMalloc(&m->st, sizeof(LinState)) == hipSuccess &&
                hipHostMalloc(&m->st_host, sizeof(LinState), hipHostMallocDefault) == hipSuccess &&
                hipEventCreate(&m->e0) == hipSuccess && hipEventCreate(&m->e1) == hipSuccess;
    okay = okay &&
           hipMemcpyAsync(m->off, ho.data(), (nrows + 1) * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess &&
           (nnz == 0 ||
            (hipMemcpyAsync(m->col, hc.data(), nnz * sizeof(int), hipMemcpyHostToDevice, s) == hipSuccess &&
             hipMemcpyAsync(m->val, vals, nnz * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess)) &&
           (nrows == 0 ||
            hipMemcpyAsync(m->inv, inv.data(), nrows * sizeof(double), hipMemcpyHostToDevice, s) == hipSuccess) &&
           hipMemsetAsync(m->st, 0, sizeof(LinState), s) == hipSuccess &&
           hipStreamSynchronize(s) == hipSuccess;
    if (!okay) {
        free_csr(m);
        delete m;
        return err(ctx, d, FL_ERR_OUT_OF_MEMORY, "device allocation or upload failed");
    }
    ctx_retain(ctx);
    *out = m;
    if (d) set_detail(d, FL_OK, "");
    return FL_OK;
}

fl_status fl_csr_set_precision(fl_csr *m, int32_t precision, double inner_tolerance) {
    if (!m) return FL_ERR_INVALID_ARGUMENT;
    if (precision != FL_PRECISION_FP64 && precision != FL_PRECISION_MIXED)
        return set_error(m->ctx, FL_ERR_INVALID_ARGUMENT, "precision must be FL_PRECISION_FP64 or FL_PRECISION_MIXED");
    if (!std::isfinite(inner_tolerance) || inner_tolerance < 0.0 || inner_tolerance >= 1.0)
        return set_error(m->ctx, FL_ERR_INVALID_ARGUMENT, "inner tolerance must lie in [0, 1)");
    m->precision = precision;
    m->inner_rel = inner_tolerance > 0.0 ? inner_tolerance : 1e-5;
    return FL_OK;
}

void fl_csr_destroy(fl_csr *m) {
    if (!m) return;
    fl_ctx *ctx = m->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    free_csr(m);
    delete m;
    ctx_release(ctx);
}

fl_status fl_spmv(fl_csr *m, const double *x, double *y, fl_err_detail *d) {
    if (!m || !x || !y) return err(m ? m->ctx : nullptr, d, FL_ERR_INVALID_ARGUMENT, "null argument");
    fl_ctx *ctx = m->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    fl_status st = ensure_vectors(m, 2);
    if (st != FL_OK) return st;
    hipStream_t s = ctx->stream;
    FL_HIP_TRY(ctx, hipMemcpyAsync(m->vec[0], x, (size_t)m->ncols * sizeof(double),
                                   hipMemcpyHostToDevice, s));
    LinState init{};
    init.bad_row = ~0ull;
    *m->st_host = init;
    FL_HIP_TRY(ctx, hipMemcpyAsync(m->st, m->st_host, sizeof(LinState), hipMemcpyHostToDevice, s));
    if (m->ncols > m->n) {
        // x is longer than the vector buffers when ncols > nrows
        double *big = nullptr;
        FL_HIP_TRY(ctx, hipMalloc(&big, (size_t)m->ncols * sizeof(double)));
        FL_HIP_TRY(ctx, hipMemcpyAsync(big, x, (size_t)m->ncols * sizeof(double), hipMemcpyHostToDevice, s));
        k_spmv<<<m->blocks, kBlock, 0, s>>>(Mat{m->off, m->col, m->val, m->n}, big, m->vec[1], m->st);
        FL_HIP_TRY(ctx, hipStreamSynchronize(s));
        (void)hipFree(big);
    } else {
        k_spmv<<<m->blocks, kBlock, 0, s>>>(Mat{m->off, m->col, m->val, m->n}, m->vec[0], m->vec[1], m->st);
    }
    FL_HIP_TRY(ctx, hipGetLastError());
    FL_HIP_TRY(ctx, hipMemcpyAsync(y, m->vec[1], (size_t)m->n * sizeof(double), hipMemcpyDeviceToHost, s));
    FL_HIP_TRY(ctx, hipMemcpyAsync(m->st_host, m->st, sizeof(LinState), hipMemcpyDeviceToHost, s));
    FL_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (m->st_host->bad_row != ~0ull)
        return err(ctx, d, FL_ERR_NON_FINITE, "SpMV output", (int64_t)m->st_host->bad_row);
    if (d) set_detail(d, FL_OK, "");
    return FL_OK;
}

fl_status fl_cg(fl_csr *m, const double *rhs, double *x, const fl_lin_opts *o, fl_lin_report *rep,
                fl_err_detail *d) {
    if (!m) return FL_ERR_INVALID_ARGUMENT;
    return solve(m, KIND_CG, rhs, x, o, rep, d);
}

fl_status fl_pcg_jacobi(fl_csr *m, const double *rhs, double *x, const fl_lin_opts *o,
                        fl_lin_report *rep, fl_err_detail *d) {
    if (!m) return FL_ERR_INVALID_ARGUMENT;
    return solve(m, KIND_PCG, rhs, x, o, rep, d);
}

fl_status fl_bicgstab(fl_csr *m, const double *rhs, double *x, const fl_lin_opts *o,
                      fl_lin_report *rep, fl_err_detail *d) {
    if (!m) return FL_ERR_INVALID_ARGUMENT;
    return solve(m, KIND_BICG, rhs, x, o, rep, d);
}

fl_status fl_pcg_jacobi_noalloc(fl_csr *m, const double *rhs, double *x, const fl_lin_opts *o,
                                fl_lin_report *rep, fl_err_detail *d) {
    if (!m) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = m->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    if ((size_t)m->n > kNoAllocMaxRows || (size_t)m->ncols > kNoAllocMaxRows)
        return err(ctx, d, FL_ERR_INVALID_ARGUMENT,
                   "matrix exceeds the allocation-free solve's limit of 1048576 rows and columns",
                   (int64_t)std::max(m->n, m->ncols));
    if (m->blocks > kNoAllocMaxBlocks)
        return err(ctx, d, FL_ERR_INVALID_ARGUMENT,
                   "launch needs more partial rows than the allocation-free buffers hold",
                   (int64_t)m->blocks);
    fl_status st = check_solve(m, KIND_PCG, rhs, x, o, d);
    if (st != FL_OK) return st;
    double *base = nullptr;
    const hipError_t e = scratch_address(ctx->device, &base);
    if (e != hipSuccess)
        return set_error(ctx, FL_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    Work w{};
    for (int k = 0; k < 5; ++k) w.vec[k] = base + k * kNoAllocStride;
    w.vec[5] = w.vec[6] = nullptr;  // BiCGSTAB only
    w.partials = base + 5 * kNoAllocStride;
    w.st = reinterpret_cast<LinState *>(w.partials + kNoAllocPartials);
    return solve_on(m, KIND_PCG, rhs, x, o, rep, d, w);
}

fl_status fl_csr_info(fl_csr *m, int64_t *nrows, int64_t *ncols, int64_t *nnz) {
    if (!m) return FL_ERR_INVALID_ARGUMENT;
    if (nrows) *nrows = m->n;
    if (ncols) *ncols = m->ncols;
    if (nnz) *nnz = m->nnz;
    return FL_OK;
}

fl_status fl_csr_download(fl_csr *m, int64_t *off, int64_t *cols, double *vals) {
    if (!m) return FL_ERR_INVALID_ARGUMENT;
    fl_ctx *ctx = m->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::vector<int> ho(m->n + 1), hc(m->nnz > 0 ? m->nnz : 1);
    FL_HIP_TRY(ctx, hipMemcpy(ho.data(), m->off, (m->n + 1) * sizeof(int), hipMemcpyDeviceToHost));
    if (m->nnz > 0)
        FL_HIP_TRY(ctx, hipMemcpy(hc.data(), m->col, m->nnz * sizeof(int), hipMemcpyDeviceToHost));
    if (off)
        for (uint32_t r = 0; r <= m->n; ++r) off[r] = ho[r];
    if (cols)
        for (int64_t e = 0; e < m->nnz; ++e) cols[e] = hc[e];
    if (vals && m->nnz > 0)
        FL_HIP_TRY(ctx, hipMemcpy(vals, m->val, m->nnz * sizeof(double), hipMemcpyDeviceToHost));
    return FL_OK;
}

}  // extern "C"
