// Shared plumbing of the gfx950 library: the context, error reporting,
// wave/block reductions and the deterministic one-pass grid reduction.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "flursys_hip.h"

struct fl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 256;
    std::string last_error;
    // Objects created from this context hold a reference; fl_destroy only
    // marks the context closed while any of them is alive, and the last one
    // to go frees it.
    int refs = 0;
    bool closed = false;
};

namespace fl {

void ctx_release(fl_ctx *ctx);  // drops one reference (context.hip)
void noalloc_init(int device);   // per-device address of the allocation-free solve buffers (csr.hip)
// Wraps device CSR arrays (int32 offsets and columns, FP64 values; sorted,
// duplicate-free, finite) in a solver matrix, taking ownership (csr.hip).
fl_status csr_adopt(fl_ctx *ctx, uint32_t n, uint32_t ncols, int64_t nnz, int *off, int *col,
                    double *val, fl_csr **out);
// The matrix's device values (same sparsity, rewritten by the caller's kernels).
double *csr_values(fl_csr *m);
// Recomputes the Jacobi inverse diagonal after the values changed (csr.hip).
fl_status csr_refresh_jacobi(fl_csr *m);
// A CG (kind 0), Jacobi-PCG (1) or BiCGSTAB (2) solve on device vectors: rhs
// and x_inout are device arrays of n doubles, produced by the library's own
// kernels (their finiteness is not re-checked on the host) (csr.hip).
fl_status csr_solve_device(fl_csr *m, int kind, const double *rhs, double *x_inout, const fl_lin_opts *opts,
                           fl_lin_report *rep, fl_err_detail *d);
inline void ctx_retain(fl_ctx *ctx) { ctx->refs += 1; }

constexpr int kBlock = 256;   // threads per workgroup: 4 wave64s
constexpr int kWaves = kBlock / 64;
constexpr int kXcds = 8;      // MI355X: 8 XCDs, workgroup b runs on XCD b % 8

inline fl_status set_error(fl_ctx *ctx, fl_status st, const char *fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return st;
}

#define FL_HIP_TRY(ctx, call)                                                              \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::fl::set_error((ctx), FL_ERR_HIP, "%s failed: %s", #call,              \
                                   hipGetErrorString(e_));                                 \
    } while (0)

// ---------------------------------------------------------------------------
// Reductions.  Sums use a fixed tree (lane butterfly, then waves in order,
// then blocks in index order), so a run is bitwise repeatable for a given
// launch shape.  Maxima use fmax, which ignores a NaN operand exactly like
// Rust's f64::max in the reference's parallel_max_abs
// (src/solver/incompressible.rs:1313-1320).
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Block-wide reduction of NS sums and NM maxima.  Result valid in thread 0.
template <int NS, int NM>
__device__ inline void block_reduce(double (&s)[NS], double (&m)[NM]) {
    __shared__ double sh[kWaves][NS + NM + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < NS; ++a) s[a] = wave_sum(s[a]);
#pragma unroll
    for (int a = 0; a < NM; ++a) m[a] = wave_max(m[a]);
    if (lane == 0) {
#pragma unroll
        for (int a = 0; a < NS; ++a) sh[wave][a] = s[a];
#pragma unroll
        for (int a = 0; a < NM; ++a) sh[wave][NS + a] = m[a];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < NS; ++a) {
            double t = sh[0][a];
            for (int w = 1; w < kWaves; ++w) t += sh[w][a];
            s[a] = t;
        }
#pragma unroll
        for (int a = 0; a < NM; ++a) {
            double t = sh[0][NS + a];
            for (int w = 1; w < kWaves; ++w) t = fmax(t, sh[w][NS + a]);
            m[a] = t;
        }
    }
}

// Reductions across a grid without fences or atomics: each block publishes
// its partial row (publish_partials: `partials` holds gridDim.x * (NS + NM)
// doubles of a 1D grid), and every block of the next kernel folds all rows
// in block order (fold_partials), so every block derives the same totals.
template <int NS, int NM>
__device__ inline void publish_partials(double (&s)[NS], double (&m)[NM], double *partials) {
    block_reduce<NS, NM>(s, m);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < NS; ++a) partials[blockIdx.x * (NS + NM) + a] = s[a];
#pragma unroll
        for (int a = 0; a < NM; ++a) partials[blockIdx.x * (NS + NM) + NS + a] = m[a];
    }
}

template <int NS, int NM>
__device__ inline void fold_partials(const double *__restrict__ partials, uint32_t rows,
                                     double (&s)[NS], double (&m)[NM]) {
    __shared__ double out[NS + NM];
#pragma unroll
    for (int a = 0; a < NS; ++a) s[a] = 0.0;
#pragma unroll
    for (int a = 0; a < NM; ++a) m[a] = 0.0;
    // rows b = t, t + blockDim, ... summed in order; the loads of kU rows are
    // issued together (a row past the end loads row 0 and is not added), so
    // the fold waits for memory once per kU rows instead of once per row
    constexpr int kU = 8, R = NS + NM;
    const uint32_t bd = blockDim.x;
    for (uint32_t b = threadIdx.x; b < rows; b += kU * bd) {
        double v[kU][R];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const uint32_t bb = b + u * bd;
            const size_t base = (size_t)(bb < rows ? bb : 0u) * R;
#pragma unroll
            for (int a = 0; a < R; ++a) v[u][a] = partials[base + a];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const bool ok = b + u * bd < rows;
#pragma unroll
            for (int a = 0; a < NS; ++a) s[a] = ok ? s[a] + v[u][a] : s[a];
#pragma unroll
            for (int a = 0; a < NM; ++a) m[a] = ok ? fmax(m[a], v[u][NS + a]) : m[a];
        }
    }
    block_reduce<NS, NM>(s, m);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < NS; ++a) out[a] = s[a];
#pragma unroll
        for (int a = 0; a < NM; ++a) out[NS + a] = m[a];
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < NS; ++a) s[a] = out[a];
#pragma unroll
    for (int a = 0; a < NM; ++a) m[a] = out[NS + a];
    __syncthreads();
}

// Unsigned 32-bit division by a run-time constant d >= 2 with a multiply-
// high and shifts (the round-down "branchfree" magic-number method of
// Granlund & Montgomery / libdivide): q = (t + ((n - t) >> 1)) >> s with
// t = mulhi(n, m), l = ceil(log2 d), m = 2^32 (2^l - d) / d + 1, s = l - 1.
struct FastDiv {
    uint32_t d = 2, m = 1, s = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t divisor) : d(divisor) {
        uint32_t l = 0;
        while ((1ull << l) < divisor) ++l;
        m = (uint32_t)(((1ull << 32) * ((1ull << l) - divisor)) / divisor + 1);
        s = l - 1;
    }
    __host__ __device__ inline uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
        const uint32_t t = __umulhi(n, m);
#else
        const uint32_t t = (uint32_t)(((uint64_t)n * m) >> 32);
#endif
        return (t + ((n - t) >> 1)) >> s;
    }
};

// Tile scheduling.  Work is cut into 256-cell tiles in linear (x-fastest)
// order.  When the grid is a multiple of 8 blocks, each XCD (blocks b with
// b % 8 == x) owns a contiguous run of tiles -- a z-slab of the grid -- so the
// z+-1 neighbours a stencil reads are, apart from slab edges, in that XCD's
// own L2.
struct TileSched {
    uint32_t first, count, step;
    __device__ inline TileSched(uint32_t ntiles) {
        if ((gridDim.x % kXcds) == 0 && gridDim.x >= 8 * kXcds) {
            const uint32_t xcd = blockIdx.x % kXcds, lb = blockIdx.x / kXcds;
            const uint32_t per = gridDim.x / kXcds;
            const uint32_t lo = (uint32_t)(((uint64_t)ntiles * xcd) / kXcds);
            const uint32_t hi = (uint32_t)(((uint64_t)ntiles * (xcd + 1)) / kXcds);
            first = lo + lb;
            count = hi;
            step = per;
        } else {
            first = blockIdx.x;
            count = ntiles;
            step = gridDim.x;
        }
    }
};

// Element-wise form of the same idea for index spaces without tiles: XCD x
// owns the contiguous range [n x / 8, n (x + 1) / 8) and its blocks stride
// through it, so stencil neighbours (the next row or plane) are read by the
// same XCD and hit its L2:  for (i = r.first; i < r.end; i += r.step).
struct XcdRange {
    uint32_t first, end, step;
    __device__ inline explicit XcdRange(uint32_t n) {
        if ((gridDim.x % kXcds) == 0 && gridDim.x >= 8 * kXcds) {
            const uint32_t xcd = blockIdx.x % kXcds, lb = blockIdx.x / kXcds;
            const uint32_t begin = (uint32_t)(((uint64_t)n * xcd) / kXcds);
            end = (uint32_t)(((uint64_t)n * (xcd + 1)) / kXcds);
            first = begin + lb * blockDim.x + threadIdx.x;
            step = (gridDim.x / kXcds) * blockDim.x;
        } else {
            first = blockIdx.x * blockDim.x + threadIdx.x;
            end = n;
            step = gridDim.x * blockDim.x;
        }
    }
};

inline uint32_t grid_blocks(const fl_ctx *ctx, uint64_t ntiles, int per_cu = 8) {
    uint64_t g = (uint64_t)ctx->num_cu * per_cu;
    if (ntiles < g) g = ntiles;
    if (g >= 64) g -= g % kXcds;
    return (uint32_t)(g < 1 ? 1 : g);
}

}  // namespace fl
