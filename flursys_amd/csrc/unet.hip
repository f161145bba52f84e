// Learned warm start: a 3D U-Net on gfx950 matrix cores (BF16 operands,
// FP32 accumulation), with native training (backward pass and Adam).
//
// The reference has no learned component; this is the warm start of the
// learned flexible CG (the cavity solver's FL_PRECOND_LEARNED).  The network
// maps the normalised right-hand side of the Poisson problem, restricted to
// an n^3 grid, to the normalised solution on that grid:
//   level 0 (n^3)      enc 16 -> 16            dec [16 | up 32] -> 16 -> 16 -> head
//   level 1 (n/2)^3    enc 16 -> 32 -> 32      dec [32 | up 48] -> 32 -> 32
//   level 2 (n/4)^3    enc 32 -> 48 -> 48      dec [48 | up 64] -> 48 -> 48
//   level 3 (n/8)^3    48 -> 64 -> 64
// 3x3x3 convolutions with zero padding and ReLU, 2x2x2 mean pooling,
// nearest upsampling fused into the next convolution's input, no biases (a
// bias-free ReLU network is positively homogeneous, as the linear solution
// operator is).  The input is padded to 16 channels and the head is a 16-
// channel convolution of which only channel 0 is used, so every layer runs
// through the same MFMA kernel.
//
// Layout: activations channels-last, [sample][voxel][channel] in BF16, voxel
// x fastest.  Weights of a layer: [cout][k] with k = tap * cin + c (tap =
// 9 (dz+1) + 3 (dy+1) + (dx+1)), K padded to a multiple of 32 with zeros;
// FP32 master copy, BF16 copy for the forward pass, and a flipped,
// transposed BF16 copy [cin][tap' * cout + n] (tap' = 26 - tap) with which
// the same kernel computes the data gradient.
//
// MFMA: v_mfma_f32_16x16x32_bf16.  Lane l = 16 g + m holds A[m][8g .. 8g+7],
// B[8g .. 8g+7][m] and D[4g .. 4g+3][m].  In the convolution A rows are 16
// output voxels and A columns (and B rows) 32 consecutive k; B columns are 16
// output channels.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "unet.hpp"

namespace fl {
namespace un {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short i16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline bf16x8 ld8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }
__device__ inline void st8(bf16 *p, bf16x8 v) { *reinterpret_cast<bf16x8 *>(p) = v; }
__device__ inline bf16x8 zero8() {
    const i16x8 z = {0, 0, 0, 0, 0, 0, 0, 0};
    return __builtin_bit_cast(bf16x8, z);
}
// a where m > 0 (the ReLU derivative taken from the layer's output), else 0
__device__ inline bf16x8 mask8(bf16x8 a, bf16x8 m) {
    const i16x8 ai = __builtin_bit_cast(i16x8, a), mi = __builtin_bit_cast(i16x8, m);
    const i16x8 zero = {0, 0, 0, 0, 0, 0, 0, 0};
    const i16x8 keep = mi > zero;
    return __builtin_bit_cast(bf16x8, ai & keep);
}
__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// voxel index -> (x, y, z) on a 2^lx cube
__device__ inline void vdec(int v, int lx, int &x, int &y, int &z) {
    const int msk = (1 << lx) - 1;
    x = v & msk;
    y = (v >> lx) & msk;
    z = v >> (2 * lx);
}
__device__ inline int venc(int x, int y, int z, int lx) { return (z << lx | y) << lx | x; }

struct ConvArgs {
    const bf16 *x1, *x2, *mask, *w;
    bf16 *y1, *y2;
    float *part;     // K-split partial sums [ksplit][batch * nvox][cout] (FP32); ksplit 1: unused
    int lx;          // log2 of the resolution of x1 and of the output
    int c1, c2;      // channels of x1 (full resolution) and x2 (half, upsampled)
    int cin, kpad;   // c1 + c2, padded 27 cin
    int cout, ysplit;  // output channels; channels >= ysplit go to y2
    int relu;
    int nsplit;      // output-channel groups of NB blocks, one per workgroup
    int ksplit;      // K ranges, one per workgroup (> 1: partial sums, then k_conv_reduce)
    int batch;
    float *hout;     // inference head (k_conv_lds<R, 1, true>): output channel 0, FP32 [batch][voxel]
    float *hsum;     // and its per-workgroup sums [batch][tile]
};

// One 3x3x3 convolution (or, with flipped transposed weights and the ReLU
// mask of the layer's output, its data gradient) as an implicit GEMM.
// Workgroup = 4 waves, each wave 16 output voxels x NB blocks of 16 output
// channels over one K range; blockIdx.x = voxel tile + tiles * (channel
// group + nsplit * K range), blockIdx.y = sample.  The coarse levels have few
// voxels, so their output channels and K are split over workgroups until
// the chip is full (k_conv_reduce sums the K ranges in a fixed order).  The
// operands of the next K step are loaded while the current one multiplies.
template <int NB>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int b = blockIdx.y;
    const int n = 1 << a.lx;
    const long long nvox = 1ll << (3 * a.lx), nvox2 = nvox >> 3;
    const int tiles = (int)(nvox >> 6);
    const int vt = blockIdx.x % tiles, rest = blockIdx.x / tiles;
    const int ns = rest % a.nsplit, ks = rest / a.nsplit;
    const int vbase = (vt * 4 + wave) * 16;
    int x, y, z;
    vdec(vbase + m, a.lx, x, y, z);
    const bf16 *x1 = a.x1 + b * nvox * a.c1;
    const bf16 *mk = a.mask ? a.mask + b * nvox * a.c1 : nullptr;
    const bf16 *x2 = a.x2 ? a.x2 + b * nvox2 * a.c2 : nullptr;
    const int K = 27 * a.cin;
    const int steps = a.kpad >> 5, per = (steps + a.ksplit - 1) / a.ksplit;
    const int kb = ks * per * 32, ke = min(steps, (ks + 1) * per) * 32;
    f32x4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    // this lane's k = k0 + 8 g: tap and channel, advanced by 32 per step
    int c = kb + g * 8, tap = c / a.cin;
    c -= tap * a.cin;
    auto load_a = [&](int k0, int tp, int ch) -> bf16x8 {
        bf16x8 av = zero8();
        if (k0 + g * 8 < K) {
            const int dz = tp / 9 - 1, dy = (tp / 3) % 3 - 1, dx = tp % 3 - 1;
            const int xx = x + dx, yy = y + dy, zz = z + dz;
            if ((unsigned)xx < (unsigned)n && (unsigned)yy < (unsigned)n && (unsigned)zz < (unsigned)n) {
                if (ch < a.c1) {
                    const long long off = (long long)venc(xx, yy, zz, a.lx) * a.c1 + ch;
                    av = ld8(x1 + off);
                    if (mk) av = mask8(av, ld8(mk + off));
                } else {
                    const long long off =
                        (long long)venc(xx >> 1, yy >> 1, zz >> 1, a.lx - 1) * a.c2 + (ch - a.c1);
                    av = ld8(x2 + off);
                }
            }
        }
        return av;
    };
    const bf16 *wp = a.w + (long long)(ns * NB * 16 + m) * a.kpad + g * 8;
    if (kb < ke) {
        bf16x8 av = load_a(kb, tap, c), bv[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) bv[nb] = ld8(wp + (long long)nb * 16 * a.kpad + kb);
        for (int k0 = kb; k0 < ke; k0 += 32) {
            c += 32;
            while (c >= a.cin) {
                c -= a.cin;
                ++tap;
            }
            bf16x8 an = zero8(), bn[NB];
            const bool more = k0 + 32 < ke;
            if (more) an = load_a(k0 + 32, tap, c);
#pragma unroll
            for (int nb = 0; nb < NB; ++nb)
                bn[nb] = more ? ld8(wp + (long long)nb * 16 * a.kpad + k0 + 32) : zero8();
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[nb] = mfma(av, bv[nb], acc[nb]);
            av = an;
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) bv[nb] = bn[nb];
        }
    }
    const int yc2 = a.cout - a.ysplit;
    const long long rows = (long long)a.batch * nvox;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int ch = (ns * NB + nb) * 16 + m;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long long vo = b * nvox + vbase + 4 * g + j;
            float val = acc[nb][j];
            if (a.ksplit > 1) {
                a.part[((long long)ks * rows + vo) * a.cout + ch] = val;
                continue;
            }
            if (a.relu) val = fmaxf(val, 0.0f);
            if (ch < a.ysplit) a.y1[vo * a.ysplit + ch] = (bf16)val;
            else a.y2[vo * yc2 + (ch - a.ysplit)] = (bf16)val;
        }
    }
}

// The forward convolution of the fine levels with its operands staged in
// LDS.  A workgroup owns a 16 x R x 4 block of output voxels (wave w: plane
// z0 + w, rows y0 .. y0 + R - 1, one 16-voxel MFMA row each) and NB
// 16-channel output blocks.  Input channels are taken 16 at a time: per
// chunk the block stages (one batch of 16-byte global loads, on the coarse
// levels issued while the previous chunk multiplies, then the LDS stores) the
// (16 + 2) x (R + 2) x 6 halo of those channels (zero outside
// the domain; channels past c1 from the half-resolution x2, the decoder's
// nearest upsampling) as two 8-channel halves, and the chunk's weights as
// [tap][half][output channel] units, so the 16 lanes of an MFMA row read
// 256 contiguous bytes of either.  K runs over the 27 taps in pairs: lanes
// g = 0, 1 take tap 2 q, channels 0-7 / 8-15, lanes g = 2, 3 tap 2 q + 1
// (a 28th zero tap completes the last pair).  Each staged input value serves
// all 27 taps and NB output blocks, each weight fragment R rows.  The MFMA
// takes the weight fragment as A (rows: output channels) and the voxel row
// as B (columns: voxels), so a lane's result is four consecutive channels of
// one voxel, stored as 8 contiguous bytes.
//
// kHead (inference): the network's head.  Only output channel 0 is used, so
// the lanes that hold it store it compactly as FP32 (the BF16 value the full
// head would store) and the workgroup's sum of it goes to hsum[sample][tile];
// the arithmetic is the full head's.
template <int R, int NB, bool kHead = false>
__global__ void __launch_bounds__(256) k_conv_lds(ConvArgs a) {
    constexpr int TZ = 4, HX = 18, HY = R + 2, HZ = TZ + 2, NU = HX * HY * HZ;
    constexpr int UA = 2 * NU, UB = 28 * 2 * NB * 16, UT = UA + UB, ITER = (UT + 255) / 256;
    __shared__ bf16x8 lds[UT];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int b = blockIdx.y;
    const int n = 1 << a.lx;
    const long long nvox = 1ll << (3 * a.lx), nvox2 = nvox >> 3;
    const int tx = n >> 4, ty = n / R, tz = n / TZ;
    const int tiles = tx * ty * tz;
    const int tile = blockIdx.x % tiles, ns = blockIdx.x / tiles;
    const int x0 = (tile % tx) * 16, y0 = ((tile / tx) % ty) * R, z0 = (tile / (tx * ty)) * TZ;
    const bf16 *x1 = a.x1 + b * nvox * a.c1;
    const bf16 *x2 = a.x2 ? a.x2 + b * nvox2 * a.c2 : nullptr;
    const bf16 *w0 = a.w + (long long)(ns * NB * 16) * a.kpad;
    f32x4 acc[R][NB];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[r][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int h = g & 1;
    // one chunk of 16 input channels: its halo tile and weights into st
    bf16x8 st[ITER];
    auto stage = [&](int c0) {
#pragma unroll
        for (int i = 0; i < ITER; ++i) {
            const int u = threadIdx.x + 256 * i;
            bf16x8 val = zero8();
            if (u < UA) {
                const int hh = u / NU, v = u - hh * NU;
                const int hx = v % HX, hy = (v / HX) % HY, hz = v / (HX * HY);
                const int xx = x0 - 1 + hx, yy = y0 - 1 + hy, zz = z0 - 1 + hz;
                if ((unsigned)xx < (unsigned)n && (unsigned)yy < (unsigned)n && (unsigned)zz < (unsigned)n) {
                    const int ch = c0 + hh * 8;
                    if (ch < a.c1)
                        val = ld8(x1 + (long long)venc(xx, yy, zz, a.lx) * a.c1 + ch);
                    else
                        val = ld8(x2 + (long long)venc(xx >> 1, yy >> 1, zz >> 1, a.lx - 1) * a.c2 + (ch - a.c1));
                }
            } else if (u < UT) {
                const int e = u - UA;
                const int oc = e % (NB * 16), hh = (e / (NB * 16)) & 1, tap = e / (NB * 32);
                if (tap < 27) val = ld8(w0 + (long long)oc * a.kpad + (long long)tap * a.cin + c0 + hh * 8);
            }
            st[i] = val;
        }
    };
    // R == 1 (the coarse levels, a few workgroups per CU at most): the next
    // chunk's loads are in flight while this chunk multiplies.  R == 4 (the
    // finest level, one or two chunks) stages each chunk when it starts,
    // which keeps st out of the MFMA loop's registers (occupancy).
    constexpr bool kPipe = R == 1;
    if (kPipe) stage(0);
    for (int c0 = 0; c0 < a.cin; c0 += 16) {
        if (!kPipe) stage(c0);
        __syncthreads();  // the previous chunk's MFMA reads are done
#pragma unroll
        for (int i = 0; i < ITER; ++i) {
            const int u = threadIdx.x + 256 * i;
            if (u < UT) lds[u] = st[i];
        }
        __syncthreads();
        if (kPipe && c0 + 16 < a.cin) stage(c0 + 16);
#pragma unroll
        for (int q = 0; q < 14; ++q) {
            const int tap = 2 * q + (g >> 1);      // 27: the zero tap
            const int tp = tap < 27 ? tap : 26;   // (its input address: any valid one)
            const int dz = tp / 9 - 1, dy = (tp / 3) % 3 - 1, dx = tp % 3 - 1;
            const int base = h * NU + ((wave + 1 + dz) * HY + (1 + dy)) * HX + (m + 1 + dx);
            bf16x8 bv[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) bv[nb] = lds[UA + (tap * 2 + h) * (NB * 16) + nb * 16 + m];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const bf16x8 av = lds[base + r * HX];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) acc[r][nb] = mfma(bv[nb], av, acc[r][nb]);
            }
        }
    }
    // D[4g + j][m]: channels 4g .. 4g + 3 of the block of voxel (x0 + m, y0 + r, z0 + wave)
    if (kHead) {
        __shared__ float hs[4];
        float sum = 0.0f;
        if (g == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float val = acc[r][0][0];
                if (a.relu) val = fmaxf(val, 0.0f);
                val = (float)(bf16)val;
                a.hout[b * nvox + venc(x0 + m, y0 + r, z0 + wave, a.lx)] = val;
                sum += val;
            }
        }
        // lanes 0-15 of each wave hold the sums: fixed-order reduction
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
        if (lane == 0) hs[wave] = sum;
        __syncthreads();
        if (threadIdx.x == 0) a.hsum[(long long)b * tiles + tile] = (hs[0] + hs[1]) + (hs[2] + hs[3]);
        return;
    }
    typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int ch = (ns * NB + nb) * 16 + 4 * g;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const long long vo = b * nvox + venc(x0 + m, y0 + r, z0 + wave, a.lx);
            bf16x4 o;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float val = acc[r][nb][j];
                if (a.relu) val = fmaxf(val, 0.0f);
                o[j] = (bf16)val;
            }
            *reinterpret_cast<bf16x4 *>(a.y1 + vo * a.cout + ch) = o;
        }
    }
}

// Sum of a K-split convolution's partial sums in K-range order, ReLU, BF16.
__global__ void __launch_bounds__(256) k_conv_reduce(ConvArgs a, long long rows) {
    const long long total = rows * a.cout;
    const int yc2 = a.cout - a.ysplit;
    for (long long e = blockIdx.x * 256ll + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        // the K ranges summed in order; eight loads in flight at a time
        float s = 0.0f;
        int ks = 0;
        for (; ks + 8 <= a.ksplit; ks += 8) {
            float v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = a.part[(long long)(ks + i) * total + e];
#pragma unroll
            for (int i = 0; i < 8; ++i) s += v[i];
        }
        for (; ks < a.ksplit; ++ks) s += a.part[(long long)ks * total + e];
        if (a.relu) s = fmaxf(s, 0.0f);
        const long long vo = e / a.cout;
        const int ch = (int)(e - vo * a.cout);
        if (ch < a.ysplit) a.y1[vo * a.ysplit + ch] = (bf16)s;
        else a.y2[vo * yc2 + (ch - a.ysplit)] = (bf16)s;
    }
}

struct WgradArgs {
    const bf16 *x1, *x2, *dy, *mask;
    float *partial;  // [gridDim.x][27][cout][cin]
    int lx, c1, c2, cin, cout, batch;
};

// Weight gradient, one tap per blockIdx.y:
//   dW_tap[n][c] = sum_{samples, voxels v} dYm[v][n] X[v + off(tap)][c],
// dYm = dY masked by the ReLU derivative.  A workgroup walks chunks of 64
// voxels; per chunk X (shifted by the tap) and dYm are staged transposed in
// LDS ([channel][voxel]) so that both MFMA operands are 16-byte LDS reads.
// Each wave owns output blocks (n-block, c-block) ob = wave, wave + 4, ...;
// per-workgroup sums go to `partial` (reduced in k_wgrad_reduce).
constexpr int kWChunk = 64;
constexpr int kMaxCin = 112, kMaxCout = 64, kMaxOB = 7;
__global__ void __launch_bounds__(256) k_wgrad(WgradArgs a) {
    // rows padded by 8 elements (16 bytes) so that the 16 rows one MFMA
    // operand read touches start in different LDS banks
    __shared__ __attribute__((aligned(16))) bf16 xs[kMaxCin][kWChunk + 8];
    __shared__ __attribute__((aligned(16))) bf16 ds[kMaxCout][kWChunk + 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int tap = blockIdx.y;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int n = 1 << a.lx;
    const long long nvox = 1ll << (3 * a.lx), nvox2 = nvox >> 3;
    const int nbk = a.cout / 16, cbk = a.cin / 16, nob = nbk * cbk;
    f32x4 acc[kMaxOB];
#pragma unroll
    for (int i = 0; i < kMaxOB; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    const long long chunks = (long long)a.batch * (nvox / kWChunk);
    const int xruns = a.cin / 8, druns = a.cout / 8;
    for (long long ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
        const int b = (int)(ch / (nvox / kWChunk));
        const int v0 = (int)(ch % (nvox / kWChunk)) * kWChunk;
        // stage X shifted by the tap: item = (voxel, 8-channel run)
        for (int it = threadIdx.x; it < kWChunk * xruns; it += 256) {
            const int vi = it / xruns, run = it % xruns, c = run * 8;
            int x, y, z;
            vdec(v0 + vi, a.lx, x, y, z);
            const int xx = x + dx, yy = y + dy, zz = z + dz;
            bf16x8 v = zero8();
            if ((unsigned)xx < (unsigned)n && (unsigned)yy < (unsigned)n && (unsigned)zz < (unsigned)n) {
                if (c < a.c1)
                    v = ld8(a.x1 + (b * nvox + venc(xx, yy, zz, a.lx)) * a.c1 + c);
                else
                    v = ld8(a.x2 + (b * nvox2 + venc(xx >> 1, yy >> 1, zz >> 1, a.lx - 1)) * a.c2 +
                            (c - a.c1));
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) xs[c + j][vi] = v[j];
        }
        for (int it = threadIdx.x; it < kWChunk * druns; it += 256) {
            const int vi = it / druns, c = (it % druns) * 8;
            const long long off = (b * nvox + v0 + vi) * a.cout + c;
            bf16x8 v = ld8(a.dy + off);
            if (a.mask) v = mask8(v, ld8(a.mask + off));
#pragma unroll
            for (int j = 0; j < 8; ++j) ds[c + j][vi] = v[j];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kMaxOB; ++i) {
            const int ob = wave + 4 * i;
            if (ob >= nob) break;
            const int nb = ob / cbk, cb = ob % cbk;
#pragma unroll
            for (int kc = 0; kc < kWChunk; kc += 32) {
                const bf16x8 av = *reinterpret_cast<const bf16x8 *>(&ds[nb * 16 + m][kc + g * 8]);
                const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(&xs[cb * 16 + m][kc + g * 8]);
                acc[i] = mfma(av, bv, acc[i]);
            }
        }
        __syncthreads();
    }
    float *out = a.partial + ((long long)blockIdx.x * 27 + tap) * a.cout * a.cin;
#pragma unroll
    for (int i = 0; i < kMaxOB; ++i) {
        const int ob = wave + 4 * i;
        if (ob >= nob) break;
        const int nb = ob / cbk, cb = ob % cbk;
#pragma unroll
        for (int j = 0; j < 4; ++j) out[(nb * 16 + 4 * g + j) * a.cin + cb * 16 + m] = acc[i][j];
    }
}

// grad[n][tap * cin + c] = sum over workgroups of partial[wg][tap][n][c]
// (fixed order: deterministic).
__global__ void k_wgrad_reduce(const float *__restrict__ partial, int groups, int cin, int cout,
                               int kpad, float *__restrict__ grad) {
    const int total = 27 * cout * cin;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int tap = e / (cout * cin), rem = e % (cout * cin), nn = rem / cin, c = rem % cin;
        float s = 0.0f;
        for (int gi = 0; gi < groups; ++gi) s += partial[(long long)gi * total + e];
        grad[(long long)nn * kpad + tap * cin + c] = s;
    }
}

// 2x2x2 mean pooling: out (level l+1) from in (level l), C channels.
__global__ void k_pool(const bf16 *__restrict__ in, bf16 *__restrict__ out, int lxc, int C,
                       long long total_runs) {
    const int runs = C / 8;
    const long long nvc = 1ll << (3 * lxc), nvf = nvc << 3;
    for (long long it = blockIdx.x * (long long)blockDim.x + threadIdx.x; it < total_runs;
         it += (long long)gridDim.x * blockDim.x) {
        const long long bv = it / runs;
        const int c = (int)(it % runs) * 8;
        const int b = (int)(bv / nvc), V = (int)(bv % nvc);
        int X, Y, Z;
        vdec(V, lxc, X, Y, Z);
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 8; ++q) {
            const int v = venc(2 * X + (q & 1), 2 * Y + (q >> 1 & 1), 2 * Z + (q >> 2), lxc + 1);
            const bf16x8 t = ld8(in + (b * nvf + v) * C + c);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += (float)t[j];
        }
        bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = (bf16)(0.125f * s[j]);
        st8(out + (b * nvc + V) * C + c, o);
    }
}

// Gradient of pooling, added: gs[v] += gp[parent(v)] / 8 (fine level lxf).
__global__ void k_pool_back_add(bf16 *__restrict__ gs, const bf16 *__restrict__ gp, int lxf, int C,
                                long long total_runs) {
    const int runs = C / 8;
    const long long nvf = 1ll << (3 * lxf), nvc = nvf >> 3;
    for (long long it = blockIdx.x * (long long)blockDim.x + threadIdx.x; it < total_runs;
         it += (long long)gridDim.x * blockDim.x) {
        const long long bv = it / runs;
        const int c = (int)(it % runs) * 8;
        const int b = (int)(bv / nvf), v = (int)(bv % nvf);
        int x, y, z;
        vdec(v, lxf, x, y, z);
        const bf16x8 p = ld8(gp + (b * nvc + venc(x >> 1, y >> 1, z >> 1, lxf - 1)) * C + c);
        bf16 *dst = gs + (b * nvf + v) * C + c;
        bf16x8 s = ld8(dst);
#pragma unroll
        for (int j = 0; j < 8; ++j) s[j] = (bf16)((float)s[j] + 0.125f * (float)p[j]);
        st8(dst, s);
    }
}

// Gradient of nearest upsampling: gc[V] = sum of the 8 children of gf
// (fine level lxc + 1), C channels.
__global__ void k_upsample_back(const bf16 *__restrict__ gf, bf16 *__restrict__ gc, int lxc, int C,
                                long long total_runs) {
    const int runs = C / 8;
    const long long nvc = 1ll << (3 * lxc), nvf = nvc << 3;
    for (long long it = blockIdx.x * (long long)blockDim.x + threadIdx.x; it < total_runs;
         it += (long long)gridDim.x * blockDim.x) {
        const long long bv = it / runs;
        const int c = (int)(it % runs) * 8;
        const int b = (int)(bv / nvc), V = (int)(bv % nvc);
        int X, Y, Z;
        vdec(V, lxc, X, Y, Z);
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 8; ++q) {
            const int v = venc(2 * X + (q & 1), 2 * Y + (q >> 1 & 1), 2 * Z + (q >> 2), lxc + 1);
            const bf16x8 t = ld8(gf + (b * nvf + v) * C + c);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += (float)t[j];
        }
        bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = (bf16)s[j];
        st8(gc + (b * nvc + V) * C + c, o);
    }
}

// Per-sample statistics are two-pass: a partials kernel on grid (kSG,
// batch) writes one row per workgroup, and every block of the kernel that
// needs the statistic folds its sample's kSG rows with the same fixed tree
// (deterministic, one row per thread).  Blocks are 256 threads.
constexpr int kSG = 256;

// Sum over a 256-thread workgroup of two values; valid in thread 0.
__device__ inline void wg_sum2(float &a, float &b) {
    __shared__ float sh[2][4];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        sh[0][w] = a;
        sh[1][w] = b;
    }
    __syncthreads();
    a = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    b = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
    __syncthreads();  // sh may be reused by the caller's next wg_sum2
}
// The sample's kSG rows of two sums, folded by the whole workgroup (every
// thread must call it): thread t loads row t, then wg_sum2's fixed tree.
__device__ inline void fold2(const float *part, int b, float &a, float &c) {
    a = part[(b * kSG + threadIdx.x) * 2];
    c = part[(b * kSG + threadIdx.x) * 2 + 1];
    wg_sum2(a, c);
}

// The first layer fused with the input normalisation: x0 = bf16(in / s_b)
// with s_b the rms of the sample (1 when zero; k_pack's rule, from the
// k_sample_partials<0> rows), then the 1 -> 16 convolution of that one
// channel (27 taps, one 32-wide MFMA step) and ReLU.  A block owns a
// 16 x 4 x 4 voxel tile and stages the normalised input with a one-voxel halo
// in LDS (zero outside the domain); wave w takes plane z0 + w, one MFMA per
// 16-voxel row.  The MFMA takes the weights as A (rows: output channels; the
// layer's weight of tap t sits at k = t * cin of its row) and the taps as B
// (columns: voxels), so a lane's D holds four consecutive channels of one
// voxel and stores 8 contiguous bytes.
__global__ void __launch_bounds__(256) k_in1(ConvArgs a, const float *__restrict__ in,
                                             const float *__restrict__ part, float *__restrict__ scale) {
    constexpr int HX = 18, HY = 6, HZ = 6, NU = HX * HY * HZ;
    __shared__ bf16 tile[NU];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, g = lane >> 4;
    const int b = blockIdx.y;
    const int n = 1 << a.lx;
    const long long nvox = 1ll << (3 * a.lx);
    const int tx = n >> 4, ty = n >> 2;
    const int t = blockIdx.x;
    const int x0 = (t % tx) * 16, y0 = ((t / tx) % ty) * 4, z0 = (t / (tx * ty)) * 4;
    // the raw input values of this thread's halo slots and its weights are
    // loaded first; the statistics fold overlaps them
    const float *inb = in + b * nvox;
    constexpr int IT = (NU + 255) / 256;
    float raw[IT];
    bool inside[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int u = threadIdx.x + 256 * i;
        const int hx = u % HX, hy = (u / HX) % HY, hz = u / (HX * HY);
        const int xx = x0 - 1 + hx, yy = y0 - 1 + hy, zz = z0 - 1 + hz;
        inside[i] = u < NU && (unsigned)xx < (unsigned)n && (unsigned)yy < (unsigned)n && (unsigned)zz < (unsigned)n;
        raw[i] = inside[i] ? inb[venc(xx, yy, zz, a.lx)] : 0.0f;
    }
    bf16x8 wv = zero8();
    const bf16 *wp = a.w + (long long)m * a.kpad;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int tp = 8 * g + j;
        if (tp < 27) wv[j] = wp[tp * a.cin];
    }
    float ss, unused;
    fold2(part, b, ss, unused);
    const float s = ss > 0.0f ? sqrtf(ss / (float)nvox) : 1.0f;
    if (blockIdx.x == 0 && threadIdx.x == 0) scale[b] = s;
    const float inv = 1.0f / s;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int u = threadIdx.x + 256 * i;
        if (u < NU) tile[u] = inside[i] ? (bf16)(raw[i] * inv) : (bf16)0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        bf16x8 xv = zero8();
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int tp = 8 * g + j;
            if (tp < 27) {
                const int dz = tp / 9 - 1, dy = (tp / 3) % 3 - 1, dx = tp % 3 - 1;
                xv[j] = tile[((wave + 1 + dz) * HY + (r + 1 + dy)) * HX + (m + 1 + dx)];
            }
        }
        const f32x4 acc = mfma(wv, xv, f32x4{0.f, 0.f, 0.f, 0.f});
        typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
        bf16x4 o;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float val = acc[j];
            if (a.relu) val = fmaxf(val, 0.0f);
            o[j] = (bf16)val;
        }
        const long long vo = b * nvox + venc(x0 + m, y0 + r, z0 + wave, a.lx);
        *reinterpret_cast<bf16x4 *>(a.y1 + vo * a.cout + 4 * g) = o;
    }
}

// Partials of sum x^2 of an FP32 input (MODE 0) or of sum o and
// sum t * d / s_b (MODE 1; o = output channel 0, t the target).
template <int MODE>
__global__ void __launch_bounds__(256) k_sample_partials(const float *__restrict__ x,
                                                         const bf16 *__restrict__ out,
                                                         const float *__restrict__ scale, float d,
                                                         int nvox, float *__restrict__ part) {
    const int b = blockIdx.y;
    float a = 0.0f, c = 0.0f;
    const float ts = (MODE == 1 && x) ? d / scale[b] : 0.0f;
#pragma unroll 4
    for (int v = blockIdx.x * 256 + threadIdx.x; v < nvox; v += kSG * 256) {
        const long long i = (long long)b * nvox + v;
        if (MODE == 0) {
            a += x[i] * x[i];
        } else {
            a += (float)out[i * 16];
            if (x) c += x[i] * ts;  // no target when unpacking
        }
    }
    wg_sum2(a, c);
    if (threadIdx.x == 0) {
        part[(b * kSG + blockIdx.x) * 2] = a;
        part[(b * kSG + blockIdx.x) * 2 + 1] = c;
    }
}

// Input packing for training: s_b = rms of the sample (1 when zero), x0 =
// in / s_b in channel 0 and channels 1..15 zero -- the 16-channel input the
// first layer's weight gradient reads (the forward pass normalises the input
// itself, k_in1).
__global__ void __launch_bounds__(256) k_pack(const float *__restrict__ in, int nvox,
                                              const float *__restrict__ part, bf16 *__restrict__ x0) {
    const int b = blockIdx.y;
    float ss, unused;
    fold2(part, b, ss, unused);
    const float s = ss > 0.0f ? sqrtf(ss / (float)nvox) : 1.0f;
    const float inv = 1.0f / s;
    for (int v = blockIdx.x * 256 + threadIdx.x; v < nvox; v += gridDim.x * 256) {
        const bf16 val = (bf16)(in[(long long)b * nvox + v] * inv);
        bf16x8 lo = zero8();
        lo[0] = val;
        bf16 *dst = x0 + ((long long)b * nvox + v) * 16;
        st8(dst, lo);
        st8(dst + 8, zero8());
    }
}

// Loss: with o = output channel 0 and t = target * d / s_b (d the target
// normalisation, tscale), e = (o - mean o) - (t - mean t); gradient
// 2 e / (nvox * batch) in channel 0, and partials of sum e^2 into part2.
__global__ void __launch_bounds__(256) k_loss(const bf16 *__restrict__ out, const float *__restrict__ tgt,
                                              const float *__restrict__ scale, float d, int nvox,
                                              int batch, const float *__restrict__ part,
                                              bf16 *__restrict__ gout, float *__restrict__ part2) {
    const int b = blockIdx.y;
    const float ts = d / scale[b];
    float so, st;
    fold2(part, b, so, st);
    const float mo = so / (float)nvox, mt = st / (float)nvox;
    const float gscale = 2.0f / ((float)nvox * (float)batch);
    float se = 0.0f, unused = 0.0f;
    for (int v = blockIdx.x * 256 + threadIdx.x; v < nvox; v += kSG * 256) {
        const long long i = (long long)b * nvox + v;
        const float e = ((float)out[i * 16] - mo) - (tgt[i] * ts - mt);
        se += e * e;
        bf16x8 lo = zero8();
        lo[0] = (bf16)(gscale * e);
        st8(gout + i * 16, lo);
        st8(gout + i * 16 + 8, zero8());
    }
    wg_sum2(se, unused);
    if (threadIdx.x == 0) {
        part2[(b * kSG + blockIdx.x) * 2] = se;
        part2[(b * kSG + blockIdx.x) * 2 + 1] = 0.0f;
    }
}

// loss_b = mean(e^2), one workgroup per sample.
__global__ void __launch_bounds__(256) k_loss_fold(const float *__restrict__ part2, int nvox,
                                                   float *__restrict__ loss) {
    const int b = blockIdx.x;
    float se, unused;
    fold2(part2, b, se, unused);
    if (threadIdx.x == 0) loss[b] = se / (float)nvox;
}

// Output unpacking from the inference head's compact channel 0: mean o from
// the head's `rows` workgroup sums of the sample (every block folds them in
// the same fixed order), then p = (o - mean o) s_b / d.
__global__ void __launch_bounds__(256) k_unpack_c(const float *__restrict__ o, const float *__restrict__ hsum,
                                                  int rows, const float *__restrict__ scale, float d, int nvox,
                                                  float *__restrict__ p) {
    __shared__ float sh[4];
    const int b = blockIdx.y;
    float a = 0.0f;
    for (int i = threadIdx.x; i < rows; i += 256) a += hsum[(long long)b * rows + i];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    const float mo = ((sh[0] + sh[1]) + (sh[2] + sh[3])) / (float)nvox, f = scale[b] / d;
    // four voxels per thread (nvox is a multiple of 1024)
    for (int v = 4 * (blockIdx.x * 256 + threadIdx.x); v < nvox; v += 4 * gridDim.x * 256) {
        const long long i = (long long)b * nvox + v;
        const float4 ov = *reinterpret_cast<const float4 *>(o + i);
        *reinterpret_cast<float4 *>(p + i) = make_float4((ov.x - mo) * f, (ov.y - mo) * f, (ov.z - mo) * f,
                                                         (ov.w - mo) * f);
    }
}

// Adam on the FP32 master weights; refreshes the BF16 copy.
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m1, float *__restrict__ m2,
                       const float *__restrict__ gr, bf16 *__restrict__ wb, long long count,
                       float lr_t, float b1, float b2, float eps) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < count;
         i += (long long)gridDim.x * blockDim.x) {
        const float gi = gr[i];
        const float a = b1 * m1[i] + (1.0f - b1) * gi;
        const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = v;
        const float wi = w[i] - lr_t * a / (sqrtf(v) + eps);
        w[i] = wi;
        wb[i] = (bf16)wi;
    }
}

// Flipped transposed weights for the data gradient:
// wt[c][tap' * cout + n] = wb[n][(26 - tap') * cin + c], zero in the padding.
__global__ void k_transpose_w(const bf16 *__restrict__ wb, int cin, int cout, int kpad,
                              bf16 *__restrict__ wt, int kpad_t) {
    const int total = cin * kpad_t;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int c = e / kpad_t, k = e % kpad_t;
        bf16 v = (bf16)0.0f;
        if (k < 27 * cout) {
            const int tp = k / cout, nn = k % cout;
            v = wb[(long long)nn * kpad + (26 - tp) * cin + c];
        }
        wt[e] = v;
    }
}

// Layout check of the MFMA operands: d = a (16x32) b (32x16), row-major FP32
// in and out, through the lane mapping the convolution kernels use.
__global__ void k_probe_mfma(const float *a, const float *b, float *d) {
    const int lane = threadIdx.x, m = lane & 15, g = lane >> 4;
    bf16x8 av, bv;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        av[j] = (bf16)a[m * 32 + g * 8 + j];
        bv[j] = (bf16)b[(g * 8 + j) * 16 + m];
    }
    const f32x4 acc = mfma(av, bv, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
    for (int j = 0; j < 4; ++j) d[(4 * g + j) * 16 + m] = acc[j];
}

// ---------------------------------------------------------------------------
// Host side: the layer table and the forward / backward schedules.
// ---------------------------------------------------------------------------
constexpr int kActs = 19, kLayers = 15;
enum Act { X0, A0, S0, P1, A1, S1, P2, A2, S2, P3, A3, B3, U2, D2, U1, D1, U0, D0, OUT };
struct ActSpec {
    int level, ch;
};
constexpr ActSpec kActSpec[kActs] = {
    {0, 16}, {0, 16}, {0, 16}, {1, 16}, {1, 32}, {1, 32}, {2, 32}, {2, 48}, {2, 48}, {3, 48},
    {3, 64}, {3, 64}, {2, 48}, {2, 48}, {1, 32}, {1, 32}, {0, 16}, {0, 16}, {0, 16}};
struct LayerSpec {
    int in1, in2, out, relu;
};
constexpr LayerSpec kLayerSpec[kLayers] = {
    {X0, -1, A0, 1}, {A0, -1, S0, 1}, {P1, -1, A1, 1}, {A1, -1, S1, 1}, {P2, -1, A2, 1},
    {A2, -1, S2, 1}, {P3, -1, A3, 1}, {A3, -1, B3, 1}, {S2, B3, U2, 1}, {U2, -1, D2, 1},
    {S1, D2, U1, 1}, {U1, -1, D1, 1}, {S0, D1, U0, 1}, {U0, -1, D0, 1}, {D0, -1, OUT, 0}};
// forward schedule: layer index >= 0, or -1 - pool index (pool s -> p)
struct Pool {
    int src, dst;
};
constexpr Pool kPools[3] = {{S0, P1}, {S1, P2}, {S2, P3}};
constexpr int kSchedule[18] = {0, 1, -1, 2, 3, -2, 4, 5, -3, 6, 7, 8, 9, 10, 11, 12, 13, 14};

}  // namespace un
}  // namespace fl

using namespace fl;
using namespace fl::un;

struct fl_unet {
    fl_ctx *ctx = nullptr;
    int n = 0, lx = 0, maxb = 0;
    struct L {
        int c1, c2, cin, cout, kpad, kpad_t, lx;
        size_t off, offt;  // offsets into the weight buffers
    } layer[kLayers]{};
    size_t P = 0, Pt = 0;  // parameter counts (padded), transposed copies
    bf16 *act[kActs] = {}, *grad[kActs] = {};
    bf16 *tmp = nullptr;           // upsample part of a concat layer's data gradient
    float *hout = nullptr;         // inference head: output channel 0 [batch][voxel]
    float *hsum = nullptr;         // and its workgroup sums [batch][head_rows]
    int head_rows = 0;
    float *w = nullptr, *m1 = nullptr, *m2 = nullptr, *gr = nullptr;
    bf16 *wb = nullptr, *wt = nullptr;
    float *part = nullptr, *part2 = nullptr;  // per-sample statistics rows
    float *partial = nullptr, *in = nullptr, *tgt = nullptr, *out = nullptr, *scale = nullptr,
          *loss = nullptr;
    float *kpart = nullptr;  // K-split partial sums of the coarse levels' convolutions
    size_t kpart_len = 0;
    float *host_loss = nullptr;  // pinned
    int wgroups = 64;
    long long step = 0;
    float gain = 1.0f;  // output gain G (see tscale)
};

namespace {

fl_status ualloc(fl_unet *u, void **p, size_t bytes) {
    FL_HIP_TRY(u->ctx, hipMalloc(p, bytes));
    FL_HIP_TRY(u->ctx, hipMemsetAsync(*p, 0, bytes, u->ctx->stream));
    return FL_OK;
}

void ufree(fl_unet *u) {
    for (int i = 0; i < kActs; ++i) {
        if (u->act[i]) (void)hipFree(u->act[i]);
        if (u->grad[i]) (void)hipFree(u->grad[i]);
    }
    void *bufs[] = {u->tmp, u->hout, u->hsum, u->w, u->m1, u->m2, u->gr, u->wb, u->wt, u->partial,
                    u->in,  u->tgt, u->out, u->scale, u->loss, u->part, u->part2, u->kpart};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (u->host_loss) (void)hipHostFree(u->host_loss);
}

long long nvox_at(const fl_unet *u, int level) { return 1ll << (3 * (u->lx - level)); }

// Target normalisation diag / (rms * G) with the network's output gain G
// (the last parameter).  The default G = n^2 / 2 makes the smoothest modes
// of the unit cube O(1) (p ~ b / pi^2, diag = 6 n^2); training sets G to the
// data's typical rms(p) diag / rms(b) so that targets are O(1).
float tscale(const fl_unet *u, double diag) { return (float)(diag / (double)u->gain); }

// The input's per-sample statistics (the first layer normalises with them);
// for training (full) also the 16-channel normalised input X0.
void launch_pack(fl_unet *u, hipStream_t stm, const float *in, int B, bool full = true) {
    const int nv = (int)nvox_at(u, 0);
    const dim3 grid(kSG, (unsigned)B);
    k_sample_partials<0><<<grid, 256, 0, stm>>>(in, nullptr, nullptr, 0.0f, nv, u->part);
    if (full) k_pack<<<grid, 256, 0, stm>>>(in, nv, u->part, u->act[X0]);
}
// out <- denormalised zero-mean prediction, from the inference head's
// compact output and workgroup sums
void launch_unpack(fl_unet *u, hipStream_t stm, double diag, float *out, int B) {
    const int nv = (int)nvox_at(u, 0);
    const dim3 grid((unsigned)(nv / 1024), (unsigned)B);
    k_unpack_c<<<grid, 256, 0, stm>>>(u->hout, u->hsum, u->head_rows, u->scale, tscale(u, diag), nv, out);
}
// grad[OUT] <- loss gradient, loss[b] <- per-sample loss (targets in u->tgt)
void launch_loss(fl_unet *u, hipStream_t stm, double diag, int B) {
    const int nv = (int)nvox_at(u, 0);
    const dim3 grid(kSG, (unsigned)B);
    k_sample_partials<1><<<grid, 256, 0, stm>>>(u->tgt, u->act[OUT], u->scale, tscale(u, diag), nv, u->part);
    k_loss<<<grid, 256, 0, stm>>>(u->act[OUT], u->tgt, u->scale, tscale(u, diag), nv, B, u->part,
                                  u->grad[OUT], u->part2);
    k_loss_fold<<<(unsigned)B, 256, 0, stm>>>(u->part2, nv, u->loss);
}

// Work split of a convolution with `cout` outputs over B * nvox voxels and
// kpad K: whole output-channel groups per wave while that fills the chip
// (8192 waves), else one 16-channel block per wave and, below 4096 waves,
// K ranges of at least two MFMA steps.
struct ConvSplit {
    int nb, nsplit, ksplit;
};
ConvSplit conv_split(long long nvox, int B, int cout, int kpad) {
    const long long base = B * nvox / 16;
    const int blocks = cout / 16, steps = kpad / 32;
    ConvSplit sp{blocks, 1, 1};
    if (base * 1 >= 8192 || blocks == 1) {
        sp = ConvSplit{blocks, 1, 1};
    } else {
        sp = ConvSplit{1, blocks, 1};
    }
    const long long waves = base * sp.nsplit;
    if (waves < 4096) {
        const long long want = (8192 + waves - 1) / waves;
        sp.ksplit = (int)std::max<long long>(1, std::min<long long>(want, steps / 2));
    }
    return sp;
}

fl_status launch_conv(fl_unet *u, hipStream_t stm, int B, ConvArgs a, bool head = false) {
    const long long nvox = 1ll << (3 * a.lx);
    if (head) {  // the inference head: compact channel 0 (cout 16, 32^3 and finer)
        const int n = 1 << a.lx, R = a.lx >= 6 ? 4 : 1;
        const unsigned tiles = (unsigned)((n / 16) * (n / R) * (n / 4));
        const dim3 grid(tiles, (unsigned)B);
        if (R == 4) k_conv_lds<4, 1, true><<<grid, 256, 0, stm>>>(a);
        else k_conv_lds<1, 1, true><<<grid, 256, 0, stm>>>(a);
        FL_HIP_TRY(u->ctx, hipGetLastError());
        return FL_OK;
    }
    // forward convolutions of the 64^3 to 16^3 levels: the LDS-staged kernel
    if (!a.mask && a.lx >= 4 && a.cin % 16 == 0 && a.c1 % 16 == 0 && a.cout % 16 == 0 &&
        a.ysplit == a.cout) {
        // 64^3 and finer: 16 x 4 x 4 voxels and one output block per
        // workgroup; 32^3 and 16^3: 16 x 1 x 4 voxels and two output blocks
        // where the channels pair up (512 workgroups at 64^3 and 32^3, 64 per
        // output block at 16^3, where the old kernel needed K splits)
        const int n = 1 << a.lx, R = a.lx >= 6 ? 4 : 1;
        const int NB = (a.lx >= 6 || a.cout % 32) ? 1 : 2;
        const unsigned tiles = (unsigned)((n / 16) * (n / R) * (n / 4));
        const dim3 grid(tiles * (unsigned)(a.cout / (16 * NB)), (unsigned)B);
        if (R == 4) k_conv_lds<4, 1><<<grid, 256, 0, stm>>>(a);
        else if (NB == 2) k_conv_lds<1, 2><<<grid, 256, 0, stm>>>(a);
        else k_conv_lds<1, 1><<<grid, 256, 0, stm>>>(a);
        FL_HIP_TRY(u->ctx, hipGetLastError());
        return FL_OK;
    }
    const ConvSplit sp = conv_split(nvox, B, a.cout, a.kpad);
    a.nsplit = sp.nsplit;
    a.ksplit = sp.ksplit;
    a.batch = B;
    a.part = u->kpart;
    const long long rows = (long long)B * nvox;
    if (sp.ksplit > 1 && (size_t)(sp.ksplit * rows * a.cout) > u->kpart_len)
        return set_error(u->ctx, FL_ERR_INVALID_ARGUMENT, "K-split buffer too small");
    const dim3 grid((unsigned)(nvox / 64 * sp.nsplit * sp.ksplit), (unsigned)B);
    switch (sp.nb) {
        case 1: k_conv<1><<<grid, 256, 0, stm>>>(a); break;
        case 2: k_conv<2><<<grid, 256, 0, stm>>>(a); break;
        case 3: k_conv<3><<<grid, 256, 0, stm>>>(a); break;
        case 4: k_conv<4><<<grid, 256, 0, stm>>>(a); break;
        case 5: k_conv<5><<<grid, 256, 0, stm>>>(a); break;
        case 6: k_conv<6><<<grid, 256, 0, stm>>>(a); break;
        case 7: k_conv<7><<<grid, 256, 0, stm>>>(a); break;
        default: return set_error(u->ctx, FL_ERR_INVALID_ARGUMENT, "unsupported channel count %d", a.cout);
    }
    if (sp.ksplit > 1) {
        const long long total = rows * a.cout;
        k_conv_reduce<<<(unsigned)std::min<long long>((total + 255) / 256, 8192), 256, 0, stm>>>(a, rows);
    }
    FL_HIP_TRY(u->ctx, hipGetLastError());
    return FL_OK;
}

unsigned ew_grid(long long items) {
    return (unsigned)std::max<long long>(1, std::min<long long>((items + 255) / 256, 8192));
}

// The forward pass of B samples of the FP32 input `in` (their statistics
// already in u->part).  infer: the head writes only channel 0, compactly
// (u->hout, u->hsum); otherwise the full 16-channel head output act[OUT]
// that training and fl_unet_forward_raw read.
fl_status forward(fl_unet *u, hipStream_t stm, int B, const float *in, bool infer) {
    for (int op : kSchedule) {
        if (op < 0) {
            const Pool &p = kPools[-1 - op];
            const int lc = kActSpec[p.dst].level, C = kActSpec[p.dst].ch;
            const long long runs = B * nvox_at(u, lc) * (C / 8);
            k_pool<<<ew_grid(runs), 256, 0, stm>>>(u->act[p.src], u->act[p.dst], u->lx - lc, C, runs);
            FL_HIP_TRY(u->ctx, hipGetLastError());
            continue;
        }
        const LayerSpec &ls = kLayerSpec[op];
        const fl_unet::L &L = u->layer[op];
        ConvArgs a{};
        a.x1 = u->act[ls.in1];
        a.x2 = ls.in2 >= 0 ? u->act[ls.in2] : nullptr;
        a.mask = nullptr;
        a.w = u->wb + L.off;
        a.y1 = u->act[ls.out];
        a.y2 = nullptr;
        a.lx = L.lx;
        a.c1 = L.c1;
        a.c2 = L.c2;
        a.cin = L.cin;
        a.kpad = L.kpad;
        a.cout = L.cout;
        a.ysplit = L.cout;
        a.relu = ls.relu;
        if (op == 0) {  // the one-channel input: 27 taps, not 27 x 16; normalised on load
            a.batch = B;
            k_in1<<<dim3((unsigned)(nvox_at(u, 0) / 256), (unsigned)B), 256, 0, stm>>>(a, in, u->part, u->scale);
            FL_HIP_TRY(u->ctx, hipGetLastError());
            continue;
        }
        const bool head = infer && op == kLayers - 1;
        if (head) {
            a.hout = u->hout;
            a.hsum = u->hsum;
        }
        fl_status st = launch_conv(u, stm, B, a, head);
        if (st != FL_OK) return st;
    }
    return FL_OK;
}

// Backward through the schedule in reverse: weight gradients into u->gr,
// activation gradients into u->grad[] (grad[OUT] holds the loss gradient).
fl_status backward(fl_unet *u, hipStream_t stm, int B) {
    for (int si = 17; si >= 0; --si) {
        const int op = kSchedule[si];
        if (op < 0) {
            const Pool &p = kPools[-1 - op];
            const int lf = kActSpec[p.src].level, C = kActSpec[p.src].ch;
            const long long runs = B * nvox_at(u, lf) * (C / 8);
            k_pool_back_add<<<ew_grid(runs), 256, 0, stm>>>(u->grad[p.src], u->grad[p.dst], u->lx - lf,
                                                            C, runs);
            FL_HIP_TRY(u->ctx, hipGetLastError());
            continue;
        }
        const LayerSpec &ls = kLayerSpec[op];
        const fl_unet::L &L = u->layer[op];
        const bf16 *mask = ls.relu ? u->act[ls.out] : nullptr;
        // weight gradient
        WgradArgs w{};
        w.x1 = u->act[ls.in1];
        w.x2 = ls.in2 >= 0 ? u->act[ls.in2] : nullptr;
        w.dy = u->grad[ls.out];
        w.mask = mask;
        w.partial = u->partial;
        w.lx = L.lx;
        w.c1 = L.c1;
        w.c2 = L.c2;
        w.cin = L.cin;
        w.cout = L.cout;
        w.batch = B;
        const long long chunks = B * (1ll << (3 * L.lx)) / kWChunk;
        const int groups = (int)std::min<long long>(u->wgroups, chunks);
        k_wgrad<<<dim3((unsigned)groups, 27), 256, 0, stm>>>(w);
        const int tot = 27 * L.cout * L.cin;
        k_wgrad_reduce<<<(tot + 255) / 256, 256, 0, stm>>>(u->partial, groups, L.cin, L.cout, L.kpad,
                                                           u->gr + L.off);
        FL_HIP_TRY(u->ctx, hipGetLastError());
        if (op == 0) continue;  // no gradient for the input
        // data gradient: conv of the masked output gradient with the flipped
        // transposed weights; a concat layer splits it into the skip part and
        // the (fine) gradient of the upsampled part
        ConvArgs a{};
        a.x1 = u->grad[ls.out];
        a.x2 = nullptr;
        a.mask = mask;
        a.w = u->wt + L.offt;
        a.lx = L.lx;
        a.c1 = L.cout;
        a.c2 = 0;
        a.cin = L.cout;
        a.kpad = L.kpad_t;
        a.cout = L.cin;
        a.relu = 0;
        a.y1 = u->grad[ls.in1];
        a.ysplit = L.c1;
        a.y2 = ls.in2 >= 0 ? u->tmp : nullptr;
        fl_status st = launch_conv(u, stm, B, a);
        if (st != FL_OK) return st;
        if (ls.in2 >= 0) {
            const int lc = kActSpec[ls.in2].level;
            const long long runs = B * nvox_at(u, lc) * (L.c2 / 8);
            k_upsample_back<<<ew_grid(runs), 256, 0, stm>>>(u->tmp, u->grad[ls.in2], u->lx - lc, L.c2,
                                                            runs);
            FL_HIP_TRY(u->ctx, hipGetLastError());
        }
    }
    return FL_OK;
}

fl_status refresh_transposed(fl_unet *u, hipStream_t stm) {
    for (int l = 1; l < kLayers; ++l) {
        const fl_unet::L &L = u->layer[l];
        const int tot = L.cin * L.kpad_t;
        k_transpose_w<<<(tot + 255) / 256, 256, 0, stm>>>(u->wb + L.off, L.cin, L.cout, L.kpad,
                                                          u->wt + L.offt, L.kpad_t);
    }
    FL_HIP_TRY(u->ctx, hipGetLastError());
    return FL_OK;
}

// splitmix64 + Box-Muller, the weight initialisation stream
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double a = uniform(), b = uniform();
        return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586 * b);
    }
};

fl_status upload_weights(fl_unet *u, const std::vector<float> &host) {
    hipStream_t stm = u->ctx->stream;
    FL_HIP_TRY(u->ctx, hipMemcpyAsync(u->w, host.data(), u->P * sizeof(float), hipMemcpyHostToDevice, stm));
    FL_HIP_TRY(u->ctx, hipMemsetAsync(u->m1, 0, u->P * sizeof(float), stm));
    FL_HIP_TRY(u->ctx, hipMemsetAsync(u->m2, 0, u->P * sizeof(float), stm));
    // the BF16 copies: one Adam step with zero learning rate
    FL_HIP_TRY(u->ctx, hipMemsetAsync(u->gr, 0, u->P * sizeof(float), stm));
    k_adam<<<ew_grid((long long)u->P), 256, 0, stm>>>(u->w, u->m1, u->m2, u->gr, u->wb, (long long)u->P,
                                                      0.0f, 0.9f, 0.999f, 1e-8f);
    FL_HIP_TRY(u->ctx, hipGetLastError());
    fl_status st = refresh_transposed(u, stm);
    if (st != FL_OK) return st;
    FL_HIP_TRY(u->ctx, hipStreamSynchronize(stm));
    u->step = 0;
    return FL_OK;
}

// Rows and columns that must stay zero: the input's padding channels (layer
// 0, c >= 1), the head's unused outputs (layer 14, n >= 1), K padding.
bool live_weight(const fl_unet *u, int l, int nn, int k) {
    const fl_unet::L &L = u->layer[l];
    if (k >= 27 * L.cin) return false;
    if (l == 0 && k % L.cin != 0) return false;
    if (l == kLayers - 1 && nn != 0) return false;
    return true;
}

}  // namespace

namespace fl {

int unet_resolution(const fl_unet *net) { return net ? net->n : 0; }
fl_ctx *unet_context(const fl_unet *net) { return net ? net->ctx : nullptr; }

fl_status unet_infer_device(fl_unet *u, hipStream_t stm, const float *in, double diag, float *out) {
    launch_pack(u, stm, in, 1, false);
    FL_HIP_TRY(u->ctx, hipGetLastError());
    fl_status st = forward(u, stm, 1, in, true);
    if (st != FL_OK) return st;
    launch_unpack(u, stm, diag, out, 1);
    FL_HIP_TRY(u->ctx, hipGetLastError());
    return FL_OK;
}

}  // namespace fl

extern "C" {

fl_status fl_unet_create(fl_ctx *ctx, const fl_unet_config *cfg, fl_unet **out) {
    if (!ctx || !cfg || !out) return set_error(ctx, FL_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    int lx = 0;
    while ((1 << lx) < cfg->n) ++lx;
    if (cfg->n < 32 || (1 << lx) != cfg->n || cfg->n > 256)
        return set_error(ctx, FL_ERR_CONFIG, "U-Net resolution must be a power of two in [32, 256], got %d",
                         cfg->n);
    if (cfg->max_batch < 1 || cfg->max_batch > 64)
        return set_error(ctx, FL_ERR_CONFIG, "U-Net max_batch must be in [1, 64], got %d", cfg->max_batch);
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    fl_unet *u = new (std::nothrow) fl_unet();
    if (!u) return set_error(ctx, FL_ERR_OUT_OF_MEMORY, "host allocation failed");
    u->ctx = ctx;
    u->n = cfg->n;
    u->lx = lx;
    u->maxb = cfg->max_batch;
    u->gain = 0.5f * (float)cfg->n * (float)cfg->n;
    fl_status st = FL_OK;
    size_t P = 0, Pt = 0, max_partial = 0, max_tmp = 0;
    for (int l = 0; l < kLayers; ++l) {
        const LayerSpec &ls = kLayerSpec[l];
        fl_unet::L &L = u->layer[l];
        L.c1 = kActSpec[ls.in1].ch;
        L.c2 = ls.in2 >= 0 ? kActSpec[ls.in2].ch : 0;
        L.cin = L.c1 + L.c2;
        L.cout = kActSpec[ls.out].ch;
        L.kpad = (27 * L.cin + 31) / 32 * 32;
        L.kpad_t = (27 * L.cout + 31) / 32 * 32;
        L.lx = lx - kActSpec[ls.out].level;
        L.off = P;
        L.offt = Pt;
        P += (size_t)L.cout * L.kpad;
        Pt += (size_t)L.cin * L.kpad_t;
        max_partial = std::max(max_partial, (size_t)27 * L.cout * L.cin);
        if (L.c2) max_tmp = std::max(max_tmp, (size_t)(1ll << (3 * L.lx)) * L.c2);
    }
    u->P = P;
    u->Pt = Pt;
    const size_t B = (size_t)u->maxb;
    // K-split buffer: the largest split forward or data-gradient convolution
    // over every batch size up to max_batch
    for (int bb = 1; bb <= u->maxb; ++bb)
        for (int l = 0; l < kLayers; ++l) {
            const fl_unet::L &L = u->layer[l];
            const long long nv = 1ll << (3 * L.lx);
            const ConvSplit f = conv_split(nv, bb, L.cout, L.kpad), t = conv_split(nv, bb, L.cin, L.kpad_t);
            if (f.ksplit > 1) u->kpart_len = std::max(u->kpart_len, (size_t)(f.ksplit * bb * nv * L.cout));
            if (t.ksplit > 1) u->kpart_len = std::max(u->kpart_len, (size_t)(t.ksplit * bb * nv * L.cin));
        }
    for (int i = 0; i < kActs && st == FL_OK; ++i) {
        const size_t bytes = B * (size_t)nvox_at(u, kActSpec[i].level) * kActSpec[i].ch * sizeof(bf16);
        if ((st = ualloc(u, (void **)&u->act[i], bytes)) == FL_OK)
            st = ualloc(u, (void **)&u->grad[i], bytes);
    }
    const size_t nv0 = (size_t)nvox_at(u, 0);
    {  // workgroups of the inference head (k_conv_lds tiles: 16 x R x 4 voxels)
        const int R = lx >= 6 ? 4 : 1;
        u->head_rows = (cfg->n / 16) * (cfg->n / R) * (cfg->n / 4);
    }
    if (st == FL_OK &&
        ((st = ualloc(u, (void **)&u->tmp, B * max_tmp * sizeof(bf16))) ||
         (st = ualloc(u, (void **)&u->w, P * 4)) || (st = ualloc(u, (void **)&u->m1, P * 4)) ||
         (st = ualloc(u, (void **)&u->m2, P * 4)) || (st = ualloc(u, (void **)&u->gr, P * 4)) ||
         (st = ualloc(u, (void **)&u->wb, P * 2)) || (st = ualloc(u, (void **)&u->wt, Pt * 2)) ||
         (st = ualloc(u, (void **)&u->partial, (size_t)u->wgroups * max_partial * 4)) ||
         (st = ualloc(u, (void **)&u->in, B * nv0 * 4)) ||
         (st = ualloc(u, (void **)&u->hout, B * nv0 * 4)) ||
         (st = ualloc(u, (void **)&u->hsum, B * (size_t)u->head_rows * 4)) ||
         (st = ualloc(u, (void **)&u->tgt, B * nv0 * 4)) ||
         (st = ualloc(u, (void **)&u->out, B * nv0 * 4)) ||
         (st = ualloc(u, (void **)&u->scale, B * 4)) || (st = ualloc(u, (void **)&u->loss, B * 4)) ||
         (st = ualloc(u, (void **)&u->part, B * kSG * 2 * 4)) ||
         (st = ualloc(u, (void **)&u->part2, B * kSG * 2 * 4)) ||
         (st = ualloc(u, (void **)&u->kpart, std::max<size_t>(u->kpart_len, 1) * 4))))
        ;
    if (st == FL_OK && hipHostMalloc(&u->host_loss, B * sizeof(float), hipHostMallocDefault) != hipSuccess)
        st = set_error(ctx, FL_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
    if (st == FL_OK) {
        // He-normal initialisation over the live fan-in; the head starts small
        std::vector<float> host(P, 0.0f);
        Rng rng{cfg->seed};
        for (int l = 0; l < kLayers; ++l) {
            const fl_unet::L &L = u->layer[l];
            const double fan = 27.0 * (l == 0 ? 1 : L.cin);
            const double sd = (l == kLayers - 1 ? 0.1 : 1.0) * std::sqrt(2.0 / fan);
            for (int nn = 0; nn < L.cout; ++nn)
                for (int k = 0; k < L.kpad; ++k)
                    if (live_weight(u, l, nn, k)) host[L.off + (size_t)nn * L.kpad + k] = (float)(sd * rng.normal());
        }
        st = upload_weights(u, host);
    }
    if (st != FL_OK) {
        ufree(u);
        delete u;
        return st;
    }
    ctx_retain(ctx);
    *out = u;
    return FL_OK;
}

void fl_unet_destroy(fl_unet *u) {
    if (!u) return;
    fl_ctx *ctx = u->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ufree(u);
    delete u;
    ctx_release(ctx);
}

// The parameter vector is the weights followed by the output gain G.
int64_t fl_unet_num_params(const fl_unet *u) { return u ? (int64_t)u->P + 1 : 0; }

fl_status fl_unet_get_params(fl_unet *u, float *out) {
    if (!u || !out) return FL_ERR_INVALID_ARGUMENT;
    FL_HIP_TRY(u->ctx, hipSetDevice(u->ctx->device));
    FL_HIP_TRY(u->ctx, hipMemcpyAsync(out, u->w, u->P * sizeof(float), hipMemcpyDeviceToHost, u->ctx->stream));
    FL_HIP_TRY(u->ctx, hipStreamSynchronize(u->ctx->stream));
    out[u->P] = u->gain;
    return FL_OK;
}

fl_status fl_unet_set_params(fl_unet *u, const float *in) {
    if (!u || !in) return FL_ERR_INVALID_ARGUMENT;
    FL_HIP_TRY(u->ctx, hipSetDevice(u->ctx->device));
    if (!(std::isfinite(in[u->P]) && in[u->P] > 0.0f))
        return set_error(u->ctx, FL_ERR_INVALID_ARGUMENT, "the output gain must be positive and finite");
    u->gain = in[u->P];
    std::vector<float> host(in, in + u->P);
    for (int l = 0; l < kLayers; ++l) {
        const fl_unet::L &L = u->layer[l];
        for (int nn = 0; nn < L.cout; ++nn)
            for (int k = 0; k < L.kpad; ++k) {
                float &v = host[L.off + (size_t)nn * L.kpad + k];
                if (!std::isfinite(v))
                    return set_error(u->ctx, FL_ERR_NON_FINITE, "non-finite parameter in layer %d", l);
                if (!live_weight(u, l, nn, k)) v = 0.0f;
            }
    }
    return upload_weights(u, host);
}

fl_status fl_unet_train_step(fl_unet *u, int32_t batch, const float *rhs, const float *p, double diag,
                             float lr, float *loss) {
    if (!u || !rhs || !p) return set_error(u ? u->ctx : nullptr, FL_ERR_INVALID_ARGUMENT, "null argument");
    if (batch < 1 || batch > u->maxb)
        return set_error(u->ctx, FL_ERR_INVALID_ARGUMENT, "batch %d outside [1, %d]", batch, u->maxb);
    if (!(diag > 0.0) || !(lr >= 0.0f))
        return set_error(u->ctx, FL_ERR_INVALID_ARGUMENT, "diag must be positive and lr non-negative");
    fl_ctx *ctx = u->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t stm = ctx->stream;
    const long long nv = nvox_at(u, 0);
    FL_HIP_TRY(ctx, hipMemcpyAsync(u->in, rhs, batch * nv * sizeof(float), hipMemcpyHostToDevice, stm));
    FL_HIP_TRY(ctx, hipMemcpyAsync(u->tgt, p, batch * nv * sizeof(float), hipMemcpyHostToDevice, stm));
    launch_pack(u, stm, u->in, batch);
    FL_HIP_TRY(ctx, hipGetLastError());
    fl_status st = forward(u, stm, batch, u->in, false);
    if (st != FL_OK) return st;
    launch_loss(u, stm, diag, batch);
    FL_HIP_TRY(ctx, hipGetLastError());
    if ((st = backward(u, stm, batch)) != FL_OK) return st;
    u->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const double t = (double)u->step;
    const float lr_t = (float)(lr * std::sqrt(1.0 - std::pow(b2, t)) / (1.0 - std::pow(b1, t)));
    k_adam<<<ew_grid((long long)u->P), 256, 0, stm>>>(u->w, u->m1, u->m2, u->gr, u->wb, (long long)u->P,
                                                      lr_t, b1, b2, 1e-8f);
    FL_HIP_TRY(ctx, hipGetLastError());
    if ((st = refresh_transposed(u, stm)) != FL_OK) return st;
    FL_HIP_TRY(ctx, hipMemcpyAsync(u->host_loss, u->loss, batch * sizeof(float), hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
    double s = 0.0;
    for (int b = 0; b < batch; ++b) s += u->host_loss[b];
    if (loss) *loss = (float)(s / batch);
    if (!std::isfinite(s)) return set_error(ctx, FL_ERR_NON_FINITE, "U-Net training loss is not finite");
    return FL_OK;
}

fl_status fl_unet_infer(fl_unet *u, int32_t batch, const float *rhs, double diag, float *p_out) {
    if (!u || !rhs || !p_out) return set_error(u ? u->ctx : nullptr, FL_ERR_INVALID_ARGUMENT, "null argument");
    if (batch < 1 || batch > u->maxb)
        return set_error(u->ctx, FL_ERR_INVALID_ARGUMENT, "batch %d outside [1, %d]", batch, u->maxb);
    if (!(diag > 0.0)) return set_error(u->ctx, FL_ERR_INVALID_ARGUMENT, "diag must be positive");
    fl_ctx *ctx = u->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t stm = ctx->stream;
    const long long nv = nvox_at(u, 0);
    FL_HIP_TRY(ctx, hipMemcpyAsync(u->in, rhs, batch * nv * sizeof(float), hipMemcpyHostToDevice, stm));
    launch_pack(u, stm, u->in, batch, false);
    FL_HIP_TRY(ctx, hipGetLastError());
    fl_status st = forward(u, stm, batch, u->in, true);
    if (st != FL_OK) return st;
    launch_unpack(u, stm, diag, u->out, batch);
    FL_HIP_TRY(ctx, hipGetLastError());
    FL_HIP_TRY(ctx, hipMemcpyAsync(p_out, u->out, batch * nv * sizeof(float), hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
    return FL_OK;
}

fl_status fl_unet_forward_raw(fl_unet *u, int32_t batch, const float *rhs, float *out16) {
    if (!u || !rhs || !out16) return set_error(u ? u->ctx : nullptr, FL_ERR_INVALID_ARGUMENT, "null argument");
    if (batch < 1 || batch > u->maxb)
        return set_error(u->ctx, FL_ERR_INVALID_ARGUMENT, "batch %d outside [1, %d]", batch, u->maxb);
    fl_ctx *ctx = u->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t stm = ctx->stream;
    const long long nv = nvox_at(u, 0);
    FL_HIP_TRY(ctx, hipMemcpyAsync(u->in, rhs, batch * nv * sizeof(float), hipMemcpyHostToDevice, stm));
    launch_pack(u, stm, u->in, batch);
    fl_status st = forward(u, stm, batch, u->in, false);
    if (st != FL_OK) return st;
    std::vector<uint16_t> raw((size_t)batch * nv * 16);
    FL_HIP_TRY(ctx, hipMemcpyAsync(raw.data(), u->act[OUT], raw.size() * 2, hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
    for (size_t i = 0; i < raw.size(); ++i) {
        const uint32_t bits = (uint32_t)raw[i] << 16;
        std::memcpy(out16 + i, &bits, 4);
    }
    return FL_OK;
}

fl_status fl_unet_grads(fl_unet *u, int32_t batch, const float *rhs, const float *p, double diag,
                        float *grad_out, float *loss) {
    if (!u || !rhs || !p || !grad_out)
        return set_error(u ? u->ctx : nullptr, FL_ERR_INVALID_ARGUMENT, "null argument");
    if (batch < 1 || batch > u->maxb)
        return set_error(u->ctx, FL_ERR_INVALID_ARGUMENT, "batch %d outside [1, %d]", batch, u->maxb);
    fl_ctx *ctx = u->ctx;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t stm = ctx->stream;
    const long long nv = nvox_at(u, 0);
    FL_HIP_TRY(ctx, hipMemcpyAsync(u->in, rhs, batch * nv * sizeof(float), hipMemcpyHostToDevice, stm));
    FL_HIP_TRY(ctx, hipMemcpyAsync(u->tgt, p, batch * nv * sizeof(float), hipMemcpyHostToDevice, stm));
    launch_pack(u, stm, u->in, batch);
    fl_status st = forward(u, stm, batch, u->in, false);
    if (st != FL_OK) return st;
    launch_loss(u, stm, diag, batch);
    if ((st = backward(u, stm, batch)) != FL_OK) return st;
    FL_HIP_TRY(ctx, hipGetLastError());
    FL_HIP_TRY(ctx, hipMemcpyAsync(grad_out, u->gr, u->P * sizeof(float), hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipMemcpyAsync(u->host_loss, u->loss, batch * sizeof(float), hipMemcpyDeviceToHost, stm));
    FL_HIP_TRY(ctx, hipStreamSynchronize(stm));
    double s = 0.0;
    for (int b = 0; b < batch; ++b) s += u->host_loss[b];
    if (loss) *loss = (float)(s / batch);
    return FL_OK;
}

fl_status fl_probe_mfma_bf16(fl_ctx *ctx, const float *a, const float *b, float *d) {
    if (!ctx || !a || !b || !d) return FL_ERR_INVALID_ARGUMENT;
    FL_HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t stm = ctx->stream;
    float *buf = nullptr;
    FL_HIP_TRY(ctx, hipMalloc(&buf, (512 + 512 + 256) * sizeof(float)));
    fl_status st = FL_OK;
    if (hipMemcpyAsync(buf, a, 512 * 4, hipMemcpyHostToDevice, stm) != hipSuccess ||
        hipMemcpyAsync(buf + 512, b, 512 * 4, hipMemcpyHostToDevice, stm) != hipSuccess)
        st = set_error(ctx, FL_ERR_HIP, "copy failed");
    if (st == FL_OK) {
        k_probe_mfma<<<1, 64, 0, stm>>>(buf, buf + 512, buf + 1024);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(d, buf + 1024, 256 * 4, hipMemcpyDeviceToHost, stm) != hipSuccess ||
            hipStreamSynchronize(stm) != hipSuccess)
            st = set_error(ctx, FL_ERR_HIP, "MFMA probe failed");
    }
    (void)hipFree(buf);
    return st;
}

}  // extern "C"
