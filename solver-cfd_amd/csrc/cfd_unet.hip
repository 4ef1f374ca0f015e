// cfd_unet.hip — the learned warm start: a 4-level convolutional U-Net (channels 32, 64,
// 128, 256; 3x3 kernels) trained and run with hand-written HIP kernels.
//
// Inputs per cell (5 channels): porous-block mask L, porosity eps, and the case's log Re,
// log Da and Ri broadcast over the grid. Outputs (4 channels): u, v, p, Z (T) on the p-cell
// index grid, from which the solver's staggered initial fields are taken.
//
// Numerics: activations and their gradients are stored in BF16 (NHWC), every convolution
// is an implicit GEMM on the CDNA matrix cores (BF16 operands, FP32 accumulation) in
// spatially tiled kernels on the gfx950 v_mfma_f32_32x32x16_bf16; layers whose channel
// counts are off the 32 grid (the first layer's 5 inputs, the head's 4 outputs) use the same
// kernels with element loads and zero fill. Master weights, weight gradients, biases and the
// Adam state are FP32.
//
//   forward   y[m][n]  = sum_k A[m][k] W[n][k] + b[n]     m = pixel, n = cout,
//                                                          k = (dy, dx, cin), A = im2col(x)
//   input grad dx = the same kernel on dy with the flipped, transposed weights
//             Wf[ci][(dy', dx', co)] = W[co][(2 - dy', 2 - dx')][ci]
//   weight grad dW[n][k] = sum_m dy[m][n] A[m][k]           (MFMA over m, split over the grid)
//   bias grad  db[n]    = sum_m dy[m][n]                      (in the weight-gradient pass)
//
// Network: enc1 (5->32->32) pool enc2 (->64->64) pool enc3 (->128->128) pool bottom
// (->256->256); up + concat enc3 -> (384->128->128); up + concat enc2 -> (192->64->64);
// up + concat enc1 -> (96->32->32); head 32->4 (no activation). ReLU after every other conv.
// Pooling is 2x2 max, upsampling nearest x2. Loss: mean squared error over all outputs.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <mutex>
#include <vector>

#include "../../include/cfd.h"

extern "C" int cfd__fail(int code, const char *fmt, ...);

namespace {

typedef uint16_t bf16;
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __host__ inline float bf2f(bf16 h)
{
    union { uint32_t u; float f; } c;
    c.u = (uint32_t)h << 16;
    return c.f;
}
__device__ __host__ inline bf16 f2bf(float f)        // round to nearest even
{
    union { uint32_t u; float f; } c;
    c.f = f;
    const uint32_t u = c.u;
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    return (bf16)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }
inline int launched() { return hipGetLastError() == hipSuccess ? CFD_OK : CFD_ERR_HIP; }

// ---------------------------------------------------------------------------------------
// 3x3 convolution, stride 1, zero padding 1, NHWC, as an implicit GEMM on the matrix cores
// x: [B][H][W][Cin] bf16; w: [Cout][9*Cin] bf16 (k = (dy*3+dx)*Cin + ci); y: [B][H][W][Cout]
//
// Spatial tiling: a 16 x 16 pixel tile plus its one-pixel halo is staged in LDS once per
// 32-channel chunk and serves all nine taps, so each input value leaves L2 about 1.3 times
// instead of nine. Operands go to v_mfma_f32_32x32x16_bf16 (gfx950: 16 k per instruction,
// 8 consecutive channels per lane); the next chunk's loads are issued into registers before
// the current chunk's MFMAs. RAGGED kernels (channel counts off the 32 grid: the first
// layer's 5 inputs, the head's 4 outputs) load element by element and zero-fill the chunk.
// ---------------------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int TT = 16;                 // output tile edge (pixels)
constexpr int TH2 = TT + 2;            // halo tile edge
constexpr int CK = 32;                 // channels per chunk
constexpr int CP = CK + 8;             // LDS pitch per pixel / per weight row: 80 B, conflict-free 16 B reads

// 8 consecutive channels c0..c0+7 of one pixel row `p` (channel count C): one 16 B load, or
// element loads with zeros past C
template <bool RAGGED>
__device__ __forceinline__ uint4 load8(const bf16 *p, int c0, int C)
{
    if (!RAGGED) return *(const uint4 *)(p + c0);
    union { uint4 v; bf16 e[8]; } u;
#pragma unroll
    for (int q = 0; q < 8; ++q) u.e[q] = c0 + q < C ? p[c0 + q] : (bf16)0;
    return u.v;
}

template <int BNT>
struct __align__(16) ConvSmem {     // 16 B aligned: operand reads are ds_read_b128
    bf16 xs[TH2 * TH2][CP];            // halo tile, one chunk of channels
    bf16 ws[9][BNT][CP];               // weights of the chunk for all taps
};

template <int BNT, bool RELU, bool RAGGED>
__global__ __launch_bounds__(256, 2) void k_conv3x3_t(const bf16 *__restrict__ x, int B, int H, int W, int Cin,
                                                      const bf16 *__restrict__ w, const float *__restrict__ bias,
                                                      int Cout, bf16 *__restrict__ y, float *__restrict__ y32)
{
    constexpr int NB = BNT / 32;                        // 32-wide cout blocks per wave
    constexpr int NXV = TH2 * TH2 * (CK / 8);           // 16 B vectors of the halo tile
    constexpr int NWV = 9 * BNT * (CK / 8);             // 16 B vectors of the weights
    constexpr int RX = (NXV + 255) / 256, RW = (NWV + 255) / 256;
    __shared__ ConvSmem<BNT> sm;
    const int tw = (W + TT - 1) / TT, th = (H + TT - 1) / TT;
    int tile = blockIdx.x;
    const int tx = tile % tw; tile /= tw;
    const int ty = tile % th;
    const int b = tile / th;
    const int h0 = ty * TT, w0 = tx * TT, n0 = blockIdx.y * BNT;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    const size_t K = (size_t)9 * Cin;
    const int nchunk = (Cin + CK - 1) / CK;

    uint4 rx[RX], rw[RW];
    auto fetch = [&](int c) {
        const int ci0 = c * CK;
#pragma unroll
        for (int i = 0; i < RX; ++i) {
            const int v = t + 256 * i;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (v < NXV) {
                const int pix = v >> 2, qd = v & 3;
                const int hh = h0 - 1 + pix / TH2, ww = w0 - 1 + pix % TH2;
                if (hh >= 0 && hh < H && ww >= 0 && ww < W)
                    val = load8<RAGGED>(x + ((size_t)(b * H + hh) * W + ww) * Cin, ci0 + 8 * qd, Cin);
            }
            rx[i] = val;
        }
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const int v = t + 256 * i;
            const int n = v / 36, rem = v - n * 36, tap = rem >> 2, qd = rem & 3;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (v < NWV && (!RAGGED || n0 + n < Cout))
                val = load8<RAGGED>(w + (size_t)(n0 + n) * K + (size_t)tap * Cin, ci0 + 8 * qd, Cin);
            rw[i] = val;
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int i = 0; i < RX; ++i) {
            const int v = t + 256 * i;
            if (v < NXV) *(uint4 *)&sm.xs[v >> 2][8 * (v & 3)] = rx[i];
        }
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const int v = t + 256 * i;
            const int n = v / 36, rem = v - n * 36, tap = rem >> 2, qd = rem & 3;
            if (v < NWV) *(uint4 *)&sm.ws[tap][n][8 * qd] = rw[i];
        }
    };

    f32x16 acc[2][NB];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[mb][nb][q] = 0.f;

    // this lane's A pixel of m-block mb: tile row 4 wv + 2 mb + l32 / 16, column l32 % 16
    const int prow = 4 * wv + (l32 >> 4), pcol = l32 & 15;
    fetch(0);
    for (int c = 0; c < nchunk; ++c) {
        __syncthreads();                 // the previous chunk's reads are done
        stash();
        __syncthreads();
        if (c + 1 < nchunk) fetch(c + 1);
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap - 3 * (tap / 3);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int kc = 16 * s + 8 * half;
                bf16x8 a[2], bb[NB];
#pragma unroll
                for (int mb = 0; mb < 2; ++mb)
                    a[mb] = *(const bf16x8 *)&sm.xs[(prow + 2 * mb + dy) * TH2 + pcol + dx][kc];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) bb[nb] = *(const bf16x8 *)&sm.ws[tap][32 * nb + l32][kc];
#pragma unroll
                for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[mb][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[mb], bb[nb], acc[mb][nb], 0, 0, 0);
            }
        }
    }
    // D layout of the 32x32 MFMA: lane l, element r -> row 8 (r/4) + 4 (l/32) + r%4, col l%32
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
        const int n = n0 + 32 * nb + l32;
        if (RAGGED && n >= Cout) continue;
        const float bn = bias ? bias[n] : 0.f;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = 8 * (r >> 2) + 4 * half + (r & 3);
                const int hh = h0 + 4 * wv + 2 * mb + (m >> 4), ww = w0 + (m & 15);
                if (hh < H && ww < W) {
                    float v = acc[mb][nb][r] + bn;
                    if (RELU) v = v > 0.f ? v : 0.f;
                    const size_t o = ((size_t)(b * H + hh) * W + ww) * Cout + n;
                    if (y) y[o] = f2bf(v);
                    if (y32) y32[o] = v;
                }
            }
    }
}

// ---------------------------------------------------------------------------------------
// spatially tiled weight gradient:
//   dW[n][tap][ci] = sum over pixels p of dy[p][n] * x[p + offset(tap)][ci]
// The reduction runs over pixels, which are strided in NHWC, so the block stages both
// operands transposed in LDS: dy of a 16 x 16 pixel tile as dyT[n][pixel], and the 18 x 18
// halo tile of x as xs[ci][halo pixel]. A tap's operand is then 8 consecutive halo pixels
// from column 8h + dx of a halo row: a 16 B LDS read at a 2 B-aligned address, which gfx950
// serves in one ds_read_b128 (unaligned LDS access).
// Block (tile group, ci chunk, n block) of three waves: wave w owns the three taps of column
// shift dx = w. It walks the 18 halo rows of x once: halo row rr meets dy rows rr, rr-1 and
// rr-2 (taps dy = 0, 1, 2), so each x operand read serves three MFMAs and each dy operand
// (kept in a three-row register window) three more: 34 LDS reads per 48 MFMAs per tile,
// where one tap per read would take 96. The next tile's loads are in flight during the
// current tile's MFMAs; the waves' outputs are disjoint, so each writes its sums straight to
// the block's slice of a partial buffer, which k_wsum reduces over the groups (no atomics).
// ---------------------------------------------------------------------------------------
constexpr int PPD = TT * TT + 8;            // dyT pitch (bf16): 528 B, conflict-free 16 B reads
constexpr int PPX = TH2 * TH2 + 4;          // xs pitch per channel (bf16): 656 B = 164 dwords
constexpr int WG_THREADS = 192;             // three waves
#ifndef WG_WAVES_PER_SIMD
#define WG_WAVES_PER_SIMD 2                 // 202 registers; at 3 the operand prefetch spills
#endif
#ifndef WG_TARGET_BLOCKS
#define WG_TARGET_BLOCKS 1024               // weight-gradient blocks per launch (tile groups x channel blocks)
#endif

struct __align__(16) WgradSmem {
    struct {
        bf16 dyT[32][PPD];
        bf16 xs[CK][PPX];
    } t;
};

// 16 B from LDS at a 2 B-aligned address
__device__ __forceinline__ bf16x8 lds_read8_unaligned(const bf16 *p)
{
    bf16x8 v;
    __builtin_memcpy(&v, __builtin_assume_aligned(p, 2), 16);
    return v;
}

template <bool RAGGED>
__global__ __launch_bounds__(WG_THREADS, WG_WAVES_PER_SIMD) void k_conv3x3_wgrad_t(const bf16 *__restrict__ x,
                                                                   const bf16 *__restrict__ dy, int B, int H, int W,
                                                                   int Cin, int Cout, float *__restrict__ part,
                                                                   long pstride, float *__restrict__ part_b,
                                                                   long bstride)
{
    // with part_b, the blocks of the first channel chunk also sum dy over their pixels (the
    // bias gradient): thread t < 128 adds pixels 64 (t / 32) .. + 63 of cout t % 32 of each tile
    const bool do_bias = part_b != nullptr && blockIdx.y == 0;
    float bsum = 0.f;
    constexpr int NXP = TH2 * TH2;
    // operand moves: a unit is one 8-channel group of 64 consecutive pixels (one 16 B load
    // per lane); dy has 4 groups x 4 pixel steps, x 4 groups x 6 pixel steps (324 pixels),
    // dealt round-robin to the three waves
    constexpr int NUY = 16, NUX = 24, RY = (NUY + 2) / 3, RX = (NUX + 2) / 3;
    __shared__ WgradSmem sm;
    const int tw = (W + TT - 1) / TT, th = (H + TT - 1) / TT, ntiles = B * th * tw;
    const int ci0 = blockIdx.y * CK, n0 = blockIdx.z * 32;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    f32x16 acc[3];                               // taps (0, wv), (1, wv), (2, wv)
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[q][e] = 0.f;

    uint4 ry[RY], rx[RX];
    auto fetch = [&](int tile) {
        const int tx = tile % tw, ty = (tile / tw) % th, b = tile / (tw * th);
        const int h0 = ty * TT, w0 = tx * TT;
#pragma unroll
        for (int k = 0; k < RY; ++k) {           // dy: 256 pixels x 32 couts
            const int u = wv + 3 * k, qd = u & 3, pix = 64 * (u >> 2) + lane;
            const int hh = h0 + (pix >> 4), ww = w0 + (pix & 15);
            uint4 val = make_uint4(0, 0, 0, 0);
            if (u < NUY && hh < H && ww < W)
                val = load8<RAGGED>(dy + ((size_t)(b * H + hh) * W + ww) * Cout, n0 + 8 * qd, Cout);
            ry[k] = val;
        }
#pragma unroll
        for (int k = 0; k < RX; ++k) {           // x: 18 x 18 halo pixels x 32 channels
            const int u = wv + 3 * k, qd = u & 3, pix = 64 * (u >> 2) + lane;
            const int hh = h0 - 1 + pix / TH2, ww = w0 - 1 + pix % TH2;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (u < NUX && pix < NXP && hh >= 0 && hh < H && ww >= 0 && ww < W)
                val = load8<RAGGED>(x + ((size_t)(b * H + hh) * W + ww) * Cin, ci0 + 8 * qd, Cin);
            rx[k] = val;
        }
    };
    // transposed stores: the 64 lanes of a unit write 64 consecutive pixels of one LDS row
    auto stash = [&]() {
#pragma unroll
        for (int k = 0; k < RY; ++k) {
            const int u = wv + 3 * k, qd = u & 3, pix = 64 * (u >> 2) + lane;
            if (u >= NUY) continue;
            union { uint4 v; bf16 e[8]; } w;
            w.v = ry[k];
#pragma unroll
            for (int q = 0; q < 8; ++q) sm.t.dyT[8 * qd + q][pix] = w.e[q];
        }
#pragma unroll
        for (int k = 0; k < RX; ++k) {
            const int u = wv + 3 * k, qd = u & 3, pix = 64 * (u >> 2) + lane;
            if (u >= NUX || pix >= NXP) continue;
            union { uint4 v; bf16 e[8]; } w;
            w.v = rx[k];
#pragma unroll
            for (int q = 0; q < 8; ++q) sm.t.xs[8 * qd + q][pix] = w.e[q];
        }
    };

    if (blockIdx.x < ntiles) fetch(blockIdx.x);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        __syncthreads();                    // the previous tile's operand reads are done
        stash();
        __syncthreads();
        if (tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
        if (do_bias && t < 128) {
            const uint4 *row = (const uint4 *)&sm.t.dyT[t & 31][64 * (t >> 5)];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const uint4 q = row[v];
                const bf16 *e = (const bf16 *)&q;
#pragma unroll
                for (int k = 0; k < 8; ++k) bsum += bf2f(e[k]);
            }
        }
        // halo row rr pairs with dy rows rr - dy for dy = 0, 1, 2 (those inside 0..15); the
        // window holds dy rows rr, rr - 1, rr - 2
        bf16x8 a0, a1, a2;
#pragma unroll
        for (int rr = 0; rr < TH2; ++rr) {
            const bf16x8 bb = lds_read8_unaligned(&sm.t.xs[l32][rr * TH2 + 8 * half + wv]);
            a2 = a1; a1 = a0;
            if (rr < TT) a0 = *(const bf16x8 *)&sm.t.dyT[l32][16 * rr + 8 * half];
            if (rr < TT) acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, bb, acc[0], 0, 0, 0);
            if (rr >= 1 && rr - 1 < TT) acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, bb, acc[1], 0, 0, 0);
            if (rr >= 2) acc[2] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, bb, acc[2], 0, 0, 0);
        }
    }
    if (do_bias) {                          // the four pixel groups of each cout meet in LDS
        __syncthreads();
        float *red = (float *)&sm.t.dyT[0][0];
        if (t < 128) red[t] = bsum;
        __syncthreads();
        if (t < 32 && (!RAGGED || n0 + t < Cout)) {
            float sb = 0.f;
#pragma unroll
            for (int g = 0; g < 4; ++g) sb += red[32 * g + t];
            part_b[(size_t)blockIdx.x * bstride + n0 + t] = sb;
        }
    }
    // the block's sums go to its own slice of the partial buffer, part + blockIdx.x pstride,
    // laid out [Cout][9 Cin] (reduced over the tile groups later); D rows are couts, columns
    // input channels
    const size_t K = (size_t)9 * Cin;
    float *out = part + (size_t)blockIdx.x * pstride;
    const int ci = ci0 + l32;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int tap = 3 * q + wv;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int n = n0 + 8 * (e >> 2) + 4 * half + (e & 3);
            if (RAGGED && (n >= Cout || ci >= Cin)) continue;
            out[(size_t)n * K + (size_t)tap * Cin + ci] = acc[q][e];
        }
    }
}

// dw[i] += sum over g < G of part[g stride + i], i < count, in a fixed order (the training
// is reproducible run to run). blockIdx.y takes a slice of the groups; with several slices
// each one leaves its sum in the first row of its slice (only this thread reads column i
// of those rows) and k_wsum_final adds the slice sums in order.
__global__ void k_wsum(float *__restrict__ part, int G, int gper, long count, long stride, float *__restrict__ dw)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int g0 = blockIdx.y * gper, g1 = min(G, g0 + gper);
    float s = 0.f;
    for (int g = g0; g < g1; ++g) s += part[(size_t)g * stride + i];
    if (gridDim.y == 1) dw[i] += s;
    else part[(size_t)g0 * stride + i] = s;
}

__global__ void k_wsum_final(const float *__restrict__ part, int nslices, int gper, long count, long stride,
                             float *__restrict__ dw)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    // the slice sums are independent loads: eight in flight, added in slice order
    float s = 0.f;
    int q = 0;
    for (; q + 8 <= nslices; q += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = part[(size_t)(q + k) * gper * stride + i];
#pragma unroll
        for (int k = 0; k < 8; ++k) s += v[k];
    }
    for (; q < nslices; ++q) s += part[(size_t)q * gper * stride + i];
    dw[i] += s;
}

// ---------------------------------------------------------------------------------------
// elementwise and data-movement kernels
// ---------------------------------------------------------------------------------------
// G channels per thread: 8 (one 16 B vector of BF16, two of FP32) when the channel counts
// and offsets are multiples of 8, else 1. Index math is 32-bit: activation tensors are
// limited to 2^31 elements (checked by the callers).
template <int G>
struct VB {                                      // G BF16 values
    union { uint4 v; bf16 e[8]; } u;
    __device__ __forceinline__ void load(const bf16 *p)
    {
        if (G == 8) u.v = *(const uint4 *)p;
        else u.e[0] = *p;
    }
    __device__ __forceinline__ void store(bf16 *p) const
    {
        if (G == 8) *(uint4 *)p = u.v;
        else *p = u.e[0];
    }
};

template <int G>
__global__ void k_relu_bwd(bf16 *g, const bf16 *y, int n)          // g *= (y > 0)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n / G; i += gridDim.x * blockDim.x) {
        VB<G> a, m;
        a.load(g + G * i); m.load(y + G * i);
#pragma unroll
        for (int k = 0; k < G; ++k)
            if (!(bf2f(m.u.e[k]) > 0.f)) a.u.e[k] = 0;
        a.store(g + G * i);
    }
}

template <int G>
__global__ void k_maxpool2(const bf16 *x, int B, int H, int W, int C, bf16 *y)
{
    const int Ho = H / 2, Wo = W / 2, CG = C / G, n = B * Ho * Wo * CG;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int cg = i % CG;
        int p = i / CG;
        const int wo = p % Wo; p /= Wo;
        const int ho = p % Ho, b = p / Ho;
        const size_t base = ((size_t)(b * H + 2 * ho) * W + 2 * wo) * C + G * cg;
        VB<G> a0, a1, a2, a3, o;
        a0.load(x + base); a1.load(x + base + C); a2.load(x + base + (size_t)W * C); a3.load(x + base + (size_t)W * C + C);
#pragma unroll
        for (int k = 0; k < G; ++k) {
            float m = bf2f(a0.u.e[k]);
            m = fmaxf(m, bf2f(a1.u.e[k]));
            m = fmaxf(m, bf2f(a2.u.e[k]));
            m = fmaxf(m, bf2f(a3.u.e[k]));
            o.u.e[k] = f2bf(m);
        }
        o.store(y + (size_t)G * i);
    }
}

// gradient to the first of the four inputs that holds the maximum (row-major order)
template <int G>
__global__ void k_maxpool2_bwd(const bf16 *x, const bf16 *y, const bf16 *gy, int B, int H, int W, int C, float *gx)
{
    const int Ho = H / 2, Wo = W / 2, CG = C / G, n = B * Ho * Wo * CG;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int cg = i % CG;
        int p = i / CG;
        const int wo = p % Wo; p /= Wo;
        const int ho = p % Ho, b = p / Ho;
        const size_t base = ((size_t)(b * H + 2 * ho) * W + 2 * wo) * C + G * cg;
        const size_t off[4] = {0, (size_t)C, (size_t)W * C, (size_t)W * C + C};
        VB<G> a[4], m, gg;
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q].load(x + base + off[q]);
        m.load(y + (size_t)G * i); gg.load(gy + (size_t)G * i);
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const float mv = bf2f(m.u.e[k]);           // compared as values: -0 equals +0
            for (int q = 0; q < 4; ++q)
                if (bf2f(a[q].u.e[k]) == mv) { gx[base + off[q] + k] += bf2f(gg.u.e[k]); break; }
        }
    }
}

// x: [B][H][W][C] -> y[B][2H][2W][Cy] channels cofs..cofs+C-1
template <int G>
__global__ void k_upsample2(const bf16 *x, int B, int H, int W, int C, bf16 *y, int Cy, int cofs)
{
    const int CG = C / G, n = B * (2 * H) * (2 * W) * CG;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int cg = i % CG;
        int p = i / CG;
        const int wo = p % (2 * W); p /= 2 * W;
        const int ho = p % (2 * H), b = p / (2 * H);
        VB<G> a;
        a.load(x + ((size_t)(b * H + ho / 2) * W + wo / 2) * C + G * cg);
        a.store(y + ((size_t)(b * 2 * H + ho) * 2 * W + wo) * Cy + cofs + G * cg);
    }
}

// gx[B][H][W][C] += sum over the 2x2 children of gy channels cofs.. (gy: [B][2H][2W][Cy])
template <int G>
__global__ void k_upsample2_bwd(const bf16 *gy, int B, int H, int W, int C, int Cy, int cofs, float *gx)
{
    const int CG = C / G, n = B * H * W * CG;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int cg = i % CG;
        int p = i / CG;
        const int w = p % W; p /= W;
        const int h = p % H, b = p / H;
        float s[G];
#pragma unroll
        for (int k = 0; k < G; ++k) s[k] = 0.f;
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                VB<G> a;
                a.load(gy + ((size_t)(b * 2 * H + 2 * h + dy) * 2 * W + 2 * w + dx) * Cy + cofs + G * cg);
#pragma unroll
                for (int k = 0; k < G; ++k) s[k] += bf2f(a.u.e[k]);
            }
        float *o = gx + (size_t)G * i;
#pragma unroll
        for (int k = 0; k < G; ++k) o[k] += s[k];
    }
}

// cat[B][2H][2W][Cl + Cs] = [upsample2(low) | skip]: low [B][H][W][Cl], skip [B][2H][2W][Cs]
template <int G>
__global__ void k_up_cat(const bf16 *low, const bf16 *skip, int B, int H, int W, int Cl, int Cs, bf16 *cat)
{
    const int Cc = Cl + Cs, CG = Cc / G, n = B * (2 * H) * (2 * W) * CG;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int c = G * (i % CG);
        const int p = i / CG;                          // output pixel (b, ho, wo)
        const int wo = p % (2 * W), q = p / (2 * W);
        const int ho = q % (2 * H), b = q / (2 * H);
        VB<G> a;
        if (c < Cl) a.load(low + ((size_t)(b * H + ho / 2) * W + wo / 2) * Cl + c);
        else a.load(skip + (size_t)p * Cs + (c - Cl));
        a.store(cat + (size_t)G * i);
    }
}

// Backward of [pool -> ... -> upsample, concat] around a skip connection, in one pass:
// g[B][H][W][C] = relu'(x) * (gc[..][cofs + c] + the pooled gradient gy routed to the first of
// the four inputs that holds the maximum), rounded once to BF16. x is the ReLU output that
// was pooled (y = maxpool(x)) and concatenated (channels cofs.. of the Cc-channel gc).
template <int G>
__global__ void k_pool_skip_bwd(const bf16 *x, const bf16 *y, const bf16 *gy, const bf16 *gc, int Cc, int cofs, int B,
                                int H, int W, int C, bf16 *g)
{
    const int Ho = H / 2, Wo = W / 2, CG = C / G, n = B * Ho * Wo * CG;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int cg = i % CG;
        int p = i / CG;
        const int wo = p % Wo; p /= Wo;
        const int ho = p % Ho, b = p / Ho;
        const int pix = (b * H + 2 * ho) * W + 2 * wo;
        const int dp[4] = {0, 1, W, W + 1};
        VB<G> a[4], sk[4], m, gg;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a[q].load(x + (size_t)(pix + dp[q]) * C + G * cg);
            sk[q].load(gc + (size_t)(pix + dp[q]) * Cc + cofs + G * cg);
        }
        m.load(y + (size_t)G * i); gg.load(gy + (size_t)G * i);
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const float mv = bf2f(m.u.e[k]);           // compared as values: -0 equals +0
            bool routed = false;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float xv = bf2f(a[q].u.e[k]);
                float v = bf2f(sk[q].u.e[k]);
                if (!routed && xv == mv) { v += bf2f(gg.u.e[k]); routed = true; }
                sk[q].u.e[k] = xv > 0.f ? f2bf(v) : (bf16)0;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) sk[q].store(g + (size_t)(pix + dp[q]) * C + G * cg);
    }
}

// g[B][H][W][C] = relu'(x) * (sum over the 2x2 children of gy channels cofs..), rounded once
// to BF16 (gy: [B][2H][2W][Cy]; x the ReLU output that was upsampled)
template <int G>
__global__ void k_up_relu_bwd(const bf16 *gy, int B, int H, int W, int C, int Cy, int cofs, const bf16 *x, bf16 *g)
{
    const int CG = C / G, n = B * H * W * CG;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int cg = i % CG;
        int p = i / CG;
        const int w = p % W; p /= W;
        const int h = p % H, b = p / H;
        float s[G];
#pragma unroll
        for (int k = 0; k < G; ++k) s[k] = 0.f;
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                VB<G> a;
                a.load(gy + ((size_t)(b * 2 * H + 2 * h + dy) * 2 * W + 2 * w + dx) * Cy + cofs + G * cg);
#pragma unroll
                for (int k = 0; k < G; ++k) s[k] += bf2f(a.u.e[k]);
            }
        VB<G> m;
        m.load(x + (size_t)G * i);
#pragma unroll
        for (int k = 0; k < G; ++k) m.u.e[k] = bf2f(m.u.e[k]) > 0.f ? f2bf(s[k]) : (bf16)0;
        m.store(g + (size_t)G * i);
    }
}

// y[i] = bf16(x[i]) for i in [i0, n): vectors of 8 from i0 (a multiple of 8) when G == 8
template <int G>
__global__ void k_f32_to_bf16(const float *x, long i0, long n, bf16 *y)
{
    const long nv = (n - i0) / G;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (long)gridDim.x * blockDim.x) {
        const long e = i0 + G * i;
        VB<G> o;
#pragma unroll
        for (int k = 0; k < G; ++k) o.u.e[k] = f2bf(x[e + k]);
        o.store(y + e);
    }
}

__global__ void k_fill_f32(float *x, long n, float v)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}

// Wf[ci][(dy', dx', co)] = W[co][((2-dy')*3 + 2-dx')][ci], from the FP32 master
__global__ void k_flip_weights(const float *w, int Cin, int Cout, bf16 *wf)
{
    const long n = (long)Cin * 9 * Cout;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int ci = (int)(i / (9 * Cout));
        const int rem = (int)(i % (9 * Cout));
        const int r = rem / Cout, co = rem % Cout;
        const int rf = 8 - r;                  // (2-dy')*3 + (2-dx')
        wf[i] = f2bf(w[((size_t)co * 9 + rf) * Cin + ci]);
    }
}

// The flipped weights of every convolution of the network in one launch: layer q covers
// elements foff[q] .. foff[q+1] of wf and reads its FP32 master at woff[q]
struct FlipTable {
    long foff[16], woff[15];
    int cin[15], cout[15];
};

__global__ void k_flip_all(const float *w, FlipTable t, bf16 *wf)
{
    const long n = t.foff[15];
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        int q = 0;
        while (i >= t.foff[q + 1]) ++q;
        const long j = i - t.foff[q];
        const int Cin = t.cin[q], Cout = t.cout[q];
        const int ci = (int)(j / (9 * Cout));
        const int rem = (int)(j % (9 * Cout));
        const int r = rem / Cout, co = rem % Cout;
        wf[i] = f2bf(w[t.woff[q] + ((size_t)co * 9 + (8 - r)) * Cin + ci]);
    }
}

// loss = mean (p - t)^2 (partial sums into *loss), g = 2 (p - t) / n in bf16
__global__ void k_mse(const float *p, const float *t, long n, bf16 *g, double *loss)
{
    __shared__ double part[256];
    double s = 0.0;
    const float inv = 2.f / (float)n;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float d = p[i] - t[i];
        s += (double)d * d;
        g[i] = f2bf(inv * d);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0] / (double)n);
}

// Adam on the FP32 masters, which also leave as the BF16 weights of the forward pass
__global__ void k_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1, float b2,
                       float eps, float c1, float c2, bf16 *wb)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi; v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        wb[i] = f2bf(wi);
    }
}

inline int eb(long n) { const long b = (n + 255) / 256; return (int)(b < 8192 ? (b > 0 ? b : 1) : 8192); }

// Partial sums of the weight gradients: one device buffer for the process, grown on
// demand. The nn operations run in order on the NULL stream, so consecutive calls share it;
// callers that issue them from several host threads at once must serialise them.
std::mutex g_ws_mu;
float *g_ws = nullptr;
size_t g_ws_n = 0;

float *workspace(size_t n)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    if (n > g_ws_n) {
        if (g_ws) {
            (void)hipDeviceSynchronize();
            (void)hipFree(g_ws);
            g_ws = nullptr; g_ws_n = 0;
        }
        if (hipMalloc((void **)&g_ws, n * sizeof(float)) != hipSuccess) { g_ws = nullptr; return nullptr; }
        g_ws_n = n;
    }
    return g_ws;
}

}  // namespace

// ---------------------------------------------------------------------------------------
// op-level C-ABI (device pointers, NULL stream) — the network is built from these
// ---------------------------------------------------------------------------------------
extern "C" {

int cfd_nn_conv3x3(const uint16_t *x, int32_t B, int32_t H, int32_t W, int32_t Cin, const uint16_t *w,
                   const float *bias, int32_t Cout, int32_t relu, uint16_t *y, float *y32)
{
    if (B < 1 || H < 1 || W < 1 || Cin < 1 || Cout < 1) return cfd__fail(CFD_ERR_ARG, "cfd_nn_conv3x3: bad shape");
    const int tiles = B * cdiv(H, TT) * cdiv(W, TT);
    const bool ragged = (Cin % CK) != 0 || (Cout % 32) != 0;
#define CONV_LAUNCH(BNT, RAG)                                                                               \
    {                                                                                                       \
        const dim3 g(tiles, cdiv(Cout, BNT));                                                               \
        if (relu) hipLaunchKernelGGL((k_conv3x3_t<BNT, true, RAG>), g, dim3(256), 0, 0, x, B, H, W, Cin, w, bias, \
                                     Cout, y, y32);                                                         \
        else hipLaunchKernelGGL((k_conv3x3_t<BNT, false, RAG>), g, dim3(256), 0, 0, x, B, H, W, Cin, w, bias, \
                                Cout, y, y32);                                                              \
    }
    if (ragged) CONV_LAUNCH(32, true)
    else if (Cout % 64 == 0) CONV_LAUNCH(64, false)
    else CONV_LAUNCH(32, false)
#undef CONV_LAUNCH
    return launched();
}

}  // extern "C"

namespace {

// dw += weight gradient; with db, db += bias gradient (sum of dy over pixels) in the same pass.
// When db follows dw in memory (db == dw + 9 Cin Cout, the network's parameter layout), each
// block's weight and bias partials share one slice and one reduction launch covers both.
int wgrad(const uint16_t *x, const uint16_t *dy, int32_t B, int32_t H, int32_t W, int32_t Cin, int32_t Cout,
          float *dw, float *db)
{
    if (B < 1 || H < 1 || W < 1 || Cin < 1 || Cout < 1) return cfd__fail(CFD_ERR_ARG, "cfd_nn_conv3x3_wgrad: bad shape");
    const long total = 9L * Cin * Cout;
    const bool ragged = (Cin % CK) != 0 || (Cout % 32) != 0;
    // tile groups so that the grid holds about WG_TARGET_BLOCKS blocks
    const int ntiles = B * cdiv(H, TT) * cdiv(W, TT);
    const int per = cdiv(Cin, CK) * cdiv(Cout, 32);
    int G = cdiv(WG_TARGET_BLOCKS, per);
    if (G > ntiles) G = ntiles;
    if (G < 1) G = 1;
    float *part = workspace((size_t)G * (total + Cout));
    if (!part) return cfd__fail(CFD_ERR_HIP, "wgrad workspace of %ld floats", (long)G * (total + Cout));
    const bool joint = db && db == dw + total;
    const long pstride = joint ? total + Cout : total;
    float *part_b = !db ? nullptr : (joint ? part + total : part + (size_t)G * total);
    const long bstride = joint ? pstride : Cout;
    const dim3 grid(G, cdiv(Cin, CK), cdiv(Cout, 32));
    if (ragged)
        hipLaunchKernelGGL(k_conv3x3_wgrad_t<true>, grid, dim3(WG_THREADS), 0, 0, x, dy, B, H, W, Cin, Cout, part,
                           pstride, part_b, bstride);
    else
        hipLaunchKernelGGL(k_conv3x3_wgrad_t<false>, grid, dim3(WG_THREADS), 0, 0, x, dy, B, H, W, Cin, Cout, part,
                           pstride, part_b, bstride);
    int r = launched();
    if (r != CFD_OK) return r;
    if (db && !joint) {
        const int gyb = G < 64 ? G : 64, gperb = cdiv(G, gyb);
        const int gyb2 = cdiv(G, gperb);
        hipLaunchKernelGGL(k_wsum, dim3(cdiv(Cout, 256), gyb2), dim3(256), 0, 0, part_b, G, gperb, (long)Cout,
                           (long)Cout, db);
        if (gyb2 > 1)
            hipLaunchKernelGGL(k_wsum_final, dim3(cdiv(Cout, 256)), dim3(256), 0, 0, (const float *)part_b, gyb2,
                               gperb, (long)Cout, (long)Cout, db);
        r = launched();
        if (r != CFD_OK) return r;
    }
    // reduce over the groups: about 2048 blocks, each summing a slice of the groups
    const long count = joint ? total + Cout : total;
    const int nbx = cdiv(count, 256);
    int gy = 2048 / nbx;
    if (gy > G) gy = G;
    if (gy < 1) gy = 1;
    const int gper = cdiv(G, gy);
    gy = cdiv(G, gper);
    hipLaunchKernelGGL(k_wsum, dim3(nbx, gy), dim3(256), 0, 0, part, G, gper, count, pstride, dw);
    if (gy > 1)
        hipLaunchKernelGGL(k_wsum_final, dim3(nbx), dim3(256), 0, 0, (const float *)part, gy, gper, count, pstride, dw);
    return launched();
}

}  // namespace

extern "C" {

int cfd_nn_conv3x3_wgrad(const uint16_t *x, const uint16_t *dy, int32_t B, int32_t H, int32_t W, int32_t Cin,
                         int32_t Cout, float *dw, float *db)
{
    return wgrad(x, dy, B, H, W, Cin, Cout, dw, db);
}

int cfd_nn_flip_weights(const float *w, int32_t Cin, int32_t Cout, uint16_t *wf)
{
    hipLaunchKernelGGL(k_flip_weights, dim3(eb((long)Cin * 9 * Cout)), dim3(256), 0, 0, w, Cin, Cout, wf);
    return launched();
}

#define ELEMS_OK(n, what)                                                                                  \
    if ((n) <= 0 || (n) >= 2147483648L) return cfd__fail(CFD_ERR_ARG, what ": %ld elements (1 .. 2^31-1)", (long)(n))

int cfd_nn_maxpool2(const uint16_t *x, int32_t B, int32_t H, int32_t W, int32_t C, uint16_t *y)
{
    const long n = (long)B * H * W * C;
    ELEMS_OK(n, "cfd_nn_maxpool2");
    if (C % 8 == 0) hipLaunchKernelGGL(k_maxpool2<8>, dim3(eb(n / 32)), dim3(256), 0, 0, x, B, H, W, C, y);
    else hipLaunchKernelGGL(k_maxpool2<1>, dim3(eb(n / 4)), dim3(256), 0, 0, x, B, H, W, C, y);
    return launched();
}

int cfd_nn_maxpool2_bwd(const uint16_t *x, const uint16_t *y, const uint16_t *gy, int32_t B, int32_t H, int32_t W,
                        int32_t C, float *gx)
{
    const long n = (long)B * H * W * C;
    ELEMS_OK(n, "cfd_nn_maxpool2_bwd");
    if (C % 8 == 0) hipLaunchKernelGGL(k_maxpool2_bwd<8>, dim3(eb(n / 32)), dim3(256), 0, 0, x, y, gy, B, H, W, C, gx);
    else hipLaunchKernelGGL(k_maxpool2_bwd<1>, dim3(eb(n / 4)), dim3(256), 0, 0, x, y, gy, B, H, W, C, gx);
    return launched();
}

int cfd_nn_upsample2(const uint16_t *x, int32_t B, int32_t H, int32_t W, int32_t C, uint16_t *y, int32_t Cy,
                     int32_t cofs)
{
    const long n = (long)B * 4 * H * W * Cy;
    ELEMS_OK(n, "cfd_nn_upsample2");
    if (C % 8 == 0 && Cy % 8 == 0 && cofs % 8 == 0)
        hipLaunchKernelGGL(k_upsample2<8>, dim3(eb(n / 8)), dim3(256), 0, 0, x, B, H, W, C, y, Cy, cofs);
    else hipLaunchKernelGGL(k_upsample2<1>, dim3(eb(n)), dim3(256), 0, 0, x, B, H, W, C, y, Cy, cofs);
    return launched();
}

int cfd_nn_upsample2_bwd(const uint16_t *gy, int32_t B, int32_t H, int32_t W, int32_t C, int32_t Cy, int32_t cofs,
                         float *gx)
{
    const long n = (long)B * 4 * H * W * Cy;
    ELEMS_OK(n, "cfd_nn_upsample2_bwd");
    if (C % 8 == 0 && Cy % 8 == 0 && cofs % 8 == 0)
        hipLaunchKernelGGL(k_upsample2_bwd<8>, dim3(eb(n / 32)), dim3(256), 0, 0, gy, B, H, W, C, Cy, cofs, gx);
    else hipLaunchKernelGGL(k_upsample2_bwd<1>, dim3(eb(n / 4)), dim3(256), 0, 0, gy, B, H, W, C, Cy, cofs, gx);
    return launched();
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// the network
// ---------------------------------------------------------------------------------------
namespace {

struct Conv {
    int cin, cout;
    bool relu;
    size_t woff, boff;          // offsets into the flat FP32 parameter vector
};

template <typename T>
struct Dev {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t count)
    {
        n = count;
        return hipMalloc((void **)&p, count * sizeof(T) + 256) == hipSuccess ? CFD_OK : CFD_ERR_HIP;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; }
};

}  // namespace

struct cfd_unet {
    int H, W, Bmax;
    std::vector<Conv> convs;
    size_t nparam = 0;
    Dev<float> w32, g32, m32, v32;        // master parameters, gradients, Adam moments
    Dev<bf16> wb, wf;                     // BF16 weights [Cout][9 Cin] and flipped for dgrad
    size_t wf_off[15];
    // activations (bf16) per stage, and their gradients
    Dev<bf16> act[40];
    Dev<bf16> gbuf[40];
    Dev<float> out32, target;
    Dev<bf16> gout;
    Dev<double> loss;
    Dev<float> xin32;
    Dev<bf16> xin;
    long step = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};   // inputs on the device, forward, backward, update
};

namespace {

// conv indices: enc1a enc1b enc2a enc2b enc3a enc3b bota botb dec3a dec3b dec2a dec2b dec1a dec1b head
void build_convs(cfd_unet *u)
{
    const int spec[15][3] = {
        {5, 32, 1}, {32, 32, 1}, {32, 64, 1}, {64, 64, 1}, {64, 128, 1}, {128, 128, 1}, {128, 256, 1}, {256, 256, 1},
        {256 + 128, 128, 1}, {128, 128, 1}, {128 + 64, 64, 1}, {64, 64, 1}, {64 + 32, 32, 1}, {32, 32, 1}, {32, 4, 0}};
    size_t off = 0, foff = 0;
    for (int q = 0; q < 15; ++q) {
        Conv c;
        c.cin = spec[q][0]; c.cout = spec[q][1]; c.relu = spec[q][2] != 0;
        c.woff = off; off += (size_t)c.cout * 9 * c.cin;
        c.boff = off; off += (size_t)c.cout;
        u->convs.push_back(c);
        u->wf_off[q] = foff; foff += (size_t)c.cin * 9 * c.cout;
    }
    u->nparam = off;
}

// activation slots
enum {
    A_IN, A_E1A, A_E1, A_P1, A_E2A, A_E2, A_P2, A_E3A, A_E3, A_P3, A_BA, A_B,
    A_C3, A_D3A, A_D3, A_C2, A_D2A, A_D2, A_C1, A_D1A, A_D1, A_NACT
};

struct Shape { int h, w, c; };

Shape act_shape(const cfd_unet *u, int a)
{
    const int H = u->H, W = u->W;
    switch (a) {
    case A_IN: return {H, W, 5};
    case A_E1A: case A_E1: case A_D1A: case A_D1: return {H, W, 32};
    case A_P1: return {H / 2, W / 2, 32};
    case A_E2A: case A_E2: case A_D2A: case A_D2: return {H / 2, W / 2, 64};
    case A_P2: return {H / 4, W / 4, 64};
    case A_E3A: case A_E3: case A_D3A: case A_D3: return {H / 4, W / 4, 128};
    case A_P3: return {H / 8, W / 8, 128};
    case A_BA: case A_B: return {H / 8, W / 8, 256};
    case A_C3: return {H / 4, W / 4, 256 + 128};
    case A_C2: return {H / 2, W / 2, 128 + 64};
    case A_C1: return {H, W, 64 + 32};
    }
    return {0, 0, 0};
}

long numel(cfd_unet *u, int a, int B)
{
    const Shape s = act_shape(u, a);
    return (long)B * s.h * s.w * s.c;
}

int conv_fwd(cfd_unet *u, int q, int B, int ain, int aout, float *out32)
{
    const Conv &c = u->convs[q];
    const Shape s = act_shape(u, ain);
    return cfd_nn_conv3x3(u->act[ain].p, B, s.h, s.w, c.cin, u->wb.p + c.woff, u->w32.p + c.boff, c.cout,
                          c.relu ? 1 : 0, aout >= 0 ? u->act[aout].p : nullptr, out32);
}

int fill0(float *p, long n)
{
    hipLaunchKernelGGL(k_fill_f32, dim3(eb(n)), dim3(256), 0, 0, p, n, 0.f);
    return launched();
}

int forward(cfd_unet *u, int B)
{
    int r = CFD_OK;
    auto pool = [&](int a, int b) {
        const Shape s = act_shape(u, a);
        return cfd_nn_maxpool2(u->act[a].p, B, s.h, s.w, s.c, u->act[b].p);
    };
    auto up_cat = [&](int low, int skip, int cat) {
        const Shape sl = act_shape(u, low), ss = act_shape(u, skip);
        // channel counts 32..256: multiples of 8
        hipLaunchKernelGGL(k_up_cat<8>, dim3(eb(numel(u, cat, B) / 8)), dim3(256), 0, 0, u->act[low].p,
                           u->act[skip].p, B, sl.h, sl.w, sl.c, ss.c, u->act[cat].p);
        return launched();
    };
    if (r == CFD_OK) r = conv_fwd(u, 0, B, A_IN, A_E1A, nullptr);
    if (r == CFD_OK) r = conv_fwd(u, 1, B, A_E1A, A_E1, nullptr);
    if (r == CFD_OK) r = pool(A_E1, A_P1);
    if (r == CFD_OK) r = conv_fwd(u, 2, B, A_P1, A_E2A, nullptr);
    if (r == CFD_OK) r = conv_fwd(u, 3, B, A_E2A, A_E2, nullptr);
    if (r == CFD_OK) r = pool(A_E2, A_P2);
    if (r == CFD_OK) r = conv_fwd(u, 4, B, A_P2, A_E3A, nullptr);
    if (r == CFD_OK) r = conv_fwd(u, 5, B, A_E3A, A_E3, nullptr);
    if (r == CFD_OK) r = pool(A_E3, A_P3);
    if (r == CFD_OK) r = conv_fwd(u, 6, B, A_P3, A_BA, nullptr);
    if (r == CFD_OK) r = conv_fwd(u, 7, B, A_BA, A_B, nullptr);
    if (r == CFD_OK) r = up_cat(A_B, A_E3, A_C3);
    if (r == CFD_OK) r = conv_fwd(u, 8, B, A_C3, A_D3A, nullptr);
    if (r == CFD_OK) r = conv_fwd(u, 9, B, A_D3A, A_D3, nullptr);
    if (r == CFD_OK) r = up_cat(A_D3, A_E2, A_C2);
    if (r == CFD_OK) r = conv_fwd(u, 10, B, A_C2, A_D2A, nullptr);
    if (r == CFD_OK) r = conv_fwd(u, 11, B, A_D2A, A_D2, nullptr);
    if (r == CFD_OK) r = up_cat(A_D2, A_E1, A_C1);
    if (r == CFD_OK) r = conv_fwd(u, 12, B, A_C1, A_D1A, nullptr);
    if (r == CFD_OK) r = conv_fwd(u, 13, B, A_D1A, A_D1, nullptr);
    if (r == CFD_OK) r = conv_fwd(u, 14, B, A_D1, -1, u->out32.p);
    return r;
}

// Backward through conv q: gy (bf16, [M][cout]) -> dW, db accumulated; gx (bf16 input grad)
// written into gbuf[ain] unless ain == A_IN. For ReLU convs gy must already be masked.
// With relu_in (ain is a ReLU output), the input gradient is then masked by relu'(ain): a pass
// of its own, since a masked store in the convolution's epilogue costs the convolution more
// (a 2 B gather per output element after the last MFMA) than the pass does.
int conv_bwd(cfd_unet *u, int q, int B, int ain, const bf16 *gy, bool relu_in)
{
    const Conv &c = u->convs[q];
    const Shape s = act_shape(u, ain);
    int r = wgrad(u->act[ain].p, gy, B, s.h, s.w, c.cin, c.cout, u->g32.p + c.woff, u->g32.p + c.boff);
    if (r != CFD_OK || ain == A_IN) return r;
    r = cfd_nn_conv3x3(gy, B, s.h, s.w, c.cout, u->wf.p + u->wf_off[q], nullptr, c.cin, 0, u->gbuf[ain].p, nullptr);
    if (r != CFD_OK || !relu_in) return r;
    const long n = numel(u, ain, B);                   // 32..256 channels: a multiple of 8
    hipLaunchKernelGGL(k_relu_bwd<8>, dim3(eb(n / 8)), dim3(256), 0, 0, u->gbuf[ain].p, u->act[ain].p, (int)n);
    return launched();
}

// y = bf16(x) over n values: 16 B vectors, then the tail
int f32_to_bf16(const float *x, long n, bf16 *y)
{
    const long n8 = n & ~7L;
    if (n8) hipLaunchKernelGGL(k_f32_to_bf16<8>, dim3(eb(n8 / 8)), dim3(256), 0, 0, x, 0L, n8, y);
    if (n > n8) hipLaunchKernelGGL(k_f32_to_bf16<1>, dim3(1), dim3(256), 0, 0, x, n8, n, y);
    return launched();
}


// Every activation gradient is BF16 and multiplied by the ReLU derivative of its activation
// before it is used: after the input-gradient convolution by a mask pass, or, where two
// branches meet, inside the fused upsample/pool/skip kernels (one FP32 sum, one rounding).
int backward(cfd_unet *u, int B)
{
    int r = fill0(u->g32.p, (long)u->nparam);
    // head (no activation on its output): the gradient of D1 leaves masked
    if (r == CFD_OK) r = conv_bwd(u, 14, B, A_D1, u->gout.p, true);
    // two convolutions of a level: aout = conv qb (amid), amid = conv qa (ain)
    auto seq = [&](int qb, int qa, int aout, int amid, int ain, bool relu_in) {
        int rr = conv_bwd(u, qb, B, amid, u->gbuf[aout].p, true);
        if (rr == CFD_OK) rr = conv_bwd(u, qa, B, ain, u->gbuf[amid].p, relu_in);
        return rr;
    };
    // the low (upsampled) part of a concat gradient -> gradient of the level below
    auto up_bwd = [&](int cat, int low) {
        const Shape sl = act_shape(u, low), sc = act_shape(u, cat);
        const long n = numel(u, low, B);               // channel counts 64..256: multiples of 8
        hipLaunchKernelGGL(k_up_relu_bwd<8>, dim3(eb(n / 8)), dim3(256), 0, 0, u->gbuf[cat].p, B, sl.h, sl.w, sl.c,
                           sc.c, 0, u->act[low].p, u->gbuf[low].p);
        return launched();
    };
    // skip part of the concat gradient plus the pooled gradient -> gradient of the encoder level
    auto pool_skip_bwd = [&](int a, int p, int cat, int low) {
        const Shape s = act_shape(u, a), sc = act_shape(u, cat), sl = act_shape(u, low);
        const long n = numel(u, p, B);                 // 32..128 channels
        hipLaunchKernelGGL(k_pool_skip_bwd<8>, dim3(eb(n / 8)), dim3(256), 0, 0, u->act[a].p, u->act[p].p,
                           u->gbuf[p].p, u->gbuf[cat].p, sc.c, sl.c, B, s.h, s.w, s.c, u->gbuf[a].p);
        return launched();
    };
    if (r == CFD_OK) r = seq(13, 12, A_D1, A_D1A, A_C1, false);
    if (r == CFD_OK) r = up_bwd(A_C1, A_D2);
    if (r == CFD_OK) r = seq(11, 10, A_D2, A_D2A, A_C2, false);
    if (r == CFD_OK) r = up_bwd(A_C2, A_D3);
    if (r == CFD_OK) r = seq(9, 8, A_D3, A_D3A, A_C3, false);
    if (r == CFD_OK) r = up_bwd(A_C3, A_B);
    if (r == CFD_OK) r = seq(7, 6, A_B, A_BA, A_P3, false);
    if (r == CFD_OK) r = pool_skip_bwd(A_E3, A_P3, A_C3, A_B);
    if (r == CFD_OK) r = seq(5, 4, A_E3, A_E3A, A_P2, false);
    if (r == CFD_OK) r = pool_skip_bwd(A_E2, A_P2, A_C2, A_D3);
    if (r == CFD_OK) r = seq(3, 2, A_E2, A_E2A, A_P1, false);
    if (r == CFD_OK) r = pool_skip_bwd(A_E1, A_P1, A_C1, A_D2);
    if (r == CFD_OK) r = seq(1, 0, A_E1, A_E1A, A_IN, false);
    return r;
}

// BF16 copies of the FP32 masters (unless Adam already wrote them) and the flipped weights
int refresh_weights(cfd_unet *u, bool bf16_copy = true)
{
    int r = bf16_copy ? f32_to_bf16(u->w32.p, (long)u->nparam, u->wb.p) : CFD_OK;
    FlipTable t;
    for (int q = 0; q < 15; ++q) {
        const Conv &c = u->convs[q];
        t.foff[q] = (long)u->wf_off[q]; t.woff[q] = (long)c.woff; t.cin[q] = c.cin; t.cout[q] = c.cout;
    }
    t.foff[15] = t.foff[14] + (long)u->convs[14].cin * 9 * u->convs[14].cout;
    if (r == CFD_OK) {
        hipLaunchKernelGGL(k_flip_all, dim3(eb(t.foff[15])), dim3(256), 0, 0, (const float *)u->w32.p, t, u->wf.p);
        r = launched();
    }
    return r;
}

int load_inputs(cfd_unet *u, const float *x, int B)
{
    const long n = (long)B * u->H * u->W * 5;
    if (hipMemcpy(u->xin32.p, x, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return CFD_ERR_HIP;
    return f32_to_bf16(u->xin32.p, n, u->act[A_IN].p);
}

uint64_t sm64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int cfd_unet_create(int32_t height, int32_t width, int32_t batch, uint64_t seed, cfd_unet **out)
{
    if (!out) return cfd__fail(CFD_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (height < 8 || width < 8 || height % 8 || width % 8 || batch < 1)
        return cfd__fail(CFD_ERR_ARG, "U-Net grid must be a multiple of 8 (>= 8) and batch >= 1, got %dx%d, %d",
                         height, width, batch);
    // the widest activation (96 channels at full resolution) must index in 32 bits
    if ((long)batch * height * width * 96 >= 2147483648L)
        return cfd__fail(CFD_ERR_ARG, "U-Net batch x grid too large: %d x %dx%d", batch, height, width);
    cfd_unet *u = new cfd_unet();
    u->H = height; u->W = width; u->Bmax = batch;
    build_convs(u);
    int r = CFD_OK;
    for (int k = 0; k < 4 && r == CFD_OK; ++k)
        if (hipEventCreate(&u->ev[k]) != hipSuccess) r = CFD_ERR_HIP;
    r = u->w32.alloc(u->nparam); if (r == CFD_OK) r = u->g32.alloc(u->nparam);
    if (r == CFD_OK) r = u->m32.alloc(u->nparam); if (r == CFD_OK) r = u->v32.alloc(u->nparam);
    if (r == CFD_OK) r = u->wb.alloc(u->nparam);
    size_t nf = 0;
    for (int q = 0; q < 15; ++q) nf += (size_t)u->convs[q].cin * 9 * u->convs[q].cout;
    if (r == CFD_OK) r = u->wf.alloc(nf);
    for (int a = 0; a < A_NACT && r == CFD_OK; ++a) {
        const long n = numel(u, a, batch);
        r = u->act[a].alloc(n);
        if (r == CFD_OK) r = u->gbuf[a].alloc(n);
    }
    const long no = (long)batch * height * width * 4;
    if (r == CFD_OK) r = u->out32.alloc(no);
    if (r == CFD_OK) r = u->target.alloc(no);
    if (r == CFD_OK) r = u->gout.alloc(no);
    if (r == CFD_OK) r = u->loss.alloc(1);
    if (r == CFD_OK) r = u->xin32.alloc((long)batch * height * width * 5);
    if (r != CFD_OK) { cfd_unet_destroy(u); return cfd__fail(CFD_ERR_HIP, "U-Net allocation failed"); }
    // He-uniform initial weights (seeded SplitMix64), zero biases
    std::vector<float> h(u->nparam, 0.f);
    uint64_t s = seed;
    for (int q = 0; q < 15; ++q) {
        const Conv &c = u->convs[q];
        const float lim = sqrtf(6.f / (9.f * c.cin)) * (c.relu ? 1.f : 0.5f);
        for (size_t i = 0; i < (size_t)c.cout * 9 * c.cin; ++i) {
            const double v = (double)(sm64(&s) >> 11) * (1.0 / 9007199254740992.0);
            h[c.woff + i] = (float)((2.0 * v - 1.0) * lim);
        }
    }
    r = cfd_unet_set_params(u, h.data());
    if (r != CFD_OK) { cfd_unet_destroy(u); return r; }
    *out = u;
    return CFD_OK;
}

int cfd_unet_destroy(cfd_unet *u)
{
    if (!u) return CFD_OK;
    (void)hipDeviceSynchronize();
    u->w32.release(); u->g32.release(); u->m32.release(); u->v32.release(); u->wb.release(); u->wf.release();
    for (int a = 0; a < 40; ++a) { u->act[a].release(); u->gbuf[a].release(); }
    u->out32.release(); u->target.release(); u->gout.release(); u->loss.release(); u->xin32.release(); u->xin.release();
    for (int k = 0; k < 4; ++k)
        if (u->ev[k]) (void)hipEventDestroy(u->ev[k]);
    delete u;
    return CFD_OK;
}

int cfd_unet_num_params(const cfd_unet *u, int64_t *n)
{
    if (!u || !n) return cfd__fail(CFD_ERR_ARG, "NULL argument");
    *n = (int64_t)u->nparam;
    return CFD_OK;
}

int cfd_unet_get_params(cfd_unet *u, float *host)
{
    if (!u || !host) return cfd__fail(CFD_ERR_ARG, "NULL argument");
    if (hipMemcpy(host, u->w32.p, u->nparam * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return cfd__fail(CFD_ERR_HIP, "parameter download failed");
    return CFD_OK;
}

int cfd_unet_set_params(cfd_unet *u, const float *host)
{
    if (!u || !host) return cfd__fail(CFD_ERR_ARG, "NULL argument");
    if (hipMemcpy(u->w32.p, host, u->nparam * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return cfd__fail(CFD_ERR_HIP, "parameter upload failed");
    int r = fill0(u->m32.p, (long)u->nparam);
    if (r == CFD_OK) r = fill0(u->v32.p, (long)u->nparam);
    if (r == CFD_OK) r = refresh_weights(u);
    u->step = 0;
    if (r == CFD_OK && hipDeviceSynchronize() != hipSuccess) r = CFD_ERR_HIP;
    return r == CFD_OK ? CFD_OK : cfd__fail(r, "U-Net parameter setup failed");
}

int cfd_unet_predict(cfd_unet *u, const float *x, int32_t B, float *y)
{
    if (!u || !x || !y) return cfd__fail(CFD_ERR_ARG, "NULL argument");
    if (B < 1 || B > u->Bmax) return cfd__fail(CFD_ERR_ARG, "batch %d outside 1..%d", B, u->Bmax);
    int r = load_inputs(u, x, B);
    if (r == CFD_OK) r = forward(u, B);
    if (r != CFD_OK) return cfd__fail(r, "U-Net forward failed");
    if (hipMemcpy(y, u->out32.p, (size_t)B * u->H * u->W * 4 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return cfd__fail(CFD_ERR_HIP, "U-Net output download failed");
    return CFD_OK;
}

// Loss and parameter gradients for one batch (gradients left on the device; copied to
// grad_host when it is not NULL).
int cfd_unet_loss_grad(cfd_unet *u, const float *x, const float *t, int32_t B, double *loss, float *grad_host)
{
    if (!u || !x || !t) return cfd__fail(CFD_ERR_ARG, "NULL argument");
    if (B < 1 || B > u->Bmax) return cfd__fail(CFD_ERR_ARG, "batch %d outside 1..%d", B, u->Bmax);
    const long no = (long)B * u->H * u->W * 4;
    int r = load_inputs(u, x, B);
    if (r == CFD_OK && hipMemcpy(u->target.p, t, no * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) r = CFD_ERR_HIP;
    if (r == CFD_OK && u->ev[0]) (void)hipEventRecord(u->ev[0], 0);
    if (r == CFD_OK) r = forward(u, B);
    if (r == CFD_OK && hipMemset(u->loss.p, 0, sizeof(double)) != hipSuccess) r = CFD_ERR_HIP;
    if (r == CFD_OK) {
        hipLaunchKernelGGL(k_mse, dim3(eb(no) < 1024 ? eb(no) : 1024), dim3(256), 0, 0, u->out32.p, u->target.p, no,
                           u->gout.p, u->loss.p);
        r = launched();
    }
    if (r == CFD_OK && u->ev[1]) (void)hipEventRecord(u->ev[1], 0);
    if (r == CFD_OK) r = backward(u, B);
    if (r == CFD_OK && u->ev[2]) (void)hipEventRecord(u->ev[2], 0);
    double l = 0.0;
    if (r == CFD_OK && hipMemcpy(&l, u->loss.p, sizeof l, hipMemcpyDeviceToHost) != hipSuccess) r = CFD_ERR_HIP;
    if (r == CFD_OK && grad_host &&
        hipMemcpy(grad_host, u->g32.p, u->nparam * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        r = CFD_ERR_HIP;
    if (r != CFD_OK) return cfd__fail(r, "U-Net loss/gradient failed");
    if (loss) *loss = l;
    return CFD_OK;
}

int cfd_unet_train_step(cfd_unet *u, const float *x, const float *t, int32_t B, float lr, double *loss)
{
    int r = cfd_unet_loss_grad(u, x, t, B, loss, nullptr);
    if (r != CFD_OK) return r;
    u->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)u->step), c2 = 1.f - powf(b2, (float)u->step);
    hipLaunchKernelGGL(k_adam, dim3(eb((long)u->nparam)), dim3(256), 0, 0, u->w32.p, u->g32.p, u->m32.p, u->v32.p,
                       (long)u->nparam, lr, b1, b2, 1e-8f, c1, c2, u->wb.p);
    r = launched();
    if (r == CFD_OK) r = refresh_weights(u, false);
    if (r == CFD_OK && u->ev[3]) (void)hipEventRecord(u->ev[3], 0);
    return r == CFD_OK ? CFD_OK : cfd__fail(r, "U-Net update failed");
}

int cfd_unet_last_timing(cfd_unet *u, double ms[3])
{
    if (!u || !ms) return cfd__fail(CFD_ERR_ARG, "NULL argument");
    if (!u->ev[3] || u->step == 0) return cfd__fail(CFD_ERR_STATE, "no training step timed yet");
    float a = 0.f, b = 0.f, c = 0.f;
    if (hipEventSynchronize(u->ev[3]) != hipSuccess || hipEventElapsedTime(&a, u->ev[0], u->ev[1]) != hipSuccess ||
        hipEventElapsedTime(&b, u->ev[1], u->ev[2]) != hipSuccess || hipEventElapsedTime(&c, u->ev[2], u->ev[3]) != hipSuccess)
        return cfd__fail(CFD_ERR_HIP, "U-Net timing events");
    ms[0] = a; ms[1] = b; ms[2] = c;
    return CFD_OK;
}

}  // extern "C"
