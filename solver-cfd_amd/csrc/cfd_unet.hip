// cfd_unet.hip — the learned warm start: a 4-level convolutional U-Net (channels 32, 64,
// 128, 256; 3x3 kernels) trained and run with hand-written HIP kernels.
//
// Inputs per cell (5 channels): porous-block mask L, porosity eps, and the case's log Re,
// log Da and Ri broadcast over the grid. Outputs (4 channels): u, v, p, Z (T) on the p-cell
// index grid, from which the solver's staggered initial fields are taken.
//
// Numerics: activations and their gradients are stored in BF16 (NHWC), every convolution
// is an implicit GEMM on the CDNA matrix cores (v_mfma_f32_32x32x8bf16_1k: BF16 operands,
// FP32 accumulation), master weights, weight gradients, biases and the Adam state are FP32.
//
//   forward   y[m][n]  = sum_k A[m][k] W[n][k] + b[n]     m = pixel, n = cout,
//                                                          k = (dy, dx, cin), A = im2col(x)
//   input grad dx = the same kernel on dy with the flipped, transposed weights
//             Wf[ci][(dy', dx', co)] = W[co][(2 - dy', 2 - dx')][ci]
//   weight grad dW[n][k] = sum_m dy[m][n] A[m][k]           (MFMA over m, split over the grid)
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
#include <vector>

#include "../../include/cfd.h"

extern "C" int cfd__fail(int code, const char *fmt, ...);

namespace {

typedef uint16_t bf16;
typedef short bf16x4 __attribute__((ext_vector_type(4)));
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

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;   // LDS row pitch avoids bank conflicts

// ---------------------------------------------------------------------------------------
// implicit-GEMM 3x3 convolution, stride 1, zero padding 1, NHWC
// ---------------------------------------------------------------------------------------
// x: [B][H][W][Cin] bf16; w: [Cout][9*Cin] bf16 (k = (dy*3+dx)*Cin + ci); y: [B][H][W][Cout]
template <bool RELU>
__global__ __launch_bounds__(256) void k_conv3x3(const bf16 *__restrict__ x, int B, int H, int W, int Cin,
                                                 const bf16 *__restrict__ w, const float *__restrict__ bias,
                                                 int Cout, bf16 *__restrict__ y, float *__restrict__ y32)
{
    __shared__ bf16 As[BM][LDK];
    __shared__ bf16 Bs[BN][LDK];
    const int M = B * H * W, K = 9 * Cin;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    // this thread's A row and k slice
    const int ar = t >> 2, ak = (t & 3) * 8;
    const int am = m0 + ar;
    int ab = 0, ah = 0, aw = 0;
    if (am < M) { ab = am / (H * W); const int rem = am - ab * H * W; ah = rem / W; aw = rem - ah * W; }
    f32x16 acc;
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
    for (int k0 = 0; k0 < K; k0 += BK) {
        // A tile: im2col gather with zero padding
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int k = k0 + ak + q;
            bf16 v = 0;
            if (am < M && k < K) {
                const int r = k / Cin, ci = k - r * Cin;
                const int hh = ah + r / 3 - 1, ww = aw + r % 3 - 1;
                if (hh >= 0 && hh < H && ww >= 0 && ww < W) v = x[((size_t)(ab * H + hh) * W + ww) * Cin + ci];
            }
            As[ar][ak + q] = v;
        }
        // B tile: Bs[n][k] = w[n][k]
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int n = n0 + ar, k = k0 + ak + q;
            Bs[ar][ak + q] = (n < Cout && k < K) ? w[(size_t)n * K + k] : (bf16)0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 8) {
            const int kk = ks + 4 * (lane >> 5);
            bf16x4 a, b;
            const bf16 *pa = &As[wm + (lane & 31)][kk];
            const bf16 *pb = &Bs[wn + (lane & 31)][kk];
            a[0] = (short)pa[0]; a[1] = (short)pa[1]; a[2] = (short)pa[2]; a[3] = (short)pa[3];
            b[0] = (short)pb[0]; b[1] = (short)pb[1]; b[2] = (short)pb[2]; b[3] = (short)pb[3];
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout of the 32x32 MFMA: lane l, element r -> row 8 (r/4) + 4 (l/32) + r%4, col l%32
    const int n = n0 + wn + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (m < M && n < Cout) {
            float v = acc[r] + (bias ? bias[n] : 0.f);
            if (RELU) v = v > 0.f ? v : 0.f;
            if (y) y[(size_t)m * Cout + n] = f2bf(v);
            if (y32) y32[(size_t)m * Cout + n] = v;
        }
    }
}

// dW[n][k] += sum_m dy[m][n] * im2col(x)[m][k] over this block's slice of m
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const bf16 *__restrict__ x, const bf16 *__restrict__ dy,
                                                       int B, int H, int W, int Cin, int Cout, int m_per_block,
                                                       float *__restrict__ dw)
{
    __shared__ bf16 Ys[BN][LDK];     // [n][m]
    __shared__ bf16 Xs[BM][LDK];     // [k][m]
    const int M = B * H * W, K = 9 * Cin;
    const int n0 = blockIdx.x * BN, kt0 = blockIdx.y * BM;
    const int mb = blockIdx.z * m_per_block, me = min(M, mb + m_per_block);
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;        // wave tile: rows n, cols k
    const int lr = t >> 2, lm = (t & 3) * 8;                   // loader: row (n or k), m slice
    f32x16 acc;
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
    for (int m0 = mb; m0 < me; m0 += BK) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int m = m0 + lm + q, n = n0 + lr;
            Ys[lr][lm + q] = (m < me && n < Cout) ? dy[(size_t)m * Cout + n] : (bf16)0;
        }
        const int k = kt0 + lr;
        int r = 0, ci = 0;
        if (k < K) { r = k / Cin; ci = k - r * Cin; }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int m = m0 + lm + q;
            bf16 v = 0;
            if (m < me && k < K) {
                const int b = m / (H * W), rem = m - b * H * W, h = rem / W, wq = rem - h * W;
                const int hh = h + r / 3 - 1, ww = wq + r % 3 - 1;
                if (hh >= 0 && hh < H && ww >= 0 && ww < W) v = x[((size_t)(b * H + hh) * W + ww) * Cin + ci];
            }
            Xs[lr][lm + q] = v;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 8) {
            const int kk = ks + 4 * (lane >> 5);
            bf16x4 a, b;
            const bf16 *pa = &Ys[wr + (lane & 31)][kk];
            const bf16 *pb = &Xs[wc + (lane & 31)][kk];
            a[0] = (short)pa[0]; a[1] = (short)pa[1]; a[2] = (short)pa[2]; a[3] = (short)pa[3];
            b[0] = (short)pb[0]; b[1] = (short)pb[1]; b[2] = (short)pb[2]; b[3] = (short)pb[3];
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int kc = kt0 + wc + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int n = n0 + wr + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (n < Cout && kc < K && acc[r] != 0.f) atomicAdd(&dw[(size_t)n * K + kc], acc[r]);
    }
}

// ---------------------------------------------------------------------------------------
// elementwise and data-movement kernels
// ---------------------------------------------------------------------------------------
__global__ void k_relu_bwd(bf16 *g, const bf16 *y, long n)          // g *= (y > 0)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        if (!(bf2f(y[i]) > 0.f)) g[i] = 0;
}

__global__ void k_bias_grad(const bf16 *g, long M, int C, float *db)
{
    // one block per channel chunk of 64 lanes x 4 row groups
    const int c = blockIdx.x * 64 + (threadIdx.x & 63);
    const int grp = threadIdx.x >> 6;
    float s = 0.f;
    if (c < C)
        for (long m = (long)blockIdx.y * 4 + grp; m < M; m += 4L * gridDim.y) s += bf2f(g[m * C + c]);
    if (c < C && s != 0.f) atomicAdd(&db[c], s);
}

__global__ void k_maxpool2(const bf16 *x, int B, int H, int W, int C, bf16 *y)
{
    const long n = (long)B * (H / 2) * (W / 2) * C;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long p = i / C;
        const int wo = (int)(p % (W / 2)); p /= (W / 2);
        const int ho = (int)(p % (H / 2));
        const int b = (int)(p / (H / 2));
        const size_t base = ((size_t)(b * H + 2 * ho) * W + 2 * wo) * C + c;
        float m = bf2f(x[base]);
        m = fmaxf(m, bf2f(x[base + C]));
        m = fmaxf(m, bf2f(x[base + (size_t)W * C]));
        m = fmaxf(m, bf2f(x[base + (size_t)W * C + C]));
        y[i] = f2bf(m);
    }
}

// gradient to the first of the four inputs that holds the maximum (row-major order)
__global__ void k_maxpool2_bwd(const bf16 *x, const bf16 *y, const bf16 *gy, int B, int H, int W, int C, float *gx)
{
    const long n = (long)B * (H / 2) * (W / 2) * C;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long p = i / C;
        const int wo = (int)(p % (W / 2)); p /= (W / 2);
        const int ho = (int)(p % (H / 2));
        const int b = (int)(p / (H / 2));
        const size_t base = ((size_t)(b * H + 2 * ho) * W + 2 * wo) * C + c;
        const size_t off[4] = {0, (size_t)C, (size_t)W * C, (size_t)W * C + C};
        const float mv = bf2f(y[i]);           // compared as values: -0 equals +0
        for (int q = 0; q < 4; ++q)
            if (bf2f(x[base + off[q]]) == mv) { gx[base + off[q]] += bf2f(gy[i]); break; }
    }
}

__global__ void k_upsample2(const bf16 *x, int B, int H, int W, int C, bf16 *y, int Cy, int cofs)
{
    // x: [B][H][W][C] -> y[B][2H][2W][Cy] channels cofs..cofs+C-1
    const long n = (long)B * (2 * H) * (2 * W) * C;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long p = i / C;
        const int wo = (int)(p % (2 * W)); p /= (2 * W);
        const int ho = (int)(p % (2 * H));
        const int b = (int)(p / (2 * H));
        y[((size_t)(b * 2 * H + ho) * 2 * W + wo) * Cy + cofs + c] = x[((size_t)(b * H + ho / 2) * W + wo / 2) * C + c];
    }
}

// gx[B][H][W][C] = sum over the 2x2 children of gy channels cofs.. (gy: [B][2H][2W][Cy])
__global__ void k_upsample2_bwd(const bf16 *gy, int B, int H, int W, int C, int Cy, int cofs, float *gx)
{
    const long n = (long)B * H * W * C;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long p = i / C;
        const int w = (int)(p % W); p /= W;
        const int h = (int)(p % H);
        const int b = (int)(p / H);
        float s = 0.f;
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                s += bf2f(gy[((size_t)(b * 2 * H + 2 * h + dy) * 2 * W + 2 * w + dx) * Cy + cofs + c]);
        gx[i] += s;
    }
}

// copy channels: y[..][cofs + c] = x[..][c], for npix pixels
__global__ void k_chan_copy(const bf16 *x, long npix, int C, bf16 *y, int Cy, int cofs)
{
    const long n = npix * C;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[(i / C) * Cy + cofs + (i % C)] = x[i];
}

// gx[..][c] += g[..][cofs + c]
__global__ void k_chan_add(const bf16 *g, long npix, int C, int Cg, int cofs, float *gx)
{
    const long n = npix * C;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        gx[i] += bf2f(g[(i / C) * Cg + cofs + (i % C)]);
}

__global__ void k_f32_to_bf16(const float *x, long n, bf16 *y)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) y[i] = f2bf(x[i]);
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

__global__ void k_adam(float *w, const float *g, float *m, float *v, long n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi; v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

inline int eb(long n) { const long b = (n + 255) / 256; return (int)(b < 8192 ? (b > 0 ? b : 1) : 8192); }

}  // namespace

// ---------------------------------------------------------------------------------------
// op-level C-ABI (device pointers, NULL stream) — the network is built from these
// ---------------------------------------------------------------------------------------
extern "C" {

int cfd_nn_conv3x3(const uint16_t *x, int32_t B, int32_t H, int32_t W, int32_t Cin, const uint16_t *w,
                   const float *bias, int32_t Cout, int32_t relu, uint16_t *y, float *y32)
{
    const int M = B * H * W;
    const dim3 g(cdiv(M, BM), cdiv(Cout, BN));
    if (relu) hipLaunchKernelGGL(k_conv3x3<true>, g, dim3(256), 0, 0, x, B, H, W, Cin, w, bias, Cout, y, y32);
    else hipLaunchKernelGGL(k_conv3x3<false>, g, dim3(256), 0, 0, x, B, H, W, Cin, w, bias, Cout, y, y32);
    return launched();
}

int cfd_nn_conv3x3_wgrad(const uint16_t *x, const uint16_t *dy, int32_t B, int32_t H, int32_t W, int32_t Cin,
                         int32_t Cout, float *dw)
{
    const int M = B * H * W, K = 9 * Cin;
    // split m so that the grid has a few hundred blocks
    const int tiles = cdiv(Cout, BN) * cdiv(K, BM);
    int splits = cdiv(512, tiles);
    const int maxs = cdiv(M, BK);
    if (splits > maxs) splits = maxs;
    if (splits < 1) splits = 1;
    const int mpb = cdiv(cdiv(M, splits), BK) * BK;
    splits = cdiv(M, mpb);
    hipLaunchKernelGGL(k_conv3x3_wgrad, dim3(cdiv(Cout, BN), cdiv(K, BM), splits), dim3(256), 0, 0, x, dy, B, H, W,
                       Cin, Cout, mpb, dw);
    return launched();
}

int cfd_nn_flip_weights(const float *w, int32_t Cin, int32_t Cout, uint16_t *wf)
{
    hipLaunchKernelGGL(k_flip_weights, dim3(eb((long)Cin * 9 * Cout)), dim3(256), 0, 0, w, Cin, Cout, wf);
    return launched();
}

int cfd_nn_maxpool2(const uint16_t *x, int32_t B, int32_t H, int32_t W, int32_t C, uint16_t *y)
{
    hipLaunchKernelGGL(k_maxpool2, dim3(eb((long)B * H * W * C / 4)), dim3(256), 0, 0, x, B, H, W, C, y);
    return launched();
}

int cfd_nn_maxpool2_bwd(const uint16_t *x, const uint16_t *y, const uint16_t *gy, int32_t B, int32_t H, int32_t W,
                        int32_t C, float *gx)
{
    hipLaunchKernelGGL(k_maxpool2_bwd, dim3(eb((long)B * H * W * C / 4)), dim3(256), 0, 0, x, y, gy, B, H, W, C, gx);
    return launched();
}

int cfd_nn_upsample2(const uint16_t *x, int32_t B, int32_t H, int32_t W, int32_t C, uint16_t *y, int32_t Cy,
                     int32_t cofs)
{
    hipLaunchKernelGGL(k_upsample2, dim3(eb((long)B * H * W * C * 4)), dim3(256), 0, 0, x, B, H, W, C, y, Cy, cofs);
    return launched();
}

int cfd_nn_upsample2_bwd(const uint16_t *gy, int32_t B, int32_t H, int32_t W, int32_t C, int32_t Cy, int32_t cofs,
                         float *gx)
{
    hipLaunchKernelGGL(k_upsample2_bwd, dim3(eb((long)B * H * W * C)), dim3(256), 0, 0, gy, B, H, W, C, Cy, cofs, gx);
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
    void release() { if (p) hipFree(p); p = nullptr; }
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
    Dev<float> gacc[40];                  // FP32 gradient accumulators where two paths meet
    Dev<bf16> gbuf[40];
    Dev<float> out32, target;
    Dev<bf16> gout;
    Dev<double> loss;
    Dev<float> xin32;
    Dev<bf16> xin;
    long step = 0;
};

namespace {

const int CH[4] = {32, 64, 128, 256};

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
        const Shape sl = act_shape(u, low), ss = act_shape(u, skip), sc = act_shape(u, cat);
        int rr = cfd_nn_upsample2(u->act[low].p, B, sl.h, sl.w, sl.c, u->act[cat].p, sc.c, 0);
        if (rr == CFD_OK) {
            hipLaunchKernelGGL(k_chan_copy, dim3(eb((long)B * ss.h * ss.w * ss.c)), dim3(256), 0, 0, u->act[skip].p,
                               (long)B * ss.h * ss.w, ss.c, u->act[cat].p, sc.c, sl.c);
            rr = launched();
        }
        return rr;
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
int conv_bwd(cfd_unet *u, int q, int B, int ain, const bf16 *gy)
{
    const Conv &c = u->convs[q];
    const Shape s = act_shape(u, ain);
    const long M = (long)B * s.h * s.w;
    int r = cfd_nn_conv3x3_wgrad(u->act[ain].p, gy, B, s.h, s.w, c.cin, c.cout, u->g32.p + c.woff);
    if (r != CFD_OK) return r;
    const int by = (int)((M + 255) / 256 < 64 ? (M + 255) / 256 : 64);
    hipLaunchKernelGGL(k_bias_grad, dim3(cdiv(c.cout, 64), by > 0 ? by : 1), dim3(256), 0, 0, gy, M, c.cout,
                       u->g32.p + c.boff);
    r = launched();
    if (r != CFD_OK || ain == A_IN) return r;
    return cfd_nn_conv3x3(gy, B, s.h, s.w, c.cout, u->wf.p + u->wf_off[q], nullptr, c.cin, 0, u->gbuf[ain].p, nullptr);
}

int relu_mask(cfd_unet *u, int a, long n)
{
    hipLaunchKernelGGL(k_relu_bwd, dim3(eb(n)), dim3(256), 0, 0, u->gbuf[a].p, u->act[a].p, n);
    return launched();
}

int to_bf(cfd_unet *u, int a, long n)
{
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(eb(n)), dim3(256), 0, 0, u->gacc[a].p, n, u->gbuf[a].p);
    return launched();
}

long numel(cfd_unet *u, int a, int B)
{
    const Shape s = act_shape(u, a);
    return (long)B * s.h * s.w * s.c;
}

int backward(cfd_unet *u, int B)
{
    int r = fill0(u->g32.p, (long)u->nparam);
    // head: gy = gout (no activation)
    if (r == CFD_OK) r = conv_bwd(u, 14, B, A_D1, u->gout.p);
    // dec1: D1 (relu) <- conv13 <- D1A (relu) <- conv12 <- C1
    auto seq = [&](int qb, int qa, int aout, int amid, int ain) {
        int rr = relu_mask(u, aout, numel(u, aout, B));
        if (rr == CFD_OK) rr = conv_bwd(u, qb, B, amid, u->gbuf[aout].p);
        if (rr == CFD_OK) rr = relu_mask(u, amid, numel(u, amid, B));
        if (rr == CFD_OK) rr = conv_bwd(u, qa, B, ain, u->gbuf[amid].p);
        return rr;
    };
    // split the concat gradient: low part -> upsample bwd into gacc[low]; skip part -> gacc[skip]
    auto split = [&](int cat, int low, int skip) {
        const Shape sl = act_shape(u, low), ss = act_shape(u, skip), sc = act_shape(u, cat);
        int rr = cfd_nn_upsample2_bwd(u->gbuf[cat].p, B, sl.h, sl.w, sl.c, sc.c, 0, u->gacc[low].p);
        if (rr == CFD_OK) {
            hipLaunchKernelGGL(k_chan_add, dim3(eb((long)B * ss.h * ss.w * ss.c)), dim3(256), 0, 0, u->gbuf[cat].p,
                               (long)B * ss.h * ss.w, ss.c, sc.c, sl.c, u->gacc[skip].p);
            rr = launched();
        }
        return rr;
    };
    auto pool_bwd = [&](int a, int p) {
        const Shape s = act_shape(u, a);
        return cfd_nn_maxpool2_bwd(u->act[a].p, u->act[p].p, u->gbuf[p].p, B, s.h, s.w, s.c, u->gacc[a].p);
    };
    const int zero[] = {A_E1, A_E2, A_E3, A_B, A_D3, A_D2};
    for (int q = 0; q < 6 && r == CFD_OK; ++q) r = fill0(u->gacc[zero[q]].p, numel(u, zero[q], B));
    if (r == CFD_OK) r = seq(13, 12, A_D1, A_D1A, A_C1);
    if (r == CFD_OK) r = split(A_C1, A_D2, A_E1);
    if (r == CFD_OK) r = to_bf(u, A_D2, numel(u, A_D2, B));
    if (r == CFD_OK) r = seq(11, 10, A_D2, A_D2A, A_C2);
    if (r == CFD_OK) r = split(A_C2, A_D3, A_E2);
    if (r == CFD_OK) r = to_bf(u, A_D3, numel(u, A_D3, B));
    if (r == CFD_OK) r = seq(9, 8, A_D3, A_D3A, A_C3);
    if (r == CFD_OK) r = split(A_C3, A_B, A_E3);
    if (r == CFD_OK) r = to_bf(u, A_B, numel(u, A_B, B));
    if (r == CFD_OK) r = seq(7, 6, A_B, A_BA, A_P3);
    if (r == CFD_OK) r = pool_bwd(A_E3, A_P3);
    if (r == CFD_OK) r = to_bf(u, A_E3, numel(u, A_E3, B));
    if (r == CFD_OK) r = seq(5, 4, A_E3, A_E3A, A_P2);
    if (r == CFD_OK) r = pool_bwd(A_E2, A_P2);
    if (r == CFD_OK) r = to_bf(u, A_E2, numel(u, A_E2, B));
    if (r == CFD_OK) r = seq(3, 2, A_E2, A_E2A, A_P1);
    if (r == CFD_OK) r = pool_bwd(A_E1, A_P1);
    if (r == CFD_OK) r = to_bf(u, A_E1, numel(u, A_E1, B));
    if (r == CFD_OK) r = seq(1, 0, A_E1, A_E1A, A_IN);
    return r;
}

int refresh_weights(cfd_unet *u)
{
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(eb((long)u->nparam)), dim3(256), 0, 0, u->w32.p, (long)u->nparam, u->wb.p);
    int r = launched();
    for (int q = 0; q < 15 && r == CFD_OK; ++q) {
        const Conv &c = u->convs[q];
        r = cfd_nn_flip_weights(u->w32.p + c.woff, c.cin, c.cout, u->wf.p + u->wf_off[q]);
    }
    return r;
}

int load_inputs(cfd_unet *u, const float *x, int B)
{
    const long n = (long)B * u->H * u->W * 5;
    if (hipMemcpy(u->xin32.p, x, n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) return CFD_ERR_HIP;
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(eb(n)), dim3(256), 0, 0, u->xin32.p, n, u->act[A_IN].p);
    return launched();
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
    cfd_unet *u = new cfd_unet();
    u->H = height; u->W = width; u->Bmax = batch;
    build_convs(u);
    int r = CFD_OK;
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
        if (r == CFD_OK) r = u->gacc[a].alloc(n);
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
    hipDeviceSynchronize();
    u->w32.release(); u->g32.release(); u->m32.release(); u->v32.release(); u->wb.release(); u->wf.release();
    for (int a = 0; a < 40; ++a) { u->act[a].release(); u->gacc[a].release(); u->gbuf[a].release(); }
    u->out32.release(); u->target.release(); u->gout.release(); u->loss.release(); u->xin32.release(); u->xin.release();
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
    if (r == CFD_OK) r = forward(u, B);
    if (r == CFD_OK && hipMemset(u->loss.p, 0, sizeof(double)) != hipSuccess) r = CFD_ERR_HIP;
    if (r == CFD_OK) {
        hipLaunchKernelGGL(k_mse, dim3(eb(no) < 1024 ? eb(no) : 1024), dim3(256), 0, 0, u->out32.p, u->target.p, no,
                           u->gout.p, u->loss.p);
        r = launched();
    }
    if (r == CFD_OK) r = backward(u, B);
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
                       (long)u->nparam, lr, b1, b2, 1e-8f, c1, c2);
    r = launched();
    if (r == CFD_OK) r = refresh_weights(u);
    return r == CFD_OK ? CFD_OK : cfd__fail(r, "U-Net update failed");
}

}  // extern "C"
