// cfd_ops.hip — per-operator kernels of the drop-in entry points (RESU ... bcC, max_reduce).
//
// Each launcher reproduces one reference routine over exactly the reference's index
// region (OpenMp_Multicore/equations.f90, boundary.f90; c_cuda/*.cu for the names), with
// the literal arithmetic of cfd_device.h. Thread x walks j (the contiguous axis), so loads
// and stores are coalesced; the reference's CUDA kernels map x to the strided axis.
#include "cfd_device.h"
#include "cfd_internal.h"

using namespace cfd;

namespace {
constexpr int BX = 64, BY = 4;

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

inline Geo mk(const cfd_devgeo *g)
{
    Geo G;
    G.x = g->x; G.y = g->y; G.xm = g->xm; G.ym = g->ym;
    G.aun = g->aun; G.aus = g->aus; G.aue = g->aue; G.auw = g->auw;
    G.avn = g->avn; G.avs = g->avs; G.ave = g->ave; G.avw = g->avw;
    G.eps = g->eps; G.lp = g->lp; G.imax = g->imax; G.jmax = g->jmax;
    return G;
}

inline int launched() { return hipGetLastError() == hipSuccess ? CFD_OK : CFD_ERR_HIP; }

// RESU: i = 2..imax, j = 2..jmax-1. QUICK on i = 3..imax-1, j = 3..jmax-2; the rest is the
// reference's upwind strips (equations.f90:109-113), which do not overlap the interior.
__global__ void k_resu(Geo g, cfd_params P, int mode, const double *um, const double *vm, const double *p,
                       const double *T, double *ru)
{
    const int j = blockIdx.x * BX + threadIdx.x + 2;
    const int i = blockIdx.y * BY + threadIdx.y + 2;
    if (i > g.imax || j > g.jmax - 1) return;
    const bool inner = i >= 3 && i <= g.imax - 1 && j >= 3 && j <= g.jmax - 2;
    if (mode == 1 && inner) return;
    const bool quick = P.scheme == CFD_SCHEME_QUICK && inner;
    ru[iu(g, i, j)] = resu_lit(g, P, um, vm, p, T, i, j, quick);
}

// RESV: i = 2..imax-1, j = 2..jmax (Fortran strips rows 2 and jmax, columns 2 and imax-1 over
// j = 2..jmax, equations.f90:413-417). CUDA: strips rows 2 and jmax-1, columns over j <= jmax-1
// (upwind_V_pair.cu:22,114-117), so row jmax is not written.
__global__ void k_resv(Geo g, cfd_params P, int mode, const double *um, const double *vm, const double *p,
                       const double *T, double *rv)
{
    const int j = blockIdx.x * BX + threadIdx.x + 2;
    const int i = blockIdx.y * BY + threadIdx.y + 2;
    const bool cuda = P.compat == CFD_COMPAT_CUDA;
    const int jend = cuda ? g.jmax - 1 : g.jmax;
    if (i > g.imax - 1 || j > jend) return;
    const int jq = cuda ? g.jmax - 2 : g.jmax - 1;
    const bool inner = i >= 3 && i <= g.imax - 2 && j >= 3 && j <= jq;
    if (mode == 1 && inner) return;
    const bool quick = P.scheme == CFD_SCHEME_QUICK && inner;
    rv[iv(g, i, j)] = resv_lit(g, P, um, vm, p, T, i, j, quick);
}

// kT (may be null): physical-level temperature of the k(T) diffusivity, k(T)/Pe per cell
__global__ void k_ressc(Geo g, double k, int species, const double *um, const double *vm,
                        const double *z, const double *kT, cfd_params P, double *rz)
{
    const int j = blockIdx.x * BX + threadIdx.x + 2;
    const int i = blockIdx.y * BY + threadIdx.y + 2;
    if (i > g.imax - 1 || j > g.jmax - 1) return;
    const double kk = kT ? cond_lit(P, kT[iu(g, i, j)]) : k;
    rz[iu(g, i, j)] = ressc_lit(g, kk, species != 0, um, vm, z, i, j);
}

// region 0: u (i=3..imax-1, j=2..jmax-1, stride jmax+1); 1: v (i=2..imax-1, j=3..jmax-1,
// stride jmax+2); 2: scalar (i=2..imax-1, j=2..jmax-1). equations.f90:262-293 etc.
__global__ void k_stage(Geo g, int region, int stage, double dt, double dtau, const double *phys,
                        const double *tau, const double *r, const double *prev, double *out, double *res)
{
    const int i0 = region == 0 ? 3 : 2, j0 = region == 1 ? 3 : 2;
    const int j = blockIdx.x * BX + threadIdx.x + j0;
    const int i = blockIdx.y * BY + threadIdx.y + i0;
    if (i > g.imax - 1 || j > g.jmax - 1) return;
    const size_t k = region == 1 ? iv(g, i, j) : iu(g, i, j);
    const double rs = ((phys[k] - tau[k]) + r[k] * dt) * dtau;
    if (res) res[k] = rs;
    double o;
    if (stage == 1) o = tau[k] + rs;
    else if (stage == 2) o = 0.75 * tau[k] + 0.25 * (prev[k] + rs);
    else o = 1.0 / 3.0 * tau[k] + 2.0 / 3.0 * (prev[k] + rs);
    out[k] = o;
}

// equations.f90:620-658; c_cuda/solve_P.cu:18 starts stage 2 at j = 3.
__global__ void k_pstage(Geo g, int stage, int j0, double dtau, double beta, const double *p,
                         const double *um_n, const double *vm_n, double *rp, double *pi, double *pn,
                         double *res)
{
    const int j = blockIdx.x * BX + threadIdx.x + j0;
    const int i = blockIdx.y * BY + threadIdx.y + 2;
    if (i > g.imax - 1 || j > g.jmax - 1) return;
    const size_t k = iu(g, i, j);
    if (stage == 1) {
        const double dudx = um_n[iu(g, i + 1, j)] * g.aue[j] - um_n[iu(g, i, j)] * g.auw[j];
        const double dvdy = vm_n[iv(g, i, j + 1)] * g.avn[i] - vm_n[iv(g, i, j)] * g.avs[i];
        const double r = -(dudx + dvdy);
        rp[k] = r;
        pi[k] = p[k] + dtau * r * beta;
    } else if (stage == 2) {
        pi[k] = 0.75 * p[k] + 0.25 * (pi[k] + dtau * rp[k] * beta);
    } else {
        const double rs = dtau * rp[k] * beta;
        res[k] = rs;
        pn[k] = 1.0 / 3.0 * p[k] + 2.0 / 3.0 * (pi[k] + rs);
    }
}

// u boundary cells: rows j=1, jmax for i=1..imax+1, then columns i=1,2,imax,imax+1 for
// j=2..jmax-1. Final values from cfd_device.h (reads interior cells only).
__global__ void k_bc_u(Geo g, cfd_params P, double *u)
{
    const int im = g.imax, jm = g.jmax;
    const int nrow = im + 1, ncol = jm - 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int i, j;
    if (t < 2 * nrow) { i = 1 + t % nrow; j = t < nrow ? 1 : jm; }
    else if (t < 2 * nrow + 4 * ncol) {
        const int q = t - 2 * nrow, c = q / ncol;
        j = 2 + q % ncol;
        i = c == 0 ? 1 : (c == 1 ? 2 : (c == 2 ? im : im + 1));
    } else return;
    u[iu(g, i, j)] = P.bc == CFD_BC_CHANNEL ? bc_u_chan(g, P, u, i, j) : bc_u_ref(g, u, i, j);
}

// v boundary cells: rows j=1,2,jmax,jmax+1 for i=1..imax, then columns i=1, imax for
// j=3..jmax-1. Reference set: the corner (imax, jmax+1) reads the pre-BC cells (imax, jmax-2)
// and (imax, jmax-1), so one thread reads all three and writes them after (boundary.f90:19-22,44-46).
__global__ void k_bc_v(Geo g, cfd_params P, double *v)
{
    const int im = g.imax, jm = g.jmax;
    const int ncol = jm - 3;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool chan = P.bc == CFD_BC_CHANNEL;
    int i, j;
    if (t < 4 * im) {
        const int r = t / im;
        i = 1 + t % im;
        j = r == 0 ? 1 : (r == 1 ? 2 : (r == 2 ? jm : jm + 1));
        if (!chan && i == im && j == jm + 1) {
            const double a = v[iv(g, im, jm - 2)], b = v[iv(g, im, jm - 1)];
            const double corner = 2.0 * b - (2.0 * a - b);
            v[iv(g, im, jm - 2)] = v[iv(g, im - 1, jm - 2)];
            v[iv(g, im, jm - 1)] = v[iv(g, im - 1, jm - 1)];
            v[iv(g, im, jm + 1)] = corner;
            return;
        }
    } else if (t < 4 * im + 2 * ncol) {
        const int q = t - 4 * im;
        i = q < ncol ? 1 : im;
        j = 3 + q % ncol;
        if (!chan && i == im && j >= jm - 2) return;   // written by the corner thread
    } else return;
    v[iv(g, i, j)] = chan ? bc_v_chan(g, v, i, j) : bc_v_ref(g, P, v, i, j);
}

__global__ void k_bc_s(Geo g, cfd_params P, int kind, double *a)
{
    const int im = g.imax, jm = g.jmax;
    const int ncol = jm - 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int i, j;
    if (t < 2 * im) { i = 1 + t % im; j = t < im ? 1 : jm; }
    else if (t < 2 * im + 2 * ncol) {
        const int q = t - 2 * im;
        i = q < ncol ? 1 : im;
        j = 2 + q % ncol;
    } else return;
    a[iu(g, i, j)] = P.bc == CFD_BC_CHANNEL ? bc_s_chan(g, P, kind, a, i, j) : bc_s_ref(g, P, kind, a, i, j);
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long m)
{
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(m, o, 64);
        m = other > m ? other : m;
    }
    return m;
}

__global__ void k_maxabs(const double *a, int ld, int i0, int j0, int ni, int nj, unsigned long long *out)
{
    const long n = (long)ni * nj;
    unsigned long long m = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = i0 + (int)(t / nj), j = j0 + (int)(t % nj);
        const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(a[(size_t)i * ld + j]));
        m = b > m ? b : m;
    }
    m = wave_max_u64(m);
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_maxabs_linear(const double *a, long n, unsigned long long *out)
{
    unsigned long long m = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(a[t]));
        m = b > m ? b : m;
    }
    m = wave_max_u64(m);
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0ull;
}

__global__ void k_zero_slots(unsigned long long *p, int n, double *l2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0ull;
    if (l2 && t < 3) l2[t] = 0.0;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ void k_sumsq(const double *a, int ld, int i0, int j0, int ni, int nj, double *out)
{
    double s = 0.0;
    const long n = (long)ni * nj;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const double v = a[(size_t)(i0 + t / nj) * ld + (j0 + t % nj)];
        s += v * v;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0 && s != 0.0) atomicAdd(out, s);
}

__global__ void k_copy(double *dst, const double *src, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        dst[t] = src[t];
}

// Stream copy for the bandwidth probe: 16-byte elements, each 256-lane block moves 4 KB per
// lane group (four coalesced 16 B loads in flight per lane, then four stores).
constexpr int SC_PER = 4;
__global__ __launch_bounds__(256) void k_stream_copy(uint4 *__restrict__ dst, const uint4 *__restrict__ src, long n)
{
    const long base = (long)blockIdx.x * (256 * SC_PER) + threadIdx.x;
    uint4 v[SC_PER];
#pragma unroll
    for (int k = 0; k < SC_PER; ++k) {
        const long t = base + (long)k * 256;
        if (t < n) v[k] = src[t];
    }
#pragma unroll
    for (int k = 0; k < SC_PER; ++k) {
        const long t = base + (long)k * 256;
        if (t < n) dst[t] = v[k];
    }
}

__global__ void k_fill(double *dst, double v, long n)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        dst[t] = v;
}

// dst_c(i, j) = src_f(i, j), i = 1..n, j = 1..m; C row stride ld, Fortran column-major.
__global__ void k_from_fortran(double *dst, const double *src, int n, int m, int ld)
{
    const int j = blockIdx.x * BX + threadIdx.x + 1;
    const int i = blockIdx.y * BY + threadIdx.y + 1;
    if (i > n || j > m) return;
    dst[(size_t)i * ld + j] = src[(size_t)(i - 1) + (size_t)(j - 1) * n];
}

__global__ void k_to_fortran(double *dst, const double *src, int n, int m, int ld)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    if (i > n || j > m) return;
    dst[(size_t)(i - 1) + (size_t)(j - 1) * n] = src[(size_t)i * ld + j];
}

int reduce_blocks(long n)
{
    long b = (n + 255) / 256;
    return (int)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}
}  // namespace

extern "C" {

int cfdk_resu(const cfd_devgeo *g, const cfd_params *P, int mode, const double *um, const double *vm,
              const double *p, const double *T, double *ru, hipStream_t s)
{
    dim3 grid(cdiv(g->jmax - 2, BX), cdiv(g->imax - 1, BY));
    hipLaunchKernelGGL(k_resu, grid, dim3(BX, BY), 0, s, mk(g), *P, mode, um, vm, p, T, ru);
    return launched();
}

int cfdk_resv(const cfd_devgeo *g, const cfd_params *P, int mode, const double *um, const double *vm,
              const double *p, const double *T, double *rv, hipStream_t s)
{
    dim3 grid(cdiv(g->jmax - 1, BX), cdiv(g->imax - 2, BY));
    hipLaunchKernelGGL(k_resv, grid, dim3(BX, BY), 0, s, mk(g), *P, mode, um, vm, p, T, rv);
    return launched();
}

int cfdk_ressc(const cfd_devgeo *g, const cfd_params *P, int species, const double *um, const double *vm,
               const double *z, const double *kT, double *rz, hipStream_t s)
{
    const double k = species ? 1.0 / P->re / P->sc : 1.0 / P->pe;
    const double *kt = (!species && P->k_model == CFD_K_POLY4) ? kT : nullptr;
    dim3 grid(cdiv(g->jmax - 2, BX), cdiv(g->imax - 2, BY));
    hipLaunchKernelGGL(k_ressc, grid, dim3(BX, BY), 0, s, mk(g), k, species, um, vm, z, kt, *P, rz);
    return launched();
}

int cfdk_stage(const cfd_devgeo *g, const cfd_params *P, int region, int stage,
               const double *phys, const double *tau, const double *r, const double *prev,
               double *out, double *res, hipStream_t s)
{
    const int ni = g->imax - 1 - (region == 0 ? 3 : 2) + 1;
    const int nj = g->jmax - 1 - (region == 1 ? 3 : 2) + 1;
    dim3 grid(cdiv(nj, BX), cdiv(ni, BY));
    hipLaunchKernelGGL(k_stage, grid, dim3(BX, BY), 0, s, mk(g), region, stage, P->dt, P->dtau,
                       phys, tau, r, prev, out, res);
    return launched();
}

int cfdk_pstage(const cfd_devgeo *g, const cfd_params *P, int stage, const double *p,
                const double *um_n, const double *vm_n, double *rp, double *pi, double *pn,
                double *res, hipStream_t s)
{
    const int j0 = (stage == 2 && P->compat == CFD_COMPAT_CUDA) ? 3 : 2;
    dim3 grid(cdiv(g->jmax - j0, BX), cdiv(g->imax - 2, BY));
    hipLaunchKernelGGL(k_pstage, grid, dim3(BX, BY), 0, s, mk(g), stage, j0, P->dtau, P->beta,
                       p, um_n, vm_n, rp, pi, pn, res);
    return launched();
}

int cfdk_bc_uv(const cfd_devgeo *g, const cfd_params *P, double *u, double *v, hipStream_t s)
{
    const Geo G = mk(g);
    if (u) {
        const int n = 2 * (g->imax + 1) + 4 * (g->jmax - 2);
        hipLaunchKernelGGL(k_bc_u, dim3(cdiv(n, 256)), dim3(256), 0, s, G, *P, u);
    }
    if (v) {
        const int n = 4 * g->imax + 2 * (g->jmax - 3);
        hipLaunchKernelGGL(k_bc_v, dim3(cdiv(n, 256)), dim3(256), 0, s, G, *P, v);
    }
    return launched();
}

int cfdk_bc_s(const cfd_devgeo *g, const cfd_params *P, int kind, double *a, hipStream_t s)
{
    const int n = 2 * g->imax + 2 * (g->jmax - 2);
    hipLaunchKernelGGL(k_bc_s, dim3(cdiv(n, 256)), dim3(256), 0, s, mk(g), *P, kind, a);
    return launched();
}

int cfdk_zero_u64(uint64_t *out, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(cdiv(n, 64)), dim3(64), 0, s, (unsigned long long *)out, n);
    return launched();
}

int cfdk_maxabs(const double *a, int ld, int i0, int i1, int j0, int j1, uint64_t *out, hipStream_t s)
{
    const int ni = i1 - i0 + 1, nj = j1 - j0 + 1;
    if (ni <= 0 || nj <= 0) return CFD_OK;
    hipLaunchKernelGGL(k_maxabs, dim3(reduce_blocks((long)ni * nj)), dim3(256), 0, s, a, ld, i0, j0, ni, nj,
                       (unsigned long long *)out);
    return launched();
}

int cfdk_zero_slots(uint64_t *slot, int n, double *l2, hipStream_t s)
{
    const int m = n > 3 ? n : 3;
    hipLaunchKernelGGL(k_zero_slots, dim3(cdiv(m, 256)), dim3(256), 0, s, (unsigned long long *)slot, n, l2);
    return launched();
}

int cfdk_sumsq(const double *a, int ld, int i0, int i1, int j0, int j1, double *out, hipStream_t s)
{
    const int ni = i1 - i0 + 1, nj = j1 - j0 + 1;
    if (ni <= 0 || nj <= 0) return CFD_OK;
    hipLaunchKernelGGL(k_sumsq, dim3(reduce_blocks((long)ni * nj)), dim3(256), 0, s, a, ld, i0, j0, ni, nj, out);
    return launched();
}

int cfdk_maxabs_linear(const double *a, long n, uint64_t *out, hipStream_t s)
{
    if (n <= 0) return CFD_OK;
    hipLaunchKernelGGL(k_maxabs_linear, dim3(reduce_blocks(n)), dim3(256), 0, s, a, n, (unsigned long long *)out);
    return launched();
}

int cfdk_copy(double *dst, const double *src, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy, dim3(reduce_blocks(n)), dim3(256), 0, s, dst, src, n);
    return launched();
}

/* HBM bandwidth probe: a plain copy of `bytes` (rounded down to 16 B) between two fresh
 * device buffers, `reps` timed launches after one warm-up; *gbs = 2 * bytes * reps / time. */
int cfd_hbm_copy_bandwidth(int64_t bytes, int32_t reps, double *gbs)
{
    if (bytes < 16 || reps < 1 || !gbs) return CFD_ERR_ARG;
    const long n = (long)(bytes / 16);
    uint4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int r = CFD_ERR_HIP;
    if (hipMalloc((void **)&a, (size_t)n * 16) == hipSuccess && hipMalloc((void **)&b, (size_t)n * 16) == hipSuccess &&
        hipMemset(a, 0, (size_t)n * 16) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess) {
        const dim3 grid(cdiv(n, 256 * SC_PER));
        hipLaunchKernelGGL(k_stream_copy, grid, dim3(256), 0, 0, b, a, n);
        hipEventRecord(e0, 0);
        for (int q = 0; q < reps; ++q) hipLaunchKernelGGL(k_stream_copy, grid, dim3(256), 0, 0, b, a, n);
        hipEventRecord(e1, 0);
        float ms = 0.f;
        if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.f) {
            *gbs = 2.0 * (double)n * 16.0 * reps / ((double)ms * 1e-3) * 1e-9;
            r = CFD_OK;
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (a) hipFree(a);
    if (b) hipFree(b);
    return r;
}

int cfdk_fill(double *dst, double v, long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill, dim3(reduce_blocks(n)), dim3(256), 0, s, dst, v, n);
    return launched();
}

int cfdk_from_fortran(double *dst_c, const double *src_f, int n, int m, int ld, hipStream_t s)
{
    dim3 grid(cdiv(m, BX), cdiv(n, BY));
    hipLaunchKernelGGL(k_from_fortran, grid, dim3(BX, BY), 0, s, dst_c, src_f, n, m, ld);
    return launched();
}

int cfdk_to_fortran(double *dst_f, const double *src_c, int n, int m, int ld, hipStream_t s)
{
    dim3 grid(cdiv(n, BX), cdiv(m, BY));
    hipLaunchKernelGGL(k_to_fortran, grid, dim3(BX, BY), 0, s, dst_f, src_c, n, m, ld);
    return launched();
}

}  // extern "C"
