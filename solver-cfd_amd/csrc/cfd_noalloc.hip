// cfd_noalloc.hip -- max_reduce with no device allocation, for callers that cannot allocate
// (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
//
// The legacy max_reduce (max_reduce.cu:14-41: max |a| over the first n_i*n_j linear
// entries) folds into a residual slot of the library's default context and returns a host
// double. This variant needs no context: a first pass writes one partial maximum per
// workgroup into `scratch`, a file-scope device array of NOALLOC_PARTIALS entries (8 KiB,
// held by every GPU the library runs on until the process ends), and a second pass folds
// those partials into the caller's device double, on the caller's stream. Every partial the
// second pass reads was written by the first pass of the same call, so nothing has to be
// cleared between calls. Calls that overlap on one GPU (two streams, two host threads) share
// that GPU's scratch and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/cfd_legacy.h"

extern "C" int cfd__fail(int code, const char *fmt, ...);   // cfd_host.c: thread-local last error

namespace {
constexpr int NOALLOC_PARTIALS = 1024;     // partial maxima: one per workgroup of the first pass
constexpr int NOALLOC_THREADS = 256;
constexpr int NOALLOC_MAX_DEVICES = 64;
// the largest input: 2^36 doubles (512 GiB, more than one GPU's memory)
constexpr long long NOALLOC_MAX_ENTRIES = 1LL << 36;
}  // namespace

static __device__ __align__(256) unsigned long long scratch[NOALLOC_PARTIALS];

// scratch's device address on each GPU (the array's name is a host-side stand-in in host
// code); filled once per device by hipGetSymbolAddress with that device current
static unsigned long long *g_scratch[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_lock;

// max of non-negative doubles through their bit patterns (they order as unsigned integers;
// a NaN's pattern lies above every finite value and infinity, so a NaN entry wins)
__device__ inline unsigned long long block_max_u64(unsigned long long m)
{
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = __shfl_xor(m, o);
        m = other > m ? other : m;
    }
    __shared__ unsigned long long wave_max[NOALLOC_THREADS / 64];
    if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < NOALLOC_THREADS / 64; ++w) m = wave_max[w] > m ? wave_max[w] : m;
    return m;   // valid in thread 0
}

__global__ __launch_bounds__(NOALLOC_THREADS) void k_max_partials(const double *a, long long n,
                                                                  unsigned long long *part)
{
    unsigned long long m = 0;
    for (long long t = (long long)blockIdx.x * NOALLOC_THREADS + threadIdx.x; t < n;
         t += (long long)gridDim.x * NOALLOC_THREADS) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(a[t]));
        m = b > m ? b : m;
    }
    m = block_max_u64(m);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ __launch_bounds__(NOALLOC_THREADS) void k_max_final(const unsigned long long *part, int nparts,
                                                               double *out)
{
    unsigned long long m = 0;
    for (int t = threadIdx.x; t < nparts; t += NOALLOC_THREADS) m = part[t] > m ? part[t] : m;
    m = block_max_u64(m);
    if (threadIdx.x == 0) out[0] = __longlong_as_double((long long)m);
}

static int scratch_for_current_device(unsigned long long **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return cfd__fail(CFD_ERR_HIP, "max_reduce_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return cfd__fail(CFD_ERR_UNSUPPORTED, "max_reduce_noalloc: device %d beyond the %d this build tracks", dev,
                         NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return cfd__fail(CFD_ERR_HIP, "max_reduce_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                             hipGetErrorString(e));
        g_scratch[dev] = (unsigned long long *)p;
    }
    *out = g_scratch[dev];
    return CFD_OK;
}

extern "C" int max_reduce_noalloc(const double *dev_a, int n_i, int n_j, double *dev_out, void *stream_)
{
    const hipStream_t stream = (hipStream_t)stream_;
    if (!dev_out) return cfd__fail(CFD_ERR_ARG, "max_reduce_noalloc: dev_out is NULL");
    if (n_i < 0 || n_j < 0) return cfd__fail(CFD_ERR_ARG, "max_reduce_noalloc: negative size %d x %d", n_i, n_j);
    const long long n = (long long)n_i * n_j;
    if (n > NOALLOC_MAX_ENTRIES)
        return cfd__fail(CFD_ERR_ARG, "max_reduce_noalloc: %lld entries exceed the maximum %lld", n,
                         NOALLOC_MAX_ENTRIES);
    if (n > 0 && !dev_a) return cfd__fail(CFD_ERR_ARG, "max_reduce_noalloc: dev_a is NULL");
    unsigned long long *part = nullptr;
    const int r = scratch_for_current_device(&part);
    if (r != CFD_OK) return r;
    // about eight entries per lane before the grid strides, at most one partial per scratch slot
    long long nb = (n + NOALLOC_THREADS * 8LL - 1) / (NOALLOC_THREADS * 8LL);
    if (nb > NOALLOC_PARTIALS) nb = NOALLOC_PARTIALS;
    if (nb > 0) hipLaunchKernelGGL(k_max_partials, dim3((unsigned)nb), dim3(NOALLOC_THREADS), 0, stream, dev_a, n, part);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(NOALLOC_THREADS), 0, stream, (const unsigned long long *)part,
                       (int)nb, dev_out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cfd__fail(CFD_ERR_HIP, "max_reduce_noalloc: launch: %s", hipGetErrorString(e));
    return CFD_OK;
}
