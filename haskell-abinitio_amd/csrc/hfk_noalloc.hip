// calcIntegrals into caller-owned device memory without any allocation
// (hfk_eri_flat_noalloc, include/hfk.h).  For callers that cannot allocate
// device memory: code under HIP graph capture, serving loops.  The one
// scratch buffer of the flat export -- the table base[i][j] = rank of
// (i,j,i,j) in calcIntegrals order, n x n 64-bit integers -- lives in a
// file-scope device array of HFK_NOALLOC_MAX_NBF^2 entries (512 KB), which
// every GPU the library runs on holds until the process ends.  Its device
// address is looked up once per device, at context creation, with that
// device current; host code uses only that address, never the array's name.
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "hfk.h"
#include "hfk_host.hpp"
#include "hfk_internal.hpp"

static __device__ __align__(256) long long scratch[HFK_NOALLOC_MAX_NBF * HFK_NOALLOC_MAX_NBF];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
long long* g_scratch[MAX_DEVICES] = {nullptr};

// base[i][j] (i <= j) = rank of (i,j,i,j) among the canonical quartets in
// calcIntegrals order (HartreeFock.hs:153-181): with P the row-major rank of
// the pair (i,j) and M = n(n+1)/2 pairs, the pairs before it start M - q
// quartets each, so base = sum_{q<P} (M - q) = P M - P(P-1)/2.
__global__ void k_flat_base(int n, long long* base)
{
    const long long m = (long long)n * (n + 1) / 2;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < (long long)n * n;
         t += (long long)gridDim.x * blockDim.x) {
        const long long i = t / n, j = t - i * n;
        if (j < i) {
            base[t] = 0;
            continue;
        }
        const long long P = i * n - i * (i - 1) / 2 + (j - i);
        base[t] = P * m - P * (P - 1) / 2;
    }
}

}  // namespace

// Look up this device's copy of the scratch array (once per device; the
// caller has made the device current).
int hfk_noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return hfk_set_error(HFK_EUNSUP, "device %d out of range", device);
    std::lock_guard<std::mutex> g(g_mu);
    if (g_scratch[device]) return HFK_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return hfk_set_error(HFK_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch[device] = (long long*)p;
    return HFK_OK;
}

extern "C" int hfk_eri_flat_noalloc(hfk_ctx* c, double* out_dev, size_t out_len)
{
    if (!c || !out_dev) return hfk_set_error(HFK_EINVAL, "hfk_eri_flat_noalloc: null argument");
    std::lock_guard<std::mutex> g(c->mu);
    const long long n = c->n, m = n * (n + 1) / 2, len = m * (m + 1) / 2;
    if (n > HFK_NOALLOC_MAX_NBF)
        return hfk_set_error(HFK_EINVAL, "hfk_eri_flat_noalloc: %lld basis functions, at most %d", n, HFK_NOALLOC_MAX_NBF);
    if ((long long)out_len != len)
        return hfk_set_error(HFK_EINVAL, "hfk_eri_flat_noalloc: out_len %zu, need %lld", out_len, len);
    long long* base = nullptr;
    {
        std::lock_guard<std::mutex> gg(g_mu);
        if (c->device >= 0 && c->device < MAX_DEVICES) base = g_scratch[c->device];
    }
    if (!base) return hfk_set_error(HFK_EHIP, "hfk_eri_flat_noalloc: scratch of device %d not set up", c->device);
    hipLaunchKernelGGL(k_flat_base, dim3((unsigned)std::min<long long>(1024, (n * n + 255) / 256)), dim3(256), 0,
                       c->stream, (int)n, base);
    hfk_launch_fill(out_dev, 0.0, len, c->stream);
    QuartetArgs a = hfk_ctx_args(c);
    a.flat = out_dev;
    a.flat_base = base;
    HFK_TRY(hfk_ctx_run_quartets(c, a, HFK_OP_FLAT));
    HFK_HIP(hipStreamSynchronize(c->stream));
    return HFK_OK;
}
