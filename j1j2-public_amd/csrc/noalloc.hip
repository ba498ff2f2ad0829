// noalloc.hip -- device memory and fill kernels of j1j2_run_cycles_noalloc,
// the entry point that allocates nothing.
//
// A cycle's sweep writes each lattice's integer observables (OBS_N int64) and
// the cycle's swap step reads them: per-cycle partial results.  The no-alloc
// entry point keeps them in one static device array per GPU instead of the
// engine's buffer, and clears the phase accumulators with fill kernels rather
// than hipMemsetAsync, so a caller may also capture its calls in a graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "kernels.h"

namespace j1j2 {

// NOALLOC_MAX_LATTICES x OBS_N int64 = 1 MiB on every GPU the library runs
// on, held until the process ends.  Every entry a cycle reads is written by
// that cycle's sweep first, so the array never needs clearing.  In host code
// `scratch` is only a stand-in for the symbol: kernels get the address that
// noalloc_setup looked up on the device.
static __device__ __align__(256) int64_t scratch[NOALLOC_MAX_LATTICES * OBS_N];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
int64_t *g_scratch[MAX_DEVICES];   // per device: the device address of `scratch`

__global__ __launch_bounds__(256) void k_fill_u64(uint64_t *p, int64_t n, uint64_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_u8(uint8_t *p, int64_t n, uint8_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = v;
}

int fill_blocks(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 2048); }

}  // namespace

hipError_t noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;      // the caller makes `device` current
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<int64_t *>(p);
    return hipSuccess;
}

int64_t *noalloc_scratch(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_scratch[device];
}

hipError_t launch_fill_u64(void *p, int64_t n, uint64_t v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(fill_blocks(n)), dim3(256), 0, s, static_cast<uint64_t *>(p), n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u8(void *p, int64_t n, uint8_t v, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u8, dim3(fill_blocks(n)), dim3(256), 0, s, static_cast<uint8_t *>(p), n, v);
    return hipGetLastError();
}

}  // namespace j1j2
