// dem_noalloc.hip - the kinetic-energy observable for callers that cannot allocate.
//
// dem_kinetic_energy stages per-particle energies in a buffer the context allocated at
// creation and sums them on the host in a vector it allocates per call.  Code under HIP
// graph capture, tight serving loops and harnesses that forbid allocation need the same
// observable with no allocation at all: dem_kinetic_energy_noalloc keeps its partial
// results in a file-scope device array (`scratch`, one copy per GPU, held until the
// process ends) and reduces them on the device.
//
// The array's address on each device is taken once, when the first context is created
// on that device (hipGetSymbolAddress with the device current), and kept per device;
// host code never passes the array's name to a kernel.  Every call on one GPU shares
// that GPU's array: calls that overlap on one GPU (two threads, or two contexts on
// different streams) race and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/dem.h"
#include "dem_engine.h"

namespace psim {

// Largest scene (particles, padding slots included) whose energies fit: 4,194,304.
constexpr int64_t KE_STATIC_MAX = DEM_KE_NOALLOC_MAX;
constexpr int KE_CHUNK = 1024;                                  // ids summed by one block
constexpr int64_t KE_PARTIALS = KE_STATIC_MAX / KE_CHUNK;       // one partial per block
static_assert(KE_STATIC_MAX % 256 == 0, "whole blocks of slots fit");

// [0, KE_STATIC_MAX): energy of particle id i;  [KE_STATIC_MAX, +KE_PARTIALS): block
// partials;  [KE_STATIC_MAX + KE_PARTIALS]: the total.  33.6 MB per GPU.
static __device__ __align__(256) double scratch[KE_STATIC_MAX + KE_PARTIALS + 1];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};
}  // namespace

// Resolves the device address of `scratch` on `device`, which must be the current
// device; idempotent per device.
int init_static_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) {
        set_error("device index out of range for the static scratch table");
        return DEM_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return DEM_OK;
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return DEM_EDEVICE;
    }
    g_scratch_ptr[device] = p;
    return DEM_OK;
}

double* static_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

// Sum of e[lo, hi) in a fixed order: each thread adds four consecutive entries, then a
// binary tree over the block's threads.  `out` gets the block's sum.
__device__ __forceinline__ void block_sum(const double* e, int64_t lo, int64_t hi, double* out) {
    __shared__ double s[256];
    const int t = (int)threadIdx.x;
    double v = 0.0;
    for (int k = 0; k < 4; k++) {
        const int64_t i = lo + 4 * (int64_t)t + k;
        if (i < hi) v += e[i];
    }
    s[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) s[t] = s[t] + s[t + w];
        __syncthreads();
    }
    if (t == 0) *out = s[0];
}

// Block b sums ids [b * KE_CHUNK, min(n, (b + 1) * KE_CHUNK)) into partial b.
__global__ void __launch_bounds__(256) k_ke_partials(double* buf, int n) {
    const int64_t lo = (int64_t)blockIdx.x * KE_CHUNK;
    const int64_t hi = lo + KE_CHUNK < n ? lo + KE_CHUNK : n;
    block_sum(buf, lo, hi, buf + KE_STATIC_MAX + blockIdx.x);
}

// One block: the partials, 256 strided runs, then the tree.
__global__ void __launch_bounds__(256) k_ke_total(double* buf, int nparts) {
    __shared__ double s[256];
    const int t = (int)threadIdx.x;
    const double* p = buf + KE_STATIC_MAX;
    double v = 0.0;
    for (int i = t; i < nparts; i += 256) v += p[i];
    s[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) s[t] = s[t] + s[t + w];
        __syncthreads();
    }
    if (t == 0) buf[KE_STATIC_MAX + KE_PARTIALS] = s[0];
}

// Reduces the energies of ids [0, n) held in buf[0, n) and copies the total to the host.
hipError_t launch_ke_reduce(double* buf, int n, double* host_out, hipStream_t st) {
    const int nparts = n > 0 ? (n + KE_CHUNK - 1) / KE_CHUNK : 0;
    if (nparts > 0) hipLaunchKernelGGL(k_ke_partials, dim3(nparts), dim3(256), 0, st, buf, n);
    hipLaunchKernelGGL(k_ke_total, dim3(1), dim3(256), 0, st, buf, nparts);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(host_out, buf + KE_STATIC_MAX + KE_PARTIALS, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

}  // namespace psim
