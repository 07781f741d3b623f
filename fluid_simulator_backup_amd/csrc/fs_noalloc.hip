// fs_noalloc.hip -- fs_cfl_dt_noalloc: the CFL time step of device-resident
// MAC velocities with no allocation, for callers that cannot allocate
// (code under HIP graph capture, tight serving loops).
//
// Each call reduces max|u|, max|v|, max|w| into per-workgroup partials held
// in a file-scope device array (`scratch`, FS_NOALLOC_SCRATCH doubles on
// every GPU the library runs on, kept until the process ends), then one
// workgroup reduces the partials and writes dt to device memory.  The host
// takes the array's device address once per device (hipGetSymbolAddress)
// and passes that address to the kernels.  Every partial a call reads is
// written earlier in the same call, so nothing needs clearing.  Calls that
// overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <mutex>

#include "../../include/fs_abi.h"
#include "fs_cfl.h"

#define FS_NOALLOC_BLOCKS 2048               // partials per velocity component
#define FS_NOALLOC_CHUNK (1L << 17)          // faces one workgroup reduces at most
#define FS_NOALLOC_SCRATCH (3 * FS_NOALLOC_BLOCKS)
#define FS_NOALLOC_MAX_FACES ((long)FS_NOALLOC_BLOCKS * FS_NOALLOC_CHUNK)  // per component
#define NT 256

static __device__ __align__(256) double scratch[FS_NOALLOC_SCRATCH];

namespace {

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_mu;

// the device address of `scratch` on the current device
int scratch_address(double** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FS_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return FS_ERR_HIP;
        g_scratch[dev] = (double*)ptr;
    }
    *out = g_scratch[dev];
    return FS_OK;
}

// max of |a| with NaN kept (a NaN anywhere makes the result NaN)
__device__ __forceinline__ double nan_max(double a, double b) { return (a != a || a > b) ? a : b; }

__device__ double block_max(double v, double* lds) {
    for (int o = 32; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o));
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    v = lds[0];
    for (int w = 1; w < NT / 64; w++) v = nan_max(v, lds[w]);
    __syncthreads();
    return v;
}

struct Comp {
    const double* a;
    long n;      // faces
    long chunk;  // faces per workgroup
    int nb;      // workgroups
};

// blockIdx.y: component; workgroup b reduces faces [b*chunk, (b+1)*chunk)
__global__ void __launch_bounds__(NT) k_partial_abs_max(Comp c0, Comp c1, Comp c2, double* part) {
    __shared__ double lds[NT / 64];
    const Comp c = blockIdx.y == 0 ? c0 : blockIdx.y == 1 ? c1 : c2;
    if ((int)blockIdx.x >= c.nb) return;  // a whole workgroup leaves together
    const long lo = (long)blockIdx.x * c.chunk, hi = min(lo + c.chunk, c.n);
    double v = 0;
    for (long i = lo + threadIdx.x; i < hi; i += NT) v = nan_max(v, fabs(c.a[i]));
    v = block_max(v, lds);
    if (threadIdx.x == 0) part[blockIdx.y * FS_NOALLOC_BLOCKS + blockIdx.x] = v;
}

__global__ void __launch_bounds__(NT) k_cfl_finish(const double* part, int nb0, int nb1, int nb2, FsCflTerms t,
                                                   double cfl, double* dt) {
    __shared__ double lds[NT / 64];
    double m[3];
    const int nb[3] = {nb0, nb1, nb2};
    for (int c = 0; c < 3; c++) {
        double v = 0;
        for (int b = threadIdx.x; b < nb[c]; b += NT) v = nan_max(v, part[c * FS_NOALLOC_BLOCKS + b]);
        m[c] = block_max(v, lds);
    }
    if (threadIdx.x == 0) {
        const bool finite = isfinite(m[0]) && isfinite(m[1]) && isfinite(m[2]);
        *dt = finite ? cfl * ((double)1 / fs_cfl_rate(t, m[0], m[1], m[2])) : __builtin_nan("");
    }
}

Comp make_comp(const double* a, long n) {
    Comp c;
    c.a = a;
    c.n = n;
    // at least 4096 faces per workgroup, at most FS_NOALLOC_BLOCKS workgroups
    long chunk = std::max((n + FS_NOALLOC_BLOCKS - 1) / FS_NOALLOC_BLOCKS, 4096L);
    chunk = (chunk + NT - 1) / NT * NT;
    c.chunk = chunk;
    c.nb = (int)((n + chunk - 1) / chunk);
    return c;
}

}  // namespace

extern "C" int fs_cfl_dt_noalloc(const fs_params* params, const fs_grid* grid, const double* u, const double* v,
                                 const double* w, double cfl, int gravity_norm, double* dt, void* stream) {
    if (!params || !grid || !u || !v || !w || !dt || !(cfl > 0) || !std::isfinite(cfl)) return FS_ERR_BAD_ARG;
    const fs_grid& G = *grid;
    if (G.nx < 3 || G.ny < 3 || G.nz < 3) return FS_ERR_BAD_ARG;
    const long nu = (long)(G.nx + 1) * G.ny * G.nz, nv = (long)G.nx * (G.ny + 1) * G.nz,
               nw = (long)G.nx * G.ny * (G.nz + 1);
    // the scratch array holds FS_NOALLOC_BLOCKS partials of at most
    // FS_NOALLOC_CHUNK faces each per component
    if (nu > FS_NOALLOC_MAX_FACES || nv > FS_NOALLOC_MAX_FACES || nw > FS_NOALLOC_MAX_FACES) return FS_ERR_OOM;
    if (params->surface_tension != 0.0) return FS_ERR_UNSUPPORTED;  // its term needs the level set's curvature
    double* part = nullptr;
    int rc = scratch_address(&part);
    if (rc != FS_OK) return rc;
    const FsCflTerms t = fs_cfl_terms(*params, G, gravity_norm);
    const Comp c0 = make_comp(u, nu), c1 = make_comp(v, nv), c2 = make_comp(w, nw);
    const int nb = std::max(c0.nb, std::max(c1.nb, c2.nb));
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_partial_abs_max, dim3(nb, 3), dim3(NT), 0, st, c0, c1, c2, part);
    hipLaunchKernelGGL(k_cfl_finish, dim3(1), dim3(NT), 0, st, part, c0.nb, c1.nb, c2.nb, t, cfl, dt);
    return hipGetLastError() == hipSuccess ? FS_OK : FS_ERR_HIP;
}
