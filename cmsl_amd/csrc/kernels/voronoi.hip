// voronoi.hip -- nearest-centroid labelling of one layer of a layered
// periodic Voronoi structure on the GPU (the labelling loop of
// voronoi.c:118-153, restated).  Centroids are integer voxel coordinates, so
// the minimal-image offsets and squared distances are exact integers; the
// strict '<' keeps the lowest index on ties, which makes the labels equal to
// the CPU labeller's bit for bit.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>
#include <vector>

#include "../host/errors.h"
#include "../hzo_internal.h"

namespace hzo {

constexpr int VOR_TILE = 1024;  // centroids staged in LDS per pass
// Grain capacity of the no-allocation entry point's centroid scratch
// (3 doubles per grain: 1.5 MiB per device).
constexpr int VOR_NOALLOC_MAX_GRAINS = HZO_VORONOI_NOALLOC_MAX_GRAINS;

// One thread per voxel of the slab [z0, z0 + span): omega = goff + index of
// the nearest of the nL centroids (x, y periodic on nx, ny; z periodic on
// the slab's span, relative to z0).  cent holds the layer's centroids as
// integer-valued doubles (x, y, z absolute), converted while staged in LDS.
__global__ __launch_bounds__(256) void k_voronoi_label(int nx, int ny, int z0, int span, int nL, int goff,
                                                       const double* __restrict__ cent, int32_t* __restrict__ omega) {
    __shared__ int3 sc[VOR_TILE];
    const long long nslab = (long long)nx * ny * span;
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = v < nslab;
    const int i = live ? (int)(v % nx) : 0;
    const int j = live ? (int)((v / nx) % ny) : 0;
    const int h = live ? (int)(v / ((long long)nx * ny)) : 0;  // slab-relative z
    long long rmin = -1;
    int nmin = 0;
    for (int base = 0; base < nL; base += VOR_TILE) {
        const int cnt = min(VOR_TILE, nL - base);
        __syncthreads();
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
            const double* c = cent + 3 * (long long)(base + t);
            sc[t] = make_int3((int)c[0], (int)c[1], (int)c[2] - z0);
        }
        __syncthreads();
        for (int t = 0; t < cnt; ++t) {
            const int3 c = sc[t];
            int dx = c.x - i, dy = c.y - j, dz = c.z - h;
            // minimal image: d > n/2 -> d - n, d < -n/2 -> d + n (2d against n keeps it exact)
            if (2 * dx > nx) dx -= nx;
            else if (2 * dx < -nx) dx += nx;
            if (2 * dy > ny) dy -= ny;
            else if (2 * dy < -ny) dy += ny;
            if (2 * dz > span) dz -= span;
            else if (2 * dz < -span) dz += span;
            const long long r2 = (long long)dx * dx + (long long)dy * dy + (long long)dz * dz;
            if (rmin < 0 || r2 < rmin) {
                rmin = r2;
                nmin = base + t;
            }
        }
    }
    if (live) omega[(long long)z0 * nx * ny + v] = goff + nmin;
}

namespace {

// Uploads the centroids to d_cent and labels every layer into d_omega on
// the null stream, then waits for it.
int label_layers(int nx, int ny, int n_layers, const int* goffs, const int* nls, const int* z0s, const int* spans,
                 const double* cent, int ng, double* d_cent, int32_t* d_omega, const char* who) {
    if (hipMemcpyAsync(d_cent, cent, sizeof(double) * 3 * (size_t)ng, hipMemcpyHostToDevice, 0) != hipSuccess)
        return set_error(HZO_ERR_HIP, "%s: copy failed", who);
    for (int L = 0; L < n_layers; ++L) {
        if (nls[L] <= 0) continue;
        const long long nslab = (long long)nx * ny * spans[L];
        const unsigned blocks = (unsigned)((nslab + 255) / 256);
        hipLaunchKernelGGL(k_voronoi_label, dim3(blocks), dim3(256), 0, 0, nx, ny, z0s[L], spans[L], nls[L], goffs[L],
                           d_cent + 3 * (size_t)goffs[L], d_omega);
        if (hipGetLastError() != hipSuccess) return set_error(HZO_ERR_HIP, "%s: launch failed", who);
    }
    if (hipStreamSynchronize(0) != hipSuccess) return set_error(HZO_ERR_HIP, "%s: labelling failed", who);
    return HZO_OK;
}

int select_device(int device, const char* who) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return set_error(HZO_ERR_HIP, "%s: device %d not available", who, device);
    if (hipSetDevice(device) != hipSuccess) return set_error(HZO_ERR_HIP, "%s: hipSetDevice failed", who);
    return HZO_OK;
}

// Centroid staging of the entry point that allocates nothing: a file-scope
// device array (each GPU holds its own copy until the process ends).  Host
// code never passes its name to a kernel or a copy: it uses the device
// address hipGetSymbolAddress returns, looked up once per device with that
// device current.  Calls that overlap on one GPU share this array and race.
__device__ __align__(256) double scratch[3 * VOR_NOALLOC_MAX_GRAINS];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

int scratch_for(int device, double** out) {
    if (device >= MAX_DEVICES) return set_error(HZO_ERR_HIP, "hzo_voronoi_gpu_noalloc: device index %d too high", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return set_error(HZO_ERR_HIP, "hzo_voronoi_gpu_noalloc: no device address for the centroid scratch");
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return HZO_OK;
}

}  // namespace

// Labels every layer; cent holds all ng centroids (double, integer-valued),
// layers give (first grain, grain count, first plane, planes) per layer.
int voronoi_label_gpu(int device, int nx, int ny, int nz, int n_layers, const int* goffs, const int* nls,
                      const int* z0s, const int* spans, const double* cent, int ng, int32_t* omega_host) {
    const long long N = (long long)nx * ny * nz;
    int rc = select_device(device, "hzo_voronoi_gpu");
    if (rc) return rc;
    int32_t* d_omega = nullptr;
    double* d_cent = nullptr;
    if (hipMalloc((void**)&d_omega, sizeof(int32_t) * N) != hipSuccess ||
        hipMalloc((void**)&d_cent, sizeof(double) * 3 * (size_t)ng) != hipSuccess)
        rc = set_error(HZO_ERR_NOMEM, "hzo_voronoi_gpu: device allocation failed");
    if (rc == HZO_OK) rc = label_layers(nx, ny, n_layers, goffs, nls, z0s, spans, cent, ng, d_cent, d_omega,
                                        "hzo_voronoi_gpu");
    if (rc == HZO_OK && hipMemcpy(omega_host, d_omega, sizeof(int32_t) * N, hipMemcpyDeviceToHost) != hipSuccess)
        rc = set_error(HZO_ERR_HIP, "hzo_voronoi_gpu: copy back failed");
    (void)hipFree(d_omega);
    (void)hipFree(d_cent);
    return rc;
}

// The same labelling into the caller's device array, with the centroids
// staged in the file-scope scratch: no device allocation.
int voronoi_label_gpu_noalloc(int device, int nx, int ny, int n_layers, const int* goffs, const int* nls,
                              const int* z0s, const int* spans, const double* cent, int ng, int32_t* omega_dev) {
    if (ng > VOR_NOALLOC_MAX_GRAINS)
        return set_error(HZO_ERR_ARG, "hzo_voronoi_gpu_noalloc: %d grains exceed the scratch limit of %d", ng,
                         VOR_NOALLOC_MAX_GRAINS);
    int rc = select_device(device, "hzo_voronoi_gpu_noalloc");
    if (rc) return rc;
    double* d_cent = nullptr;
    rc = scratch_for(device, &d_cent);
    if (rc) return rc;
    return label_layers(nx, ny, n_layers, goffs, nls, z0s, spans, cent, ng, d_cent, omega_dev,
                        "hzo_voronoi_gpu_noalloc");
}

}  // namespace hzo
