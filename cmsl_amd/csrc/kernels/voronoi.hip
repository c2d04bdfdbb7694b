// voronoi.hip -- nearest-centroid labelling of one layer of a layered
// periodic Voronoi structure on the GPU (the labelling loop of
// voronoi.c:118-153, restated).  Centroids are integer voxel coordinates, so
// the minimal-image offsets and squared distances are exact integers; the
// strict '<' keeps the lowest index on ties, which makes the labels equal to
// the CPU labeller's bit for bit.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "../host/errors.h"
#include "../hzo_internal.h"

namespace hzo {

constexpr int VOR_TILE = 1024;  // centroids staged in LDS per pass

// One thread per voxel of the slab [z0, z0 + span): omega = goff + index of
// the nearest of the nL centroids (x, y periodic on nx, ny; z periodic on
// the slab's span, relative to z0).
__global__ __launch_bounds__(256) void k_voronoi_label(int nx, int ny, int z0, int span, int nL, int goff,
                                                       const int3* __restrict__ cent, int32_t* __restrict__ omega) {
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
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) sc[t] = cent[base + t];
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

// Labels every layer; cent holds all ng centroids (double, integer-valued),
// layers give (first grain, grain count, first plane, planes) per layer.
int voronoi_label_gpu(int device, int nx, int ny, int nz, int n_layers, const int* goffs, const int* nls,
                      const int* z0s, const int* spans, const double* cent, int ng, int32_t* omega_host) {
    const long long N = (long long)nx * ny * nz;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device >= ndev)
        return set_error(HZO_ERR_HIP, "hzo_voronoi_gpu: device %d not available", device);
    if (hipSetDevice(device) != hipSuccess) return set_error(HZO_ERR_HIP, "hzo_voronoi_gpu: hipSetDevice failed");
    int32_t* d_omega = nullptr;
    int3* d_cent = nullptr;
    std::vector<int3> hc(ng > 0 ? ng : 1);
    for (int g = 0; g < ng; ++g)
        hc[g] = make_int3((int)cent[3 * g], (int)cent[3 * g + 1], (int)cent[3 * g + 2]);
    int rc = HZO_OK;
    if (hipMalloc((void**)&d_omega, sizeof(int32_t) * N) != hipSuccess ||
        hipMalloc((void**)&d_cent, sizeof(int3) * hc.size()) != hipSuccess)
        rc = set_error(HZO_ERR_NOMEM, "hzo_voronoi_gpu: device allocation failed");
    if (rc == HZO_OK && hipMemcpy(d_cent, hc.data(), sizeof(int3) * hc.size(), hipMemcpyHostToDevice) != hipSuccess)
        rc = set_error(HZO_ERR_HIP, "hzo_voronoi_gpu: copy failed");
    for (int L = 0; L < n_layers && rc == HZO_OK; ++L) {
        if (nls[L] <= 0) continue;
        // slab-relative z of this layer's centroids
        std::vector<int3> lc(nls[L]);
        for (int k = 0; k < nls[L]; ++k) {
            lc[k] = hc[goffs[L] + k];
            lc[k].z -= z0s[L];
        }
        if (hipMemcpy(d_cent + goffs[L], lc.data(), sizeof(int3) * lc.size(), hipMemcpyHostToDevice) != hipSuccess) {
            rc = set_error(HZO_ERR_HIP, "hzo_voronoi_gpu: copy failed");
            break;
        }
        const long long nslab = (long long)nx * ny * spans[L];
        const unsigned blocks = (unsigned)((nslab + 255) / 256);
        hipLaunchKernelGGL(k_voronoi_label, dim3(blocks), dim3(256), 0, 0, nx, ny, z0s[L], spans[L], nls[L], goffs[L],
                           d_cent + goffs[L], d_omega);
        if (hipGetLastError() != hipSuccess) rc = set_error(HZO_ERR_HIP, "hzo_voronoi_gpu: launch failed");
    }
    if (rc == HZO_OK && (hipDeviceSynchronize() != hipSuccess ||
                         hipMemcpy(omega_host, d_omega, sizeof(int32_t) * N, hipMemcpyDeviceToHost) != hipSuccess))
        rc = set_error(HZO_ERR_HIP, "hzo_voronoi_gpu: labelling failed");
    (void)hipFree(d_omega);
    (void)hipFree(d_cent);
    return rc;
}

}  // namespace hzo
