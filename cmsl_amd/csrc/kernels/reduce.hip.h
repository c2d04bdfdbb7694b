// reduce.hip.h -- deterministic wave64 / workgroup / grid reductions.
//
// A grid reduction is finished by the last workgroup to arrive (completion
// counter), which adds the per-workgroup partial sums in a fixed order.  The
// result therefore does not depend on workgroup scheduling, and the scalar
// never leaves the device.  Partials and the counter use agent-scope atomics
// so that workgroups on different XCDs (separate L2s) see each other's data.
#pragma once
#include <hip/hip_runtime.h>

namespace hzo {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int linear_block() {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

// Reduces NQ per-thread values over the whole grid.  Returns true in every
// thread of the last workgroup, with tot[] holding the grid totals.
template <int NQ>
__device__ bool grid_reduce_last(const double (&v)[NQ], double* partials, int nb,
                                 unsigned* counter, double (&tot)[NQ]) {
    __shared__ double sred[NQ][16];
    __shared__ int s_last;
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int lane = tid & 63, wid = tid >> 6, nw = nthreads >> 6;
    const int b = linear_block();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double w = wave_sum(v[q]);
        if (lane == 0) sred[q][wid] = w;
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += sred[q][w];
            __hip_atomic_store(&partials[(long long)q * nb + b], s, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        __threadfence();
        const unsigned prev =
            __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (prev == (unsigned)nb - 1u);
    }
    __syncthreads();
    if (!s_last) return false;
    __threadfence();
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    for (int t = tid; t < nb; t += nthreads) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            acc[q] += __hip_atomic_load(&partials[(long long)q * nb + t], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double w = wave_sum(acc[q]);
        if (lane == 0) sred[q][wid] = w;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += sred[q][w];
        tot[q] = s;
    }
    if (tid == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

}  // namespace hzo
