// reduce.hip.h -- deterministic wave64 / workgroup reductions and the
// fence-free grid reduction protocol.
//
// A kernel that produces an inner product writes one partial sum per
// workgroup (plain stores).  The NEXT kernel on the stream finishes the
// reduction: every one of its workgroups adds the partials in the same fixed
// order, so all workgroups obtain bit-identical totals without atomics or
// memory fences (the kernel boundary makes the partials visible on every
// XCD).  Partials and the PCG control block are double-buffered by launch
// parity, so a kernel never overwrites what its own workgroups still read.
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

__device__ __forceinline__ int linear_tid() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}

// Workgroup sum of NQ per-thread values; thread 0 stores them at
// out[q*nb + block].
template <int NQ>
__device__ __forceinline__ void write_block_partials(const double (&v)[NQ], double* out, int nb) {
    __shared__ double sred[NQ][16];
    const int tid = linear_tid();
    const int lane = tid & 63, wid = tid >> 6;
    const int nw = (blockDim.x * blockDim.y * blockDim.z) >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double w = wave_sum(v[q]);
        if (lane == 0) sred[q][wid] = w;
    }
    __syncthreads();
    if (tid == 0) {
        const int b = linear_block();
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += sred[q][w];
            out[(long long)q * nb + b] = s;
        }
    }
}

// Totals of the NQ partial arrays written by the previous kernel; identical
// in every thread of every workgroup.
template <int NQ>
__device__ __forceinline__ void sum_partials(const double* in, int nb, double (&tot)[NQ]) {
    __shared__ double sred[NQ][16];
    const int tid = linear_tid();
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6, nw = nthreads >> 6;
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = 0.0;
    for (int t = tid; t < nb; t += nthreads) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) acc[q] += in[(long long)q * nb + t];
    }
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
    __syncthreads();
}

}  // namespace hzo
