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
// out[q*stride + block].
template <int NQ>
__device__ __forceinline__ void write_block_partials(const double (&v)[NQ], double* out, int stride) {
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
            out[(long long)q * stride + b] = s;
        }
    }
}

// Cross-kernel reduction of the PCG partials.  Partial arrays have a padded
// stride `ps` = PM * (threads per workgroup) whose tail beyond the number of
// producing workgroups is zero for the life of the buffer, so thread t of
// every consuming workgroup loads entries t, t + NT, ... (PM of them) with
// no bounds test, and every workgroup adds them in the same fixed order.
// PM = 0 selects a loop over `nb` entries for very large films.
//
// The raw values are loaded first (load_partials) and summed later
// (finish_partials), so the loads stay in flight behind the tile loads.
template <int NQ, int PM>
struct Partials {
    double v[NQ][PM > 0 ? PM : 1];
};

template <int NQ, int PM>
__device__ __forceinline__ void load_partials(const double* in, int nb, int ps, Partials<NQ, PM>& P) {
    const int tid = linear_tid();
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    if (PM > 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int m = 0; m < (PM > 0 ? PM : 1); ++m) P.v[q][m] = in[(long long)q * ps + tid + m * nthreads];
    } else {
#pragma unroll
        for (int q = 0; q < NQ; ++q) P.v[q][0] = 0.0;
        for (int t = tid; t < nb; t += nthreads) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) P.v[q][0] += in[(long long)q * ps + t];
        }
    }
}

template <int NQ, int PM>
__device__ __forceinline__ void finish_partials(const Partials<NQ, PM>& P, double (&tot)[NQ]) {
    __shared__ double sred[NQ][16];
    const int tid = linear_tid();
    const int nthreads = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6, nw = nthreads >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double a = P.v[q][0];
#pragma unroll
        for (int m = 1; m < (PM > 0 ? PM : 1); ++m) a += P.v[q][m];
        const double w = wave_sum(a);
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

template <int NQ, int PM>
__device__ __forceinline__ void sum_partials(const double* in, int nb, int ps, double (&tot)[NQ]) {
    Partials<NQ, PM> P;
    load_partials<NQ, PM>(in, nb, ps, P);
    finish_partials<NQ, PM>(P, tot);
}

}  // namespace hzo
