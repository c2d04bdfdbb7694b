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

// One 32-bit DPP move of each half of a double (all lanes active).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double lane_d(double v, int lane) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Wave64 sum, the same bits in every lane: butterfly within each 16-lane row
// by DPP (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror; every lane of a row ends with the identical row sum because
// IEEE addition is commutative), then the four row sums in a fixed order.
// Must be called with the whole wave active.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return ((lane_d(v, 0) + lane_d(v, 16)) + lane_d(v, 32)) + lane_d(v, 48);
}

__device__ __forceinline__ int linear_block() {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

__device__ __forceinline__ int linear_tid() {
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}

// Per-thread values of a 512-thread workgroup staged in LDS for wave 0.
template <int NQ>
struct RawLDS {
    double v[NQ][512];
};

// Wave 0 sums the NQ columns of a RawLDS: lane l adds entries l, l+64, ...,
// l+448 in that order, then the wave sum; the same bits in every lane of
// wave 0 (the other waves must not call it).
template <int NQ>
__device__ __forceinline__ void wave0_total(const RawLDS<NQ>& R, double (&tot)[NQ]) {
    const int lane = linear_tid() & 63;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double a = R.v[q][lane];
#pragma unroll
        for (int m = 1; m < 8; ++m) a += R.v[q][lane + 64 * m];
        tot[q] = wave_sum(a);
    }
}

// Workgroup sum of NQ per-thread values, stored by thread 0 at
// out[q*stride + blk].  Uses R as scratch (one barrier inside); the caller
// must not reuse R before a later barrier.
template <int NQ>
__device__ __forceinline__ void write_block_partials(const double (&v)[NQ], double* out, int stride, int blk,
                                                     RawLDS<NQ>& R) {
    const int tid = linear_tid();
#pragma unroll
    for (int q = 0; q < NQ; ++q) R.v[q][tid] = v[q];
    __syncthreads();
    if (tid < 64) {
        double tot[NQ];
        wave0_total<NQ>(R, tot);
        if (tid == 0) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) out[(long long)q * stride + blk] = tot[q];
        }
    }
}

template <int NQ>
__device__ __forceinline__ void write_block_partials(const double (&v)[NQ], double* out, int stride, int blk) {
    __shared__ RawLDS<NQ> R;
    write_block_partials<NQ>(v, out, stride, blk, R);
}

// Cross-kernel reduction of the PCG partials.  Partial arrays have a padded
// stride `ps` = PM * 512 whose tail beyond the number of producing
// workgroups is zero for the life of the buffer, so thread t of every
// consuming workgroup loads entries t, t + 512, ... (PM of them) with no
// bounds test; PM = 0 selects a loop over `nb` entries for very large films.
// The raw values are loaded first (load_partials) and staged for wave 0
// later (stage_partials), so the loads stay in flight behind the tile loads.
// Every workgroup adds the partials in the same fixed order.
template <int NQ, int PM>
struct Partials {
    double v[NQ][PM > 0 ? PM : 1];
};

// An empty asm statement that takes v as an operand: the value must exist
// at this point, which keeps the load that produces it ahead of a branch.
template <class V>
__device__ __forceinline__ void keep_live(V v) {
    asm volatile("" ::"v"(v));
}
template <int NQ, int PM>
__device__ __forceinline__ void keep_live_partials(const Partials<NQ, PM>& P) {
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int m = 0; m < (PM > 0 ? PM : 1); ++m) keep_live(P.v[q][m]);
}

template <int NQ, int PM>
__device__ __forceinline__ void load_partials(const double* in, int nb, int ps, Partials<NQ, PM>& P) {
    const int tid = linear_tid();
    if (PM > 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int m = 0; m < (PM > 0 ? PM : 1); ++m) P.v[q][m] = in[(long long)q * ps + tid + m * 512];
    } else {
#pragma unroll
        for (int q = 0; q < NQ; ++q) P.v[q][0] = 0.0;
        for (int t = tid; t < nb; t += 512) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) P.v[q][0] += in[(long long)q * ps + t];
        }
    }
}

template <int NQ, int PM>
__device__ __forceinline__ void stage_partials(const Partials<NQ, PM>& P, RawLDS<NQ>& R) {
    const int tid = linear_tid();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double a = P.v[q][0];
#pragma unroll
        for (int m = 1; m < (PM > 0 ? PM : 1); ++m) a += P.v[q][m];
        R.v[q][tid] = a;
    }
}


// Workgroup totals by waves (512 threads = 8 waves): each wave's DPP sum of
// one value per lane goes to slot [q][wave]; after a barrier the 8 slots
// are added in wave order.  No 512-entry staging and no single-wave
// bottleneck; every thread that adds the slots gets the same bits.
template <int NQ>
struct WaveLDS {
    double v[NQ][8];
};

// Consuming side: thread sums of the PM partials (in order), then wave sums
// into W.  Call with the whole workgroup active, then barrier, then
// wave_lds_total.
template <int NQ, int PM>
__device__ __forceinline__ void wave_partials(const Partials<NQ, PM>& P, WaveLDS<NQ>& W) {
    const int tid = linear_tid();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double a = P.v[q][0];
#pragma unroll
        for (int m = 1; m < (PM > 0 ? PM : 1); ++m) a += P.v[q][m];
        a = wave_sum(a);
        if ((tid & 63) == 0) W.v[q][tid >> 6] = a;
    }
}

template <int NQ>
__device__ __forceinline__ void wave_lds_total(const WaveLDS<NQ>& W, double (&tot)[NQ]) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double t = W.v[q][0];
#pragma unroll
        for (int w = 1; w < 8; ++w) t += W.v[q][w];
        tot[q] = t;
    }
}

// Producing side: workgroup sum of NQ per-thread values, stored by thread 0
// at out[q*stride + blk] (one barrier inside; W must not be read by anyone
// else in this launch).
template <int NQ>
__device__ __forceinline__ void write_block_partials_w(const double (&v)[NQ], double* out, int stride, int blk,
                                                       WaveLDS<NQ>& W) {
    const int tid = linear_tid();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double a = wave_sum(v[q]);
        if ((tid & 63) == 0) W.v[q][tid >> 6] = a;
    }
    __syncthreads();
    if (tid == 0) {
        double tot[NQ];
        wave_lds_total<NQ>(W, tot);
#pragma unroll
        for (int q = 0; q < NQ; ++q) out[(long long)q * stride + blk] = tot[q];
    }
}

}  // namespace hzo
