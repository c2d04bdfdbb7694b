// tile.hip.h -- 2.5D LDS tiling for the 7-point Poisson operator.
//
// A workgroup of 64 x 8 threads owns a 64 x 8 column tile and walks kz planes
// of it.  The operand field of each plane is staged in LDS with a one-voxel
// x-y halo (66 x 10 values), in a ring of three planes (k-1, k, k+1), so every
// voxel value is read from global memory once per plane and the stencil
// takes its six neighbours from LDS.  The loads of plane k+2 are issued
// before the stencil of plane k is evaluated, which keeps two planes of
// memory traffic in flight per workgroup.
//
// Thread (tx, ty) stages its own voxel (ty+1, tx+1); threads 0..147 also
// stage one halo voxel each (top row, bottom row, left and right column).
// Edge tiles wrap their loads periodically, which is exactly the periodic
// neighbour a boundary voxel needs.
#pragma once
#include "common.hip.h"

namespace hzo {

constexpr int TX = BX, TY = BY;
constexpr int HX = TX + 2, HY = TY + 2;
constexpr int NHALO = 2 * HX + 2 * TY;  // 148

__device__ __forceinline__ int wrapi(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

struct TilePos {
    int i, j;     // this thread's voxel column (may lie beyond the grid in edge tiles)
    int ci, cj;   // the same wrapped into the grid (always a valid address)
    bool live;    // i < nx && j < ny
    bool has_h;   // this thread stages a halo voxel
    int hx, hy;   // its LDS position
    int gi, gj;   // its grid column (wrapped)
};

__device__ __forceinline__ TilePos tile_pos(const StepConsts& c, const TileId& T) {
    TilePos t;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = T.bx * TX, j0 = T.by * TY;
    t.i = i0 + tx;
    t.j = j0 + ty;
    t.live = t.i < c.nx && t.j < c.ny;
    t.ci = wrapi(t.i, c.nx);
    t.cj = wrapi(t.j, c.ny);
    const int h = tx + TX * ty;
    t.has_h = h < NHALO;
    if (h < HX) {
        t.hy = 0;
        t.hx = h;
    } else if (h < 2 * HX) {
        t.hy = HY - 1;
        t.hx = h - HX;
    } else if (h < 2 * HX + TY) {
        t.hx = 0;
        t.hy = 1 + (h - 2 * HX);
    } else {
        t.hx = HX - 1;
        t.hy = 1 + (h - 2 * HX - TY);
    }
    t.gi = wrapi(i0 + t.hx - 1, c.nx);
    t.gj = wrapi(j0 + t.hy - 1, c.ny);
    return t;
}

// ---------------------------------------------------------------- buffer access
// Raw buffer loads/stores: a 128-bit resource (base, byte size) in SGPRs, a
// 32-bit per-lane byte offset (constant down a column) and a workgroup-
// uniform byte offset (the z plane) in an SGPR, so walking planes costs no
// per-lane address arithmetic.  The host limits a film to 2^28 voxels, which
// keeps every byte offset of an FP64 array below 2^31.
using rsrc_t = __amdgpu_buffer_rsrc_t;

__device__ __forceinline__ rsrc_t make_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

template <class T>
__device__ __forceinline__ T bload(rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float bload<float>(rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ double bload<double>(rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ unsigned bload_code(rsrc_t r, unsigned voff, unsigned soff) {
    return (unsigned)__builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
}
// Cache policy of the solver stores: device scope (the SC1 bit), so each
// line is written through the XCD's L2 as it is stored.  The next launch
// reads these vectors on every XCD, and a launch whose stores were written
// through leaves no dirty L2 lines for the write-back that ends it.  On a
// 256x256x32 film this took the single-reduction CG launch from 16.8 to
// 16.2 us and the FP64 PCG launches from 16.7 to 15.7 (direction) and
// 14.7 us (update).
constexpr int ST_DEVICE = 16;
// Nontemporal (the NT bit): films whose vectors exceed the last-level cache.
constexpr int ST_STREAM = 2;

__device__ __forceinline__ void bstore(rsrc_t r, unsigned voff, unsigned soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, ST_DEVICE);
}
__device__ __forceinline__ void bstore(rsrc_t r, unsigned voff, unsigned soff, double v) {
    using u2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)voff, (int)soff, ST_DEVICE);
}

// Column geometry of one thread of a tile: element offsets within a z plane
// of its own column and of the halo column it stages.  Every wave issues the
// halo loads (no branch, so the compiler keeps precise memory-counter waits);
// lanes that stage no halo voxel get the offset nvox, one past the end of
// every per-voxel array, which the buffer unit answers with 0 and no memory
// access.
struct TileCols {
    unsigned cofs, hofs;
};

__device__ __forceinline__ TileCols tile_cols(const TilePos& tp, const StepConsts& c) {
    TileCols g;
    g.cofs = (unsigned)tp.ci + (unsigned)c.nx * (unsigned)tp.cj;
    g.hofs = tp.has_h ? (unsigned)tp.gi + (unsigned)c.nx * (unsigned)tp.gj : c.nvox;
    return g;
}

// Paired harmonic face means: entry (c << 4) | (a | b << 2) holds
// (H[c][a], H[c][b]), so one LDS read serves the + and - face of one axis
// (stencil-code bits 2-5: e,w; 6-9: n,s; 10-13: u,d).  Every wave loads the
// 64 pairs (lane e loads entry e; two L2-resident values) as its first
// memory operation, so waiting for them never waits for a tile load; wave 0
// writes them to LDS after the kernel's first barrier point.
template <class T>
struct HPair {
    T p, m;
};

template <class T>
__device__ __forceinline__ HPair<T> load_hpair(const T* lut) {
    const unsigned e = (threadIdx.x + BX * threadIdx.y) & 63u;
    const unsigned row = (e >> 4) * 4u;
    HPair<T> h;
    h.p = lut[LUT_HARM + row + (e & 3u)];
    h.m = lut[LUT_HARM + row + ((e >> 2) & 3u)];
    return h;
}

// The z planes whose operand values the tile's planes [k0, k1) need: k0-1 and
// k1 when they exist (Neumann mirror rows also stay inside k-1..k+1).
__device__ __forceinline__ void plane_window(const StepConsts& c, int k0, int k1, int& lo, int& hi) {
    lo = k0 > 0 ? k0 - 1 : 0;
    hi = k1 < c.nz ? k1 : c.nz - 1;  // inclusive
}

}  // namespace hzo
