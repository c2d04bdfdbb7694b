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

__device__ __forceinline__ TilePos tile_pos(const StepConsts& c) {
    TilePos t;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
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

// The z planes whose operand values the tile's planes [k0, k1) need: k0-1 and
// k1 when they exist (Neumann mirror rows also stay inside k-1..k+1).
__device__ __forceinline__ void plane_window(const StepConsts& c, int k0, int k1, int& lo, int& hi) {
    lo = k0 > 0 ? k0 - 1 : 0;
    hi = k1 < c.nz ? k1 : c.nz - 1;  // inclusive
}

}  // namespace hzo
