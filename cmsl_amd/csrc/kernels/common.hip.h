// common.hip.h -- launch geometry and index helpers shared by the FP64 and
// FP32-refined kernels.
#pragma once
#include "../hzo_internal.h"

namespace hzo {

constexpr int BX = 64, BY = 4, KZ = 4;

inline int blocks_for_impl(const Grid& g) {
    return ((g.nx + BX - 1) / BX) * ((g.ny + BY - 1) / BY) * ((g.nz + KZ - 1) / KZ);
}

inline dim3 grid3(const Grid& g) {
    return dim3((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, (g.nz + KZ - 1) / KZ);
}

__device__ __forceinline__ long long vid(int i, int j, int k, int nx, int ny) {
    return (long long)i + (long long)nx * ((long long)j + (long long)ny * k);
}

// ---------------------------------------------------------------- operator
struct Nbr {
    long long c, e, w, n, s, u, d;
};

__device__ __forceinline__ Nbr op_nbrs(int i, int j, int k, const StepConsts& c) {
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const int e = (i + 1 == nx) ? 0 : i + 1, w = (i == 0) ? nx - 1 : i - 1;
    const int n = (j + 1 == ny) ? 0 : j + 1, s = (j == 0) ? ny - 1 : j - 1;
    int up, dn;
    if (c.zbc == 0) {
        up = k + 1;
        dn = k - 1;
    } else if (k == 0) {
        up = 1;
        dn = 1;
    } else if (k == nz - 1) {
        up = nz - 2;
        dn = nz - 2;
    } else {
        up = k + 1;
        dn = k - 1;
    }
    Nbr r;
    r.c = vid(i, j, k, nx, ny);
    r.e = vid(e, j, k, nx, ny);
    r.w = vid(w, j, k, nx, ny);
    r.n = vid(i, n, k, nx, ny);
    r.s = vid(i, s, k, nx, ny);
    r.u = vid(i, j, up, nx, ny);
    r.d = vid(i, j, dn, nx, ny);
    return r;
}

// identity rows: y = x - g on Dirichlet planes, y = x at the Neumann gauge node
template <class T>
__device__ __forceinline__ T op_identity(T xc, int k, const StepConsts& c) {
    if (c.zbc == 0) return xc - (T)((k == 0) ? c.g_bot : c.g_top);
    return xc;
}

__device__ __forceinline__ bool keep_going(const PcgState& st, double rz, double rr) {
    if (st.iters >= st.max_iter) return false;
    if (st.rule == 0) return rz > st.tol_abs2 && (st.rz0 <= 0.0 || (rz / st.rz0) > st.tol_rel2);
    return rr > st.tol_rel2 * st.bb && rr > st.tol_abs2;
}

__device__ __forceinline__ bool finite(double v) { return __builtin_isfinite(v); }

template <class T>
__device__ __forceinline__ void load_lut(T* sH, const T* lut) {
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    if (tid < 16) sH[tid] = lut[LUT_HARM + tid];
    __syncthreads();
}

}  // namespace hzo
