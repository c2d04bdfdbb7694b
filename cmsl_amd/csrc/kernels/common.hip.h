// common.hip.h -- launch geometry and index helpers shared by the FP64 and
// FP32-refined kernels.
#pragma once
#include "../hzo_internal.h"
#include <hip/hip_ext.h>

namespace hzo {

inline int blocks_for_impl(const Grid& g) {
    return ((g.nx + BX - 1) / BX) * ((g.ny + BY - 1) / BY) * ((g.nz + g.kz - 1) / g.kz);
}

// PCG partial sums: PM entries per thread of a consuming workgroup, over a
// zero-padded stride of PM*NT (PM = 0: plain stride nb, loop reduction).
inline int part_pm(int nb) { return nb <= NT ? 1 : nb <= 2 * NT ? 2 : nb <= 4 * NT ? 4 : 0; }
inline int part_stride_impl(int nb) {
    const int pm = part_pm(nb);
    return pm ? pm * NT : nb;
}

// Launch of a PCG kernel template <int KZ, int PM, EXTRA...> with the
// instance that matches the grid's tile depth and the layout of the partials
// it reads (nbin of them).  `X` is the
// kernel's trailing template argument list (or nothing).
#define HZO_PCG_LAUNCH_PM(kern, KZ, pm, gr, bl, s, ...)                                   \
    switch (pm) {                                                                         \
        case 1: hipExtLaunchKernelGGL((kern<KZ, 1 HZO_X>), gr, bl, 0, s, ev.start, ev.stop, 0, __VA_ARGS__); break;  \
        case 2: hipExtLaunchKernelGGL((kern<KZ, 2 HZO_X>), gr, bl, 0, s, ev.start, ev.stop, 0, __VA_ARGS__); break;  \
        case 4: hipExtLaunchKernelGGL((kern<KZ, 4 HZO_X>), gr, bl, 0, s, ev.start, ev.stop, 0, __VA_ARGS__); break;  \
        default: hipExtLaunchKernelGGL((kern<KZ, 0 HZO_X>), gr, bl, 0, s, ev.start, ev.stop, 0, __VA_ARGS__); break; \
    }
#define HZO_PCG_LAUNCH_X(kern, g, nbin, s, ...)                                           \
    do {                                                                                  \
        const int pm_ = part_pm(nbin);                                                    \
        const dim3 gr_ = grid3(g), bl_(BX, BY, 1);                                        \
        switch ((g).kz) {                                                                 \
            case 2: HZO_PCG_LAUNCH_PM(kern, 2, pm_, gr_, bl_, s, __VA_ARGS__); break;     \
            case 4: HZO_PCG_LAUNCH_PM(kern, 4, pm_, gr_, bl_, s, __VA_ARGS__); break;     \
            default: HZO_PCG_LAUNCH_PM(kern, 8, pm_, gr_, bl_, s, __VA_ARGS__); break;    \
        }                                                                                 \
    } while (0)

inline dim3 grid3(const Grid& g) {
    return dim3((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, (g.nz + g.kz - 1) / g.kz);
}

// XCD-aware tile order.  MI355X dispatches workgroup w (linear index) to XCD
// w % 8, and each XCD has its own L2.  Tiles are numbered x fastest, then z,
// then y, and each XCD receives one contiguous run of that numbering (a slab
// in y), so the halo columns, rows and planes a tile shares with its
// neighbours are mostly fetched by workgroups of the same XCD and served from
// that L2 instead of crossing the fabric once per neighbour.
constexpr int N_XCD = 8;

struct TileId {
    int bx, by, bz;
    int blk;    // logical tile index within the film: its partial-sum slot
    bool lead;  // the film's tile 0, which passes the control block on
};

// Tile coordinates of logical tile t of a film with gx x gy x gz tiles,
// numbered x fastest, then z, then y.
__device__ __forceinline__ TileId tile_of(int t, int gx, int gy, int gz) {
    TileId T;
    T.bx = t % gx;
    const int r = t / gx;
    T.bz = r % gz;
    T.by = r / gz;
    T.blk = t;
    T.lead = t == 0;
    (void)gy;
    return T;
}

__device__ __forceinline__ TileId tile_id() {
    const int gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const int nbk = gx * gy * gz;
    const int w = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int xcd = w % N_XCD, loc = w / N_XCD;
    const int per = nbk / N_XCD, rem = nbk % N_XCD;
    const int t = xcd * per + min(xcd, rem) + loc;
    return tile_of(t, gx, gy, gz);
}

// Batched launches: workgroup w of a 1-D grid over all tiles of all films
// takes tile t of the XCD-aware order (contiguous runs per XCD, as above);
// tiles[t] = (film, logical tile within the film).  Returns the film index.
__device__ __forceinline__ int batch_tile(const int2* tiles, const BatchFilm* films, TileId& T) {
    const int nbk = gridDim.x;
    const int w = blockIdx.x;
    const int xcd = w % N_XCD, loc = w / N_XCD;
    const int per = nbk / N_XCD, rem = nbk % N_XCD;
    const int t = xcd * per + min(xcd, rem) + loc;
    const int2 e = tiles[t];
    const BatchFilm& F = films[e.x];
    T = tile_of(e.y, F.gx, F.gy, F.gz);
    return e.x;
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

// Inner-product term of plane k in the FP32 solves' CG sums (r.z, p.Ap,
// w.u): acc += w_k x y.  Under Neumann electrodes a mirrored plane (k = 0 or
// nz - 1) couples to its one neighbour plane with twice the coefficient that
// plane couples back with: A = S A_s, S = 2 on those planes and 1 elsewhere,
// A_s symmetric.  Single-reduction CG needs a symmetric operator (its p.Ap
// recurrence assumes u.Ap = p.Au) and stalls on A itself, so the FP32 solves
// run CG on A_s x = S^-1 b, which has the same solution: with Jacobi
// D = S D_s the iterates stay those of A and only these sums weight the
// mirrored planes by w = 1/2.  The half is taken as (2 acc + x y) / 2, two
// power-of-two scalings of the running sum, so the result is exactly the
// weighted sum and no product is held beside it.  r.r and b.b, which the
// relative-residual rule reads, stay unweighted.  MIR = false compiles the
// plain sum, for kernels launched only on films with Dirichlet electrodes.
// The FP64 form accumulates in FP64 (setup, residual replacement, the
// two-kernel PCG); cg_dot_addf accumulates a thread's own planes in FP32
// (the single-reduction CG iterations, at most 32 terms per thread), whose
// sums are then added across threads, waves and workgroups in FP64.
template <bool MIR>
__device__ __forceinline__ void cg_dot_add(double& acc, float x, float y, const StepConsts& c, int k) {
    if constexpr (MIR) {
        const bool mirrored = c.zbc != 0 && (k == 0 || k == c.nz - 1);
        const double up = mirrored ? 2.0 : 1.0, down = mirrored ? 0.5 : 1.0;  // uniform
        acc = __builtin_fma((double)x, (double)y, acc * up) * down;
    } else {
        acc = __builtin_fma((double)x, (double)y, acc);
    }
}

template <bool MIR>
__device__ __forceinline__ void cg_dot_addf(float& acc, float x, float y, const StepConsts& c, int k) {
    if constexpr (MIR) {
        const bool mirrored = c.zbc != 0 && (k == 0 || k == c.nz - 1);
        const float up = mirrored ? 2.0f : 1.0f, down = mirrored ? 0.5f : 1.0f;  // uniform
        acc = __builtin_fmaf(x, y, acc * up) * down;
    } else {
        acc = __builtin_fmaf(x, y, acc);
    }
}

// v as the wave's first active lane holds it: for values every lane
// computes alike, which the compiler then keeps in a scalar register.
__device__ __forceinline__ float uniform_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

__device__ __forceinline__ bool keep_going(const PcgState& st, double rz, double rr) {
    if (st.iters >= st.max_iter) return false;
    if (st.rule == 0) return rz > st.tol_abs2 && (st.rz0 <= 0.0 || (rz / st.rz0) > st.tol_rel2);
    return rr > st.tol_rel2 * st.bb && rr > st.tol_abs2;
}

__device__ __forceinline__ bool finite(double v) { return __builtin_isfinite(v); }

__device__ __forceinline__ bool block0_thread0() {
    return blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0;
}

// Field-wise copies of the control block.  Aggregate copies would be lowered
// to memcpy through private memory; field by field it stays in registers.
__device__ __forceinline__ PcgState load_state(const PcgState* p) {
    PcgState s;
    s.rz_old = p->rz_old;
    s.rz0 = p->rz0;
    s.rr = p->rr;
    s.rr0 = p->rr0;
    s.bb = p->bb;
    s.alpha = p->alpha;
    s.beta = p->beta;
    s.denom = p->denom;
    s.iters = p->iters;
    s.status = p->status;
    s.done = p->done;
    s.refinements = p->refinements;
    s.max_iter = p->max_iter;
    s.rule = p->rule;
    s.kind = p->kind;
    s.pbuf = p->pbuf;
    s.tol_abs2 = p->tol_abs2;
    s.tol_rel2 = p->tol_rel2;
    s.refine_every = p->refine_every;
    s.xpend = p->xpend;
    s.xadds = p->xadds;
    s.passes = p->passes;
    s.alpha_prev = p->alpha_prev;
    return s;
}

__device__ __forceinline__ void store_state(const PcgState& s, PcgState* p) {
    p->rz_old = s.rz_old;
    p->rz0 = s.rz0;
    p->rr = s.rr;
    p->rr0 = s.rr0;
    p->bb = s.bb;
    p->alpha = s.alpha;
    p->beta = s.beta;
    p->denom = s.denom;
    p->iters = s.iters;
    p->status = s.status;
    p->done = s.done;
    p->refinements = s.refinements;
    p->max_iter = s.max_iter;
    p->rule = s.rule;
    p->kind = s.kind;
    p->pbuf = s.pbuf;
    p->tol_abs2 = s.tol_abs2;
    p->tol_rel2 = s.tol_rel2;
    p->refine_every = s.refine_every;
    p->xpend = s.xpend;
    p->xadds = s.xadds;
    p->passes = s.passes;
    p->alpha_prev = s.alpha_prev;
}

// Copies the control block forward (thread 0 of the film's lead tile).
__device__ __forceinline__ void pass_state(const PcgState& s, PcgState* out, bool lead) {
    if (lead && threadIdx.x == 0 && threadIdx.y == 0) store_state(s, out);
}

// The stopping test after a completed reduction of r.z and r.r (the
// reference loop head, depol_varEPS.cu:472-474, or ||r|| <= tol ||b||).
__device__ __forceinline__ void judge(PcgState& s, double rz, double rr) {
    if (!keep_going(s, rz, rr)) {
        s.done = 1;
        if (s.iters >= s.max_iter) s.status = HZO_PCG_MAXITER;
    }
}

// Jacobi inverse diagonal of a row from its stencil code and the harmonic
// face table H (depol_varEPS.cu:177-179: D = sum of the six face means / h^2,
// D^-1 = 1/D, or 1 when D <= 1e-30 and on identity rows).  In FP64 this is
// the same sequence of IEEE operations as the reference.
__device__ __forceinline__ double dinv_of(unsigned code, const double* H, const StepConsts& c) {
    if (code & SC_IDENTITY) return 1.0;
    const double* Hr = H + (code & 3u) * 4u;
    const double sum = Hr[(code >> 2) & 3u] + Hr[(code >> 4) & 3u] + Hr[(code >> 6) & 3u] + Hr[(code >> 8) & 3u] +
                       Hr[(code >> 10) & 3u] + Hr[(code >> 12) & 3u];
    const double D = c.h2_pow2 ? sum * c.inv_h2 : sum / c.h2;
    return D > 1e-30 ? 1.0 / D : 1.0;
}

// FP32 form: the hardware reciprocal (1 ulp) replaces the division; every
// FP32 kernel uses this same function, so the preconditioner is one fixed
// diagonal matrix throughout a solve.
__device__ __forceinline__ float dinv_of(unsigned code, const float* H, const StepConsts& c) {
    if (code & SC_IDENTITY) return 1.0f;
    const float* Hr = H + (code & 3u) * 4u;
    const float sum = Hr[(code >> 2) & 3u] + Hr[(code >> 4) & 3u] + Hr[(code >> 6) & 3u] + Hr[(code >> 8) & 3u] +
                      Hr[(code >> 10) & 3u] + Hr[(code >> 12) & 3u];
    const float D = sum * (float)c.inv_h2;
    return D > 1e-30f ? __builtin_amdgcn_rcpf(D) : 1.0f;
}

template <class T>
__device__ __forceinline__ void load_lut(T* sH, const T* lut) {
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    if (tid < 16) sH[tid] = lut[LUT_HARM + tid];
    __syncthreads();
}

}  // namespace hzo
