// kernels_cg32.hip -- single-reduction CG for the FP32-refined Poisson solve.
//
// The Jacobi-preconditioned CG of the FP64 form needs two grid-wide
// reductions per iteration (p.Ap, then r.z).  On one MI355X each reduction
// costs a kernel boundary, and at 2 M voxels those boundaries and the
// latency they expose dominate the solve.  The FP32-refined form therefore
// runs the Chronopoulos-Gear recurrence (Chronopoulos & Gear 1989), which
// carries s = A p alongside p and needs one reduction of (gamma, delta) per
// iteration:
//   u = M r, w = A u, gamma = (r, u), delta = (w, u)
//   beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old)
//   p = u + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s
// One launch performs a whole iteration: it forms s and the new r for its
// tile and x-y halo (a recompute of the neighbours' update, which needs w,
// s, r of the halo), u = M r, and w = A u on the tile, so r, w and s are
// double-buffered by the control block's parity.  It is mathematically the
// same Krylov iteration as the FP64 form's PCG, with the same stopping rules.
//
// The FP64 residual replacement of the FP32-refined form (reliable updates)
// stays: every refine_every iterations and whenever the recursive residual
// claims convergence, one launch recomputes r = b - A (x_hi + x_lo) in FP64
// (RESID) and the next folds x_lo into x_hi and re-forms w = A M r and
// (gamma, delta) from the true residual (WCOMP); the recurrence then
// continues with p and s.  The first launch after the setup is a WCOMP.
//
// Every launch writes three partial sums: (gamma, delta, r.r) after ITER and
// WCOMP, (gamma, r.r, 0) after RESID, (r.z, r.r, b.b) after the setup.
#include "../hzo_internal.h"
#include "common.hip.h"
#include "tile.hip.h"
#include "reduce.hip.h"

namespace hzo {

struct Bcast32 {
    double alpha, beta;
    int flag;  // 0 work, 1 done, 2 residual replacement
};

__device__ __forceinline__ float row_Ax32(unsigned code, const HPair<float>* sHP, float ih2, float xc, float xe,
                                          float xw, float xn, float xs, float xu, float xd) {
    const unsigned cb = (code & 3u) << 4;
    const HPair<float> hx = sHP[cb | ((code >> 2) & 15u)];
    const HPair<float> hy = sHP[cb | ((code >> 6) & 15u)];
    const HPair<float> hz = sHP[cb | ((code >> 10) & 15u)];
    const float num = hx.p * (xe - xc) - hx.m * (xc - xw) + hy.p * (xn - xc) - hy.m * (xc - xs) +
                      hz.p * (xu - xc) - hz.m * (xc - xd);
    return -(num * ih2);
}

// r = b - A (x_hi + x_lo) in FP64 for one interior voxel (direct loads).
__device__ __forceinline__ double resid64(const CG32Bufs& b, const double* sH, const StepConsts& c, const Nbr& v,
                                          unsigned code, int k) {
    auto X = [&](long long q) { return b.xhi[q] + (double)b.xlo[q]; };
    double Ax;
    if (code & SC_IDENTITY) {
        Ax = op_identity(X(v.c), k, c);
    } else {
        const double* Hr = sH + (code & 3u) * 4u;
        const double xc = X(v.c);
        const double num = Hr[(code >> 2) & 3u] * (X(v.e) - xc) - Hr[(code >> 4) & 3u] * (xc - X(v.w)) +
                           Hr[(code >> 6) & 3u] * (X(v.n) - xc) - Hr[(code >> 8) & 3u] * (xc - X(v.s)) +
                           Hr[(code >> 10) & 3u] * (X(v.u) - xc) - Hr[(code >> 12) & 3u] * (xc - X(v.d));
        Ax = -(num * c.inv_h2);
    }
    return b.b[v.c] - Ax;
}

// The loop logic after the previous launch's totals (wave 0 only; every lane
// computes the same values).  Returns the work mode: 0 ITER, 1 done,
// 2 RESID, 3 WCOMP.  Sets alpha/beta for ITER.
__device__ __forceinline__ int cg_decide(PcgState& s, const double (&tot)[3], double& alpha, double& beta) {
    const int prev = s.kind;
    if (prev == K_SETUP || prev == K_RESID) {
        const double g = tot[0], rr = tot[1];
        if (prev == K_SETUP) {
            s.rz0 = g;
            s.rr0 = rr;
            s.bb = tot[2];
        } else {
            s.refinements += 1;
        }
        s.rr = rr;
        if (!finite(g) || g < 0.0) {
            s.status = prev == K_SETUP ? HZO_PCG_BAD_INITIAL : HZO_PCG_BAD_RZ;
            s.done = 1;
            return 1;
        }
        judge(s, g, rr);
        return s.done ? 1 : 3;
    }
    const double g = tot[0], dl = tot[1], rr = tot[2];
    s.rr = rr;
    if (!finite(g) || g < 0.0) {
        s.status = HZO_PCG_BAD_RZ;
        s.done = 1;
        return 1;
    }
    if (prev == K_ITER) {
        if (s.iters >= s.max_iter) {
            s.done = 1;
            s.status = HZO_PCG_MAXITER;
            return 1;
        }
        if (!keep_going(s, g, rr) || (s.refine_every > 0 && s.iters % s.refine_every == 0)) return 2;
    }
    // prev == K_WCOMP (fresh start when iters == 0) or K_ITER
    double den;
    if (prev == K_WCOMP && s.iters == 0) {
        beta = 0.0;
        den = dl;
    } else {
        beta = g / s.rz_old;
        den = dl - beta * g / s.alpha;
    }
    if (!finite(den) || !(den > 0.0)) {
        s.status = HZO_PCG_BREAKDOWN;
        s.done = 1;
        return 1;
    }
    alpha = g / den;
    s.alpha = alpha;
    s.beta = beta;
    s.denom = den;
    s.rz_old = g;
    return 0;
}

template <int KZ, int PM>
__device__ __forceinline__ void cg32_body(const CG32Bufs& b, const StepConsts& c, int nb, int ps, int t,
                                          const TileId& T) {
    const int ip = (t - 1) & 1, io = t & 1;
    PcgState s = load_state(b.st + ip);
    if (s.done) {
        pass_state(s, b.st + io, T.lead);
        return;
    }
    __shared__ HPair<float> sHP[64];
    __shared__ float pl[KZ][HY][HX];
    __shared__ RawLDS<3> sraw;
    __shared__ Bcast32 bc;
    const int nz = c.nz;
    const int tid = linear_tid();
    
    const int k0 = T.bz * KZ;
    const TilePos tp = tile_pos(c, T);
    const TileCols G = tile_cols(tp, c);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const unsigned pb4 = c.plane * 4u, pb2 = c.plane * 2u;
    const int cur = s.pbuf;
    const rsrc_t Rr = make_rsrc(cur ? b.r[1] : b.r[0], c.nvox * 4u);
    const rsrc_t Rw = make_rsrc(cur ? b.w[1] : b.w[0], c.nvox * 4u);
    const rsrc_t Rs = make_rsrc(cur ? b.s[1] : b.s[0], c.nvox * 4u);
    const rsrc_t Rsc = make_rsrc(b.sc, c.nvox * 2u);
    const rsrc_t Rdi = make_rsrc(b.lut32 + LUT_DINV, SC_LUT_SIZE * 4u);
    const bool after_true = s.kind == K_SETUP || s.kind == K_RESID;  // this launch re-forms w (WCOMP)
    const float ih2 = (float)c.inv_h2;
    const unsigned oob = c.nvox * 4u;

    const HPair<float> hpv = load_hpair(b.lut32);
    Partials<3, PM> P;
    load_partials<3, PM>(ip ? b.part[1] : b.part[0], nb, ps, P);
    unsigned so[KZ + 2];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) so[q] = (unsigned)min(max(k0 - 1 + q, 0), nz - 1);
    unsigned cc[KZ + 2], ch[KZ];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) cc[q] = bload_code(Rsc, G.cofs * 2u, so[q] * pb2);
#pragma unroll
    for (int q = 0; q < KZ; ++q) ch[q] = bload_code(Rsc, G.hofs * 2u, so[q + 1] * pb2);
    float rc[KZ + 2], rh[KZ];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) rc[q] = bload<float>(Rr, G.cofs * 4u, so[q] * pb4);
#pragma unroll
    for (int q = 0; q < KZ; ++q) rh[q] = bload<float>(Rr, G.hofs * 4u, so[q + 1] * pb4);
    float wc[KZ + 2], sc_[KZ + 2], wh[KZ], sh[KZ];
    const rsrc_t Rp = make_rsrc(b.p, c.nvox * 4u);
    const rsrc_t Rx = make_rsrc(b.xlo, c.nvox * 4u);
    if (!after_true) {
#pragma unroll
        for (int q = 0; q < KZ + 2; ++q) {
            wc[q] = bload<float>(Rw, G.cofs * 4u, so[q] * pb4);
            sc_[q] = bload<float>(Rs, G.cofs * 4u, so[q] * pb4);
        }
#pragma unroll
        for (int q = 0; q < KZ; ++q) {
            wh[q] = bload<float>(Rw, G.hofs * 4u, so[q + 1] * pb4);
            sh[q] = bload<float>(Rs, G.hofs * 4u, so[q + 1] * pb4);
        }
    }
    float dc[KZ + 2], dh[KZ];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) dc[q] = bload<float>(Rdi, cc[q] * 4u, 0u);
#pragma unroll
    for (int q = 0; q < KZ; ++q) dh[q] = bload<float>(Rdi, ch[q] * 4u, 0u);

    stage_partials<3, PM>(P, sraw);
    __syncthreads();
    if (tid < 64) {
        sHP[tid] = hpv;
        double tot[3];
        wave0_total<3>(sraw, tot);
        double alpha = 0.0, beta = 0.0;
        const int mode = cg_decide(s, tot, alpha, beta);
        if (mode == 1) {
            pass_state(s, b.st + io, T.lead);
        } else if (mode == 2) {
            s.kind = K_RESID;
        } else if (mode == 3) {
            s.kind = K_WCOMP;
        } else {
            s.iters += 1;
            s.pbuf ^= 1;
            s.kind = K_ITER;
        }
        if (tid == 0) {
            bc.alpha = alpha;
            bc.beta = beta;
            bc.flag = mode;
        }
    }
    __syncthreads();
    const int mode = bc.flag;
    if (mode == 1) return;
    double acc[3] = {0.0, 0.0, 0.0};
    if (mode == 2) {
        // RESID: r = b - A (x_hi + x_lo) in FP64 over the tile (direct loads)
        __shared__ double sH64[16];
        load_lut(sH64, b.lut);
        float* rout = cur ? b.r[1] : b.r[0];
        if (tp.live) {
            for (int k = k0; k < min(k0 + KZ, nz); ++k) {
                const Nbr v = op_nbrs(tp.i, tp.j, k, c);
                const unsigned code = b.sc[v.c];
                const float r = (float)resid64(b, sH64, c, v, code, k);
                rout[v.c] = r;
                const float z = r * b.lut32[LUT_DINV + code];
                acc[0] += (double)r * (double)z;
                acc[1] += (double)r * (double)r;
            }
        }
    } else if (mode == 3) {
        // WCOMP: fold x_lo into x_hi, u = M r, w = A u, (gamma, delta, r.r)
        if (tp.live) {
#pragma unroll
            for (int q = 0; q < KZ; ++q) {
                const int k = k0 + q;
                if (k < nz) {
                    const long long id = (long long)G.cofs + (long long)c.plane * k;
                    b.xhi[id] += (double)b.xlo[id];
                    b.xlo[id] = 0.0f;
                }
            }
        }
        float uc[KZ + 2];
#pragma unroll
        for (int q = 0; q < KZ + 2; ++q) {
            uc[q] = rc[q] * dc[q];
            if (q >= 1 && q <= KZ) pl[q - 1][ty + 1][tx + 1] = uc[q];
        }
#pragma unroll
        for (int q = 0; q < KZ; ++q)
            if (tp.has_h) pl[q][tp.hy][tp.hx] = rh[q] * dh[q];
        __syncthreads();
#pragma unroll
        for (int q = 1; q <= KZ; ++q) {
            const int k = k0 - 1 + q;
            const float xc = uc[q];
            const unsigned code = cc[q];
            float xu = uc[q + 1], xd = uc[q - 1];
            if (c.zbc != 0) {
                if (k == 0) xd = xu;
                else if (k == nz - 1) xu = xd;
            }
            const float qs = row_Ax32(code, sHP, ih2, xc, pl[q - 1][ty + 1][tx + 2], pl[q - 1][ty + 1][tx],
                                      pl[q - 1][ty + 2][tx + 1], pl[q - 1][ty][tx + 1], xu, xd);
            const float wv = (code & SC_IDENTITY) ? op_identity(xc, k, c) : qs;
            const bool ok = tp.live && k < nz;
            bstore(Rw, ok ? G.cofs * 4u : oob, so[q] * pb4, wv);
            if (ok) {
                acc[0] += (double)rc[q] * (double)xc;
                acc[1] += (double)wv * (double)xc;
                acc[2] += (double)rc[q] * (double)rc[q];
            }
        }
    } else {
        // ITER: s = w + beta s, r -= alpha s on tile + halo; p = M r_old + beta p, x += alpha p;
        // u = M r, w = A u; (gamma, delta, r.r)
        const float alpha = (float)bc.alpha, beta = (float)bc.beta;
        const bool fresh = bc.beta == 0.0;  // first iteration of a solve: no previous p, s
        const rsrc_t Rr2 = make_rsrc(cur ? b.r[0] : b.r[1], c.nvox * 4u);
        const rsrc_t Rw2 = make_rsrc(cur ? b.w[0] : b.w[1], c.nvox * 4u);
        const rsrc_t Rs2 = make_rsrc(cur ? b.s[0] : b.s[1], c.nvox * 4u);
        float uc[KZ + 2], uo[KZ];
#pragma unroll
        for (int q = 0; q < KZ + 2; ++q) {
            const int k = k0 - 1 + q;
            const float sn = fresh ? wc[q] : wc[q] + beta * sc_[q];
            const float rn = rc[q] - alpha * sn;
            uc[q] = rn * dc[q];
            if (q >= 1 && q <= KZ) {
                uo[q - 1] = rc[q] * dc[q];
                pl[q - 1][ty + 1][tx + 1] = uc[q];
                const bool ok = tp.live && k < nz;
                const unsigned vo = ok ? G.cofs * 4u : oob;
                bstore(Rs2, vo, so[q] * pb4, sn);
                bstore(Rr2, vo, so[q] * pb4, rn);
                if (ok) {
                    acc[0] += (double)rn * (double)uc[q];
                    acc[2] += (double)rn * (double)rn;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < KZ; ++q) {
            const float shn = fresh ? wh[q] : wh[q] + beta * sh[q];
            const float rhn = rh[q] - alpha * shn;
            if (tp.has_h) pl[q][tp.hy][tp.hx] = rhn * dh[q];
        }
        __syncthreads();
        // p and x are needed only by the interior update after the stencil:
        // their loads are issued here and land while w = A u is formed
        float pv[KZ], xv[KZ];
#pragma unroll
        for (int q = 0; q < KZ; ++q) {
            pv[q] = bload<float>(Rp, G.cofs * 4u, so[q + 1] * pb4);
            xv[q] = bload<float>(Rx, G.cofs * 4u, so[q + 1] * pb4);
        }
#pragma unroll
        for (int q = 1; q <= KZ; ++q) {
            const int k = k0 - 1 + q;
            const float xc = uc[q];
            const unsigned code = cc[q];
            float xu = uc[q + 1], xd = uc[q - 1];
            if (c.zbc != 0) {
                if (k == 0) xd = xu;
                else if (k == nz - 1) xu = xd;
            }
            const float qs = row_Ax32(code, sHP, ih2, xc, pl[q - 1][ty + 1][tx + 2], pl[q - 1][ty + 1][tx],
                                      pl[q - 1][ty + 2][tx + 1], pl[q - 1][ty][tx + 1], xu, xd);
            const float wv = (code & SC_IDENTITY) ? op_identity(xc, k, c) : qs;
            const bool ok = tp.live && k < nz;
            bstore(Rw2, ok ? G.cofs * 4u : oob, so[q] * pb4, wv);
            if (ok) acc[1] += (double)wv * (double)xc;
        }
#pragma unroll
        for (int q = 1; q <= KZ; ++q) {
            const int k = k0 - 1 + q;
            const float pn = fresh ? uo[q - 1] : uo[q - 1] + beta * pv[q - 1];
            const float xn = xv[q - 1] + alpha * pn;
            const bool ok = tp.live && k < nz;
            const unsigned vo = ok ? G.cofs * 4u : oob;
            bstore(Rp, vo, so[q] * pb4, pn);
            bstore(Rx, vo, so[q] * pb4, xn);
        }
    }
    write_block_partials<3>(acc, io ? b.part[1] : b.part[0], ps, T.blk, sraw);
    pass_state(s, b.st + io, T.lead);
}

template <int KZ, int PM>
__global__ __launch_bounds__(NT, KZ >= 16 ? 2 : 4) void k_cg32(CG32Bufs b, StepConsts c, int nb, int ps, int t) {
    cg32_body<KZ, PM>(b, c, nb, ps, t, tile_id());
}

// All films of a batch in one launch: the tile table maps each workgroup to
// (film, logical tile); every film runs its own control block, and the tiles
// of a film whose solve has finished return at once.
template <int KZ, int PM>
__global__ __launch_bounds__(NT, KZ >= 16 ? 2 : 4) void k_cg32_batch(const BatchFilm* __restrict__ films, const int2* __restrict__ tiles,
                                                                      int t) {
    TileId T;
    const BatchFilm& F = films[batch_tile(tiles, films, T)];
    // descriptor fields read once, before any store (scalar loads)
    const CG32Bufs b = F.cg;
    const StepConsts c = F.c;
    cg32_body<KZ, PM>(b, c, F.nb, F.ps, t, T);
}

void launch_cg32_batch(hipStream_t s, int ntiles, int kz, int pm, const BatchFilm* films, const int2* tiles,
                       int t) {
    const dim3 gr(ntiles), bl(BX, BY, 1);
#define HZO_CGB(KZ)                                                                                   \
    switch (pm) {                                                                                     \
        case 1: hipLaunchKernelGGL((k_cg32_batch<KZ, 1>), gr, bl, 0, s, films, tiles, t); break;       \
        case 2: hipLaunchKernelGGL((k_cg32_batch<KZ, 2>), gr, bl, 0, s, films, tiles, t); break;       \
        case 4: hipLaunchKernelGGL((k_cg32_batch<KZ, 4>), gr, bl, 0, s, films, tiles, t); break;       \
        default: hipLaunchKernelGGL((k_cg32_batch<KZ, 0>), gr, bl, 0, s, films, tiles, t); break;      \
    }
    if (kz == 2) {
        HZO_CGB(2)
    } else {
        HZO_CGB(4)
    }
#undef HZO_CGB
}

#define HZO_X
void launch_cg32(hipStream_t s, const Grid& g, const StepConsts& c, const CG32Bufs& b, int t,
                 const LaunchEv& ev) {
    if (g.kz == 16) {  // one workgroup per CU, up to 256 VGPRs: the whole film in one round
        const int nb = blocks_for_impl(g), pm = part_pm(nb);
        const dim3 gr = grid3(g), bl(BX, BY, 1);
        HZO_PCG_LAUNCH_PM(k_cg32, 16, pm, gr, bl, s, b, c, nb, part_stride_impl(nb), t);
        return;
    }
    HZO_PCG_LAUNCH_X(k_cg32, g, s, b, c, blocks_for_impl(g), part_stride_impl(blocks_for_impl(g)), t);
}
#undef HZO_X

}  // namespace hzo

namespace hzo {

// Batch bookkeeping kernels (one small launch each).
__global__ void k_batch_set_ez(BatchFilm* films, const double* ez, int n) {
    for (int f = threadIdx.x; f < n; f += blockDim.x) films[f].Ez = (float)ez[f];
}

// info[4f..4f+3] = done flag, iterations, status and residual replacements of
// film f's control block after launch t.
__global__ void k_batch_done(const BatchFilm* films, int n, int t, int* info) {
    for (int f = threadIdx.x; f < n; f += blockDim.x) {
        const PcgState& st = films[f].cg.st[t & 1];
        info[4 * f + 0] = st.done;
        info[4 * f + 1] = st.iters;
        info[4 * f + 2] = st.status;
        info[4 * f + 3] = st.refinements;
    }
}

// m[f] = sum of the first diagnostics column of film f (one wave per film).
__global__ void k_batch_sum_pmz(const BatchFilm* films, double* m) {
    const BatchFilm& F = films[blockIdx.x];
    const int lane = threadIdx.x;
    double s = 0.0;
    for (int t = lane; t < F.nb; t += 64) s += F.diag_part[t];
    s = wave_sum(s);
    if (lane == 0) m[blockIdx.x] = s;
}

void launch_batch_set_ez(hipStream_t s, BatchFilm* films, const double* ez, int n) {
    hipLaunchKernelGGL(k_batch_set_ez, dim3(1), dim3(256), 0, s, films, ez, n);
}
void launch_batch_done(hipStream_t s, const BatchFilm* films, int n, int t, int* info) {
    hipLaunchKernelGGL(k_batch_done, dim3(1), dim3(256), 0, s, films, n, t, info);
}
void launch_batch_sum_pmz(hipStream_t s, const BatchFilm* films, int n, double* m) {
    hipLaunchKernelGGL(k_batch_sum_pmz, dim3(n), dim3(64), 0, s, films, m);
}

}  // namespace hzo
