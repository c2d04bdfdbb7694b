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
//
// Two kernels run this iteration: k_cg32 carries w and s in memory (42
// bytes per voxel and iteration), k_cg32r forms s = A p and w = A u by two
// stencils in the launch and carries only r and p (26 bytes).  The host
// picks by film size; both share the control logic and the RESID mode.
#include "../hzo_internal.h"
#include "common.hip.h"
#include "tile.hip.h"
#include "reduce.hip.h"
#include "cg_ctl.hip.h"

namespace hzo {


// acc += w x y in FP32 over a thread's planes (the launch's sums across
// threads are FP64), with the mirrored-plane weight w = 1/2 taken as
// (2 acc + x y) / 2 (the terms of cg_dot_addf), chosen by a uniform branch
// instead of multiplications by 1 on every other plane.
__device__ __forceinline__ void dot_w(float& acc, float x, float y, bool mirrored) {
    if (mirrored)
        acc = __builtin_fmaf(x, y, acc * 2.0f) * 0.5f;
    else
        acc = __builtin_fmaf(x, y, acc);
}

// How an iteration updates x_lo (from the control block before the launch's
// decision): 0 leaves its alpha p to the next iteration, 1 adds alpha p,
// 2 adds the deferred term first, (x_lo + alpha_prev p_prev) + alpha p, the
// same two roundings as two separate updates.  Iterations defer in turn, so
// half of them neither read nor write x_lo; the one before a periodic
// residual replacement never defers, so the replacement sees all of x.
__device__ __forceinline__ int cg_x_mode(const PcgState& s) {
    if (s.xpend) return 2;
    return (s.refine_every > 0 && multiple_of(s.iters + 1, s.refine_every)) ? 1 : 0;
}

// The control-block update after cg_decide.  A solve that stops, or replaces
// its residual, while an x_lo term is deferred first spends one launch
// (mode 4, XADD) adding it; after a replacement that launch hands on to the
// RESID launch, after a stop it keeps `done`.
__device__ __forceinline__ int cg_advance(PcgState& s, int mode, int xm, double alpha) {
    if ((mode == 1 || mode == 2) && s.xpend) {
        s.kind = K_XADD;
        s.xpend = 0;
        s.xadds += 1;
        return 4;
    }
    if (mode == 2) {
        s.kind = K_RESID;
    } else if (mode == 3) {
        s.kind = K_WCOMP;
    } else if (mode == 0) {
        s.iters += 1;
        s.pbuf ^= 1;
        s.kind = K_ITER;
        s.xpend = xm == 0;
        if (xm == 0) s.alpha_prev = alpha;
    }
    return mode;
}

// One plane of a tile as the z-streaming loop holds it: the centre column's
// r, w, s (w and s only for an iteration), D^-1 and, on interior planes, p
// and x_lo; the halo column's r, w, s and D^-1.
struct Plane32 {
    float r, w, s, d, p, x;
    float rh, wh, sh, dh;
};

// RESID: r = b - A (x_hi + x_lo) in FP64 over the tile's planes, written as
// FP32 to the r array at byte offset o_r of RA; partials (r.z, r.r).  Each
// plane's potential is staged in LDS with the x-y halo (double-buffered, one
// barrier per plane), the column's planes below and above stay in registers
// (z neighbours, clamped into the film; identity rows, where the clamp
// applies under Dirichlet electrodes, use only the centre), and the next
// plane's values are loaded while the current one is evaluated.  Lanes of
// an edge tile beyond the film stage the periodic image and store nothing.
template <int KZ, bool MIR>
__device__ __forceinline__ void cg32_resid(const CG32Bufs& b, const StepConsts& c, const TilePos& tp,
                                           const TileCols& G, int k0, rsrc_t RA, unsigned o_r, rsrc_t Rsc,
                                           rsrc_t Rdi, double (&acc)[3]) {
    const int nz = c.nz;
    const int tx = threadIdx.x, ty = threadIdx.y;
    __shared__ double sH64[16];
    __shared__ double sx[2][HY][HX];
    load_lut(sH64, b.lut);
    const rsrc_t Rxh = make_rsrc(b.xhi, c.nvox * 8u);
    const rsrc_t Rxl = make_rsrc(b.xlo, c.nvox * 4u);
    const rsrc_t Rb = make_rsrc(b.b, c.nvox * 8u);
    // halo lanes stage their cell; the others load at offset N (answered
    // with 0, never staged)
    const unsigned oc = G.cofs, oh = G.hofs, ost = tp.live ? G.cofs : 4u * c.nvox;
    auto X = [&](unsigned col, int k) {
        const unsigned pofs = (unsigned)k * c.plane;
        return bload<double>(Rxh, col * 8u, pofs * 8u) + (double)bload<float>(Rxl, col * 4u, pofs * 4u);
    };
    auto B = [&](int k) { return bload<double>(Rb, oc * 8u, (unsigned)k * c.plane * 8u); };
    auto CODE = [&](int k) { return bload_code(Rsc, oc * 2u, (unsigned)k * c.plane * 2u); };
    const int k1 = min(k0 + KZ, nz);
    double xm = X(oc, max(k0 - 1, 0)), x0 = X(oc, k0), xp = X(oc, min(k0 + 1, nz - 1)), xh = X(oh, k0);
    double bv = B(k0);
    unsigned code = CODE(k0);
#pragma unroll 1
    for (int k = k0; k < k1; ++k) {
        const int sb = (k - k0) & 1;
        sx[sb][ty + 1][tx + 1] = x0;
        if (tp.has_h) sx[sb][tp.hy][tp.hx] = xh;
        const int kn = min(k + 1, k1 - 1);
        // (the last plane reloads the ghost plane above the tile, unused)
        const double xpp = X(oc, min(k + 2, min(k1, nz - 1))), xhn = X(oh, kn), bn = B(kn);
        const unsigned coden = CODE(kn);
        __syncthreads();
        double xu = xp, xd = xm;
        if (c.zbc != 0) {  // mirror: planes 1 and nz - 2 stand on both sides of the faces
            if (k == 0) xd = xu;
            else if (k == nz - 1) xu = xd;
        }
        double Ax;
        if (code & SC_IDENTITY) {
            Ax = op_identity(x0, k, c);
        } else {
            const double(*S)[HX] = sx[sb];
            const double* Hr = sH64 + (code & 3u) * 4u;
            const double num = Hr[(code >> 2) & 3u] * (S[ty + 1][tx + 2] - x0) -
                               Hr[(code >> 4) & 3u] * (x0 - S[ty + 1][tx]) +
                               Hr[(code >> 6) & 3u] * (S[ty + 2][tx + 1] - x0) -
                               Hr[(code >> 8) & 3u] * (x0 - S[ty][tx + 1]) + Hr[(code >> 10) & 3u] * (xu - x0) -
                               Hr[(code >> 12) & 3u] * (x0 - xd);
            Ax = -(num * c.inv_h2);
        }
        const float r = (float)(bv - Ax);
        bstore(RA, ost * 4u, o_r + (unsigned)k * c.plane * 4u, r);
        if (tp.live) {
            const float z = r * bload<float>(Rdi, code * 4u, 0u);
            cg_dot_add<MIR>(acc[0], r, z, c, k);
            acc[1] += (double)r * (double)r;
        }
        xm = x0;
        x0 = xp;
        xp = xpp;
        xh = xhn;
        bv = bn;
        code = coden;
    }
}

template <int KZ, int PM, bool MIR>
__device__ __forceinline__ void cg32_body(const CG32Bufs& b, const StepConsts& c, int nb, int ps, int t,
                                          const TileId& T) {
    static_assert(KZ % 2 == 0, "the plane loop closes its stencil pairs on an odd last plane");
    const int ip = (t - 1) & 1, io = t & 1;
    __shared__ HPair<float> sHP[64];
    // u of plane q (centre + x-y halo) in buffer q % 6: the stencils of two
    // planes follow each barrier, and a buffer is rewritten only after the
    // barrier that ends its last reads
    __shared__ float pl[6][HY][HX];
    __shared__ WaveLDS<3> sred, spart;  // previous launch's totals; this launch's partials
    const int nz = c.nz;
    const int tid = linear_tid();

    // Loads that do not depend on the control block go out first, so that
    // their latency overlaps the control-block load: the face-pair table,
    // the previous launch's partial sums and the stencil codes.
    const int k0 = T.bz * KZ;
    const TilePos tp = tile_pos(c, T);
    const TileCols G = tile_cols(tp, c);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const unsigned pb4 = c.plane * 4u, pb2 = c.plane * 2u;
    const rsrc_t Rsc = make_rsrc(b.sc, c.nvox * 2u);
    const rsrc_t Rdi = make_rsrc(b.lut32 + LUT_DINV, SC_LUT_SIZE * 4u);
    const float ih2 = (float)c.inv_h2;
    const HPair<float> hpv = load_hpair(b.lut32);
    Partials<3, PM> P;
    load_partials<3, PM>(ip ? b.pin[1] : b.pin[0], b.nb_in, b.ps_in, P);
    // plane q of the tile's window is z = so[q]: q = 0 and KZ + 1 are the
    // ghost planes below and above (clamped at the film's faces)
    unsigned so[KZ + 2];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) so[q] = (unsigned)min(max(k0 - 1 + q, 0), nz - 1);
    unsigned cc[KZ + 2], ch[KZ + 2];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) {
        cc[q] = bload_code(Rsc, G.cofs * 2u, so[q] * pb2);
        ch[q] = (q >= 1 && q <= KZ) ? bload_code(Rsc, G.hofs * 2u, so[q] * pb2) : 0u;
    }

    PcgState s = load_state(b.st + ip);
    if (s.done) {
        // the values are consumed here too, so the loads above stay ahead of
        // the branch instead of being sunk behind the control-block wait
        keep_live(hpv.p);
#pragma unroll
        for (int q = 0; q < KZ + 2; ++q) keep_live(cc[q]);
        keep_live_partials(P);
        pass_state(s, b.st + io, T.lead);
        return;
    }
    // two resources cover seven vectors: RA = r[0] r[1] w[0] w[1], RB =
    // s[0] s[1] p; an array is chosen by a workgroup-uniform byte offset.
    // x_lo has its own (an adjoint solve brings its own solution vector).
    // `aend` (one past an arena) is out of range of every resource here.
    const unsigned cur = (unsigned)s.pbuf;
    const unsigned seg = c.nvox * 4u, aend = 4u * seg;
    const rsrc_t RA = make_rsrc(b.va, aend);
    const rsrc_t RB = make_rsrc(b.vb, aend);
    const rsrc_t RX = make_rsrc(b.xlo, seg);
    const unsigned o_r = cur * seg, o_w = (2u + cur) * seg, o_s = cur * seg;
    const unsigned o_p = 2u * seg;
    const unsigned cv = G.cofs * 4u, hv = tp.has_h ? G.hofs * 4u : aend;
    // an iteration follows unless the last launch set up or replaced the
    // residual (then this launch re-forms w and needs neither w nor s)
    const bool after_true = s.kind == K_SETUP || s.kind == K_RESID;
    // x_lo is read and written only by iterations that update it; the lanes
    // of the others address it out of range (no access and no branch)
    const int xm = cg_x_mode(s);
    const float aprev = (float)s.alpha_prev;
    const unsigned xo = (after_true || xm == 0) ? aend : cv;

    // Plane loads, two planes ahead of their use.
    Plane32 L[KZ + 2];
    auto issue = [&](int q) {
        const unsigned sv = so[q] * pb4;
        L[q].r = bload<float>(RA, cv, o_r + sv);
        L[q].d = bload<float>(Rdi, cc[q] * 4u, 0u);
        L[q].w = after_true ? 0.0f : bload<float>(RA, cv, o_w + sv);
        L[q].s = after_true ? 0.0f : bload<float>(RB, cv, o_s + sv);
        if (q >= 1 && q <= KZ) {
            L[q].p = after_true ? 0.0f : bload<float>(RB, cv, o_p + sv);
            L[q].x = bload<float>(RX, xo, sv);
            L[q].rh = bload<float>(RA, hv, o_r + sv);
            L[q].dh = bload<float>(Rdi, ch[q] * 4u, 0u);
            L[q].wh = after_true ? 0.0f : bload<float>(RA, hv, o_w + sv);
            L[q].sh = after_true ? 0.0f : bload<float>(RB, hv, o_s + sv);
        }
    };
    issue(0);
    issue(1);

    // The previous launch's totals: wave sums through LDS, then every wave
    // adds the 8 slots and runs the (deterministic) loop logic itself, so no
    // broadcast barrier follows.
    if (tid < 64) sHP[tid] = hpv;
    wave_partials<3, PM>(P, sred);
    __syncthreads();
    double tot[3];
    wave_lds_total<3>(sred, tot);
    double alpha_d = 0.0, beta_d = 0.0;
    // every lane holds the same decision; made wave-uniform, the branches on
    // it are scalar and alpha, beta sit in scalar registers
    const int mode = __builtin_amdgcn_readfirstlane(cg_advance(s, cg_decide(s, tot, alpha_d, beta_d), xm, alpha_d));
    // handed on at once: nothing later in the launch reads the block, so it
    // need not stay live in registers through the tile work
    pass_state(s, b.st + io, T.lead);
    if (mode == 1) return;
    double acc[3] = {0.0, 0.0, 0.0};
    if (mode == 2) {
        cg32_resid<KZ, MIR>(b, c, tp, G, k0, RA, o_r, Rsc, Rdi, acc);
    } else if (mode == 4) {
        // XADD: the deferred x_lo += alpha_prev p of the last iteration
#pragma unroll
        for (int q = 1; q <= KZ; ++q) {
            const int k = k0 - 1 + q;
            const unsigned sv = so[q] * pb4;
            const float xv = bload<float>(RX, cv, sv), pv = bload<float>(RB, cv, o_p + sv);
            bstore(RX, (tp.live && k < nz) ? cv : aend, sv, xv + aprev * pv);
        }
    } else {
        // ITER (mode 0): s = w + beta s and r -= alpha s on the tile and its
        // halo, p = M r_old + beta p, x += alpha p, then u = M r and w = A u;
        // partials (gamma, delta, r.r).  WCOMP (mode 3): x_lo folded into
        // x_hi, u = M r, w = A u; partials (r.u, w.u, r.r).  The planes are
        // streamed bottom to top: plane q's update and u, then the stencil of
        // plane q - 1, whose z neighbours are u of planes q - 2 and q.
        const float alpha = uniform_f((float)alpha_d), beta = uniform_f((float)beta_d);
        const bool fresh = __builtin_amdgcn_readfirstlane(beta_d == 0.0 ? 1 : 0) != 0;  // first iteration: no previous p, s
        const unsigned o_r2 = (cur ^ 1u) * seg, o_s2 = (cur ^ 1u) * seg;
        float af[3] = {0.0f, 0.0f, 0.0f};  // the thread's sums over its planes (FP32)
        // The plane loop, built twice: an iteration (ITER) and the
        // re-formation after a residual replacement (WCOMP), so neither
        // carries the other's selects.
        auto planes = [&](auto iter_c) {
            constexpr bool iter = decltype(iter_c)::value;
            // w goes to the other buffer after an iteration, in place after WCOMP
            const unsigned o_wo = iter ? (3u - cur) * seg : o_w;
            // planes that can hold identity rows: the electrode planes under
            // Dirichlet electrodes, the gauge node's plane under Neumann
            auto ident_plane = [&](int kk) { return c.zbc == 0 ? (kk == 0 || kk == nz - 1) : (kk == nz / 2); };
            auto mirrored = [&](int kk) { return MIR && c.zbc != 0 && (kk == 0 || kk == nz - 1); };
            float u[KZ + 2];
#pragma unroll
            for (int q = 0; q < KZ + 2; ++q) {
                if (q + 2 < KZ + 2) issue(q + 2);
                const int k = k0 - 1 + q;
                const bool interior = q >= 1 && q <= KZ;
                const bool ok = interior && tp.live && k < nz;
                const unsigned vo = ok ? cv : aend;
                const unsigned sv = so[q] * pb4;
                // A(q): update and u of plane q
                float rn = L[q].r;
                if constexpr (iter) {
                    const float sn = fresh ? L[q].w : L[q].w + beta * L[q].s;
                    rn = L[q].r - alpha * sn;
                    if (interior) {
                        bstore(RB, vo, o_s2 + sv, sn);
                        bstore(RA, vo, o_r2 + sv, rn);
                        const float uo = L[q].r * L[q].d;
                        const float pn = fresh ? uo : uo + beta * L[q].p;
                        bstore(RB, vo, o_p + sv, pn);
                        const float xa = xm == 2 ? L[q].x + aprev * L[q].p : L[q].x;
                        bstore(RX, xm != 0 ? vo : aend, sv, xa + alpha * pn);
                    }
                } else if (interior && ok) {
                    const long long id = (long long)G.cofs + (long long)c.plane * k;
                    b.xhi[id] += (double)b.xlo[id];
                    b.xlo[id] = 0.0f;
                }
                u[q] = rn * L[q].d;
                if (ok) {
                    dot_w(af[0], rn, u[q], mirrored(k));
                    af[2] = __builtin_fmaf(rn, rn, af[2]);
                }
                if (interior) {
                    pl[q % 6][ty + 1][tx + 1] = u[q];
                    float rhn = L[q].rh;
                    if constexpr (iter) {
                        const float shn = fresh ? L[q].wh : L[q].wh + beta * L[q].sh;
                        rhn = L[q].rh - alpha * shn;
                    }
                    if (tp.has_h) pl[q % 6][tp.hy][tp.hx] = rhn * L[q].dh;
                }
                // B(q - 2), B(q - 1): stencils of planes q - 2 and q - 1 after
                // every second plane (KZ is even, so the last plane q = KZ + 1 is
                // odd and closes the loop)
                if (q >= 3 && (q & 1)) {
                    __syncthreads();
#pragma unroll
                    for (int m = q - 2; m <= q - 1; ++m) {
                        const int km = k0 - 1 + m;
                        const float xc = u[m];
                        const unsigned code = cc[m];
                        float xu = u[m + 1], xd = u[m - 1];
                        if (c.zbc != 0) {
                            if (km == 0) xd = xu;
                            else if (km == nz - 1) xu = xd;
                        }
                        const float(*pm)[HX] = pl[m % 6];
                        const float qs = row_Ax32(code, sHP, ih2, xc, pm[ty + 1][tx + 2], pm[ty + 1][tx],
                                                  pm[ty + 2][tx + 1], pm[ty][tx + 1], xu, xd);
                        float wv = qs;
                        if (ident_plane(km) && (code & SC_IDENTITY)) wv = op_identity(xc, km, c);
                        const bool okm = tp.live && km < nz;
                        bstore(RA, okm ? cv : aend, o_wo + so[m] * pb4, wv);
                        if (okm) dot_w(af[1], wv, xc, mirrored(km));
                    }
                }
            }
        };
        if (mode == 0)
            planes(std::true_type{});
        else
            planes(std::false_type{});
#pragma unroll
        for (int qq = 0; qq < 3; ++qq) acc[qq] = af[qq];
    }
    write_block_partials_w<3>(acc, io ? b.part[1] : b.part[0], ps, T.blk, spart);
}

// ---------------------------------------------------------------- two-stencil iteration
// The same single-reduction iteration with s = A p and w = A u formed in the
// launch instead of carried in memory: a launch reads r_i and p_(i-1) and
// writes r_(i+1) and p_i (both double-buffered by the control block's
// parity) and x_lo, 26 bytes per voxel instead of 42.  In exact arithmetic
// s_i = A p_i is the vector the recurrence s = w + beta s carries, so the
// iterates are those of the same Krylov method; the stencil is applied
// twice per voxel instead of once, on a wider halo.
//   p_i = M r_i + beta p_(i-1)   tile and a two-voxel x-y ring
//   s_i = A p_i                   tile and its one-voxel ring
//   r_(i+1) = r_i - alpha s_i, u = M r_(i+1)   same cells
//   w = A u                       tile; x_lo += alpha p_i on the tile
// Partials (gamma, delta, r.r) as in the carried form; RESID and WCOMP are
// the same launch modes (WCOMP forms u and w from the true residual).
//
// Ring cells: thread h < R1 owns a cell of the one-voxel ring without its
// corners (rows above and below the tile, then the left and right columns),
// R1 <= h < R2 one of the four corners or of the outer ring's rows and
// columns (the outer ring's corners are never read).  A thread keeps its own
// column and ring cell on every plane, so z neighbours come from registers
// and only in-plane neighbours go through LDS.
constexpr int R1 = 2 * TX + 2 * TY;           // 144
constexpr int R2 = R1 + 4 + 2 * TX + 2 * TY;  // 292
constexpr int PX = TX + 4, PY = TY + 4;

struct Ring2 {
    bool has, inner;
    int lx, ly;  // position relative to the tile's first voxel: -2 .. TX + 1, -2 .. TY + 1
};

__device__ __forceinline__ Ring2 ring2_cell(int h) {
    Ring2 r;
    r.has = h < R2;
    r.inner = h < R1;
    int x = 0, y = 0;
    if (h < TX) {
        x = h, y = -1;
    } else if (h < 2 * TX) {
        x = h - TX, y = TY;
    } else if (h < 2 * TX + TY) {
        x = -1, y = h - 2 * TX;
    } else if (h < R1) {
        x = TX, y = h - 2 * TX - TY;
    } else if (h < R1 + 4) {
        x = ((h - R1) & 1) ? TX : -1, y = ((h - R1) & 2) ? TY : -1;
    } else if (h < R1 + 4 + TX) {
        x = h - R1 - 4, y = -2;
    } else if (h < R1 + 4 + 2 * TX) {
        x = h - R1 - 4 - TX, y = TY + 1;
    } else if (h < R1 + 4 + 2 * TX + TY) {
        x = -2, y = h - R1 - 4 - 2 * TX;
    } else if (h < R2) {
        x = TX + 1, y = h - R1 - 4 - 2 * TX - TY;
    }
    r.lx = x;
    r.ly = y;
    return r;
}

// One window plane: the centre column's r, p_(i-1), x_lo and D^-1, and the
// ring cell's r, p_(i-1) and D^-1.
struct Plane2 {
    float r, p, x, d;
    float rh, ph, dh;
};

// The mirrored-plane weights of Neumann films are applied at run time (a
// weight of exactly 1 elsewhere leaves every sum unchanged); a build without
// them allocates more registers at deep tiles and runs slower.
template <int KZ, int PM, bool NTS>
__device__ __forceinline__ void cg32r_body(const CG32Bufs& b, const StepConsts& c, int nb, int ps, int t,
                                           const TileId& T) {
    constexpr int NW = KZ + 4;  // window plane m is z = k0 - 2 + m; the tile's planes are m = 2 .. KZ + 1
    const int ip = (t - 1) & 1, io = t & 1;
    __shared__ HPair<float> sHP[64];
    // p of plane m (tile + two-voxel ring) in sp[m % 2], u (tile + one-voxel
    // ring) in su[m % 2]: each is written in one phase of the plane loop and
    // read only in the next
    __shared__ float sp[2][PY][PX];
    __shared__ float su[2][HY][HX];
    __shared__ WaveLDS<3> sred, spart;
    const int nz = c.nz;
    const int tid = linear_tid();
    const int k0 = T.bz * KZ;
    const TilePos tp = tile_pos(c, T);
    const TileCols G = tile_cols(tp, c);
    const Ring2 R = ring2_cell(tid);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const unsigned rofs = R.has ? (unsigned)wrapi(T.bx * TX + R.lx, c.nx) +
                                      (unsigned)c.nx * (unsigned)wrapi(T.by * TY + R.ly, c.ny)
                                : c.nvox;
    // the ring stencil of a lane without a one-voxel-ring cell runs on its
    // own column (in range, result unused)
    const int bx = R.inner ? R.lx : tx, by = R.inner ? R.ly : ty;
    // A wave is one tile row (64 x 8 threads); the one-voxel ring's cells
    // sit on waves 0-2, so the ring stencil is skipped by the other waves
    // (a scalar branch with no loads inside).
    const bool winner = __builtin_amdgcn_readfirstlane(ty) * TX < R1;
    const unsigned pb4 = c.plane * 4u, pb2 = c.plane * 2u;
    const rsrc_t Rsc = make_rsrc(b.sc, c.nvox * 2u);
    const rsrc_t Rdi = make_rsrc(b.lut32 + LUT_DINV, SC_LUT_SIZE * 4u);
    const float ih2 = (float)c.inv_h2;
    const HPair<float> hpv = load_hpair(b.lut32);
    Partials<3, PM> P;
    load_partials<3, PM>(ip ? b.pin[1] : b.pin[0], b.nb_in, b.ps_in, P);
    // Plane z of window plane m, clamped into the film.  It is formed from a
    // copy of k0 that each phase of the plane loop renews (an empty asm that
    // claims to change it), so each plane offset is computed where it is
    // used: formed all at once, the 36 offsets of a depth-32 window and their
    // products were held in scalar registers, which spilled (125 of them at
    // depth 32) and were read back with a VALU instruction per use.
    int k0p = k0;
    auto so = [&](int m) { return (unsigned)min(max(k0p - 2 + m, 0), nz - 1); };
    // stencil codes (centre | ring << 16), each loaded two phases before its
    // plane's data, whose D^-1 gathers need it
    unsigned cd[NW];
    auto load_code = [&](int m) {
        cd[m] = bload_code(Rsc, G.cofs * 2u, so(m) * pb2) | (bload_code(Rsc, rofs * 2u, so(m) * pb2) << 16);
    };
#pragma unroll
    for (int m = 0; m < 4; ++m) load_code(m);

    PcgState s = load_state(b.st + ip);
    if (s.done) {
        keep_live(hpv.p);
#pragma unroll
        for (int m = 0; m < 4; ++m) keep_live(cd[m]);
        keep_live_partials(P);
        pass_state(s, b.st + io, T.lead);
        return;
    }
    // RA = r[0] r[1] (w slots unused), RB = p[0] p[1] (in the s slots)
    const unsigned cur = (unsigned)s.pbuf;
    const unsigned seg = c.nvox * 4u, aend = 4u * seg;
    const rsrc_t RA = make_rsrc(b.va, aend);
    const rsrc_t RB = make_rsrc(b.vb, aend);
    const rsrc_t RX = make_rsrc(b.xlo, seg);
    const unsigned o_r = cur * seg, o_rn = (cur ^ 1u) * seg;
    const unsigned o_p = cur * seg, o_pn = (cur ^ 1u) * seg;
    const unsigned cv = G.cofs * 4u, rv = R.has ? rofs * 4u : aend;
    const bool after_true = s.kind == K_SETUP || s.kind == K_RESID;  // no p needed: WCOMP or done
    // x_lo is read and written only by iterations that update it; the lanes
    // of the others address it out of range (no access and no branch)
    const int xm = cg_x_mode(s);
    const float aprev = (float)s.alpha_prev;
    const unsigned xo = (after_true || xm == 0) ? aend : cv;
    // NTS: the iteration's vectors are stored nontemporal (films whose CG
    // vectors exceed the last-level cache, where nothing they store is read
    // again from it anyway); otherwise with device scope, like every solver
    // store
    auto vst = [](rsrc_t r, unsigned voff, unsigned soff, float v) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff,
                                              NTS ? ST_STREAM : ST_DEVICE);
    };

    Plane2 L[NW];
    auto issue = [&](int m) {
        const unsigned sv = so(m) * pb4;
        L[m].r = bload<float>(RA, cv, o_r + sv);
        L[m].d = bload<float>(Rdi, (cd[m] & 0xFFFFu) * 4u, 0u);
        L[m].rh = bload<float>(RA, rv, o_r + sv);
        L[m].dh = bload<float>(Rdi, (cd[m] >> 16) * 4u, 0u);
        L[m].p = after_true ? 0.0f : bload<float>(RB, cv, o_p + sv);
        L[m].ph = after_true ? 0.0f : bload<float>(RB, rv, o_p + sv);
        L[m].x = (m < 2 || m > KZ + 1) ? 0.0f : bload<float>(RX, xo, sv);
    };
    issue(0);
    issue(1);

    if (tid < 64) sHP[tid] = hpv;
    wave_partials<3, PM>(P, sred);
    __syncthreads();
    double tot[3];
    wave_lds_total<3>(sred, tot);
    double alpha_d = 0.0, beta_d = 0.0;
    // every lane holds the same decision; made wave-uniform, the branches on
    // it are scalar and alpha, beta sit in scalar registers
    const int mode = __builtin_amdgcn_readfirstlane(cg_advance(s, cg_decide(s, tot, alpha_d, beta_d), xm, alpha_d));
    pass_state(s, b.st + io, T.lead);
    if (mode == 1) return;
    double acc[3] = {0.0, 0.0, 0.0};
    if (mode == 2) {
        cg32_resid<KZ, true>(b, c, tp, G, k0, RA, o_r, Rsc, Rdi, acc);
    } else if (mode == 4) {
        // XADD: the deferred x_lo += alpha_prev p of the last iteration
#pragma unroll
        for (int m = 2; m <= KZ + 1; ++m) {
            const int k = k0 - 2 + m;
            const unsigned sv = so(m) * pb4;
            const float xv = bload<float>(RX, cv, sv), pv = bload<float>(RB, cv, o_p + sv);
            vst(RX, (tp.live && k < nz) ? cv : aend, sv, xv + aprev * pv);
        }
    } else {
        const float alpha = uniform_f((float)alpha_d), beta = uniform_f((float)beta_d);
        const bool fresh = __builtin_amdgcn_readfirstlane(beta_d == 0.0 ? 1 : 0) != 0;
        float af[3] = {0.0f, 0.0f, 0.0f};  // the thread's sums over its planes (FP32)
        // The plane loop, built twice: an iteration (ITER) and the
        // re-formation after a residual replacement (WCOMP), so neither
        // carries the other's selects.
        auto planes = [&](auto iter_c) {
            constexpr bool iter = decltype(iter_c)::value;
            float pc[NW], ph[NW], uc[NW];
            Coef3 cf[NW];
            // planes that can hold identity rows: the electrode planes under
            // Dirichlet electrodes, the gauge node's plane under Neumann
            auto ident_plane = [&](int k) { return c.zbc == 0 ? (k == 0 || k == nz - 1) : (k == nz / 2); };
            auto mirrored = [&](int k) { return c.zbc != 0 && (k == 0 || k == nz - 1); };
            // A(m): p of plane m (tile and two-voxel ring) into LDS; on the
            // tile p_i and x_lo += alpha p_i are stored (WCOMP: x_lo folded)
            auto stageA = [&](int m) {
                const int k = k0 - 2 + m;
                const bool tile_plane = m >= 2 && m <= KZ + 1;
                const bool ok = tile_plane && tp.live && k < nz;
                if constexpr (iter) {
                    const float uo = L[m].r * L[m].d, uoh = L[m].rh * L[m].dh;
                    pc[m] = fresh ? uo : uo + beta * L[m].p;
                    ph[m] = fresh ? uoh : uoh + beta * L[m].ph;
                    sp[m & 1][ty + 2][tx + 2] = pc[m];
                    if (R.has) sp[m & 1][R.ly + 2][R.lx + 2] = ph[m];
                    if (tile_plane) {
                        const unsigned vo = ok ? cv : aend, sv = so(m) * pb4;
                        vst(RB, vo, o_pn + sv, pc[m]);
                        const float xa = xm == 2 ? L[m].x + aprev * L[m].p : L[m].x;
                        vst(RX, xm != 0 ? vo : aend, sv, xa + alpha * pc[m]);
                    }
                } else if (ok) {
                    const long long id = (long long)G.cofs + (long long)c.plane * k;
                    b.xhi[id] += (double)b.xlo[id];
                    b.xlo[id] = 0.0f;
                }
            };
            // B(m): s = A p, r -= alpha s, u = M r on the tile and the
            // one-voxel ring of plane m (in-plane p from LDS, z neighbours
            // from registers); u into LDS, the new r stored on the tile.  The
            // centre row's face coefficients are kept for C(m).
            auto stageB = [&](int m) {
                const int k = k0 - 2 + m;
                const unsigned cc = cd[m] & 0xFFFFu;
                cf[m] = coef_of(cc, sHP);
                float sc = 0.0f, sh = 0.0f;
                if constexpr (iter) {
                    const unsigned ch = cd[m] >> 16;
                    float xu = pc[m + 1], xd = pc[m - 1], hu = ph[m + 1], hd = ph[m - 1];
                    if (c.zbc != 0) {
                        if (k == 0) xd = xu, hd = hu;
                        else if (k == nz - 1) xu = xd, hu = hd;
                    }
                    const float(*Pm)[PX] = sp[m & 1];
                    sc = row_apply32(cf[m], ih2, pc[m], Pm[ty + 2][tx + 3], Pm[ty + 2][tx + 1], Pm[ty + 3][tx + 2],
                                     Pm[ty + 1][tx + 2], xu, xd);
                    if (winner)
                        sh = row_Ax32(ch, sHP, ih2, ph[m], Pm[by + 2][bx + 3], Pm[by + 2][bx + 1], Pm[by + 3][bx + 2],
                                      Pm[by + 1][bx + 2], hu, hd);
                    if (ident_plane(k)) {
                        if (cc & SC_IDENTITY) sc = op_identity(pc[m], k, c);
                        if (ch & SC_IDENTITY) sh = op_identity(ph[m], k, c);
                    }
                }
                const float rc = iter ? L[m].r - alpha * sc : L[m].r;
                uc[m] = rc * L[m].d;
                su[m & 1][ty + 1][tx + 1] = uc[m];
                if (R.inner) {
                    const float rh = iter ? L[m].rh - alpha * sh : L[m].rh;
                    su[m & 1][R.ly + 1][R.lx + 1] = rh * L[m].dh;
                }
                if (m >= 2 && m <= KZ + 1) {
                    const bool ok = tp.live && k < nz;
                    if constexpr (iter) vst(RA, ok ? cv : aend, o_rn + so(m) * pb4, rc);
                    if (ok) {
                        dot_w(af[0], rc, uc[m], mirrored(k));
                        af[2] = __builtin_fmaf(rc, rc, af[2]);
                    }
                }
            };
            // C(m): w = A u on the tile of plane m and its term of w.u
            auto stageC = [&](int m) {
                const int k = k0 - 2 + m;
                float xu = uc[m + 1], xd = uc[m - 1];
                if (c.zbc != 0) {
                    if (k == 0) xd = xu;
                    else if (k == nz - 1) xu = xd;
                }
                const float(*Um)[HX] = su[m & 1];
                float wv = row_apply32(cf[m], ih2, uc[m], Um[ty + 1][tx + 2], Um[ty + 1][tx], Um[ty + 2][tx + 1],
                                       Um[ty][tx + 1], xu, xd);
                if (ident_plane(k) && (cd[m] & SC_IDENTITY)) wv = op_identity(uc[m], k, c);
                if (tp.live && k < nz) {
                    dot_w(af[1], wv, uc[m], mirrored(k));
                }
            };
            // Phase n: A(n); B(n - 1), whose in-plane p was written in phase
            // n - 1 and whose upper z neighbour A(n) has just formed; C(n - 2),
            // whose u was written in phase n - 1.  One barrier per phase; a
            // plane's LDS buffer is rewritten two phases after it was filled.
#pragma unroll
            for (int n = 0; n < NW; ++n) {
                k0p = k0;
                asm volatile("" : "+s"(k0p));
                if (n + 4 < NW) load_code(n + 4);
                if (n + 2 < NW) issue(n + 2);
                stageA(n);
                if (n >= 2) stageB(n - 1);
                if (n >= 4) stageC(n - 2);
                // the sums are formed phase by phase (left to the compiler,
                // their FMAs were gathered late in the loop and the kernel
                // spilled at depth 32)
                asm volatile("" : "+v"(af[0]), "+v"(af[1]), "+v"(af[2]));
                if (n + 1 < NW) __syncthreads();
            }
        };
        if (mode == 0)
            planes(std::true_type{});
        else
            planes(std::false_type{});
#pragma unroll
        for (int qq = 0; qq < 3; ++qq) acc[qq] = af[qq];
    }
    write_block_partials_w<3>(acc, io ? b.part[1] : b.part[0], ps, T.blk, spart);
}

// Waves per SIMD the register budget is set for: 6 (3 workgroups per CU,
// <= 80 VGPRs) at two planes per tile, 5 at four and eight, 4 at sixteen.
// The automatic choice runs this form only on films of at most 512
// depth-8 workgroups, two per CU, which 5 waves per SIMD still hold.  The
// batch kernel, which also holds its film's descriptor, is set for 4.
constexpr int cg32_waves(int kz) { return kz >= 16 ? 4 : (kz <= 2 ? 6 : 5); }

// MIR: the film has Neumann electrodes (mirrored boundary planes weighted in
// the CG sums); Dirichlet films run the MIR = false build.
template <int KZ, int PM, bool MIR>
__global__ __launch_bounds__(NT, cg32_waves(KZ)) void k_cg32(CG32Bufs b, StepConsts c, int nb, int ps, int t) {
    cg32_body<KZ, PM, MIR>(b, c, nb, ps, t, tile_id());
}

// The two-stencil form holds p, u and r of three planes besides the loads in
// flight: 6 waves per SIMD at two planes per tile, 5 at four, 4 at eight,
// sixteen and thirty-two.
constexpr int cg32r_waves(int kz) { return kz <= 2 ? 6 : (kz == 4 ? 5 : 4); }  // 4 at 16 and 32

template <int KZ, int PM, bool NTS>
__global__ __launch_bounds__(NT, cg32r_waves(KZ)) void k_cg32r(CG32Bufs b, StepConsts c, int nb, int ps, int t) {
    cg32r_body<KZ, PM, NTS>(b, c, nb, ps, t, tile_id());
}

// All films of a batch in one launch: the tile table maps each workgroup to
// (film, logical tile); every film runs its own control block, and the tiles
// of a film whose solve has finished return at once.
template <int KZ, int PM, bool MIR>
__global__ __launch_bounds__(NT, 4) void k_cg32_batch(const BatchFilm* __restrict__ films, const int2* __restrict__ tiles,
                                                                      int t) {
    TileId T;
    const BatchFilm& F = films[batch_tile(tiles, films, T)];
    // descriptor fields read once, before any store (scalar loads)
    const CG32Bufs b = F.cg;
    const StepConsts c = F.c;
    cg32_body<KZ, PM, MIR>(b, c, F.nb, F.ps, t, T);
}

void launch_cg32_batch(hipStream_t s, int ntiles, int kz, int pm, bool mirrored, const BatchFilm* films,
                       const int2* tiles, int t) {
    const dim3 gr(ntiles), bl(BX, BY, 1);
#define HZO_CGB_PM(KZ, M)                                                                                \
    switch (pm) {                                                                                        \
        case 1: hipLaunchKernelGGL((k_cg32_batch<KZ, 1, M>), gr, bl, 0, s, films, tiles, t); break;       \
        case 2: hipLaunchKernelGGL((k_cg32_batch<KZ, 2, M>), gr, bl, 0, s, films, tiles, t); break;       \
        case 4: hipLaunchKernelGGL((k_cg32_batch<KZ, 4, M>), gr, bl, 0, s, films, tiles, t); break;       \
        default: hipLaunchKernelGGL((k_cg32_batch<KZ, 0, M>), gr, bl, 0, s, films, tiles, t); break;      \
    }
#define HZO_CGB(KZ)             \
    if (mirrored) {             \
        HZO_CGB_PM(KZ, true)    \
    } else {                    \
        HZO_CGB_PM(KZ, false)   \
    }
    if (kz == 2) {
        HZO_CGB(2)
    } else if (kz == 4) {
        HZO_CGB(4)
    } else {
        HZO_CGB(8)
    }
#undef HZO_CGB
#undef HZO_CGB_PM
}

template <bool MIR>
static void launch_cg32_as(hipStream_t s, const Grid& g, const StepConsts& c, const CG32Bufs& b, int t,
                           const LaunchEv& ev) {
#define HZO_X , MIR
    const int nb = blocks_for_impl(g), ps = part_stride_impl(nb);
    if (g.kz == 16) {  // deep tiles for large films
        const int pm = part_pm(b.nb_in);
        const dim3 gr = grid3(g), bl(BX, BY, 1);
        HZO_PCG_LAUNCH_PM(k_cg32, 16, pm, gr, bl, s, b, c, nb, ps, t);
        return;
    }
    HZO_PCG_LAUNCH_X(k_cg32, g, b.nb_in, s, b, c, nb, ps, t);
#undef HZO_X
}

static void launch_cg32r(hipStream_t s, const Grid& g, const StepConsts& c, const CG32Bufs& b, int t,
                         bool stream_stores, const LaunchEv& ev) {
    const int nb = blocks_for_impl(g), ps = part_stride_impl(nb);
    if (g.kz >= 16) {  // deep tiles for large films
        const int pm = part_pm(b.nb_in);
        const dim3 gr = grid3(g), bl(BX, BY, 1);
#define HZO_X , true
        if (stream_stores) {
            if (g.kz == 32) {
                HZO_PCG_LAUNCH_PM(k_cg32r, 32, pm, gr, bl, s, b, c, nb, ps, t);
            } else {
                HZO_PCG_LAUNCH_PM(k_cg32r, 16, pm, gr, bl, s, b, c, nb, ps, t);
            }
            return;
        }
#undef HZO_X
#define HZO_X , false
        if (g.kz == 32) {
            HZO_PCG_LAUNCH_PM(k_cg32r, 32, pm, gr, bl, s, b, c, nb, ps, t);
        } else {
            HZO_PCG_LAUNCH_PM(k_cg32r, 16, pm, gr, bl, s, b, c, nb, ps, t);
        }
        return;
    }
    HZO_PCG_LAUNCH_X(k_cg32r, g, b.nb_in, s, b, c, nb, ps, t);
#undef HZO_X
}

void launch_cg32(hipStream_t s, const Grid& g, const StepConsts& c, const CG32Bufs& b, int t, bool two_stencil,
                 bool stream_stores, const LaunchEv& ev) {
    if (two_stencil)
        launch_cg32r(s, g, c, b, t, stream_stores, ev);
    else if (c.zbc != 0)
        launch_cg32_as<true>(s, g, c, b, t, ev);
    else
        launch_cg32_as<false>(s, g, c, b, t, ev);
}
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
        int* o = info + BATCH_INFO * f;
        o[0] = st.done;
        o[1] = st.iters;
        o[2] = st.status;
        o[3] = st.refinements;
        o[4] = st.xadds;
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
