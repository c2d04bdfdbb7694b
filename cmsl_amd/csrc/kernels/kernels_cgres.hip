// kernels_cgres.hip -- register-resident single-reduction CG: the whole
// FP32-refined Poisson solve in one persistent launch.
//
// The single-reduction CG of kernels_cg32.hip streams r, w, s, p and x_lo
// through memory every iteration (38 bytes per voxel); on a film that fits
// the GPU's register files that traffic is not needed at all.  Here each
// workgroup owns a 16 x 16 column tile over the film's full depth, two
// threads per column (the lower and the upper half of the planes), and
// keeps the tile's r, w, s, p, x_lo and Jacobi D^-1 in registers for the
// whole solve (an MI355X holds 256 x 512 KB of vector registers; a
// 256 x 256 x 32 film needs 50 MB of them).  Per
// iteration only the tile's edge columns of w cross to the neighbouring
// tiles, and one grid-wide reduction of (gamma, delta, r.r) ends it:
//   after the barrier: every workgroup sums the published partials in the
//   same fixed order, runs the same loop decision (cg_decide), and updates
//   s = w + beta s, r -= alpha s, p = M r_old + beta p, x_lo += alpha p on
//   its columns and, for s and r, on its one-column ring (the ring keeps
//   the neighbours' s and r by the same recurrence and the same arithmetic,
//   given their published w, so it holds their values bit for bit);
//   then u = M r on tile and ring (LDS), w = A u on the tile, partials,
//   publication of the edge columns of w, and the barrier.
// The residual replacement of the FP32-refined form (every refine_every
// iterations and when the recursive residual claims convergence): x_lo is
// folded into x_hi and the tile's edge columns of x go out in FP64 (done by
// the iteration itself when the replacement is the periodic one, otherwise
// by a pass of its own); then r = b - A x in FP64 on the tile (plane by
// plane through LDS) and its edge columns go out; then u and w are
// re-formed from the true residual (WCOMP).  The iterates are those of the
// streaming kernel's recurrence; only the grouping of the inner products
// differs.
//
// Grid barrier: cooperative launch (every workgroup resident), one arrival
// counter per XCD (workgroup w runs on XCD w % 8) whose last arriver bumps a
// global counter that all workgroups poll.  Data crossing workgroups is
// stored with device scope (written through the XCD's L2) and loaded with
// device-scope coherence, so no cache-wide write-back or invalidate is
// needed; every wait is bounded and a timeout raises an error flag that
// ends every workgroup at its next barrier.
#include "../hzo_internal.h"
#include "common.hip.h"
#include "tile.hip.h"
#include "reduce.hip.h"
#include "cg_ctl.hip.h"

namespace hzo {

namespace {

constexpr int RT = 16;        // tile edge in x and y
constexpr int RTH = RT + 2;   // with the one-column ring
constexpr int RNT = 2 * RT * RT;  // threads per workgroup: two per column (z halves)
constexpr int RNW = RNT / 64;     // waves per workgroup
constexpr unsigned RES_SPIN_LIMIT = 1u << 22;
constexpr int LD_COH = 17;  // buffer-load cache policy: device-scope coherent (SC0 | SC1)

__device__ __forceinline__ float ld_coh(rsrc_t R, unsigned voff, unsigned soff = 0u) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(R, (int)voff, (int)soff, LD_COH));
}
__device__ __forceinline__ double ld_coh_d(rsrc_t R, unsigned voff, unsigned soff = 0u) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(R, (int)voff, (int)soff, LD_COH));
}

// Grid barrier number q (0, 1, ...) of the G workgroups blk0 .. blk0 + G - 1
// of the launch.  The caller's stores have all been issued; each wave waits
// for them to complete before the workgroup arrives.  Returns false after a
// timeout (the error flag is set).
__device__ __forceinline__ bool res_barrier(unsigned* sync, int* err, int q, int G, int blk0, int& sok) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int ngrp = G < N_XCD ? G : N_XCD;  // G consecutive workgroups cover min(G, 8) XCDs
        const int xg = blockIdx.x % N_XCD;
        // workgroups of the range on this XCD: those of [0, blk0 + G) less those of [0, blk0)
        const unsigned nxg = (unsigned)((blk0 + G - xg + N_XCD - 1) / N_XCD - (blk0 - xg + N_XCD - 1) / N_XCD);
        const unsigned old = __hip_atomic_fetch_add(sync + 64 * (1 + xg), 1u, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1u == (unsigned)(q + 1) * nxg)
            __hip_atomic_fetch_add(sync, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned target = (unsigned)(q + 1) * (unsigned)ngrp;
        unsigned spins = 0;
        int ok = 1;
        auto give_up = [&]() {
            if (++spins > RES_SPIN_LIMIT ||
                ((spins & 1023u) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                return true;
            }
            __builtin_amdgcn_s_sleep(1);
            return false;
        };
        // Two polls in flight: each poll's load goes out before the previous
        // one's value is tested (the loop alternates between two values, so
        // no copy waits for the younger load), and the release is seen about
        // half a load round trip sooner than by polls issued one after
        // another.
        unsigned va = __hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (true) {
            const unsigned vb = __hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (va >= target || give_up()) break;
            va = __hip_atomic_load(sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (vb >= target || give_up()) break;
        }
        sok = ok;
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(sok) != 0;
}

// The control-block update after cg_decide for the resident solve: an
// iteration counts, the other modes only record what the pass does (x_lo
// lives in registers, so no x_lo term is ever deferred).
__device__ __forceinline__ int res_advance(PcgState& s, int mode) {
    if (mode == 0) {
        s.iters += 1;
        s.kind = K_ITER;
    } else if (mode == 2) {
        s.kind = K_RESID;
    } else if (mode == 3) {
        s.kind = K_WCOMP;
    }
    return mode;
}

}  // namespace

// The solve of one film by the G workgroups blk0 .. blk0 + G - 1 of the
// launch; this workgroup owns tile blk (0 .. G - 1).  NZ = nz of the film
// (each thread holds NZ/2 planes of a column); MIR: the build that weights
// mirrored planes in the inner products (Neumann electrodes; exact weight 1
// on Dirichlet films).
template <int NZ, bool MIR>
__device__ __forceinline__ void cgres_body(const ResBufs& b, const StepConsts& c, int gxt, int G, int blk,
                                           int blk0) {
    static_assert(NZ % 8 == 0, "ring cells are spread over the workgroup in whole rounds");
    constexpr int NH = NZ / 2;       // planes per thread
    constexpr int NR = 4 * NZ * RT;  // ring cells: 4 sides x NZ planes x 16
    constexpr int RPT = NR / RNT;    // ring cells per thread
    __shared__ __align__(16) float uT[NZ][RTH][RTH];  // RESID's second pass holds the ring's x here
    // x_lo of the tile stays in LDS (each thread touches only its own
    // voxels, once per iteration), which leaves the registers to r, w, s, p,
    // D^-1 and the codes; eT stages the vector whose edge columns a pass
    // publishes, so that the stores go out coalesced and branch-free
    __shared__ float eT[NZ][RT][RT], xT[NZ][RT][RT];
    __shared__ float rR[NR], sR[NR], dR[NR];
    __shared__ HPair<float> sHP[64];
    __shared__ double sH64[16];
    __shared__ double xP[2][4][RTH][RTH];  // RESID: [half][plane % 4]
    __shared__ double sredA[3][RNW], sredB[3][RNW];
    __shared__ int sok;

    const int tid = threadIdx.x, tx = tid & (RT - 1), ty = (tid / RT) & (RT - 1), tz = tid / (RT * RT);
    // a wave lies in one z half: its first plane is wave-uniform, so plane
    // offsets can go in the buffer instructions' scalar offset
    const int wv = tid >> 6, k0 = __builtin_amdgcn_readfirstlane(tz * NH);
    const int bxt = blk % gxt, byt = blk / gxt, gyt = G / gxt;
    const int nx = c.nx, ny = c.ny;
    const unsigned pl = c.plane, nv = c.nvox;
    const int i0 = bxt * RT, j0 = byt * RT;
    const unsigned col = (unsigned)(i0 + tx) + (unsigned)nx * (unsigned)(j0 + ty);
    // neighbour tiles across the x-, x+, y-, y+ sides (periodic)
    int nbt[4];
    nbt[0] = (bxt + gxt - 1) % gxt + byt * gxt;
    nbt[1] = (bxt + 1) % gxt + byt * gxt;
    nbt[2] = bxt + ((byt + gyt - 1) % gyt) * gxt;
    nbt[3] = bxt + ((byt + 1) % gyt) * gxt;

    const rsrc_t Rsc = make_rsrc(b.sc, nv * 2u);
    const rsrc_t Rr0 = make_rsrc(b.r0, nv * 4u);
    const rsrc_t Rxl = make_rsrc(b.xlo, nv * 4u);
    const rsrc_t Rxh = make_rsrc(b.xhi, nv * 8u);
    const rsrc_t RXD = make_rsrc(b.xchd, (unsigned)G * NR * 8u);
    const rsrc_t Rb = make_rsrc(b.b, nv * 8u);
    const rsrc_t Rdi = make_rsrc(b.lut32 + LUT_DINV, SC_LUT_SIZE * 4u);
    const rsrc_t RX = make_rsrc(b.xch, 2u * (unsigned)G * NR * 4u);
    const rsrc_t RP = make_rsrc(b.part, 2u * (unsigned)G * 4u * 8u);
    // edge-column slot: [parity][tile][side][plane][position]
    auto xoff = [&](unsigned slot, int tile, int side, int k, int pos) {
        return ((((slot * (unsigned)G + (unsigned)tile) * 4u + (unsigned)side) * NZ + (unsigned)k) * RT +
                (unsigned)pos) * 4u;
    };
    // FP64 edge-column slot of the residual replacement: [tile][side][plane][position]
    auto xdoff = [&](int tile, int side, int k, int pos) {
        return ((((unsigned)tile * 4u + (unsigned)side) * NZ + (unsigned)k) * RT + (unsigned)pos) * 8u;
    };
    // ring cell e = (side * NZ + k) * 16 + pos: its LDS place and film
    // column.  Unsigned arithmetic: with the thread index known to be below
    // RNT, the side of each of a thread's cells folds to a constant where
    // NZ * 16 divides RNT.
    auto ring_lds = [&](unsigned e, int& k, int& ly, int& lx) {
        const int side = (int)(e / (NZ * RT)), pos = (int)(e % RT);
        k = (int)((e / RT) % NZ);
        ly = side < 2 ? pos + 1 : (side == 2 ? 0 : RTH - 1);
        lx = side < 2 ? (side == 0 ? 0 : RTH - 1) : pos + 1;
    };
    auto ring_col = [&](unsigned e) {
        const int side = (int)(e / (NZ * RT)), pos = (int)(e % RT);
        int gi, gj;
        if (side < 2) {
            gi = side == 0 ? (i0 + nx - 1) % nx : (i0 + RT) % nx;
            gj = j0 + pos;
        } else {
            gi = i0 + pos;
            gj = side == 2 ? (j0 + ny - 1) % ny : (j0 + RT) % ny;
        }
        return (unsigned)gi + (unsigned)nx * (unsigned)gj;
    };

    if (tid < 64) sHP[tid] = load_hpair(b.lut32);
    if (tid < 16) sH64[tid] = b.lut[LUT_HARM + tid];

    // The thread's planes of its column: vectors and codes (two per register).
    float r[NH], w[NH], s[NH], p[NH], d[NH];
    unsigned cdp[NH / 2];
#pragma unroll
    for (int kk = 0; kk < NH; ++kk) {
        const unsigned so = (unsigned)(k0 + kk) * pl;
        const unsigned code = bload_code(Rsc, col * 2u, so * 2u);
        if (kk & 1) cdp[kk / 2] |= code << 16;
        else cdp[kk / 2] = code;
        d[kk] = bload<float>(Rdi, code * 4u, 0u);
        r[kk] = bload<float>(Rr0, col * 4u, so * 4u);
        xT[k0 + kk][ty][tx] = bload<float>(Rxl, col * 4u, so * 4u);
        p[kk] = 0.0f;
        w[kk] = 0.0f;
        s[kk] = 0.0f;
    }
    auto code_of = [&](int kk) { return (cdp[kk / 2] >> (16 * (kk & 1))) & 0xFFFFu; };
#pragma unroll
    for (int m = 0; m < RPT; ++m) {
        const unsigned e = (unsigned)tid + RNT * m;
        int k, ly, lx;
        ring_lds(e, k, ly, lx);
        const unsigned rv = ring_col(e) + (unsigned)k * pl;  // the plane differs across lanes
        const unsigned code = bload_code(Rsc, rv * 2u, 0u);
        dR[e] = bload<float>(Rdi, code * 4u, 0u);
        rR[e] = bload<float>(Rr0, rv * 4u, 0u);
        sR[e] = 0.0f;
    }

    // A pass's partial sums: each wave's sums go to its LDS slots here; the
    // workgroup barrier that starts the edge-column publication makes them
    // visible, and thread 0 adds the slots in wave order after it (the
    // operations of wg_total below), so the pass spends no barrier of its own
    // on them.
    auto wave_slots = [&](double (&a)[3], double (&S)[3][RNW]) {
#pragma unroll
        for (int qq = 0; qq < 3; ++qq) {
            const double ws = wave_sum(a[qq]);
            if ((tid & 63) == 0) S[qq][wv] = ws;
        }
    };
    // Workgroup totals of per-thread values in a fixed order (waves, then
    // the wave slots in order): the same bits in every thread.
    auto wg_total = [&](double (&a)[3], double (&S)[3][RNW], double (&t)[3]) {
#pragma unroll
        for (int qq = 0; qq < 3; ++qq) {
            const double ws = wave_sum(a[qq]);
            if ((tid & 63) == 0) S[qq][wv] = ws;
        }
        __syncthreads();
#pragma unroll
        for (int qq = 0; qq < 3; ++qq) {
            double v = S[qq][0];
#pragma unroll
            for (int u = 1; u < RNW; ++u) v += S[qq][u];
            t[qq] = v;
        }
    };

    // The control block lives in LDS between decisions, double-buffered:
    // every thread runs the (deterministic) decision on the same totals from
    // copy dpar, thread 0 writes the result to the other copy, which the
    // next decision reads after at least one workgroup barrier.  So no
    // broadcast barrier follows a decision, and the block is live in
    // registers only around it.
    __shared__ PcgState sst[2];
    int dpar = 0;
    if (tid == 0) store_state(load_state(b.st), &sst[0]);
    int mode = 1;
    float af = 0.0f, bf = 0.0f;
    bool fresh = false;
    bool fold_now = false;  // an iteration that a periodic residual replacement follows
    auto decide = [&](const double (&t)[3], bool cap, PcgState st) {
        double a = 0.0, bb = 0.0;
        int m = res_advance(st, cg_decide(st, t, a, bb));
        if (m != 1 && cap) {
            st.done = 1;
            st.status = HZO_PCG_MAXITER;
            m = 1;
        }
        if (tid == 0) store_state(st, &sst[dpar ^ 1]);
        dpar ^= 1;
        // every thread holds the same decision: made wave-uniform, the
        // branches on it are scalar and alpha, beta sit in scalar registers
        mode = __builtin_amdgcn_readfirstlane(m);
        af = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, (float)a)));
        bf = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, (float)bb)));
        fresh = __builtin_amdgcn_readfirstlane(bb == 0.0 ? 1 : 0) != 0;  // first iteration of a solve: no previous p, s
        // the decision after this iteration will be a replacement (or the
        // end) when its count is a multiple of the replacement period
        fold_now = __builtin_amdgcn_readfirstlane(
                       (m == 0 && st.refine_every > 0 && multiple_of(st.iters, st.refine_every)) ? 1 : 0) != 0;
    };

    // The setup's partial sums (r.z, r.r, b.b) and the first decision.
    double tot[3];
    {
        double a[3] = {0.0, 0.0, 0.0};
        for (int t = tid; t < b.nb_in; t += RNT) {
#pragma unroll
            for (int qq = 0; qq < 3; ++qq) a[qq] += b.pin[(long long)qq * b.ps_in + t];
        }
        wg_total(a, sredB, tot);
    }
    decide(tot, false, load_state(&sst[dpar]));
    const float ih2 = (float)c.inv_h2;
    const int qcap = 4 * sst[0].max_iter + 16;
    int q = 0;             // grid barriers passed
    bool ring_r = false;   // the ring's r comes from the last pass's edge columns (after RESID)
    bool resid_x = false;  // RESID: x_lo is folded and the edge columns of x are out
    // the ring cells' values the neighbours published in the last pass
    // (loaded right after each grid barrier, beside the partial sums)
    float rin[RPT];
#pragma unroll
    for (int m = 0; m < RPT; ++m) rin[m] = 0.0f;

    while (mode != 1) {
        // the codes and the thread index enter each pass as fresh values, so
        // the addresses derived from them are formed where they are used
        // instead of being hoisted out of the loop and held live through it
#pragma unroll
        for (int i = 0; i < NH / 2; ++i) asm volatile("" : "+v"(cdp[i]));
        unsigned tidp = (unsigned)tid;
        int txp = tx, typ = ty;
        unsigned colp = col;
        asm volatile("" : "+v"(tidp), "+v"(txp), "+v"(typ), "+v"(colp));
        __builtin_assume(tidp < (unsigned)RNT);
        const unsigned slot = (unsigned)q & 1u;
        // the ring cells' values the neighbours published in pass slot sl
        auto ring_load = [&](unsigned sl) {
#pragma unroll
            for (int m = 0; m < RPT; ++m) {
                const unsigned e = tidp + RNT * m;
                const int side = (int)(e / (NZ * RT)), k = (int)((e / RT) % NZ), pos = (int)(e % RT);
                rin[m] = ld_coh(RX, xoff(sl, nbt[side], side ^ 1, k, pos));
            }
        };
        // The tile's edge columns of the vector staged in eT go out, one
        // ring-ordered slot per thread and round (contiguous stores).
        auto publish_staged = [&](unsigned sl) {
            __syncthreads();
#pragma unroll
            for (int m = 0; m < RPT; ++m) {
                const unsigned e = tidp + RNT * m;
                const int side = (int)(e / (NZ * RT)), k = (int)((e / RT) % NZ), pos = (int)(e % RT);
                const int ex = side < 2 ? (side == 0 ? 0 : RT - 1) : pos;
                const int ey = side < 2 ? pos : (side == 2 ? 0 : RT - 1);
                bstore(RX, (((sl * (unsigned)G + (unsigned)blk) * (unsigned)NR) + (unsigned)e) * 4u, 0u, eT[k][ey][ex]);
            }
        };

        // x_hi + x_lo of the thread's voxels, x_lo = 0.  Every load is
        // issued before the first store: a load behind a store waits for
        // the store's write-through to complete.
        auto fold_load = [&](rsrc_t Rfrom, double (&xh)[NH]) {
#pragma unroll
            for (int kk = 0; kk < NH; ++kk) xh[kk] = ld_coh_d(Rfrom, colp * 8u, (unsigned)(k0 + kk) * pl * 8u);
        };
        auto fold_store = [&](rsrc_t Rto, const double (&xh)[NH]) {
#pragma unroll
            for (int kk = 0; kk < NH; ++kk) {
                float& xv = xT[k0 + kk][typ][txp];
                bstore(Rto, colp * 8u, (unsigned)(k0 + kk) * pl * 8u, xh[kk] + (double)xv);
                xv = 0.0f;
            }
        };
        // x_lo folded into x_hi in place (no other thread touches these
        // voxels of x_hi during the solve), and the tile's edge columns of
        // x = x_hi go out in FP64 for the residual replacement
        auto fold_publish_x = [&](const double (&xh)[NH]) {
#pragma unroll
            for (int kk = 0; kk < NH; ++kk) {
                const int k = k0 + kk;
                float& xl = xT[k][typ][txp];
                const double xv = xh[kk] + (double)xl;
                xl = 0.0f;
                bstore(Rxh, colp * 8u, (unsigned)k * pl * 8u, xv);
                if (txp == 0) bstore(RXD, xdoff(blk, 0, k, typ), 0u, xv);
                if (txp == RT - 1) bstore(RXD, xdoff(blk, 1, k, typ), 0u, xv);
                if (typ == 0) bstore(RXD, xdoff(blk, 2, k, txp), 0u, xv);
                if (typ == RT - 1) bstore(RXD, xdoff(blk, 3, k, txp), 0u, xv);
            }
        };

        double acc[3] = {0.0, 0.0, 0.0};
        bool partials = true;
        if (mode == 0 || mode == 3) {
            // ITER (0): s = w + beta s and r -= alpha s on tile and ring,
            // p = M r_old + beta p, x_lo += alpha p; WCOMP (3): x_lo folded
            // into x_hi.  Then u = M r (LDS), w = A u, partials (gamma, delta,
            // r.r), edge columns of w.
            const bool iter = mode == 0;
#pragma unroll
            for (int m = 0; m < RPT; ++m) {
                const unsigned e = tidp + RNT * m;
                int k, ly, lx;
                ring_lds(e, k, ly, lx);
                float rn = rR[e];
                if (iter) {
                    const float wn = rin[m];
                    const float sn = fresh ? wn : wn + bf * sR[e];
                    sR[e] = sn;
                    rn = rn - af * sn;
                    rR[e] = rn;
                } else if (ring_r) {
                    rn = rin[m];
                    rR[e] = rn;
                }
                uT[k][ly][lx] = rn * dR[e];
            }
            if (iter) {
#pragma unroll
                for (int kk = 0; kk < NH; ++kk) {
                    const float sn = fresh ? w[kk] : w[kk] + bf * s[kk];
                    s[kk] = sn;
                    const float uo = r[kk] * d[kk];
                    const float pn = fresh ? uo : uo + bf * p[kk];
                    p[kk] = pn;
                    float& xv = xT[k0 + kk][typ][txp];
                    xv = xv + af * pn;
                    r[kk] = r[kk] - af * sn;
                }
            } else if (!ring_r) {
                // WCOMP after the setup: x_lo is folded into x_hi (after a
                // replacement the RESID pass has folded it already)
                double xh[NH];
                fold_load(Rxh, xh);
                fold_store(Rxh, xh);
            }
#pragma unroll
            for (int kk = 0; kk < NH; ++kk) uT[k0 + kk][typ + 1][txp + 1] = r[kk] * d[kk];
            __syncthreads();
            // the thread's sums over its planes in FP32, across threads in FP64
            float af[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kk = 0; kk < NH; ++kk) {
                const int k = k0 + kk;
                // u of the column's own planes from registers (the same
                // product that went to LDS), of the other half's from LDS
                const float xc = r[kk] * d[kk];
                float xu = kk + 1 < NH ? r[kk + 1] * d[kk + 1] : (k + 1 < NZ ? uT[k + 1][typ + 1][txp + 1] : xc);
                float xd = kk > 0 ? r[kk - 1] * d[kk - 1] : (k > 0 ? uT[k - 1][typ + 1][txp + 1] : xc);
                if (c.zbc != 0) {  // mirror: planes 1 and nz - 2 stand on both sides of the faces
                    if (k == 0) xd = xu;
                    else if (k == NZ - 1) xu = xd;
                }
                const unsigned code = code_of(kk);
                const float qs = row_Ax32(code, sHP, ih2, xc, uT[k][typ + 1][txp + 2], uT[k][typ + 1][txp],
                                          uT[k][typ + 2][txp + 1], uT[k][typ][txp + 1], xu, xd);
                // identity rows lie on the electrode planes (Dirichlet) or
                // at the gauge voxel (Neumann): interior planes of a
                // Dirichlet film skip the test (a scalar branch)
                float wvl = qs;
                if (c.zbc != 0 || k == 0 || k == NZ - 1) {
                    if (code & SC_IDENTITY) wvl = op_identity(xc, k, c);
                }
                w[kk] = wvl;
                cg_dot_addf<MIR>(af[0], r[kk], xc, c, k);
                cg_dot_addf<MIR>(af[1], wvl, xc, c, k);
                af[2] = __builtin_fmaf(r[kk], r[kk], af[2]);
                eT[k][typ][txp] = wvl;
            }
#pragma unroll
            for (int qq = 0; qq < 3; ++qq) acc[qq] = af[qq];
            wave_slots(acc, sredA);
            if (fold_now) {
                // the iteration a periodic replacement follows: the
                // replacement's first pass is done here (x_hi loaded ahead
                // of the edge stores), so the replacement starts with its
                // second pass
                double xh[NH];
                fold_load(Rxh, xh);
                publish_staged(slot);
                fold_publish_x(xh);
            } else {
                publish_staged(slot);
            }
        } else if (!resid_x) {
            // RESID, first pass (a replacement the convergence test asked
            // for): x_lo folded and the edges of x published
            double xh[NH];
            fold_load(Rxh, xh);
            fold_publish_x(xh);
            partials = false;
        } else {
            // RESID, second pass: r = b - A x in FP64 (x = x_hi after the
            // first pass's fold).  The ring's x (the neighbours' published
            // edge columns) is staged in LDS as doubles over uT, which this
            // pass does not use; the column's x on planes k0 - 1 .. k0 + NH
            // (clamped into the film) is loaded up front, b two planes ahead.
            // Both halves walk their planes together, two planes per
            // barrier, each plane's x on tile and ring staged in
            // xP[half][plane % 4].
            double* xR = reinterpret_cast<double*>(&uT[0][0][0]);
            const unsigned o8 = colp * 8u;
            double rx[RPT];
#pragma unroll
            for (int m = 0; m < RPT; ++m) {
                const unsigned e = tidp + RNT * m;
                const int side = (int)(e / (NZ * RT)), k = (int)((e / RT) % NZ), pos = (int)(e % RT);
                rx[m] = ld_coh_d(RXD, xdoff(nbt[side], side ^ 1, k, pos));
            }
            double xq[NH + 2];
#pragma unroll
            for (int i = 0; i < NH + 2; ++i) {
                const int k = k0 - 1 + i, kc = k < 0 ? 0 : (k >= NZ ? NZ - 1 : k);
                xq[i] = ld_coh_d(Rxh, o8, (unsigned)kc * pl * 8u);
            }
            double bq[NH];
            bq[0] = bload<double>(Rb, o8, (unsigned)k0 * pl * 8u);
            bq[1] = bload<double>(Rb, o8, (unsigned)(k0 + 1) * pl * 8u);
#pragma unroll
            for (int m = 0; m < RPT; ++m) xR[tidp + RNT * m] = rx[m];
            __syncthreads();
            const int rl = (int)(tidp & (RT * RT - 1));  // ring loaders: 64 threads of each half
            static_assert(NH % 2 == 0, "the replacement stages its planes in pairs");
#pragma unroll
            for (int kk = 0; kk < NH; kk += 2) {
                // planes kk and kk + 1 of both halves staged (four plane
                // buffers: a pair's buffers are written again two pairs
                // later, after the next pair's barrier), b two pairs ahead
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int kj = kk + j, k = k0 + kj;
                    double (*S)[RTH] = xP[tz][kj & 3];
                    S[typ + 1][txp + 1] = xq[kj + 1];
                    if (rl < 4 * RT) {  // the plane's 64 ring cells
                        const int side = rl / RT, pos = rl % RT;
                        const unsigned e = (unsigned)((side * NZ + k) * RT + pos);
                        int kq, ly, lx;
                        ring_lds(e, kq, ly, lx);
                        S[ly][lx] = xR[e];
                    }
                    if (kj + 2 < NH) bq[kj + 2] = bload<double>(Rb, o8, (unsigned)(k + 2) * pl * 8u);
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int kj = kk + j, k = k0 + kj;
                    const double(*S)[RTH] = xP[tz][kj & 3];
                    const double xm = xq[kj], x0 = xq[kj + 1], xp = xq[kj + 2];
                    const double bv = bq[kj];
                    double xu = xp, xd = xm;
                    if (c.zbc != 0) {
                        if (k == 0) xd = xu;
                        else if (k == NZ - 1) xu = xd;
                    }
                    const unsigned code = code_of(kj);
                    // the row as a stencil (an identity row's code reads table
                    // entry 0, and its value is replaced below)
                    const double* Hr = sH64 + (code & 3u) * 4u;
                    const double num = Hr[(code >> 2) & 3u] * (S[typ + 1][txp + 2] - x0) -
                                       Hr[(code >> 4) & 3u] * (x0 - S[typ + 1][txp]) +
                                       Hr[(code >> 6) & 3u] * (S[typ + 2][txp + 1] - x0) -
                                       Hr[(code >> 8) & 3u] * (x0 - S[typ][txp + 1]) +
                                       Hr[(code >> 10) & 3u] * (xu - x0) - Hr[(code >> 12) & 3u] * (x0 - xd);
                    double Ax = -(num * c.inv_h2);
                    if (c.zbc != 0 || k == 0 || k == NZ - 1) {  // planes that can hold identity rows
                        if (code & SC_IDENTITY) Ax = op_identity(x0, k, c);
                    }
                    const float rv = (float)(bv - Ax);
                    r[kj] = rv;
                    cg_dot_add<MIR>(acc[0], rv, rv * d[kj], c, k);
                    acc[1] += (double)rv * (double)rv;
                    eT[k][typ][txp] = rv;
                }
            }
            wave_slots(acc, sredA);
            publish_staged(slot);
        }
        if (partials && tid == 0) {
#pragma unroll
            for (int qq = 0; qq < 3; ++qq) {
                double v = sredA[qq][0];
#pragma unroll
                for (int u = 1; u < RNW; ++u) v += sredA[qq][u];
                bstore(RP, ((slot * (unsigned)G + (unsigned)blk) * 4u + (unsigned)qq) * 8u, 0u, v);
            }
        }
        // the control block for the decision after the barrier, read before
        // the barrier is waited for (a pass with partials has passed a
        // workgroup barrier since the last decision wrote it; the first pass
        // of a replacement has not, and makes no decision)
        PcgState st_next;
        if (partials) st_next = load_state(&sst[dpar]);
        if (!res_barrier(b.sync, b.err, q, G, blk0, sok)) break;
        ++q;
        if (!partials) {  // RESID continues with its second pass
            resid_x = true;
            continue;
        }
        // every workgroup adds the partials in the same order; the ring's
        // values for the next pass are loaded behind them, in flight while
        // the partials are summed and the decision runs
        {
            double a[3] = {0.0, 0.0, 0.0};
            if (tid < G) {
#pragma unroll
                for (int qq = 0; qq < 3; ++qq)
                    a[qq] = ld_coh_d(RP, ((slot * (unsigned)G + (unsigned)tid) * 4u + (unsigned)qq) * 8u);
            }
            ring_load(slot);
            wg_total(a, sredB, tot);
        }
        ring_r = mode == 2;
        const bool folded = fold_now;  // the pass just ended did a replacement's first pass
        if (mode == 2) tot[2] = 0.0;
        decide(tot, q >= qcap, st_next);
        resid_x = mode == 2 && folded;
    }
    // the solution's correction goes back; x_hi is already in memory
#pragma unroll
    for (int kk = 0; kk < NH; ++kk) bstore(Rxl, col * 4u, (unsigned)(k0 + kk) * pl * 4u, xT[k0 + kk][ty][tx]);

    if (blk == 0 && tid == 0) {
        PcgState fin = load_state(&sst[dpar]);
        fin.passes = q;
        store_state(fin, b.st);
    }
}

template <int NZ, bool MIR>
__global__ __launch_bounds__(RNT, 1) void k_cgres(ResBufs b, StepConsts c, int gxt, int G) {
    cgres_body<NZ, MIR>(b, c, gxt, G, blockIdx.x, 0);
}

// Several films of one depth in one launch: each takes the block range
// films[f].blk0 .. + G - 1 (ascending, contiguous) and runs its own solve,
// with its own barriers, partial sums and control block.
template <int NZ, bool MIR>
__global__ __launch_bounds__(RNT, 1) void k_cgres_films(const ResFilm* __restrict__ films, int nf) {
    int f = 0;
    while (f + 1 < nf && films[f + 1].blk0 <= (int)blockIdx.x) ++f;
    const ResBufs b = films[f].b;
    const StepConsts c = films[f].c;
    const int blk0 = films[f].blk0;
    cgres_body<NZ, MIR>(b, c, films[f].gxt, films[f].G, (int)blockIdx.x - blk0, blk0);
}

__global__ void k_fill_u32(unsigned* p, int n, unsigned v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_u32(hipStream_t s, unsigned* p, int n, unsigned v) {
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}

// Resident solve of the film in one cooperative launch on stream s (the
// setup kernel has run).  Returns the HIP status of the launch.
hipError_t launch_cgres(hipStream_t s, const StepConsts& c, const ResBufs& b, int gxt, int G) {
    void* args[] = {(void*)&b, (void*)&c, (void*)&gxt, (void*)&G};
    const void* fn = nullptr;
    const bool mir = c.zbc != 0;
    switch (c.nz) {
        case 16: fn = mir ? (const void*)k_cgres<16, true> : (const void*)k_cgres<16, false>; break;
        case 24: fn = mir ? (const void*)k_cgres<24, true> : (const void*)k_cgres<24, false>; break;
        case 32: fn = mir ? (const void*)k_cgres<32, true> : (const void*)k_cgres<32, false>; break;
        default: return hipErrorInvalidValue;
    }
    return hipLaunchCooperativeKernel(fn, dim3(G), dim3(RNT), args, 0, s);
}

// The films' resident solves in one cooperative launch of their total tiles
// (films on device, nf of them, every one of depth nz; mir: some film has
// Neumann electrodes).
hipError_t launch_cgres_films(hipStream_t s, const ResFilm* films, int nf, int tiles, int nz, bool mir) {
    void* args[] = {(void*)&films, (void*)&nf};
    const void* fn = nullptr;
    switch (nz) {
        case 16: fn = mir ? (const void*)k_cgres_films<16, true> : (const void*)k_cgres_films<16, false>; break;
        case 24: fn = mir ? (const void*)k_cgres_films<24, true> : (const void*)k_cgres_films<24, false>; break;
        case 32: fn = mir ? (const void*)k_cgres_films<32, true> : (const void*)k_cgres_films<32, false>; break;
        default: return hipErrorInvalidValue;
    }
    return hipLaunchCooperativeKernel(fn, dim3(tiles), dim3(RNT), args, 0, s);
}

// Workgroups of the resident kernel one device can hold at once (0 if the
// film's depth has no instance or the device has no cooperative launch).
int cgres_capacity(int device, int nz, bool mir) {
    int coop = 0, ncu = 0, occ = 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess || !coop) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    const void* fn = nullptr;
    switch (nz) {
        case 16: fn = mir ? (const void*)k_cgres<16, true> : (const void*)k_cgres<16, false>; break;
        case 24: fn = mir ? (const void*)k_cgres<24, true> : (const void*)k_cgres<24, false>; break;
        case 32: fn = mir ? (const void*)k_cgres<32, true> : (const void*)k_cgres<32, false>; break;
        default: return 0;
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, RNT, 0) != hipSuccess) return 0;
    return occ * ncu;
}

}  // namespace hzo
