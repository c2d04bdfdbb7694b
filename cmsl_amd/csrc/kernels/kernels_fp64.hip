// kernels_fp64.hip -- plain FP64 form of the HZO step on CDNA4 (gfx950).
//
//   k_tdgl_f64        kernelCalc_antiferro + kernelUpdate_P (kernels_evol.cu:47-242)
//                     with E = -grad(phi) (depol_varEPS.cu:223-276) evaluated
//                     in-kernel from the previous solve's potential.
//   k_pcg_setup_f64   divergence + RHS + r = b - A x + z = D^-1 r and the
//                     inner products r.z, r.r, b.b (depol_varEPS.cu:279-329,
//                     :400-460)
//   k_pcg_dir_f64     p = z + beta p fused with q = A p and p.q (:479-485, :515-519)
//   k_pcg_update_f64  x += alpha p, r -= alpha A p, z = D^-1 r, r.z, r.r (:492-507)
//   k_diag_f64        the eight lab-frame diagnostics sums (evolve.cu:259-357)
//
// Every floating-point expression keeps the reference's operand order, and
// the file is compiled with -ffp-contract=off, so each voxel value is the
// double the reference expression denotes.  Inner products are
// deterministic (fixed-order grid reduction, reduce.hip.h).
//
// Launch shape: 64x8 workgroups over x-y tiles (x fastest: each wave reads a
// contiguous x-row segment), each walking kz z-planes (tile depth chosen per
// grid so that about a thousand workgroups cover the film).
#include "../hzo_internal.h"
#include "reduce.hip.h"
#include "common.hip.h"
#include "tile.hip.h"

namespace hzo {

int blocks_for(const Grid& g) { return blocks_for_impl(g); }
int part_stride(const Grid& g) { return part_stride_impl(blocks_for_impl(g)); }

// ---------------------------------------------------------------- TDGL
// z-streaming, as the FP32 form: each plane's P (6 components) and phi are
// staged in LDS with a one-voxel x-y halo (one buffer of 37 KB, two
// barriers per plane), the column's planes below and above stay in
// registers (z neighbours, clamped at the faces).  The phase table sits in
// LDS.  The FP64 arithmetic needs about 176 VGPRs, so one workgroup runs
// per CU and the plane loads are not issued a plane ahead (a 128-VGPR build
// for two workgroups per CU spilled and ran 2.4x slower; measured 69 us
// against 104 for the per-voxel form at 256x256x32).  The operands reach
// each expression in the reference's order, so every voxel value is
// bit-identical to the per-voxel evaluation.
struct Col64 {
    double P[6];
    double phi;
};

__global__ __launch_bounds__(NT) void k_tdgl_f64(Fp64Bufs b, StepConsts c, double Ez_ext, int nb) {
    __shared__ double sP[1][6][HY][HX];
    __shared__ double sphi[1][HY][HX];
    __shared__ PhaseC sph[PTAB_ENTRIES];
    const int nz = c.nz;
    const long long N = (long long)c.nx * c.ny * nz;
    const TileId T = tile_id();
    const TilePos tp = tile_pos(c, T);
    const TileCols G = tile_cols(tp, c);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = linear_tid();
    if (tid < PTAB_ENTRIES) sph[tid] = b.ptab[tid];
    const unsigned cc = G.cofs, hc = tp.has_h ? G.hofs : G.cofs, pb = c.plane;
    // buffer loads: one resource per array (each below 2^31 bytes for the
    // 2^28-voxel limit), the column as a lane offset, the plane uniform
    rsrc_t RP[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) RP[q] = make_rsrc(b.Pold + q * N, c.nvox * 8u);
    const rsrc_t Rph = make_rsrc(b.phi, c.nvox * 8u);
    auto ld = [&](unsigned col, int k, Col64& v) {
        const unsigned po = (unsigned)k * pb;
#pragma unroll
        for (int q = 0; q < 6; ++q) v.P[q] = bload<double>(RP[q], col * 8u, po * 8u);
        v.phi = c.dep_flag ? bload<double>(Rph, col * 8u, po * 8u) : 0.0;
    };
    const int k0 = T.bz * c.kz, k1 = min(k0 + c.kz, nz);
    Col64 below, cur;
    ld(cc, max(k0 - 1, 0), below);
    ld(cc, k0, cur);
    double loc[2] = {0.0, 0.0};  // sum Pmz_local, sum Psz_local (evolve.cu:269-276)
    for (int k = k0; k < k1; ++k) {
        const int sb = 0;
        // this plane's halo cell and the column's plane above: loaded here,
        // not a plane ahead (FP64 registers); other workgroups on the CU
        // cover the latency
        Col64 ring, above;
        ld(hc, k, ring);
        ld(cc, min(k + 1, nz - 1), above);
        const uint32_t cw = b.cw[cc + (long long)pb * k];
        if (k > k0) __syncthreads();  // the previous plane's reads are done
#pragma unroll
        for (int q = 0; q < 6; ++q) sP[sb][q][ty + 1][tx + 1] = cur.P[q];
        sphi[sb][ty + 1][tx + 1] = cur.phi;
        if (tp.has_h) {
#pragma unroll
            for (int q = 0; q < 6; ++q) sP[sb][q][tp.hy][tp.hx] = ring.P[q];
            sphi[sb][tp.hy][tp.hx] = ring.phi;
        }
        __syncthreads();
        if (tp.live) {
            const long long id = (long long)cc + (long long)pb * k;
            if (cw & CW_VOID) {  // never written by the reference: zero buffer
#pragma unroll
                for (int q = 0; q < 6; ++q) b.Pnew[q * N + id] = 0.0;
            } else {
                // neighbour sum in the reference order: e, w, n, s, u, d
                double sum[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const double(*S)[HX] = sP[sb][q];
                    sum[q] = S[ty + 1][tx + 2] + S[ty + 1][tx] + S[ty + 2][tx + 1] + S[ty][tx + 1] + above.P[q] +
                             below.P[q];
                }
                const PhaseC pc = sph[((cw >> CW_PHASE_SHIFT) & 3u) * 2u + ((cw & CW_IFACE) ? 1u : 0u)];
                const double* Tm = b.TR + (long long)(cw & CW_GID_MASK) * 9;
                const double tr11 = Tm[0], tr12 = Tm[1], tr13 = Tm[2];
                const double tr21 = Tm[3], tr22 = Tm[4], tr23 = Tm[5];
                const double tr31 = Tm[6], tr32 = Tm[7], tr33 = Tm[8];
                const double pMx0 = cur.P[0], pMy0 = cur.P[1], pMz0 = cur.P[2];
                const double pSx0 = cur.P[3], pSy0 = cur.P[4], pSz0 = cur.P[5];

                const double pMXl = tr11 * pMx0 + tr12 * pMy0 + tr13 * pMz0;
                const double pMYl = tr21 * pMx0 + tr22 * pMy0 + tr23 * pMz0;
                const double pMZl = tr31 * pMx0 + tr32 * pMy0 + tr33 * pMz0;
                const double pSXl = tr11 * pSx0 + tr12 * pSy0 + tr13 * pSz0;
                const double pSYl = tr21 * pSx0 + tr22 * pSy0 + tr23 * pSz0;
                const double pSZl = tr31 * pSx0 + tr32 * pSy0 + tr33 * pSz0;
                loc[0] += pMZl;
                loc[1] += pSZl;

                const double dfdPmxl = pc.a0_2 * pMXl;
                const double dfdPmyl = pc.a0_2 * pMYl;
                const double dfdPsxl = pc.a0_2 * pSXl;
                const double dfdPsyl = pc.a0_2 * pSYl;
                const double phiMz2 = pMZl * pMZl;
                const double phiMz3 = phiMz2 * pMZl;
                const double phiMz4 = phiMz3 * pMZl;
                const double phiMz5 = phiMz4 * pMZl;
                const double phiSz2 = pSZl * pSZl;
                const double phiSz3 = phiSz2 * pSZl;
                const double phiSz4 = phiSz3 * pSZl;
                const double phiSz5 = phiSz4 * pSZl;
                const double dfdPmzl = pc.ag * pMZl + pc.beta_4 * phiMz3 + pc.beta_12 * pMZl * phiSz2 +
                                       pc.gamma_6 * (phiMz5 + 10.0 * phiMz3 * phiSz2 + 5.0 * pMZl * phiSz4);
                const double dfdPszl = pc.ags * pSZl + pc.beta_4 * phiSz3 + pc.beta_12 * pSZl * phiMz2 +
                                       pc.gamma_6 * (phiSz5 + 10.0 * phiMz2 * phiSz3 + 5.0 * pSZl * phiMz4);
                double dfdPmx_b = tr11 * dfdPmxl + tr21 * dfdPmyl + tr31 * dfdPmzl;
                double dfdPmy_b = tr12 * dfdPmxl + tr22 * dfdPmyl + tr32 * dfdPmzl;
                double dfdPmz_b = tr13 * dfdPmxl + tr23 * dfdPmyl + tr33 * dfdPmzl;
                const double dfdPsx_b = tr11 * dfdPsxl + tr21 * dfdPsyl + tr31 * dfdPszl;
                const double dfdPsy_b = tr12 * dfdPsxl + tr22 * dfdPsyl + tr32 * dfdPszl;
                const double dfdPsz_b = tr13 * dfdPsxl + tr23 * dfdPsyl + tr33 * dfdPszl;

                double Ex = 0.0, Ey = 0.0, Ez = 0.0;
                if (c.dep_flag) {  // E = -grad phi, depol_varEPS.cu:246-273
                    const double(*F)[HX] = sphi[sb];
                    const double dx = -(F[ty + 1][tx + 2] - F[ty + 1][tx]);
                    const double dy = -(F[ty + 2][tx + 1] - F[ty][tx + 1]);
                    double up_v, dn_v;
                    if (c.zbc == 0) {
                        up_v = (k == nz - 1) ? c.g_top : above.phi;
                        dn_v = (k == 0) ? c.g_bot : below.phi;
                    } else {  // mirror: planes 1 and nz - 2 stand on both sides of the faces
                        up_v = (k == nz - 1) ? below.phi : above.phi;
                        dn_v = (k == 0) ? above.phi : below.phi;
                    }
                    const double dz = -(up_v - dn_v);
                    if (c.two_h_pow2) {
                        Ex = dx * c.inv_two_h;
                        Ey = dy * c.inv_two_h;
                        Ez = dz * c.inv_two_h;
                    } else {
                        Ex = dx / c.two_h;
                        Ey = dy / c.two_h;
                        Ez = dz / c.two_h;
                    }
                }
                const double dfdPmz_app = -Ez_ext;
                const double dfdPmx_e = c.dep_flag ? -1.0 * Ex : 0.0;
                const double dfdPmy_e = c.dep_flag ? -1.0 * Ey : 0.0;
                const double dfdPmz_e = c.dep_flag ? -1.0 * Ez : 0.0;
                dfdPmx_b += dfdPmx_e;
                dfdPmy_b += dfdPmy_e;
                dfdPmz_b += dfdPmz_app + dfdPmz_e;

                const double A = pc.A, om = pc.one_m6A, cL = pc.cL;
                b.Pnew[0 * N + id] = pMx0 * om + A * sum[0] - cL * dfdPmx_b;
                b.Pnew[1 * N + id] = pMy0 * om + A * sum[1] - cL * dfdPmy_b;
                b.Pnew[2 * N + id] = pMz0 * om + A * sum[2] - cL * dfdPmz_b;
                b.Pnew[3 * N + id] = pSx0 * om + A * sum[3] - cL * dfdPsx_b;
                b.Pnew[4 * N + id] = pSy0 * om + A * sum[4] - cL * dfdPsy_b;
                b.Pnew[5 * N + id] = pSz0 * om + A * sum[5] - cL * dfdPsz_b;
            }
        }
        below = cur;
        cur = above;
    }
    write_block_partials<TDGL_Q>(loc, b.tdgl_part, nb, T.blk);
}

// -(sum of harmonic-mean face fluxes)/h^2, depol_varEPS.cu:76-123
__device__ __forceinline__ double op_row(unsigned code, const double* H, const StepConsts& c,
                                         double xc, double xe, double xw, double xn, double xs,
                                         double xu, double xd) {
    const double* Hr = H + (code & 3u) * 4u;
    const double ex_p = Hr[(code >> 2) & 3u], ex_m = Hr[(code >> 4) & 3u];
    const double ey_p = Hr[(code >> 6) & 3u], ey_m = Hr[(code >> 8) & 3u];
    const double ez_p = Hr[(code >> 10) & 3u], ez_m = Hr[(code >> 12) & 3u];
    const double num = ex_p * (xe - xc) - ex_m * (xc - xw) + ey_p * (xn - xc) - ey_m * (xc - xs) +
                       ez_p * (xu - xc) - ez_m * (xc - xd);
    const double Ax = c.h2_pow2 ? num * c.inv_h2 : num / c.h2;
    return -Ax;
}

// ---------------------------------------------------------------- PCG setup
// rhs == nullptr: b = -eps0 div P (the forward solve); otherwise b = rhs
// (0 on identity rows), the adjoint solve of the same operator.
__global__ __launch_bounds__(NT) void k_pcg_setup_f64(Fp64Bufs b, StepConsts c, int ps, PcgState init,
                                                      const double* rhs) {
    __shared__ double sH[16];
    load_lut(sH, b.lut);
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const long long N = (long long)nx * ny * nz;
    const TileId T = tile_id();
    const int i = T.bx * BX + threadIdx.x;
    const int j = T.by * BY + threadIdx.y;
    const int k0 = T.bz * c.kz;
    double acc[3] = {0.0, 0.0, 0.0};  // r.z, r.r, b.b
    if (i < nx && j < ny) {
        const double* Px = b.Pcur;
        const double* Py = b.Pcur + N;
        const double* Pz = b.Pcur + 2 * N;
        const int k1 = min(k0 + c.kz, nz);
        for (int k = k0; k < k1; ++k) {
            const Nbr nb7 = op_nbrs(i, j, k, c);
            const unsigned code = b.sc[nb7.c];
            const double* x = b.phi;
            double bval, Ax;
            if (code & SC_IDENTITY) {
                bval = 0.0;  // rhs_zero_on_planes / gauge pin (depol_varEPS.cu:192-201, 404-414)
                Ax = op_identity(x[nb7.c], k, c);
            } else if (rhs) {
                bval = rhs[nb7.c];
                Ax = op_row(code, sH, c, x[nb7.c], x[nb7.e], x[nb7.w], x[nb7.n], x[nb7.s], x[nb7.u],
                            x[nb7.d]);
            } else {
                // divergence with clamped z (depol_varEPS.cu:293-325)
                const int uz = (k == nz - 1) ? k : k + 1;
                const int dz = (k == 0) ? k : k - 1;
                const double v = (Px[nb7.e] - Px[nb7.w]) + (Py[nb7.n] - Py[nb7.s]) +
                                 (Pz[vid(i, j, uz, nx, ny)] - Pz[vid(i, j, dz, nx, ny)]);
                double divp = c.two_h_pow2 ? v * c.inv_two_h : v / c.two_h;
                divp *= c.eps0;
                bval = -1.0 * divp;
                Ax = op_row(code, sH, c, x[nb7.c], x[nb7.e], x[nb7.w], x[nb7.n], x[nb7.s], x[nb7.u],
                            x[nb7.d]);
            }
            const double r = -1.0 * Ax + bval;
            const double z = r * dinv_of(code, sH, c);
            b.r[nb7.c] = r;
            acc[0] += r * z;
            acc[1] += r * r;
            acc[2] += bval * bval;
        }
    }
    write_block_partials<3>(acc, b.part[0], ps, T.blk);
    init.kind = K_SETUP;
    pass_state(init, b.st, T.lead);
}

// A x of one non-identity row from the paired face table, in the reference's
// operand order (depol_varEPS.cu:76-123): num / h^2, as a multiplication by
// the exact reciprocal when h^2 is a power of two (P2).
template <bool P2>
__device__ __forceinline__ double row_num(unsigned code, const HPair<double>* sHP, const StepConsts& c, double xc,
                                          double xe, double xw, double xn, double xs, double xu, double xd) {
    const unsigned cb = (code & 3u) << 4;
    const HPair<double> hx = sHP[cb | ((code >> 2) & 15u)];
    const HPair<double> hy = sHP[cb | ((code >> 6) & 15u)];
    const HPair<double> hz = sHP[cb | ((code >> 10) & 15u)];
    const double num = hx.p * (xe - xc) - hx.m * (xc - xw) + hy.p * (xn - xc) - hy.m * (xc - xs) +
                       hz.p * (xu - xc) - hz.m * (xc - xd);
    return P2 ? num * c.inv_h2 : num / c.h2;
}

// Scalars wave 0 hands to the rest of the workgroup after the loop test.
struct Bcast {
    double coef;  // beta (direction kernels) or alpha (update kernels)
    int flag;     // 0 continue, 1 done, 2 fold (FP32-refined)
};

// The direction kernels' loop test on the totals of the previous launch
// (setup: r.z, r.r, b.b; update: r.z, r.r) and beta (depol_varEPS.cu:459-476,
// 506-521).  A stopping solve keeps s.done = 1; a continuing one counts the
// iteration and flips the direction buffer.
__device__ __forceinline__ void dir_decide(PcgState& s, const double (&tot)[3], bool first) {
    if (first) {
        s.rz0 = s.rz_old = tot[0];
        s.rr0 = s.rr = tot[1];
        s.bb = tot[2];
        if (!finite(tot[0]) || tot[0] < 0.0) {  // :466-469
            s.status = HZO_PCG_BAD_INITIAL;
            s.done = 1;
        } else {
            judge(s, tot[0], tot[1]);
        }
    } else {
        s.rr = tot[1];
        if (!finite(tot[0]) || tot[0] < 0.0) {  // :508-513
            s.status = HZO_PCG_BAD_RZ;
            s.rz_old = tot[0];
            s.done = 1;
        } else {
            s.beta = tot[0] / s.rz_old;  // :515
            s.rz_old = tot[0];
            judge(s, tot[0], tot[1]);
        }
    }
    if (!s.done) {
        s.iters += 1;  // ++kiter (:476)
        s.pbuf ^= 1;
        s.kind = K_DIR;
    }
}

// The update kernels' breakdown test on p.Ap and alpha (:485-492).
__device__ __forceinline__ void update_decide(PcgState& s, double pap) {
    s.denom = pap;
    if (!finite(pap) || !(pap > 0.0)) {  // :486-490
        s.status = HZO_PCG_BREAKDOWN;
        s.done = 1;
    } else {
        s.alpha = s.rz_old / pap;  // :492
        s.kind = K_UPDATE;
    }
}

// ---------------------------------------------------------------- PCG direction + SpMV
// p_new = z + beta*p_old (first iteration: p_new = z), q = A p_new, p_new.q.
// z = D^-1 r is formed from r and the exact per-code D^-1 table wherever it
// is staged, so z is never stored.
//
// Issue order per workgroup: the face-pair table, the previous kernel's
// partials, the stencil codes, r and p_old of the KZ planes and of the two
// z-ghost planes (centre column only: ghost planes feed just the u/d
// neighbours), the x-y halo of the KZ planes, then the D^-1 gathers (which
// need only the codes).  Wave 0 finishes the reduction and applies the loop
// test while the tile loads are in flight and broadcasts beta through LDS.
// Plane indices are clamped into the film instead of branched on (values of
// clamped planes feed no live row) and masked stores use an out-of-range
// buffer offset, so the tile code is straight-line.  The thread's own
// column of p stays in registers; e/w/n/s come from LDS.
template <int KZ, int PM, bool P2>
__global__ __launch_bounds__(NT, 4) void k_pcg_dir_f64(Fp64Bufs b, StepConsts c, int nb, int ps, int t) {
    const int ip = (t - 1) & 1, io = t & 1;
    const TileId T = tile_id();
    PcgState s = load_state(b.st + ip);
    if (s.done) {
        pass_state(s, b.st + io, T.lead);
        return;
    }
    const bool first = s.kind == K_SETUP;
    __shared__ HPair<double> sHP[64];
    __shared__ double pl[KZ][HY][HX];
    __shared__ RawLDS<3> sraw;
    __shared__ Bcast bc;
    const int nz = c.nz;
    const int tid = linear_tid();
    
    const int k0 = T.bz * KZ;
    const TilePos tp = tile_pos(c, T);
    const TileCols G = tile_cols(tp, c);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const unsigned pb8 = c.plane * 8u, pb2 = c.plane * 2u;
    const rsrc_t Rr = make_rsrc(b.r, c.nvox * 8u);
    const rsrc_t Rpo = make_rsrc(s.pbuf ? b.p[1] : b.p[0], c.nvox * 8u);
    const rsrc_t Rpn = make_rsrc(s.pbuf ? b.p[0] : b.p[1], c.nvox * 8u);
    const rsrc_t Rsc = make_rsrc(b.sc, c.nvox * 2u);
    const rsrc_t Rdi = make_rsrc(b.lut + LUT_DINV, SC_LUT_SIZE * 8u);
    const double* pin = ip ? b.pin[1] : b.pin[0];

    const HPair<double> hpv = load_hpair(b.lut);
    Partials<3, PM> P;
    if (first) {
        load_partials<3, PM>(pin, b.nb_in, b.ps_in, P);
    } else {
        Partials<2, PM> Q2;
        load_partials<2, PM>(pin, b.nb_in, b.ps_in, Q2);
#pragma unroll
        for (int m = 0; m < (PM > 0 ? PM : 1); ++m) {
            P.v[0][m] = Q2.v[0][m];
            P.v[1][m] = Q2.v[1][m];
            P.v[2][m] = 0.0;
        }
    }
    unsigned so[KZ + 2];  // clamped plane of each staged slot
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) so[q] = (unsigned)min(max(k0 - 1 + q, 0), nz - 1);
    unsigned cc[KZ + 2], ch[KZ];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) cc[q] = bload_code(Rsc, G.cofs * 2u, so[q] * pb2);
#pragma unroll
    for (int q = 0; q < KZ; ++q) ch[q] = bload_code(Rsc, G.hofs * 2u, so[q + 1] * pb2);
    double rc[KZ + 2], pc[KZ + 2], rh[KZ], ph[KZ];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) {
        rc[q] = bload<double>(Rr, G.cofs * 8u, so[q] * pb8);
        pc[q] = bload<double>(Rpo, G.cofs * 8u, so[q] * pb8);
    }
#pragma unroll
    for (int q = 0; q < KZ; ++q) {
        rh[q] = bload<double>(Rr, G.hofs * 8u, so[q + 1] * pb8);
        ph[q] = bload<double>(Rpo, G.hofs * 8u, so[q + 1] * pb8);
    }
    double dc[KZ + 2], dh[KZ];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) dc[q] = bload<double>(Rdi, cc[q] * 8u, 0u);
#pragma unroll
    for (int q = 0; q < KZ; ++q) dh[q] = bload<double>(Rdi, ch[q] * 8u, 0u);

    stage_partials<3, PM>(P, sraw);
    __syncthreads();
    if (tid < 64) {
        // the previous kernel's inner products (setup: r.z, r.r, b.b; update:
        // r.z, r.r) and the loop test (depol_varEPS.cu:459-474, 506-521)
        sHP[tid] = hpv;
        double tot[3];
        wave0_total<3>(sraw, tot);
        dir_decide(s, tot, first);
        if (s.done) pass_state(s, b.st + io, T.lead);
        if (tid == 0) {
            bc.coef = s.beta;
            bc.flag = s.done;
        }
    }
    __syncthreads();
    if (bc.flag) return;
    const double beta = bc.coef;
    double pcol[KZ + 2];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) {
        const double z = rc[q] * dc[q];  // :502
        pcol[q] = first ? z : z + beta * pc[q];  // :518-519
        if (q >= 1 && q <= KZ) pl[q - 1][ty + 1][tx + 1] = pcol[q];
    }
#pragma unroll
    for (int q = 0; q < KZ; ++q) {
        const double zh = rh[q] * dh[q];
        const double v = first ? zh : zh + beta * ph[q];
        if (tp.has_h) pl[q][tp.hy][tp.hx] = v;
    }
    __syncthreads();
    const unsigned oob = c.nvox * 8u;  // store offset of masked lanes (dropped by the buffer unit)
    double acc[1] = {0.0};
#pragma unroll
    for (int q = 1; q <= KZ; ++q) {
        const int k = k0 - 1 + q;
        const double xc = pcol[q];
        const unsigned code = cc[q];
        double xu = pcol[q + 1], xd = pcol[q - 1];
        if (c.zbc != 0) {  // mirror ghosts (depol_varEPS.cu:94-98)
            if (k == 0) xd = xu;
            else if (k == nz - 1) xu = xd;
        }
        const double Ax = row_num<P2>(code, sHP, c, xc, pl[q - 1][ty + 1][tx + 2], pl[q - 1][ty + 1][tx],
                                      pl[q - 1][ty + 2][tx + 1], pl[q - 1][ty][tx + 1], xu, xd);
        const double qv = (code & SC_IDENTITY) ? op_identity(xc, k, c) : -Ax;
        const bool ok = tp.live && k < nz;
        bstore(Rpn, ok ? G.cofs * 8u : oob, so[q] * pb8, xc);
        if (ok) acc[0] += xc * qv;  // p.Ap (:485)
    }
    RawLDS<1>& sacc = *reinterpret_cast<RawLDS<1>*>(&sraw);
    write_block_partials<1>(acc, io ? b.part[1] : b.part[0], ps, T.blk, sacc);
    pass_state(s, b.st + io, T.lead);
}

// ---------------------------------------------------------------- PCG update
// x += alpha p, r -= alpha A p (A p re-evaluated from p instead of stored),
// z = D^-1 r, r.z and r.r (depol_varEPS.cu:492-507).
template <int KZ, int PM, bool P2>
__global__ __launch_bounds__(NT, 4) void k_pcg_update_f64(Fp64Bufs b, StepConsts c, int nb, int ps, int t) {
    const int ip = (t - 1) & 1, io = t & 1;
    const TileId T = tile_id();
    PcgState s = load_state(b.st + ip);
    if (s.done) {
        pass_state(s, b.st + io, T.lead);
        return;
    }
    __shared__ HPair<double> sHP[64];
    __shared__ double pl[KZ][HY][HX];
    __shared__ RawLDS<2> sraw;
    __shared__ Bcast bc;
    const int nz = c.nz;
    const int tid = linear_tid();
    
    const int k0 = T.bz * KZ;
    const TilePos tp = tile_pos(c, T);
    const TileCols G = tile_cols(tp, c);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const unsigned pb8 = c.plane * 8u, pb2 = c.plane * 2u;
    const rsrc_t Rp = make_rsrc(s.pbuf ? b.p[1] : b.p[0], c.nvox * 8u);
    const rsrc_t Rr = make_rsrc(b.r, c.nvox * 8u);
    const rsrc_t Rx = make_rsrc(b.phi, c.nvox * 8u);
    const rsrc_t Rsc = make_rsrc(b.sc, c.nvox * 2u);
    const rsrc_t Rdi = make_rsrc(b.lut + LUT_DINV, SC_LUT_SIZE * 8u);

    const HPair<double> hpv = load_hpair(b.lut);
    Partials<1, PM> P;
    load_partials<1, PM>(ip ? b.pin[1] : b.pin[0], b.nb_in, b.ps_in, P);
    unsigned so[KZ + 2];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) so[q] = (unsigned)min(max(k0 - 1 + q, 0), nz - 1);
    unsigned cv[KZ];
#pragma unroll
    for (int q = 0; q < KZ; ++q) cv[q] = bload_code(Rsc, G.cofs * 2u, so[q + 1] * pb2);
    double pcol[KZ + 2], ph[KZ];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) pcol[q] = bload<double>(Rp, G.cofs * 8u, so[q] * pb8);
#pragma unroll
    for (int q = 0; q < KZ; ++q) ph[q] = bload<double>(Rp, G.hofs * 8u, so[q + 1] * pb8);
    double xv[KZ], rv[KZ];
#pragma unroll
    for (int q = 0; q < KZ; ++q) {
        xv[q] = bload<double>(Rx, G.cofs * 8u, so[q + 1] * pb8);
        rv[q] = bload<double>(Rr, G.cofs * 8u, so[q + 1] * pb8);
    }
    double dv[KZ];
#pragma unroll
    for (int q = 0; q < KZ; ++q) dv[q] = bload<double>(Rdi, cv[q] * 8u, 0u);

    {
        RawLDS<1>& s1 = *reinterpret_cast<RawLDS<1>*>(&sraw);
        stage_partials<1, PM>(P, s1);
    }
    __syncthreads();
    if (tid < 64) {
        sHP[tid] = hpv;
        double tot[1];
        wave0_total<1>(*reinterpret_cast<RawLDS<1>*>(&sraw), tot);  // p.Ap (:485)
        update_decide(s, tot[0]);
        if (s.done) pass_state(s, b.st + io, T.lead);
        if (tid == 0) {
            bc.coef = s.alpha;
            bc.flag = s.done;
        }
    }
#pragma unroll
    for (int q = 0; q < KZ; ++q) pl[q][ty + 1][tx + 1] = pcol[q + 1];
    if (tp.has_h) {
#pragma unroll
        for (int q = 0; q < KZ; ++q) pl[q][tp.hy][tp.hx] = ph[q];
    }
    __syncthreads();
    if (bc.flag) return;
    const double alpha = bc.coef;
    const double nalpha = -alpha;
    const unsigned oob = c.nvox * 8u;
    double acc[2] = {0.0, 0.0};  // r.z, r.r
#pragma unroll
    for (int q = 1; q <= KZ; ++q) {
        const int k = k0 - 1 + q;
        const double xc = pcol[q];
        const unsigned code = cv[q - 1];
        double xu = pcol[q + 1], xd = pcol[q - 1];
        if (c.zbc != 0) {
            if (k == 0) xd = xu;
            else if (k == nz - 1) xu = xd;
        }
        const double Ax = row_num<P2>(code, sHP, c, xc, pl[q - 1][ty + 1][tx + 2], pl[q - 1][ty + 1][tx],
                                      pl[q - 1][ty + 2][tx + 1], pl[q - 1][ty][tx + 1], xu, xd);
        const double qv = (code & SC_IDENTITY) ? op_identity(xc, k, c) : -Ax;
        const bool ok = tp.live && k < nz;
        const unsigned vo = ok ? G.cofs * 8u : oob;
        bstore(Rx, vo, so[q] * pb8, alpha * xc + xv[q - 1]);  // Daxpy x += alpha p (:495)
        const double r = nalpha * qv + rv[q - 1];             // Daxpy r -= alpha Ap (:499)
        bstore(Rr, vo, so[q] * pb8, r);
        const double z = r * dv[q - 1];  // :502
        if (ok) {
            acc[0] += r * z;
            acc[1] += r * r;
        }
    }
    write_block_partials<2>(acc, io ? b.part[1] : b.part[0], ps, T.blk, sraw);
    pass_state(s, b.st + io, T.lead);
}

// ---------------------------------------------------------------- PCG, z-streaming deep tiles
// The direction and update kernels above stage all KZ planes of a tile in
// LDS at once, which limits them to shallow tiles: a large film then needs
// thousands of workgroups (8192 at 512x512x64, depth 4), each spending most
// of its life on the load-compute-store latency of its few planes, plus a
// one-workgroup fold of the partial sums after every launch.  These kernels
// run the same per-voxel arithmetic, operand for operand, over deep tiles
// (KZ 16 or 32, at most 2048 workgroups, so no fold) that walk their planes
// bottom to top.  Window plane m is z = k0 - 1 + m (m = 0 and KZ + 1 are the
// z-ghost planes, which feed only the centre column's u/d neighbours).
// Phase n forms plane n's operand of the stencil on the tile and its
// one-voxel x-y halo into LDS buffer n % 2, then evaluates the stencil of
// plane n - 1 (in-plane neighbours from buffer (n - 1) % 2, z neighbours
// from registers); one barrier ends each phase, so a buffer is rewritten
// only after every wave has finished reading it.  Loads run two planes
// ahead, stencil codes three.  Only the grouping of the inner products'
// partial sums follows the tile depth.

// Loads run F64Z_LA planes ahead of their use, stencil codes F64Z_CL (two
// and three planes; three and five measured the same at 512x512x64).
constexpr int F64Z_LA = 2, F64Z_CL = 3;

// A streamed FP64 load of the tile's own column: NTL (films whose PCG
// vectors exceed the last-level cache) with the nontemporal hint, which took
// the 512x512x64 film's direction and update launches from 97.3 and 120.2 to
// 93.9 and 114.8 us; on films that fit the cache the hint cost 20-25 %.
template <bool NTL>
__device__ __forceinline__ double bload_s(rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, NTL ? 2 : 0));
}

// D^-1 of a row from the paired face table in LDS: the operations of
// dinv_of and of the host's per-code table (depol_varEPS.cu:177-179), so the
// same double.  The direction kernel forms D^-1 of its tile and halo cells
// this way instead of gathering them from the 16385-entry table: 122.9 ->
// 99.4 us per launch at 512x512x64 (the update kernel, which needs D^-1 of
// the centre only, keeps the gather: forming it there took 128 -> 443 us).
template <bool P2>
__device__ __forceinline__ double dinv_pairs(unsigned code, const HPair<double>* sHP, const StepConsts& c) {
    const unsigned cb = (code & 3u) << 4;
    const HPair<double> hx = sHP[cb | ((code >> 2) & 15u)];
    const HPair<double> hy = sHP[cb | ((code >> 6) & 15u)];
    const HPair<double> hz = sHP[cb | ((code >> 10) & 15u)];
    const double sum = hx.p + hx.m + hy.p + hy.m + hz.p + hz.m;
    const double D = P2 ? sum * c.inv_h2 : sum / c.h2;
    const double d = D > 1e-30 ? 1.0 / D : 1.0;
    return (code & SC_IDENTITY) ? 1.0 : d;
}

// p_new = z + beta p_old, q = A p_new, p_new.q (as k_pcg_dir_f64).
template <int KZ, int PM, bool P2, bool NTL>
__global__ __launch_bounds__(NT, 4) void k_pcg_dir_f64z(Fp64Bufs b, StepConsts c, int nb, int ps, int t) {
    constexpr int NW = KZ + 2;
    const int ip = (t - 1) & 1, io = t & 1;
    const TileId T = tile_id();
    __shared__ HPair<double> sHP[64];
    __shared__ double pl[2][HY][HX];
    __shared__ RawLDS<3> sraw;
    __shared__ Bcast bc;
    const int nz = c.nz;
    const int tid = linear_tid();
    const int k0 = T.bz * KZ;
    const TilePos tp = tile_pos(c, T);
    const TileCols G = tile_cols(tp, c);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const unsigned pb8 = c.plane * 8u, pb2 = c.plane * 2u;
    const unsigned cv = G.cofs * 8u, hv = G.hofs * 8u;
    const rsrc_t Rr = make_rsrc(b.r, c.nvox * 8u);
    const rsrc_t Rsc = make_rsrc(b.sc, c.nvox * 2u);
    const double* pin = ip ? b.pin[1] : b.pin[0];

    // What does not depend on the control block goes out first: the
    // face-pair table, the partials (all three columns; after an update
    // launch the third is not used), the first planes' codes and r.
    const HPair<double> hpv = load_hpair(b.lut);
    Partials<3, PM> P;
    load_partials<3, PM>(pin, b.nb_in, b.ps_in, P);
    // Plane z of window plane m, clamped into the film, formed from a copy of
    // k0 that each phase renews (an empty asm that claims to change it): the
    // loads of a phase then cannot be hoisted above it, which would hold the
    // whole window's loads in registers.
    int k0p = k0;
    auto so = [&](int m) { return (unsigned)min(max(k0p - 1 + m, 0), nz - 1); };
    unsigned cc[NW], ch[NW];
    auto load_code = [&](int m) {
        cc[m] = bload_code(Rsc, G.cofs * 2u, so(m) * pb2);
        ch[m] = (m == 0 || m == NW - 1) ? 0u : bload_code(Rsc, G.hofs * 2u, so(m) * pb2);
    };
    double rc[NW], rh[NW], pc_old[NW], ph_old[NW], dc[NW], dh[NW];
    auto load_r = [&](int m) {
        rc[m] = bload_s<NTL>(Rr, cv, so(m) * pb8);
        rh[m] = (m == 0 || m == NW - 1) ? 0.0 : bload<double>(Rr, hv, so(m) * pb8);
    };
#pragma unroll
    for (int m = 0; m < F64Z_CL; ++m) load_code(m);
#pragma unroll
    for (int m = 0; m < F64Z_LA; ++m) load_r(m);

    PcgState s = load_state(b.st + ip);
    if (s.done) {
        keep_live(hpv.p);
        keep_live_partials(P);
#pragma unroll
        for (int m = 0; m < F64Z_LA; ++m) {
            keep_live(rc[m]);
            keep_live(rh[m]);
        }
#pragma unroll
        for (int m = F64Z_LA; m < F64Z_CL; ++m) {
            keep_live(cc[m]);
            keep_live(ch[m]);
        }
        pass_state(s, b.st + io, T.lead);
        return;
    }
    const bool first = s.kind == K_SETUP;
    const rsrc_t Rpo = make_rsrc(s.pbuf ? b.p[1] : b.p[0], c.nvox * 8u);
    const rsrc_t Rpn = make_rsrc(s.pbuf ? b.p[0] : b.p[1], c.nvox * 8u);
    auto load_p = [&](int m) {
        pc_old[m] = first ? 0.0 : bload_s<NTL>(Rpo, cv, so(m) * pb8);
        ph_old[m] = (first || m == 0 || m == NW - 1) ? 0.0 : bload<double>(Rpo, hv, so(m) * pb8);
    };
#pragma unroll
    for (int m = 0; m < F64Z_LA; ++m) load_p(m);

    stage_partials<3, PM>(P, sraw);
    __syncthreads();
    if (tid < 64) {
        sHP[tid] = hpv;
        double tot[3];
        wave0_total<3>(sraw, tot);
        dir_decide(s, tot, first);
        if (s.done) pass_state(s, b.st + io, T.lead);
        if (tid == 0) {
            bc.coef = s.beta;
            bc.flag = s.done;
        }
    }
    __syncthreads();
    if (bc.flag) return;
    const double beta = bc.coef;
    const unsigned oob = c.nvox * 8u;
    double pc[NW];
    double acc[1] = {0.0};
#pragma unroll
    for (int n = 0; n < NW; ++n) {
        k0p = k0;
        asm volatile("" : "+s"(k0p));
        if (n + F64Z_CL < NW) load_code(n + F64Z_CL);
        if (n + F64Z_LA < NW) {
            load_r(n + F64Z_LA);
            load_p(n + F64Z_LA);
        }
        // plane n: p_new on the tile (and, on tile planes, its halo)
        dc[n] = dinv_pairs<P2>(cc[n], sHP, c);
        if (n >= 1 && n <= KZ) dh[n] = dinv_pairs<P2>(ch[n], sHP, c);
        const double z = rc[n] * dc[n];  // :502
        pc[n] = first ? z : z + beta * pc_old[n];  // :518-519
        if (n >= 1 && n <= KZ) {
            pl[n & 1][ty + 1][tx + 1] = pc[n];
            const double zh = rh[n] * dh[n];
            const double v = first ? zh : zh + beta * ph_old[n];
            if (tp.has_h) pl[n & 1][tp.hy][tp.hx] = v;
            const bool ok = tp.live && k0 - 1 + n < nz;
            bstore(Rpn, ok ? cv : oob, so(n) * pb8, pc[n]);
        }
        // plane n - 1: q = A p_new and p_new.q
        if (n >= 2) {
            const int m = n - 1, k = k0 - 1 + m;
            const double xc = pc[m];
            const unsigned code = cc[m];
            double xu = pc[m + 1], xd = pc[m - 1];
            if (c.zbc != 0) {  // mirror ghosts (depol_varEPS.cu:94-98)
                if (k == 0) xd = xu;
                else if (k == nz - 1) xu = xd;
            }
            const double(*S)[HX] = pl[m & 1];
            const double Ax = row_num<P2>(code, sHP, c, xc, S[ty + 1][tx + 2], S[ty + 1][tx], S[ty + 2][tx + 1],
                                          S[ty][tx + 1], xu, xd);
            const double qv = (code & SC_IDENTITY) ? op_identity(xc, k, c) : -Ax;
            if (tp.live && k < nz) acc[0] += xc * qv;  // p.Ap (:485)
            // the sum must exist here: otherwise the compiler keeps every
            // phase's LDS operands (in scratch) and evaluates all stencils
            // after the loop, where their only use is
            asm volatile("" : "+v"(acc[0]));
        }
        if (n >= 1 && n + 1 < NW) __syncthreads();
    }
    RawLDS<1>& sacc = *reinterpret_cast<RawLDS<1>*>(&sraw);
    write_block_partials<1>(acc, io ? b.part[1] : b.part[0], ps, T.blk, sacc);
    pass_state(s, b.st + io, T.lead);
}

// x += alpha p, r -= alpha A p, z = D^-1 r, r.z and r.r (as k_pcg_update_f64).
template <int KZ, int PM, bool P2, bool NTL>
__global__ __launch_bounds__(NT, 4) void k_pcg_update_f64z(Fp64Bufs b, StepConsts c, int nb, int ps, int t) {
    constexpr int NW = KZ + 2;
    const int ip = (t - 1) & 1, io = t & 1;
    const TileId T = tile_id();
    __shared__ HPair<double> sHP[64];
    __shared__ double pl[2][HY][HX];
    __shared__ RawLDS<2> sraw;
    __shared__ Bcast bc;
    const int nz = c.nz;
    const int tid = linear_tid();
    const int k0 = T.bz * KZ;
    const TilePos tp = tile_pos(c, T);
    const TileCols G = tile_cols(tp, c);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const unsigned pb8 = c.plane * 8u, pb2 = c.plane * 2u;
    const unsigned cv = G.cofs * 8u, hv = G.hofs * 8u;
    const rsrc_t Rr = make_rsrc(b.r, c.nvox * 8u);
    const rsrc_t Rx = make_rsrc(b.phi, c.nvox * 8u);
    const rsrc_t Rsc = make_rsrc(b.sc, c.nvox * 2u);
    const rsrc_t Rdi = make_rsrc(b.lut + LUT_DINV, SC_LUT_SIZE * 8u);

    const HPair<double> hpv = load_hpair(b.lut);
    Partials<1, PM> P;
    load_partials<1, PM>(ip ? b.pin[1] : b.pin[0], b.nb_in, b.ps_in, P);
    // Plane z of window plane m, clamped into the film, formed from a copy of
    // k0 that each phase renews (an empty asm that claims to change it): the
    // loads of a phase then cannot be hoisted above it, which would hold the
    // whole window's loads in registers.
    int k0p = k0;
    auto so = [&](int m) { return (unsigned)min(max(k0p - 1 + m, 0), nz - 1); };
    // the centre column's code, D^-1, x and r of tile planes
    unsigned cc[NW];
    double dc[NW], xv[NW], rv[NW];
    auto load_code = [&](int m) {
        if (m >= 1 && m <= KZ) cc[m] = bload_code(Rsc, G.cofs * 2u, so(m) * pb2);
    };
    auto load_xr = [&](int m) {
        if (m >= 1 && m <= KZ) {
            dc[m] = bload<double>(Rdi, cc[m] * 8u, 0u);
            xv[m] = bload_s<NTL>(Rx, cv, so(m) * pb8);
            rv[m] = bload_s<NTL>(Rr, cv, so(m) * pb8);
        }
    };
#pragma unroll
    for (int m = 0; m < F64Z_CL; ++m) load_code(m);
#pragma unroll
    for (int m = 0; m < F64Z_LA; ++m) load_xr(m);

    PcgState s = load_state(b.st + ip);
    if (s.done) {
        keep_live(hpv.p);
        keep_live_partials(P);
#pragma unroll
        for (int m = 1; m < F64Z_LA; ++m) {
            keep_live(dc[m]);
            keep_live(xv[m]);
            keep_live(rv[m]);
        }
#pragma unroll
        for (int m = F64Z_LA; m < F64Z_CL; ++m)
            if (m >= 1 && m <= KZ) keep_live(cc[m]);
        pass_state(s, b.st + io, T.lead);
        return;
    }
    const rsrc_t Rp = make_rsrc(s.pbuf ? b.p[1] : b.p[0], c.nvox * 8u);
    double pc[NW], ph[NW];
    auto load_p = [&](int m) {
        pc[m] = bload_s<NTL>(Rp, cv, so(m) * pb8);
        ph[m] = (m == 0 || m == NW - 1) ? 0.0 : bload<double>(Rp, hv, so(m) * pb8);
    };
#pragma unroll
    for (int m = 0; m < F64Z_LA; ++m) load_p(m);

    {
        RawLDS<1>& s1 = *reinterpret_cast<RawLDS<1>*>(&sraw);
        stage_partials<1, PM>(P, s1);
    }
    __syncthreads();
    if (tid < 64) {
        sHP[tid] = hpv;
        double tot[1];
        wave0_total<1>(*reinterpret_cast<RawLDS<1>*>(&sraw), tot);  // p.Ap (:485)
        update_decide(s, tot[0]);
        if (s.done) pass_state(s, b.st + io, T.lead);
        if (tid == 0) {
            bc.coef = s.alpha;
            bc.flag = s.done;
        }
    }
    __syncthreads();
    if (bc.flag) return;
    const double alpha = bc.coef;
    const double nalpha = -alpha;
    const unsigned oob = c.nvox * 8u;
    double acc[2] = {0.0, 0.0};  // r.z, r.r
#pragma unroll
    for (int n = 0; n < NW; ++n) {
        k0p = k0;
        asm volatile("" : "+s"(k0p));
        if (n + F64Z_CL < NW) load_code(n + F64Z_CL);
        if (n + F64Z_LA < NW) {
            load_xr(n + F64Z_LA);
            load_p(n + F64Z_LA);
        }
        if (n >= 1 && n <= KZ) {
            pl[n & 1][ty + 1][tx + 1] = pc[n];
            if (tp.has_h) pl[n & 1][tp.hy][tp.hx] = ph[n];
        }
        if (n >= 2) {
            const int m = n - 1, k = k0 - 1 + m;
            const double xc = pc[m];
            const unsigned code = cc[m];
            double xu = pc[m + 1], xd = pc[m - 1];
            if (c.zbc != 0) {
                if (k == 0) xd = xu;
                else if (k == nz - 1) xu = xd;
            }
            const double(*S)[HX] = pl[m & 1];
            const double Ax = row_num<P2>(code, sHP, c, xc, S[ty + 1][tx + 2], S[ty + 1][tx], S[ty + 2][tx + 1],
                                          S[ty][tx + 1], xu, xd);
            const double qv = (code & SC_IDENTITY) ? op_identity(xc, k, c) : -Ax;
            const bool ok = tp.live && k < nz;
            const unsigned vo = ok ? cv : oob;
            bstore(Rx, vo, so(m) * pb8, alpha * xc + xv[m]);  // Daxpy x += alpha p (:495)
            const double r = nalpha * qv + rv[m];               // Daxpy r -= alpha Ap (:499)
            bstore(Rr, vo, so(m) * pb8, r);
            const double z = r * dc[m];  // :502
            if (ok) {
                acc[0] += r * z;
                acc[1] += r * r;
            }
        }
        if (n >= 1 && n + 1 < NW) __syncthreads();
    }
    write_block_partials<2>(acc, io ? b.part[1] : b.part[0], ps, T.blk, sraw);
    pass_state(s, b.st + io, T.lead);
}

// ---------------------------------------------------------------- diagnostics
// out8 = sum Pmz, sum Psz, then FE, DE, AFE: sum(mask*Pmz), sum(mask)
// (evolve.cu:259-266, 280-357; build_mat_mask_and_p kernels_evol.cu:23-44).
__global__ __launch_bounds__(NT) void k_diag_f64(const double* __restrict__ P,
                                                  const uint32_t* __restrict__ cw, Grid g, double* part,
                                                  int nb) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long N = g.N;
    const TileId T = tile_id();
    const int i = T.bx * BX + threadIdx.x;
    const int j = T.by * BY + threadIdx.y;
    const int k0 = T.bz * g.kz;
    // sum Pmz, sum Psz, FE: mask*Pmz, mask; DE: mask*Pmz, mask; AFE: mask*Pmz, mask
    double a[DIAG_Q] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < nx && j < ny) {
        const int k1 = min(k0 + g.kz, nz);
        for (int k = k0; k < k1; ++k) {
            const long long id = vid(i, j, k, nx, ny);
            const double pmz = P[2 * N + id];
            a[0] += pmz;
            a[1] += P[5 * N + id];
            const uint32_t w = cw[id];
            const int phase = (w & CW_VOID) ? -1 : (int)((w >> CW_PHASE_SHIFT) & 3u);
            const double mfe = (phase == 1) ? 1.0 : 0.0;
            const double mde = (phase == 2) ? 1.0 : 0.0;
            const double mafe = (phase == 0) ? 1.0 : 0.0;
            a[2] += mfe * pmz;
            a[3] += mfe;
            a[4] += mde * pmz;
            a[5] += mde;
            a[6] += mafe * pmz;
            a[7] += mafe;
        }
    }
    write_block_partials<DIAG_Q>(a, part, nb, T.blk);
}

// ---------------------------------------------------------------- launchers
void launch_tdgl_f64(hipStream_t s, const Grid& g, const StepConsts& c, double Ez_ext, const Fp64Bufs& b,
                  const LaunchEv& ev) {
    hipExtLaunchKernelGGL(k_tdgl_f64, grid3(g), dim3(BX, BY, 1), 0, s, ev.start, ev.stop, 0, b, c, Ez_ext, blocks_for(g));
}

void launch_pcg_setup_f64(hipStream_t s, const Grid& g, const StepConsts& c, const Fp64Bufs& b,
                          const PcgState& init, const double* rhs,
                  const LaunchEv& ev) {
    hipExtLaunchKernelGGL(k_pcg_setup_f64, grid3(g), dim3(BX, BY, 1), 0, s, ev.start, ev.stop, 0, b, c, part_stride(g), init,
                          rhs);
}

// Tile depths 8, 16 and 32 run the z-streaming kernels.
template <bool P2>
static void launch_dir64(hipStream_t s, const Grid& g, const StepConsts& c, const Fp64Bufs& b, int t, const LaunchEv& ev) {
    if (g.kz >= 8) {
        const int pm = part_pm(b.nb_in);
        const dim3 gr = grid3(g), bl(BX, BY, 1);
        if (b.stream_loads) {
#define HZO_X , P2, true
            if (g.kz == 32) { HZO_PCG_LAUNCH_PM(k_pcg_dir_f64z, 32, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
            else if (g.kz == 16) { HZO_PCG_LAUNCH_PM(k_pcg_dir_f64z, 16, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
            else { HZO_PCG_LAUNCH_PM(k_pcg_dir_f64z, 8, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
#undef HZO_X
        } else {
#define HZO_X , P2, false
            if (g.kz == 32) { HZO_PCG_LAUNCH_PM(k_pcg_dir_f64z, 32, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
            else if (g.kz == 16) { HZO_PCG_LAUNCH_PM(k_pcg_dir_f64z, 16, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
            else { HZO_PCG_LAUNCH_PM(k_pcg_dir_f64z, 8, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
#undef HZO_X
        }
        return;
    }
#define HZO_X , P2
    HZO_PCG_LAUNCH_X(k_pcg_dir_f64, g, b.nb_in, s, b, c, blocks_for(g), part_stride(g), t);
#undef HZO_X
}
template <bool P2>
static void launch_update64(hipStream_t s, const Grid& g, const StepConsts& c, const Fp64Bufs& b, int t, const LaunchEv& ev) {
    if (g.kz >= 8) {
        const int pm = part_pm(b.nb_in);
        const dim3 gr = grid3(g), bl(BX, BY, 1);
        if (b.stream_loads) {
#define HZO_X , P2, true
            if (g.kz == 32) { HZO_PCG_LAUNCH_PM(k_pcg_update_f64z, 32, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
            else if (g.kz == 16) { HZO_PCG_LAUNCH_PM(k_pcg_update_f64z, 16, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
            else { HZO_PCG_LAUNCH_PM(k_pcg_update_f64z, 8, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
#undef HZO_X
        } else {
#define HZO_X , P2, false
            if (g.kz == 32) { HZO_PCG_LAUNCH_PM(k_pcg_update_f64z, 32, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
            else if (g.kz == 16) { HZO_PCG_LAUNCH_PM(k_pcg_update_f64z, 16, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
            else { HZO_PCG_LAUNCH_PM(k_pcg_update_f64z, 8, pm, gr, bl, s, b, c, blocks_for(g), part_stride(g), t); }
#undef HZO_X
        }
        return;
    }
#define HZO_X , P2
    HZO_PCG_LAUNCH_X(k_pcg_update_f64, g, b.nb_in, s, b, c, blocks_for(g), part_stride(g), t);
#undef HZO_X
}

#undef HZO_X

void launch_pcg_dir_f64(hipStream_t s, const Grid& g, const StepConsts& c, const Fp64Bufs& b, int t,
                  const LaunchEv& ev) {
    if (c.h2_pow2) launch_dir64<true>(s, g, c, b, t, ev);
    else launch_dir64<false>(s, g, c, b, t, ev);
}

void launch_pcg_update_f64(hipStream_t s, const Grid& g, const StepConsts& c, const Fp64Bufs& b, int t,
                  const LaunchEv& ev) {
    if (c.h2_pow2) launch_update64<true>(s, g, c, b, t, ev);
    else launch_update64<false>(s, g, c, b, t, ev);
}

void launch_diag_f64(hipStream_t s, const Grid& g, const double* P, const uint32_t* cw, double* part,
                  const LaunchEv& ev) {
    hipExtLaunchKernelGGL(k_diag_f64, grid3(g), dim3(BX, BY, 1), 0, s, ev.start, ev.stop, 0, P, cw, g, part, blocks_for(g));
}

// The partials were written by the previous launch, mostly on other XCDs, so
// each load is a round trip to memory: every thread issues FOLD_U loads per
// column before it adds any, instead of one dependent load per step.
constexpr int FOLD_U = 8;

__global__ __launch_bounds__(NT) void k_fold_partials(const double* __restrict__ in, int nb, int ps,
                                                      double* __restrict__ out) {
    __shared__ WaveLDS<RED_MAXQ> W;
    const int tid = linear_tid();
    double a[RED_MAXQ];
#pragma unroll
    for (int q = 0; q < RED_MAXQ; ++q) a[q] = 0.0;
    for (int t0 = tid; t0 < nb; t0 += FOLD_U * NT) {
        double v[RED_MAXQ][FOLD_U];
#pragma unroll
        for (int q = 0; q < RED_MAXQ; ++q)
#pragma unroll
            for (int u = 0; u < FOLD_U; ++u) {
                const int t = t0 + u * NT;
                v[q][u] = t < nb ? in[(long long)q * ps + t] : 0.0;
            }
#pragma unroll
        for (int q = 0; q < RED_MAXQ; ++q)
#pragma unroll
            for (int u = 0; u < FOLD_U; ++u) a[q] += v[q][u];
    }
#pragma unroll
    for (int q = 0; q < RED_MAXQ; ++q) {
        const double w = wave_sum(a[q]);
        if ((tid & 63) == 0) W.v[q][tid >> 6] = w;
    }
    __syncthreads();
    if (tid == 0) {
        double tot[RED_MAXQ];
        wave_lds_total<RED_MAXQ>(W, tot);
#pragma unroll
        for (int q = 0; q < RED_MAXQ; ++q) out[q * NT] = tot[q];
    }
}

void launch_fold_partials(hipStream_t s, const double* in, int nb, int ps, double* out) {
    hipLaunchKernelGGL(k_fold_partials, dim3(1), dim3(BX, BY, 1), 0, s, in, nb, ps, out);
}

}  // namespace hzo
