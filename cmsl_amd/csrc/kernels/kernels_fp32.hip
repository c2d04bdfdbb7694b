// kernels_fp32.hip -- FP32-refined form of the HZO step on CDNA4 (gfx950).
//
// Polarisation, TDGL arithmetic and the CG vectors (r, p, correction x_lo)
// are FP32; FP64 is kept for the inner products (FP32 products are exact in
// FP64 and accumulated there), for the right-hand side b and the potential
// x_hi, and for the residual replacement r = b - A (x_hi) that runs every
// `refine_every` iterations and whenever the recursively updated residual
// claims convergence (reliable updates, Sleijpen & van der Vorst 1996).
// The potential is always phi = x_hi + x_lo.
//
// The iteration is the same Jacobi PCG as the FP64 form (depol_varEPS.cu:
// 442-523); the launch pattern is one (dir, update) pair per host step and
// the device decides, from the control block, whether a pair is a CG
// iteration or a residual replacement (fold x_lo into x_hi, recompute r).
//
// This file is compiled with -ffp-contract=fast: FP32 kernels may fuse.
#include "../hzo_internal.h"
#include "common.hip.h"
#include "tile.hip.h"
#include "reduce.hip.h"

namespace hzo {

// ---------------------------------------------------------------- TDGL
// z-streaming: a 64x8 workgroup walks its kz planes bottom to top.  Each
// plane's P (6 components) and potential phi = x_hi + x_lo are staged in LDS
// with a one-voxel x-y halo (the in-plane neighbours of the Laplacian and
// of E = -grad phi), double-buffered so one barrier per plane suffices; a
// thread keeps its own column's values of the planes below and above in
// registers (the z neighbours, clamped at the film's faces as the reference
// clamps them), and the next plane's loads, including its voxels' grain
// rotations, are issued before the current plane is evaluated.  The phase
// table sits in LDS.  Per voxel the arithmetic and its order are those of
// kernels_evol.cu:47-242.
struct Col32 {
    float P[6];
    double phi;
};

__device__ __forceinline__ void tdgl_f32_body(const Fp32Bufs& b, const StepConsts& c, float Ez_ext, int nb,
                                              const TileId& T) {
    __shared__ float sP[2][6][HY][HX];
    __shared__ double sphi[2][HY][HX];
    __shared__ PhaseC32 sph[PTAB_ENTRIES];
    const int nz = c.nz;
    const long long N = (long long)c.nx * c.ny * nz, plane = (long long)c.nx * c.ny;
    const TilePos tp = tile_pos(c, T);
    const TileCols G = tile_cols(tp, c);
    const int tx = threadIdx.x, ty = threadIdx.y, tid = linear_tid();
    if (tid < PTAB_ENTRIES) sph[tid] = b.ptab32[tid];
    const long long cc = G.cofs, hc = tp.has_h ? G.hofs : G.cofs;
    // buffer loads: one resource per array (each below 2^32 bytes), the
    // column as a 32-bit lane offset, the plane as a uniform offset
    rsrc_t RP[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) RP[q] = make_rsrc(b.Pold + q * N, c.nvox * 4u);
    const rsrc_t Rxh = make_rsrc(b.xhi, c.nvox * 8u), Rxl = make_rsrc(b.xlo, c.nvox * 4u);
    const unsigned pb = c.plane;
    auto ld = [&](long long col, int k, Col32& v) {
        const unsigned o = (unsigned)col, po = (unsigned)k * pb;
#pragma unroll
        for (int q = 0; q < 6; ++q) v.P[q] = bload<float>(RP[q], o * 4u, po * 4u);
        v.phi = c.dep_flag ? bload<double>(Rxh, o * 8u, po * 8u) + (double)bload<float>(Rxl, o * 4u, po * 4u) : 0.0;
    };
    const int k0 = T.bz * c.kz, k1 = min(k0 + c.kz, nz);
    Col32 below, cur, above, ring;
    ld(cc, max(k0 - 1, 0), below);
    ld(cc, k0, cur);
    ld(cc, min(k0 + 1, nz - 1), above);
    ld(hc, k0, ring);
    uint32_t cw = b.cw[cc + plane * k0];
    float Tn[9];
    {
        const float* Tr = b.TR32 + (long long)(cw & CW_GID_MASK) * 9;
#pragma unroll
        for (int e = 0; e < 9; ++e) Tn[e] = Tr[e];
    }
    double loc[2] = {0.0, 0.0};
    for (int k = k0; k < k1; ++k) {
        const int sb = (k - k0) & 1;
#pragma unroll
        for (int q = 0; q < 6; ++q) sP[sb][q][ty + 1][tx + 1] = cur.P[q];
        sphi[sb][ty + 1][tx + 1] = cur.phi;
        if (tp.has_h) {
#pragma unroll
            for (int q = 0; q < 6; ++q) sP[sb][q][tp.hy][tp.hx] = ring.P[q];
            sphi[sb][tp.hy][tp.hx] = ring.phi;
        }
        // the next plane: the column two planes up and its cell word now, its
        // halo cell (an L2 hit, loaded after this plane's evaluation so its
        // registers are free during it).  The tile's last plane loads the
        // ghost plane above the tile again (unused) rather than a plane
        // beyond it.
        Col32 above_n;
        const int kn = min(k + 1, k1 - 1);
        ld(cc, min(k + 2, min(k1, nz - 1)), above_n);
        const uint32_t cw_n = b.cw[cc + plane * kn];
        float T9[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) T9[e] = Tn[e];
        __syncthreads();
        if (tp.live) {
            const long long id = cc + plane * k;
            if (cw & CW_VOID) {
#pragma unroll
                for (int q = 0; q < 6; ++q) b.Pnew[q * N + id] = 0.0f;
            } else {
                float sum[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const float(*S)[HX] = sP[sb][q];
                    sum[q] = S[ty + 1][tx + 2] + S[ty + 1][tx] + S[ty + 2][tx + 1] + S[ty][tx + 1] + above.P[q] +
                             below.P[q];
                }
                const PhaseC32 pc = sph[((cw >> CW_PHASE_SHIFT) & 3u) * 2u + ((cw & CW_IFACE) ? 1u : 0u)];
                const float t11 = T9[0], t12 = T9[1], t13 = T9[2], t21 = T9[3], t22 = T9[4], t23 = T9[5];
                const float t31 = T9[6], t32 = T9[7], t33 = T9[8];
                const float mx = cur.P[0], my = cur.P[1], mz = cur.P[2], sx = cur.P[3], sy = cur.P[4], sz = cur.P[5];
                const float mXl = t11 * mx + t12 * my + t13 * mz;
                const float mYl = t21 * mx + t22 * my + t23 * mz;
                const float mZl = t31 * mx + t32 * my + t33 * mz;
                const float sXl = t11 * sx + t12 * sy + t13 * sz;
                const float sYl = t21 * sx + t22 * sy + t23 * sz;
                const float sZl = t31 * sx + t32 * sy + t33 * sz;
                loc[0] += (double)mZl;
                loc[1] += (double)sZl;
                const float m2 = mZl * mZl, m3 = m2 * mZl, m4 = m3 * mZl, m5 = m4 * mZl;
                const float s2 = sZl * sZl, s3 = s2 * sZl, s4 = s3 * sZl, s5 = s4 * sZl;
                const float fmz = pc.ag * mZl + pc.beta_4 * m3 + pc.beta_12 * mZl * s2 +
                                  pc.gamma_6 * (m5 + 10.0f * m3 * s2 + 5.0f * mZl * s4);
                const float fsz = pc.ags * sZl + pc.beta_4 * s3 + pc.beta_12 * sZl * m2 +
                                  pc.gamma_6 * (s5 + 10.0f * m2 * s3 + 5.0f * sZl * m4);
                const float fmx = pc.a0_2 * mXl, fmy = pc.a0_2 * mYl, fsx = pc.a0_2 * sXl, fsy = pc.a0_2 * sYl;
                float dmx = t11 * fmx + t21 * fmy + t31 * fmz;
                float dmy = t12 * fmx + t22 * fmy + t32 * fmz;
                float dmz = t13 * fmx + t23 * fmy + t33 * fmz;
                const float dsx = t11 * fsx + t21 * fsy + t31 * fsz;
                const float dsy = t12 * fsx + t22 * fsy + t32 * fsz;
                const float dsz = t13 * fsx + t23 * fsy + t33 * fsz;
                if (c.dep_flag) {  // E = -grad(x_hi + x_lo)
                    const double(*F)[HX] = sphi[sb];
                    double up_v, dn_v;
                    if (c.zbc == 0) {
                        up_v = (k == nz - 1) ? c.g_top : above.phi;
                        dn_v = (k == 0) ? c.g_bot : below.phi;
                    } else {  // mirror: planes 1 and nz - 2 stand on both sides of the faces
                        up_v = (k == nz - 1) ? below.phi : above.phi;
                        dn_v = (k == 0) ? above.phi : below.phi;
                    }
                    const float Ex = (float)(-(F[ty + 1][tx + 2] - F[ty + 1][tx]) * c.inv_two_h);
                    const float Ey = (float)(-(F[ty + 2][tx + 1] - F[ty][tx + 1]) * c.inv_two_h);
                    const float Ez = (float)(-(up_v - dn_v) * c.inv_two_h);
                    dmx -= Ex;
                    dmy -= Ey;
                    dmz -= Ez;
                }
                dmz -= Ez_ext;
                const float A = pc.A, om = pc.one_m6A, cL = pc.cL;
                b.Pnew[0 * N + id] = mx * om + A * sum[0] - cL * dmx;
                b.Pnew[1 * N + id] = my * om + A * sum[1] - cL * dmy;
                b.Pnew[2 * N + id] = mz * om + A * sum[2] - cL * dmz;
                b.Pnew[3 * N + id] = sx * om + A * sum[3] - cL * dsx;
                b.Pnew[4 * N + id] = sy * om + A * sum[4] - cL * dsy;
                b.Pnew[5 * N + id] = sz * om + A * sum[5] - cL * dsz;
            }
        }
        ld(hc, kn, ring);
        below = cur;
        cur = above;
        above = above_n;
        cw = cw_n;
        {
            const float* Tr = b.TR32 + (long long)(cw & CW_GID_MASK) * 9;
#pragma unroll
            for (int e = 0; e < 9; ++e) Tn[e] = Tr[e];
        }
    }
    write_block_partials<TDGL_Q>(loc, b.tdgl_part, nb, T.blk);
}

__global__ __launch_bounds__(NT, 4) void k_tdgl_f32(Fp32Bufs b, StepConsts c, float Ez_ext, int nb) {
    tdgl_f32_body(b, c, Ez_ext, nb, tile_id());
}

// The same launch with the applied field read from device memory (*ez), so
// one captured launch serves every step of a run.
__global__ __launch_bounds__(NT, 4) void k_tdgl_f32_dev(Fp32Bufs b, StepConsts c, const double* __restrict__ ez,
                                                     int nb) {
    tdgl_f32_body(b, c, (float)*ez, nb, tile_id());
}

__global__ __launch_bounds__(NT, 4) void k_tdgl_f32_batch(const BatchFilm* __restrict__ films, const int2* __restrict__ tiles,
                                                       int parity) {
    TileId T;
    const BatchFilm& F = films[batch_tile(tiles, films, T)];
    tdgl_f32_body(F.b[parity], F.c, F.Ez, F.nb, T);
}

// ---------------------------------------------------------------- operator rows
template <class T>
__device__ __forceinline__ T row_Ax(unsigned code, const T* H, const StepConsts& c, T xc, T xe, T xw, T xn,
                                    T xs, T xu, T xd) {
    const T* Hr = H + (code & 3u) * 4u;
    const T num = Hr[(code >> 2) & 3u] * (xe - xc) - Hr[(code >> 4) & 3u] * (xc - xw) +
                  Hr[(code >> 6) & 3u] * (xn - xc) - Hr[(code >> 8) & 3u] * (xc - xs) +
                  Hr[(code >> 10) & 3u] * (xu - xc) - Hr[(code >> 12) & 3u] * (xc - xd);
    return -(num * (T)c.inv_h2);
}

// FP32 row from the paired face table: -(sum_d H_d (x_d - x_c)) / h^2.
__device__ __forceinline__ float row_Ax_pairs(unsigned code, const HPair<float>* sHP, float ih2, float xc,
                                              float xe, float xw, float xn, float xs, float xu, float xd) {
    const unsigned cb = (code & 3u) << 4;
    const HPair<float> hx = sHP[cb | ((code >> 2) & 15u)];
    const HPair<float> hy = sHP[cb | ((code >> 6) & 15u)];
    const HPair<float> hz = sHP[cb | ((code >> 10) & 15u)];
    const float num = hx.p * (xe - xc) - hx.m * (xc - xw) + hy.p * (xn - xc) - hy.m * (xc - xs) +
                      hz.p * (xu - xc) - hz.m * (xc - xd);
    return -(num * ih2);
}

// r = b - A x with x = x_hi (+ x_lo), all in FP64; shared by the PCG setup and
// the residual replacement.
__device__ __forceinline__ double true_residual(const Fp32Bufs& b, const double* sH, const StepConsts& c,
                                                const Nbr& v, unsigned code, int k, double bval,
                                                bool with_lo) {
    auto X = [&](long long q) { return with_lo ? b.xhi[q] + (double)b.xlo[q] : b.xhi[q]; };
    double Ax;
    if (code & SC_IDENTITY)
        Ax = op_identity(X(v.c), k, c);
    else
        Ax = row_Ax<double>(code, sH, c, X(v.c), X(v.e), X(v.w), X(v.n), X(v.s), X(v.u), X(v.d));
    return bval - Ax;
}

// ---------------------------------------------------------------- PCG setup
// rhs == nullptr: b = -eps0 div P (the forward solve); otherwise b = rhs
// (0 on identity rows), the adjoint solve of the same operator.
__device__ __forceinline__ void setup_f32_body(const Fp32Bufs& b, const StepConsts& c, int ps, PcgState init,
                                               const double* rhs, const TileId& T) {
    __shared__ double sH[16];
    load_lut(sH, b.lut);
    const float* dinv = b.lut32 + LUT_DINV;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const long long N = (long long)nx * ny * nz;
    const int i = T.bx * BX + threadIdx.x;
    const int j = T.by * BY + threadIdx.y;
    const int k0 = T.bz * c.kz;
    double acc[3] = {0.0, 0.0, 0.0};
    if (i < nx && j < ny) {
        const float* Px = b.Pcur;
        const float* Py = b.Pcur + N;
        const float* Pz = b.Pcur + 2 * N;
        const int k1 = min(k0 + c.kz, nz);
        for (int k = k0; k < k1; ++k) {
            const Nbr v = op_nbrs(i, j, k, c);
            const unsigned code = b.sc[v.c];
            double bval = 0.0;
            if (rhs) {
                if (!(code & SC_IDENTITY)) bval = rhs[v.c];
            } else if (!(code & SC_IDENTITY)) {
                const int uz = (k == nz - 1) ? k : k + 1;
                const int dz = (k == 0) ? k : k - 1;
                const double dv = ((double)Px[v.e] - (double)Px[v.w]) + ((double)Py[v.n] - (double)Py[v.s]) +
                                  ((double)Pz[vid(i, j, uz, nx, ny)] - (double)Pz[vid(i, j, dz, nx, ny)]);
                bval = -(dv * c.inv_two_h * c.eps0);
            }
            b.b[v.c] = bval;
            const float r = (float)true_residual(b, sH, c, v, code, k, bval, true);
            const float z = r * dinv[code];
            b.r[v.c] = r;
            cg_dot_add<true>(acc[0], r, z, c, k);
            acc[1] += (double)r * (double)r;
            acc[2] += bval * bval;
        }
    }
    write_block_partials<3>(acc, b.part[0], ps, T.blk);
    init.kind = K_SETUP;
    pass_state(init, b.st, T.lead);
}

__global__ __launch_bounds__(NT) void k_pcg_setup_f32(Fp32Bufs b, StepConsts c, int ps, PcgState init,
                                                      const double* rhs) {
    setup_f32_body(b, c, ps, init, rhs, tile_id());
}

__global__ __launch_bounds__(NT) void k_pcg_setup_f32_batch(const BatchFilm* __restrict__ films, const int2* __restrict__ tiles,
                                                            int parity) {
    TileId T;
    const BatchFilm& F = films[batch_tile(tiles, films, T)];
    setup_f32_body(F.b[parity], F.c, F.ps, F.init, nullptr, T);
}

// Scalars wave 0 hands to the rest of the workgroup after the loop test.
struct Bcast {
    double coef;  // beta (direction kernels) or alpha (update kernels)
    int flag;     // 0 continue, 1 done, 2 fold
};

// ---------------------------------------------------------------- direction / fold
// Finishes the previous kernel's inner products and decides what this launch
// is: a CG direction update (K_DIR: p = z + beta p, q = A p, p.q) or the fold
// x_hi += x_lo that starts a residual replacement (K_FOLD).  Load order, the
// wave-0 control logic and the tile staging are those of the FP64 form; D^-1
// is the FP32 rounding of the exact per-code table, the same matrix in every
// FP32 kernel.
template <int KZ, int PM>
__global__ __launch_bounds__(NT, 4) void k_pcg_dir_f32(Fp32Bufs b, StepConsts c, int nb, int ps, int t) {
    const int ip = (t - 1) & 1, io = t & 1;
    const TileId T = tile_id();
    PcgState s = load_state(b.st + ip);
    if (s.done) {
        pass_state(s, b.st + io, T.lead);
        return;
    }
    const bool first = s.kind == K_SETUP;
    __shared__ HPair<float> sHP[64];
    __shared__ float pl[KZ][HY][HX];
    __shared__ RawLDS<3> sraw;
    __shared__ Bcast bc;
    const int nz = c.nz;
    const int tid = linear_tid();
    
    const int k0 = T.bz * KZ;
    const TilePos tp = tile_pos(c, T);
    const TileCols G = tile_cols(tp, c);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const unsigned pb4 = c.plane * 4u, pb2 = c.plane * 2u;
    const rsrc_t Rr = make_rsrc(b.r, c.nvox * 4u);
    const rsrc_t Rpo = make_rsrc(s.pbuf ? b.p[1] : b.p[0], c.nvox * 4u);
    const rsrc_t Rpn = make_rsrc(s.pbuf ? b.p[0] : b.p[1], c.nvox * 4u);
    const rsrc_t Rsc = make_rsrc(b.sc, c.nvox * 2u);
    const rsrc_t Rdi = make_rsrc(b.lut32 + LUT_DINV, SC_LUT_SIZE * 4u);
    const double* pin = ip ? b.pin[1] : b.pin[0];

    const HPair<float> hpv = load_hpair(b.lut32);
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
    unsigned so[KZ + 2];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) so[q] = (unsigned)min(max(k0 - 1 + q, 0), nz - 1);
    unsigned cc[KZ + 2], ch[KZ];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) cc[q] = bload_code(Rsc, G.cofs * 2u, so[q] * pb2);
#pragma unroll
    for (int q = 0; q < KZ; ++q) ch[q] = bload_code(Rsc, G.hofs * 2u, so[q + 1] * pb2);
    float rc[KZ + 2], pc[KZ + 2], rh[KZ], ph[KZ];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) {
        rc[q] = bload<float>(Rr, G.cofs * 4u, so[q] * pb4);
        pc[q] = bload<float>(Rpo, G.cofs * 4u, so[q] * pb4);
    }
#pragma unroll
    for (int q = 0; q < KZ; ++q) {
        rh[q] = bload<float>(Rr, G.hofs * 4u, so[q + 1] * pb4);
        ph[q] = bload<float>(Rpo, G.hofs * 4u, so[q + 1] * pb4);
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
        bool fold = false;
        if (first) {
            s.rz0 = s.rz_old = tot[0];
            s.rr0 = s.rr = tot[1];
            s.bb = tot[2];
            if (!finite(tot[0]) || tot[0] < 0.0) {
                s.status = HZO_PCG_BAD_INITIAL;
                s.done = 1;
            } else {
                judge(s, tot[0], tot[1]);
            }
        } else {
            s.rr = tot[1];
            if (!finite(tot[0]) || tot[0] < 0.0) {
                s.status = HZO_PCG_BAD_RZ;
                s.rz_old = tot[0];
                s.done = 1;
            } else if (s.kind == K_RESID) {  // true residual: trust the test
                s.refinements += 1;
                s.beta = tot[0] / s.rz_old;
                s.rz_old = tot[0];
                judge(s, tot[0], tot[1]);
            } else {  // K_UPDATE: recursively updated residual
                if (s.iters >= s.max_iter) {
                    s.rz_old = tot[0];
                    s.done = 1;
                    s.status = HZO_PCG_MAXITER;
                } else if (!keep_going(s, tot[0], tot[1]) ||
                           (s.refine_every > 0 && s.iters % s.refine_every == 0)) {
                    fold = true;  // replace the residual before trusting or continuing it
                } else {
                    s.beta = tot[0] / s.rz_old;
                    s.rz_old = tot[0];
                }
            }
        }
        if (s.done) {
            pass_state(s, b.st + io, T.lead);
        } else if (fold) {
            s.kind = K_FOLD;
            pass_state(s, b.st + io, T.lead);
        } else {
            s.iters += 1;
            s.pbuf ^= 1;
            s.kind = K_DIR;
        }
        if (tid == 0) {
            bc.coef = s.beta;
            bc.flag = s.done ? 1 : (fold ? 2 : 0);
        }
    }
    __syncthreads();
    const int flag = bc.flag;
    if (flag == 1) return;
    if (flag == 2) {
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
        return;
    }
    const float beta = (float)bc.coef;
    float pcol[KZ + 2];
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) {
        const float z = rc[q] * dc[q];
        pcol[q] = first ? z : z + beta * pc[q];
        if (q >= 1 && q <= KZ) pl[q - 1][ty + 1][tx + 1] = pcol[q];
    }
#pragma unroll
    for (int q = 0; q < KZ; ++q) {
        const float zh = rh[q] * dh[q];
        const float v = first ? zh : zh + beta * ph[q];
        if (tp.has_h) pl[q][tp.hy][tp.hx] = v;
    }
    __syncthreads();
    const float ih2 = (float)c.inv_h2;
    const unsigned oob = c.nvox * 4u;  // store offset of masked lanes (dropped by the buffer unit)
    double acc[1] = {0.0};
#pragma unroll
    for (int q = 1; q <= KZ; ++q) {
        const int k = k0 - 1 + q;
        const float xc = pcol[q];
        const unsigned code = cc[q];
        float xu = pcol[q + 1], xd = pcol[q - 1];
        if (c.zbc != 0) {
            if (k == 0) xd = xu;
            else if (k == nz - 1) xu = xd;
        }
        const float qs = row_Ax_pairs(code, sHP, ih2, xc, pl[q - 1][ty + 1][tx + 2], pl[q - 1][ty + 1][tx],
                                      pl[q - 1][ty + 2][tx + 1], pl[q - 1][ty][tx + 1], xu, xd);
        const float qv = (code & SC_IDENTITY) ? op_identity(xc, k, c) : qs;
        const bool ok = tp.live && k < nz;
        bstore(Rpn, ok ? G.cofs * 4u : oob, so[q] * pb4, xc);
        if (ok) cg_dot_add<true>(acc[0], xc, qv, c, k);
    }
    RawLDS<1>& sacc = *reinterpret_cast<RawLDS<1>*>(&sraw);
    write_block_partials<1>(acc, io ? b.part[1] : b.part[0], ps, T.blk, sacc);
    pass_state(s, b.st + io, T.lead);
}

// ---------------------------------------------------------------- update / residual replacement
template <int KZ, int PM>
__global__ __launch_bounds__(NT, 4) void k_pcg_update_f32(Fp32Bufs b, StepConsts c, int nb, int ps, int t) {
    const int ip = (t - 1) & 1, io = t & 1;
    const TileId T = tile_id();
    PcgState s = load_state(b.st + ip);
    if (s.done) {
        pass_state(s, b.st + io, T.lead);
        return;
    }
    __shared__ RawLDS<2> sraw;
    const bool resid = s.kind == K_FOLD;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    
    const int k0 = T.bz * KZ;
    double acc[2] = {0.0, 0.0};
    if (resid) {  // r = b - A x_hi in FP64 (x_lo was folded by the previous launch)
        __shared__ double sH64[16];
        load_lut(sH64, b.lut);
        const float* dinv = b.lut32 + LUT_DINV;
        const int i = T.bx * BX + threadIdx.x;
        const int j = T.by * BY + threadIdx.y;
        if (i < nx && j < ny) {
            const int k1 = min(k0 + KZ, nz);
            for (int k = k0; k < k1; ++k) {
                const Nbr v = op_nbrs(i, j, k, c);
                const unsigned code = b.sc[v.c];
                const float r = (float)true_residual(b, sH64, c, v, code, k, b.b[v.c], false);
                b.r[v.c] = r;
                const float z = r * dinv[code];
                cg_dot_add<true>(acc[0], r, z, c, k);
                acc[1] += (double)r * (double)r;
            }
        }
        s.kind = K_RESID;
    } else {
        __shared__ HPair<float> sHP[64];
        __shared__ float pl[KZ][HY][HX];
        __shared__ Bcast bc;
        const int tid = linear_tid();
        const TilePos tp = tile_pos(c, T);
        const TileCols G = tile_cols(tp, c);
        const int tx = threadIdx.x, ty = threadIdx.y;
        const unsigned pb4 = c.plane * 4u, pb2 = c.plane * 2u;
        const rsrc_t Rp = make_rsrc(s.pbuf ? b.p[1] : b.p[0], c.nvox * 4u);
        const rsrc_t Rr = make_rsrc(b.r, c.nvox * 4u);
        const rsrc_t Rx = make_rsrc(b.xlo, c.nvox * 4u);
        const rsrc_t Rsc = make_rsrc(b.sc, c.nvox * 2u);
        const rsrc_t Rdi = make_rsrc(b.lut32 + LUT_DINV, SC_LUT_SIZE * 4u);

        const HPair<float> hpv = load_hpair(b.lut32);
        Partials<1, PM> P;
        load_partials<1, PM>(ip ? b.pin[1] : b.pin[0], b.nb_in, b.ps_in, P);
        unsigned so[KZ + 2];
#pragma unroll
        for (int q = 0; q < KZ + 2; ++q) so[q] = (unsigned)min(max(k0 - 1 + q, 0), nz - 1);
        unsigned cv[KZ];
#pragma unroll
        for (int q = 0; q < KZ; ++q) cv[q] = bload_code(Rsc, G.cofs * 2u, so[q + 1] * pb2);
        float pcol[KZ + 2], ph[KZ];
#pragma unroll
        for (int q = 0; q < KZ + 2; ++q) pcol[q] = bload<float>(Rp, G.cofs * 4u, so[q] * pb4);
#pragma unroll
        for (int q = 0; q < KZ; ++q) ph[q] = bload<float>(Rp, G.hofs * 4u, so[q + 1] * pb4);
        float xv[KZ], rv[KZ];
#pragma unroll
        for (int q = 0; q < KZ; ++q) {
            xv[q] = bload<float>(Rx, G.cofs * 4u, so[q + 1] * pb4);
            rv[q] = bload<float>(Rr, G.cofs * 4u, so[q + 1] * pb4);
        }
        float dv[KZ];
#pragma unroll
        for (int q = 0; q < KZ; ++q) dv[q] = bload<float>(Rdi, cv[q] * 4u, 0u);

        {
            RawLDS<1>& s1 = *reinterpret_cast<RawLDS<1>*>(&sraw);
            stage_partials<1, PM>(P, s1);
        }
        __syncthreads();
        if (tid < 64) {
            sHP[tid] = hpv;
            double tot[1];
            wave0_total<1>(*reinterpret_cast<RawLDS<1>*>(&sraw), tot);
            s.denom = tot[0];
            if (!finite(tot[0]) || !(tot[0] > 0.0)) {
                s.status = HZO_PCG_BREAKDOWN;
                s.done = 1;
                pass_state(s, b.st + io, T.lead);
            } else {
                s.alpha = s.rz_old / tot[0];
            }
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
        const float alpha = (float)bc.coef;
        const float ih2 = (float)c.inv_h2;
        const unsigned oob = c.nvox * 4u;
#pragma unroll
        for (int q = 1; q <= KZ; ++q) {
            const int k = k0 - 1 + q;
            const float xc = pcol[q];
            const unsigned code = cv[q - 1];
            float xu = pcol[q + 1], xd = pcol[q - 1];
            if (c.zbc != 0) {
                if (k == 0) xd = xu;
                else if (k == nz - 1) xu = xd;
            }
            const float qs = row_Ax_pairs(code, sHP, ih2, xc, pl[q - 1][ty + 1][tx + 2], pl[q - 1][ty + 1][tx],
                                          pl[q - 1][ty + 2][tx + 1], pl[q - 1][ty][tx + 1], xu, xd);
            const float qv = (code & SC_IDENTITY) ? op_identity(xc, k, c) : qs;
            const bool ok = tp.live && k < nz;
            const unsigned vo = ok ? G.cofs * 4u : oob;
            bstore(Rx, vo, so[q] * pb4, xv[q - 1] + alpha * xc);
            const float r = rv[q - 1] - alpha * qv;
            bstore(Rr, vo, so[q] * pb4, r);
            const float z = r * dv[q - 1];
            if (ok) {
                cg_dot_add<true>(acc[0], r, z, c, k);
                acc[1] += (double)r * (double)r;
            }
        }
        s.kind = K_UPDATE;
    }
    write_block_partials<2>(acc, io ? b.part[1] : b.part[0], ps, T.blk, sraw);
    pass_state(s, b.st + io, T.lead);
}

// ---------------------------------------------------------------- diagnostics
__device__ __forceinline__ void diag_f32_body(const float* __restrict__ P, const uint32_t* __restrict__ cw,
                                              const Grid& g, double* part, int nb, const TileId& T) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long N = g.N;
    const int i = T.bx * BX + threadIdx.x;
    const int j = T.by * BY + threadIdx.y;
    const int k0 = T.bz * g.kz;
    double a[DIAG_Q] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < nx && j < ny) {
        const int k1 = min(k0 + g.kz, nz);
        for (int k = k0; k < k1; ++k) {
            const long long id = vid(i, j, k, nx, ny);
            const double pmz = P[2 * N + id];
            a[0] += pmz;
            a[1] += (double)P[5 * N + id];
            const uint32_t w = cw[id];
            const int phase = (w & CW_VOID) ? -1 : (int)((w >> CW_PHASE_SHIFT) & 3u);
            a[2] += (phase == 1) ? pmz : 0.0;
            a[3] += (phase == 1) ? 1.0 : 0.0;
            a[4] += (phase == 2) ? pmz : 0.0;
            a[5] += (phase == 2) ? 1.0 : 0.0;
            a[6] += (phase == 0) ? pmz : 0.0;
            a[7] += (phase == 0) ? 1.0 : 0.0;
        }
    }
    write_block_partials<DIAG_Q>(a, part, nb, T.blk);
}

__global__ __launch_bounds__(NT) void k_diag_f32(const float* __restrict__ P, const uint32_t* __restrict__ cw,
                                                  Grid g, double* part, int nb) {
    diag_f32_body(P, cw, g, part, nb, tile_id());
}

__global__ __launch_bounds__(NT) void k_diag_f32_batch(const BatchFilm* __restrict__ films, const int2* __restrict__ tiles,
                                                       int parity) {
    TileId T;
    const BatchFilm& F = films[batch_tile(tiles, films, T)];
    diag_f32_body(F.b[parity].Pcur, F.b[parity].cw, F.g, F.diag_part, F.nb, T);
}

void launch_tdgl_f32(hipStream_t s, const Grid& g, const StepConsts& c, double Ez_ext, const Fp32Bufs& b,
                  const LaunchEv& ev) {
    hipExtLaunchKernelGGL(k_tdgl_f32, grid3(g), dim3(BX, BY, 1), 0, s, ev.start, ev.stop, 0, b, c, (float)Ez_ext, blocks_for_impl(g));
}
void launch_tdgl_f32_dev(hipStream_t s, const Grid& g, const StepConsts& c, const double* ez, const Fp32Bufs& b) {
    hipLaunchKernelGGL(k_tdgl_f32_dev, grid3(g), dim3(BX, BY, 1), 0, s, b, c, ez, blocks_for_impl(g));
}
void launch_pcg_setup_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b,
                          const PcgState& init, const double* rhs,
                  const LaunchEv& ev) {
    hipExtLaunchKernelGGL(k_pcg_setup_f32, grid3(g), dim3(BX, BY, 1), 0, s, ev.start, ev.stop, 0, b, c, part_stride(g), init,
                          rhs);
}
#define HZO_X
void launch_pcg_dir_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b, int t,
                  const LaunchEv& ev) {
    HZO_PCG_LAUNCH_X(k_pcg_dir_f32, g, b.nb_in, s, b, c, blocks_for(g), part_stride(g), t);
}
void launch_pcg_update_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b, int t,
                  const LaunchEv& ev) {
    HZO_PCG_LAUNCH_X(k_pcg_update_f32, g, b.nb_in, s, b, c, blocks_for(g), part_stride(g), t);
}
#undef HZO_X
void launch_tdgl_f32_batch(hipStream_t s, int ntiles, const BatchFilm* films, const int2* tiles, int parity) {
    hipLaunchKernelGGL(k_tdgl_f32_batch, dim3(ntiles), dim3(BX, BY, 1), 0, s, films, tiles, parity);
}
void launch_setup_f32_batch(hipStream_t s, int ntiles, const BatchFilm* films, const int2* tiles, int parity) {
    hipLaunchKernelGGL(k_pcg_setup_f32_batch, dim3(ntiles), dim3(BX, BY, 1), 0, s, films, tiles, parity);
}
void launch_diag_f32_batch(hipStream_t s, int ntiles, const BatchFilm* films, const int2* tiles, int parity) {
    hipLaunchKernelGGL(k_diag_f32_batch, dim3(ntiles), dim3(BX, BY, 1), 0, s, films, tiles, parity);
}
void launch_diag_f32(hipStream_t s, const Grid& g, const float* P, const uint32_t* cw, double* part,
                  const LaunchEv& ev) {
    hipExtLaunchKernelGGL(k_diag_f32, grid3(g), dim3(BX, BY, 1), 0, s, ev.start, ev.stop, 0, P, cw, g, part, blocks_for_impl(g));
}

}  // namespace hzo
