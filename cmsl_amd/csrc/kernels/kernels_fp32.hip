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
__global__ __launch_bounds__(NT) void k_tdgl_f32(Fp32Bufs b, StepConsts c, float Ez_ext, int nb) {
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const long long N = (long long)nx * ny * nz;
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k0 = blockIdx.z * c.kz;
    double loc[2] = {0.0, 0.0};
    if (i < nx && j < ny) {
        const int e = (i + 1 == nx) ? 0 : i + 1;
        const int w = (i == 0) ? nx - 1 : i - 1;
        const int n = (j + 1 == ny) ? 0 : j + 1;
        const int s = (j == 0) ? ny - 1 : j - 1;
        const int k1 = min(k0 + c.kz, nz);
        for (int k = k0; k < k1; ++k) {
            const long long id = vid(i, j, k, nx, ny);
            const uint32_t cw = b.cw[id];
            if (cw & CW_VOID) {
#pragma unroll
                for (int q = 0; q < 6; ++q) b.Pnew[q * N + id] = 0.0f;
                continue;
            }
            const int u = (k == nz - 1) ? k : k + 1;
            const int d = (k == 0) ? k : k - 1;
            const long long ie = vid(e, j, k, nx, ny), iw = vid(w, j, k, nx, ny);
            const long long in = vid(i, n, k, nx, ny), is = vid(i, s, k, nx, ny);
            const long long iu = vid(i, j, u, nx, ny), idn = vid(i, j, d, nx, ny);
            float sum[6], ctr[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const float* P = b.Pold + q * N;
                sum[q] = P[ie] + P[iw] + P[in] + P[is] + P[iu] + P[idn];
                ctr[q] = P[id];
            }
            const PhaseC32 pc = b.ptab32[((cw >> CW_PHASE_SHIFT) & 3u) * 2u + ((cw & CW_IFACE) ? 1u : 0u)];
            const float* T = b.TR32 + (long long)(cw & CW_GID_MASK) * 9;
            const float t11 = T[0], t12 = T[1], t13 = T[2], t21 = T[3], t22 = T[4], t23 = T[5];
            const float t31 = T[6], t32 = T[7], t33 = T[8];
            const float mx = ctr[0], my = ctr[1], mz = ctr[2], sx = ctr[3], sy = ctr[4], sz = ctr[5];
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
                auto phi = [&](long long v) { return b.xhi[v] + (double)b.xlo[v]; };
                double up_v, dn_v;
                if (c.zbc == 0) {
                    up_v = (k == nz - 1) ? c.g_top : phi(vid(i, j, k + 1, nx, ny));
                    dn_v = (k == 0) ? c.g_bot : phi(vid(i, j, k - 1, nx, ny));
                } else {
                    const int uu = (k == 0) ? 1 : ((k == nz - 1) ? nz - 2 : k + 1);
                    const int dd = (k == 0) ? 1 : ((k == nz - 1) ? nz - 2 : k - 1);
                    up_v = phi(vid(i, j, uu, nx, ny));
                    dn_v = phi(vid(i, j, dd, nx, ny));
                }
                const float Ex = (float)(-(phi(ie) - phi(iw)) * c.inv_two_h);
                const float Ey = (float)(-(phi(in) - phi(is)) * c.inv_two_h);
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
    write_block_partials<TDGL_Q>(loc, b.tdgl_part, nb);
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
__global__ __launch_bounds__(NT) void k_pcg_setup_f32(Fp32Bufs b, StepConsts c, int nb, PcgState init) {
    __shared__ double sH[16];
    __shared__ float sH32[16];
    load_lut(sH, b.lut);
    load_lut(sH32, b.lut32);
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const long long N = (long long)nx * ny * nz;
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k0 = blockIdx.z * c.kz;
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
            if (!(code & SC_IDENTITY)) {
                const int uz = (k == nz - 1) ? k : k + 1;
                const int dz = (k == 0) ? k : k - 1;
                const double dv = ((double)Px[v.e] - (double)Px[v.w]) + ((double)Py[v.n] - (double)Py[v.s]) +
                                  ((double)Pz[vid(i, j, uz, nx, ny)] - (double)Pz[vid(i, j, dz, nx, ny)]);
                bval = -(dv * c.inv_two_h * c.eps0);
            }
            b.b[v.c] = bval;
            const float r = (float)true_residual(b, sH, c, v, code, k, bval, true);
            const float z = r * dinv_of(code, sH32, c);
            b.r[v.c] = r;
            acc[0] += (double)r * (double)z;
            acc[1] += (double)r * (double)r;
            acc[2] += bval * bval;
        }
    }
    write_block_partials<3>(acc, b.part[0], nb);
    init.kind = K_SETUP;
    pass_state(init, b.st);
}

// ---------------------------------------------------------------- direction / fold
// Finishes the previous kernel's inner products and decides what this launch
// is: a CG direction update (K_DIR) or the fold x_hi += x_lo that starts a
// residual replacement (K_FOLD).  Loads of all KZ+2 planes of the tile are
// issued before the reduction prologue (one memory round trip per workgroup).
template <int KZ>
__global__ __launch_bounds__(NT) void k_pcg_dir_f32(Fp32Bufs b, StepConsts c, int nb, int t) {
    const int ip = (t - 1) & 1, io = t & 1;
    PcgState s = load_state(b.st + ip);
    if (s.done) {
        pass_state(s, b.st + io);
        return;
    }
    const bool first = s.kind == K_SETUP;
    __shared__ float sH[16];
    __shared__ float pl[KZ + 2][HY][HX];
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const int k0 = blockIdx.z * KZ;
    const TilePos tp = tile_pos(c);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const float* pold = b.p[s.pbuf];
    float* pnew = b.p[s.pbuf ^ 1];
    float rc[KZ + 2], pc[KZ + 2], rh[KZ + 2], ph[KZ + 2];
    unsigned cc[KZ + 2], ch[KZ + 2];
    const long long plane = (long long)nx * ny;
    const long long cbase = vid(tp.ci, tp.cj, k0 - 1, nx, ny), hbase = vid(tp.gi, tp.gj, k0 - 1, nx, ny);
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) {
        const int kk = k0 - 1 + q;
        rc[q] = pc[q] = rh[q] = ph[q] = 0.f;
        cc[q] = ch[q] = SC_IDENTITY;
        if (kk >= 0 && kk < nz) {
            const long long id = cbase + q * plane;
            rc[q] = b.r[id];
            cc[q] = b.sc[id];
            if (!first) pc[q] = pold[id];
            if (tp.has_h) {
                const long long h = hbase + q * plane;
                rh[q] = b.r[h];
                ch[q] = b.sc[h];
                if (!first) ph[q] = pold[h];
            }
        }
    }
    bool fold = false;
    if (first) {
        double tot[3];
        sum_partials<3>(b.part[ip], nb, tot);
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
        double tot[2];
        sum_partials<2>(b.part[ip], nb, tot);
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
        pass_state(s, b.st + io);
        return;
    }
    if (fold) {
        if (tp.live) {
#pragma unroll
            for (int q = 0; q < KZ; ++q) {
                const int k = k0 + q;
                if (k < nz) {
                    const long long id = vid(tp.i, tp.j, k, nx, ny);
                    b.xhi[id] += (double)b.xlo[id];
                    b.xlo[id] = 0.0f;
                }
            }
        }
        s.kind = K_FOLD;
        pass_state(s, b.st + io);
        return;
    }
    const float beta = (float)s.beta;
    load_lut(sH, b.lut32);
#pragma unroll
    for (int q = 0; q < KZ + 2; ++q) {
        const int kk = k0 - 1 + q;
        if (kk >= 0 && kk < nz) {
            const float z = rc[q] * dinv_of(cc[q], sH, c);
            pl[q][ty + 1][tx + 1] = first ? z : z + beta * pc[q];
            if (tp.has_h) {
                const float zh = rh[q] * dinv_of(ch[q], sH, c);
                pl[q][tp.hy][tp.hx] = first ? zh : zh + beta * ph[q];
            }
        }
    }
    __syncthreads();
    double acc[1] = {0.0};
#pragma unroll
    for (int q = 1; q <= KZ; ++q) {
        const int k = k0 - 1 + q;
        if (k < nz) {
            const float xc = pl[q][ty + 1][tx + 1];
            float qv;
            if (cc[q] & SC_IDENTITY) {
                qv = op_identity(xc, k, c);
            } else {
                int qu = q + 1, qd = q - 1;
                if (c.zbc != 0) {
                    if (k == 0) qu = qd = q + 1;
                    else if (k == nz - 1) qu = qd = q - 1;
                }
                qv = row_Ax<float>(cc[q], sH, c, xc, pl[q][ty + 1][tx + 2], pl[q][ty + 1][tx], pl[q][ty + 2][tx + 1],
                                   pl[q][ty][tx + 1], pl[qu][ty + 1][tx + 1], pl[qd][ty + 1][tx + 1]);
            }
            if (tp.live) {
                pnew[cbase + q * plane] = xc;
                acc[0] += (double)xc * (double)qv;
            }
        }
    }
    write_block_partials<1>(acc, b.part[io], nb);
    s.iters += 1;
    s.pbuf ^= 1;
    s.kind = K_DIR;
    pass_state(s, b.st + io);
}

// ---------------------------------------------------------------- update / residual replacement
template <int KZ>
__global__ __launch_bounds__(NT) void k_pcg_update_f32(Fp32Bufs b, StepConsts c, int nb, int t) {
    const int ip = (t - 1) & 1, io = t & 1;
    PcgState s = load_state(b.st + ip);
    if (s.done) {
        pass_state(s, b.st + io);
        return;
    }
    const bool resid = s.kind == K_FOLD;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const int k0 = blockIdx.z * KZ;
    double acc[2] = {0.0, 0.0};
    if (resid) {  // r = b - A x_hi in FP64 (x_lo was folded by the previous launch)
        __shared__ double sH64[16];
        __shared__ float sH32r[16];
        load_lut(sH64, b.lut);
        load_lut(sH32r, b.lut32);
        const int i = blockIdx.x * BX + threadIdx.x;
        const int j = blockIdx.y * BY + threadIdx.y;
        if (i < nx && j < ny) {
            const int k1 = min(k0 + KZ, nz);
            for (int k = k0; k < k1; ++k) {
                const Nbr v = op_nbrs(i, j, k, c);
                const unsigned code = b.sc[v.c];
                const float r = (float)true_residual(b, sH64, c, v, code, k, b.b[v.c], false);
                b.r[v.c] = r;
                const float z = r * dinv_of(code, sH32r, c);
                acc[0] += (double)r * (double)z;
                acc[1] += (double)r * (double)r;
            }
        }
    } else {
        __shared__ float sH32[16];
        __shared__ float pl[KZ + 2][HY][HX];
        const TilePos tp = tile_pos(c);
        const int tx = threadIdx.x, ty = threadIdx.y;
        const float* p = b.p[s.pbuf];
        float pc[KZ + 2], ph[KZ + 2], xv[KZ], rv[KZ];
        unsigned cv[KZ];
        const long long plane = (long long)nx * ny;
        const long long cbase = vid(tp.ci, tp.cj, k0 - 1, nx, ny), hbase = vid(tp.gi, tp.gj, k0 - 1, nx, ny);
#pragma unroll
        for (int q = 0; q < KZ + 2; ++q) {
            const int kk = k0 - 1 + q;
            pc[q] = ph[q] = 0.f;
            if (kk >= 0 && kk < nz) {
                pc[q] = p[cbase + q * plane];
                if (tp.has_h) ph[q] = p[hbase + q * plane];
            }
        }
#pragma unroll
        for (int q = 0; q < KZ; ++q) {
            const int kk = k0 + q;
            xv[q] = rv[q] = 0.f;
            cv[q] = SC_IDENTITY;
            if (kk < nz) {
                const long long id = cbase + (q + 1) * plane;
                xv[q] = b.xlo[id];
                rv[q] = b.r[id];
                cv[q] = b.sc[id];
            }
        }
        double tot[1];
        sum_partials<1>(b.part[ip], nb, tot);
        s.denom = tot[0];
        if (!finite(tot[0]) || !(tot[0] > 0.0)) {
            s.status = HZO_PCG_BREAKDOWN;
            s.done = 1;
            pass_state(s, b.st + io);
            return;
        }
        s.alpha = s.rz_old / tot[0];
        const float alpha = (float)s.alpha;
        load_lut(sH32, b.lut32);
#pragma unroll
        for (int q = 0; q < KZ + 2; ++q) {
            const int kk = k0 - 1 + q;
            if (kk >= 0 && kk < nz) {
                pl[q][ty + 1][tx + 1] = pc[q];
                if (tp.has_h) pl[q][tp.hy][tp.hx] = ph[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 1; q <= KZ; ++q) {
            const int k = k0 - 1 + q;
            if (k < nz) {
                const float xc = pl[q][ty + 1][tx + 1];
                const unsigned code = cv[q - 1];
                float qv;
                if (code & SC_IDENTITY) {
                    qv = op_identity(xc, k, c);
                } else {
                    int qu = q + 1, qd = q - 1;
                    if (c.zbc != 0) {
                        if (k == 0) qu = qd = q + 1;
                        else if (k == nz - 1) qu = qd = q - 1;
                    }
                    qv = row_Ax<float>(code, sH32, c, xc, pl[q][ty + 1][tx + 2], pl[q][ty + 1][tx],
                                       pl[q][ty + 2][tx + 1], pl[q][ty][tx + 1], pl[qu][ty + 1][tx + 1],
                                       pl[qd][ty + 1][tx + 1]);
                }
                if (tp.live) {
                    const long long id = cbase + q * plane;
                    b.xlo[id] = xv[q - 1] + alpha * xc;
                    const float r = rv[q - 1] - alpha * qv;
                    b.r[id] = r;
                    const float z = r * dinv_of(code, sH32, c);
                    acc[0] += (double)r * (double)z;
                    acc[1] += (double)r * (double)r;
                }
            }
        }
    }
    write_block_partials<2>(acc, b.part[io], nb);
    s.kind = resid ? K_RESID : K_UPDATE;
    pass_state(s, b.st + io);
}

// ---------------------------------------------------------------- diagnostics
__global__ __launch_bounds__(NT) void k_diag_f32(const float* __restrict__ P, const uint32_t* __restrict__ cw,
                                                  Grid g, double* part, int nb) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long N = g.N;
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k0 = blockIdx.z * g.kz;
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
    write_block_partials<DIAG_Q>(a, part, nb);
}

void launch_tdgl_f32(hipStream_t s, const Grid& g, const StepConsts& c, double Ez_ext, const Fp32Bufs& b) {
    hipLaunchKernelGGL(k_tdgl_f32, grid3(g), dim3(BX, BY, 1), 0, s, b, c, (float)Ez_ext, blocks_for_impl(g));
}
void launch_pcg_setup_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b,
                          const PcgState& init) {
    hipLaunchKernelGGL(k_pcg_setup_f32, grid3(g), dim3(BX, BY, 1), 0, s, b, c, blocks_for_impl(g), init);
}
void launch_pcg_dir_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b, int t) {
    const dim3 gr = grid3(g), bl(BX, BY, 1);
    const int nb = blocks_for_impl(g);
    switch (g.kz) {
        case 2: hipLaunchKernelGGL(k_pcg_dir_f32<2>, gr, bl, 0, s, b, c, nb, t); break;
        case 4: hipLaunchKernelGGL(k_pcg_dir_f32<4>, gr, bl, 0, s, b, c, nb, t); break;
        default: hipLaunchKernelGGL(k_pcg_dir_f32<8>, gr, bl, 0, s, b, c, nb, t); break;
    }
}
void launch_pcg_update_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b, int t) {
    const dim3 gr = grid3(g), bl(BX, BY, 1);
    const int nb = blocks_for_impl(g);
    switch (g.kz) {
        case 2: hipLaunchKernelGGL(k_pcg_update_f32<2>, gr, bl, 0, s, b, c, nb, t); break;
        case 4: hipLaunchKernelGGL(k_pcg_update_f32<4>, gr, bl, 0, s, b, c, nb, t); break;
        default: hipLaunchKernelGGL(k_pcg_update_f32<8>, gr, bl, 0, s, b, c, nb, t); break;
    }
}
void launch_diag_f32(hipStream_t s, const Grid& g, const float* P, const uint32_t* cw, double* part) {
    hipLaunchKernelGGL(k_diag_f32, grid3(g), dim3(BX, BY, 1), 0, s, P, cw, g, part, blocks_for_impl(g));
}

}  // namespace hzo
