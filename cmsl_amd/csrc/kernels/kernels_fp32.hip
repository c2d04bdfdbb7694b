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
    if (block0_thread0()) {
        init.kind = K_SETUP;
        b.st[0] = init;
    }
}

// ---------------------------------------------------------------- direction / fold
// Finishes the previous kernel's inner products and decides what this launch
// is: a CG direction update (K_DIR) or the fold x_hi += x_lo that starts a
// residual replacement (K_FOLD).
__global__ __launch_bounds__(NT) void k_pcg_dir_f32(Fp32Bufs b, StepConsts c, int nb, int t) {
    const int ip = (t - 1) & 1, io = t & 1;
    PcgState s = b.st[ip];
    if (s.done) {
        pass_state(s, b.st + io);
        return;
    }
    const bool first = s.kind == K_SETUP;
    bool fold = false;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k0 = blockIdx.z * c.kz;
    const int k1 = min(k0 + c.kz, nz);
    __shared__ float sH[16];
    __shared__ float ring[3][HY][HX];
    const TilePos tp = tile_pos(c);
    const int tx = threadIdx.x, ty = threadIdx.y;
    const float* pold = b.p[s.pbuf];
    float* pnew = b.p[s.pbuf ^ 1];
    float beta = 0.f;  // set once the previous inner products are finished
    float rc_ = 0.f, pc_ = 0.f, rh = 0.f, ph = 0.f;
    unsigned cc_ = 0, ch = 0;
    auto fetch = [&](int kk) {
        const long long id = vid(tp.ci, tp.cj, kk, nx, ny);
        rc_ = b.r[id];
        cc_ = b.sc[id];
        if (!first) pc_ = pold[id];
        if (tp.has_h) {
            const long long h = vid(tp.gi, tp.gj, kk, nx, ny);
            rh = b.r[h];
            ch = b.sc[h];
            if (!first) ph = pold[h];
        }
    };
    auto make = [&](int kk) -> unsigned {
        float(*S)[HX] = ring[kk % 3];
        const float z = rc_ * dinv_of(cc_, sH, c);
        S[ty + 1][tx + 1] = first ? z : z + beta * pc_;
        if (tp.has_h) {
            const float zh = rh * dinv_of(ch, sH, c);
            S[tp.hy][tp.hx] = first ? zh : zh + beta * ph;
        }
        return cc_;
    };
    int lo, hi;
    plane_window(c, k0, k1, lo, hi);
    double acc[1] = {0.0};
    fetch(lo);  // the first plane's loads overlap the reduction below
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
        if (i < nx && j < ny) {
            for (int k = k0; k < k1; ++k) {
                const long long id = vid(i, j, k, nx, ny);
                b.xhi[id] += (double)b.xlo[id];
                b.xlo[id] = 0.0f;
            }
        }
        s.kind = K_FOLD;
        pass_state(s, b.st + io);
        return;
    }
    beta = (float)s.beta;
    load_lut(sH, b.lut32);
    unsigned code_k = make(lo), code_next = 0;
    int kn = lo + 1;
    if (kn <= hi) fetch(kn);
    for (int k = k0; k < k1; ++k) {
        const int need = min(k + 1, hi);
        while (kn <= need) {
            const unsigned cm = make(kn);
            if (kn == k) code_k = cm;
            else code_next = cm;
            ++kn;
            if (kn <= hi) fetch(kn);
        }
        __syncthreads();
        float(*S)[HX] = ring[k % 3];
        const float pc = S[ty + 1][tx + 1];
        float q;
        if (code_k & SC_IDENTITY) {
            q = op_identity(pc, k, c);
        } else {
            int up = k + 1, dn = k - 1;
            if (c.zbc != 0) {
                if (k == 0) up = dn = 1;
                else if (k == nz - 1) up = dn = nz - 2;
            }
            q = row_Ax<float>(code_k, sH, c, pc, S[ty + 1][tx + 2], S[ty + 1][tx], S[ty + 2][tx + 1],
                              S[ty][tx + 1], ring[up % 3][ty + 1][tx + 1], ring[dn % 3][ty + 1][tx + 1]);
        }
        if (tp.live) {
            pnew[vid(tp.i, tp.j, k, nx, ny)] = pc;
            acc[0] += (double)pc * (double)q;
        }
        __syncthreads();
        code_k = code_next;
    }
    write_block_partials<1>(acc, b.part[io], nb);
    s.iters += 1;
    s.pbuf ^= 1;
    s.kind = K_DIR;
    pass_state(s, b.st + io);
}

// ---------------------------------------------------------------- update / residual replacement
__global__ __launch_bounds__(NT) void k_pcg_update_f32(Fp32Bufs b, StepConsts c, int nb, int t) {
    const int ip = (t - 1) & 1, io = t & 1;
    PcgState s = b.st[ip];
    if (s.done) {
        pass_state(s, b.st + io);
        return;
    }
    const bool resid = s.kind == K_FOLD;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const int k0 = blockIdx.z * c.kz, k1 = min(k0 + c.kz, nz);
    double acc[2] = {0.0, 0.0};
    if (resid) {  // r = b - A x_hi in FP64 (x_lo was folded by the previous launch)
        __shared__ double sH64[16];
        __shared__ float sH32r[16];
        load_lut(sH64, b.lut);
        load_lut(sH32r, b.lut32);
        const int i = blockIdx.x * BX + threadIdx.x;
        const int j = blockIdx.y * BY + threadIdx.y;
        if (i < nx && j < ny) {
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
        __shared__ float ring[3][HY][HX];
        const TilePos tp = tile_pos(c);
        const int tx = threadIdx.x, ty = threadIdx.y;
        const float* p = b.p[s.pbuf];
        float pcv = 0.f, phv = 0.f;
        auto fetch = [&](int kk) {
            pcv = p[vid(tp.ci, tp.cj, kk, nx, ny)];
            if (tp.has_h) phv = p[vid(tp.gi, tp.gj, kk, nx, ny)];
        };
        auto make = [&](int kk) {
            ring[kk % 3][ty + 1][tx + 1] = pcv;
            if (tp.has_h) ring[kk % 3][tp.hy][tp.hx] = phv;
        };
        struct Ctr {
            float x, r;
            unsigned code;
        };
        auto fetch_c = [&](int kk) {
            const long long id = vid(tp.ci, tp.cj, kk, nx, ny);
            Ctr v;
            v.x = b.xlo[id];
            v.r = b.r[id];
            v.code = b.sc[id];
            return v;
        };
        int lo, hi;
        plane_window(c, k0, k1, lo, hi);
        fetch(lo);  // loads of the first planes overlap the reduction below
        Ctr cur = fetch_c(k0), nxt = cur;
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
        make(lo);
        int kn = lo + 1;
        if (kn <= hi) fetch(kn);
        for (int k = k0; k < k1; ++k) {
            const int need = min(k + 1, hi);
            while (kn <= need) {
                make(kn);
                ++kn;
                if (kn <= hi) fetch(kn);
            }
            if (k + 1 < k1) nxt = fetch_c(k + 1);
            __syncthreads();
            float(*S)[HX] = ring[k % 3];
            const float pc = S[ty + 1][tx + 1];
            float q;
            if (cur.code & SC_IDENTITY) {
                q = op_identity(pc, k, c);
            } else {
                int up = k + 1, dn = k - 1;
                if (c.zbc != 0) {
                    if (k == 0) up = dn = 1;
                    else if (k == nz - 1) up = dn = nz - 2;
                }
                q = row_Ax<float>(cur.code, sH32, c, pc, S[ty + 1][tx + 2], S[ty + 1][tx], S[ty + 2][tx + 1],
                                  S[ty][tx + 1], ring[up % 3][ty + 1][tx + 1], ring[dn % 3][ty + 1][tx + 1]);
            }
            if (tp.live) {
                const long long id = vid(tp.i, tp.j, k, nx, ny);
                b.xlo[id] = cur.x + alpha * pc;
                const float r = cur.r - alpha * q;
                b.r[id] = r;
                const float z = r * dinv_of(cur.code, sH32, c);
                acc[0] += (double)r * (double)z;
                acc[1] += (double)r * (double)r;
            }
            __syncthreads();
            cur = nxt;
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
    hipLaunchKernelGGL(k_pcg_dir_f32, grid3(g), dim3(BX, BY, 1), 0, s, b, c, blocks_for_impl(g), t);
}
void launch_pcg_update_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b, int t) {
    hipLaunchKernelGGL(k_pcg_update_f32, grid3(g), dim3(BX, BY, 1), 0, s, b, c, blocks_for_impl(g), t);
}
void launch_diag_f32(hipStream_t s, const Grid& g, const float* P, const uint32_t* cw, double* part) {
    hipLaunchKernelGGL(k_diag_f32, grid3(g), dim3(BX, BY, 1), 0, s, P, cw, g, part, blocks_for_impl(g));
}

}  // namespace hzo
