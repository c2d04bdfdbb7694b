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
// the device decides, through PcgState.mode, whether a pair is a CG
// iteration or a residual replacement (fold x_lo into x_hi, recompute r).
//
// This file is compiled with -ffp-contract=fast: FP32 kernels may fuse.
#include "../hzo_internal.h"
#include "common.hip.h"
#include "reduce.hip.h"

namespace hzo {

// ---------------------------------------------------------------- TDGL
__global__ __launch_bounds__(256) void k_tdgl_f32(Fp32Bufs b, StepConsts c, float Ez_ext, int nb) {
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const long long N = (long long)nx * ny * nz;
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k0 = blockIdx.z * KZ;
    double loc[2] = {0.0, 0.0};
    if (i < nx && j < ny) {
        const int e = (i + 1 == nx) ? 0 : i + 1;
        const int w = (i == 0) ? nx - 1 : i - 1;
        const int n = (j + 1 == ny) ? 0 : j + 1;
        const int s = (j == 0) ? ny - 1 : j - 1;
        const int k1 = min(k0 + KZ, nz);
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
    double tot[2];
    if (grid_reduce_last<2>(loc, b.partials, nb, b.counters + 0, tot)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            b.diag[2] = tot[0];
            b.diag[3] = tot[1];
        }
    }
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
__global__ __launch_bounds__(256) void k_pcg_setup_f32(Fp32Bufs b, StepConsts c, int nb) {
    __shared__ double sH[16];
    load_lut(sH, b.lut);
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const long long N = (long long)nx * ny * nz;
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k0 = blockIdx.z * KZ;
    double acc[3] = {0.0, 0.0, 0.0};
    if (i < nx && j < ny) {
        const float* dinv = b.lut32 + LUT_DINV;
        const float* Px = b.Pcur;
        const float* Py = b.Pcur + N;
        const float* Pz = b.Pcur + 2 * N;
        const int k1 = min(k0 + KZ, nz);
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
            const float z = r * dinv[code];
            b.r[v.c] = r;
            acc[0] += (double)r * (double)z;
            acc[1] += (double)r * (double)r;
            acc[2] += bval * bval;
        }
    }
    double tot[3];
    if (grid_reduce_last<3>(acc, b.partials, nb, b.counters + 1, tot)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            PcgState& st = *b.st;
            st.rz0 = st.rz_old = tot[0];
            st.rr0 = st.rr = tot[1];
            st.bb = tot[2];
            st.iters = 0;
            st.status = 0;
            st.done = 0;
            st.beta = 0.0;
            st.mode = 0;
            st.refinements = 0;
            if (!finite(tot[0]) || tot[0] < 0.0) {
                st.status = HZO_PCG_BAD_INITIAL;
                st.done = 1;
            } else if (!keep_going(st, tot[0], tot[1])) {
                st.done = 1;
                if (st.iters >= st.max_iter) st.status = HZO_PCG_MAXITER;
            }
        }
    }
}

// ---------------------------------------------------------------- direction / fold
__global__ __launch_bounds__(256) void k_pcg_dir_f32(Fp32Bufs b, StepConsts c, int nb) {
    const PcgState* sp = b.st;
    if (sp->done) return;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k0 = blockIdx.z * KZ;
    if (sp->mode == 1) {  // fold the FP32 correction into the FP64 potential
        if (i < nx && j < ny) {
            const int k1 = min(k0 + KZ, nz);
            for (int k = k0; k < k1; ++k) {
                const long long id = vid(i, j, k, nx, ny);
                b.xhi[id] += (double)b.xlo[id];
                b.xlo[id] = 0.0f;
            }
        }
        return;
    }
    __shared__ float sH[16];
    load_lut(sH, b.lut32);
    const int pb = sp->pbuf;
    const float* pold = b.p[pb];
    float* pnew = b.p[pb ^ 1];
    const bool first = sp->iters == 0;
    const float beta = (float)sp->beta;
    double acc[1] = {0.0};
    if (i < nx && j < ny) {
        const float* dinv = b.lut32 + LUT_DINV;
        auto pn = [&](long long q) {
            const float z = b.r[q] * dinv[b.sc[q]];
            return first ? z : z + beta * pold[q];
        };
        const int k1 = min(k0 + KZ, nz);
        for (int k = k0; k < k1; ++k) {
            const Nbr v = op_nbrs(i, j, k, c);
            const unsigned code = b.sc[v.c];
            const float pc = pn(v.c);
            float q;
            if (code & SC_IDENTITY)
                q = op_identity(pc, k, c);
            else
                q = row_Ax<float>(code, sH, c, pc, pn(v.e), pn(v.w), pn(v.n), pn(v.s), pn(v.u), pn(v.d));
            pnew[v.c] = pc;
            acc[0] += (double)pc * (double)q;
        }
    }
    double tot[1];
    if (grid_reduce_last<1>(acc, b.partials, nb, b.counters + 2, tot)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            PcgState& st = *b.st;
            st.iters += 1;
            st.denom = tot[0];
            st.pbuf = pb ^ 1;
            if (!finite(tot[0]) || !(tot[0] > 0.0)) {
                st.status = HZO_PCG_BREAKDOWN;
                st.done = 1;
            } else {
                st.alpha = st.rz_old / tot[0];
            }
        }
    }
}

// ---------------------------------------------------------------- update / residual replacement
__global__ __launch_bounds__(256) void k_pcg_update_f32(Fp32Bufs b, StepConsts c, int nb) {
    const PcgState* sp = b.st;
    if (sp->done) return;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k0 = blockIdx.z * KZ;
    const int mode = sp->mode;
    __shared__ float sH32[16];
    __shared__ double sH64[16];
    if (mode == 0)
        load_lut(sH32, b.lut32);
    else
        load_lut(sH64, b.lut);
    const float alpha = (float)sp->alpha;
    const float* p = b.p[sp->pbuf];
    double acc[2] = {0.0, 0.0};
    if (i < nx && j < ny) {
        const float* dinv = b.lut32 + LUT_DINV;
        const int k1 = min(k0 + KZ, nz);
        for (int k = k0; k < k1; ++k) {
            const Nbr v = op_nbrs(i, j, k, c);
            const unsigned code = b.sc[v.c];
            float r;
            if (mode == 0) {
                const float pc = p[v.c];
                float q;
                if (code & SC_IDENTITY)
                    q = op_identity(pc, k, c);
                else
                    q = row_Ax<float>(code, sH32, c, pc, p[v.e], p[v.w], p[v.n], p[v.s], p[v.u], p[v.d]);
                b.xlo[v.c] += alpha * pc;
                r = b.r[v.c] - alpha * q;
            } else {
                r = (float)true_residual(b, sH64, c, v, code, k, b.b[v.c], false);
            }
            b.r[v.c] = r;
            const float z = r * dinv[code];
            acc[0] += (double)r * (double)z;
            acc[1] += (double)r * (double)r;
        }
    }
    double tot[2];
    if (grid_reduce_last<2>(acc, b.partials, nb, b.counters + 3, tot)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            PcgState& st = *b.st;
            const double rz = tot[0];
            st.rr = tot[1];
            if (!finite(rz) || rz < 0.0) {
                st.status = HZO_PCG_BAD_RZ;
                st.rz_old = rz;
                st.done = 1;
            } else if (mode == 0) {
                const bool conv = !keep_going(st, rz, tot[1]);
                if (st.iters >= st.max_iter) {
                    st.rz_old = rz;
                    st.done = 1;
                    st.status = HZO_PCG_MAXITER;
                } else if (conv || (st.refine_every > 0 && st.iters % st.refine_every == 0)) {
                    st.mode = 1;  // replace the recursive residual before trusting it
                } else {
                    st.beta = rz / st.rz_old;
                    st.rz_old = rz;
                }
            } else {
                st.refinements += 1;
                st.mode = 0;
                st.beta = rz / st.rz_old;
                st.rz_old = rz;
                if (!keep_going(st, rz, tot[1])) {
                    st.done = 1;
                    if (st.iters >= st.max_iter) st.status = HZO_PCG_MAXITER;
                }
            }
        }
    }
}

// ---------------------------------------------------------------- diagnostics
__global__ __launch_bounds__(256) void k_diag_f32(const float* __restrict__ P, const uint32_t* __restrict__ cw,
                                                  Grid g, double* partials, unsigned* counter, double* out,
                                                  int nb) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long N = g.N;
    const int i = blockIdx.x * BX + threadIdx.x;
    const int j = blockIdx.y * BY + threadIdx.y;
    const int k0 = blockIdx.z * KZ;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < nx && j < ny) {
        const int k1 = min(k0 + KZ, nz);
        for (int k = k0; k < k1; ++k) {
            const long long id = vid(i, j, k, nx, ny);
            const double pmz = P[2 * N + id];
            a[0] += pmz;
            a[1] += (double)P[5 * N + id];
            const uint32_t w = cw[id];
            const int phase = (w & CW_VOID) ? -1 : (int)((w >> CW_PHASE_SHIFT) & 3u);
            a[2] += (phase == 1) ? pmz : 0.0;
            a[3] += (phase == 2) ? pmz : 0.0;
            m[0] += (phase == 0) ? pmz : 0.0;
            m[1] += (phase == 1) ? 1.0 : 0.0;
            m[2] += (phase == 2) ? 1.0 : 0.0;
            m[3] += (phase == 0) ? 1.0 : 0.0;
        }
    }
    double t1[4];
    if (grid_reduce_last<4>(a, partials, nb, counter, t1)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            out[0] = t1[0];
            out[1] = t1[1];
            out[2] = t1[2];
            out[4] = t1[3];
        }
    }
    double t2[4];
    if (grid_reduce_last<4>(m, partials + 4 * (long long)nb, nb, counter + 1, t2)) {
        if (threadIdx.x == 0 && threadIdx.y == 0) {
            out[6] = t2[0];
            out[3] = t2[1];
            out[5] = t2[2];
            out[7] = t2[3];
        }
    }
}

void launch_tdgl_f32(hipStream_t s, const Grid& g, const StepConsts& c, double Ez_ext, const Fp32Bufs& b) {
    hipLaunchKernelGGL(k_tdgl_f32, grid3(g), dim3(BX, BY, 1), 0, s, b, c, (float)Ez_ext, blocks_for_impl(g));
}
void launch_pcg_setup_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b) {
    hipLaunchKernelGGL(k_pcg_setup_f32, grid3(g), dim3(BX, BY, 1), 0, s, b, c, blocks_for_impl(g));
}
void launch_pcg_dir_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b) {
    hipLaunchKernelGGL(k_pcg_dir_f32, grid3(g), dim3(BX, BY, 1), 0, s, b, c, blocks_for_impl(g));
}
void launch_pcg_update_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b) {
    hipLaunchKernelGGL(k_pcg_update_f32, grid3(g), dim3(BX, BY, 1), 0, s, b, c, blocks_for_impl(g));
}
void launch_diag_f32(hipStream_t s, const Grid& g, const float* P, const uint32_t* cw, double* partials,
                     unsigned* counters, double* out8) {
    hipLaunchKernelGGL(k_diag_f32, grid3(g), dim3(BX, BY, 1), 0, s, P, cw, g, partials, counters, out8,
                       blocks_for_impl(g));
}

}  // namespace hzo
