// kernels_adjoint.hip -- reverse-mode (adjoint) kernels of the HZO step.
//
// Forward step n (FP64 and FP32-refined forms alike):
//   P_{n+1} = T(P_n, phi_n; theta, E_n)        TDGL, kernels_evol.cu:47-231
//   phi_{n+1} = S(P_{n+1})                      A phi = b(P), depol_varEPS.cu:332-595
// with theta = (alpha, beta, gamma, kappa) of each phase.  Given the adjoint
// a = dL/dP_{n+1}, one backward step evaluates
//   w = (dT/dP_n)^T a       (local Landau Hessian + transposed Laplacian)
//   v = (dT/dphi_n)^T a     (transposed E = -grad phi)
//   g += (dT/dtheta)^T a    (per-phase parameter gradients, FP64 sums)
// then solves A mu = v (the interior operator is symmetric for Dirichlet
// electrodes, so the same PCG serves) and forms
//   dL/dP_n = w + (db/dP)^T mu + dl_{n-1}/dP_n.
//
// The TDGL update, per non-void voxel c with phase coefficients of c:
//   P'_c = P_c (1 - 6A_c) + A_c sum_{j in nb(c)} P_j - cL_c (R^T f(R P_c) - E_c - E_ext z)
// (only the Pm components see the field).  x/y neighbours are periodic; in z
// the neighbour beyond the film is the voxel itself (ghost = self), so the
// transpose adds A_c a_c on the top and bottom planes.  Void voxels write 0
// and read nothing, so they neither receive nor pass adjoint.
#include "../hzo_internal.h"
#include "common.hip.h"
#include "reduce.hip.h"

namespace hzo {

template <class T>
struct AdjTDGLBufs {
    const T* P;        // forward input state P_n, 6 planes of N
    const T* a;        // adjoint of P_{n+1}, 6 planes
    T* w;              // out: (dT/dP)^T a, 6 planes
    double* v;         // out: (dT/dphi)^T a, 0 on identity rows (adjoint Poisson right-hand side)
    const uint32_t* cw;
    const uint16_t* sc;
    const T* TR;
    double* gpart;     // out: 12 x nb workgroup partials of the parameter gradients
};

template <class T>
struct PhaseTab;
template <>
struct PhaseTab<double> {
    using type = PhaseC;
};
template <>
struct PhaseTab<float> {
    using type = PhaseC32;
};

// 12 per-thread values -> one partial per workgroup (wave sums by DPP, then
// the 8 wave totals in order); out[q*nb + block].
__device__ __forceinline__ void block_partials12(const double (&v)[12], double* out, int nb) {
    __shared__ double sw[8][12];
    const int tid = linear_tid();
    const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        const double s = wave_sum(v[q]);
        if (lane == 0) sw[wid][q] = s;
    }
    __syncthreads();
    if (tid < 12) {
        double s = 0.0;
        for (int w = 0; w < 8; ++w) s += sw[w][tid];
        out[(long long)tid * nb + linear_block()] = s;
    }
}

template <class T>
__global__ __launch_bounds__(NT) void k_adj_tdgl(AdjTDGLBufs<T> b, const typename PhaseTab<T>::type* ptab,
                                                 StepConsts c, double dAdk, int nb) {
    using PC = typename PhaseTab<T>::type;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const long long N = (long long)nx * ny * nz;
    const TileId Tl = tile_id();
    const int i = Tl.bx * BX + threadIdx.x;
    const int j = Tl.by * BY + threadIdx.y;
    const int k0 = Tl.bz * c.kz;
    const double inv2h = 1.0 / c.two_h;
    double g[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) g[q] = 0.0;
    auto coef = [&](uint32_t w) -> const PC& {
        return ptab[((w >> CW_PHASE_SHIFT) & 3u) * 2u + ((w & CW_IFACE) ? 1u : 0u)];
    };
    if (i < nx && j < ny) {
        const int e = (i + 1 == nx) ? 0 : i + 1;
        const int wq = (i == 0) ? nx - 1 : i - 1;
        const int n = (j + 1 == ny) ? 0 : j + 1;
        const int s = (j == 0) ? ny - 1 : j - 1;
        const int k1 = min(k0 + c.kz, nz);
        for (int k = k0; k < k1; ++k) {
            const long long id = vid(i, j, k, nx, ny);
            const long long nbr[4] = {vid(e, j, k, nx, ny), vid(wq, j, k, nx, ny), vid(i, n, k, nx, ny),
                                      vid(i, s, k, nx, ny)};
            const uint32_t cwc = b.cw[id];
            // ---- E transpose: m'_j gains cL_j E_j with E = -grad phi (central, /2h)
            if (c.dep_flag) {
                double vv = 0.0;
                if (!(b.sc[id] & SC_IDENTITY)) {
                    auto cla = [&](long long jj, int comp) -> double {
                        const uint32_t wj = b.cw[jj];
                        if (wj & CW_VOID) return 0.0;
                        return (double)coef(wj).cL * (double)b.a[comp * N + jj];
                    };
                    vv = cla(nbr[0], 0) - cla(nbr[1], 0) + cla(nbr[2], 1) - cla(nbr[3], 1);
                    if (k <= nz - 2) vv += cla(vid(i, j, k + 1, nx, ny), 2);
                    if (k >= 1) vv -= cla(vid(i, j, k - 1, nx, ny), 2);
                    vv *= inv2h;
                }
                b.v[id] = vv;
            }
            if (cwc & CW_VOID) {
#pragma unroll
                for (int q = 0; q < 6; ++q) b.w[q * N + id] = (T)0;
                continue;
            }
            const PC& pc = coef(cwc);
            const int ph = (int)((cwc >> CW_PHASE_SHIFT) & 3u);
            const T* R = b.TR + (long long)(cwc & CW_GID_MASK) * 9;
            const T r11 = R[0], r12 = R[1], r13 = R[2], r21 = R[3], r22 = R[4], r23 = R[5], r31 = R[6],
                    r32 = R[7], r33 = R[8];
            T P[6], A[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                P[q] = b.P[q * N + id];
                A[q] = b.a[q * N + id];
            }
            // local frame
            const T x = r31 * P[0] + r32 * P[1] + r33 * P[2];
            const T y = r31 * P[3] + r32 * P[4] + r33 * P[5];
            const T amx = r11 * A[0] + r12 * A[1] + r13 * A[2];
            const T amy = r21 * A[0] + r22 * A[1] + r23 * A[2];
            const T amz = r31 * A[0] + r32 * A[1] + r33 * A[2];
            const T asx = r11 * A[3] + r12 * A[4] + r13 * A[5];
            const T asy = r21 * A[3] + r22 * A[4] + r23 * A[5];
            const T asz = r31 * A[3] + r32 * A[4] + r33 * A[5];
            const T x2 = x * x, y2 = y * y, x3 = x2 * x, y3 = y2 * y, x4 = x2 * x2, y4 = y2 * y2;
            // second derivatives of the local free energy (z-z block)
            const T hmm = pc.ag + (T)3 * pc.beta_4 * x2 + pc.beta_12 * y2 +
                          pc.gamma_6 * ((T)5 * x4 + (T)30 * x2 * y2 + (T)5 * y4);
            const T hms = (T)2 * pc.beta_12 * x * y + pc.gamma_6 * ((T)20 * x3 * y + (T)20 * x * y3);
            const T hss = pc.ags + (T)3 * pc.beta_4 * y2 + pc.beta_12 * x2 +
                          pc.gamma_6 * ((T)5 * y4 + (T)30 * x2 * y2 + (T)5 * x4);
            const T gmx = pc.a0_2 * amx, gmy = pc.a0_2 * amy, gmz = hmm * amz + hms * asz;
            const T gsx = pc.a0_2 * asx, gsy = pc.a0_2 * asy, gsz = hms * amz + hss * asz;
            const T cL = pc.cL, om = pc.one_m6A;
            T W[6];
            W[0] = om * A[0] - cL * (r11 * gmx + r21 * gmy + r31 * gmz);
            W[1] = om * A[1] - cL * (r12 * gmx + r22 * gmy + r32 * gmz);
            W[2] = om * A[2] - cL * (r13 * gmx + r23 * gmy + r33 * gmz);
            W[3] = om * A[3] - cL * (r11 * gsx + r21 * gsy + r31 * gsz);
            W[4] = om * A[4] - cL * (r12 * gsx + r22 * gsy + r32 * gsz);
            W[5] = om * A[5] - cL * (r13 * gsx + r23 * gsy + r33 * gsz);
            // transposed Laplacian: every non-void voxel j that reads P_c adds A_j a_j
            auto addnb = [&](long long jj) {
                const uint32_t wj = b.cw[jj];
                if (wj & CW_VOID) return;
                const T Aj = coef(wj).A;
#pragma unroll
                for (int q = 0; q < 6; ++q) W[q] += Aj * b.a[q * N + jj];
            };
#pragma unroll
            for (int t = 0; t < 4; ++t) addnb(nbr[t]);
            if (k >= 1) addnb(vid(i, j, k - 1, nx, ny));
            if (k <= nz - 2) addnb(vid(i, j, k + 1, nx, ny));
            if (k == nz - 1) {
#pragma unroll
                for (int q = 0; q < 6; ++q) W[q] += pc.A * A[q];
            }
            if (k == 0) {
#pragma unroll
                for (int q = 0; q < 6; ++q) W[q] += pc.A * A[q];
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) b.w[q * N + id] = W[q];
            // ---- parameter gradients (FP64 accumulation)
            const double dx = (double)x, dy = (double)y, dx2 = dx * dx, dy2 = dy * dy;
            const double am = (double)amz, as = (double)asz, dcL = (double)cL;
            const double ga = -dcL * (am * 2.0 * dx + as * 2.0 * dy);
            const double gb = -dcL * (am * (4.0 * dx2 * dx + 12.0 * dx * dy2) + as * (4.0 * dy2 * dy + 12.0 * dy * dx2));
            const double gg = -dcL * (am * 6.0 * (dx2 * dx2 * dx + 10.0 * dx2 * dx * dy2 + 5.0 * dx * dy2 * dy2) +
                                      as * 6.0 * (dy2 * dy2 * dy + 10.0 * dx2 * dy2 * dy + 5.0 * dy * dx2 * dx2));
            // kappa: d/dA of P(1-6A) + A sum_nb P = sum_nb P - 6P (forward neighbour rule)
            const long long iu = vid(i, j, (k == nz - 1) ? k : k + 1, nx, ny);
            const long long idn = vid(i, j, (k == 0) ? k : k - 1, nx, ny);
            double gk = 0.0;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const T* Pq = b.P + q * N;
                const double lap = (double)Pq[nbr[0]] + (double)Pq[nbr[1]] + (double)Pq[nbr[2]] +
                                   (double)Pq[nbr[3]] + (double)Pq[iu] + (double)Pq[idn] - 6.0 * (double)P[q];
                gk += (double)A[q] * lap;
            }
            gk *= dAdk;
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                const bool on = ph == p;
                g[p * 4 + 0] += on ? ga : 0.0;
                g[p * 4 + 1] += on ? gb : 0.0;
                g[p * 4 + 2] += on ? gg : 0.0;
                g[p * 4 + 3] += on ? gk : 0.0;
            }
        }
    }
    block_partials12(g, b.gpart, nb);
}

// grad[q] += sum of the 12 x nb partials, one wave, fixed order.
__global__ __launch_bounds__(64) void k_adj_gradacc(const double* gpart, int nb, double* grad) {
    const int lane = threadIdx.x;
    for (int q = 0; q < 12; ++q) {
        double s = 0.0;
        for (int t = lane; t < nb; t += 64) s += gpart[(long long)q * nb + t];
        s = wave_sum(s);
        if (lane == 0) grad[q] += s;
    }
}

// a_n = w + (db/dP)^T mu + seed/N on Pmz, with b = -eps0 div Pm / (2h)
// (z clamped: the neighbour beyond the film is the voxel itself) and b = 0 on
// identity rows, where mu is 0.  mu = mu_hi (+ mu_lo in the FP32 form).
template <class T>
__global__ __launch_bounds__(NT) void k_adj_divt(const T* w, const double* mu_hi, const float* mu_lo, T* a,
                                                 StepConsts c, double seed_over_N, int with_mu) {
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const long long N = (long long)nx * ny * nz;
    const TileId Tl = tile_id();
    const int i = Tl.bx * BX + threadIdx.x;
    const int j = Tl.by * BY + threadIdx.y;
    const int k0 = Tl.bz * c.kz;
    if (i >= nx || j >= ny) return;
    const int e = (i + 1 == nx) ? 0 : i + 1;
    const int wq = (i == 0) ? nx - 1 : i - 1;
    const int n = (j + 1 == ny) ? 0 : j + 1;
    const int s = (j == 0) ? ny - 1 : j - 1;
    const double f = -c.eps0 / c.two_h;
    auto MU = [&](long long q) { return mu_hi[q] + (mu_lo ? (double)mu_lo[q] : 0.0); };
    const int k1 = min(k0 + c.kz, nz);
    for (int k = k0; k < k1; ++k) {
        const long long id = vid(i, j, k, nx, ny);
        double dx = 0.0, dy = 0.0, dz = 0.0;
        if (with_mu) {
            dx = f * (MU(vid(wq, j, k, nx, ny)) - MU(vid(e, j, k, nx, ny)));
            dy = f * (MU(vid(i, s, k, nx, ny)) - MU(vid(i, n, k, nx, ny)));
            double z = 0.0;
            if (k >= 1) z += MU(vid(i, j, k - 1, nx, ny));
            if (k == nz - 1) z += MU(id);
            if (k <= nz - 2) z -= MU(vid(i, j, k + 1, nx, ny));
            if (k == 0) z -= MU(id);
            dz = f * z;
        }
        a[id] = (T)((double)w[id] + dx);
        a[N + id] = (T)((double)w[N + id] + dy);
        a[2 * N + id] = (T)((double)w[2 * N + id] + dz + seed_over_N);
#pragma unroll
        for (int q = 3; q < 6; ++q) a[q * N + id] = w[q * N + id];
    }
}

// a = seed/N on Pmz, 0 elsewhere (adjoint of the last state).
template <class T>
__global__ __launch_bounds__(NT) void k_adj_seed(T* a, long long N, double seed_over_N) {
    const long long id = (long long)blockIdx.x * NT + threadIdx.x + BX * threadIdx.y;
    if (id >= N) return;
#pragma unroll
    for (int q = 0; q < 6; ++q) a[q * N + id] = (T)(q == 2 ? seed_over_N : 0.0);
}

// out[0] = sum of the first column of a partial array (nb entries, fixed order).
__global__ __launch_bounds__(64) void k_sum_column(const double* part, int nb, double* out) {
    const int lane = threadIdx.x;
    double s = 0.0;
    for (int t = lane; t < nb; t += 64) s += part[t];
    s = wave_sum(s);
    if (lane == 0) out[0] = s;
}

// ---------------------------------------------------------------- launchers
template <class T>
static void launch_adj_tdgl_t(hipStream_t s, const Grid& g, const StepConsts& c, const AdjTDGLBufsAny& b) {
    AdjTDGLBufs<T> k;
    k.P = (const T*)b.P;
    k.a = (const T*)b.a;
    k.w = (T*)b.w;
    k.v = b.v;
    k.cw = b.cw;
    k.sc = b.sc;
    k.TR = (const T*)b.TR;
    k.gpart = b.gpart;
    hipLaunchKernelGGL(k_adj_tdgl<T>, grid3(g), dim3(BX, BY, 1), 0, s, k,
                       (const typename PhaseTab<T>::type*)b.ptab, c, b.dAdk, blocks_for_impl(g));
}

void launch_adj_tdgl(hipStream_t s, const Grid& g, const StepConsts& c, const AdjTDGLBufsAny& b) {
    if (b.fp32) launch_adj_tdgl_t<float>(s, g, c, b);
    else launch_adj_tdgl_t<double>(s, g, c, b);
}

void launch_adj_gradacc(hipStream_t s, const Grid& g, const double* gpart, double* grad) {
    hipLaunchKernelGGL(k_adj_gradacc, dim3(1), dim3(64), 0, s, gpart, blocks_for_impl(g), grad);
}

void launch_adj_divt(hipStream_t s, const Grid& g, const StepConsts& c, bool fp32, const void* w,
                     const double* mu_hi, const float* mu_lo, void* a, double seed_over_N, bool with_mu) {
    if (fp32)
        hipLaunchKernelGGL(k_adj_divt<float>, grid3(g), dim3(BX, BY, 1), 0, s, (const float*)w, mu_hi, mu_lo,
                           (float*)a, c, seed_over_N, (int)with_mu);
    else
        hipLaunchKernelGGL(k_adj_divt<double>, grid3(g), dim3(BX, BY, 1), 0, s, (const double*)w, mu_hi, mu_lo,
                           (double*)a, c, seed_over_N, (int)with_mu);
}

void launch_adj_seed(hipStream_t s, const Grid& g, bool fp32, void* a, double seed_over_N) {
    const unsigned nblk = (unsigned)((g.N + NT - 1) / NT);
    if (fp32) hipLaunchKernelGGL(k_adj_seed<float>, dim3(nblk), dim3(BX, BY, 1), 0, s, (float*)a, g.N, seed_over_N);
    else hipLaunchKernelGGL(k_adj_seed<double>, dim3(nblk), dim3(BX, BY, 1), 0, s, (double*)a, g.N, seed_over_N);
}

void launch_sum_column(hipStream_t s, const Grid& g, const double* part, double* out) {
    hipLaunchKernelGGL(k_sum_column, dim3(1), dim3(64), 0, s, part, blocks_for_impl(g), out);
}

}  // namespace hzo
