// hzo_oracle.cpp -- CPU FP64 restatement of the HZO TDGL + Jacobi-PCG step.
//
// TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load this library; the product never does.
//
// Every arithmetic expression below keeps the operand order of the reference
// CUDA source it cites, and the library is compiled with -ffp-contract=off, so
// each value is the IEEE-754 double the reference expression denotes when
// evaluated without FMA contraction.  The reference itself cannot run here
// (CUDA only, and its HEAD does not compile); see oracle/Makefile.
#include "hzo_oracle.h"

#include <cmath>
#include <omp.h>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

inline long long cid(int i, int j, int k, int nx, int ny) {
    return (long long)i + (long long)nx * ((long long)j + (long long)ny * k);
}

// depol_varEPS.cu:10-15
inline double harm(double a, double b) {
    const double eps = 1e-20;
    double denom = a + b;
    if (denom < eps) denom = eps;
    return (2.0 * a * b) / denom;
}

// Sum over k-planes of plane-sequential partial sums: deterministic for any
// thread count.
template <class F>
double plane_sum(int nx, int ny, int nz, F f) {
    std::vector<double> part(nz, 0.0);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        double s = 0.0;
        const long long base = (long long)nx * ny * k;
        for (long long t = 0; t < (long long)nx * ny; ++t) s += f(base + t);
        part[k] = s;
    }
    double s = 0.0;
    for (int k = 0; k < nz; ++k) s += part[k];
    return s;
}

}  // namespace

extern "C" {

// kernels_evol.cu:47-231, one thread per voxel.  Voxels with gid<0 or
// gid>=ng are skipped by the reference (:67-68); their "new" value is the
// never-written, zero-initialised buffer (allocateVars.cu:76-88), so 0 here.
void hzo_oracle_tdgl(const hzo_oracle_cfg* c, const double* Pold, double* Pnew,
                     double* Plocal, const int32_t* omega, int ng,
                     const int32_t* grain_phase, const double* TR,
                     const hzo_oracle_phase ph[3], const int32_t* iface,
                     const double* E3, double Ez_ext) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const long long N = (long long)nx * ny * nz;
    const double* Pmx_o = Pold;
    const double* Pmy_o = Pold + N;
    const double* Pmz_o = Pold + 2 * N;
    const double* Psx_o = Pold + 3 * N;
    const double* Psy_o = Pold + 4 * N;
    const double* Psz_o = Pold + 5 * N;
    const double del_h = c->del_h, del_t = c->del_t;
    const int dep_flag = c->dep_flag;

#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const long long id = cid(i, j, k, nx, ny);
                const int gid = omega[id];
                if (gid < 0 || gid >= ng) {
                    for (int q = 0; q < 6; ++q) {
                        Pnew[q * N + id] = 0.0;
                        if (Plocal) Plocal[q * N + id] = 0.0;
                    }
                    continue;
                }
                // :71-83 neighbours: x,y periodic, z clamped
                const int e = (i + 1) % nx;
                const int w = (i - 1 + nx) % nx;
                const int n = (j + 1) % ny;
                const int s = (j - 1 + ny) % ny;
                const int u = (k == nz - 1) ? k : (k + 1);
                const int d = (k == 0) ? k : (k - 1);
                const long long ie = cid(e, j, k, nx, ny), iw = cid(w, j, k, nx, ny);
                const long long in = cid(i, n, k, nx, ny), is = cid(i, s, k, nx, ny);
                const long long iu = cid(i, j, u, nx, ny), idd = cid(i, j, d, nx, ny);

                // :87-103, :106-122 neighbour sums in e,w,n,s,u,d order
                const double phiMxsum = Pmx_o[ie] + Pmx_o[iw] + Pmx_o[in] + Pmx_o[is] + Pmx_o[iu] + Pmx_o[idd];
                const double phiMysum = Pmy_o[ie] + Pmy_o[iw] + Pmy_o[in] + Pmy_o[is] + Pmy_o[iu] + Pmy_o[idd];
                const double phiMzsum = Pmz_o[ie] + Pmz_o[iw] + Pmz_o[in] + Pmz_o[is] + Pmz_o[iu] + Pmz_o[idd];
                const double phiSxsum = Psx_o[ie] + Psx_o[iw] + Psx_o[in] + Psx_o[is] + Psx_o[iu] + Psx_o[idd];
                const double phiSysum = Psy_o[ie] + Psy_o[iw] + Psy_o[in] + Psy_o[is] + Psy_o[iu] + Psy_o[idd];
                const double phiSzsum = Psz_o[ie] + Psz_o[iw] + Psz_o[in] + Psz_o[is] + Psz_o[iu] + Psz_o[idd];

                const double pMx0 = Pmx_o[id], pMy0 = Pmy_o[id], pMz0 = Pmz_o[id];
                const double pSx0 = Psx_o[id], pSy0 = Psy_o[id], pSz0 = Psz_o[id];

                // init_conf.cu:91-94: 2 -> DE, 1 -> FE, anything else -> AFE
                const int mat = grain_phase[gid];
                const hzo_oracle_phase& P = ph[mat == 2 ? 2 : (mat == 1 ? 1 : 0)];
                const double alpha_nd = P.alpha, beta_nd = P.beta, gamma_nd = P.gamma;
                double g_loc = P.g;
                const double a0_nd = P.a0, kappa_nd = P.kappa;
                double Lfac = P.L;
                if (iface && iface[id]) {  // :142-146
                    g_loc *= 0.4;
                    Lfac *= (1.0 + 10.0);
                }
                const double alpha_g_nd = 2.0 * (alpha_nd + g_loc);
                const double alpha_g_nd_s = 2.0 * (alpha_nd - g_loc);
                const double beta_4 = 4.0 * beta_nd;
                const double beta_12 = 12.0 * beta_nd;
                const double gamma_6 = 6.0 * gamma_nd;
                const double a0_2 = 2.0 * a0_nd;

                const double* T = TR + (long long)gid * 9;  // :157-159 row-major
                const double tr11 = T[0], tr12 = T[1], tr13 = T[2];
                const double tr21 = T[3], tr22 = T[4], tr23 = T[5];
                const double tr31 = T[6], tr32 = T[7], tr33 = T[8];

                // :162-168
                const double pMXl = tr11 * pMx0 + tr12 * pMy0 + tr13 * pMz0;
                const double pMYl = tr21 * pMx0 + tr22 * pMy0 + tr23 * pMz0;
                const double pMZl = tr31 * pMx0 + tr32 * pMy0 + tr33 * pMz0;
                const double pSXl = tr11 * pSx0 + tr12 * pSy0 + tr13 * pSz0;
                const double pSYl = tr21 * pSx0 + tr22 * pSy0 + tr23 * pSz0;
                const double pSZl = tr31 * pSx0 + tr32 * pSy0 + tr33 * pSz0;
                if (Plocal) {  // :170-175
                    Plocal[0 * N + id] = pMXl;
                    Plocal[1 * N + id] = pMYl;
                    Plocal[2 * N + id] = pMZl;
                    Plocal[3 * N + id] = pSXl;
                    Plocal[4 * N + id] = pSYl;
                    Plocal[5 * N + id] = pSZl;
                }
                // :178-199
                const double dfdPmxl = a0_2 * pMXl;
                const double dfdPmyl = a0_2 * pMYl;
                const double dfdPsxl = a0_2 * pSXl;
                const double dfdPsyl = a0_2 * pSYl;
                const double phiMz2 = pMZl * pMZl;
                const double phiMz3 = phiMz2 * pMZl;
                const double phiMz4 = phiMz3 * pMZl;
                const double phiMz5 = phiMz4 * pMZl;
                const double phiSz2 = pSZl * pSZl;
                const double phiSz3 = phiSz2 * pSZl;
                const double phiSz4 = phiSz3 * pSZl;
                const double phiSz5 = phiSz4 * pSZl;
                const double dfdPmzl = alpha_g_nd * pMZl + beta_4 * phiMz3 + beta_12 * pMZl * phiSz2 +
                                       gamma_6 * (phiMz5 + 10.0 * phiMz3 * phiSz2 + 5.0 * pMZl * phiSz4);
                const double dfdPszl = alpha_g_nd_s * pSZl + beta_4 * phiSz3 + beta_12 * pSZl * phiMz2 +
                                       gamma_6 * (phiSz5 + 10.0 * phiMz2 * phiSz3 + 5.0 * pSZl * phiMz4);
                // :202-208 back to the lab frame with TR^T
                double dfdPmx_b = tr11 * dfdPmxl + tr21 * dfdPmyl + tr31 * dfdPmzl;
                double dfdPmy_b = tr12 * dfdPmxl + tr22 * dfdPmyl + tr32 * dfdPmzl;
                double dfdPmz_b = tr13 * dfdPmxl + tr23 * dfdPmyl + tr33 * dfdPmzl;
                const double dfdPsx_b = tr11 * dfdPsxl + tr21 * dfdPsyl + tr31 * dfdPszl;
                const double dfdPsy_b = tr12 * dfdPsxl + tr22 * dfdPsyl + tr32 * dfdPszl;
                const double dfdPsz_b = tr13 * dfdPsxl + tr23 * dfdPsyl + tr33 * dfdPszl;
                // :210-220
                const double dfdPmz_app = -Ez_ext;
                const double A_loc = kappa_nd * del_t / (del_h * del_h);
                const double dfdPmx_e = (dep_flag != 0) ? -1.0 * E3[id] : 0.0;
                const double dfdPmy_e = (dep_flag != 0) ? -1.0 * E3[N + id] : 0.0;
                const double dfdPmz_e = (dep_flag != 0) ? -1.0 * E3[2 * N + id] : 0.0;
                dfdPmx_b += dfdPmx_e;
                dfdPmy_b += dfdPmy_e;
                dfdPmz_b += dfdPmz_app + dfdPmz_e;
                // :223-229
                Pnew[0 * N + id] = pMx0 * (1.0 - 6.0 * A_loc) + A_loc * phiMxsum - Lfac * del_t * dfdPmx_b;
                Pnew[1 * N + id] = pMy0 * (1.0 - 6.0 * A_loc) + A_loc * phiMysum - Lfac * del_t * dfdPmy_b;
                Pnew[2 * N + id] = pMz0 * (1.0 - 6.0 * A_loc) + A_loc * phiMzsum - Lfac * del_t * dfdPmz_b;
                Pnew[3 * N + id] = pSx0 * (1.0 - 6.0 * A_loc) + A_loc * phiSxsum - Lfac * del_t * dfdPsx_b;
                Pnew[4 * N + id] = pSy0 * (1.0 - 6.0 * A_loc) + A_loc * phiSysum - Lfac * del_t * dfdPsy_b;
                Pnew[5 * N + id] = pSz0 * (1.0 - 6.0 * A_loc) + A_loc * phiSzsum - Lfac * del_t * dfdPsz_b;
            }
}

// depol_varEPS.cu:279-329
void hzo_oracle_divergence(const hzo_oracle_cfg* c, const double* P, double* divP) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const long long N = (long long)nx * ny * nz;
    const double *Px = P, *Py = P + N, *Pz = P + 2 * N;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int e = (i + 1) % nx, w = (i - 1 + nx) % nx;
                const int n = (j + 1) % ny, s = (j - 1 + ny) % ny;
                const int u = (k == nz - 1) ? k : (k + 1);
                const int d = (k == 0) ? k : (k - 1);
                const long long id = cid(i, j, k, nx, ny);
                double v = ((Px[cid(e, j, k, nx, ny)] - Px[cid(w, j, k, nx, ny)]) +
                            (Py[cid(i, n, k, nx, ny)] - Py[cid(i, s, k, nx, ny)]) +
                            (Pz[cid(i, j, u, nx, ny)] - Pz[cid(i, j, d, nx, ny)])) /
                           (2.0 * c->del_h);
                v *= c->eps0_nd;
                divP[id] = v;
            }
}

// init_conf.cu:92-105 (eps_m = N.eps_r of the voxel's phase); void voxels
// keep the cudaMemset zero of allocateVars.cu:24.
void hzo_oracle_eps_map(const hzo_oracle_cfg* c, const int32_t* omega, int ng,
                        const int32_t* grain_phase, const hzo_oracle_phase ph[3],
                        double* eps) {
    const long long N = (long long)c->nx * c->ny * c->nz;
    for (long long id = 0; id < N; ++id) {
        const int gid = omega[id];
        if (gid < 0 || gid >= ng) {
            eps[id] = 0.0;
            continue;
        }
        const int mat = grain_phase[gid];
        eps[id] = ph[mat == 2 ? 2 : (mat == 1 ? 1 : 0)].eps_r;
    }
}

static inline void z_nb(const hzo_oracle_cfg* c, int k, int* up, int* dn) {
    const int nz = c->nz;
    if (c->zbc == 0) {
        *up = k + 1;
        *dn = k - 1;
    } else {
        if (k == 0) {
            *up = 1;
            *dn = 1;
        } else if (k == nz - 1) {
            *up = nz - 2;
            *dn = nz - 2;
        } else {
            *up = k + 1;
            *dn = k - 1;
        }
    }
}

static inline long long ref_node(const hzo_oracle_cfg* c) {  // depol_varEPS.cu:341-345
    return cid(c->nx / 2, c->ny / 2, c->nz / 2, c->nx, c->ny);
}

// depol_varEPS.cu:37-124
void hzo_oracle_apply_A(const hzo_oracle_cfg* c, const double* eps, const double* x, double* y) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const double h = c->del_h;
    const long long ref = ref_node(c);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const long long id = cid(i, j, k, nx, ny);
                if (c->zbc == 1 && id == ref) {
                    y[id] = x[id];
                    continue;
                }
                if (c->zbc == 0) {
                    if (k == 0) {
                        y[id] = x[id] - c->g_bot;
                        continue;
                    }
                    if (k == nz - 1) {
                        y[id] = x[id] - c->g_top;
                        continue;
                    }
                }
                const int ip = (i + 1) % nx, im = (i - 1 + nx) % nx;
                const int jp = (j + 1) % ny, jm = (j - 1 + ny) % ny;
                const double ec = eps[id];
                const double ee = eps[cid(ip, j, k, nx, ny)], ew = eps[cid(im, j, k, nx, ny)];
                const double en = eps[cid(i, jp, k, nx, ny)], es = eps[cid(i, jm, k, nx, ny)];
                const double ex_p = harm(ec, ee), ex_m = harm(ec, ew);
                const double ey_p = harm(ec, en), ey_m = harm(ec, es);
                int up, dn;
                z_nb(c, k, &up, &dn);
                const long long iu = cid(i, j, up, nx, ny), idn = cid(i, j, dn, nx, ny);
                const double ez_p = harm(ec, eps[iu]), ez_m = harm(ec, eps[idn]);
                const double xc = x[id];
                const double xe = x[cid(ip, j, k, nx, ny)], xw = x[cid(im, j, k, nx, ny)];
                const double xn = x[cid(i, jp, k, nx, ny)], xs = x[cid(i, jm, k, nx, ny)];
                const double xu = x[iu], xd = x[idn];
                const double Ax = (ex_p * (xe - xc) - ex_m * (xc - xw) + ey_p * (xn - xc) - ey_m * (xc - xs) +
                                   ez_p * (xu - xc) - ez_m * (xc - xd)) /
                                  (h * h);
                y[id] = -Ax;
            }
}

// depol_varEPS.cu:126-180
void hzo_oracle_diag_inv(const hzo_oracle_cfg* c, const double* eps, double* dinv) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const double h = c->del_h;
    const long long ref = ref_node(c);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const long long id = cid(i, j, k, nx, ny);
                if (c->zbc == 1 && id == ref) {
                    dinv[id] = 1.0;
                    continue;
                }
                if (c->zbc == 0 && (k == 0 || k == nz - 1)) {
                    dinv[id] = 1.0;
                    continue;
                }
                const int ip = (i + 1) % nx, im = (i - 1 + nx) % nx;
                const int jp = (j + 1) % ny, jm = (j - 1 + ny) % ny;
                const double ec = eps[id];
                const double ex_p = harm(ec, eps[cid(ip, j, k, nx, ny)]);
                const double ex_m = harm(ec, eps[cid(im, j, k, nx, ny)]);
                const double ey_p = harm(ec, eps[cid(i, jp, k, nx, ny)]);
                const double ey_m = harm(ec, eps[cid(i, jm, k, nx, ny)]);
                int up, dn;
                z_nb(c, k, &up, &dn);
                const double ez_p = harm(ec, eps[cid(i, j, up, nx, ny)]);
                const double ez_m = harm(ec, eps[cid(i, j, dn, nx, ny)]);
                const double D = (ex_p + ex_m + ey_p + ey_m + ez_p + ez_m) / (h * h);
                dinv[id] = (D > 1e-30 ? 1.0 / D : 1.0);
            }
}

// depol_varEPS.cu:223-276
void hzo_oracle_efield(const hzo_oracle_cfg* c, const double* phi, double* E3) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const long long N = (long long)nx * ny * nz;
    const double del_h = c->del_h;
    double *Ex = E3, *Ey = E3 + N, *Ez = E3 + 2 * N;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const long long id = cid(i, j, k, nx, ny);
                const int e = (i + 1) % nx, w = (i - 1 + nx) % nx;
                const int n = (j + 1) % ny, s = (j - 1 + ny) % ny;
                Ex[id] = -(phi[cid(e, j, k, nx, ny)] - phi[cid(w, j, k, nx, ny)]) / (2.0 * del_h);
                Ey[id] = -(phi[cid(i, n, k, nx, ny)] - phi[cid(i, s, k, nx, ny)]) / (2.0 * del_h);
                if (c->zbc == 0) {
                    const int up = (k == nz - 1 ? k : k + 1);
                    const int dn = (k == 0 ? k : k - 1);
                    const double phi_up = (k == nz - 1 ? c->g_top : phi[cid(i, j, up, nx, ny)]);
                    const double phi_dn = (k == 0 ? c->g_bot : phi[cid(i, j, dn, nx, ny)]);
                    Ez[id] = -(phi_up - phi_dn) / (2 * del_h);
                } else {
                    int up, dn;
                    if (k == 0) {
                        up = 1;
                        dn = 1;
                    } else if (k == nz - 1) {
                        up = nz - 2;
                        dn = nz - 2;
                    } else {
                        up = k + 1;
                        dn = k - 1;
                    }
                    Ez[id] = -(phi[cid(i, j, up, nx, ny)] - phi[cid(i, j, dn, nx, ny)]) / (2 * del_h);
                }
            }
}

// depol_varEPS.cu:332-595 (the work of Depol_field, without its per-call
// allocations, cuBLAS handle and ones vector, which do not affect results).
void hzo_oracle_depol(const hzo_oracle_cfg* c, const double* eps, const double* P,
                      double* phi, double* E3, double* divP_out,
                      hzo_oracle_pcg_stats* st, double* rz_hist, int max_hist) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const long long N = (long long)nx * ny * nz;
    std::vector<double> b(N), r(N), z(N), p(N), Ap(N), dinv(N);
    double* x = phi;  // warm start (:429-430), result copied back (:543)

    // (1) b = -(eps0 * div Pm) (:382, :401-402)
    hzo_oracle_divergence(c, P, b.data());
    if (divP_out) std::memcpy(divP_out, b.data(), sizeof(double) * N);
    for (long long t = 0; t < N; ++t) b[t] = -1.0 * b[t];
    if (c->zbc == 1) {
        b[ref_node(c)] = 0.0;  // :405-414
    } else {
        for (int j = 0; j < ny; ++j)  // :417-424
            for (int i = 0; i < nx; ++i) {
                b[cid(i, j, 0, nx, ny)] = 0.0;
                b[cid(i, j, nz - 1, nx, ny)] = 0.0;
            }
    }

    // (2) PCG (:433-523)
    hzo_oracle_diag_inv(c, eps, dinv.data());
    hzo_oracle_apply_A(c, eps, x, Ap.data());
    for (long long t = 0; t < N; ++t) r[t] = -1.0 * Ap[t] + b[t];  // Dcopy + Daxpy(-1)
    for (long long t = 0; t < N; ++t) z[t] = r[t] * dinv[t];
    std::memcpy(p.data(), z.data(), sizeof(double) * N);
    double rz_old = plane_sum(nx, ny, nz, [&](long long t) { return r[t] * z[t]; });
    double rr = plane_sum(nx, ny, nz, [&](long long t) { return r[t] * r[t]; });
    const double bb = plane_sum(nx, ny, nz, [&](long long t) { return b[t] * b[t]; });
    const double rz0 = rz_old;
    st->rz0 = rz0;
    st->rr0 = rr;
    st->bb = bb;
    st->status = 0;
    if (rz_hist && max_hist > 0) rz_hist[0] = rz0;

    int kiter = 0;
    const double tabs2 = c->tol_abs * c->tol_abs;
    const double trel2 = c->tol_rel * c->tol_rel;
    auto keep_going = [&](double rz, double rrv) {
        if (kiter >= c->max_iter) return false;
        if (c->pcg_rule == 0) return rz > tabs2 && (rz0 <= 0.0 || (rz / rz0) > trel2);
        return rrv > trel2 * bb && rrv > tabs2;
    };
    if (!std::isfinite(rz_old) || rz_old < 0.0) {
        st->status = -1;  // :466-468
    } else {
        while (keep_going(rz_old, rr)) {
            ++kiter;
            hzo_oracle_apply_A(c, eps, p.data(), Ap.data());
            const double denom = plane_sum(nx, ny, nz, [&](long long t) { return p[t] * Ap[t]; });
            if (!std::isfinite(denom) || !(denom > 0.0)) {  // :486-490
                st->status = -2;
                break;
            }
            const double alpha = rz_old / denom;
            for (long long t = 0; t < N; ++t) x[t] = alpha * p[t] + x[t];
            const double nalpha = -alpha;
            for (long long t = 0; t < N; ++t) r[t] = nalpha * Ap[t] + r[t];
            for (long long t = 0; t < N; ++t) z[t] = r[t] * dinv[t];
            const double rz_new = plane_sum(nx, ny, nz, [&](long long t) { return r[t] * z[t]; });
            rr = plane_sum(nx, ny, nz, [&](long long t) { return r[t] * r[t]; });
            if (rz_hist && kiter < max_hist) rz_hist[kiter] = rz_new;
            if (!std::isfinite(rz_new) || rz_new < 0.0) {  // :508-513
                st->status = -3;
                rz_old = rz_new;
                break;
            }
            const double beta = rz_new / rz_old;
            for (long long t = 0; t < N; ++t) p[t] = beta * p[t];  // Dscal
            for (long long t = 0; t < N; ++t) p[t] = 1.0 * z[t] + p[t];  // Daxpy(1)
            rz_old = rz_new;
        }
        if (st->status == 0 && kiter >= c->max_iter) st->status = 1;
    }
    st->iters = kiter;
    st->rz = rz_old;
    st->rr = rr;
    // (3) E = -grad phi (:547-550)
    hzo_oracle_efield(c, x, E3);
}

// evolve.cu:220-237: kernelCalc_antiferro, 2x kernelUpdate_P, Depol_field.
void hzo_oracle_step(const hzo_oracle_cfg* c, double* P, double* work, double* Plocal,
                     const int32_t* omega, int ng, const int32_t* grain_phase,
                     const double* TR, const hzo_oracle_phase ph[3],
                     const int32_t* iface, const double* eps, double* phi, double* E3,
                     double Ez_ext, hzo_oracle_pcg_stats* st) {
    const long long N = (long long)c->nx * c->ny * c->nz;
    hzo_oracle_tdgl(c, P, work, Plocal, omega, ng, grain_phase, TR, ph, iface, E3, Ez_ext);
    std::memcpy(P, work, sizeof(double) * 6 * N);
    if (c->dep_flag) {
        hzo_oracle_depol(c, eps, P, phi, E3, nullptr, st, nullptr, 0);
    } else if (st) {
        std::memset(st, 0, sizeof(*st));
    }
}

void hzo_oracle_run(const hzo_oracle_cfg* c, double* P, double* Plocal,
                    const int32_t* omega, int ng, const int32_t* grain_phase,
                    const double* TR, const hzo_oracle_phase ph[3],
                    const int32_t* iface, double* phi, double* E3,
                    const double* Ez, int nsteps, int* its_out) {
    const long long N = (long long)c->nx * c->ny * c->nz;
    std::vector<double> work(6 * N), eps(N);
    hzo_oracle_eps_map(c, omega, ng, grain_phase, ph, eps.data());
    for (int s = 0; s < nsteps; ++s) {
        hzo_oracle_pcg_stats st;
        std::memset(&st, 0, sizeof(st));
        hzo_oracle_step(c, P, work.data(), Plocal, omega, ng, grain_phase, TR, ph, iface,
                        eps.data(), phi, E3, Ez[s], &st);
        if (its_out) its_out[s] = st.iters;
    }
}

// main.cu:13-42 (grain ids >= the grain count are treated as void).
void hzo_oracle_interface_flag(const hzo_oracle_cfg* c, const int32_t* omega,
                               const int32_t* grain_phase, int32_t* iface) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const long long N = (long long)nx * ny * nz;
    const long long plane = (long long)nx * ny;
    for (long long id = 0; id < N; ++id) {
        const int gid = omega[id];
        if (gid < 0) {
            iface[id] = 0;
            continue;
        }
        if (grain_phase[gid] != 1) {
            iface[id] = 0;
            continue;
        }
        const long long k = id / plane;
        if (k == 0 || k == nz - 1) {
            iface[id] = 0;
            continue;
        }
        const int gu = omega[id + plane], gd = omega[id - plane];
        const int mu = (gu >= 0) ? grain_phase[gu] : -1;
        const int md = (gd >= 0) ? grain_phase[gd] : -1;
        iface[id] = (mu != 1 || md != 1);
    }
}

// evolve.cu:259-357 with build_mat_mask_and_p (kernels_evol.cu:23-44).
void hzo_oracle_diagnostics(const hzo_oracle_cfg* c, const double* P, const double* Plocal,
                            const int32_t* omega, const int32_t* grain_phase,
                            double out[10]) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const long long N = (long long)nx * ny * nz;
    const double* Pmz = P + 2 * N;
    const double* Psz = P + 5 * N;
    out[0] = plane_sum(nx, ny, nz, [&](long long t) { return Pmz[t]; });
    out[1] = plane_sum(nx, ny, nz, [&](long long t) { return Psz[t]; });
    out[2] = plane_sum(nx, ny, nz, [&](long long t) { return Plocal[2 * N + t]; });
    out[3] = plane_sum(nx, ny, nz, [&](long long t) { return Plocal[5 * N + t]; });
    const int mats[3] = {1, 2, 0};  // FE, DE, AFE (evolve.cu:285, 313, 338)
    for (int q = 0; q < 3; ++q) {
        const int target = mats[q];
        auto mask = [&](long long t) {
            const int gid = omega[t];
            return (gid >= 0 && grain_phase[gid] == target) ? 1.0 : 0.0;
        };
        out[4 + 2 * q] = plane_sum(nx, ny, nz, [&](long long t) { return mask(t) * Pmz[t]; });
        out[5 + 2 * q] = plane_sum(nx, ny, nz, [&](long long t) { return mask(t); });
    }
}

// voronoi.c:118-153 for one layer.
void hzo_oracle_voronoi_layer(int nx, int ny, int nz, int z0, int span,
                              const double* cx, const double* cy, const double* cz,
                              int nL, int global_offset, int32_t* omega_zf) {
    const int z1 = z0 + span;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int h = z0; h < z1; ++h) {
                double rmin = 1e300;
                int nmin = 0;
                for (int k = 0; k < nL; ++k) {
                    const double ccx = cx[global_offset + k];
                    const double ccy = cy[global_offset + k];
                    const double ccz = cz[global_offset + k];
                    double dx = ccx - (double)i;
                    if (dx > 0.5 * nx) dx -= nx;
                    else if (dx < -0.5 * nx) dx += nx;
                    double dy = ccy - (double)j;
                    if (dy > 0.5 * ny) dy -= ny;
                    else if (dy < -0.5 * ny) dy += ny;
                    const double hz = (double)h - (double)z0;
                    const double cz_local = ccz - (double)z0;
                    double dz = cz_local - hz;
                    if (dz > 0.5 * span) dz -= span;
                    else if (dz < -0.5 * span) dz += span;
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 < rmin) {
                        rmin = r2;
                        nmin = k;
                    }
                }
                omega_zf[(long long)h + ((long long)j + (long long)i * ny) * nz] = global_offset + nmin;
            }
}

// grain_io.cu:61-69
void hzo_oracle_reorder_zf_to_c(int nx, int ny, int nz, const int32_t* in_zf, int32_t* out_c) {
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const long long id_f = k + (long long)nz * (j + (long long)i * ny);
                const long long id_c = i + (long long)nx * (j + (long long)k * ny);
                out_c[id_c] = in_zf[id_f];
            }
}

// Thread count of the oracle's parallel loops (the CPU baseline times it on
// all cores and on one); n <= 0 leaves the OpenMP default.
void hzo_oracle_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

int hzo_oracle_get_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
