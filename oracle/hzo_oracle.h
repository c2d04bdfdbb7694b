/*
 * hzo_oracle.h -- CPU FP64 restatement of the HZO phase-field time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in cmsl_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product links it.
 *
 * It restates, in the reference's evaluation order, the CUDA kernels and the
 * host control flow of one simulator step of jhb6617-arch/cmsl:
 *   HZO_Code_github/src/kernels_evol.cu:47-242   (TDGL update + new->old copy)
 *   HZO_Code_github/src/depol_varEPS.cu:6-595    (div P, RHS, variable-eps
 *                                                 operator, Jacobi PCG, E=-grad phi)
 *   HZO_Code_github/src/evolve.cu:217-241,259-357 (step order, diagnostics)
 *   HZO_Code_github/src/main.cu:13-42            (interface flag)
 *   HZO_Code_github/src/grain_io.cu:61-69        (z-fastest -> C reorder)
 *   genpolycr_layered/voronoi.c:118-153          (nearest-centroid labels)
 *
 * Conventions (shared with include/hzo.h):
 *   cell index id = i + nx*(j + ny*k)  (x fastest, C order)
 *   P buffer  = 6 planes of N doubles: Pmx, Pmy, Pmz, Psx, Psy, Psz
 *   E buffer  = 3 planes of N doubles: Ex, Ey, Ez
 *   phases    = 0 AFE, 1 FE, 2 DE (reference types.h:12)
 *
 * Reductions (PCG inner products, diagnostics) are evaluated as per-k-plane
 * sequential partial sums added in k order, so the result does not depend on
 * the OpenMP thread count.  The reference's cuBLAS/CUB summation order is not
 * reproducible (closed library); results agree to rounding only.
 */
#ifndef HZO_ORACLE_H
#define HZO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Non-dimensional per-phase coefficients (reference MaterialNorm, types.h:22-27). */
typedef struct hzo_oracle_phase {
    double alpha, beta, gamma, g, a0, kappa, L, eps_r;
} hzo_oracle_phase;

typedef struct hzo_oracle_cfg {
    int nx, ny, nz;
    double del_h, del_t;
    int dep_flag;       /* 0: no depolarisation solve (evolve.cu:236) */
    int zbc;            /* 0 Dirichlet ("diri"), 1 Neumann ("neu") (main.cu:424-425) */
    double eps0_nd;     /* main.cu:344 */
    double g_bot, g_top;/* electrode potentials, global.cu:140-141 */
    int pcg_rule;       /* 0: reference rule (depol_varEPS.cu:472-474); 1: ||r||2 <= tol_rel*||b||2 */
    double tol_abs, tol_rel;
    int max_iter;
} hzo_oracle_cfg;

typedef struct hzo_oracle_pcg_stats {
    int iters;          /* PCG iterations performed */
    int status;         /* 0 converged/stopped by rule, 1 max_iter, -1 invalid initial residual,
                           -2 breakdown p.Ap, -3 invalid r.z */
    double rz0, rz;     /* first and last preconditioned residual r.z */
    double rr0, rr;     /* first and last ||r||^2 */
    double bb;          /* ||b||^2 */
} hzo_oracle_pcg_stats;

/* a1: one explicit-Euler TDGL step (kernels_evol.cu:47-231).  Pold/Pnew: 6N.
 * Plocal (6N, may be NULL) receives the grain-frame copies (:170-175).
 * E3 (3N) is the depolarisation field used when dep_flag != 0.
 * iface (N) may be NULL (no interface voxels). */
void hzo_oracle_tdgl(const hzo_oracle_cfg* c, const double* Pold, double* Pnew,
                     double* Plocal, const int32_t* omega, int ng,
                     const int32_t* grain_phase, const double* TR,
                     const hzo_oracle_phase ph[3], const int32_t* iface,
                     const double* E3, double Ez_ext);

/* a4: divP = eps0_nd * div(Pm) with clamped z (depol_varEPS.cu:279-329). */
void hzo_oracle_divergence(const hzo_oracle_cfg* c, const double* P, double* divP);

/* Per-voxel permittivity map (init_conf.cu:105; 0 for void voxels, allocateVars.cu:24). */
void hzo_oracle_eps_map(const hzo_oracle_cfg* c, const int32_t* omega, int ng,
                        const int32_t* grain_phase, const hzo_oracle_phase ph[3],
                        double* eps);

/* a8: y = A x (depol_varEPS.cu:37-124). */
void hzo_oracle_apply_A(const hzo_oracle_cfg* c, const double* eps, const double* x, double* y);

/* a7: Jacobi inverse diagonal (depol_varEPS.cu:126-180). */
void hzo_oracle_diag_inv(const hzo_oracle_cfg* c, const double* eps, double* dinv);

/* a11: E = -grad phi (depol_varEPS.cu:223-276). */
void hzo_oracle_efield(const hzo_oracle_cfg* c, const double* phi, double* E3);

/* a3: Depol_field (depol_varEPS.cu:332-595).  phi is the warm start on input
 * and the solution on output; E3 receives -grad phi; divP (N, may be NULL)
 * receives eps0*div P.  rz_hist (max_hist entries, may be NULL) receives r.z
 * after each iteration (entry 0 = rz0). */
void hzo_oracle_depol(const hzo_oracle_cfg* c, const double* eps, const double* P,
                      double* phi, double* E3, double* divP,
                      hzo_oracle_pcg_stats* st, double* rz_hist, int max_hist);

/* Full step (evolve.cu:220-237): TDGL, swap, Depol if dep_flag.
 * P (6N) is updated in place; work (6N) is scratch. */
void hzo_oracle_step(const hzo_oracle_cfg* c, double* P, double* work, double* Plocal,
                     const int32_t* omega, int ng, const int32_t* grain_phase,
                     const double* TR, const hzo_oracle_phase ph[3],
                     const int32_t* iface, const double* eps, double* phi, double* E3,
                     double Ez_ext, hzo_oracle_pcg_stats* st);

/* n steps with per-step applied field Ez[n]; its_out[n] (may be NULL) receives
 * PCG iterations per step.  Plocal (6N, may be NULL) holds the last step's
 * grain-frame copies. */
void hzo_oracle_run(const hzo_oracle_cfg* c, double* P, double* Plocal,
                    const int32_t* omega, int ng, const int32_t* grain_phase,
                    const double* TR, const hzo_oracle_phase ph[3],
                    const int32_t* iface, double* phi, double* E3,
                    const double* Ez, int nsteps, int* its_out);

/* a14: FE voxels with a non-FE z neighbour (main.cu:13-42). */
void hzo_oracle_interface_flag(const hzo_oracle_cfg* c, const int32_t* omega,
                               const int32_t* grain_phase, int32_t* iface);

/* a19: 10 sums (evolve.cu:259-357): sum Pmz, sum Psz, sum Pmz_local, sum Psz_local,
 * then for FE, DE, AFE: sum(mask*Pmz), sum(mask). */
void hzo_oracle_diagnostics(const hzo_oracle_cfg* c, const double* P, const double* Plocal,
                            const int32_t* omega, const int32_t* grain_phase,
                            double out[10]);

/* Nearest-centroid labels of one layer slab [z0, z0+span) (voronoi.c:118-153),
 * written into omega_zf (z-fastest order, id = h + nz*(j + ny*i)). */
void hzo_oracle_voronoi_layer(int nx, int ny, int nz, int z0, int span,
                              const double* cx, const double* cy, const double* cz,
                              int nL, int global_offset, int32_t* omega_zf);

/* z-fastest (generator file order) -> C order (grain_io.cu:61-69). */
void hzo_oracle_reorder_zf_to_c(int nx, int ny, int nz, const int32_t* in_zf, int32_t* out_c);

/* OpenMP threads of the oracle's loops (n <= 0: unchanged) and the current count. */
void hzo_oracle_set_threads(int n);
int hzo_oracle_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
