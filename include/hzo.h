/*
 * hzo.h -- C ABI of the MI355X-native HZO phase-field time step.
 *
 * The library replaces one simulator step of jhb6617-arch/cmsl (HZO
 * ferroelectric thin-film phase-field code): the explicit-Euler TDGL update
 * and the Jacobi-preconditioned CG Poisson solve for the depolarisation
 * potential.  In the reference the step is three call sites inside the
 * Evolve() loop (HZO_Code_github/src/evolve.cu:220-237):
 *   kernelCalc_antiferro<<<>>>(37 pointer/scalar args)   kernels_evol.h:9-18
 *   kernelUpdate_P<<<>>>(..) x2                          kernels_evol.h:20-21
 *   void Depol_field(void)                               header.h:32
 * operating on ~40 process-global device pointers (cuda_deps.h:13-41).
 * Here one opaque context owns all device state of one film; the caller hands
 * over host arrays once and then steps.  Every call returns HZO_OK (0) or a
 * negative status; hzo_last_error() gives the text (the reference prints and
 * calls exit(1) instead, check.h:9-83).  PCG breakdowns are not errors: they
 * are reported in hzo_pcg_stats.status, as the reference prints and continues
 * (depol_varEPS.cu:466-513).
 *
 * Layout conventions (identical to the reference, kernels_evol.cu:65):
 *   cell id = i + nx*(j + ny*k), x fastest ("C order");
 *   P       = 6 planes of N doubles: Pmx, Pmy, Pmz, Psx, Psy, Psz (non-dimensional);
 *   E       = 3 planes of N doubles: Ex, Ey, Ez;
 *   phases  = 0 AFE, 1 FE, 2 DE (types.h:12).
 *
 * Threading: a context is used by one host thread at a time.  Different
 * contexts are independent.
 */
#ifndef HZO_H
#define HZO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HZO_OK 0
#define HZO_ERR_ARG (-1)      /* invalid argument / inconsistent sizes */
#define HZO_ERR_HIP (-2)      /* HIP runtime error (no device, launch failure) */
#define HZO_ERR_IO (-3)       /* file missing, short read, parse error */
#define HZO_ERR_STATE (-4)    /* call order violated (e.g. step before set_state) */
#define HZO_ERR_NOMEM (-5)    /* device or host allocation failed */
#define HZO_ERR_DATA (-6)     /* input data inconsistent (layer sums, grain ids) */

enum { HZO_AFE = 0, HZO_FE = 1, HZO_DE = 2 };
enum { HZO_ZBC_DIRICHLET = 0, HZO_ZBC_NEUMANN = 1 };
enum { HZO_PREC_FP64 = 0, HZO_PREC_FP32_REFINED = 1 };
enum { HZO_PCG_RULE_REFERENCE = 0, HZO_PCG_RULE_RELRES = 1 };

/* PCG status codes (hzo_pcg_stats.status). */
#define HZO_PCG_CONVERGED 0
#define HZO_PCG_MAXITER 1
#define HZO_PCG_BAD_INITIAL (-1)   /* non-finite or negative r.z at start */
#define HZO_PCG_BREAKDOWN (-2)     /* p.Ap non-finite or <= 0 */
#define HZO_PCG_BAD_RZ (-3)        /* non-finite or negative r.z during iteration */

/* Non-dimensional coefficients of one phase (reference MaterialNorm,
 * types.h:22-27, as produced by make_nd_wrt_reference, main.cu:232-256). */
typedef struct hzo_phase_nd {
    double alpha, beta, gamma, g, a0, kappa, L, eps_r;
} hzo_phase_nd;

/* Raw (SI) coefficients of one phase (reference MaterialParamsRaw, types.h:15-20). */
typedef struct hzo_phase_raw {
    double alpha, beta, gamma, g, a0;
    int transition_type;
    double L, kappa, eps_r;
} hzo_phase_raw;

/* Step configuration.  hzo_config_default() fills the reference constants
 * (global.cu:30-31, 140-144). */
typedef struct hzo_config {
    int nx, ny, nz;
    double del_h, del_t;
    int dep_flag;        /* 1: depolarisation solve each step (InputParams dep_flag) */
    int zbc;             /* HZO_ZBC_* (InputParams z_bc_type diri|neu) */
    double eps0_nd;      /* 4/(8.854e-12*|alpha_ref|), main.cu:344 */
    double g_bot, g_top; /* electrode potentials, global.cu:140-141 */
    int precision;       /* HZO_PREC_* */
    int pcg_rule;        /* HZO_PCG_RULE_* */
    double tol_abs, tol_rel;
    int max_iter;
    int refine_every;    /* FP32-refined: FP64 residual recompute period (iterations) */
    int tile_kz;         /* z planes per workgroup of the stencil kernels: 2, 4, 8 (16: single-
                            reduction CG only; 32: its two-stencil form only); 0 = automatic */
    int cg_fused;        /* FP32-refined: 0 = the two-kernel PCG of the FP64 form; otherwise the
                            single-reduction CG (Chronopoulos-Gear): 2 = one fused kernel per
                            iteration carrying w = A u and s = A p in memory (42 bytes per voxel
                            and iteration), 3 = the same forming both by stencils in the launch
                            (26 bytes, more arithmetic), 4 = the whole solve in one cooperative
                            launch with the vectors held in registers and LDS (16 x 16 column
                            tiles: nx, ny multiples of 16, nz of 16, 24 or 32, every tile
                            resident at once), 1 = automatic (4 where it applies and tile_kz is
                            0; else 3 when the tiles are 16 planes deep, else 2).
                            Films of 2^27 voxels or more always run the two-kernel PCG. */
} hzo_config;

typedef struct hzo_pcg_stats {
    int iters;           /* PCG iterations of the last solve */
    int status;          /* HZO_PCG_* */
    int refinements;     /* FP64 residual replacements (FP32-refined path) */
    int reserved;
    double rz0, rz;      /* preconditioned residual r.z at start / end */
    double rr0, rr;      /* ||r||^2 at start / end */
    double bb;           /* ||b||^2 */
} hzo_pcg_stats;

typedef struct hzo_ctx hzo_ctx;

/* ---------------------------------------------------------------- library */
const char* hzo_last_error(void);
int hzo_version(void);
/* Number of visible HIP devices (0 without a GPU); returns HZO_OK. */
int hzo_device_count(int* n);
void hzo_config_default(hzo_config* cfg);

/* ---------------------------------------------------------------- context
 * Replaces AllocateData (allocateVars.cu:5-113) and the device globals. */
int hzo_create(const hzo_config* cfg, hzo_ctx** out);
void hzo_destroy(hzo_ctx* ctx);

/* Grain structure: omega[N] grain id per voxel (C order, <0 = void),
 * grain_phase[ng] (0 AFE,1 FE,2 DE), TR[9*ng] row-major grain rotations
 * (grain_io.cu:102-126).  iface[N] may be NULL: it is then built by the
 * reference rule (main.cu:13-42). */
int hzo_set_microstructure(hzo_ctx* ctx, const int32_t* omega, int ng,
                           const int32_t* grain_phase, const double* TR,
                           const int32_t* iface);
/* Per-phase non-dimensional coefficients, indexed by phase (replaces the
 * per-voxel maps of InitConfKernel, init_conf.cu:68-106, whose spread is 0). */
int hzo_set_phase_table(hzo_ctx* ctx, const hzo_phase_nd ph[3]);
/* Polarisation P[6N] and potential phi[N] (phi NULL: init_phi_linear,
 * depol_varEPS.cu:203-221).  E is always -grad phi of the current phi. */
int hzo_set_state(hzo_ctx* ctx, const double* P, const double* phi);
/* Any pointer may be NULL. */
int hzo_get_state(hzo_ctx* ctx, double* P, double* phi, double* E);

/* ---------------------------------------------------------------- step */
/* kernelCalc_antiferro + kernelUpdate_P (kernels_evol.cu:47-242). */
int hzo_tdgl_step(hzo_ctx* ctx, double Ez_ext);
/* Depol_field (depol_varEPS.cu:332-595); stats may be NULL. */
int hzo_depol_solve(hzo_ctx* ctx, hzo_pcg_stats* stats);
/* One step in reference order (evolve.cu:220-237); stats may be NULL. */
int hzo_step(hzo_ctx* ctx, double Ez_ext, hzo_pcg_stats* stats);
/* nsteps steps with applied field Ez[s]; its[s] (may be NULL) receives the PCG
 * iteration count of step s.  Issues no host synchronisation per step beyond
 * what the PCG stopping test needs: a film that runs the register-resident
 * solve enqueues up to 64 steps before it waits, so a failure of one of them
 * (a grid barrier that timed out) is returned when its group of steps is
 * collected, with the film's state then undefined. */
int hzo_run(hzo_ctx* ctx, const double* Ez, int nsteps, int* its);
/* Reference diagnostics (evolve.cu:259-357): out[10] = sum Pmz, sum Psz,
 * sum Pmz_local, sum Psz_local (grain frame, from the last TDGL call's input),
 * then FE, DE, AFE: sum(mask*Pmz), sum(mask). */
int hzo_diagnostics(hzo_ctx* ctx, double out[10]);
int hzo_synchronize(hzo_ctx* ctx);
/* Grid dimensions of a context: dims[3] = nx, ny, nz. */
int hzo_grid_dims(hzo_ctx* ctx, int dims[3]);
/* Tile depth (z planes per workgroup) the context launches with. */
int hzo_tile_depth(hzo_ctx* ctx);
/* Poisson solver the context runs: 0 two-kernel PCG (FP64 form or FP32
 * cg_fused = 0), 1 single-reduction CG carrying w and s, 2 single-reduction
 * CG forming them by stencils, 3 the register-resident single-reduction CG
 * (one launch per solve); -1 for a null context.  No reference counterpart
 * (the reference has one solver). */
int hzo_cg_form(hzo_ctx* ctx);

/* ---------------------------------------------------------------- sweep driver
 * The staircase P-E loop of Evolve() (evolve.cu:213-410): the field starts at
 * 0, changes by delEz every num_steps steps and turns at +-maxEz; the run ends
 * after one full loop (peloop == 2).  One 12-column row per level from the
 * first turn on, in the sweep%03d.txt format (evolve.cu:391-400): E [MV/cm],
 * <Pmz>, <Psz>, <Pmz_local>, <Psz_local>, <Pmz>_FE, f_FE, <Pmz>_AFE, f_AFE,
 * <Pmz>_DE, f_DE (P in uC/cm^2), step count. */
typedef struct hzo_sweep_params {
    double maxEz, delEz;   /* non-dimensional: Eap/E0_ref, delEap/E0_ref (main.cu:347-348) */
    int num_steps;         /* steps per field level (time_to_change) */
    int max_levels;        /* safety cap on levels, 0 = none */
    double Ez_scale;       /* E0_ref/1e8: non-dimensional -> MV/cm (evolve.cu:117) */
    double P0_scale;       /* P0_ref*100: non-dimensional -> uC/cm^2 (evolve.cu:118) */
} hzo_sweep_params;
/* Host-only: field (MV/cm) and step count of every row the sweep writes. */
int hzo_sweep_schedule(const hzo_sweep_params* p, int max_rows, double* E_MVcm, int* counts, int* nrows);
/* Runs the sweep on a context.  path (may be NULL) is opened in append mode;
 * rows_out (may be NULL) receives up to max_rows rows of 12 doubles. */
int hzo_sweep_run(hzo_ctx* ctx, const hzo_sweep_params* p, const char* path, int max_rows,
                  double* rows_out, int* nrows);

/* ---------------------------------------------------------------- timing
 * Kernel-class timing with HIP events attached to the kernels' own
 * dispatches on the context's stream.  While enabled, the first
 * HZO_TIMING_MAX launches of every class are timed; hzo_kernel_times returns,
 * per class, the number of working launches and their mean duration in ms
 * (launches that found the solve already finished, under a quarter of the
 * class median, are left out).  Classes: 0 tdgl, 1 pcg_setup, 2 pcg_dir_spmv,
 * 3 pcg_update (two-kernel PCG), 4 pcg_fused (single-reduction CG launch of
 * the FP32-refined form), 5 pcg_resident (one resident solve). */
#define HZO_TIMING_MAX 8192
#define HZO_KCLASSES 6
int hzo_set_timing(hzo_ctx* ctx, int enable);
int hzo_kernel_times(hzo_ctx* ctx, int counts[HZO_KCLASSES], double mean_ms[HZO_KCLASSES]);
/* Resident solves since hzo_set_timing: out[0] solves, out[1] CG iterations,
 * out[2] grid barriers (one per iteration, three per residual replacement,
 * one to re-form w after the setup). */
int hzo_cg_counters(hzo_ctx* ctx, long long out[3]);

/* ---------------------------------------------------------------- reverse mode
 * Gradients of a P-E-loop loss with respect to the per-phase Landau and
 * gradient coefficients (the alpha, beta, gamma, kappa of the hzo_phase_nd
 * table), by hand-written adjoint kernels of the TDGL step and the same
 * (self-adjoint) PCG for the Poisson solve.  No counterpart in the
 * reference, which has no derivative path; it serves the Landau-coefficient
 * fitting side (torch.autograd.Function in cmsl_amd/adjoint.py).
 *
 * Loop sample of step n: m_n = <Pmz> (mean over all voxels) after step n.
 * Forward states are checkpointed every ckpt_every steps (P and phi) and the
 * steps of each segment are recomputed during the backward pass, so device
 * memory is (nsteps/ckpt_every + 1 + ckpt_every) states.  Needs Dirichlet
 * electrodes when the depolarisation solve is on. */
/* Allocates the checkpoint, segment and adjoint buffers for up to nsteps
 * steps; tol is the relative residual of the adjoint Poisson solves. */
int hzo_adjoint_reserve(hzo_ctx* ctx, int nsteps, int ckpt_every, double tol);
/* Runs nsteps steps from the current state with applied field Ez[n],
 * records checkpoints and writes m_out[n] (may be NULL). */
int hzo_forward_record(hzo_ctx* ctx, const double* Ez, int nsteps, double* m_out);
/* dL_dm[n] = dL/dm_n for the recorded run.  grad[p*4 + q]: phase p (0 AFE,
 * 1 FE, 2 DE), q = 0 alpha, 1 beta, 2 gamma, 3 kappa (non-dimensional).
 * adj_pcg_iters (may be NULL) receives the adjoint PCG iterations.  Leaves
 * the context at the recorded final state. */
int hzo_backward(hzo_ctx* ctx, const double* dL_dm, double grad[12], long long* adj_pcg_iters);

/* ---------------------------------------------------------------- dumps and restart */
/* save_snapshot (generate.cu:266-303): Pm and Ps of the current state as
 * interleaved (x, y, z) f64 per voxel, multiplied by P0_scale. */
int hzo_save_snapshot(hzo_ctx* ctx, const char* pm_path, const char* ps_path, double P0_scale);
/* write_checkpoint (generate.cu:90-168): header nx, ny, nz, peloop, count
 * (int32), Ez_ext, incEz (f64), then Pmx..Psz and Ex, Ey, Ez of the current
 * state (N f64 each, non-dimensional). */
int hzo_checkpoint_save(hzo_ctx* ctx, const char* path, int peloop, int count, double Ez_ext, double incEz);
/* read_checkpoint (generate.cu:170-264) as a restart: sets P from the file
 * and re-solves phi from it (the file holds E, not phi).  Header values and
 * E_file[3N] (the stored field) may be NULL. */
int hzo_checkpoint_load(hzo_ctx* ctx, const char* path, int* peloop, int* count, double* Ez_ext, double* incEz,
                        double* E_file);

/* ---------------------------------------------------------------- batched films
 * Many films of different sizes and coefficients stepped together: each
 * kernel of the FP32-refined step (TDGL, PCG setup, single-reduction CG
 * iteration, diagnostics) runs once per group of films over a tile table
 * that maps every workgroup to (film, tile); each film keeps its own solver
 * control block.  A group's step (TDGL, PCG setup, first chunk of CG
 * launches) is replayed from a captured hipGraph, the further CG launches
 * from a chunk graph until every film of the group has converged.  Films are packed
 * into groups whose CG working set fits a cache budget, and each group runs
 * all steps of a call before the next group starts; a group of one film runs
 * that film's own single-film kernels at its own tile depth.  No counterpart in the
 * reference, which steps one film per process (evolve.cu:220-237 in a loop
 * over decks); it serves parameter sweeps over many microstructures.
 *
 * The films must be FP32-refined (HZO_PREC_FP32_REFINED, cg_fused != 0), on
 * one device, with the same dep_flag, and ready (microstructure, phase table
 * and state set).  A batch holds the films' device buffers by reference: the
 * films must outlive it and must not be stepped concurrently with it.
 * Films whose own solve is the register-resident one (hzo_cg_form 3) are
 * packed by depth into groups that fit the device together, and each such
 * group's solves run as one cooperative launch; each of these films ends a
 * batched run in exactly the state its own hzo_run leaves.  The other films
 * go into groups by CG working set: a film alone in its group runs its own
 * kernels, the films of a larger group run the batch kernels at the batch's
 * tile depth, and each ends in exactly the state its own hzo_run leaves at
 * that tile depth. */
typedef struct hzo_batch hzo_batch;
/* tile_kz: 2, 4 or 8 (planes per workgroup tile); all films in one group. */
int hzo_batch_create(hzo_ctx* const* films, int n, int tile_kz, hzo_batch** out);
void hzo_batch_destroy(hzo_batch* batch);
/* use_graphs: replay each step and its CG launches from captured hipGraphs
 * (default 1); chunk: CG launches per replay, even, >= 2 (default 16). */
int hzo_batch_config(hzo_batch* batch, int use_graphs, int chunk);
/* lanes (1 .. 4, default 2): host threads, each with its own stream, over
 * which the groups that do not run the resident solve are shared out, so
 * that their launches overlap on the GPU (resident groups always run alone).
 * Every film ends in the same state whatever the number of lanes. */
int hzo_batch_set_lanes(hzo_batch* batch, int lanes);
/* Tile depth (2, 4 or 8) of the groups of several films, and group budget in
 * MiB of CG working set (34 bytes per voxel; 0: one group of all films). */
int hzo_batch_set_shape(hzo_batch* batch, int tile_kz, double group_mib);
/* out[3] = tile depth, group budget (MiB), number of groups. */
int hzo_batch_shape(const hzo_batch* batch, int out[3]);
int hzo_batch_tile_depth(hzo_batch* batch);
/* out[3] = group index, films in that group, and how the group runs: 0 batch
 * kernels over several films, 1 the film's own kernels, 2 the resident solve
 * (one cooperative launch for the group). */
int hzo_batch_film_group(const hzo_batch* batch, int film, int out[3]);
/* nsteps steps; Ez[s*n + f] is the applied field of film f at step s.
 * iters[s*n + f] (may be NULL): PCG iterations of film f at step s.
 * m_out[s*n + f] (may be NULL): <Pmz> of film f after step s. */
int hzo_batch_run(hzo_batch* batch, const double* Ez, int nsteps, int* iters, double* m_out);
/* Times one step (applied fields Ez[f], best of reps after a warm-up step)
 * for each shape c = 4*k + g: tile depth {2, 4, 8}[k] x group budget
 * {0, 32, 64, 128}[g] MiB, restores the films' states after every trial and
 * keeps the fastest shape.  kz_out and ms[12] (per-shape times) may be NULL. */
int hzo_batch_autotune(hzo_batch* batch, const double* Ez, int reps, int* kz_out, double ms[12]);
/* out[0]: CG launches issued by all batched runs; out[1]: film solves that
 * stopped at the launch cap without converging. */
int hzo_batch_counters(const hzo_batch* batch, long long out[2]);

/* ---------------------------------------------------------------- host side
 * These run on the CPU only (no device needed). */

/* Reference deck "inputs/InputParams" (get_InputParams.cu:8-151). */
typedef struct hzo_deck {
    int device_flag, nx, ny, nz, dep_flag, num_steps, out_flag;
    int dimx, dimy, dimz, do_pre_relax, n_layers;
    int layer_counts[16];
    double del_h, del_t, delEap, Eap, PsNoise, G0, omg;
    double kappa_fe, kappa_afe, kappa_de, L_fe, L_afe, L_de;
    double eps_fe, eps_afe, eps_de;
    char crystal_type[20];
    char z_bc_type[20];
} hzo_deck;
int hzo_read_deck(const char* path, hzo_deck* out);

/* Landau-coefficient tables (SI values).  Replaces LoadLandauPool
 * (main.cu:112-151), which reads 13-token rows (12 coefficients and the
 * reference phase label AFE|FE|DE); this reader also takes the 12-column rows
 * of inputs/FreeEnergyParams.txt (no label: ref_type = -1) and the 16-token
 * rows of src/inputs/SingleL_input.txt (12 coefficients, AFE/FE/DE fractions
 * in percent, label).  Blank and '#' lines are comments; rows of any other
 * token count or with another label are dropped and counted in *nskipped (may
 * be NULL), as the reference drops them.  *nrows = rows read; the first
 * max_rows are stored in rows (rows may be NULL when max_rows is 0). */
typedef struct hzo_landau_row {
    double afe[5];    /* AFE alpha, beta, gamma, g, a0 */
    double fe[5];     /* FE alpha, beta, gamma, g, a0 */
    double de[2];     /* DE alpha, a0 */
    double fracs[3];  /* AFE, FE, DE fractions in percent (16-token rows), else -1 */
    int ref_type;     /* reference phase of the label (HZO_AFE/FE/DE), -1 without one */
    int ncols;        /* 12, 13 or 16 */
} hzo_landau_row;
int hzo_read_landau_table(const char* path, int max_rows, hzo_landau_row* rows, int* nrows, int* nskipped);
/* PickLandauSet (main.cu:155-178): writes alpha, beta, gamma, g, a0 of raw[AFE]
 * and raw[FE] from the row and raw[DE] = (alpha, 0, 0, 0, a0); L, kappa, eps_r
 * and transition_type of raw[] are kept. */
int hzo_landau_to_raw(const hzo_landau_row* row, hzo_phase_raw raw[3]);

/* Grain count file "gnumb_<tag>.txt" (grain_io.cu:10-25). */
int hzo_read_grain_count(const char* path, int* ng);
/* "gdata_<tag>.bin": int32 omega[N] then f64 eta[N], z-fastest; reordered to
 * C order (grain_io.cu:26-69).  eta may be NULL. */
int hzo_read_gdata(const char* path, int nx, int ny, int nz, int32_t* omega, double* eta);
/* "gchar_<tag>.txt": gid cx cy cz phi1 Phi phi2 (radians) per line
 * (grain_io.cu:85-99).  centroids[3*ng], euler[3*ng]; missing ids keep -1/0. */
int hzo_read_gchar(const char* path, int ng, double* centroids, double* euler);
/* Dump formats (generate.cu:11-88), host arrays in voxel order. */
/* write_vector_field_binary: (vx, vy, vz) * scale interleaved, f64. */
int hzo_write_vector_field(const char* path, long long n, const double* vx, const double* vy, const double* vz,
                           double scale);
/* write_field: text rows "i\tj\tk\tv" (v = (int)(map * scale)), i outer, k
 * inner, a blank line after each j and each i. */
int hzo_write_grain_map(const char* path, int nx, int ny, int nz, const int32_t* map, double scale);
/* write_midplane_slice: text rows "i j %.16e" of plane nz/2, j outer, a blank
 * line after each row. */
int hzo_write_midplane_slice(const char* path, int nx, int ny, int nz, const double* v);
/* Checkpoint file (generate.cu:90-264): P[6N] SoA (Pmx..Psz), E[3N] SoA. */
int hzo_write_checkpoint(const char* path, int nx, int ny, int nz, int peloop, int count, double Ez_ext,
                         double incEz, const double* P, const double* E);
/* Reads a checkpoint of an nx x ny x nz grid (HZO_ERR_DATA on a grid
 * mismatch); every output may be NULL. */
int hzo_read_checkpoint(const char* path, int nx, int ny, int nz, int* peloop, int* count, double* Ez_ext,
                        double* incEz, double* P, double* E);
/* Raw f64 dump of n values (the potential, which the reference never writes). */
int hzo_write_raw_f64(const char* path, long long n, const double* v);
/* Rotation matrices from Bunge Euler angles (grain_io.cu:102-126). */
int hzo_euler_to_TR(int ng, const double* euler, double* TR);
/* Dielectric grain-boundary cells (no counterpart in the reference; the
 * large-film workload puts a dielectric Landau term on 1 % of the cells,
 * in grain boundaries).  Boundary cells are those with a face neighbour of
 * another grain (x, y periodic, z inside the film); round(fraction * N) of
 * them (all if fewer), chosen by a seeded SplitMix64 shuffle, move in place
 * to new grains ng, ng+1, ... (one per parent grain, in order of first
 * appearance), parent[new - ng] = parent grain (capacity ng).  The caller
 * gives the new grains the dielectric phase and the parents' orientations;
 * the per-grain model then carries the per-cell phase exactly. */
int hzo_grain_boundary_cells(int nx, int ny, int nz, int32_t* omega, int ng, double fraction, uint64_t seed,
                             int32_t* parent, int* ng_new);
/* Multi-structure layout (main.cu:430-440, README_changes:9-26): n = the
 * number of consecutive directories inputs_dir/struc1, struc2, ...; each
 * holds gchar_poly.txt and gdata_poly.bin, gnumb_poly.txt is in inputs_dir. */
int hzo_count_structures(const char* inputs_dir, int* n);
/* Writers of the generator formats (voronoi.c:162-223). */
int hzo_write_gdata(const char* path, int nx, int ny, int nz, const int32_t* omega_c);
int hzo_write_gchar(const char* path, int ng, const double* centroids, const double* euler);

/* Seeded layered Voronoi structure (voronoi.c:55-153 with the rand() stream
 * replaced by SplitMix64 and orientations uniform on SO(3)).
 * layer_counts[n_layers] must sum to nz.  Outputs: omega[N] (C order),
 * centroids[3*ng], euler[3*ng]. */
int hzo_voronoi(int nx, int ny, int nz, int n_layers, const int* layer_counts,
                int ng, uint64_t seed, int32_t* omega, double* centroids, double* euler);
/* The same structure with the nearest-centroid labelling on GPU `device`
 * (integer distances; labels equal to hzo_voronoi's bit for bit). */
int hzo_voronoi_gpu(int device, int nx, int ny, int nz, int n_layers, const int* layer_counts, int ng,
                    uint64_t seed, int32_t* omega, double* cent, double* euler);
/* hzo_voronoi_gpu for callers that cannot allocate device memory (code under
 * graph capture, serving loops, harnesses that forbid hipMalloc); it has no
 * counterpart in the reference.  The labels go to the caller's device array
 * omega_dev[N] (C order) on GPU `device`; the centroids are staged in a
 * file-scope device array of HZO_VORONOI_NOALLOC_MAX_GRAINS grains that each
 * GPU holds until the process ends, and no device memory is allocated or
 * freed.  ng above that limit is HZO_ERR_ARG.  Runs on the null stream and
 * returns when the labels are written; calls that overlap on one GPU share
 * the staging array and race. */
#define HZO_VORONOI_NOALLOC_MAX_GRAINS 65536
int hzo_voronoi_gpu_noalloc(int device, int nx, int ny, int nz, int n_layers, const int* layer_counts, int ng,
                            uint64_t seed, int32_t* omega_dev, double* cent, double* euler);

/* Non-dimensionalisation (main.cu:200-256, 268-274, 343-349).
 * ref_type selects the reference phase.  out_nd[3] indexed by phase;
 * scales[4] = {P0_ref, E0_ref (V/m), eps0_nd, alpha_ref}. */
int hzo_nondim(const hzo_phase_raw raw[3], int ref_type, hzo_phase_nd out_nd[3], double scales[4]);

/* Layer-wise greedy phase assignment (init_conf.cu:108-285), glibc srand/rand
 * seeded with `seed` (the reference uses 123456).  fracs[3*n_layers] =
 * (afe, fe, de) per layer. */
int hzo_assign_phases(int nx, int ny, int nz, const int32_t* omega, int ng,
                      const double* centz, int n_layers, const int* layer_counts,
                      const double* fracs, unsigned seed, int32_t* grain_phase);
/* Interface flag (main.cu:13-42). */
int hzo_interface_flag(int nx, int ny, int nz, const int32_t* omega, int ng,
                       const int32_t* grain_phase, int32_t* iface);
/* Initial polarisation by phase (init_conf.cu:14-66) with the per-voxel
 * uniform draws u = SplitMix64(seed) stream (the reference's cuRAND stream is
 * not reproducible): FE Pm = amp_fe*(2u-1) on all three components, AFE
 * Ps = (1,1,1), DE and void 0.  P[6N]. */
int hzo_init_polarization(int nx, int ny, int nz, const int32_t* omega, int ng,
                          const int32_t* grain_phase, double amp_fe, uint64_t seed, double* P);
/* n outputs of SplitMix64 started at `seed` (Steele, Lea, Flood 2014). */
int hzo_splitmix64(uint64_t seed, int64_t n, uint64_t* out);

#ifdef __cplusplus
}
#endif
#endif
