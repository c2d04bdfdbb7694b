// hzo_internal.h -- device data layout and launch interface shared by the
// HIP kernels and the C-ABI layer.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hzo.h"

namespace hzo {

// ---------------------------------------------------------------------------
// Per-voxel words.
//
// cell word (uint32, TDGL): grain id in bits 0-27, phase in bits 28-29,
// interface flag in bit 30, void (omega < 0 or >= ng) in bit 31.  Replaces
// the reference's omega_d + is_interface_d + eight per-voxel parameter maps
// (kernels_evol.cu:51-58, init_conf.cu:98-105): the maps hold per-phase
// constants, so the phase index is all the TDGL needs per voxel.
constexpr uint32_t CW_GID_MASK = 0x0FFFFFFFu;
constexpr int CW_PHASE_SHIFT = 28;
constexpr uint32_t CW_IFACE = 1u << 30;
constexpr uint32_t CW_VOID = 1u << 31;

// stencil code (uint16, Poisson operator): permittivity class (0 AFE, 1 FE,
// 2 DE, 3 void with eps = 0) of the voxel and of its e,w,n,s,u,d neighbours
// under the operator's index rules, 2 bits each (bits 0-13); bit 14 marks an
// identity row (Dirichlet electrode plane or Neumann gauge node,
// depol_varEPS.cu:53-71).  The harmonic face means and the Jacobi inverse
// diagonal of every row are functions of this code alone and come from a
// lookup table built on the host with the reference formulas
// (depol_varEPS.cu:10-15, 126-180), so they are bit-identical to
// per-voxel evaluation.
constexpr uint16_t SC_IDENTITY = 1u << 14;
constexpr int SC_LUT_SIZE = (1 << 14) + 1;  // index SC_IDENTITY maps to the last entry
// lut layout (doubles): [0,16) harm table H[a*4+b]; [16, 16+SC_LUT_SIZE) dinv.
constexpr int LUT_HARM = 0;
constexpr int LUT_DINV = 16;
constexpr int LUT_DOUBLES = 16 + SC_LUT_SIZE;

// TDGL coefficients of one (phase, interface-flag) pair, precomputed on the
// host with the reference's own expressions (kernels_evol.cu:140-153, 212):
// identical doubles to the per-voxel evaluation.  Table index phase*2+iface.
struct PhaseC {
    double A;        // kappa*del_t/(del_h*del_h)
    double one_m6A;  // 1.0 - 6.0*A
    double a0_2;     // 2.0*a0
    double beta_4;   // 4.0*beta
    double beta_12;  // 12.0*beta
    double gamma_6;  // 6.0*gamma
    double ag;       // 2.0*(alpha + g_loc)
    double ags;      // 2.0*(alpha - g_loc)
    double cL;       // Lfac*del_t
};
constexpr int PTAB_ENTRIES = 6;

// Device-resident PCG control block.  Scalars never travel to the host
// inside the iteration; the host polls `done` between launch batches.  Two
// copies alternate by launch parity: the kernel launched t-th after the setup
// reads copy (t-1)&1 and its workgroup 0 writes copy t&1 (see reduce.hip.h).
enum : int { K_SETUP = 0, K_DIR = 1, K_FOLD = 2, K_UPDATE = 3, K_RESID = 4, K_WCOMP = 5, K_ITER = 6, K_XADD = 7 };
struct PcgState {
    double rz_old, rz0, rr, rr0, bb, alpha, beta, denom;
    int iters, status, done, refinements;
    int max_iter, rule, kind, pbuf;  // kind: K_* of the kernel that wrote this copy
    double tol_abs2, tol_rel2;
    int refine_every;  // FP32-refined: residual replacement period
    // single-reduction CG: an iteration may leave its x_lo += alpha p to the
    // next one (xpend = 1, alpha_prev = its alpha); xadds counts the launches
    // that only add such a term (before a residual replacement or the end)
    int xpend, xadds;
    int passes;  // resident solve: grid barriers passed
    double alpha_prev;
};

// Per-workgroup partial sums: two parity slots of RED_MAXQ arrays each for
// the PCG, plus the TDGL local-frame sums and the diagnostics sums.
constexpr int RED_MAXQ = 4;
// Workgroup tile: BX x BY columns (one thread each) x kz planes.
constexpr int BX = 64, BY = 8, NT = BX * BY;
constexpr int DIAG_Q = 8;
constexpr int TDGL_Q = 2;

struct Grid {
    int nx, ny, nz;
    long long N;
    int kz;  // z planes per workgroup (tile depth)
};

struct StepConsts {
    int nx, ny, nz;
    int kz;  // z planes per workgroup (tile depth), equal to Grid::kz
    int dep_flag, zbc;
    double del_h, del_t;
    double g_bot, g_top;
    double eps0;
    // divisions by 2h and h^2 in the reference become multiplications by the
    // exact reciprocal only when the divisor is a power of two (then both
    // give the same correctly rounded value).
    double two_h, inv_two_h, h2, inv_h2;
    int two_h_pow2, h2_pow2;
    long long ref_id;  // Neumann gauge node
    unsigned plane;    // nx*ny
    unsigned nvox;     // nx*ny*nz (at most 2^28)
};

// Launch helpers (kernels_fp64.hip).
struct Fp64Bufs {
    const double* Pold;  // TDGL input
    double* Pnew;        // TDGL output
    const double* Pcur;  // current polarisation (PCG right-hand side)
    double* phi;    // PCG x
    double* r;
    double* p[2];   // ping-pong search direction
    const uint32_t* cw;
    const uint16_t* sc;
    const double* TR;
    const PhaseC* ptab;
    const double* lut;
    double* part[2];    // PCG partials by launch parity, RED_MAXQ * nb each
    const double* pin[2];  // partials a launch reads, by parity: part[] itself, or their folded totals
    int nb_in, ps_in;      // count and stride of pin (1 and 512 when folded)
    PcgState* st;       // [2] control blocks by launch parity
    double* tdgl_part;  // TDGL_Q * nb: grain-frame z sums of the last TDGL call
    int stream_loads;   // deep-tile PCG kernels load their streamed vectors nontemporal
};

// FP32-refined form: FP32 polarisation and CG vectors, FP64 potential x_hi
// with an FP32 correction x_lo (phi = x_hi + x_lo), FP64 right-hand side b.
struct PhaseC32 {
    float A, one_m6A, a0_2, beta_4, beta_12, gamma_6, ag, ags, cL;
};
struct Fp32Bufs {
    const float* Pold;
    float* Pnew;
    const float* Pcur;
    double* xhi;
    float* xlo;
    float* r;
    float* p[2];
    double* b;
    const uint32_t* cw;
    const uint16_t* sc;
    const float* TR32;
    const PhaseC32* ptab32;
    const double* lut;    // FP64 harm + dinv (residual replacement)
    const float* lut32;   // FP32 harm + dinv (CG iteration)
    double* part[2];
    const double* pin[2];  // partials a launch reads, by parity: part[] itself, or their folded totals
    int nb_in, ps_in;      // count and stride of pin (1 and 512 when folded)
    PcgState* st;
    double* tdgl_part;
};

// Single-reduction CG of the FP32-refined form (kernels_cg32.hip).  The form
// carrying w = A M r and s = A p double-buffers r, w, s by the control
// block's parity; the two-stencil form (k_cg32r) carries only r and p, both
// double-buffered (p in the s slots), and leaves the w slots and the p slot
// unused.
struct CG32Bufs {
    // FP32 vectors in arenas of arrays of N (one buffer resource each):
    // va = r[0] r[1] w[0] w[1], vb = s[0] s[1] p (a film keeps its x_lo as
    // the fourth array of vb; a solve may bring another)
    float* va;
    float* vb;
    float* xlo;
    double* xhi;
    const double* b;
    const uint16_t* sc;
    const double* lut;
    const float* lut32;
    double* part[2];
    const double* pin[2];  // partials a launch reads, by parity: part[] itself, or their folded totals
    int nb_in, ps_in;      // count and stride of pin (1 and 512 when folded)
    PcgState* st;
};

// Register-resident single-reduction CG (kernels_cgres.hip): the whole
// FP32-refined solve in one cooperative launch over 16 x 16 column tiles.
struct ResBufs {
    const float* r0;       // r = b - A x after the setup kernel
    float* xlo;            // x_lo: read at the start, written at the end
    double* xhi;           // x_hi: x_lo is folded into it at every residual replacement
    const double* b;       // right-hand side
    const uint16_t* sc;    // stencil codes
    const double* lut;     // FP64 face means (residual replacement)
    const float* lut32;    // FP32 face means and D^-1
    const double* pin;     // the setup's partial sums (r.z, r.r, b.b): nb_in each, stride ps_in
    int nb_in, ps_in;
    PcgState* st;          // in: the setup's control block; out: the final one
    float* xch;            // edge columns: [2][tiles][4][nz][16]
    double* xchd;          // FP64 edge columns of x (residual replacement): [tiles][4][nz][16]
    double* part;          // workgroup partial sums: [2][tiles][4]
    unsigned* sync;        // barrier counters (global, then one per XCD, 64 apart), zero at launch
    int* err;              // barrier timeout flag, zero at launch
};
constexpr int RES_SYNC_WORDS = 64 * 9;
hipError_t launch_cgres(hipStream_t s, const StepConsts& c, const ResBufs& b, int gxt, int G);
// One film of a multi-film resident launch: its buffers and constants, tiles
// in x, tiles, and the first workgroup of its block range.
struct ResFilm {
    ResBufs b;
    StepConsts c;
    int gxt, G, blk0, pad;
};
hipError_t launch_cgres_films(hipStream_t s, const ResFilm* films, int nf, int tiles, int nz, bool mir);
int cgres_capacity(int device, int nz, bool mir);
void launch_fill_u32(hipStream_t s, unsigned* p, int n, unsigned v);

// Solver summary of one batch film after a poll (k_batch_done): done,
// iterations, status, residual replacements, deferred-x_lo launches.
constexpr int BATCH_INFO = 5;

// One film of a batch (device table, hzo_batch_*): the buffers of its
// context by P parity, its single-reduction CG vectors, step constants, tile
// grid and partial layout (batch-owned partial arrays, common stride).
struct BatchFilm {
    Fp32Bufs b[2];
    CG32Bufs cg;
    StepConsts c;
    Grid g;
    PcgState init;
    int gx, gy, gz;  // tile grid (x, y, z)
    int nb, ps;      // tiles and partial stride
    float Ez;        // applied field of the current step
    double* diag_part;
};

int blocks_for(const Grid& g);
// Stride of one PCG partial-sum array (>= blocks_for; the tail is kept zero).
int part_stride(const Grid& g);

// Optional start/stop events attached to a kernel's own dispatch
// (hipExtLaunchKernelGGL): their elapsed time is the kernel's execution
// time, without the gaps separate event records would add.
struct LaunchEv {
    hipEvent_t start = nullptr, stop = nullptr;
};

// Launchers.  `t` is the launch index since the last PCG setup (setup = 0).
void launch_tdgl_f64(hipStream_t s, const Grid& g, const StepConsts& c, double Ez_ext, const Fp64Bufs& b,
                 const LaunchEv& ev = {});
void launch_pcg_setup_f64(hipStream_t s, const Grid& g, const StepConsts& c, const Fp64Bufs& b,
                          const PcgState& init, const double* rhs,
                 const LaunchEv& ev = {});
void launch_pcg_dir_f64(hipStream_t s, const Grid& g, const StepConsts& c, const Fp64Bufs& b, int t,
                 const LaunchEv& ev = {});
void launch_pcg_update_f64(hipStream_t s, const Grid& g, const StepConsts& c, const Fp64Bufs& b, int t,
                 const LaunchEv& ev = {});
void launch_diag_f64(hipStream_t s, const Grid& g, const double* P, const uint32_t* cw, double* part,
                 const LaunchEv& ev = {});
void launch_tdgl_f32(hipStream_t s, const Grid& g, const StepConsts& c, double Ez_ext, const Fp32Bufs& b,
                 const LaunchEv& ev = {});
// The FP32 TDGL with the applied field at ez (device memory), for captured steps.
void launch_tdgl_f32_dev(hipStream_t s, const Grid& g, const StepConsts& c, const double* ez, const Fp32Bufs& b);
void launch_pcg_setup_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b,
                          const PcgState& init, const double* rhs,
                 const LaunchEv& ev = {});
void launch_pcg_dir_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b, int t,
                 const LaunchEv& ev = {});
void launch_pcg_update_f32(hipStream_t s, const Grid& g, const StepConsts& c, const Fp32Bufs& b, int t,
                 const LaunchEv& ev = {});
void launch_diag_f32(hipStream_t s, const Grid& g, const float* P, const uint32_t* cw, double* part,
                 const LaunchEv& ev = {});
// two_stencil: the form that recomputes s = A p and w = A u in the launch
// (reads r, p; writes r, p, x_lo) instead of carrying w and s.
// stream_stores: that form at depth 16 or 32 stores its vectors nontemporal
// (films whose CG vectors exceed the last-level cache).
void launch_cg32(hipStream_t s, const Grid& g, const StepConsts& c, const CG32Bufs& b, int t, bool two_stencil,
                 bool stream_stores, const LaunchEv& ev = {});
// Batched FP32-refined launches over every tile of every film (1-D grid of
// ntiles workgroups; tiles[] maps each to (film, logical tile)).
void launch_tdgl_f32_batch(hipStream_t s, int ntiles, const BatchFilm* films, const int2* tiles, int parity);
void launch_setup_f32_batch(hipStream_t s, int ntiles, const BatchFilm* films, const int2* tiles, int parity);
void launch_diag_f32_batch(hipStream_t s, int ntiles, const BatchFilm* films, const int2* tiles, int parity);
// mirrored: some film of the launch has Neumann electrodes.
void launch_cg32_batch(hipStream_t s, int ntiles, int kz, int pm, bool mirrored, const BatchFilm* films,
                       const int2* tiles, int t);
void launch_batch_set_ez(hipStream_t s, BatchFilm* films, const double* ez, int n);
// Sum of each of the RED_MAXQ partial arrays (nb entries, stride ps) into
// out[q*NT] (the rest of out stays zero), one workgroup: for films whose
// partial sums outnumber what a consuming thread loads directly.
void launch_fold_partials(hipStream_t s, const double* in, int nb, int ps, double* out);
// info: done, iters, status, refinements per film
void launch_batch_done(hipStream_t s, const BatchFilm* films, int n, int t, int* info);
void launch_batch_sum_pmz(hipStream_t s, const BatchFilm* films, int n, double* m);

// Adjoint kernels (kernels_adjoint.hip).  Pointers are float or double per
// fp32; ptab is PhaseC32 or PhaseC accordingly.
struct AdjTDGLBufsAny {
    bool fp32;
    const void* P;   // forward input state of the step (6N)
    const void* a;   // adjoint of the step's output state (6N)
    void* w;         // out: adjoint through the TDGL map (6N)
    double* v;       // out: adjoint Poisson right-hand side (N)
    const uint32_t* cw;
    const uint16_t* sc;
    const void* TR;
    const void* ptab;
    double* gpart;   // 12 x nb parameter-gradient partials
    double dAdk;     // dA/dkappa = del_t / del_h^2
};
void launch_adj_tdgl(hipStream_t s, const Grid& g, const StepConsts& c, const AdjTDGLBufsAny& b);
void launch_adj_gradacc(hipStream_t s, const Grid& g, const double* gpart, double* grad);
void launch_adj_divt(hipStream_t s, const Grid& g, const StepConsts& c, bool fp32, const void* w,
                     const double* mu_hi, const float* mu_lo, void* a, double seed_over_N, bool with_mu);
void launch_adj_seed(hipStream_t s, const Grid& g, bool fp32, void* a, double seed_over_N);
void launch_sum_column(hipStream_t s, const Grid& g, const double* part, double* out);

}  // namespace hzo
