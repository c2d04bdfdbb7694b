// hzo_api.hip -- C-ABI context layer: owns the device state of one film and
// drives the step (replaces AllocateData/Cleanup, the device globals, the
// step loop body of evolve.cu:220-237 and the host side of Depol_field).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "host/errors.h"
#include "hzo_internal.h"

using hzo::set_error;

#define HZO_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_error(HZO_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call,          \
                             hipGetErrorString(e_));                                           \
    } while (0)

// Steps hzo_run enqueues on a resident film before it waits for them.
constexpr int RUN_RING = 64;

struct hzo_ctx {
    hzo_config cfg;
    hzo::Grid g;
    hzo::StepConsts c;
    int device = 0;
    hipStream_t stream = nullptr;
    int nb = 0;

    bool fp32 = false;  // HZO_PREC_FP32_REFINED
    bool cg_two = false;  // single-reduction CG forming s = A p and w = A u in the launch
    bool cg_nt = false;   // its iteration stores nontemporal
    bool pcg_ntl = false; // FP64 deep-tile PCG: streamed loads nontemporal
    // register-resident single-reduction CG: the whole solve in one
    // cooperative launch over 16 x 16 column tiles (films whose tiles all fit
    // the device at once)
    bool cg_res = false;
    int res_gxt = 0, res_G = 0;  // tiles in x, tiles
    float* res_xch = nullptr;    // edge columns [2][G][4][nz][16]
    double* res_xchd = nullptr;  // FP64 edge columns [G][4][nz][16]
    double* res_part = nullptr;  // partial sums [2][G][4]
    unsigned* res_sync = nullptr;  // barrier counters, then the timeout flag
    int* res_err_host = nullptr;   // pinned copy of the timeout flag
    // hzo_run on a resident film: each step's control block and timeout flag
    // land in a pinned ring (RUN_RING entries) that the host reads once per
    // RUN_RING steps instead of waiting for every solve
    hzo::PcgState* res_ring = nullptr;
    int* res_err_ring = nullptr;
    long long res_counts[3] = {0, 0, 0};  // solves, iterations, grid barriers (since hzo_set_timing)
    double* P[2] = {nullptr, nullptr};
    float* P32[2] = {nullptr, nullptr};
    float* xlo = nullptr;
    float* r32 = nullptr;
    float* p32[2] = {nullptr, nullptr};
    double* b64 = nullptr;
    float* cg_r1 = nullptr;  // single-reduction CG: second r buffer, w and s pairs
    float* cg_w[2] = {nullptr, nullptr};
    float* cg_s[2] = {nullptr, nullptr};
    // single-reduction CG arenas: r32 cg_r1 cg_w[0] cg_w[1] | cg_s[0] cg_s[1] p32[0] xlo
    float* cg_arena[2] = {nullptr, nullptr};
    float* TR32 = nullptr;
    hzo::PhaseC32* ptab32 = nullptr;
    float* lut32 = nullptr;
    int cur = 0;
    double* phi = nullptr;
    double* r = nullptr;
    double* p[2] = {nullptr, nullptr};
    uint32_t* cw = nullptr;
    uint16_t* sc = nullptr;
    double* TR = nullptr;
    hzo::PhaseC* ptab = nullptr;
    double* lut = nullptr;
    double* part = nullptr;       // 2 parity slots x RED_MAXQ x ps (zero tail beyond nb)
    int ps = 0;                   // stride of one partial array
    double* fold = nullptr;       // more than 4 x 512 partials: 2 x RED_MAXQ x 512 folded totals (zero tail)
    double* tdgl_part = nullptr;  // TDGL_Q x nb
    double* diag_part = nullptr;  // DIAG_Q x nb
    hzo::PcgState* st = nullptr;  // [2]
    hzo::PcgState* st_host = nullptr;  // pinned
    std::vector<double> diag_host;

    int ng = 0;
    std::vector<int32_t> omega, grain_phase, iface;
    hzo_phase_nd ph[3];
    bool have_micro = false, have_phase = false, have_state = false;
    int last_iters = 8;

    bool timing = false;
    std::vector<hipEvent_t> ev[HZO_KCLASSES][2];
    int nt[HZO_KCLASSES] = {};

    // reverse mode: checkpoints, recompute segment and adjoint vectors
    // (hzo_adjoint_reserve / hzo_forward_record / hzo_backward)
    struct Adjoint {
        int nsteps = 0;    // steps reserved
        int K = 0;         // checkpoint period
        int nck = 0;       // checkpoint slots (the last one holds the final state)
        int recorded = 0;  // steps of the last forward_record
        int last_iters = 8;
        double tol = 1e-8;
        std::vector<double> Ez;
        void* ckP = nullptr;      // nck x 6N polarisation (float or double)
        double* ckphi = nullptr;  // nck x N potential (x_hi)
        float* cklo = nullptr;    // nck x N potential correction x_lo (FP32-refined)
        void* seg = nullptr;      // K x 6N forward inputs of one segment
        void* a = nullptr;        // 6N adjoint of the state
        void* w = nullptr;        // 6N adjoint through the TDGL map
        double* mu = nullptr;     // N adjoint potential (x_hi)
        float* mulo = nullptr;    // N (FP32-refined x_lo)
        double* rhs = nullptr;    // N adjoint Poisson right-hand side (FP64; FP32 uses b64)
        double* gpart = nullptr;  // 12 x nb
        double* grad = nullptr;   // 12
        double* m = nullptr;      // nsteps loop samples (sum of Pmz)
    } adj;
};

namespace {

bool is_pow2(double v) {
    if (!(v > 0.0) || !std::isfinite(v)) return false;
    int e;
    const double m = frexp(v, &e);
    return m == 0.5;
}

// depol_varEPS.cu:10-15
double harm(double a, double b) {
    const double eps = 1e-20;
    double denom = a + b;
    if (denom < eps) denom = eps;
    return (2.0 * a * b) / denom;
}

// The partials the next launch reads: the arrays themselves, or (films with
// more than 4 x 512 workgroups) their totals folded into 512-entry arrays.
template <class B>
void set_pin(hzo_ctx* x, B& b) {
    for (int q = 0; q < 2; ++q)
        b.pin[q] = x->fold ? x->fold + (size_t)q * hzo::RED_MAXQ * hzo::NT : b.part[q];
    b.nb_in = x->fold ? 1 : x->nb;
    b.ps_in = x->fold ? hzo::NT : x->ps;
}

hzo::Fp64Bufs bufs(hzo_ctx* x) {
    hzo::Fp64Bufs b;
    b.Pold = x->P[x->cur];
    b.Pnew = x->P[x->cur ^ 1];
    b.Pcur = x->P[x->cur];
    b.phi = x->phi;
    b.r = x->r;
    b.p[0] = x->p[0];
    b.p[1] = x->p[1];
    b.cw = x->cw;
    b.sc = x->sc;
    b.TR = x->TR;
    b.ptab = x->ptab;
    b.lut = x->lut;
    b.part[0] = x->part;
    b.part[1] = x->part + (size_t)hzo::RED_MAXQ * x->ps;
    set_pin(x, b);
    b.st = x->st;
    b.tdgl_part = x->tdgl_part;
    b.stream_loads = x->pcg_ntl;
    return b;
}

hzo::Fp32Bufs bufs32(hzo_ctx* x) {
    hzo::Fp32Bufs b;
    b.Pold = x->P32[x->cur];
    b.Pnew = x->P32[x->cur ^ 1];
    b.Pcur = x->P32[x->cur];
    b.xhi = x->phi;
    b.xlo = x->xlo;
    b.r = x->r32;
    b.p[0] = x->p32[0];
    b.p[1] = x->p32[1];
    b.b = x->b64;
    b.cw = x->cw;
    b.sc = x->sc;
    b.TR32 = x->TR32;
    b.ptab32 = x->ptab32;
    b.lut = x->lut;
    b.lut32 = x->lut32;
    b.part[0] = x->part;
    b.part[1] = x->part + (size_t)hzo::RED_MAXQ * x->ps;
    set_pin(x, b);
    b.st = x->st;
    b.tdgl_part = x->tdgl_part;
    return b;
}

// After a launch that wrote the partials of parity t & 1 (the setup is t =
// 0): fold them when the film has more than a consuming thread loads.
void fold_after(hzo_ctx* x, hipStream_t s, int t) {
    if (!x->fold) return;
    const size_t q = (size_t)(t & 1);
    hzo::launch_fold_partials(s, x->part + q * hzo::RED_MAXQ * x->ps, x->nb, x->ps,
                              x->fold + q * hzo::RED_MAXQ * hzo::NT);
}

template <class F>
int timed(hzo_ctx* x, int cls, F launch) {
    const bool t = x->timing && x->nt[cls] < HZO_TIMING_MAX;
    if (t) {
        auto& s = x->ev[cls][0];
        auto& e = x->ev[cls][1];
        const int i = x->nt[cls];
        if ((int)s.size() <= i) {
            hipEvent_t a, b;
            HZO_HIP(hipEventCreate(&a));
            HZO_HIP(hipEventCreate(&b));
            s.push_back(a);
            e.push_back(b);
        }
        hzo::LaunchEv ev;
        ev.start = s[i];
        ev.stop = e[i];
        launch(ev);
        HZO_HIP(hipGetLastError());
        x->nt[cls]++;
    } else {
        launch(hzo::LaunchEv{});
        HZO_HIP(hipGetLastError());
    }
    return HZO_OK;
}

int check_ready(hzo_ctx* x) {
    if (!x) return set_error(HZO_ERR_ARG, "null context");
    if (!x->have_micro || !x->have_phase || !x->have_state)
        return set_error(HZO_ERR_STATE, "context needs set_microstructure, set_phase_table and set_state first");
    return HZO_OK;
}

// Stencil code per voxel (see hzo_internal.h): permittivity class of the
// voxel and of its operator neighbours, identity rows flagged.
void build_codes(hzo_ctx* x, std::vector<uint16_t>& sc) {
    const int nx = x->g.nx, ny = x->g.ny, nz = x->g.nz;
    const long long N = x->g.N;
    sc.assign(N, 0);
    auto cls = [&](long long id) -> unsigned {
        const int gid = x->omega[id];
        if (gid < 0 || gid >= x->ng) return 3u;
        const int m = x->grain_phase[gid];
        return m == 2 ? 2u : (m == 1 ? 1u : 0u);
    };
    const long long ref = x->c.ref_id;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const long long id = (long long)i + (long long)nx * ((long long)j + (long long)ny * k);
                if ((x->c.zbc == 0 && (k == 0 || k == nz - 1)) || (x->c.zbc == 1 && id == ref)) {
                    sc[id] = hzo::SC_IDENTITY;
                    continue;
                }
                const int e = (i + 1) % nx, w = (i - 1 + nx) % nx;
                const int n = (j + 1) % ny, s = (j - 1 + ny) % ny;
                int up, dn;
                if (x->c.zbc == 0) {
                    up = k + 1;
                    dn = k - 1;
                } else if (k == 0) {
                    up = dn = 1;
                } else if (k == nz - 1) {
                    up = dn = nz - 2;
                } else {
                    up = k + 1;
                    dn = k - 1;
                }
                auto at = [&](int a, int b2, int c2) {
                    return (long long)a + (long long)nx * ((long long)b2 + (long long)ny * c2);
                };
                const unsigned code = cls(id) | (cls(at(e, j, k)) << 2) | (cls(at(w, j, k)) << 4) |
                                      (cls(at(i, n, k)) << 6) | (cls(at(i, s, k)) << 8) |
                                      (cls(at(i, j, up)) << 10) | (cls(at(i, j, dn)) << 12);
                sc[id] = (uint16_t)code;
            }
}

int upload_micro(hzo_ctx* x) {
    const long long N = x->g.N;
    std::vector<uint32_t> cw(N);
    for (long long id = 0; id < N; ++id) {
        const int gid = x->omega[id];
        if (gid < 0 || gid >= x->ng) {
            cw[id] = hzo::CW_VOID;
            continue;
        }
        const int m = x->grain_phase[gid];
        const uint32_t ph = m == 2 ? 2u : (m == 1 ? 1u : 0u);
        cw[id] = ((uint32_t)gid & hzo::CW_GID_MASK) | (ph << hzo::CW_PHASE_SHIFT) |
                 (x->iface[id] ? hzo::CW_IFACE : 0u);
    }
    std::vector<uint16_t> sc;
    build_codes(x, sc);
    HZO_HIP(hipMemcpy(x->cw, cw.data(), sizeof(uint32_t) * N, hipMemcpyHostToDevice));
    HZO_HIP(hipMemcpy(x->sc, sc.data(), sizeof(uint16_t) * N, hipMemcpyHostToDevice));
    return HZO_OK;
}

int upload_phase(hzo_ctx* x) {
    const double del_h = x->cfg.del_h, del_t = x->cfg.del_t;
    hzo::PhaseC pt[hzo::PTAB_ENTRIES];
    for (int q = 0; q < 3; ++q)
        for (int f = 0; f < 2; ++f) {
            const hzo_phase_nd& P = x->ph[q];
            double g_loc = P.g, Lfac = P.L;
            if (f) {  // kernels_evol.cu:142-146
                g_loc *= 0.4;
                Lfac *= (1.0 + 10.0);
            }
            hzo::PhaseC& c = pt[q * 2 + f];
            c.A = P.kappa * del_t / (del_h * del_h);
            c.one_m6A = 1.0 - 6.0 * c.A;
            c.a0_2 = 2.0 * P.a0;
            c.beta_4 = 4.0 * P.beta;
            c.beta_12 = 12.0 * P.beta;
            c.gamma_6 = 6.0 * P.gamma;
            c.ag = 2.0 * (P.alpha + g_loc);
            c.ags = 2.0 * (P.alpha - g_loc);
            c.cL = Lfac * del_t;
        }
    std::vector<double> lut(hzo::LUT_DOUBLES, 0.0);
    double eps[4] = {x->ph[0].eps_r, x->ph[1].eps_r, x->ph[2].eps_r, 0.0};
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) lut[hzo::LUT_HARM + a * 4 + b] = harm(eps[a], eps[b]);
    const double h = del_h;
    for (unsigned code = 0; code < (1u << 14); ++code) {
        const double* Hr = &lut[hzo::LUT_HARM + (code & 3u) * 4];
        const double ex_p = Hr[(code >> 2) & 3], ex_m = Hr[(code >> 4) & 3];
        const double ey_p = Hr[(code >> 6) & 3], ey_m = Hr[(code >> 8) & 3];
        const double ez_p = Hr[(code >> 10) & 3], ez_m = Hr[(code >> 12) & 3];
        const double D = (ex_p + ex_m + ey_p + ey_m + ez_p + ez_m) / (h * h);  // depol_varEPS.cu:177-179
        lut[hzo::LUT_DINV + code] = (D > 1e-30 ? 1.0 / D : 1.0);
    }
    lut[hzo::LUT_DINV + hzo::SC_IDENTITY] = 1.0;

    HZO_HIP(hipMemcpy(x->ptab, pt, sizeof(pt), hipMemcpyHostToDevice));
    HZO_HIP(hipMemcpy(x->lut, lut.data(), sizeof(double) * hzo::LUT_DOUBLES, hipMemcpyHostToDevice));
    if (x->fp32) {
        hzo::PhaseC32 p32[hzo::PTAB_ENTRIES];
        for (int q = 0; q < hzo::PTAB_ENTRIES; ++q) {
            p32[q].A = (float)pt[q].A;
            p32[q].one_m6A = (float)pt[q].one_m6A;
            p32[q].a0_2 = (float)pt[q].a0_2;
            p32[q].beta_4 = (float)pt[q].beta_4;
            p32[q].beta_12 = (float)pt[q].beta_12;
            p32[q].gamma_6 = (float)pt[q].gamma_6;
            p32[q].ag = (float)pt[q].ag;
            p32[q].ags = (float)pt[q].ags;
            p32[q].cL = (float)pt[q].cL;
        }
        std::vector<float> l32(lut.begin(), lut.end());
        HZO_HIP(hipMemcpy(x->ptab32, p32, sizeof(p32), hipMemcpyHostToDevice));
        HZO_HIP(hipMemcpy(x->lut32, l32.data(), sizeof(float) * hzo::LUT_DOUBLES, hipMemcpyHostToDevice));
    }
    return HZO_OK;
}

hzo::PcgState pcg_init(const hzo_ctx* x) {
    hzo::PcgState s;
    memset(&s, 0, sizeof(s));
    s.max_iter = x->cfg.max_iter;
    s.rule = x->cfg.pcg_rule;
    s.tol_abs2 = x->cfg.tol_abs * x->cfg.tol_abs;
    s.tol_rel2 = x->cfg.tol_rel * x->cfg.tol_rel;
    s.refine_every = x->cfg.refine_every;
    s.kind = hzo::K_SETUP;
    return s;
}

int tdgl(hzo_ctx* x, double Ez_ext) {
    int rc;
    if (x->fp32) {
        const hzo::Fp32Bufs b = bufs32(x);
        rc = timed(x, 0, [&](const hzo::LaunchEv& ev) { hzo::launch_tdgl_f32(x->stream, x->g, x->c, Ez_ext, b, ev); });
    } else {
        const hzo::Fp64Bufs b = bufs(x);
        rc = timed(x, 0, [&](const hzo::LaunchEv& ev) { hzo::launch_tdgl_f64(x->stream, x->g, x->c, Ez_ext, b, ev); });
    }
    if (rc) return rc;
    x->cur ^= 1;  // kernelUpdate_P becomes a buffer swap
    return HZO_OK;
}

// One PCG solve on the given vectors: the forward depolarisation solve
// (rhs == nullptr: b = -eps0 div P) or an adjoint solve of the same operator
// (rhs = the right-hand side, 0 on identity rows).  last_iters sizes the
// launch batches from the previous solve of the same kind.
int copy_stats(const hzo::PcgState& s, hzo_pcg_stats* stats) {
    if (stats) {
        stats->iters = s.iters;
        stats->status = s.status;
        stats->refinements = s.refinements;
        stats->reserved = 0;
        stats->rz0 = s.rz0;
        stats->rz = s.rz_old;
        stats->rr0 = s.rr0;
        stats->rr = s.rr;
        stats->bb = s.bb;
    }
    return HZO_OK;
}

// Vectors of the single-reduction CG of a film (the solution, right-hand side,
// partial sums and control block come from b32).
hzo::CG32Bufs cg_bufs(hzo_ctx* x, const hzo::Fp32Bufs& b32) {
    hzo::CG32Bufs cb;
    cb.va = x->cg_arena[0];
    cb.vb = x->cg_arena[1];
    cb.xlo = b32.xlo;
    cb.xhi = b32.xhi;
    cb.b = b32.b;
    cb.sc = x->sc;
    cb.lut = x->lut;
    cb.lut32 = x->lut32;
    cb.part[0] = b32.part[0];
    cb.part[1] = b32.part[1];
    cb.pin[0] = b32.pin[0];
    cb.pin[1] = b32.pin[1];
    cb.nb_in = b32.nb_in;
    cb.ps_in = b32.ps_in;
    cb.st = b32.st;
    return cb;
}

// Size of the next follow-up round of launches after a solve outlasts the
// round sized from the previous solve.  Iteration counts drift slowly from
// step to step, so such a solve usually needs only a few more launches:
// the rounds start at 4 and double up to an eighth of the previous solve
// (at least 8), which bounds both the surplus launches (each returns at
// once but costs a dispatch) and the number of host polls.
int follow_up_round(int& next, int last_iters) {
    const int cap = last_iters / 8 > 8 ? last_iters / 8 : 8;
    const int r = next < cap ? next : cap;
    next = 2 * r;
    return r;
}

// The resident solve after the setup: one cooperative launch runs every
// iteration and residual replacement; the host waits for it once.
// The resident solve's buffers for film x after its setup kernel (cb: the
// film's CG buffers), with the film's own barrier words and timeout flag.
hzo::ResBufs res_bufs(hzo_ctx* x, const hzo::CG32Bufs& cb) {
    hzo::ResBufs rb;
    rb.r0 = x->r32;
    rb.xlo = cb.xlo;
    rb.xhi = cb.xhi;
    rb.xchd = x->res_xchd;
    rb.b = cb.b;
    rb.sc = cb.sc;
    rb.lut = cb.lut;
    rb.lut32 = cb.lut32;
    rb.pin = cb.pin[0];
    rb.nb_in = cb.nb_in;
    rb.ps_in = cb.ps_in;
    rb.st = cb.st;
    rb.xch = x->res_xch;
    rb.part = x->res_part;
    rb.sync = x->res_sync;
    rb.err = reinterpret_cast<int*>(x->res_sync + hzo::RES_SYNC_WORDS);
    return rb;
}

// Control-block bookkeeping after a resident solve of film x (its state in
// x->st_host): a solve the kernel ended at its pass cap reports max_iter.
void res_after(hzo_ctx* x) {
    hzo::PcgState& s = *x->st_host;
    if (!s.done) s.status = HZO_PCG_MAXITER;
    x->res_counts[0] += 1;
    x->res_counts[1] += s.iters;
    x->res_counts[2] += s.passes;
}

// Enqueues the resident solve of film x on its stream; its control block
// and timeout flag are copied to st_out and err_out (pinned) when it ends.
int cgres_enqueue(hzo_ctx* x, const hzo::StepConsts& c, const hzo::CG32Bufs& cb, hzo::PcgState* st_out,
                  int* err_out) {
    const hzo::ResBufs rb = res_bufs(x, cb);
    // counters and flag start at zero (a fill kernel, not a memset node, in
    // case the caller captures the stream)
    hzo::launch_fill_u32(x->stream, x->res_sync, hzo::RES_SYNC_WORDS + 64, 0u);
    HZO_HIP(hipGetLastError());
    hipError_t le = hipSuccess;
    int rc = timed(x, 5, [&](const hzo::LaunchEv& ev) {
        if (ev.start) (void)hipEventRecord(ev.start, x->stream);
        le = hzo::launch_cgres(x->stream, c, rb, x->res_gxt, x->res_G);
        if (ev.stop) (void)hipEventRecord(ev.stop, x->stream);
    });
    if (rc) return rc;
    if (le != hipSuccess)
        return set_error(HZO_ERR_HIP, "resident CG launch failed: %s", hipGetErrorString(le));
    HZO_HIP(hipMemcpyAsync(st_out, cb.st, sizeof(hzo::PcgState), hipMemcpyDeviceToHost, x->stream));
    HZO_HIP(hipMemcpyAsync(err_out, rb.err, sizeof(int), hipMemcpyDeviceToHost, x->stream));
    return HZO_OK;
}

// The host's side of a finished resident solve whose control block and
// timeout flag are st and err.
int cgres_collect(hzo_ctx* x, const hzo::PcgState& st, int err, int& last_iters, hzo_pcg_stats* stats) {
    if (err) return set_error(HZO_ERR_HIP, "resident CG: a grid barrier timed out");
    if (&st != x->st_host) *x->st_host = st;
    res_after(x);
    const hzo::PcgState& s = *x->st_host;
    last_iters = s.iters + 2 * s.refinements + 1;
    return copy_stats(s, stats);
}

int cgres_solve(hzo_ctx* x, const hzo::StepConsts& c, const hzo::CG32Bufs& cb, int& last_iters,
                hzo_pcg_stats* stats) {
    int rc = cgres_enqueue(x, c, cb, x->st_host, x->res_err_host);
    if (rc) return rc;
    HZO_HIP(hipStreamSynchronize(x->stream));
    return cgres_collect(x, *x->st_host, *x->res_err_host, last_iters, stats);
}

// FP32-refined solve by the single-reduction CG: the setup (r = b - A x,
// partials r.z, r.r, b.b) and then one fused launch per iteration, the
// device deciding from the control block whether a launch iterates,
// recomputes the true residual in FP64, or re-forms w after it.
int cg32_solve(hzo_ctx* x, const hzo::Fp32Bufs& b32in, const hzo::StepConsts& c, const hzo::PcgState& init,
               const double* rhs, int& last_iters, hzo_pcg_stats* stats, int ring_slot = -1) {
    hzo::Fp32Bufs b32 = b32in;
    b32.r = x->r32;
    const hzo::CG32Bufs cb = cg_bufs(x, b32);
    int rc = timed(x, 1, [&](const hzo::LaunchEv& ev) { hzo::launch_pcg_setup_f32(x->stream, x->g, c, b32, init, rhs, ev); });
    if (rc) return rc;
    fold_after(x, x->stream, 0);
    // ring_slot >= 0: a resident solve is only enqueued (hzo_run collects it)
    if (x->cg_res && ring_slot >= 0)
        return cgres_enqueue(x, c, cb, x->res_ring + ring_slot, x->res_err_ring + ring_slot);
    if (x->cg_res) return cgres_solve(x, c, cb, last_iters, stats);
    // launches: one WCOMP, one per iteration, two per residual replacement
    // and one before it or the end when an iteration's x_lo term is still
    // deferred; at most four per iteration (add, replacement, re-form,
    // iterate)
    const long long maxl = 4LL * init.max_iter + 8;
    long long launches = 0;
    int t = 0;
    int batch = last_iters + 2, next = 4;
    while (true) {
        for (int m = 0; m < batch && launches < maxl; ++m, ++launches) {
            const int tt = ++t;
            rc = timed(x, 4, [&](const hzo::LaunchEv& ev) { hzo::launch_cg32(x->stream, x->g, c, cb, tt, x->cg_two, x->cg_nt, ev); });
            if (rc) return rc;
            fold_after(x, x->stream, tt);
        }
        HZO_HIP(hipMemcpyAsync(x->st_host, x->st + (t & 1), sizeof(hzo::PcgState), hipMemcpyDeviceToHost,
                               x->stream));
        HZO_HIP(hipStreamSynchronize(x->stream));
        if (x->st_host->done || launches >= maxl) break;
        batch = follow_up_round(next, last_iters);
    }
    hzo::PcgState& s = *x->st_host;
    if (!s.done) s.status = HZO_PCG_MAXITER;
    last_iters = s.iters + 2 * s.refinements + s.xadds + 1;
    return copy_stats(s, stats);
}

int pcg_solve(hzo_ctx* x, const hzo::Fp64Bufs& b, const hzo::Fp32Bufs& b32, const hzo::StepConsts& c,
              const hzo::PcgState& init, const double* rhs, int& last_iters, hzo_pcg_stats* stats) {
    if (x->fp32 && x->cfg.cg_fused) return cg32_solve(x, b32, c, init, rhs, last_iters, stats);
    int rc = x->fp32 ? timed(x, 1, [&](const hzo::LaunchEv& ev) { hzo::launch_pcg_setup_f32(x->stream, x->g, c, b32, init, rhs, ev); })
                     : timed(x, 1, [&](const hzo::LaunchEv& ev) { hzo::launch_pcg_setup_f64(x->stream, x->g, c, b, init, rhs, ev); });
    if (rc) return rc;
    fold_after(x, x->stream, 0);
    // One (dir, update) launch pair per CG iteration; the pair that follows
    // the last iteration finishes its inner products and raises `done`.  The
    // FP32-refined form also spends pairs on residual replacements, so up to
    // max_iter + max_iter/refine_every + 2 pairs can be needed.
    // FP32-refined: a replacement costs a pair too, at most one per iteration
    const long long maxpairs = (long long)init.max_iter * (x->fp32 ? 2 : 1) + 4;
    long long pairs = 0;
    int t = 0;  // launch index since the setup
    int batch = last_iters + 1, next = 4;
    while (true) {
        for (int m = 0; m < batch && pairs < maxpairs; ++m, ++pairs) {
            const int ta = ++t, tb = ++t;
            if (x->fp32) {
                rc = timed(x, 2, [&](const hzo::LaunchEv& ev) { hzo::launch_pcg_dir_f32(x->stream, x->g, c, b32, ta, ev); });
                if (rc) return rc;
                fold_after(x, x->stream, ta);
                rc = timed(x, 3, [&](const hzo::LaunchEv& ev) { hzo::launch_pcg_update_f32(x->stream, x->g, c, b32, tb, ev); });
            } else {
                rc = timed(x, 2, [&](const hzo::LaunchEv& ev) { hzo::launch_pcg_dir_f64(x->stream, x->g, c, b, ta, ev); });
                if (rc) return rc;
                fold_after(x, x->stream, ta);
                rc = timed(x, 3, [&](const hzo::LaunchEv& ev) { hzo::launch_pcg_update_f64(x->stream, x->g, c, b, tb, ev); });
            }
            if (rc) return rc;
            fold_after(x, x->stream, tb);
        }
        HZO_HIP(hipMemcpyAsync(x->st_host, x->st + (t & 1), sizeof(hzo::PcgState), hipMemcpyDeviceToHost,
                               x->stream));
        HZO_HIP(hipStreamSynchronize(x->stream));
        if (x->st_host->done || pairs >= maxpairs) break;
        batch = follow_up_round(next, last_iters);
    }
    hzo::PcgState& s = *x->st_host;
    if (!s.done) s.status = HZO_PCG_MAXITER;
    last_iters = s.iters + s.refinements > 0 ? s.iters + s.refinements : 1;
    return copy_stats(s, stats);
}

int depol(hzo_ctx* x, hzo_pcg_stats* stats) {
    return pcg_solve(x, bufs(x), bufs32(x), x->c, pcg_init(x), nullptr, x->last_iters, stats);
}

void free_adj(hzo_ctx* x) {
    hzo_ctx::Adjoint& A = x->adj;
    (void)hipFree(A.ckP);
    (void)hipFree(A.ckphi);
    (void)hipFree(A.cklo);
    (void)hipFree(A.seg);
    (void)hipFree(A.a);
    (void)hipFree(A.w);
    (void)hipFree(A.mu);
    (void)hipFree(A.mulo);
    (void)hipFree(A.rhs);
    (void)hipFree(A.gpart);
    (void)hipFree(A.grad);
    (void)hipFree(A.m);
    A = hzo_ctx::Adjoint();
}

void free_all(hzo_ctx* x) {
    free_adj(x);
    (void)hipFree(x->P[0]);
    (void)hipFree(x->P[1]);
    (void)hipFree(x->phi);
    (void)hipFree(x->r);
    (void)hipFree(x->p[0]);
    (void)hipFree(x->p[1]);
    (void)hipFree(x->cw);
    (void)hipFree(x->sc);
    (void)hipFree(x->TR);
    (void)hipFree(x->ptab);
    (void)hipFree(x->lut);
    (void)hipFree(x->part);
    (void)hipFree(x->fold);
    (void)hipFree(x->tdgl_part);
    (void)hipFree(x->diag_part);
    (void)hipFree(x->st);
    (void)hipFree(x->P32[0]);
    (void)hipFree(x->P32[1]);
    if (x->cg_arena[0] || x->cg_arena[1]) {  // the vectors are views into the arenas
        (void)hipFree(x->cg_arena[0]);
        (void)hipFree(x->cg_arena[1]);
    } else {
        (void)hipFree(x->xlo);
        (void)hipFree(x->r32);
        (void)hipFree(x->p32[0]);
    }
    (void)hipFree(x->p32[1]);
    (void)hipFree(x->b64);
    (void)hipFree(x->TR32);
    (void)hipFree(x->ptab32);
    (void)hipFree(x->lut32);
    if (x->st_host) (void)hipHostFree(x->st_host);
    (void)hipFree(x->res_xch);
    (void)hipFree(x->res_xchd);
    (void)hipFree(x->res_part);
    (void)hipFree(x->res_sync);
    if (x->res_err_host) (void)hipHostFree(x->res_err_host);
    if (x->res_ring) (void)hipHostFree(x->res_ring);
    if (x->res_err_ring) (void)hipHostFree(x->res_err_ring);
    for (int c = 0; c < HZO_KCLASSES; ++c)
        for (int q = 0; q < 2; ++q)
            for (auto e : x->ev[c][q]) (void)hipEventDestroy(e);
    if (x->stream) (void)hipStreamDestroy(x->stream);
}

}  // namespace

extern "C" {

int hzo_device_count(int* n) {
    if (!n) return set_error(HZO_ERR_ARG, "null argument");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return HZO_OK;
}

void hzo_config_default(hzo_config* cfg) {
    if (!cfg) return;
    memset(cfg, 0, sizeof(*cfg));
    cfg->nx = cfg->ny = cfg->nz = 0;
    cfg->del_h = 1.0;
    cfg->del_t = 0.01;
    cfg->dep_flag = 1;
    cfg->zbc = HZO_ZBC_DIRICHLET;
    cfg->eps0_nd = 1.0;
    cfg->g_bot = 0.0;  // global.cu:140-144
    cfg->g_top = 0.0;
    cfg->precision = HZO_PREC_FP64;
    cfg->pcg_rule = HZO_PCG_RULE_REFERENCE;
    cfg->tol_abs = 1.0e-3;
    cfg->tol_rel = 1.0e-2;
    cfg->max_iter = 4000;
    cfg->refine_every = 20;
    cfg->tile_kz = 0;
    cfg->cg_fused = 1;
}

int hzo_create(const hzo_config* cfg, hzo_ctx** out) {
    if (!cfg || !out) return set_error(HZO_ERR_ARG, "hzo_create: null argument");
    *out = nullptr;
    if (cfg->nx <= 0 || cfg->ny <= 0 || cfg->nz <= 0)
        return set_error(HZO_ERR_ARG, "hzo_create: grid %dx%dx%d", cfg->nx, cfg->ny, cfg->nz);
    if (!(cfg->del_h > 0.0) || !(cfg->del_t > 0.0)) return set_error(HZO_ERR_ARG, "hzo_create: del_h, del_t must be > 0");
    if (cfg->zbc != HZO_ZBC_DIRICHLET && cfg->zbc != HZO_ZBC_NEUMANN)
        return set_error(HZO_ERR_ARG, "hzo_create: zbc %d", cfg->zbc);
    if (cfg->zbc == HZO_ZBC_NEUMANN && cfg->nz < 3)
        return set_error(HZO_ERR_ARG, "hzo_create: Neumann mirror needs nz >= 3");
    if (cfg->zbc == HZO_ZBC_DIRICHLET && cfg->nz < 2)
        return set_error(HZO_ERR_ARG, "hzo_create: Dirichlet electrodes need nz >= 2");
    if (cfg->precision != HZO_PREC_FP64 && cfg->precision != HZO_PREC_FP32_REFINED)
        return set_error(HZO_ERR_ARG, "hzo_create: precision %d", cfg->precision);
    if (cfg->refine_every < 0) return set_error(HZO_ERR_ARG, "hzo_create: refine_every < 0");
    if (cfg->tile_kz != 0 && cfg->tile_kz != 2 && cfg->tile_kz != 4 && cfg->tile_kz != 8 &&
        !(cfg->tile_kz == 16 && (cfg->precision == HZO_PREC_FP64 ||
                                 (cfg->precision == HZO_PREC_FP32_REFINED && cfg->cg_fused))) &&
        !(cfg->tile_kz == 32 && (cfg->precision == HZO_PREC_FP64 ||
                                 (cfg->precision == HZO_PREC_FP32_REFINED &&
                                  (cfg->cg_fused == 1 || cfg->cg_fused == 3 || cfg->cg_fused == 4)))))
        return set_error(HZO_ERR_ARG, "hzo_create: tile_kz must be 0, 2, 4 or 8 (16: FP64 or the single-reduction "
                                      "FP32 CG; 32: FP64 or that CG's two-stencil form)");
    if (cfg->pcg_rule != 0 && cfg->pcg_rule != 1) return set_error(HZO_ERR_ARG, "hzo_create: pcg_rule %d", cfg->pcg_rule);
    if (cfg->max_iter < 0) return set_error(HZO_ERR_ARG, "hzo_create: max_iter < 0");
    const long long N = (long long)cfg->nx * cfg->ny * cfg->nz;
    if (N > (1LL << 28)) return set_error(HZO_ERR_ARG, "hzo_create: grid too large (at most 2^28 voxels)");
    // the single-reduction CG addresses four of its vectors through one
    // buffer range of 16 N bytes, which must stay below 2^31: larger films
    // run the two-kernel FP32 PCG
    if (cfg->cg_fused < 0 || cfg->cg_fused > 4) return set_error(HZO_ERR_ARG, "hzo_create: cg_fused %d", cfg->cg_fused);
    const bool cg_fused = cfg->cg_fused && 16 * N < (1LL << 31);
    if (cfg->tile_kz >= 16 && cfg->precision == HZO_PREC_FP32_REFINED && !cg_fused)
        return set_error(HZO_ERR_ARG, "hzo_create: tile_kz %d needs the single-reduction CG (at most 2^27 - 1 voxels)",
                         cfg->tile_kz);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(HZO_ERR_HIP, "hzo_create: no HIP device available");

    hzo_ctx* x = new hzo_ctx();
    x->cfg = *cfg;
    x->cfg.cg_fused = cg_fused ? cfg->cg_fused : 0;
    x->g.nx = cfg->nx;
    x->g.ny = cfg->ny;
    x->g.nz = cfg->nz;
    x->g.N = N;
    // Tile depth (z planes per 64x8 workgroup) unless the caller fixes it;
    // the inner-product grouping, hence the last bits of the dot products,
    // follows the tiling.  Single-reduction FP32 CG (z-streaming), measured
    // on MI355X: once depth 8 would need more than one round of workgroups
    // (512: two per CU), the two-stencil form at the deepest of {32, 16}
    // that still gives 512 workgroups and pads at most an eighth of the
    // planes (automatic choice or form 3); otherwise the carried form at the
    // deepest of {8, 4} giving at least 512 workgroups, else 2, and 16 when
    // depth 8 would leave more than 2048 partial sums.  Two-kernel FP32 PCG:
    // the smallest depth giving at most 512 workgroups (one partial per
    // consuming thread, two workgroups per CU), else 8.  FP64: films that
    // fill 512 workgroups of depth 8 run the z-streaming PCG kernels at the
    // deepest of {32, 16, 8} that still gives 512 workgroups and pads at
    // most an eighth of the planes (measured per CG iteration: 256x256x32
    // depth 8 27.7 us against 32.3 at the shallow kernels' depth 4;
    // 512x512x64 depth 32 about 200 against 300 us); smaller films the
    // shallow kernels at the smallest of {2, 4} giving at most 1024
    // workgroups, else 4.
    {
        const long long txy = (long long)((cfg->nx + 63) / 64) * ((cfg->ny + 7) / 8);
        const bool f32 = cfg->precision == HZO_PREC_FP32_REFINED && !x->cfg.cg_fused;
        const bool fused = cfg->precision == HZO_PREC_FP32_REFINED && x->cfg.cg_fused;
        auto tiles_at = [&](int k) { return txy * ((cfg->nz + k - 1) / k); };
        int kz = cfg->tile_kz;
        if (kz == 0 && fused && cfg->cg_fused != 2 && tiles_at(8) > 512)
            for (int cand : {32, 16})
                if (tiles_at(cand) >= 512 && 8LL * ((cfg->nz + cand - 1) / cand) * cand <= 9LL * cfg->nz) {
                    kz = cand;
                    break;
                }
        if (kz == 0 && fused) {
            kz = 2;
            for (int cand : {8, 4})
                if (tiles_at(cand) >= 512) {
                    kz = cand;
                    break;
                }
            // very large films: depth 16 keeps the partial sums at most
            // 2048 (four per consuming thread) instead of a fold launch
            if (kz == 8 && tiles_at(8) > 2048 && tiles_at(16) >= 512) kz = 16;
        } else if (kz == 0 && !f32 && tiles_at(8) >= 512) {
            for (int cand : {32, 16, 8})
                if (tiles_at(cand) >= 512 && 8LL * ((cfg->nz + cand - 1) / cand) * cand <= 9LL * cfg->nz) {
                    kz = cand;
                    break;
                }
            if (kz == 0) kz = 8;
        } else if (kz == 0) {
            kz = f32 ? 8 : 4;
            const long long cap = f32 ? 512 : 1024;
            for (int cand : {2, 4, 8}) {
                if (!f32 && cand == 8) break;
                if (txy * ((cfg->nz + cand - 1) / cand) <= cap) {
                    kz = cand;
                    break;
                }
            }
        }
        x->g.kz = kz;
        // the two-stencil form moves 26 instead of 42 bytes per voxel and
        // iteration but does more work per byte: it runs the deep tiles
        // (16, 32) of films beyond one round of depth-8 workgroups, the
        // carried form the shallow tiles of smaller ones
        x->cg_two = fused && (cfg->cg_fused == 3 || (cfg->cg_fused == 1 && kz >= 16));
        // its vectors r, p (twice), x_lo and the codes, 22 bytes per voxel:
        // beyond the 256 MB last-level cache nothing it stores is read back
        // from there, and nontemporal stores measured 3 % faster at
        // 512x512x64 (slower at every film whose vectors fit)
        x->cg_nt = x->cg_two && kz >= 16 && 22LL * N > (256LL << 20);
        // FP64 deep tiles: r, p (twice), x and the codes, 34 bytes per voxel;
        // beyond the last-level cache their streamed loads are nontemporal
        x->pcg_ntl = cfg->precision == HZO_PREC_FP64 && kz >= 8 && 34LL * N > (256LL << 20);
    }
    (void)hipGetDevice(&x->device);
    // Register-resident CG: automatic (cg_fused 1, tile depth automatic) or
    // forced (4) when every 16 x 16 column tile of the film fits the device
    // at once, for film depths with a kernel instance (16, 24, 32).
    {
        const bool fused = cfg->precision == HZO_PREC_FP32_REFINED && x->cfg.cg_fused;
        const bool shape = fused && cfg->nx % 16 == 0 && cfg->ny % 16 == 0 &&
                           (cfg->nz == 16 || cfg->nz == 24 || cfg->nz == 32);
        const int G = (cfg->nx / 16) * (cfg->ny / 16);
        const int cap = shape ? hzo::cgres_capacity(x->device, cfg->nz, cfg->zbc != 0) : 0;
        x->cg_res = ((cfg->cg_fused == 1 && cfg->tile_kz == 0) || cfg->cg_fused == 4) && shape && G <= cap;
        if (cfg->cg_fused == 4 && !x->cg_res) {
            delete x;
            return set_error(HZO_ERR_ARG, "hzo_create: the resident CG needs FP32-refined, nx and ny multiples of 16, "
                                          "nz of 16, 24 or 32 and at most %d tiles of 16 x 16 columns (%d here)",
                             cap, G);
        }
        x->res_gxt = cfg->nx / 16;
        x->res_G = G;
    }
    hzo::StepConsts& c = x->c;
    c.nx = cfg->nx;
    c.ny = cfg->ny;
    c.nz = cfg->nz;
    c.kz = x->g.kz;
    c.dep_flag = cfg->dep_flag;
    c.zbc = cfg->zbc;
    c.del_h = cfg->del_h;
    c.del_t = cfg->del_t;
    c.g_bot = cfg->g_bot;
    c.g_top = cfg->g_top;
    c.eps0 = cfg->eps0_nd;
    c.two_h = 2.0 * cfg->del_h;
    c.inv_two_h = 1.0 / c.two_h;
    c.two_h_pow2 = is_pow2(c.two_h);
    c.h2 = cfg->del_h * cfg->del_h;
    c.inv_h2 = 1.0 / c.h2;
    c.h2_pow2 = is_pow2(c.h2);
    c.plane = (unsigned)cfg->nx * (unsigned)cfg->ny;
    c.nvox = (unsigned)N;
    c.ref_id = (long long)(cfg->nx / 2) + (long long)cfg->nx * ((long long)(cfg->ny / 2) + (long long)cfg->ny * (cfg->nz / 2));

    x->nb = hzo::blocks_for(x->g);
    x->ps = hzo::part_stride(x->g);
    auto fail = [&](int rc) {
        free_all(x);
        delete x;
        return rc;
    };
#define HZO_ALLOC(ptr, bytes)                                                                        \
    do {                                                                                             \
        if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess)                                       \
            return fail(set_error(HZO_ERR_NOMEM, "hzo_create: hipMalloc of %zu bytes failed", (size_t)(bytes))); \
    } while (0)
    if (hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_error(HZO_ERR_HIP, "hzo_create: stream creation failed"));
    x->fp32 = cfg->precision == HZO_PREC_FP32_REFINED;
    HZO_ALLOC(x->phi, sizeof(double) * N);
    if (x->fp32) {
        HZO_ALLOC(x->P32[0], sizeof(float) * 6 * N);
        HZO_ALLOC(x->P32[1], sizeof(float) * 6 * N);
        if (x->cfg.cg_fused) {
            // one buffer resource addresses four arrays: 16 N bytes must
            // stay below 2^31
            HZO_ALLOC(x->cg_arena[0], sizeof(float) * 4 * N);
            HZO_ALLOC(x->cg_arena[1], sizeof(float) * 4 * N);
            x->r32 = x->cg_arena[0];
            x->cg_r1 = x->cg_arena[0] + N;
            x->cg_w[0] = x->cg_arena[0] + 2 * N;
            x->cg_w[1] = x->cg_arena[0] + 3 * N;
            x->cg_s[0] = x->cg_arena[1];
            x->cg_s[1] = x->cg_arena[1] + N;
            x->p32[0] = x->cg_arena[1] + 2 * N;
            x->xlo = x->cg_arena[1] + 3 * N;
        } else {
            HZO_ALLOC(x->xlo, sizeof(float) * N);
            HZO_ALLOC(x->r32, sizeof(float) * N);
            HZO_ALLOC(x->p32[0], sizeof(float) * N);
        }
        HZO_ALLOC(x->p32[1], sizeof(float) * N);
        HZO_ALLOC(x->b64, sizeof(double) * N);
        HZO_ALLOC(x->ptab32, sizeof(hzo::PhaseC32) * hzo::PTAB_ENTRIES);
        HZO_ALLOC(x->lut32, sizeof(float) * hzo::LUT_DOUBLES);
    } else {
        HZO_ALLOC(x->P[0], sizeof(double) * 6 * N);
        HZO_ALLOC(x->P[1], sizeof(double) * 6 * N);
        HZO_ALLOC(x->r, sizeof(double) * N);
        HZO_ALLOC(x->p[0], sizeof(double) * N);
        HZO_ALLOC(x->p[1], sizeof(double) * N);
    }
    HZO_ALLOC(x->cw, sizeof(uint32_t) * N);
    HZO_ALLOC(x->sc, sizeof(uint16_t) * N);
    HZO_ALLOC(x->ptab, sizeof(hzo::PhaseC) * hzo::PTAB_ENTRIES);
    HZO_ALLOC(x->lut, sizeof(double) * hzo::LUT_DOUBLES);
    HZO_ALLOC(x->part, sizeof(double) * 2 * hzo::RED_MAXQ * (size_t)x->ps);
    if (x->nb > 4 * hzo::NT) HZO_ALLOC(x->fold, sizeof(double) * 2 * hzo::RED_MAXQ * hzo::NT);
    HZO_ALLOC(x->tdgl_part, sizeof(double) * hzo::TDGL_Q * (size_t)x->nb);
    HZO_ALLOC(x->diag_part, sizeof(double) * hzo::DIAG_Q * (size_t)x->nb);
    HZO_ALLOC(x->st, sizeof(hzo::PcgState) * 2);
    if (x->cg_res) {
        HZO_ALLOC(x->res_xch, sizeof(float) * 2 * (size_t)x->res_G * 4 * cfg->nz * 16);
        HZO_ALLOC(x->res_xchd, sizeof(double) * (size_t)x->res_G * 4 * cfg->nz * 16);
        HZO_ALLOC(x->res_part, sizeof(double) * 2 * (size_t)x->res_G * 4);
        HZO_ALLOC(x->res_sync, sizeof(unsigned) * (hzo::RES_SYNC_WORDS + 64));
        if (hipHostMalloc((void**)&x->res_err_host, sizeof(int), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&x->res_ring, sizeof(hzo::PcgState) * RUN_RING, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&x->res_err_ring, sizeof(int) * RUN_RING, hipHostMallocDefault) != hipSuccess)
            return fail(set_error(HZO_ERR_NOMEM, "hzo_create: pinned host allocation failed"));
    }
#undef HZO_ALLOC
    if (hipHostMalloc((void**)&x->st_host, sizeof(hzo::PcgState), hipHostMallocDefault) != hipSuccess)
        return fail(set_error(HZO_ERR_NOMEM, "hzo_create: pinned host allocation failed"));
    if (hipMemset(x->part, 0, sizeof(double) * 2 * hzo::RED_MAXQ * (size_t)x->ps) != hipSuccess)
        return fail(set_error(HZO_ERR_HIP, "hzo_create: memset failed"));
    if (x->fold && hipMemset(x->fold, 0, sizeof(double) * 2 * hzo::RED_MAXQ * hzo::NT) != hipSuccess)
        return fail(set_error(HZO_ERR_HIP, "hzo_create: memset failed"));
    if (hipMemset(x->tdgl_part, 0, sizeof(double) * hzo::TDGL_Q * (size_t)x->nb) != hipSuccess)
        return fail(set_error(HZO_ERR_HIP, "hzo_create: memset failed"));
    *out = x;
    return HZO_OK;
}

void hzo_destroy(hzo_ctx* x) {
    if (!x) return;
    (void)hipSetDevice(x->device);
    (void)hipStreamSynchronize(x->stream);
    free_all(x);
    delete x;
}

int hzo_set_microstructure(hzo_ctx* x, const int32_t* omega, int ng, const int32_t* grain_phase,
                           const double* TR, const int32_t* iface) {
    if (!x || !omega || !grain_phase || !TR || ng <= 0)
        return set_error(HZO_ERR_ARG, "hzo_set_microstructure: bad argument");
    if (ng > (int)hzo::CW_GID_MASK) return set_error(HZO_ERR_ARG, "hzo_set_microstructure: too many grains");
    HZO_HIP(hipSetDevice(x->device));
    const long long N = x->g.N;
    x->ng = ng;
    x->omega.assign(omega, omega + N);
    x->grain_phase.assign(grain_phase, grain_phase + ng);
    x->iface.assign(N, 0);
    if (iface) {
        for (long long i = 0; i < N; ++i) x->iface[i] = iface[i] ? 1 : 0;
    } else {
        int rc = hzo_interface_flag(x->g.nx, x->g.ny, x->g.nz, omega, ng, grain_phase, x->iface.data());
        if (rc) return rc;
    }
    (void)hipFree(x->TR);
    x->TR = nullptr;
    if (hipMalloc((void**)&x->TR, sizeof(double) * 9 * (size_t)ng) != hipSuccess)
        return set_error(HZO_ERR_NOMEM, "hzo_set_microstructure: TR allocation failed");
    HZO_HIP(hipMemcpy(x->TR, TR, sizeof(double) * 9 * (size_t)ng, hipMemcpyHostToDevice));
    if (x->fp32) {
        (void)hipFree(x->TR32);
        x->TR32 = nullptr;
        if (hipMalloc((void**)&x->TR32, sizeof(float) * 9 * (size_t)ng) != hipSuccess)
            return set_error(HZO_ERR_NOMEM, "hzo_set_microstructure: TR allocation failed");
        std::vector<float> t32(TR, TR + 9 * (size_t)ng);
        HZO_HIP(hipMemcpy(x->TR32, t32.data(), sizeof(float) * 9 * (size_t)ng, hipMemcpyHostToDevice));
    }
    int rc = upload_micro(x);
    if (rc) return rc;
    x->have_micro = true;
    return HZO_OK;
}

int hzo_set_phase_table(hzo_ctx* x, const hzo_phase_nd ph[3]) {
    if (!x || !ph) return set_error(HZO_ERR_ARG, "hzo_set_phase_table: null argument");
    HZO_HIP(hipSetDevice(x->device));
    for (int q = 0; q < 3; ++q) x->ph[q] = ph[q];
    int rc = upload_phase(x);
    if (rc) return rc;
    x->have_phase = true;
    return HZO_OK;
}

int hzo_set_state(hzo_ctx* x, const double* P, const double* phi) {
    if (!x || !P) return set_error(HZO_ERR_ARG, "hzo_set_state: null argument");
    HZO_HIP(hipSetDevice(x->device));
    const long long N = x->g.N;
    x->cur = 0;
    if (x->fp32) {
        std::vector<float> p32(P, P + 6 * N);
        HZO_HIP(hipMemcpy(x->P32[0], p32.data(), sizeof(float) * 6 * N, hipMemcpyHostToDevice));
        HZO_HIP(hipMemset(x->xlo, 0, sizeof(float) * N));
    } else {
        HZO_HIP(hipMemcpy(x->P[0], P, sizeof(double) * 6 * N, hipMemcpyHostToDevice));
    }
    if (phi) {
        HZO_HIP(hipMemcpy(x->phi, phi, sizeof(double) * N, hipMemcpyHostToDevice));
    } else {  // init_phi_linear (depol_varEPS.cu:203-221)
        std::vector<double> h(N);
        const int nx = x->g.nx, ny = x->g.ny, nz = x->g.nz;
        for (int k = 0; k < nz; ++k) {
            double val = 0.0;
            if (x->c.zbc == 0) {
                const double t = (nz <= 1) ? 0.0 : (double)k / (double)(nz - 1);
                val = (1.0 - t) * x->c.g_bot + t * x->c.g_top;
            }
            for (long long q = 0; q < (long long)nx * ny; ++q) h[(long long)nx * ny * k + q] = val;
        }
        HZO_HIP(hipMemcpy(x->phi, h.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    }
    x->have_state = true;
    x->last_iters = 8;
    return HZO_OK;
}

int hzo_get_state(hzo_ctx* x, double* P, double* phi, double* E) {
    if (!x) return set_error(HZO_ERR_ARG, "hzo_get_state: null context");
    HZO_HIP(hipSetDevice(x->device));
    HZO_HIP(hipStreamSynchronize(x->stream));
    const long long N = x->g.N;
    if (P) {
        if (x->fp32) {
            std::vector<float> p32(6 * N);
            HZO_HIP(hipMemcpy(p32.data(), x->P32[x->cur], sizeof(float) * 6 * N, hipMemcpyDeviceToHost));
            for (long long q = 0; q < 6 * N; ++q) P[q] = p32[q];
        } else {
            HZO_HIP(hipMemcpy(P, x->P[x->cur], sizeof(double) * 6 * N, hipMemcpyDeviceToHost));
        }
    }
    std::vector<double> ph;
    if (phi || E) {
        ph.resize(N);
        HZO_HIP(hipMemcpy(ph.data(), x->phi, sizeof(double) * N, hipMemcpyDeviceToHost));
        if (x->fp32) {  // phi = x_hi + x_lo
            std::vector<float> lo(N);
            HZO_HIP(hipMemcpy(lo.data(), x->xlo, sizeof(float) * N, hipMemcpyDeviceToHost));
            for (long long q = 0; q < N; ++q) ph[q] += (double)lo[q];
        }
        if (phi) memcpy(phi, ph.data(), sizeof(double) * N);
    }
    if (E) {  // E = -grad phi, depol_varEPS.cu:223-276 (host evaluation of the same expressions)
        const int nx = x->g.nx, ny = x->g.ny, nz = x->g.nz;
        const double del_h = x->cfg.del_h;
        auto at = [&](int i, int j, int k) { return (long long)i + (long long)nx * ((long long)j + (long long)ny * k); };
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const long long id = at(i, j, k);
                    const int e = (i + 1) % nx, w = (i - 1 + nx) % nx, n = (j + 1) % ny, s = (j - 1 + ny) % ny;
                    E[id] = -(ph[at(e, j, k)] - ph[at(w, j, k)]) / (2.0 * del_h);
                    E[N + id] = -(ph[at(i, n, k)] - ph[at(i, s, k)]) / (2.0 * del_h);
                    double up, dn;
                    if (x->c.zbc == 0) {
                        up = (k == nz - 1) ? x->c.g_top : ph[at(i, j, k + 1)];
                        dn = (k == 0) ? x->c.g_bot : ph[at(i, j, k - 1)];
                    } else {
                        const int uu = (k == 0) ? 1 : ((k == nz - 1) ? nz - 2 : k + 1);
                        const int dd = (k == 0) ? 1 : ((k == nz - 1) ? nz - 2 : k - 1);
                        up = ph[at(i, j, uu)];
                        dn = ph[at(i, j, dd)];
                    }
                    E[2 * N + id] = -(up - dn) / (2 * del_h);
                }
    }
    return HZO_OK;
}

int hzo_tdgl_step(hzo_ctx* x, double Ez_ext) {
    int rc = check_ready(x);
    if (rc) return rc;
    HZO_HIP(hipSetDevice(x->device));
    return tdgl(x, Ez_ext);
}

int hzo_depol_solve(hzo_ctx* x, hzo_pcg_stats* stats) {
    int rc = check_ready(x);
    if (rc) return rc;
    HZO_HIP(hipSetDevice(x->device));
    return depol(x, stats);
}

int hzo_step(hzo_ctx* x, double Ez_ext, hzo_pcg_stats* stats) {
    int rc = check_ready(x);
    if (rc) return rc;
    HZO_HIP(hipSetDevice(x->device));
    rc = tdgl(x, Ez_ext);
    if (rc) return rc;
    if (x->cfg.dep_flag) return depol(x, stats);
    if (stats) memset(stats, 0, sizeof(*stats));
    return HZO_OK;
}

int hzo_run(hzo_ctx* x, const double* Ez, int nsteps, int* its) {
    int rc = check_ready(x);
    if (rc) return rc;
    if (nsteps < 0 || (nsteps > 0 && !Ez)) return set_error(HZO_ERR_ARG, "hzo_run: bad argument");
    HZO_HIP(hipSetDevice(x->device));
    if (x->cg_res && x->cfg.dep_flag) {
        // Resident film: the steps are enqueued back to back (TDGL, setup,
        // resident solve, the copies of its control block and timeout flag
        // into a ring slot), and the host waits once per RUN_RING steps
        // instead of after every solve.  A timed-out grid barrier is
        // reported when its group of steps is collected.
        for (int s0 = 0; s0 < nsteps; s0 += RUN_RING) {
            const int m = nsteps - s0 < RUN_RING ? nsteps - s0 : RUN_RING;
            for (int j = 0; j < m; ++j) {
                rc = tdgl(x, Ez[s0 + j]);
                if (rc) return rc;
                rc = cg32_solve(x, bufs32(x), x->c, pcg_init(x), nullptr, x->last_iters, nullptr, j);
                if (rc) return rc;
            }
            HZO_HIP(hipStreamSynchronize(x->stream));
            for (int j = 0; j < m; ++j) {
                hzo_pcg_stats st;
                memset(&st, 0, sizeof(st));
                rc = cgres_collect(x, x->res_ring[j], x->res_err_ring[j], x->last_iters, &st);
                if (rc) return rc;
                if (its) its[s0 + j] = st.iters;
            }
        }
        return HZO_OK;
    }
    for (int s = 0; s < nsteps; ++s) {
        hzo_pcg_stats st;
        memset(&st, 0, sizeof(st));
        rc = tdgl(x, Ez[s]);
        if (rc) return rc;
        if (x->cfg.dep_flag) {
            rc = depol(x, &st);
            if (rc) return rc;
        }
        if (its) its[s] = st.iters;
    }
    return HZO_OK;
}

int hzo_diagnostics(hzo_ctx* x, double out[10]) {
    int rc = check_ready(x);
    if (rc) return rc;
    if (!out) return set_error(HZO_ERR_ARG, "hzo_diagnostics: null output");
    HZO_HIP(hipSetDevice(x->device));
    if (x->fp32)
        hzo::launch_diag_f32(x->stream, x->g, x->P32[x->cur], x->cw, x->diag_part);
    else
        hzo::launch_diag_f64(x->stream, x->g, x->P[x->cur], x->cw, x->diag_part);
    HZO_HIP(hipGetLastError());
    const size_t nb = (size_t)x->nb;
    x->diag_host.resize((hzo::DIAG_Q + hzo::TDGL_Q) * nb);
    HZO_HIP(hipMemcpyAsync(x->diag_host.data(), x->diag_part, sizeof(double) * hzo::DIAG_Q * nb,
                           hipMemcpyDeviceToHost, x->stream));
    HZO_HIP(hipMemcpyAsync(x->diag_host.data() + hzo::DIAG_Q * nb, x->tdgl_part, sizeof(double) * hzo::TDGL_Q * nb,
                           hipMemcpyDeviceToHost, x->stream));
    HZO_HIP(hipStreamSynchronize(x->stream));
    // fixed-order sums of the per-workgroup partials
    double s[hzo::DIAG_Q + hzo::TDGL_Q];
    for (int q = 0; q < hzo::DIAG_Q + hzo::TDGL_Q; ++q) {
        double a = 0.0;
        for (size_t i = 0; i < nb; ++i) a += x->diag_host[q * nb + i];
        s[q] = a;
    }
    out[0] = s[0];
    out[1] = s[1];
    out[2] = s[hzo::DIAG_Q];      // grain-frame Pmz of the last TDGL input
    out[3] = s[hzo::DIAG_Q + 1];  // grain-frame Psz
    for (int q = 0; q < 6; ++q) out[4 + q] = s[2 + q];
    return HZO_OK;
}

int hzo_synchronize(hzo_ctx* x) {
    if (!x) return set_error(HZO_ERR_ARG, "null context");
    HZO_HIP(hipSetDevice(x->device));
    HZO_HIP(hipStreamSynchronize(x->stream));
    return HZO_OK;
}

int hzo_grid_dims(hzo_ctx* x, int dims[3]) {
    if (!x || !dims) return set_error(HZO_ERR_ARG, "hzo_grid_dims: null argument");
    dims[0] = x->g.nx;
    dims[1] = x->g.ny;
    dims[2] = x->g.nz;
    return HZO_OK;
}

int hzo_tile_depth(hzo_ctx* x) { return x ? x->g.kz : 0; }
int hzo_cg_form(hzo_ctx* x) {
    return !x ? -1 : (!x->fp32 || !x->cfg.cg_fused) ? 0 : x->cg_res ? 3 : (x->cg_two ? 2 : 1);
}

int hzo_set_timing(hzo_ctx* x, int enable) {
    if (!x) return set_error(HZO_ERR_ARG, "null context");
    HZO_HIP(hipSetDevice(x->device));
    HZO_HIP(hipStreamSynchronize(x->stream));
    x->timing = enable != 0;
    for (int c = 0; c < HZO_KCLASSES; ++c) x->nt[c] = 0;
    for (long long& v : x->res_counts) v = 0;
    return HZO_OK;
}

int hzo_cg_counters(hzo_ctx* x, long long out[3]) {
    if (!x || !out) return set_error(HZO_ERR_ARG, "hzo_cg_counters: null argument");
    for (int i = 0; i < 3; ++i) out[i] = x->res_counts[i];
    return HZO_OK;
}

int hzo_kernel_times(hzo_ctx* x, int counts[HZO_KCLASSES], double mean_ms[HZO_KCLASSES]) {
    if (!x || !counts || !mean_ms) return set_error(HZO_ERR_ARG, "hzo_kernel_times: null argument");
    HZO_HIP(hipSetDevice(x->device));
    HZO_HIP(hipStreamSynchronize(x->stream));
    for (int c = 0; c < HZO_KCLASSES; ++c) {
        std::vector<double> t(x->nt[c]);
        for (int i = 0; i < x->nt[c]; ++i) {
            float ms = 0.f;
            HZO_HIP(hipEventElapsedTime(&ms, x->ev[c][0][i], x->ev[c][1][i]));
            t[i] = ms;
        }
        // launches that found the solve already finished return at once;
        // they take under a quarter of the median and are left out
        double med = 0.0;
        if (!t.empty()) {
            std::vector<double> srt(t);
            std::nth_element(srt.begin(), srt.begin() + srt.size() / 2, srt.end());
            med = srt[srt.size() / 2];
        }
        double tot = 0.0;
        int n = 0;
        for (double v : t)
            if (v >= 0.25 * med) {
                tot += v;
                ++n;
            }
        counts[c] = n;
        mean_ms[c] = n ? tot / n : 0.0;
    }
    return HZO_OK;
}


/* ---------------------------------------------------------------- reverse mode */

int hzo_adjoint_reserve(hzo_ctx* x, int nsteps, int ckpt_every, double tol) {
    if (!x) return set_error(HZO_ERR_ARG, "hzo_adjoint_reserve: null context");
    if (nsteps < 1 || ckpt_every < 1 || !(tol > 0.0))
        return set_error(HZO_ERR_ARG, "hzo_adjoint_reserve: nsteps, ckpt_every and tol must be positive");
    if (x->cfg.dep_flag && x->c.zbc != 0)
        return set_error(HZO_ERR_ARG,
                         "hzo_adjoint_reserve: the adjoint needs Dirichlet electrodes (the Neumann gauge row makes "
                         "the Poisson operator non-symmetric)");
    HZO_HIP(hipSetDevice(x->device));
    HZO_HIP(hipStreamSynchronize(x->stream));
    free_adj(x);
    hzo_ctx::Adjoint& A = x->adj;
    const long long N = x->g.N;
    const size_t ts = x->fp32 ? sizeof(float) : sizeof(double);
    const int K = ckpt_every < nsteps ? ckpt_every : nsteps;
    A.nsteps = nsteps;
    A.K = K;
    A.nck = (nsteps + K - 1) / K + 1;
    A.tol = tol;
#define HZO_AALLOC(ptr, bytes)                                                                         \
    do {                                                                                               \
        if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                                        \
            free_adj(x);                                                                               \
            return set_error(HZO_ERR_NOMEM, "hzo_adjoint_reserve: hipMalloc of %zu bytes failed", (size_t)(bytes)); \
        }                                                                                              \
    } while (0)
    HZO_AALLOC(A.ckP, ts * 6 * N * A.nck);
    HZO_AALLOC(A.ckphi, sizeof(double) * N * A.nck);
    if (x->fp32) HZO_AALLOC(A.cklo, sizeof(float) * N * A.nck);
    HZO_AALLOC(A.seg, ts * 6 * N * K);
    HZO_AALLOC(A.a, ts * 6 * N);
    HZO_AALLOC(A.w, ts * 6 * N);
    HZO_AALLOC(A.mu, sizeof(double) * N);
    if (x->fp32) HZO_AALLOC(A.mulo, sizeof(float) * N);
    else HZO_AALLOC(A.rhs, sizeof(double) * N);
    HZO_AALLOC(A.gpart, sizeof(double) * 12 * (size_t)x->nb);
    HZO_AALLOC(A.grad, sizeof(double) * 12);
    HZO_AALLOC(A.m, sizeof(double) * nsteps);
#undef HZO_AALLOC
    return HZO_OK;
}

namespace {

const void* cur_P(hzo_ctx* x) { return x->fp32 ? (const void*)x->P32[x->cur] : (const void*)x->P[x->cur]; }

int save_ck(hzo_ctx* x, int slot) {
    hzo_ctx::Adjoint& A = x->adj;
    const long long N = x->g.N;
    const size_t ts = x->fp32 ? sizeof(float) : sizeof(double);
    HZO_HIP(hipMemcpyAsync((char*)A.ckP + ts * 6 * N * slot, cur_P(x), ts * 6 * N, hipMemcpyDeviceToDevice, x->stream));
    HZO_HIP(hipMemcpyAsync(A.ckphi + N * slot, x->phi, sizeof(double) * N, hipMemcpyDeviceToDevice, x->stream));
    if (x->fp32)
        HZO_HIP(hipMemcpyAsync(A.cklo + N * slot, x->xlo, sizeof(float) * N, hipMemcpyDeviceToDevice, x->stream));
    return HZO_OK;
}

int load_ck(hzo_ctx* x, int slot) {
    hzo_ctx::Adjoint& A = x->adj;
    const long long N = x->g.N;
    const size_t ts = x->fp32 ? sizeof(float) : sizeof(double);
    x->cur = 0;
    void* P = x->fp32 ? (void*)x->P32[0] : (void*)x->P[0];
    HZO_HIP(hipMemcpyAsync(P, (char*)A.ckP + ts * 6 * N * slot, ts * 6 * N, hipMemcpyDeviceToDevice, x->stream));
    HZO_HIP(hipMemcpyAsync(x->phi, A.ckphi + N * slot, sizeof(double) * N, hipMemcpyDeviceToDevice, x->stream));
    if (x->fp32)
        HZO_HIP(hipMemcpyAsync(x->xlo, A.cklo + N * slot, sizeof(float) * N, hipMemcpyDeviceToDevice, x->stream));
    return HZO_OK;
}

int fwd_step(hzo_ctx* x, double Ez) {
    int rc = tdgl(x, Ez);
    if (rc) return rc;
    if (x->cfg.dep_flag) return depol(x, nullptr);
    return HZO_OK;
}

}  // namespace

int hzo_forward_record(hzo_ctx* x, const double* Ez, int nsteps, double* m_out) {
    int rc = check_ready(x);
    if (rc) return rc;
    hzo_ctx::Adjoint& A = x->adj;
    if (!A.ckP) return set_error(HZO_ERR_STATE, "hzo_forward_record: call hzo_adjoint_reserve first");
    if (nsteps < 1 || nsteps > A.nsteps || !Ez)
        return set_error(HZO_ERR_ARG, "hzo_forward_record: nsteps must be 1..%d (reserved)", A.nsteps);
    HZO_HIP(hipSetDevice(x->device));
    A.Ez.assign(Ez, Ez + nsteps);
    A.recorded = 0;
    for (int n = 0; n < nsteps; ++n) {
        if (n % A.K == 0) {
            rc = save_ck(x, n / A.K);
            if (rc) return rc;
        }
        rc = fwd_step(x, Ez[n]);
        if (rc) return rc;
        if (x->fp32)
            hzo::launch_diag_f32(x->stream, x->g, x->P32[x->cur], x->cw, x->diag_part);
        else
            hzo::launch_diag_f64(x->stream, x->g, x->P[x->cur], x->cw, x->diag_part);
        hzo::launch_sum_column(x->stream, x->g, x->diag_part, A.m + n);
        HZO_HIP(hipGetLastError());
    }
    rc = save_ck(x, A.nck - 1);
    if (rc) return rc;
    std::vector<double> m(nsteps);
    HZO_HIP(hipMemcpyAsync(m.data(), A.m, sizeof(double) * nsteps, hipMemcpyDeviceToHost, x->stream));
    HZO_HIP(hipStreamSynchronize(x->stream));
    const double N = (double)x->g.N;
    if (m_out)
        for (int n = 0; n < nsteps; ++n) m_out[n] = m[n] / N;
    A.recorded = nsteps;
    return HZO_OK;
}

int hzo_backward(hzo_ctx* x, const double* dL_dm, double grad[12], long long* adj_pcg_iters) {
    int rc = check_ready(x);
    if (rc) return rc;
    hzo_ctx::Adjoint& A = x->adj;
    if (!A.recorded) return set_error(HZO_ERR_STATE, "hzo_backward: call hzo_forward_record first");
    if (!dL_dm || !grad) return set_error(HZO_ERR_ARG, "hzo_backward: null argument");
    HZO_HIP(hipSetDevice(x->device));
    const int S = A.recorded;
    const long long N = x->g.N;
    const size_t ts = x->fp32 ? sizeof(float) : sizeof(double);
    const bool dep = x->cfg.dep_flag != 0;
    // the adjoint operator: the same matrix without the electrode offsets
    hzo::StepConsts ca = x->c;
    ca.g_bot = 0.0;
    ca.g_top = 0.0;
    hzo::PcgState init = pcg_init(x);
    init.rule = HZO_PCG_RULE_RELRES;
    init.tol_rel2 = A.tol * A.tol;
    init.tol_abs2 = 0.0;
    long long iters = 0;

    HZO_HIP(hipMemsetAsync(A.grad, 0, sizeof(double) * 12, x->stream));
    HZO_HIP(hipMemsetAsync(A.mu, 0, sizeof(double) * N, x->stream));
    if (x->fp32) HZO_HIP(hipMemsetAsync(A.mulo, 0, sizeof(float) * N, x->stream));
    hzo::launch_adj_seed(x->stream, x->g, x->fp32, A.a, dL_dm[S - 1] / (double)N);
    HZO_HIP(hipGetLastError());

    hzo::AdjTDGLBufsAny kb;
    kb.fp32 = x->fp32;
    kb.a = A.a;
    kb.w = A.w;
    kb.v = x->fp32 ? x->b64 : A.rhs;
    kb.cw = x->cw;
    kb.sc = x->sc;
    kb.TR = x->fp32 ? (const void*)x->TR32 : (const void*)x->TR;
    kb.ptab = x->fp32 ? (const void*)x->ptab32 : (const void*)x->ptab;
    kb.gpart = A.gpart;
    kb.dAdk = x->cfg.del_t / (x->cfg.del_h * x->cfg.del_h);

    const int nseg = (S + A.K - 1) / A.K;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * A.K, s1 = (s0 + A.K < S) ? s0 + A.K : S;
        // recompute the forward inputs P_n, n in [s0, s1), from the checkpoint
        rc = load_ck(x, sg);
        if (rc) return rc;
        HZO_HIP(hipMemcpyAsync(A.seg, cur_P(x), ts * 6 * N, hipMemcpyDeviceToDevice, x->stream));
        for (int n = s0; n < s1 - 1; ++n) {
            rc = fwd_step(x, A.Ez[n]);
            if (rc) return rc;
            HZO_HIP(hipMemcpyAsync((char*)A.seg + ts * 6 * N * (n + 1 - s0), cur_P(x), ts * 6 * N,
                                   hipMemcpyDeviceToDevice, x->stream));
        }
        for (int n = s1 - 1; n >= s0; --n) {
            kb.P = (const char*)A.seg + ts * 6 * N * (n - s0);
            hzo::launch_adj_tdgl(x->stream, x->g, x->c, kb);
            hzo::launch_adj_gradacc(x->stream, x->g, A.gpart, A.grad);
            HZO_HIP(hipGetLastError());
            if (n == 0) break;  // P_0 and phi_0 are inputs: no adjoint needed
            if (dep) {
                hzo_pcg_stats st;
                if (x->fp32) {
                    hzo::Fp32Bufs b32 = bufs32(x);
                    b32.xhi = A.mu;
                    b32.xlo = A.mulo;
                    rc = pcg_solve(x, bufs(x), b32, ca, init, x->b64, A.last_iters, &st);
                } else {
                    hzo::Fp64Bufs b = bufs(x);
                    b.phi = A.mu;
                    rc = pcg_solve(x, b, bufs32(x), ca, init, A.rhs, A.last_iters, &st);
                }
                if (rc) return rc;
                iters += st.iters;
            }
            hzo::launch_adj_divt(x->stream, x->g, x->c, x->fp32, A.w, A.mu, A.mulo, A.a,
                                 dL_dm[n - 1] / (double)N, dep);
            HZO_HIP(hipGetLastError());
        }
    }
    rc = load_ck(x, A.nck - 1);  // leave the context at the recorded final state
    if (rc) return rc;
    HZO_HIP(hipMemcpyAsync(grad, A.grad, sizeof(double) * 12, hipMemcpyDeviceToHost, x->stream));
    HZO_HIP(hipStreamSynchronize(x->stream));
    if (adj_pcg_iters) *adj_pcg_iters = iters;
    return HZO_OK;
}

/* ---------------------------------------------------------------- dumps and restart */

// save_snapshot (generate.cu:266-303): Pm and Ps as interleaved xyz f64
// dumps scaled by P0.
int hzo_save_snapshot(hzo_ctx* x, const char* pm_path, const char* ps_path, double P0_scale) {
    int rc = check_ready(x);
    if (rc) return rc;
    if (!pm_path || !ps_path) return set_error(HZO_ERR_ARG, "hzo_save_snapshot: null path");
    const size_t N = (size_t)x->g.N;
    std::vector<double> P(6 * N);
    if ((rc = hzo_get_state(x, P.data(), nullptr, nullptr)) != HZO_OK) return rc;
    const double* p = P.data();
    if ((rc = hzo_write_vector_field(pm_path, (long long)N, p, p + N, p + 2 * N, P0_scale)) != HZO_OK) return rc;
    return hzo_write_vector_field(ps_path, (long long)N, p + 3 * N, p + 4 * N, p + 5 * N, P0_scale);
}

// write_checkpoint (generate.cu:90-168): the current P and E = -grad phi.
int hzo_checkpoint_save(hzo_ctx* x, const char* path, int peloop, int count, double Ez_ext, double incEz) {
    int rc = check_ready(x);
    if (rc) return rc;
    const size_t N = (size_t)x->g.N;
    std::vector<double> P(6 * N), E(3 * N);
    if ((rc = hzo_get_state(x, P.data(), nullptr, E.data())) != HZO_OK) return rc;
    return hzo_write_checkpoint(path, x->g.nx, x->g.ny, x->g.nz, peloop, count, Ez_ext, incEz, P.data(), E.data());
}

// read_checkpoint (generate.cu:170-264) as a restart: P from the file; the
// file holds E but no potential, so phi is re-solved from P (linear initial
// guess, then the depolarisation solve when it is on), which reproduces the
// stored E to the solver tolerance.  E_file (3N, may be NULL) receives the
// stored field.
int hzo_checkpoint_load(hzo_ctx* x, const char* path, int* peloop, int* count, double* Ez_ext, double* incEz,
                        double* E_file) {
    if (!x) return set_error(HZO_ERR_ARG, "null context");
    const size_t N = (size_t)x->g.N;
    std::vector<double> P(6 * N);
    int rc = hzo_read_checkpoint(path, x->g.nx, x->g.ny, x->g.nz, peloop, count, Ez_ext, incEz, P.data(), E_file);
    if (rc) return rc;
    if ((rc = hzo_set_state(x, P.data(), nullptr)) != HZO_OK) return rc;
    if (x->cfg.dep_flag) return hzo_depol_solve(x, nullptr);
    return HZO_OK;
}

/* ---------------------------------------------------------------- batched films */

// Films are packed into groups whose solver working set fits a cache budget;
// each group's kernels run as one launch over a (film, tile) table, and the
// groups run one after another, each through all steps of a call, so that a
// group's CG vectors stay in the 256 MB last-level cache across iterations.
// Host threads (each with its own stream) a batch may run its streaming
// groups on side by side.
constexpr int BATCH_MAX_LANES = 4;

struct hzo_batch {
    struct Group {
        std::vector<int> idx;  // batch indices of the films
        int single = -1;       // a group of one film runs that film's own kernels and tile depth
        std::vector<hzo::BatchFilm> hf;
        std::vector<double*> part, tdgl_part, diag_part;
        hzo::BatchFilm* dfilms = nullptr;
        int2* dtiles = nullptr;
        int ntiles = 0, pm = 1, ps = 0;
        hipGraphExec_t gexec = nullptr;
        int g_chunk = -1;
        // a streaming group's whole step (TDGL, setup, the first CG chunks)
        // captured once per P parity and number of chunks, then replayed
        struct StepGraph {
            int par, nchunks, chunk;
            hipGraphExec_t exec;
        };
        std::vector<StepGraph> sgraphs;
        int need = 16;  // CG launches the group's last solve needed
        bool mirrored = false;  // some film has Neumann electrodes
        // a resident group: films of one depth whose tiles fit the device
        // together, each stepped by its own TDGL and setup kernels and solved
        // by one cooperative launch for all of them
        bool res = false;
        int res_nz = 0, res_tiles = 0;
        hzo::ResFilm* dres = nullptr;  // per-film descriptors (device)
        unsigned* rsync = nullptr;     // per-film barrier words, then one timeout flag
    };
    // A lane: a stream and the staging buffers with which one host thread
    // runs its share of the streaming groups; lane 0 also runs the resident
    // groups and the batch's own copies.
    struct Lane {
        hipStream_t stream = nullptr;
        double* dez = nullptr;
        double* hez = nullptr;  // pinned
        int* dinfo = nullptr;   // BATCH_INFO ints per film: done, iters, status, refinements, xadds
        int* hinfo = nullptr;   // pinned
        double* dm = nullptr;
        double* hm = nullptr;   // pinned
        long long launches = 0, unconverged = 0;  // of the lane's groups in the current call
    };
    std::vector<hzo_ctx*> films;
    int n = 0;
    int device = 0;
    hipStream_t stream = nullptr;  // lanes[0].stream
    int kz = 2;
    double budget = 0.0;  // bytes of CG working set per group (0: one group)
    std::vector<Group> groups;
    std::vector<Lane> lanes;  // BATCH_MAX_LANES of them
    int nlanes = 2;           // lanes that run streaming groups side by side
    int chunk = 16;           // CG launches per graph replay (even)
    bool graphs = true;
    long long launches_total = 0;  // CG launches of all batched runs so far
    long long unconverged = 0;     // film solves that hit the launch cap
};

namespace {

// Bytes per voxel the single-reduction CG keeps live across iterations:
// r, w, s (two buffers each), p, x_lo (float) and the stencil code (uint16).
constexpr double CG_SET_BYTES = 8 * 4 + 2;

void group_free(hzo_batch::Group& G) {
    if (G.gexec) (void)hipGraphExecDestroy(G.gexec);
    G.gexec = nullptr;
    for (hzo_batch::Group::StepGraph& sg : G.sgraphs) (void)hipGraphExecDestroy(sg.exec);
    G.sgraphs.clear();
    for (double* p : G.part) (void)hipFree(p);
    for (double* p : G.tdgl_part) (void)hipFree(p);
    for (double* p : G.diag_part) (void)hipFree(p);
    G.part.clear();
    G.tdgl_part.clear();
    G.diag_part.clear();
    (void)hipFree(G.dfilms);
    (void)hipFree(G.dtiles);
    (void)hipFree(G.dres);
    (void)hipFree(G.rsync);
    G.dfilms = nullptr;
    G.dtiles = nullptr;
    G.dres = nullptr;
    G.rsync = nullptr;
}

void batch_free_groups(hzo_batch* B) {
    for (hzo_batch::Group& G : B->groups) group_free(G);
    B->groups.clear();
}

void batch_free(hzo_batch* B) {
    batch_free_groups(B);
    for (hzo_batch::Lane& Ln : B->lanes) {
        (void)hipFree(Ln.dez);
        (void)hipFree(Ln.dinfo);
        (void)hipFree(Ln.dm);
        if (Ln.hez) (void)hipHostFree(Ln.hez);
        if (Ln.hinfo) (void)hipHostFree(Ln.hinfo);
        if (Ln.hm) (void)hipHostFree(Ln.hm);
        if (Ln.stream) (void)hipStreamDestroy(Ln.stream);
    }
    B->lanes.clear();
    B->stream = nullptr;
}

hzo::Fp32Bufs film_bufs32(hzo_ctx* x, int cur) {
    const int keep = x->cur;
    x->cur = cur;
    hzo::Fp32Bufs b = bufs32(x);
    x->cur = keep;
    return b;
}

// Descriptor table, tile table and group-owned partial arrays of one group
// at tile depth kz.  P parity of a film at step s of a run is (cur + s) & 1.
// A group of one film needs none of them: it runs the single-film kernels.
int group_tables(hzo_batch* B, hzo_batch::Group& G, int kz) {
    const int n = (int)G.idx.size();
    if (G.res) {
        std::vector<hzo::ResFilm> rf(n);
        const size_t words = (size_t)n * hzo::RES_SYNC_WORDS + 64;
        HZO_HIP(hipMalloc((void**)&G.rsync, sizeof(unsigned) * words));
        int blk0 = 0;
        G.mirrored = false;
        for (int j = 0; j < n; ++j) {
            hzo_ctx* x = B->films[G.idx[j]];
            hzo::Fp32Bufs b32 = bufs32(x);
            b32.r = x->r32;
            rf[j].b = res_bufs(x, cg_bufs(x, b32));
            rf[j].b.sync = G.rsync + (size_t)j * hzo::RES_SYNC_WORDS;
            rf[j].b.err = reinterpret_cast<int*>(G.rsync + (size_t)n * hzo::RES_SYNC_WORDS);
            rf[j].c = x->c;
            rf[j].gxt = x->res_gxt;
            rf[j].G = x->res_G;
            rf[j].blk0 = blk0;
            rf[j].pad = 0;
            blk0 += x->res_G;
            G.mirrored = G.mirrored || x->c.zbc != 0;
        }
        G.res_nz = B->films[G.idx[0]]->cfg.nz;
        G.res_tiles = blk0;
        HZO_HIP(hipMalloc((void**)&G.dres, sizeof(hzo::ResFilm) * n));
        HZO_HIP(hipMemcpy(G.dres, rf.data(), sizeof(hzo::ResFilm) * n, hipMemcpyHostToDevice));
        return HZO_OK;
    }
    if (n == 1) {
        G.single = G.idx[0];
        G.ntiles = B->films[G.single]->nb;
        return HZO_OK;
    }
    int maxnb = 0;
    std::vector<int> nb(n);
    G.mirrored = false;
    for (int j = 0; j < n; ++j) {
        G.mirrored = G.mirrored || B->films[G.idx[j]]->c.zbc != 0;
        hzo::Grid g = B->films[G.idx[j]]->g;
        g.kz = kz;
        nb[j] = hzo::blocks_for(g);
        maxnb = nb[j] > maxnb ? nb[j] : maxnb;
    }
    G.pm = maxnb <= hzo::NT ? 1 : maxnb <= 2 * hzo::NT ? 2 : maxnb <= 4 * hzo::NT ? 4 : 0;
    G.ps = G.pm ? G.pm * hzo::NT : maxnb;
    G.hf.assign(n, hzo::BatchFilm());
    std::vector<int2> tiles;
    for (int j = 0; j < n; ++j) {
        hzo_ctx* x = B->films[G.idx[j]];
        double *pa = nullptr, *tp = nullptr, *dp = nullptr;
        const size_t pbytes = sizeof(double) * 2 * hzo::RED_MAXQ * (size_t)G.ps;
        if (hipMalloc((void**)&pa, pbytes) != hipSuccess || hipMemset(pa, 0, pbytes) != hipSuccess ||
            hipMalloc((void**)&tp, sizeof(double) * hzo::TDGL_Q * nb[j]) != hipSuccess ||
            hipMalloc((void**)&dp, sizeof(double) * hzo::DIAG_Q * nb[j]) != hipSuccess) {
            (void)hipFree(pa);
            (void)hipFree(tp);
            return set_error(HZO_ERR_NOMEM, "hzo_batch: partial-sum allocation failed");
        }
        G.part.push_back(pa);
        G.tdgl_part.push_back(tp);
        G.diag_part.push_back(dp);
        hzo::BatchFilm& F = G.hf[j];
        for (int q = 0; q < 2; ++q) {
            F.b[q] = film_bufs32(x, (x->cur + q) & 1);
            F.b[q].part[0] = pa;
            F.b[q].part[1] = pa + (size_t)hzo::RED_MAXQ * G.ps;
            F.b[q].pin[0] = F.b[q].part[0];
            F.b[q].pin[1] = F.b[q].part[1];
            F.b[q].nb_in = nb[j];
            F.b[q].ps_in = G.ps;
            F.b[q].tdgl_part = tp;
        }
        F.cg.va = x->cg_arena[0];
        F.cg.vb = x->cg_arena[1];
        F.cg.xlo = x->xlo;
        F.cg.xhi = x->phi;
        F.cg.b = x->b64;
        F.cg.sc = x->sc;
        F.cg.lut = x->lut;
        F.cg.lut32 = x->lut32;
        F.cg.part[0] = pa;
        F.cg.part[1] = pa + (size_t)hzo::RED_MAXQ * G.ps;
        F.cg.pin[0] = F.cg.part[0];
        F.cg.pin[1] = F.cg.part[1];
        F.cg.nb_in = nb[j];
        F.cg.ps_in = G.ps;
        F.cg.st = x->st;
        F.c = x->c;
        F.c.kz = kz;
        F.g = x->g;
        F.g.kz = kz;
        F.init = pcg_init(x);
        F.gx = (x->g.nx + hzo::BX - 1) / hzo::BX;
        F.gy = (x->g.ny + hzo::BY - 1) / hzo::BY;
        F.gz = (x->g.nz + kz - 1) / kz;
        F.nb = nb[j];
        F.ps = G.ps;
        F.Ez = 0.f;
        F.diag_part = dp;
        for (int t = 0; t < nb[j]; ++t) tiles.push_back(make_int2(j, t));
    }
    G.ntiles = (int)tiles.size();
    HZO_HIP(hipMalloc((void**)&G.dfilms, sizeof(hzo::BatchFilm) * n));
    HZO_HIP(hipMalloc((void**)&G.dtiles, sizeof(int2) * tiles.size()));
    HZO_HIP(hipMemcpy(G.dfilms, G.hf.data(), sizeof(hzo::BatchFilm) * n, hipMemcpyHostToDevice));
    HZO_HIP(hipMemcpy(G.dtiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice));
    return HZO_OK;
}

// Packs the films into groups by first-fit decreasing on the CG working set
// (a film larger than the budget forms a group of its own) and builds each
// group's tables at tile depth kz.
int batch_tables(hzo_batch* B, int kz, double budget) {
    batch_free_groups(B);
    B->kz = kz;
    B->budget = budget;
    // Films that take the resident solve on their own are packed, by depth
    // and first-fit decreasing on tiles, into groups that fit the device at
    // once (the Neumann build's capacity when a group may hold such a film);
    // the others go by CG working set as below.
    std::vector<int> order;
    std::vector<std::vector<int>> rgroups;
    std::vector<int> rnz, rload;
    {
        std::vector<int> rfilms;
        for (int f = 0; f < B->n; ++f) {
            if (B->films[f]->cg_res) rfilms.push_back(f);
            else order.push_back(f);
        }
        std::stable_sort(rfilms.begin(), rfilms.end(),
                         [&](int a, int b) { return B->films[a]->res_G > B->films[b]->res_G; });
        for (int f : rfilms) {
            hzo_ctx* x = B->films[f];
            const int nz = x->cfg.nz;
            const int cap = std::min(hzo::cgres_capacity(B->device, nz, false), hzo::cgres_capacity(B->device, nz, true));
            size_t g = 0;
            while (g < rgroups.size() && (rnz[g] != nz || rload[g] + x->res_G > cap)) ++g;
            if (g == rgroups.size()) {
                rgroups.emplace_back();
                rnz.push_back(nz);
                rload.push_back(0);
            }
            rgroups[g].push_back(f);
            rload[g] += x->res_G;
        }
    }
    for (std::vector<int>& idx : rgroups) {
        B->groups.emplace_back();
        std::sort(idx.begin(), idx.end());
        B->groups.back().idx = idx;
        B->groups.back().res = true;
    }
    const size_t nres = B->groups.size();
    auto bytes = [&](int f) { return CG_SET_BYTES * (double)B->films[f]->g.N; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return bytes(a) > bytes(b); });
    std::vector<double> load;
    for (int f : order) {
        size_t g = 0;
        if (budget > 0.0)
            while (g < load.size() && load[g] + bytes(f) > budget) ++g;
        if (g == load.size()) {
            load.push_back(0.0);
            B->groups.emplace_back();
        }
        load[g] += bytes(f);
        B->groups[nres + g].idx.push_back(f);
    }
    for (hzo_batch::Group& G : B->groups) {
        std::sort(G.idx.begin(), G.idx.end());
        int rc = group_tables(B, G, kz);
        if (rc) return rc;
    }
    return HZO_OK;
}

// Re-reads each film's P parity into the group's descriptor table (contexts
// may have been stepped on their own since the last batched run).
int group_sync_parity(hzo_batch* B, hzo_batch::Group& G, hipStream_t st) {
    if (G.single >= 0 || G.res) return HZO_OK;
    for (size_t j = 0; j < G.idx.size(); ++j) {
        hzo_ctx* x = B->films[G.idx[j]];
        hzo::BatchFilm& F = G.hf[j];
        for (int q = 0; q < 2; ++q) {
            const hzo::Fp32Bufs b = film_bufs32(x, (x->cur + q) & 1);
            F.b[q].Pold = b.Pold;
            F.b[q].Pnew = b.Pnew;
            F.b[q].Pcur = b.Pcur;
        }
    }
    // ordered on the lane's stream (G.hf lives as long as the group)
    HZO_HIP(hipMemcpyAsync(G.dfilms, G.hf.data(), sizeof(hzo::BatchFilm) * G.idx.size(), hipMemcpyHostToDevice, st));
    return HZO_OK;
}

// K launches of the single-reduction CG for the group (launch index 1..K).
void group_cg_launches(hzo_batch* B, hzo_batch::Group& G, int K, hipStream_t st) {
    if (G.single >= 0) {
        hzo_ctx* x = B->films[G.single];
        const hzo::CG32Bufs cb = cg_bufs(x, bufs32(x));
        for (int t = 1; t <= K; ++t) {
            hzo::launch_cg32(st, x->g, x->c, cb, t, x->cg_two, x->cg_nt);
            fold_after(x, st, t);
        }
    } else {
        for (int t = 1; t <= K; ++t)
            hzo::launch_cg32_batch(st, G.ntiles, B->kz, G.pm, G.mirrored, G.dfilms, G.dtiles, t);
    }
}

int group_cg_chunk(hzo_batch* B, hzo_batch::Group& G, hipStream_t st) {
    const int K = B->chunk;
    if (B->graphs) {
        if (!G.gexec || G.g_chunk != K) {
            if (G.gexec) (void)hipGraphExecDestroy(G.gexec);
            G.gexec = nullptr;
            hipGraph_t g;
            HZO_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            group_cg_launches(B, G, K, st);
            HZO_HIP(hipStreamEndCapture(st, &g));
            const hipError_t e = hipGraphInstantiate(&G.gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HZO_HIP(e);
            G.g_chunk = K;
        }
        HZO_HIP(hipGraphLaunch(G.gexec, st));
    } else {
        group_cg_launches(B, G, K, st);
        HZO_HIP(hipGetLastError());
    }
    return HZO_OK;
}

// Solver statistics of the group's films after a chunk (launch index K) into
// hinfo[BATCH_INFO*j..]: done, iterations, status, residual replacements,
// launches that only added a deferred x_lo term.
int group_poll(hzo_batch* B, hzo_batch::Group& G, hzo_batch::Lane& Ln) {
    const int n = (int)G.idx.size();
    if (G.single >= 0) {
        hzo_ctx* x = B->films[G.single];
        HZO_HIP(hipMemcpyAsync(x->st_host, x->st + (B->chunk & 1), sizeof(hzo::PcgState), hipMemcpyDeviceToHost,
                               Ln.stream));
        HZO_HIP(hipStreamSynchronize(Ln.stream));
        Ln.hinfo[0] = x->st_host->done;
        Ln.hinfo[1] = x->st_host->iters;
        Ln.hinfo[2] = x->st_host->status;
        Ln.hinfo[3] = x->st_host->refinements;
        Ln.hinfo[4] = x->st_host->xadds;
        return HZO_OK;
    }
    hzo::launch_batch_done(Ln.stream, G.dfilms, n, B->chunk, Ln.dinfo);
    HZO_HIP(hipMemcpyAsync(Ln.hinfo, Ln.dinfo, sizeof(int) * hzo::BATCH_INFO * n, hipMemcpyDeviceToHost, Ln.stream));
    HZO_HIP(hipStreamSynchronize(Ln.stream));
    return HZO_OK;
}

// One step of one group: TDGL of every film, PCG setup, then chunks of
// single-reduction CG launches until every film's solve has finished.  The
// launch index t restarts at 1 in each chunk (chunks are even, so the
// control-block parity continues).  hinfo[BATCH_INFO*j..] receives film j's solver
// statistics.
int group_step_res(hzo_batch* B, hzo_batch::Group& G, hzo_batch::Lane& Ln, const double* Ez, int s, bool dep,
                   int& launches) {
    const int n = (int)G.idx.size();
    for (int f : G.idx) {
        hzo_ctx* x = B->films[f];
        hzo::launch_tdgl_f32(Ln.stream, x->g, x->c, Ez[(size_t)s * B->n + f], bufs32(x));
        x->cur ^= 1;
    }
    HZO_HIP(hipGetLastError());
    launches = 0;
    if (!dep) {
        memset(Ln.hinfo, 0, sizeof(int) * hzo::BATCH_INFO * n);
        return HZO_OK;
    }
    for (int f : G.idx) {
        hzo_ctx* x = B->films[f];
        hzo::launch_pcg_setup_f32(Ln.stream, x->g, x->c, bufs32(x), pcg_init(x), nullptr);
        fold_after(x, Ln.stream, 0);
    }
    // barrier words and the timeout flag start at zero (a fill kernel, not a
    // memset node, in case the caller captures the stream)
    hzo::launch_fill_u32(Ln.stream, G.rsync, n * hzo::RES_SYNC_WORDS + 64, 0u);
    HZO_HIP(hipGetLastError());
    const hipError_t le = hzo::launch_cgres_films(Ln.stream, G.dres, n, G.res_tiles, G.res_nz, G.mirrored);
    if (le != hipSuccess) return set_error(HZO_ERR_HIP, "resident CG launch failed: %s", hipGetErrorString(le));
    launches = 1;
    hzo_ctx* x0 = B->films[G.idx[0]];
    for (int f : G.idx) {
        hzo_ctx* x = B->films[f];
        HZO_HIP(hipMemcpyAsync(x->st_host, x->st, sizeof(hzo::PcgState), hipMemcpyDeviceToHost, Ln.stream));
    }
    HZO_HIP(hipMemcpyAsync(x0->res_err_host, G.rsync + (size_t)n * hzo::RES_SYNC_WORDS, sizeof(int),
                           hipMemcpyDeviceToHost, Ln.stream));
    HZO_HIP(hipStreamSynchronize(Ln.stream));
    if (*x0->res_err_host) return set_error(HZO_ERR_HIP, "resident CG: a grid barrier timed out");
    for (int j = 0; j < n; ++j) {
        hzo_ctx* x = B->films[G.idx[j]];
        res_after(x);
        const hzo::PcgState& st = *x->st_host;
        int* h = Ln.hinfo + hzo::BATCH_INFO * j;
        h[0] = st.done;
        h[1] = st.iters;
        h[2] = st.status;
        h[3] = st.refinements;
        h[4] = st.xadds;
    }
    return HZO_OK;
}

// The launches of one step of a streaming group whose films' P parity is
// par: TDGL of every film (the applied fields read from Ln.dez), then, with
// depolarisation, the PCG setup and nchunks chunks of single-reduction CG
// launches.  Captured into a graph or issued as they are; a single film's
// context is left at its new parity either way.
void group_step_launches(hzo_batch* B, hzo_batch::Group& G, hzo_batch::Lane& Ln, int par, bool dep, int nchunks) {
    const int n = (int)G.idx.size();
    if (G.single >= 0) {
        hzo_ctx* x = B->films[G.single];
        hzo::launch_tdgl_f32_dev(Ln.stream, x->g, x->c, Ln.dez, film_bufs32(x, par));
        x->cur = par ^ 1;
        if (dep) {
            hzo::launch_pcg_setup_f32(Ln.stream, x->g, x->c, bufs32(x), pcg_init(x), nullptr);
            fold_after(x, Ln.stream, 0);
        }
    } else {
        hzo::launch_batch_set_ez(Ln.stream, G.dfilms, Ln.dez, n);
        hzo::launch_tdgl_f32_batch(Ln.stream, G.ntiles, G.dfilms, G.dtiles, par);
        if (dep) hzo::launch_setup_f32_batch(Ln.stream, G.ntiles, G.dfilms, G.dtiles, par ^ 1);
    }
    for (int q = 0; q < nchunks; ++q) group_cg_launches(B, G, B->chunk, Ln.stream);
}

// The executable graph of a streaming group's step at parity par with
// nchunks CG chunks, captured on first use (at most eight kept per group).
int group_step_graph(hzo_batch* B, hzo_batch::Group& G, hzo_batch::Lane& Ln, int par, bool dep, int nchunks,
                     hipGraphExec_t& out) {
    for (const hzo_batch::Group::StepGraph& sg : G.sgraphs)
        if (sg.par == par && sg.nchunks == nchunks && sg.chunk == B->chunk) {
            out = sg.exec;
            return HZO_OK;
        }
    if (G.sgraphs.size() >= 8) {
        (void)hipGraphExecDestroy(G.sgraphs.front().exec);
        G.sgraphs.erase(G.sgraphs.begin());
    }
    hipGraph_t g;
    HZO_HIP(hipStreamBeginCapture(Ln.stream, hipStreamCaptureModeThreadLocal));
    group_step_launches(B, G, Ln, par, dep, nchunks);
    HZO_HIP(hipStreamEndCapture(Ln.stream, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HZO_HIP(e);
    G.sgraphs.push_back({par, nchunks, B->chunk, ex});
    out = ex;
    return HZO_OK;
}

// One step of a streaming group: the films' applied fields to the device,
// then the step's launches (TDGL, setup and as many CG chunks as the
// slowest film's last solve needed), replayed from a graph when graphs are
// on; then further chunks until every film's solve has finished.  The
// launch index t restarts at 1 in each chunk (chunks are even, so the
// control-block parity continues).  hinfo[BATCH_INFO*j..] receives film j's
// solver statistics.
int group_step(hzo_batch* B, hzo_batch::Group& G, hzo_batch::Lane& Ln, const double* Ez, int s, bool dep,
               int& launches) {
    if (G.res) return group_step_res(B, G, Ln, Ez, s, dep, launches);
    const int n = (int)G.idx.size();
    HZO_HIP(hipStreamSynchronize(Ln.stream));  // the previous step's copy out of hez has run
    for (int j = 0; j < n; ++j) Ln.hez[j] = Ez[(size_t)s * B->n + G.idx[j]];
    HZO_HIP(hipMemcpyAsync(Ln.dez, Ln.hez, sizeof(double) * n, hipMemcpyHostToDevice, Ln.stream));
    const int par = G.single >= 0 ? B->films[G.single]->cur : (s & 1);
    const int nchunks = dep ? (G.need + B->chunk - 1) / B->chunk : 0;
    // The step graph holds the step's first chunk; the others replay the
    // chunk graph one by one, so the GPU starts on the step while the host
    // still submits its later chunks (one graph of all chunks measured 3 %
    // slower on the 64-film workload).
    const int first = nchunks < 1 ? nchunks : 1;
    if (B->graphs) {
        hipGraphExec_t ex = nullptr;
        int rc = group_step_graph(B, G, Ln, par, dep, first, ex);
        if (rc) return rc;
        HZO_HIP(hipGraphLaunch(ex, Ln.stream));
    } else {
        group_step_launches(B, G, Ln, par, dep, first);
        HZO_HIP(hipGetLastError());
    }
    if (G.single >= 0) B->films[G.single]->cur = par ^ 1;
    for (int q = first; q < nchunks; ++q) {
        int rc = group_cg_chunk(B, G, Ln.stream);
        if (rc) return rc;
    }
    launches = nchunks * B->chunk;
    if (!dep) {
        memset(Ln.hinfo, 0, sizeof(int) * hzo::BATCH_INFO * n);
        return HZO_OK;
    }
    // the most launches any film may need: four per iteration
    int maxl = 0;
    for (int f : G.idx) {
        const int m = 4 * B->films[f]->cfg.max_iter + 8;
        maxl = m > maxl ? m : maxl;
    }
    while (true) {
        int rc = group_poll(B, G, Ln);
        if (rc) return rc;
        bool all = true;
        for (int j = 0; j < n; ++j) all = all && Ln.hinfo[hzo::BATCH_INFO * j];
        if (all || launches >= maxl) break;
        rc = group_cg_chunk(B, G, Ln.stream);
        if (rc) return rc;
        launches += B->chunk;
    }
    // launches the slowest film needed: setup hand-over, one per iteration,
    // two per residual replacement, one per deferred-x_lo addition
    int need = 2;
    for (int j = 0; j < n; ++j) {
        const int* h = Ln.hinfo + hzo::BATCH_INFO * j;
        const int m = h[1] + 2 * h[3] + h[4] + 2;
        need = m > need ? m : need;
    }
    G.need = need;
    return HZO_OK;
}

// nsteps steps of one group; leaves each film's P parity advanced by nsteps.
int group_run(hzo_batch* B, hzo_batch::Group& G, hzo_batch::Lane& Ln, const double* Ez, int nsteps, int* iters,
              double* m_out) {
    int rc = group_sync_parity(B, G, Ln.stream);
    if (rc) return rc;
    const int n = (int)G.idx.size();
    const bool dep = B->films[0]->cfg.dep_flag != 0;
    for (int s = 0; s < nsteps; ++s) {
        int l = 0;
        rc = group_step(B, G, Ln, Ez, s, dep, l);
        if (rc) return rc;
        Ln.launches += l;
        for (int j = 0; j < n; ++j) {
            if (iters) iters[(size_t)s * B->n + G.idx[j]] = Ln.hinfo[hzo::BATCH_INFO * j + 1];
            Ln.unconverged += dep && !Ln.hinfo[hzo::BATCH_INFO * j];
        }
        if (m_out) {
            if (G.single >= 0 || G.res) {
                for (int j = 0; j < n; ++j) {
                    hzo_ctx* x = B->films[G.idx[j]];
                    hzo::launch_diag_f32(Ln.stream, x->g, x->P32[x->cur], x->cw, x->diag_part);
                    hzo::launch_sum_column(Ln.stream, x->g, x->diag_part, Ln.dm + j);
                }
            } else {
                hzo::launch_diag_f32_batch(Ln.stream, G.ntiles, G.dfilms, G.dtiles, (s + 1) & 1);
                hzo::launch_batch_sum_pmz(Ln.stream, G.dfilms, n, Ln.dm);
            }
            HZO_HIP(hipGetLastError());
            HZO_HIP(hipMemcpyAsync(Ln.hm, Ln.dm, sizeof(double) * n, hipMemcpyDeviceToHost, Ln.stream));
            HZO_HIP(hipStreamSynchronize(Ln.stream));
            for (int j = 0; j < n; ++j)
                m_out[(size_t)s * B->n + G.idx[j]] = Ln.hm[j] / (double)B->films[G.idx[j]]->g.N;
        }
    }
    HZO_HIP(hipStreamSynchronize(Ln.stream));
    if (G.single < 0 && !G.res)
        for (int f : G.idx) B->films[f]->cur = (B->films[f]->cur + nsteps) & 1;
    return HZO_OK;
}

// nsteps steps of every group.  The resident groups run first, one after
// another on lane 0: a cooperative launch needs every workgroup of its grid
// resident at once, so nothing else runs beside it.  The streaming groups
// are then shared out over the lanes, longest first to the least loaded
// lane (cost: voxels times the CG launches of the group's last step), and
// each lane runs its share on its own stream from its own host thread
// (lane 0 from the calling thread), so one group's launches fill the CUs
// that another's leave idle (the tail of a launch, the gaps between
// launches and while the host polls).  Groups share no buffers and each
// film's work stays in order on one stream, so every film ends in the same
// state whatever the number of lanes.
int run_groups(hzo_batch* B, const double* Ez, int nsteps, int* iters, double* m_out) {
    for (hzo_batch::Lane& Ln : B->lanes) Ln.launches = Ln.unconverged = 0;
    auto tally = [&]() {
        for (hzo_batch::Lane& Ln : B->lanes) {
            B->launches_total += Ln.launches;
            B->unconverged += Ln.unconverged;
            Ln.launches = Ln.unconverged = 0;
        }
    };
    std::vector<int> streaming;
    for (size_t g = 0; g < B->groups.size(); ++g) {
        hzo_batch::Group& G = B->groups[g];
        if (!G.res) {
            streaming.push_back((int)g);
            continue;
        }
        int rc = group_run(B, G, B->lanes[0], Ez, nsteps, iters, m_out);
        if (rc) {
            tally();
            return rc;
        }
    }
    const int nl = std::max(1, std::min(B->nlanes, (int)streaming.size()));
    std::vector<double> cost(B->groups.size(), 0.0), load(nl, 0.0);
    for (int g : streaming) {
        double v = 0.0;
        for (int f : B->groups[g].idx) v += (double)B->films[f]->g.N;
        cost[g] = v * (double)std::max(B->groups[g].need, 1);
    }
    std::stable_sort(streaming.begin(), streaming.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<std::vector<int>> share(nl);
    for (int g : streaming) {
        const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        share[l].push_back(g);
        load[l] += cost[g];
    }
    std::vector<int> rcs(nl, HZO_OK);
    std::vector<std::string> msgs(nl);
    auto run_lane = [&](int l) {
        for (int g : share[l]) {
            rcs[l] = group_run(B, B->groups[g], B->lanes[l], Ez, nsteps, iters, m_out);
            if (rcs[l]) {
                msgs[l] = hzo::last_error();
                return;
            }
        }
    };
    std::vector<std::thread> th;
    for (int l = 1; l < nl; ++l) {
        try {
            th.emplace_back([&, l]() {
                if (hipSetDevice(B->device) != hipSuccess) {
                    rcs[l] = HZO_ERR_HIP;
                    msgs[l] = "batch lane: hipSetDevice failed";
                    return;
                }
                run_lane(l);
            });
        } catch (...) {
            run_lane(l);  // no thread: the lane's share runs here, after the others have been issued
        }
    }
    run_lane(0);
    for (std::thread& t : th) t.join();
    tally();
    for (int l = 0; l < nl; ++l)
        if (rcs[l]) return set_error(rcs[l], "%s", msgs[l].c_str());
    return HZO_OK;
}

}  // namespace

int hzo_batch_create(hzo_ctx* const* films, int n, int tile_kz, hzo_batch** out) {
    if (!films || n < 1 || !out) return set_error(HZO_ERR_ARG, "hzo_batch_create: bad argument");
    if (tile_kz != 2 && tile_kz != 4 && tile_kz != 8)
        return set_error(HZO_ERR_ARG, "hzo_batch_create: tile_kz must be 2, 4 or 8");
    for (int f = 0; f < n; ++f) {
        hzo_ctx* x = films[f];
        int rc = check_ready(x);
        if (rc) return rc;
        if (!x->fp32 || !x->cfg.cg_fused)
            return set_error(HZO_ERR_ARG, "hzo_batch_create: films must be FP32-refined with the single-reduction CG");
        if (x->device != films[0]->device || x->cfg.dep_flag != films[0]->cfg.dep_flag)
            return set_error(HZO_ERR_ARG, "hzo_batch_create: films must share the device and dep_flag");
        for (int e = 0; e < f; ++e)
            if (films[e] == x) return set_error(HZO_ERR_ARG, "hzo_batch_create: film %d appears twice", f);
    }
    hzo_batch* B = new hzo_batch();
    B->films.assign(films, films + n);
    B->n = n;
    B->device = films[0]->device;
    auto fail = [&](int rc) {
        batch_free(B);
        delete B;
        return rc;
    };
    if (hipSetDevice(B->device) != hipSuccess)
        return fail(set_error(HZO_ERR_HIP, "hzo_batch_create: hipSetDevice failed"));
    B->lanes.resize(BATCH_MAX_LANES);
    for (hzo_batch::Lane& Ln : B->lanes) {
        if (hipStreamCreateWithFlags(&Ln.stream, hipStreamNonBlocking) != hipSuccess)
            return fail(set_error(HZO_ERR_HIP, "hzo_batch_create: stream creation failed"));
        if (hipMalloc((void**)&Ln.dez, sizeof(double) * n) != hipSuccess ||
            hipMalloc((void**)&Ln.dinfo, sizeof(int) * hzo::BATCH_INFO * n) != hipSuccess ||
            hipMalloc((void**)&Ln.dm, sizeof(double) * n) != hipSuccess ||
            hipHostMalloc((void**)&Ln.hez, sizeof(double) * n, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&Ln.hinfo, sizeof(int) * hzo::BATCH_INFO * n, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void**)&Ln.hm, sizeof(double) * n, hipHostMallocDefault) != hipSuccess)
            return fail(set_error(HZO_ERR_NOMEM, "hzo_batch_create: allocation failed"));
    }
    B->stream = B->lanes[0].stream;
    for (hzo_ctx* x : B->films)
        if (hipStreamSynchronize(x->stream) != hipSuccess)
            return fail(set_error(HZO_ERR_HIP, "hzo_batch_create: film stream synchronisation failed"));
    int rc = batch_tables(B, tile_kz, 0.0);
    if (rc) return fail(rc);
    *out = B;
    return HZO_OK;
}

void hzo_batch_destroy(hzo_batch* B) {
    if (!B) return;
    (void)hipSetDevice(B->device);
    for (hzo_batch::Lane& Ln : B->lanes) (void)hipStreamSynchronize(Ln.stream);
    batch_free(B);
    delete B;
}

// Graph replay is refused while a rocprofiler-sdk tool is attached to the
// process (rocprofv3 names its tool library in ROCP_TOOL_LIBRARIES): a
// kernel trace of the batch with replay on ended in a segmentation fault
// inside the runtime's graph launch, called from the tool's interception,
// while the same replays run clean without a tool.  Plain launches (use_graphs
// = 0) run the same kernels and are traced normally.
static int refuse_traced_replay(const hzo_batch* B, const char* fn) {
    const char* tool = std::getenv("ROCP_TOOL_LIBRARIES");
    if (B->graphs && tool && *tool)
        return set_error(HZO_ERR_ARG,
                         "%s: hipGraph replay is refused under a profiler tool (ROCP_TOOL_LIBRARIES is set); "
                         "call hzo_batch_config(batch, 0, chunk) for plain launches",
                         fn);
    return HZO_OK;
}

int hzo_batch_config(hzo_batch* B, int use_graphs, int chunk) {
    if (!B || chunk < 2 || (chunk & 1)) return set_error(HZO_ERR_ARG, "hzo_batch_config: chunk must be even and >= 2");
    B->graphs = use_graphs != 0;
    B->chunk = chunk;
    return HZO_OK;
}

int hzo_batch_set_lanes(hzo_batch* B, int lanes) {
    if (!B || lanes < 1 || lanes > BATCH_MAX_LANES)
        return set_error(HZO_ERR_ARG, "hzo_batch_set_lanes: lanes must be 1 .. %d", BATCH_MAX_LANES);
    B->nlanes = lanes;
    return HZO_OK;
}

int hzo_batch_tile_depth(hzo_batch* B) { return B ? B->kz : 0; }

int hzo_batch_film_group(const hzo_batch* B, int film, int out[3]) {
    if (!B || !out || film < 0 || film >= B->n) return set_error(HZO_ERR_ARG, "hzo_batch_film_group: bad argument");
    for (size_t g = 0; g < B->groups.size(); ++g) {
        const hzo_batch::Group& G = B->groups[g];
        for (int f : G.idx) {
            if (f != film) continue;
            out[0] = (int)g;
            out[1] = (int)G.idx.size();
            out[2] = G.res ? 2 : (G.single >= 0 ? 1 : 0);
            return HZO_OK;
        }
    }
    return set_error(HZO_ERR_STATE, "hzo_batch_film_group: film %d is in no group", film);
}

int hzo_batch_set_shape(hzo_batch* B, int tile_kz, double group_mib) {
    if (!B || (tile_kz != 2 && tile_kz != 4 && tile_kz != 8) || !(group_mib >= 0.0))
        return set_error(HZO_ERR_ARG, "hzo_batch_set_shape: tile_kz must be 2, 4 or 8, group_mib >= 0");
    HZO_HIP(hipSetDevice(B->device));
    HZO_HIP(hipStreamSynchronize(B->stream));
    return batch_tables(B, tile_kz, group_mib * 1048576.0);
}

int hzo_batch_shape(const hzo_batch* B, int out[3]) {
    if (!B || !out) return set_error(HZO_ERR_ARG, "hzo_batch_shape: null argument");
    out[0] = B->kz;
    out[1] = (int)(B->budget / 1048576.0 + 0.5);
    out[2] = (int)B->groups.size();
    return HZO_OK;
}

int hzo_batch_run(hzo_batch* B, const double* Ez, int nsteps, int* iters, double* m_out) {
    if (!B || nsteps < 0 || (nsteps > 0 && !Ez)) return set_error(HZO_ERR_ARG, "hzo_batch_run: bad argument");
    if (int rc = refuse_traced_replay(B, "hzo_batch_run")) return rc;
    HZO_HIP(hipSetDevice(B->device));
    for (hzo_ctx* x : B->films) HZO_HIP(hipStreamSynchronize(x->stream));
    return run_groups(B, Ez, nsteps, iters, m_out);
}

int hzo_batch_autotune(hzo_batch* B, const double* Ez, int reps, int* kz_out, double ms[12]) {
    if (!B || !Ez || reps < 1) return set_error(HZO_ERR_ARG, "hzo_batch_autotune: bad argument");
    if (int rc = refuse_traced_replay(B, "hzo_batch_autotune")) return rc;
    HZO_HIP(hipSetDevice(B->device));
    for (hzo_ctx* x : B->films) HZO_HIP(hipStreamSynchronize(x->stream));
    // A trial step reads P[cur] and writes only P[cur^1], the potential and
    // the solver's work vectors, so restoring the potential (x_hi, x_lo) and
    // the parity returns every film to its state before the trial.
    std::vector<double*> hi(B->n, nullptr);
    std::vector<float*> lo(B->n, nullptr);
    std::vector<int> cur0(B->n);
    for (int f = 0; f < B->n; ++f) cur0[f] = B->films[f]->cur;
    auto release = [&]() {
        for (int f = 0; f < B->n; ++f) {
            (void)hipFree(hi[f]);
            (void)hipFree(lo[f]);
        }
    };
    for (int f = 0; f < B->n; ++f) {
        const size_t N = (size_t)B->films[f]->g.N;
        if (hipMalloc((void**)&hi[f], sizeof(double) * N) != hipSuccess ||
            hipMalloc((void**)&lo[f], sizeof(float) * N) != hipSuccess) {
            release();
            return set_error(HZO_ERR_NOMEM, "hzo_batch_autotune: snapshot allocation failed");
        }
    }
    auto copy = [&](bool save) -> int {
        for (int f = 0; f < B->n; ++f) {
            hzo_ctx* x = B->films[f];
            const size_t N = (size_t)x->g.N;
            x->cur = cur0[f];
            HZO_HIP(hipMemcpyAsync(save ? hi[f] : x->phi, save ? x->phi : hi[f], sizeof(double) * N,
                                   hipMemcpyDeviceToDevice, B->stream));
            HZO_HIP(hipMemcpyAsync(save ? lo[f] : x->xlo, save ? x->xlo : lo[f], sizeof(float) * N,
                                   hipMemcpyDeviceToDevice, B->stream));
        }
        HZO_HIP(hipStreamSynchronize(B->stream));
        return HZO_OK;
    };
    // candidates: tile depth x group budget (MiB of CG working set; 0 = all
    // films in one group)
    const int kzs[3] = {2, 4, 8};
    const double mib[4] = {0.0, 32.0, 64.0, 128.0};
    double best_ms[12];
    for (double& v : best_ms) v = 1e300;
    int rc = copy(true);
    for (int c = 0; c < 12 && rc == HZO_OK; ++c) {
        if ((rc = batch_tables(B, kzs[c / 4], mib[c % 4] * 1048576.0)) != HZO_OK) break;
        for (int r = 0; r <= reps && rc == HZO_OK; ++r) {  // r = 0 warms up (code objects, graphs)
            if ((rc = copy(false)) != HZO_OK) break;
            const auto t0 = std::chrono::steady_clock::now();
            if ((rc = run_groups(B, Ez, 1, nullptr, nullptr)) != HZO_OK) break;
            const double t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (r > 0 && t < best_ms[c]) best_ms[c] = t;
        }
    }
    if (rc == HZO_OK) rc = copy(false);
    release();
    if (rc) return rc;
    int best = 0;
    for (int c = 1; c < 12; ++c)
        if (best_ms[c] < best_ms[best]) best = c;
    if ((rc = batch_tables(B, kzs[best / 4], mib[best % 4] * 1048576.0)) != HZO_OK) return rc;
    if (kz_out) *kz_out = kzs[best / 4];
    if (ms)
        for (int c = 0; c < 12; ++c) ms[c] = best_ms[c];
    return HZO_OK;
}

int hzo_batch_counters(const hzo_batch* B, long long out[2]) {
    if (!B || !out) return set_error(HZO_ERR_ARG, "hzo_batch_counters: null argument");
    out[0] = B->launches_total;
    out[1] = B->unconverged;
    return HZO_OK;
}

}  // extern "C"
