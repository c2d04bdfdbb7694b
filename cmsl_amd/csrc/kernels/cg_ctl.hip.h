// cg_ctl.hip.h -- pieces shared by the single-reduction CG kernels of the
// FP32-refined form (kernels_cg32.hip: one launch per iteration;
// kernels_cgres.hip: one persistent launch per solve): the FP32 operator row
// from a stencil code and the loop decision after a completed reduction.
#pragma once
#include "../hzo_internal.h"
#include "common.hip.h"
#include "tile.hip.h"
#include <type_traits>

namespace hzo {

// One row of the operator (-(sum of face fluxes)/h^2) from the stencil code:
// the paired harmonic face means of the three axes come from the 64-entry
// HPair table (see load_hpair).  coef_of reads them, row_apply32 applies
// them, so a kernel that evaluates one row on two operands reads the
// coefficients once.
struct Coef3 {
    HPair<float> x, y, z;
};

__device__ __forceinline__ Coef3 coef_of(unsigned code, const HPair<float>* sHP) {
    const unsigned cb = (code & 3u) << 4;
    Coef3 h;
    h.x = sHP[cb | ((code >> 2) & 15u)];
    h.y = sHP[cb | ((code >> 6) & 15u)];
    h.z = sHP[cb | ((code >> 10) & 15u)];
    return h;
}

__device__ __forceinline__ float row_apply32(const Coef3& h, float ih2, float xc, float xe, float xw, float xn,
                                             float xs, float xu, float xd) {
    const float num = h.x.p * (xe - xc) - h.x.m * (xc - xw) + h.y.p * (xn - xc) - h.y.m * (xc - xs) +
                      h.z.p * (xu - xc) - h.z.m * (xc - xd);
    return -(num * ih2);
}

__device__ __forceinline__ float row_Ax32(unsigned code, const HPair<float>* sHP, float ih2, float xc, float xe,
                                          float xw, float xn, float xs, float xu, float xd) {
    return row_apply32(coef_of(code, sHP), ih2, xc, xe, xw, xn, xs, xu, xd);
}

// a is a multiple of b (a >= 0, b > 0), the replacement-period test of the
// single-reduction CG: a quotient from the hardware reciprocal, corrected by
// one step (exact while a and b stay below 2^22, where the reciprocal's
// error moves the quotient by at most one), instead of the integer-division
// sequence a run-time divisor costs.
__device__ __forceinline__ bool multiple_of(int a, int b) {
    if (a >= (1 << 22) || b >= (1 << 22)) return a % b == 0;
    int q = (int)((float)a * __builtin_amdgcn_rcpf((float)b));
    int r = a - q * b;
    if (r < 0) r += b;
    else if (r >= b) r -= b;
    return r == 0;
}

// The loop logic after a completed reduction (every lane of every
// workgroup computes the same values).  Returns the work mode: 0 ITER,
// 1 done, 2 RESID, 3 WCOMP.  Sets alpha/beta for ITER.
__device__ __forceinline__ int cg_decide(PcgState& s, const double (&tot)[3], double& alpha, double& beta) {
    const int prev = s.kind;
    if (prev == K_SETUP || prev == K_RESID) {
        const double g = tot[0], rr = tot[1];
        if (prev == K_SETUP) {
            s.rz0 = g;
            s.rr0 = rr;
            s.bb = tot[2];
        } else {
            s.refinements += 1;
        }
        s.rr = rr;
        if (!finite(g) || g < 0.0) {
            s.status = prev == K_SETUP ? HZO_PCG_BAD_INITIAL : HZO_PCG_BAD_RZ;
            s.done = 1;
            return 1;
        }
        judge(s, g, rr);
        return s.done ? 1 : 3;
    }
    const double g = tot[0], dl = tot[1], rr = tot[2];
    s.rr = rr;
    if (!finite(g) || g < 0.0) {
        s.status = HZO_PCG_BAD_RZ;
        s.done = 1;
        return 1;
    }
    // K_XADD: the last launch added a deferred x_lo term before a residual
    // replacement (its partials are zero; the replacement restates rr)
    if (prev == K_ITER || prev == K_XADD) {
        if (s.iters >= s.max_iter) {
            s.done = 1;
            s.status = HZO_PCG_MAXITER;
            return 1;
        }
        if (prev == K_XADD || !keep_going(s, g, rr) || (s.refine_every > 0 && multiple_of(s.iters, s.refine_every)))
            return 2;
    }
    // prev == K_WCOMP (fresh start when iters == 0) or K_ITER
    double den;
    if (prev == K_WCOMP && s.iters == 0) {
        beta = 0.0;
        den = dl;
    } else {
        beta = g / s.rz_old;
        den = dl - beta * g / s.alpha;
    }
    if (!finite(den) || !(den > 0.0)) {
        s.status = HZO_PCG_BREAKDOWN;
        s.done = 1;
        return 1;
    }
    alpha = g / den;
    s.alpha = alpha;
    s.beta = beta;
    s.denom = den;
    s.rz_old = g;
    return 0;
}

}  // namespace hzo
