"""FP32-refined path against the FP64 oracle.

The FP32 form is not bit-comparable; it is held to (for each solver:
cg_fused 1 automatic, 0 two-kernel PCG, 3 single-reduction CG forming
s = A p and w = A u by stencils in the launch):
* the Poisson solve meets ||b - A phi|| <= tol ||b|| in FP64 arithmetic
  (the residual is recomputed by the oracle from the returned phi and P);
* after a P-E sweep, P and phi stay within stated relative errors of the FP64
  oracle run, and the remanent polarisation within 1e-4 relative.
"""
import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L
from helpers import film_inputs, oracle_film, rel_err
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu]", *a, flush=True)


@pytest.mark.parametrize("fused", [1, 0, 3])
def test_fp32_refined_solve_meets_fp64_tolerance(fused):
    inp = film_inputs(64, 48, 16, 12, fracs=(0.2, 0.6, 0.2), seed=6, jitter=0.2)
    _say(f"fp32 solve 64x48x16 fused={fused}")
    g = H.Film(inp, precision=L.PREC_FP32_REFINED, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30,
               refine_every=20, cg_fused=fused)
    o = oracle_film(inp, pcg_rule=1, tol_rel=1e-8, tol_abs=1e-30)
    Ez = inp.field_nd(H.triangle_wave(200, 3.0, 5))
    its = g.run(Ez)
    st = g.step(float(inp.field_nd(0.8)))
    P, phi, E = g.get_state()
    b = -O.divergence(o.c, P)
    b.reshape(inp.nz, -1)[[0, -1]] = 0.0
    r = b - O.apply_A(o.c, o.eps, phi)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    _say(f"iters {list(its)} + {st.iters}, refinements {st.refinements}, true rel residual {rel:.3e}")
    assert st.status == 0 and st.refinements >= 1
    assert rel <= 1.0e-8 * 1.05
    g.close()


@pytest.mark.parametrize("fused", [1, 0, 3])
def test_fp32_refined_tracks_fp64_oracle_over_a_sweep(fused):
    inp = film_inputs(32, 32, 10, 8, fracs=(0.2, 0.6, 0.2), seed=9, jitter=0.2)
    nsteps = 160
    Ez = inp.field_nd(H.triangle_wave(nsteps, 3.0, nsteps))
    _say(f"fp32 sweep 32x32x10 160 steps fused={fused}")
    g = H.Film(inp, precision=L.PREC_FP32_REFINED, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30,
               cg_fused=fused)
    o = oracle_film(inp, pcg_rule=1, tol_rel=1e-8, tol_abs=1e-30)
    P = inp.P0.copy()
    phi = np.zeros(inp.N)
    E3 = np.zeros(3 * inp.N)
    mz_g, mz_o = [], []
    N = inp.N
    for s in range(nsteps):
        g.step(Ez[s])
        mz_g.append(g.diagnostics()[0] / N)
        o.run(P, phi, E3, Ez[s:s + 1])
        mz_o.append(P[2 * N:3 * N].mean())
    Pg, phig, _ = g.get_state()
    mz_g, mz_o = np.array(mz_g), np.array(mz_o)
    eP, ephi = rel_err(Pg, P), rel_err(phig, phi)
    # remanent polarisation: <Pmz> where the descending branch crosses E = 0 (step T/2)
    k = nsteps // 2
    pr_g, pr_o = mz_g[k], mz_o[k]
    _say(f"rel err P {eP:.2e} phi {ephi:.2e}; Pr gpu {pr_g:.6e} oracle {pr_o:.6e}; "
         f"loop max dev {np.max(np.abs(mz_g - mz_o)):.2e}")
    assert eP < 1e-3 and ephi < 1e-3
    assert abs(pr_g - pr_o) <= 1e-4 * abs(pr_o)
    g.close()


@pytest.mark.parametrize("fused", [1, 0, 3])
def test_fp32_refined_neumann_electrodes(fused):
    """Neumann electrodes: the mirrored planes k = 0, nz - 1 make A = S A_s
    non-symmetric, and the FP32 solves weight those planes by 1/2 in their CG
    sums.  Without that weight the single-reduction CG stalls and then
    diverges on this film (phi ~ 1e18, then NaN).  Each solve must converge
    with a true FP64 residual within the tolerance, and the sweep must track
    the FP64 oracle (plain Jacobi PCG, solved 100x tighter)."""
    inp = film_inputs(48, 40, 12, 8, fracs=(0.2, 0.6, 0.2), seed=12, jitter=0.2)
    nsteps = 12
    Ez = inp.field_nd(H.triangle_wave(nsteps, 3.0, nsteps))
    _say(f"fp32 neumann 48x40x12 {nsteps} steps fused={fused}")
    g = H.Film(inp, precision=L.PREC_FP32_REFINED, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30,
               max_iter=3000, zbc=L.ZBC_NEUMANN, cg_fused=fused)
    o = oracle_film(inp, zbc=L.ZBC_NEUMANN, pcg_rule=1, tol_rel=1e-10, tol_abs=1e-30, max_iter=20000)
    P = inp.P0.copy()
    phi = np.zeros(inp.N)
    E3 = np.zeros(3 * inp.N)
    its = []
    for s in range(nsteps):
        st = g.step(Ez[s])
        its.append((st.iters, st.status, st.refinements))
        assert st.status == 0, its
        o.run(P, phi, E3, Ez[s:s + 1])
    Pg, phig, _ = g.get_state()
    ref = inp.nx // 2 + inp.nx * (inp.ny // 2 + inp.ny * (inp.nz // 2))  # gauge node, phi = 0
    b = -O.divergence(o.c, Pg)
    b[ref] = 0.0
    r = b - O.apply_A(o.c, o.eps, phig)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    eP, ephi = rel_err(Pg, P), rel_err(phig, phi)
    _say(f"(iters, status, refinements) {its}; true rel residual {rel:.3e}; rel err P {eP:.2e} phi {ephi:.2e}")
    assert phig[ref] == 0.0
    assert rel <= 1.0e-8 * 1.05
    assert max(i for i, _, _ in its) < 1500
    assert eP < 1e-5 and ephi < 1e-5
    g.close()


@pytest.mark.parametrize("kz,zbc", [(2, 0), (4, 0), (8, 0), (16, 0), (32, 0), (4, 1), (16, 1), (32, 1)])
def test_two_stencil_form_matches_the_carried_form(kz, zbc):
    """The single-reduction CG forming s = A p and w = A u in the launch is
    the same Krylov iteration as the form carrying w and s.  With Dirichlet
    electrodes the iteration counts agree to one iteration per solve and P,
    phi to rounding.  With Neumann electrodes (mirrored planes, gauge node,
    weighted inner products) the system is near-singular and the carried
    recurrence s = w + beta s drifts from A p, so the counts differ: every
    solve must converge, the two-stencil form (s = A p exactly) may not need
    more iterations than 2 % above the carried one, and P, phi agree to the
    solver tolerance.  The automatic choice takes the two-stencil form
    exactly at tile depths 16 and 32; depth 32 exists only in that form, so
    the carried reference runs at depth 16 there."""
    inp = film_inputs(72, 40, 36, 10, fracs=(0.2, 0.6, 0.2), seed=14, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(40, 3.0, 6))
    out = {}
    for fused in (2, 3, 1):
        kf = min(kz, 16) if fused == 2 else kz
        g = H.Film(inp, precision=L.PREC_FP32_REFINED, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30,
                   tile_kz=kf, cg_fused=fused, zbc=zbc)
        assert g.tile_depth == kf
        its = []
        for e in Ez:
            st = g.step(float(e))
            assert st.status == 0, (fused, st.status)
            its.append(st.iters)
        out[fused] = (np.array(its), *g.get_state()[:2], g.cg_form)
        g.close()
    _say(f"kz={kz} zbc={zbc} iterations carried {out[2][0].tolist()} two-stencil {out[3][0].tolist()}")
    assert out[2][3] == 1 and out[3][3] == 2 and out[1][3] == (2 if kz >= 16 else 1)
    if zbc == 0:
        assert np.max(np.abs(out[2][0] - out[3][0])) <= 1
        assert rel_err(out[3][1], out[2][1]) < 1e-6 and rel_err(out[3][2], out[2][2]) < 1e-6
    else:
        assert out[3][0].sum() <= 1.02 * out[2][0].sum()
        assert rel_err(out[3][1], out[2][1]) < 1e-5 and rel_err(out[3][2], out[2][2]) < 1e-5


@pytest.mark.parametrize("shape,form,kz", [
    ((256, 256, 32), 3, 8),    # 256 column tiles, all resident at once: the resident CG
    ((256, 256, 64), 2, 16),   # two rounds at depth 8: two-stencil form, depth 32 would give 256 workgroups
    ((384, 384, 48), 2, 16),   # depth 32 would pad 16 of 48 planes
    ((512, 512, 32), 2, 32),   # 1024 column tiles exceed one per CU
    ((512, 512, 64), 2, 32),
    ((64, 48, 16), 3, 2),
    ((96, 40, 12), 1, 2),      # depth 12 has no resident instance: carried form
])
def test_automatic_solver_form_and_tile_depth_by_film_size(shape, form, kz):
    """The automatic choice (cg_fused = 1) takes the form and depth measured
    fastest for the film's size (the resident CG wherever every 16 x 16
    column tile fits the device at once); FP64 and cg_fused = 0 run the
    two-kernel PCG (form 0).  Only the context is created: no film data is
    needed."""
    import ctypes as C
    nx, ny, nz = shape
    out = {}
    for prec, fused in ((L.PREC_FP32_REFINED, 1), (L.PREC_FP32_REFINED, 0), (L.PREC_FP64, 1)):
        cfg = L.default_config(nx=nx, ny=ny, nz=nz, precision=prec, cg_fused=fused)
        h = C.c_void_p()
        assert L.lib().hzo_create(C.byref(cfg), C.byref(h)) == 0, L.lib().hzo_last_error()
        out[(prec, fused)] = (L.lib().hzo_cg_form(h), L.lib().hzo_tile_depth(h))
        L.lib().hzo_destroy(h)
    _say(shape, out)
    assert out[(L.PREC_FP32_REFINED, 1)] == (form, kz)
    assert out[(L.PREC_FP32_REFINED, 0)][0] == 0 and out[(L.PREC_FP64, 1)][0] == 0


@pytest.mark.parametrize("shape,kz", [((40, 4, 6), 32), ((20, 12, 4), 8), ((70, 9, 5), 16)])
def test_two_stencil_form_on_films_smaller_than_a_tile(shape, kz):
    """Films narrower than the 64x8 tile, with fewer rows than it, or
    shallower than its depth: the two-voxel ring wraps more than once, whole
    tile rows lie beyond the film and most window planes are clamped.  The
    two-stencil form must still follow the carried form (depth 2)."""
    nx, ny, nz = shape
    inp = film_inputs(nx, ny, nz, 3, fracs=(0.2, 0.6, 0.2), seed=17, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(30, 3.0, 5))
    out = {}
    for fused, kf in ((2, 2), (3, kz)):
        g = H.Film(inp, precision=L.PREC_FP32_REFINED, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30,
                   tile_kz=kf, cg_fused=fused)
        its = [g.step(float(e)) for e in Ez]
        assert all(st.status == 0 for st in its)
        out[fused] = (np.array([st.iters for st in its]), *g.get_state()[:2])
        g.close()
    _say(shape, kz, out[2][0].tolist(), out[3][0].tolist())
    assert np.max(np.abs(out[2][0] - out[3][0])) <= 1
    assert rel_err(out[3][1], out[2][1]) < 1e-6 and rel_err(out[3][2], out[2][2]) < 1e-6


@pytest.mark.parametrize("max_iter", [3, 4, 7])
@pytest.mark.parametrize("fused", [2, 3])
def test_single_reduction_cg_stopped_early_keeps_every_update(max_iter, fused):
    """Iterations of the single-reduction CG leave x_lo += alpha p to the
    next one in turn; a solve that stops with such a term outstanding (odd
    max_iter here) adds it in one more launch.  Stopped after a few
    iterations, the FP32 potential must then match the FP64 PCG stopped at
    the same count: the iterates agree to FP32 rounding, while one dropped
    alpha p would leave an error of the size of an early CG correction."""
    inp = film_inputs(64, 48, 16, 12, fracs=(0.2, 0.6, 0.2), seed=21, jitter=0.2)
    e = float(inp.field_nd(1.5))
    out = {}
    for prec, fu in ((L.PREC_FP64, 0), (L.PREC_FP32_REFINED, fused)):
        g = H.Film(inp, precision=prec, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-12, tol_abs=1e-30,
                   max_iter=max_iter, refine_every=20, cg_fused=fu)
        st = g.step(e)
        out[prec] = (st, g.get_state()[1])
        g.close()
    s64, phi64 = out[L.PREC_FP64]
    s32, phi32 = out[L.PREC_FP32_REFINED]
    err = rel_err(phi32, phi64)
    _say(f"max_iter={max_iter} fused={fused}: iters {s64.iters}/{s32.iters}, status {s64.status}/{s32.status}, "
         f"phi rel err {err:.2e}")
    assert s32.iters == s64.iters
    assert s32.status == s64.status == 1  # HZO_PCG_MAXITER
    assert err < 1e-5
