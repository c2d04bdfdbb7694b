"""PCG failure statuses and electrode potentials on the GPU, against the
oracle.

The reference treats a numerical failure of its PCG as non-fatal: it prints,
keeps the partial solution and continues (depol_varEPS.cu:466-513).  The
C ABI reports the same cases in hzo_pcg_stats.status:
- permittivity 0 everywhere: every face mean is 0, so p.Ap = 0 and the
  solve breaks down at its first iteration (HZO_PCG_BREAKDOWN);
- a NaN in P: the initial r.z is not finite (HZO_PCG_BAD_INITIAL);
- max_iter below what the tolerance needs (HZO_PCG_MAXITER), with the
  partial potential equal to the oracle's after the same iterations.
Nonzero g_bot / g_top exercise the affine identity rows A x = x - g of the
Dirichlet planes (:62-71), the electrode ghosts of E (:254-261) and
init_phi_linear (:203-221).
"""
import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L
from helpers import film_inputs, oracle_film, rel_err

pytestmark = pytest.mark.gpu

FORMS = [(L.PREC_FP64, 0), (L.PREC_FP32_REFINED, 0), (L.PREC_FP32_REFINED, 2), (L.PREC_FP32_REFINED, 3)]
FORM_IDS = ["fp64", "fp32-two-kernel", "fp32-carried", "fp32-two-stencil"]


def _say(*a):
    print("[gpu]", *a, flush=True)


def _zero_eps(inp):
    rows = [dict(r) for r in inp.phases]
    for r in rows:
        r["eps_r"] = 0.0
    inp.phases = rows
    return inp


@pytest.mark.parametrize("prec,form", FORMS, ids=FORM_IDS)
def test_breakdown_on_zero_permittivity(prec, form):
    inp = _zero_eps(film_inputs(24, 16, 6, 4, fracs=(0.2, 0.6, 0.2), seed=2))
    _say("zero eps 24x16x6", FORM_IDS[FORMS.index((prec, form))])
    g = H.Film(inp, precision=prec, cg_fused=form)
    st = g.step(0.3)
    assert st.status == L.HZO_PCG_BREAKDOWN
    o = oracle_film(inp)
    Po = o.tdgl(inp.P0, np.zeros(3 * inp.N), 0.3)
    _, sto, _, _ = o.depol(Po, np.zeros(inp.N))
    assert sto.status == L.HZO_PCG_BREAKDOWN and sto.iters == 1
    if prec == L.PREC_FP64:
        assert st.iters == sto.iters
    else:
        assert st.iters <= 1
    _, phi, _ = g.get_state()
    assert np.all(np.isfinite(phi))  # the partial solution is kept
    g.close()


@pytest.mark.parametrize("prec,form", FORMS, ids=FORM_IDS)
def test_bad_initial_residual_on_nan(prec, form):
    inp = film_inputs(24, 16, 6, 4, fracs=(0.2, 0.6, 0.2), seed=2)
    _say("nan in P 24x16x6", FORM_IDS[FORMS.index((prec, form))])
    P = inp.P0.copy()
    P[2 * inp.N + 3 * inp.nx * inp.ny + 17] = np.nan
    g = H.Film(inp, precision=prec, cg_fused=form, set_state=False)
    g.set_state(P)
    st = g.depol_solve()
    o = oracle_film(inp)
    _, sto, _, _ = o.depol(P, np.zeros(inp.N))
    assert sto.status == L.HZO_PCG_BAD_INITIAL and sto.iters == 0
    assert st.status == L.HZO_PCG_BAD_INITIAL and st.iters == 0
    g.close()


def test_fp64_max_iter_matches_the_oracle():
    inp = film_inputs(32, 24, 8, 6, fracs=(0.2, 0.6, 0.2), seed=9, jitter=0.2)
    _say("max_iter 5 32x24x8")
    kw = dict(pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-14, tol_abs=1e-30, max_iter=5)
    g = H.Film(inp, **kw)
    st = g.step(0.2)
    _, phig, _ = g.get_state()
    o = oracle_film(inp, **kw)
    Po = o.tdgl(inp.P0, np.zeros(3 * inp.N), 0.2)
    phio = np.zeros(inp.N)
    _, sto, _, _ = o.depol(Po, phio)
    assert sto.status == st.status == L.HZO_PCG_MAXITER
    assert sto.iters == st.iters == 5
    assert rel_err(phig, phio) < 1e-10
    g.close()


@pytest.mark.parametrize("g_bot,g_top", [(0.3, -0.2), (-0.15, 0.0)])
def test_fp64_nonzero_electrode_potentials(g_bot, g_top):
    # The reference applies the affine identity rows A x = x - g to the
    # search directions too (apply_A at depol_varEPS.cu:481), so with g != 0
    # its PCG is not a Krylov method of a linear system: r.z grows by about
    # 1e30 over 120 iterations and the solve ends by breakdown at an
    # iteration that rounding decides.  The GPU restates the same operator;
    # parity is pinned where rounding has not yet been amplified: the
    # initial guess (init_phi_linear), the TDGL step with the electrode
    # ghosts, rz0, and the first 5 iterations.
    inp = film_inputs(36, 28, 9, 7, fracs=(0.2, 0.6, 0.2), seed=31, jitter=0.2)
    _say(f"electrodes g_bot={g_bot} g_top={g_top} 36x28x9, 5 iterations")
    Ez = float(inp.field_nd(0.9))
    g = H.Film(inp, g_bot=g_bot, g_top=g_top, max_iter=5)  # phi = NULL: init_phi_linear
    _, phi0, E0 = g.get_state()
    t = np.arange(inp.nz) / (inp.nz - 1)
    lin = np.repeat((1.0 - t) * g_bot + t * g_top, inp.nx * inp.ny)
    assert np.array_equal(phi0, lin)
    o = oracle_film(inp, g_bot=g_bot, g_top=g_top, max_iter=5)
    from oracle import pyoracle as O
    E3 = O.efield(o.c, lin)
    assert np.array_equal(E0, E3)  # the electrode planes take g as their ghost (:254-261)
    g.tdgl_step(Ez)
    Pg, _, _ = g.get_state()
    Po = o.tdgl(inp.P0, E3, Ez)
    assert np.array_equal(Pg, Po)
    st = g.depol_solve()
    _, phig, _ = g.get_state()
    phio = lin.copy()
    _, sto, _, _ = o.depol(Po, phio)
    assert abs(st.rz0 - sto.rz0) <= 1e-12 * sto.rz0
    assert st.status == sto.status == L.HZO_PCG_MAXITER and st.iters == sto.iters == 5
    assert rel_err(phig, phio) < 1e-9
    # the directions' identity rows move the electrode planes off g
    assert np.max(np.abs(phig.reshape(inp.nz, -1)[0] - g_bot)) > 0.0
    g.close()
