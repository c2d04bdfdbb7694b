"""The register-resident single-reduction CG (one cooperative launch per
solve, vectors held in registers and LDS) against the FP64 oracle and
against the streaming single-reduction CG it replaces.

Bars: every solve's true FP64 residual, recomputed by the oracle's operator
from the returned P and phi, meets ||b - A phi|| <= 1e-8 ||b||; P and phi
within 1e-5 (max-norm relative) of the FP64 oracle over 12 steps; against
the streaming form of the same iteration, PCG iteration counts within 1 and
P, phi within 1e-6 (the two group the inner products differently).
"""
import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L
from helpers import film_inputs, oracle_film, rel_err
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

RELRES = dict(pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30, refine_every=20)


def _say(*a):
    print("[gpu]", *a, flush=True)


def _true_relres(o, P, phi, zbc=0):
    b = -O.divergence(o.c, P)
    if zbc == 0:
        b.reshape(o.c.nz, -1)[[0, -1]] = 0.0
    else:
        b[o.c.nx // 2 + o.c.nx * (o.c.ny // 2 + o.c.ny * (o.c.nz // 2))] = 0.0
    r = b - O.apply_A(o.c, o.eps, phi)
    return float(np.linalg.norm(r) / np.linalg.norm(b))


SHAPES = [(64, 48, 16, 10), (32, 32, 24, 6), (48, 32, 32, 8)]


@pytest.mark.parametrize("nx,ny,nz,ng", SHAPES)
def test_resident_cg_against_oracle_and_true_residual(nx, ny, nz, ng):
    inp = film_inputs(nx, ny, nz, ng, fracs=(0.2, 0.6, 0.2), seed=7 + nz, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(40, 3.0, 12))
    _say(f"resident {nx}x{ny}x{nz}, 12 steps")
    g = H.Film(inp, precision=L.PREC_FP32_REFINED, cg_fused=4, **RELRES)
    assert g.cg_form == 3
    o = oracle_film(inp, pcg_rule=1, tol_rel=1e-8, tol_abs=1e-30)
    for s in range(Ez.size):
        st = g.step(Ez[s])
        assert st.status == L.HZO_PCG_CONVERGED and st.iters > 0, (s, st.status, st.iters)
        P, phi, _ = g.get_state()
        rr = _true_relres(o, P, phi)
        assert rr <= 1e-8 * 1.001, (s, rr)
    Pg, phig, _ = g.get_state()
    Po = inp.P0.copy()
    phio = np.zeros(inp.N)
    o.run(Po, phio, np.zeros(3 * inp.N), Ez)
    eP, ephi = rel_err(Pg, Po), rel_err(phig, phio)
    _say("vs oracle: P", eP, "phi", ephi)
    assert eP < 1e-5 and ephi < 1e-5
    g.close()


def test_resident_against_streaming_form_and_counters():
    inp = film_inputs(64, 48, 16, 10, fracs=(0.2, 0.6, 0.2), seed=3, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(40, 3.0, 12))
    _say("resident vs carried 64x48x16")
    a = H.Film(inp, precision=L.PREC_FP32_REFINED, cg_fused=4, **RELRES)
    b = H.Film(inp, precision=L.PREC_FP32_REFINED, cg_fused=2, tile_kz=8, **RELRES)
    a.set_timing(True)
    its_a = a.run(Ez)
    solves, iters, passes = a.cg_counters()
    its_b = b.run(Ez)
    assert np.all(np.abs(its_a - its_b) <= 1), (its_a, its_b)
    assert solves == Ez.size and iters == int(its_a.sum())
    assert passes > iters  # one per iteration, plus the re-formations and replacements
    Pa, phia, _ = a.get_state()
    Pb, phib, _ = b.get_state()
    assert rel_err(Pa, Pb) < 1e-6 and rel_err(phia, phib) < 1e-6
    a.close()
    b.close()


def test_resident_neumann_electrodes():
    inp = film_inputs(48, 32, 16, 7, fracs=(0.2, 0.6, 0.2), seed=21, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(40, 3.0, 12))
    _say("resident Neumann 48x32x16")
    g = H.Film(inp, precision=L.PREC_FP32_REFINED, cg_fused=4, zbc=L.ZBC_NEUMANN, **RELRES)
    o = oracle_film(inp, zbc=1, pcg_rule=1, tol_rel=1e-8, tol_abs=1e-30)
    for s in range(Ez.size):
        st = g.step(Ez[s])
        assert st.status == L.HZO_PCG_CONVERGED, (s, st.status)
        P, phi, _ = g.get_state()
        assert _true_relres(o, P, phi, zbc=1) <= 1e-8 * 1.001
    Pg, phig, _ = g.get_state()
    Po = inp.P0.copy()
    phio = np.zeros(inp.N)
    o.run(Po, phio, np.zeros(3 * inp.N), Ez)
    assert rel_err(Pg, Po) < 1e-5 and rel_err(phig, phio) < 1e-5
    g.close()


def test_resident_max_iter_and_failure_statuses():
    inp = film_inputs(32, 32, 16, 5, fracs=(0.2, 0.6, 0.2), seed=2)
    _say("resident statuses 32x32x16")
    g = H.Film(inp, precision=L.PREC_FP32_REFINED, cg_fused=4, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-14,
               tol_abs=1e-30, max_iter=7)
    st = g.step(0.3)
    assert st.status == L.HZO_PCG_MAXITER and st.iters == 7
    g.close()
    P = inp.P0.copy()
    P[2 * inp.N + 5 * inp.nx * inp.ny + 3] = np.nan
    g = H.Film(inp, precision=L.PREC_FP32_REFINED, cg_fused=4, set_state=False, **RELRES)
    g.set_state(P)
    st = g.depol_solve()
    assert st.status == L.HZO_PCG_BAD_INITIAL and st.iters == 0
    g.close()
    rows = [dict(r) for r in inp.phases]
    for r in rows:
        r["eps_r"] = 0.0
    inp.phases = rows
    g = H.Film(inp, precision=L.PREC_FP32_REFINED, cg_fused=4, **RELRES)
    st = g.step(0.3)
    assert st.status == L.HZO_PCG_BREAKDOWN and st.iters <= 1
    g.close()


def test_resident_refused_for_shapes_it_does_not_cover():
    inp = film_inputs(40, 32, 16, 5, seed=2)
    with pytest.raises(L.HzoError):
        H.Film(inp, precision=L.PREC_FP32_REFINED, cg_fused=4, **RELRES)
    # automatic choice: resident only where it applies, and not with a fixed tile depth
    assert H.Film(inp, precision=L.PREC_FP32_REFINED, **RELRES).cg_form in (1, 2)
    inp2 = film_inputs(32, 32, 16, 5, seed=2)
    assert H.Film(inp2, precision=L.PREC_FP32_REFINED, **RELRES).cg_form == 3
    assert H.Film(inp2, precision=L.PREC_FP32_REFINED, tile_kz=8, **RELRES).cg_form in (1, 2)


def test_run_enqueues_resident_steps_and_matches_single_steps():
    # hzo_run on a resident film enqueues its steps 64 at a time and reads
    # their control blocks afterwards; 70 steps cross one such group.  The
    # same kernels run as with one hzo_step per step, so the states and the
    # iteration counts are bit-identical.
    inp = film_inputs(32, 32, 16, 6, fracs=(0.2, 0.6, 0.2), seed=17, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(40, 3.0, 70))
    _say("resident 32x32x16: hzo_run of 70 steps against 70 hzo_step calls")
    a = H.Film(inp, precision=L.PREC_FP32_REFINED, **RELRES)
    b = H.Film(inp, precision=L.PREC_FP32_REFINED, **RELRES)
    assert a.cg_form == 3 and b.cg_form == 3
    its_a = a.run(Ez)
    its_b = [b.step(float(e)).iters for e in Ez]
    Pa, phia, Ea = a.get_state()
    Pb, phib, Eb = b.get_state()
    a.close()
    b.close()
    assert list(its_a) == its_b
    assert np.array_equal(Pa, Pb) and np.array_equal(phia, phib) and np.array_equal(Ea, Eb)
