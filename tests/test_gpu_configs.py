"""The FP32-refined, batched and reverse-mode forms at the sizes of their
workloads, by properties that do not depend on size.

- 256x256x32 film (200 grains FE/AFE/DE 0.6/0.2/0.2, +-20 % Landau), the
  first 20 steps of the +-3 MV/cm / 2000-step triangle, FP32-refined with
  FP64 residual replacement every 20 iterations: every solve's true FP64
  residual, recomputed by the oracle's operator from the returned P and phi,
  meets ||b - A phi|| <= 1e-8 ||b||; P and phi stay within 1e-4 (max-norm
  relative) of the FP64 run of the same steps, whose TDGL is bit-exact and
  whose solves are oracle-pinned at this size.
- 64 films of uneven sizes (nx = ny from {128, 192, 256, 384}, nz from
  {16, 24, 32, 48}, 50-300 grains), 3 steps of the +-3 MV/cm sinusoid in the
  batch's measured shape: every film converges, sampled films meet the true
  residual bar, and sampled films end bit-identical to their own runs.
- 128x128x24 film (100 grains FE/AFE 0.7/0.3), 500-step triangle,
  checkpoints every 50 steps: the FP32 adjoint gradient of the loop MSE is
  within 2e-3 of the FP64 adjoint's, and a checkpoint period of 25 gives the
  same FP32 gradient bit for bit.
"""
import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L
from cmsl_amd.adjoint import coeff_matrix, rows_with
from helpers import oracle_film, rel_err
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

RELRES = dict(pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30, refine_every=20)


def _say(*a):
    print("[gpu]", *a, flush=True)


def _true_relres(o, P, phi):
    b = -O.divergence(o.c, P)
    b.reshape(o.c.nz, -1)[[0, -1]] = 0.0
    r = b - O.apply_A(o.c, o.eps, phi)
    return float(np.linalg.norm(r) / np.linalg.norm(b))


def test_fp32_refined_256x256x32_true_residual_and_fp64_agreement():
    inp = H.make_film_inputs(256, 256, 32, 200, fracs=(0.2, 0.6, 0.2), seed=1, jitter=0.2, label_device=0)
    Ez = inp.field_nd(H.triangle_wave(2000, 3.0, 20))
    o = oracle_film(inp)
    _say("256x256x32 FP32-refined, 20 steps")
    f32 = H.Film(inp, precision=L.PREC_FP32_REFINED, **RELRES)
    worst = 0.0
    for s in range(Ez.size):
        st = f32.step(Ez[s])
        assert st.status == L.HZO_PCG_CONVERGED and st.iters > 0, (s, st.status)
        P, phi, _ = f32.get_state()
        rr = _true_relres(o, P, phi)
        worst = max(worst, rr)
        assert rr <= 1.0e-8 * 1.001, (s, rr)
    _say("worst true relative residual", worst)
    f64 = H.Film(inp, **RELRES)
    f64.run(Ez)
    P64, phi64, _ = f64.get_state()
    P32, phi32, _ = f32.get_state()
    eP, ephi = rel_err(P32, P64), rel_err(phi32, phi64)
    _say("FP32 vs FP64 after 20 steps: P", eP, "phi", ephi)
    assert eP < 1e-4 and ephi < 1e-4
    f32.close()
    f64.close()


def test_batched_64_films_converge_and_match_their_own_runs():
    from concurrent.futures import ThreadPoolExecutor
    specs = H.film.batch_film_specs(64)
    with ThreadPoolExecutor(max_workers=8) as ex:
        inps = list(ex.map(lambda s: H.make_film_inputs(s[0], s[0], s[1], s[2], fracs=(0.2, 0.6, 0.2), seed=s[3],
                                                        jitter=0.2, label_device=0), specs))
    _say("64 films,", sum(i.N for i in inps), "cells, 3 steps")
    S = 3
    wave = 3.0 * np.sin(2.0 * np.pi * np.arange(S) / 1000.0 + 0.3)
    Ez = np.stack([inp.field_nd(wave) for inp in inps], axis=1)
    films = [H.Film(inp, precision=L.PREC_FP32_REFINED, **RELRES) for inp in inps]
    batch = H.Batch(films, tile_kz=8)
    batch.set_shape(8, 64.0)
    its = batch.run(Ez)
    cnt = batch.counters()
    _say("shape", batch.shape, "counters", cnt, "iterations min/max", int(its.min()), int(its.max()))
    assert cnt["unconverged"] == 0 and np.all(its > 0)
    order = np.argsort([inp.N for inp in inps])
    sample = [int(order[0]), int(order[1]), int(order[-1]), int(order[len(order) // 2])]
    for f in sample:
        P, phi, _ = films[f].get_state()
        rr = _true_relres(oracle_film(inps[f]), P, phi)
        _say("film", f, specs[f], "true relative residual", rr)
        assert rr <= 1.0e-8 * 1.001
    # each film ends as its own run: at automatic tile depth where it runs
    # its own kernels or the resident solve (groups of kind 1 and 2), at the
    # batch's depth 8 in a multi-film streaming group (kind 0); one film of
    # each kind is checked, the smallest of it, and the largest film
    kinds = [batch.film_group(f)[2] for f in range(len(films))]
    _say("films per group kind", [kinds.count(k) for k in (0, 1, 2)])
    assert set(kinds) == {0, 1, 2}
    picks = {}
    for f in [int(i) for i in order]:
        picks.setdefault(kinds[f], f)
    picks.setdefault("largest", int(order[-1]))
    batch.close()
    for f in picks.values():
        kz = 8 if kinds[f] == 0 else 0
        single = H.Film(inps[f], precision=L.PREC_FP32_REFINED, tile_kz=kz, **RELRES)
        its_s = single.run(Ez[:, f])
        Pb, phib, _ = films[f].get_state()
        Ps, phis, _ = single.get_state()
        assert np.array_equal(its[:, f], its_s), (f, its[:, f], its_s)
        assert np.array_equal(Pb, Ps) and np.array_equal(phib, phis), f
        single.close()
    for fl in films:
        fl.close()


def test_adjoint_128x128x24_500_steps_fp32_against_fp64_and_checkpoint_period():
    S = 500
    inp = H.make_film_inputs(128, 128, 24, 100, fracs=(0.3, 0.7, 0.0), seed=5, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(S, 3.0, S))
    base = coeff_matrix(inp.phases)
    hidden = base * (1.0 + 0.1 * np.random.default_rng(12345).uniform(-1.0, 1.0, base.shape))
    _say("adjoint 128x128x24, 500 steps, K=50")

    def film(prec, K):
        f = H.Film(inp, precision=prec, **RELRES)
        f.adjoint_reserve(S, K, 1e-8)
        return f

    def grad(f, target):
        f.set_phase_table(rows_with(inp.phases, base))
        f.set_state(inp.P0)
        m = f.forward_record(Ez)
        return f.backward(2.0 * (m - target) / S)[0], m

    f64 = film(L.PREC_FP64, 50)
    f64.set_phase_table(rows_with(inp.phases, hidden))
    f64.set_state(inp.P0)
    target = f64.forward_record(Ez)
    g64, m64 = grad(f64, target)
    f64.close()
    f32 = film(L.PREC_FP32_REFINED, 50)
    g32, m32 = grad(f32, target)
    f32.close()
    f32b = film(L.PREC_FP32_REFINED, 25)
    g32b, _ = grad(f32b, target)
    f32b.close()
    e = float(np.max(np.abs(g32 - g64)) / np.max(np.abs(g64)))
    _say("FP32 vs FP64 adjoint gradient", e, "loop", rel_err(m32, m64))
    assert np.all(np.isfinite(g64)) and np.max(np.abs(g64)) > 0.0
    assert e < 2e-3
    assert np.array_equal(g32, g32b)
