"""Parity of the HIP FP64 path (libhzo.so) with the CPU FP64 oracle.

Bar: the TDGL update is bit-exact (same IEEE operations, no contraction on
either side); after N steps P and phi agree within 1e-10 relative (max-norm),
with equal PCG iteration counts in every step.  Only the inner-product
summation order differs between the two sides.
"""
import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L
from helpers import film_inputs, oracle_film, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _say(*a):
    print("[gpu]", *a, flush=True)


def _pair(inp, **kw):
    okw = {k: v for k, v in kw.items() if k in ("dep_flag", "zbc", "pcg_rule", "tol_rel", "tol_abs", "max_iter")}
    return H.Film(inp, **kw), oracle_film(inp, **okw)


def test_tdgl_step_bitexact_with_field():
    inp = film_inputs(40, 36, 12, 9, fracs=(0.3, 0.5, 0.2), seed=21, jitter=0.2)
    rng = np.random.default_rng(0)
    phi = rng.uniform(-0.3, 0.3, inp.N)
    phi.reshape(inp.nz, -1)[[0, -1]] = 0.0
    _say("tdgl bitexact 40x36x12 seed 21")
    g, o = _pair(inp)
    g.set_state(inp.P0, phi)
    g.tdgl_step(0.4)
    Pg, phig, Eg = g.get_state()
    from oracle import pyoracle as O
    E3 = O.efield(o.c, phi)
    assert np.array_equal(Eg, E3)
    Po = o.tdgl(inp.P0, E3, 0.4)
    assert np.array_equal(Pg, Po), rel_err(Pg, Po)
    g.close()


def test_single_solve_parity():
    inp = film_inputs(48, 40, 10, 12, fracs=(0.2, 0.6, 0.2), seed=5, jitter=0.2)
    _say("depol 48x40x10 seed 5")
    g, o = _pair(inp)
    st = g.step(0.3)
    Pg, phig, Eg = g.get_state()
    Po = o.tdgl(inp.P0, np.zeros(3 * inp.N), 0.3)
    assert np.array_equal(Pg, Po)
    phio = np.zeros(inp.N)
    E3, sto, divP, _ = o.depol(Po, phio)
    assert st.iters == sto.iters and st.status == sto.status == 0
    assert rel_err(phig, phio) < TOL
    assert rel_err(Eg, E3) < TOL
    assert abs(st.rz0 - sto.rz0) <= 1e-12 * abs(sto.rz0)
    assert abs(st.bb - sto.bb) <= 1e-12 * abs(sto.bb)
    g.close()


def _run_pair(inp, Ez, **kw):
    g, o = _pair(inp, **kw)
    its_g = g.run(Ez)
    Pg, phig, Eg = g.get_state()
    P = inp.P0.copy()
    phi = np.zeros(inp.N)
    E3 = np.zeros(3 * inp.N)
    its_o, _ = o.run(P, phi, E3, Ez)
    g.close()
    return its_g, Pg, phig, its_o, P, phi


def test_run_parity_single_film_200_steps():
    # 64x64x16, 8 FE grains, 1 MV/cm, 200 steps with a PCG solve after each
    inp = film_inputs(64, 64, 16, 8, fracs=(0.0, 1.0, 0.0), seed=3)
    Ez = np.full(200, float(inp.field_nd(1.0)))
    _say("run 64x64x16 FE 200 steps")
    its_g, Pg, phig, its_o, Po, phio = _run_pair(inp, Ez)
    assert np.array_equal(its_g, its_o), (its_g, its_o)
    assert rel_err(Pg, Po) < TOL, rel_err(Pg, Po)
    assert rel_err(phig, phio) < TOL, rel_err(phig, phio)


def test_run_parity_three_phases_triangle_sweep():
    inp = film_inputs(32, 24, 10, 14, fracs=(0.2, 0.6, 0.2), seed=8, jitter=0.2)
    assert inp.iface.any()
    Ez = inp.field_nd(H.triangle_wave(80, 3.0, 60))
    _say("run 32x24x10 FE/AFE/DE 60 steps")
    its_g, Pg, phig, its_o, Po, phio = _run_pair(inp, Ez)
    assert np.array_equal(its_g, its_o)
    assert rel_err(Pg, Po) < TOL
    assert rel_err(phig, phio) < TOL


def test_run_parity_relative_residual_rule():
    inp = film_inputs(48, 40, 12, 10, fracs=(0.2, 0.6, 0.2), seed=12, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(2000, 3.0, 15))
    _say("run 48x40x12 relres 1e-8 15 steps")
    its_g, Pg, phig, its_o, Po, phio = _run_pair(inp, Ez, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8,
                                                 tol_abs=1e-30)
    assert np.array_equal(its_g, its_o)
    assert rel_err(Pg, Po) < TOL
    assert rel_err(phig, phio) < TOL


def test_run_parity_neumann_gauge():
    inp = film_inputs(24, 20, 8, 6, fracs=(0.2, 0.6, 0.2), seed=2, jitter=0.1)
    Ez = np.full(6, float(inp.field_nd(0.5)))
    _say("run 24x20x8 Neumann 6 steps")
    its_g, Pg, phig, its_o, Po, phio = _run_pair(inp, Ez, zbc=L.ZBC_NEUMANN, max_iter=40)
    assert np.array_equal(its_g, its_o)
    assert rel_err(Pg, Po) < 1e-9
    assert rel_err(phig, phio) < 1e-9


def test_run_parity_odd_sizes():
    # grid not a multiple of the 64x4 tile, nz odd
    inp = film_inputs(37, 29, 7, 11, fracs=(0.3, 0.4, 0.3), seed=31, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(40, 2.0, 12))
    _say("run 37x29x7 12 steps")
    its_g, Pg, phig, its_o, Po, phio = _run_pair(inp, Ez)
    assert np.array_equal(its_g, its_o)
    assert rel_err(Pg, Po) < TOL
    assert rel_err(phig, phio) < TOL


def test_void_voxels_tdgl_bitexact():
    # Void voxels (omega < 0) are skipped by the TDGL kernel and read as 0.
    # With eps = 0 they also make the Poisson matrix singular (zero rows), a
    # degenerate system the reference cannot solve either, so this case runs
    # without the depolarisation solve.
    inp = film_inputs(37, 29, 7, 11, fracs=(0.3, 0.4, 0.3), seed=31, jitter=0.2)
    om = inp.omega.copy()
    om[np.random.default_rng(4).random(inp.N) < 0.05] = -1
    inp.omega = om
    inp.iface = H.film.interface_flag(inp.nx, inp.ny, inp.nz, om, inp.grain_phase)
    inp.P0 = H.film.init_polarization(inp.nx, inp.ny, inp.nz, om, inp.grain_phase, 0.0124, 9)
    Ez = inp.field_nd(H.triangle_wave(40, 2.0, 12))
    _say("run 37x29x7 void dep_flag=0 12 steps")
    its_g, Pg, phig, its_o, Po, phio = _run_pair(inp, Ez, dep_flag=0)
    assert np.array_equal(Pg, Po)
    assert np.all(Pg.reshape(6, -1)[:, om < 0] == 0.0)


def test_no_depolarisation_path_bitexact():
    inp = film_inputs(30, 20, 6, 5, fracs=(0.3, 0.4, 0.3), seed=17, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(40, 3.0, 25))
    _say("run 30x20x6 dep_flag=0 25 steps")
    its_g, Pg, phig, its_o, Po, phio = _run_pair(inp, Ez, dep_flag=0)
    assert np.array_equal(Pg, Po)
    assert np.all(its_g == 0)


def test_diagnostics_parity():
    inp = film_inputs(40, 32, 8, 10, fracs=(0.2, 0.6, 0.2), seed=44, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(40, 3.0, 7))
    _say("diagnostics 40x32x8")
    g, o = _pair(inp)
    g.run(Ez[:-1])
    Pg_before, _, _ = g.get_state()
    g.step(Ez[-1])
    dg = g.diagnostics()
    P = inp.P0.copy()
    phi = np.zeros(inp.N)
    E3 = np.zeros(3 * inp.N)
    o.run(P, phi, E3, Ez[:-1])
    _, Pl = o.run(P, phi, E3, Ez[-1:], want_local=True)
    do = o.diagnostics(P, Pl)
    assert np.allclose(dg, do, rtol=1e-11, atol=1e-11 * inp.N)
    assert dg[5] + dg[7] + dg[9] == inp.N  # phase masks cover the film
    g.close()


def test_full_size_film_step_and_residual():
    # 256x256x32, 200 grains FE/AFE/DE 0.6/0.2/0.2: one step; TDGL bit-exact,
    # PCG residual of the returned potential meets ||r|| <= 1e-8 ||b||.
    from oracle import pyoracle as O
    inp = film_inputs(256, 256, 32, 200, fracs=(0.2, 0.6, 0.2), seed=1, jitter=0.2)
    _say("full size 256x256x32 one step")
    g, o = _pair(inp, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30)
    Ez = float(inp.field_nd(0.3))
    g.tdgl_step(Ez)
    Pg, _, _ = g.get_state()
    Po = o.tdgl(inp.P0, np.zeros(3 * inp.N), Ez)
    assert np.array_equal(Pg, Po)
    st = g.depol_solve()
    _, phig, _ = g.get_state()
    b = -O.divergence(o.c, Pg)
    b.reshape(inp.nz, -1)[[0, -1]] = 0.0
    r = b - O.apply_A(o.c, o.eps, phig)
    assert st.status == 0 and st.iters > 0
    assert np.linalg.norm(r) <= 1.0e-8 * np.linalg.norm(b) * 1.001
    g.close()
