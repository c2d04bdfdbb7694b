"""The FP64 PCG on z-streaming tiles (tile depth 8, 16 and 32).

Large FP64 films run the direction and update kernels that walk their planes
bottom to top instead of staging a shallow tile at once.  They evaluate
every voxel with the same IEEE operations as the shallow kernels, so the
parity bar is the FP64 path's own: equal PCG iterations in every step and
P, phi within 1e-10 of the oracle.  Cases: ragged films (nx not a multiple
of 64, ny not of 8, nz not of the depth), Dirichlet electrodes under both
stopping rules, Neumann electrodes, the reverse-mode solve, and the
automatic choice at 512x512x64 checked against the shallow kernels there.
"""
import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L
from cmsl_amd.adjoint import coeff_matrix, rows_with
from helpers import film_inputs, oracle_film, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _say(*a):
    print("[gpu]", *a, flush=True)


@pytest.mark.parametrize("kz,shape,rule,zbc", [
    (8, (72, 36, 20), L.PCG_RULE_RELRES, L.ZBC_DIRICHLET),
    (8, (40, 24, 12), L.PCG_RULE_REFERENCE, L.ZBC_NEUMANN),
    (16, (72, 36, 20), L.PCG_RULE_RELRES, L.ZBC_DIRICHLET),
    (32, (72, 36, 40), L.PCG_RULE_RELRES, L.ZBC_DIRICHLET),
    (16, (64, 24, 16), L.PCG_RULE_REFERENCE, L.ZBC_DIRICHLET),
    (32, (40, 20, 33), L.PCG_RULE_REFERENCE, L.ZBC_DIRICHLET),
    (16, (48, 24, 18), L.PCG_RULE_REFERENCE, L.ZBC_NEUMANN),
])
def test_deep_tiles_match_the_oracle(kz, shape, rule, zbc):
    nx, ny, nz = shape
    inp = film_inputs(nx, ny, nz, 10, fracs=(0.2, 0.6, 0.2), seed=61 + kz + nz, jitter=0.2)
    steps = 6 if zbc == L.ZBC_NEUMANN else 10
    Ez = inp.field_nd(H.triangle_wave(24, 3.0, steps))
    kw = dict(pcg_rule=rule, zbc=zbc)
    if rule == L.PCG_RULE_RELRES:
        kw.update(tol_rel=1e-10, tol_abs=1e-30)
    if zbc == L.ZBC_NEUMANN:
        kw.update(max_iter=60)
    _say(f"FP64 depth {kz} {nx}x{ny}x{nz} rule {rule} zbc {zbc}")
    g = H.Film(inp, precision=L.PREC_FP64, tile_kz=kz, **kw)
    assert g.tile_depth == kz
    its = g.run(Ez)
    P, phi, E = g.get_state()
    g.close()
    o = oracle_film(inp, **kw)
    Po = inp.P0.copy()
    phio = np.zeros(inp.N)
    Eo = np.zeros(3 * inp.N)
    its_o, _ = o.run(Po, phio, Eo, Ez)
    _say("iterations", list(its), "oracle", list(its_o))
    assert list(its) == list(its_o)
    tol = 1e-9 if zbc == L.ZBC_NEUMANN else TOL
    assert rel_err(P, Po) < tol, rel_err(P, Po)
    assert rel_err(phi, phio) < tol, rel_err(phi, phio)


def test_deep_tiles_in_the_reverse_mode_solve():
    # the adjoint solve runs the same PCG kernels with its own right-hand
    # side; depth 16 against the oracle-pinned shallow tiles
    inp = film_inputs(40, 24, 20, 6, fracs=(0.3, 0.5, 0.2), seed=4, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(12, 3.0, 12))
    grads = {}
    for kz in (4, 16):
        _say(f"FP64 adjoint 40x24x20 depth {kz}")
        f = H.Film(inp, precision=L.PREC_FP64, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-12, tol_abs=1e-30,
                   tile_kz=kz)
        f.adjoint_reserve(Ez.size, ckpt_every=5, tol=1e-12)
        f.set_phase_table(rows_with(inp.phases, coeff_matrix(inp.phases)))
        f.set_state(inp.P0)
        m = f.forward_record(Ez)
        g, its = f.backward(2.0 * (m - 0.9 * m))
        grads[kz] = (m, g, its)
        f.close()
    m4, g4, _ = grads[4]
    m16, g16, _ = grads[16]
    assert rel_err(m16, m4) < 1e-10
    assert rel_err(g16, g4) < 1e-8, (g16, g4)


def test_large_film_automatic_depth_against_shallow_tiles():
    _say("512x512x64 FP64 inputs")
    inp = H.make_film_inputs(512, 512, 64, 1000, fracs=(0.3, 0.5, 0.2), seed=41, jitter=0.2, label_device=0,
                             gb_dielectric=0.01)
    Ez = inp.field_nd(H.triangle_wave(5000, 4.0, 2))
    out = {}
    for kz in (0, 4):
        g = H.Film(inp, precision=L.PREC_FP64, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30,
                   tile_kz=kz)
        depth = g.tile_depth
        its = g.run(Ez)
        P, phi, _ = g.get_state()
        g.close()
        _say(f"512x512x64 FP64 depth {depth}: iterations {list(its)}")
        out[kz] = (depth, list(its), P, phi)
    assert out[0][0] == 32 and out[4][0] == 4
    assert out[0][1] == out[4][1]
    assert rel_err(out[0][2], out[4][2]) < TOL
    assert rel_err(out[0][3], out[4][3]) < TOL
