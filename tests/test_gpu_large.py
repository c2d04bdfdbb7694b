"""Dielectric grain-boundary cells and the large-film workload shape.

A film whose boundary cells carry the dielectric phase is an ordinary film
with more grains, so the FP64 path must still match the oracle within
1e-10 with equal PCG iterations.  At 512x512x64 (16.8 M cells, 1000 grains,
FE/AFE/DE 0.5/0.3/0.2, 1 % dielectric boundary cells, Voronoi labels from the
GPU labeller) the FP32-refined step is checked through size-independent
properties: the solve meets ||b - A phi|| <= 1e-8 ||b|| in FP64 (recomputed
by the oracle operator) and the state stays finite.
"""
import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L
from helpers import oracle_film, rel_err
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def test_fp64_parity_with_dielectric_boundary_cells():
    inp = H.make_film_inputs(40, 36, 10, 9, fracs=(0.3, 0.5, 0.2), seed=31, jitter=0.2, gb_dielectric=0.01)
    print(f"[gpu] 40x36x10 with {inp.grain_phase.size - 9} boundary grains", flush=True)
    Ez = inp.field_nd(H.triangle_wave(30, 3.0, 12))
    g = H.Film(inp, precision=L.PREC_FP64, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-10, tol_abs=1e-30)
    its = g.run(Ez)
    P, phi, _ = g.get_state()
    g.close()
    o = oracle_film(inp, pcg_rule=1, tol_rel=1e-10, tol_abs=1e-30)
    Po = inp.P0.copy()
    phio = np.zeros(inp.N)
    its_o, _ = o.run(Po, phio, np.zeros(3 * inp.N), Ez)
    assert list(its) == list(its_o)
    assert rel_err(P, Po) < 1e-10 and rel_err(phi, phio) < 1e-10


def test_fp32_step_at_512x512x64_with_dielectric_boundary_cells():
    print("[gpu] 512x512x64 inputs", flush=True)
    inp = H.make_film_inputs(512, 512, 64, 1000, fracs=(0.3, 0.5, 0.2), seed=41, jitter=0.2, label_device=0,
                             gb_dielectric=0.01)
    g = H.Film(inp, precision=L.PREC_FP32_REFINED, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30,
               refine_every=20)
    print(f"[gpu] 512x512x64 film created, tile depth {g.tile_depth}", flush=True)
    Ez = inp.field_nd(H.triangle_wave(5000, 4.0, 3))
    its = g.run(Ez)
    st = g.step(float(inp.field_nd(0.5)))
    P, phi, _ = g.get_state()
    g.close()
    print(f"[gpu] iterations {list(its)} + {st.iters}", flush=True)
    assert st.status == 0 and np.all(np.isfinite(P)) and np.all(np.isfinite(phi))
    o = oracle_film(inp, pcg_rule=1, tol_rel=1e-8, tol_abs=1e-30)
    b = -O.divergence(o.c, P)
    b.reshape(inp.nz, -1)[[0, -1]] = 0.0
    r = b - O.apply_A(o.c, o.eps, phi)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    print(f"[gpu] true relative residual {rel:.3e}", flush=True)
    assert rel <= 1.0e-8 * 1.05


@pytest.mark.parametrize("prec", [L.PREC_FP64, L.PREC_FP32_REFINED])
def test_folded_partials_beyond_2048_workgroups(prec):
    # 320x256x32 at tile depth 2: 5 x 32 x 16 = 2560 workgroups, so every
    # launch's partial sums are folded by a separate one-workgroup kernel
    inp = H.make_film_inputs(320, 256, 32, 40, fracs=(0.2, 0.6, 0.2), seed=51, jitter=0.2)
    print(f"[gpu] folded partials, 320x256x32 kz 2, precision {prec}", flush=True)
    Ez = inp.field_nd(H.triangle_wave(40, 3.0, 3))
    g = H.Film(inp, precision=prec, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-10 if prec == L.PREC_FP64 else 1e-8,
               tol_abs=1e-30, tile_kz=2)
    its = g.run(Ez)
    P, phi, _ = g.get_state()
    g.close()
    o = oracle_film(inp, pcg_rule=1, tol_rel=1e-10, tol_abs=1e-30)
    if prec == L.PREC_FP64:
        Po = inp.P0.copy()
        phio = np.zeros(inp.N)
        its_o, _ = o.run(Po, phio, np.zeros(3 * inp.N), Ez)
        print(f"[gpu] iterations {list(its)} oracle {list(its_o)}", flush=True)
        assert list(its) == list(its_o)
        assert rel_err(P, Po) < 1e-10 and rel_err(phi, phio) < 1e-10
    else:
        b = -O.divergence(o.c, P)
        b.reshape(inp.nz, -1)[[0, -1]] = 0.0
        rel = np.linalg.norm(b - O.apply_A(o.c, o.eps, phi)) / np.linalg.norm(b)
        print(f"[gpu] iterations {list(its)}, true relative residual {rel:.3e}", flush=True)
        assert rel <= 1.0e-8 * 1.05
