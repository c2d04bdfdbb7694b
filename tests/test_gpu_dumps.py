"""Snapshot dumps and checkpoint restart of a device film.

A restart reads P from the reference checkpoint file and re-solves phi from
it (the file stores E, not phi).  P must come back bit for bit, the
re-solved E must match the stored E to the solver tolerance, and stepping on
from the restart must track the uninterrupted run.
"""
import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L
from helpers import film_inputs, rel_err

pytestmark = pytest.mark.gpu


def _film(inp):
    return H.Film(inp, precision=L.PREC_FP64, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-11, tol_abs=1e-30)


def test_checkpoint_restart_tracks_the_uninterrupted_run(tmp_path):
    inp = film_inputs(24, 20, 8, 6, fracs=(0.3, 0.5, 0.2), seed=21, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(20, 3.0, 10))
    print("[gpu] checkpoint restart 24x20x8", flush=True)
    a = _film(inp)
    a.run(Ez[:6])
    path = tmp_path / "ck.bin"
    a.checkpoint_save(path, peloop=2, count=6, Ez_ext=float(Ez[5]), incEz=0.01)
    P1, _, E1 = a.get_state()
    b = _film(inp)
    hdr, E_file = b.checkpoint_load(path)
    assert hdr == {"peloop": 2, "count": 6, "Ez_ext": float(Ez[5]), "incEz": 0.01}
    P2, _, E2 = b.get_state()
    assert np.array_equal(P1, P2)
    assert np.array_equal(E_file, E1)
    assert rel_err(E2, E1) < 1e-8
    a.run(Ez[6:])
    b.run(Ez[6:])
    Pa, _, _ = a.get_state()
    Pb, _, _ = b.get_state()
    assert rel_err(Pb, Pa) < 1e-9
    a.close()
    b.close()


def test_snapshot_dumps_hold_the_scaled_state(tmp_path):
    inp = film_inputs(16, 12, 6, 4, fracs=(0.3, 0.5, 0.2), seed=22)
    f = _film(inp)
    f.run(inp.field_nd(np.full(3, 1.0)))
    P, _, _ = f.get_state()
    N = inp.N
    f.save_snapshot(tmp_path / "s_pm.bin", tmp_path / "s_ps.bin", 100.0)
    pm = np.fromfile(tmp_path / "s_pm.bin", dtype="<f8").reshape(N, 3)
    ps = np.fromfile(tmp_path / "s_ps.bin", dtype="<f8").reshape(N, 3)
    assert np.array_equal(pm, P[: 3 * N].reshape(3, N).T * 100.0)
    assert np.array_equal(ps, P[3 * N:].reshape(3, N).T * 100.0)
    f.close()
