"""The staircase P-E sweep driver (hzo_sweep_run, the Evolve loop of
evolve.cu:213-410) on the GPU, against the same schedule run on the oracle.

The schedule is restated here from the reference loop (field 0, +delEz per
level of num_steps steps, turn down above +maxEz counting a half loop, turn
up below -maxEz, stop at peloop == 2, rows from the first turn on), and each
row's columns from the diagnostics rule (evolve.cu:259-357, means over the
film, masked means per phase, P in uC/cm^2, E in MV/cm) applied in numpy to
the oracle's state; the written file must hold exactly the "%le\\t" x 11 +
"%d\\n" text of the rows (evolve.cu:391-400), appended to what was there.
"""
import re

import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L
from helpers import film_inputs, oracle_film

pytestmark = pytest.mark.gpu

FE, DE, AFE = L.HZO_FE, L.HZO_DE, L.HZO_AFE


def _say(*a):
    print("[gpu]", *a, flush=True)


def _oracle_rows(inp, maxEz, delEz, num_steps, Ez_scale, P0_scale):
    o = oracle_film(inp)
    N = inp.N
    P = inp.P0.copy()
    phi = np.zeros(N)
    E3 = np.zeros(3 * N)
    mat = np.where(inp.omega >= 0, inp.grain_phase[np.maximum(inp.omega, 0)], -1)
    Ez, inc, peloop, count = 0.0, delEz, 0, 0
    rows = []
    while peloop < 2:
        if num_steps > 1:
            o.run(P, phi, E3, np.full(num_steps - 1, Ez))
        _, Pl = o.run(P, phi, E3, np.array([Ez]), want_local=True)
        count += num_steps
        Pmz = P[2 * N:3 * N]
        cols = [Pmz.sum() / N, P[5 * N:].sum() / N, Pl[2 * N:3 * N].sum() / N, Pl[5 * N:].sum() / N]
        per = {}
        for t in (FE, AFE, DE):
            m = mat == t
            cnt = float(m.sum())
            per[t] = ((Pmz[m].sum() / cnt) if cnt > 0 else 0.0, cnt / N if cnt > 0 else 0.0)
        if Ez > maxEz:
            inc = -delEz
            peloop += 1
        if Ez < -maxEz:
            inc = delEz
        if peloop > 0:
            r = [Ez * Ez_scale] + [v * P0_scale for v in cols]
            for t in (FE, AFE, DE):  # column order of the row (evolve.cu:391-400)
                r += [per[t][0] * P0_scale, per[t][1]]
            rows.append(r + [count])
        Ez += inc
    return np.array(rows)


def test_sweep_run_rows_and_file_text(tmp_path):
    inp = film_inputs(24, 20, 8, 6, fracs=(0.3, 0.4, 0.3), seed=12, jitter=0.2)
    E0 = inp.scales["E0"]
    Eap, delEap, num_steps = 2.4e8, 1.0e8, 3
    p = H.film.sweep_params(inp.scales, Eap, delEap, num_steps)
    path = tmp_path / "sweep001.txt"
    path.write_text("# rows appended below\n")
    _say("sweep 24x20x8, 3 steps per level")
    g = H.Film(inp)
    rows = g.sweep(Eap, delEap, num_steps, path=path)
    g.close()
    exp = _oracle_rows(inp, Eap / E0, delEap / E0, num_steps, p.Ez_scale, p.P0_scale)
    # 4 levels up, 6 down, 6 up; rows from the first turn: 1 + 6 + 6
    assert rows.shape == exp.shape == (13, 12)
    sched_E, sched_cnt = H.film.sweep_schedule(p)
    assert np.array_equal(rows[:, 0], sched_E) and np.array_equal(rows[:, 11], sched_cnt)
    assert np.array_equal(rows[:, 0], exp[:, 0]) and np.array_equal(rows[:, 11], exp[:, 11])
    for c in (6, 8, 10):  # phase fractions: exact
        assert np.array_equal(rows[:, c], exp[:, c])
    for c in (1, 2, 3, 4, 5, 7, 9):
        scale = max(np.max(np.abs(exp[:, c])), 1e-300)
        assert np.max(np.abs(rows[:, c] - exp[:, c])) <= 1e-10 * scale, c
    # the file: the old line, then one line per row in the reference format
    text = path.read_text()
    lines = text.splitlines(keepends=True)
    assert lines[0] == "# rows appended below\n" and len(lines) == 1 + rows.shape[0]
    num = r"-?\d\.\d{6}e[+-]\d{2,3}"
    pat = re.compile(r"^(" + num + r"\t){11}-?\d+\n$")
    for line, r in zip(lines[1:], rows):
        assert pat.match(line), line
        assert line == "\t".join("%e" % v for v in r[:11]) + "\t%d\n" % int(r[11])
