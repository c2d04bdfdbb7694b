"""Batched films (hzo_batch_*) against the same films stepped one by one.

A batch runs every film's kernels in shared launches over a (film, tile)
table, but each film keeps its own tile numbering, partial-sum slots and
solver control block, so at the same tile depth each film must end in
exactly the state its own hzo_run leaves: P, x_hi and the PCG iteration
counts are compared bit for bit.  The films differ in every dimension, grain
count and Landau coefficients (+-20 % jitter), and run different applied
fields, so a film whose solve finishes early idles in later CG launches.
"""
import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L
from helpers import film_inputs

pytestmark = pytest.mark.gpu

SHAPES = [(64, 48, 16, 12), (40, 72, 12, 9), (96, 32, 10, 15), (32, 32, 8, 6)]


def _say(*a):
    print("[gpu]", *a, flush=True)


def _inputs():
    return [film_inputs(nx, ny, nz, ng, fracs=(0.3, 0.5, 0.2), seed=11 + f, jitter=0.2)
            for f, (nx, ny, nz, ng) in enumerate(SHAPES)]


def _film(inp, kz=2, zbc=L.ZBC_DIRICHLET):
    return H.Film(inp, precision=L.PREC_FP32_REFINED, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30,
                  refine_every=20, tile_kz=kz, zbc=zbc)


def _fields(inps, S):
    amps = [3.0, 2.0, 2.5, 1.5]
    return np.stack([inp.field_nd(H.triangle_wave(S, amps[f], S)) for f, inp in enumerate(inps)], axis=1)


@pytest.mark.parametrize("kz,mib,groups", [(2, 0.0, 1), (4, 0.0, 1), (8, 0.0, 1), (2, 1.5, 3)])
def test_batched_films_match_single_film_runs_bit_for_bit(kz, mib, groups):
    inps = _inputs()
    S = 14
    Ez = _fields(inps, S)
    _say(f"batch of {len(inps)} films, {S} steps, kz={kz}, group budget {mib} MiB")
    films = [_film(inp, kz) for inp in inps]
    batch = H.Batch(films, tile_kz=kz)
    batch.set_shape(kz, mib)
    assert batch.tile_depth == kz and batch.shape["groups"] == groups
    its_b, m_b = batch.run(Ez, want_m=True)
    cnt = batch.counters()
    _say("batched iterations per step", its_b.tolist(), cnt)
    assert cnt["unconverged"] == 0
    for f, inp in enumerate(inps):
        single = _film(inp, kz)
        assert single.tile_depth == kz
        its_s = single.run(Ez[:, f])
        Pb, phib, _ = films[f].get_state()
        Ps, phis, _ = single.get_state()
        assert np.array_equal(its_b[:, f], its_s), (f, its_b[:, f], its_s)
        assert np.array_equal(Pb, Ps), f
        assert np.array_equal(phib, phis), f
        # <Pmz> after the last step, from the single film's diagnostics
        d = single.diagnostics()
        assert abs(m_b[-1, f] - d[0] / inp.N) <= 1e-12 * max(1.0, abs(d[0] / inp.N))
        single.close()
    batch.close()
    for f in films:
        f.close()


@pytest.mark.parametrize("kz", [2, 8])
def test_batch_with_a_neumann_film_matches_single_film_runs(kz):
    # film 1 has Neumann electrodes, so the group runs the kernel build that
    # weights mirrored planes in the CG sums; the Dirichlet films of the same
    # launches take weight 1 there and must still match their own runs
    inps = _inputs()
    S = 10
    Ez = _fields(inps, S)
    zbc = [L.ZBC_DIRICHLET, L.ZBC_NEUMANN, L.ZBC_DIRICHLET, L.ZBC_DIRICHLET]
    _say(f"batch with a Neumann film, kz={kz}")
    films = [_film(inp, kz, z) for inp, z in zip(inps, zbc)]
    batch = H.Batch(films, tile_kz=kz)
    batch.set_shape(kz, 0.0)
    its_b = batch.run(Ez)
    assert batch.counters()["unconverged"] == 0
    for f, inp in enumerate(inps):
        single = _film(inp, kz, zbc[f])
        its_s = single.run(Ez[:, f])
        Pb, phib, _ = films[f].get_state()
        Ps, phis, _ = single.get_state()
        assert np.array_equal(its_b[:, f], its_s), (f, its_b[:, f], its_s)
        assert np.array_equal(Pb, Ps) and np.array_equal(phib, phis), f
        single.close()
    batch.close()
    for f in films:
        f.close()


def test_graph_replay_and_plain_launches_agree():
    inps = _inputs()[:3]
    S = 8
    Ez = _fields(inps, S)
    out = []
    for graphs, chunk in ((True, 16), (False, 6), (True, 2)):
        _say(f"graphs={graphs} chunk={chunk}")
        films = [_film(inp) for inp in inps]
        batch = H.Batch(films, tile_kz=2, use_graphs=graphs, chunk=chunk)
        its = batch.run(Ez)
        out.append((its, [f.get_state()[0] for f in films]))
        batch.close()
        for f in films:
            f.close()
    for its, P in out[1:]:
        assert np.array_equal(its, out[0][0])
        for a, b in zip(P, out[0][1]):
            assert np.array_equal(a, b)


def test_autotune_keeps_the_films_state_and_picks_a_depth():
    inps = _inputs()
    S = 6
    Ez = _fields(inps, S)
    films = [_film(inp) for inp in inps]
    batch = H.Batch(films, tile_kz=2)
    batch.run(Ez[:2])
    before = [f.get_state() for f in films]
    shape, ms = batch.autotune(Ez[2], reps=1)
    _say("autotune", shape, ms)
    assert shape["tile_kz"] in (2, 4, 8) and batch.tile_depth == shape["tile_kz"]
    assert len(ms) == 12 and min(ms.values()) > 0.0
    assert ms[(shape["tile_kz"], shape["group_mib"])] == min(ms.values())
    for f, (P, phi, _) in zip(films, before):
        P2, phi2, _ = f.get_state()
        assert np.array_equal(P, P2) and np.array_equal(phi, phi2)
    its = batch.run(Ez[2:])
    assert batch.counters()["unconverged"] == 0 and its.min() > 0
    batch.close()
    for f in films:
        f.close()


def test_batch_rejects_fp64_and_two_kernel_films():
    inp = _inputs()[3]
    f64 = H.Film(inp, precision=L.PREC_FP64)
    two = H.Film(inp, precision=L.PREC_FP32_REFINED, cg_fused=0)
    for bad in (f64, two):
        with pytest.raises(L.HzoError):
            H.Batch([bad])
    f64.close()
    two.close()


def test_graph_replay_is_refused_under_a_profiler_tool(monkeypatch):
    inps = _inputs()[:2]
    Ez = _fields(inps, 2)
    films = [_film(inp) for inp in inps]
    batch = H.Batch(films, tile_kz=2, use_graphs=True, chunk=16)
    monkeypatch.setenv("ROCP_TOOL_LIBRARIES", "librocprofiler-sdk-tool.so")
    with pytest.raises(L.HzoError, match="hipGraph replay is refused"):
        batch.run(Ez[:1])
    # plain launches run under the tool, and the refused call left no step behind
    L.call("hzo_batch_config", batch.h, 0, 16)
    its = batch.run(Ez[:1])
    assert its.shape == (1, 2) and np.all(its > 0)
    monkeypatch.delenv("ROCP_TOOL_LIBRARIES")
    batch.close()
    for f in films:
        f.close()


def test_resident_groups_match_their_own_runs():
    # films at automatic tile depth whose shapes the resident solve covers
    # share one cooperative launch per depth (the nz 16 group holds a Neumann
    # film, so it runs the build with mirrored-plane weights, exactly 1 on the
    # Dirichlet films); the 40-wide film is not covered and runs its own
    # streaming kernels; every film ends bit-identical to its own run
    shapes = [(64, 48, 16, 12), (32, 32, 16, 5), (48, 32, 16, 7), (32, 48, 24, 6), (40, 72, 12, 9)]
    zbc = [L.ZBC_DIRICHLET, L.ZBC_NEUMANN, L.ZBC_DIRICHLET, L.ZBC_DIRICHLET, L.ZBC_DIRICHLET]
    inps = [film_inputs(nx, ny, nz, ng, fracs=(0.3, 0.5, 0.2), seed=41 + f, jitter=0.2)
            for f, (nx, ny, nz, ng) in enumerate(shapes)]
    S = 8
    amps = [3.0, 2.0, 2.5, 1.5, 2.2]
    Ez = np.stack([inp.field_nd(H.triangle_wave(S, amps[f], S)) for f, inp in enumerate(inps)], axis=1)
    _say("resident groups: 3 films of depth 16 (one Neumann), 1 of depth 24, 1 streaming film,", S, "steps")
    films = [_film(inp, 0, z) for inp, z in zip(inps, zbc)]
    assert [f.cg_form for f in films] == [3, 3, 3, 3, 1]
    batch = H.Batch(films, tile_kz=2)
    groups = [batch.film_group(f) for f in range(len(films))]
    _say("groups", groups)
    assert [g[1:] for g in groups] == [(3, 2), (3, 2), (3, 2), (1, 2), (1, 1)]
    assert groups[0][0] == groups[1][0] == groups[2][0] != groups[3][0]
    its_b, m_b = batch.run(Ez, want_m=True)
    assert batch.counters()["unconverged"] == 0
    for f, inp in enumerate(inps):
        single = _film(inp, 0, zbc[f])
        its_s = single.run(Ez[:, f])
        Pb, phib, _ = films[f].get_state()
        Ps, phis, _ = single.get_state()
        assert np.array_equal(its_b[:, f], its_s), (f, its_b[:, f], its_s)
        assert np.array_equal(Pb, Ps) and np.array_equal(phib, phis), f
        d = single.diagnostics()
        assert abs(m_b[-1, f] - d[0] / inp.N) <= 1e-12 * max(1.0, abs(d[0] / inp.N))
        single.close()
    batch.close()
    for f in films:
        f.close()


def test_lanes_leave_every_film_in_the_same_state():
    # resident groups (depth 16 and 24), one streaming group of two films on
    # the batch kernels and one single-film streaming group, run with 1, 2
    # and 4 lanes (and 2 lanes with plain launches): the streaming groups
    # then run side by side on their own streams and host threads, and every
    # film's iterations, <Pmz> samples, P and potential are bit-identical to
    # the one-lane run
    shapes = [(64, 48, 16, 12), (32, 48, 24, 6), (40, 72, 12, 9), (96, 32, 10, 15), (32, 32, 8, 6)]
    kz = [0, 0, 2, 2, 2]
    inps = [film_inputs(nx, ny, nz, ng, fracs=(0.3, 0.5, 0.2), seed=61 + f, jitter=0.2)
            for f, (nx, ny, nz, ng) in enumerate(shapes)]
    S = 6
    amps = [3.0, 2.0, 2.5, 1.5, 2.2]
    Ez = np.stack([inp.field_nd(H.triangle_wave(S, amps[f], S)) for f, inp in enumerate(inps)], axis=1)
    out = []
    for lanes, graphs in ((1, True), (2, True), (4, True), (2, False)):
        _say(f"lanes={lanes} graphs={graphs}: 2 resident films, 3 streaming films in 2 groups, {S} steps")
        films = [_film(inp, k) for inp, k in zip(inps, kz)]
        batch = H.Batch(films, tile_kz=2, lanes=lanes, use_graphs=graphs)
        batch.set_shape(2, 1.5)
        kinds = [batch.film_group(f)[1:] for f in range(len(films))]
        assert kinds == [(1, 2), (1, 2), (2, 0), (1, 1), (2, 0)], kinds
        its, m = batch.run(Ez, want_m=True)
        cnt = batch.counters()
        assert cnt["unconverged"] == 0
        out.append((its, m, cnt, [f.get_state()[:2] for f in films]))
        batch.close()
        for f in films:
            f.close()
    its0, m0, cnt0, st0 = out[0]
    for its, m, cnt, st in out[1:]:
        assert np.array_equal(its, its0) and np.array_equal(m, m0) and cnt == cnt0
        for (P, phi), (P0, phi0) in zip(st, st0):
            assert np.array_equal(P, P0) and np.array_equal(phi, phi0)
    f = _film(inps[4], 2)
    with pytest.raises(L.HzoError):
        H.Batch([f], lanes=5)
    f.close()
