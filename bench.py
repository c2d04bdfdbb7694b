#!/usr/bin/env python3
"""Throughput of the HZO phase-field step on one MI355X.

Workload (the reference's two-precision comparison film): 256x256x32 voxels,
200 seeded Voronoi grains with FE/AFE/DE fractions 0.6/0.2/0.2, per-phase
Landau coefficients within +-20 % of the base values, triangular applied
field +-3 MV/cm over 2000 steps, Jacobi-PCG depolarisation solve after every
step to ||r|| <= 1e-8 ||b||.  A step = TDGL update + full PCG solve.

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NX, NY, NZ, NG = 256, 256, 32, 200
FRACS = (0.2, 0.6, 0.2)  # AFE, FE, DE
PERIOD, AMP = 2000, 3.0


def build_inputs():
    import cmsl_amd as H
    return H.make_film_inputs(NX, NY, NZ, NG, fracs=FRACS, seed=1, jitter=0.2)


def algorithmic_bytes_per_cell(kclass: str, precision: str, cg_form: int = 1) -> float:
    """Compulsory HBM bytes per voxel of one launch (each array touched once).

    fp64: pcg_dir_spmv reads r, p_old, code (8+8+2) and writes p_new (8) = 26;
          pcg_update reads x, p, r, code (8+8+8+2) and writes x, r (16) = 42;
          pcg_setup reads Pmx,Pmy,Pmz, x, code (32+2) and writes r (8) = 42;
          tdgl reads P (48), phi (8), cell word (4) and writes P (48) = 108.
    fp32: pcg_dir_spmv reads r, p_old (4+4), code (2), writes p_new (4) = 14;
          pcg_update reads x_lo, p, r (12), code (2), writes x_lo, r (8) = 22;
          pcg_setup reads Pmx,Pmy,Pmz (12), x_hi, x_lo (12), code (2), writes b (8), r (4) = 38;
          tdgl reads P (24), x_hi + x_lo (12), cell word (4), writes P (24) = 64;
          pcg_fused (single-reduction CG, one launch per iteration) carrying w and
          s reads r, w, s, p (16), code (2), writes s, r, w, p (16) = 34, and
          every second iteration also reads and writes x_lo (8, two iterations'
          alpha p at once): 38 on average; forming s = A p and w = A u by
          stencils (cg_form 2) it reads r, p (8), code (2), writes r, p (8) = 18,
          with x_lo likewise: 22 on average.
    """
    if kclass == "pcg_fused" and precision == "fp32" and cg_form == 2:
        return 22.0
    table = {
        "fp64": {"tdgl": 108.0, "pcg_setup": 42.0, "pcg_dir_spmv": 26.0, "pcg_update": 42.0, "pcg_fused": 0.0},
        "fp32": {"tdgl": 64.0, "pcg_setup": 38.0, "pcg_dir_spmv": 14.0, "pcg_update": 22.0, "pcg_fused": 38.0},
    }
    return table[precision][kclass]


def run_gpu(args, inp, precision):
    import cmsl_amd as H
    from cmsl_amd import _lib as L

    prec = L.PREC_FP64 if precision == "fp64" else L.PREC_FP32_REFINED
    kz = args.tile_kz if (precision == "fp32" or args.tile_kz != 16) else 0
    film = H.Film(inp, precision=prec, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30,
                  refine_every=20, tile_kz=kz, cg_fused=args.cg_fused if precision == "fp32" else 1)
    Ez = inp.field_nd(H.triangle_wave(PERIOD, AMP, args.warmup + args.steps))
    film.run(Ez[: args.warmup])
    film.synchronize()
    t0 = time.perf_counter()
    its = film.run(Ez[args.warmup:])
    film.synchronize()
    t1 = time.perf_counter()
    # kernel timing pass: the same steps again, each launch bracketed by HIP
    # events on the film's stream
    film.set_timing(True)
    Ez2 = inp.field_nd(H.triangle_wave(PERIOD, AMP, args.timing_steps, start=args.warmup + args.steps))
    its2 = film.run(Ez2)
    kt = film.kernel_times()
    cnt = film.cg_counters()
    film.set_timing(False)
    form = film.cg_form
    film.close()
    return t1 - t0, its, kt, form, its2, cnt


def cg_iteration_roofline(kt, its_timed, counters, form, N, precision="fp32"):
    """Time and algorithmic bytes of one CG iteration of the FP32-refined
    solve, amortizing every launch or pass of the solver class over the
    iterations of the timing pass (residual replacements, re-formations and
    the launches that only add a deferred x_lo term included).  The bytes are
    those the iteration's vectors need when streamed from HBM (38 per voxel
    for the carried single-reduction CG, 22 for its two-stencil form); the
    resident form holds them in registers and LDS and streams only edge
    columns and partial sums, so its fraction says how fast it does the
    iteration relative to a kernel that would have to move them."""
    if form == 3:
        cls, iters, bpc = "pcg_resident", counters[1], 38.0
    else:
        cls, iters = "pcg_fused", int(np.sum(its_timed))
        bpc = algorithmic_bytes_per_cell("pcg_fused", precision, 2 if form == 2 else 1)
    c, ms = kt[cls]
    ipl = iters / max(c, 1)  # iterations per launch
    bytes_per_launch = bpc * N * ipl
    achieved = bytes_per_launch / (ms * 1e-3) / 1e9
    return {"kernel": cls, "achieved": achieved, "frac": achieved / HBM_PEAK_GBS, "bytes_per_cell": bpc,
            "launch_ms": ms, "bytes_per_launch": bytes_per_launch, "iterations_per_launch": ipl,
            "ms_per_iteration": ms / max(ipl, 1e-30), "launches_timed": c, "iterations_timed": iters}


def measured_traffic(kernel, ipl):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, calibrated), rescaled from the
    profiled run's iterations per dispatch to `ipl`; None without it."""
    prof = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(prof):
        return None
    with open(prof) as fh:
        pm = json.load(fh).get("fp32", {})
    per_dispatch, its = pm.get(kernel), pm.get(kernel + "_iterations_per_dispatch")
    if per_dispatch is None or not its:
        return None
    return per_dispatch / its * ipl


def run_cpu_baseline(inp, Ez, threads):
    """The FP64 oracle (a restatement of the reference kernels; the reference
    itself is CUDA-only) on the same film and fields, on `threads` host
    threads; returns (seconds, PCG iterations per step)."""
    from oracle import pyoracle as O

    O.set_threads(threads)
    c = O.cfg(inp.nx, inp.ny, inp.nz, inp.del_h, inp.del_t, dep_flag=1, zbc=0,
              eps0_nd=inp.scales["eps0_nd"], pcg_rule=1, tol_abs=1e-30, tol_rel=1e-8)
    f = O.Film(c, inp.omega, inp.grain_phase, inp.TR, inp.phases, inp.iface)
    P = inp.P0.copy()
    phi = np.zeros(inp.N)
    E3 = np.zeros(3 * inp.N)
    t0 = time.perf_counter()
    its, _ = f.run(P, phi, E3, np.ascontiguousarray(Ez, np.float64))
    dt = time.perf_counter() - t0
    return dt, its


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baselines(inp, args):
    """The oracle on the host cores the process may use (OMP_NUM_THREADS when
    set: the GPU box allots 16 cores per GPU; else every core) and on one
    core: a bounded sample of the 256x256x32 workload (its first steps) and
    the whole 64x64x16, 8-grain, 1 MV/cm, 200-step single film."""
    import cmsl_amd as H

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    N = inp.N
    Ez1 = inp.field_nd(H.triangle_wave(PERIOD, AMP, max(args.cpu_steps, args.cpu_single_steps)))
    dt, its = run_cpu_baseline(inp, Ez1[: args.cpu_steps], threads)
    dt1, its1 = run_cpu_baseline(inp, Ez1[: args.cpu_single_steps], 1)
    inp0 = H.make_film_inputs(64, 64, 16, 8, fracs=(0.0, 1.0, 0.0), seed=3)
    Ez0 = np.full(200, float(inp0.field_nd(1.0)))
    dt0, its0 = run_cpu_baseline(inp0, Ez0, threads)
    dt01, _ = run_cpu_baseline(inp0, Ez0, 1)
    return {"value": N * args.cpu_steps / dt / 1e6, "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle, first {args.cpu_steps} steps of the {inp.nx}x{inp.ny}x{inp.nz} film and field, "
                      f"{int(np.sum(its))} PCG iterations, {dt:.1f} s on {threads} threads",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "single_core": {"value": N * args.cpu_single_steps / dt1 / 1e6, "unit": "Mcell-steps/s", "cores": 1,
                            "sample": f"first {args.cpu_single_steps} step(s), {int(np.sum(its1))} PCG "
                                      f"iterations, {dt1:.1f} s"},
            "single_film_64x64x16": {"workload": "64x64x16, 8 FE grains, 1 MV/cm, 200 steps, PCG ||r||<=1e-8||b||",
                                     "value": inp0.N * 200 / dt0 / 1e6, "cores": threads,
                                     "value_single_core": inp0.N * 200 / dt01 / 1e6,
                                     "pcg_iters_mean": float(np.mean(its0)), "s": dt0, "s_single_core": dt01}}


def loop_observables(E, m):
    """Remanent <Pmz> (descending branch at E = 0) and coercive field
    (descending branch, <Pmz> = 0, linear interpolation) of a sampled loop."""
    n = len(E)
    pr, ec = None, None
    for i in range(n // 4, 3 * n // 4):
        if E[i] >= 0.0 > E[i + 1] or E[i] == 0.0:
            f = 0.0 if E[i] == 0.0 else E[i] / (E[i] - E[i + 1])
            pr = m[i] + f * (m[i + 1] - m[i])
            break
    for i in range(n // 4, 3 * n // 4):
        if m[i] >= 0.0 > m[i + 1]:
            f = m[i] / (m[i] - m[i + 1])
            ec = E[i] + f * (E[i + 1] - E[i])
            break
    return pr, ec


def run_loop_errors(inp, every=10):
    """One full period of the triangle field at 256x256x32 in both precisions,
    <Pmz> sampled every `every` steps: errors of the FP32-refined loop, final P,
    phi, remanent polarisation and coercive field against the FP64 run (the
    FP64 GPU path is held to the oracle within 1e-10 by the parity tests)."""
    import cmsl_amd as H
    from cmsl_amd import _lib as L

    Ez = inp.field_nd(H.triangle_wave(PERIOD, AMP, PERIOD))
    E_MV = H.triangle_wave(PERIOD, AMP, PERIOD)
    out = {}
    for name, prec in (("fp64", L.PREC_FP64), ("fp32", L.PREC_FP32_REFINED)):
        film = H.Film(inp, precision=prec, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30,
                      refine_every=20)
        ms, Es = [], []
        for s0 in range(0, PERIOD, every):
            film.run(Ez[s0:s0 + every])
            ms.append(film.diagnostics()[0] / inp.N)
            Es.append(E_MV[s0 + every - 1])
        P, phi, _ = film.get_state()
        film.close()
        out[name] = (np.array(ms), np.array(Es), P, phi)
    m64, E, P64, phi64 = out["fp64"]
    m32, _, P32, phi32 = out["fp32"]
    pr64, ec64 = loop_observables(E, m64)
    pr32, ec32 = loop_observables(E, m32)
    rel = lambda a, b: float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    return {"steps": PERIOD, "sample_every": every,
            "rel_err_P_final": rel(P32, P64), "rel_err_phi_final": rel(phi32, phi64),
            "rel_err_loop_max": rel(m32, m64),
            "remanent_Pmz_fp64": pr64, "remanent_Pmz_fp32": pr32,
            "rel_err_remanent": None if pr64 in (None, 0.0) else abs(pr32 - pr64) / abs(pr64),
            "coercive_field_MVcm_fp64": ec64, "coercive_field_MVcm_fp32": ec32,
            "rel_err_coercive": None if ec64 in (None, 0.0) or ec32 is None else abs(ec32 - ec64) / abs(ec64)}


def run_adjoint(args):
    """Reverse-mode cost on the differentiable workload: 128x128x24 film, 100
    grains FE/AFE 0.7/0.3, 500-step triangle +-3 MV/cm, FP32-refined with FP64
    reductions, checkpoints every 50 steps.  Loss: MSE between the loop and a
    target loop simulated with hidden seeded coefficients (+-10 %)."""
    import cmsl_amd as H
    from cmsl_amd import _lib as L
    from cmsl_amd.adjoint import coeff_matrix, rows_with

    S, K = args.adj_steps, 50
    inp = H.make_film_inputs(128, 128, 24, 100, fracs=(0.3, 0.7, 0.0), seed=5, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(S, AMP, S))
    film = H.Film(inp, precision=L.PREC_FP32_REFINED, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8,
                  tol_abs=1e-30, refine_every=20)
    film.adjoint_reserve(S, K, 1e-8)
    base = coeff_matrix(inp.phases)
    rng = np.random.default_rng(12345)
    hidden = base * (1.0 + 0.1 * rng.uniform(-1.0, 1.0, base.shape))
    film.set_phase_table(rows_with(inp.phases, hidden))
    film.set_state(inp.P0)
    target = film.forward_record(Ez)
    film.set_phase_table(rows_with(inp.phases, base))
    film.set_state(inp.P0)
    film.run(Ez[:10])  # warm-up
    film.set_state(inp.P0)
    film.synchronize()
    t0 = time.perf_counter()
    its = film.run(Ez)
    film.synchronize()
    t_plain = time.perf_counter() - t0
    film.set_state(inp.P0)
    film.synchronize()
    t0 = time.perf_counter()
    m = film.forward_record(Ez)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g, adj_its = film.backward(2.0 * (m - target) / S)
    t_bwd = time.perf_counter() - t0
    film.close()
    return {"workload": "128x128x24 film, 100 grains FE/AFE 0.7/0.3, %d-step triangle +-3 MV/cm, FP32-refined, "
                        "checkpoints every %d steps" % (S, K),
            "t_forward_s": t_fwd, "t_backward_incl_recompute_s": t_bwd, "t_plain_run_s": t_plain,
            "adjoint_fwd_cost": t_bwd / t_fwd, "loss": float(np.mean((m - target) ** 2)),
            "fwd_pcg_iters_mean": float(np.mean(its)), "adj_pcg_iters_mean": adj_its / max(S - 1, 1),
            "grad_norm": float(np.linalg.norm(g))}


def run_batched(args):
    """Batched workload: args.batch_films films of uneven sizes, grain counts
    and Landau coefficients (+-20 %), FP32-refined, sinusoidal +-3 MV/cm over
    1000 steps; all films advance together (resident groups of small films
    in one cooperative launch each, the other films in cache-sized groups
    whose steps replay captured hipGraphs, shared over args.batch_lanes host
    threads with a stream each; tile depth and group budget picked by the
    autotuner unless args.batch_shape names them).  Times args.batch_steps
    steps of the batch, then (unless args.batch_no_seq) the same number of
    following steps with the films stepped one by one."""
    from concurrent.futures import ThreadPoolExecutor

    import cmsl_amd as H
    from cmsl_amd import _lib as L

    specs = H.film.batch_film_specs(args.batch_films)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=8) as ex:
        inps = list(ex.map(lambda s: H.make_film_inputs(s[0], s[0], s[1], s[2], fracs=FRACS, seed=s[3],
                                                        jitter=0.2), specs))
    t_inputs = time.perf_counter() - t0
    films = [H.Film(inp, precision=L.PREC_FP32_REFINED, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8,
                    tol_abs=1e-30, refine_every=20) for inp in inps]
    ntot = sum(inp.N for inp in inps)
    W, S = args.batch_warmup, args.batch_steps
    wave = AMP * np.sin(2.0 * np.pi * np.arange(W + 2 * S) / 1000.0)
    Ez = np.stack([inp.field_nd(wave) for inp in inps], axis=1)  # (steps, films)
    # under a profiler tool the library refuses graph replay; the plain
    # launches run the same kernels
    batch = H.Batch(films, tile_kz=2, use_graphs=not os.environ.get("ROCP_TOOL_LIBRARIES"), lanes=args.batch_lanes)
    if args.batch_shape:
        kz_b, mib_b = (float(v) for v in args.batch_shape.split(","))
        batch.set_shape(int(kz_b), mib_b)
        shape, tune_ms = batch.shape, {}
    else:
        shape, tune_ms = batch.autotune(Ez[0], reps=2)
    batch.run(Ez[:W])
    cnt0 = batch.counters()
    t0 = time.perf_counter()
    its = batch.run(Ez[W:W + S])
    dt_batch = time.perf_counter() - t0
    cnt = batch.counters()
    cnt = {k: cnt[k] - cnt0[k] for k in cnt}
    batch.close()
    its_seq, dt_seq = None, None
    if not args.batch_no_seq:
        t0 = time.perf_counter()
        its_seq = np.stack([f.run(Ez[W + S:W + 2 * S, q]) for q, f in enumerate(films)], axis=1)
        for f in films:
            f.synchronize()
        dt_seq = time.perf_counter() - t0
    # algorithmic bytes: TDGL + PCG setup per voxel, plus one fused CG
    # launch per iteration
    per_vox = algorithmic_bytes_per_cell("tdgl", "fp32") + algorithmic_bytes_per_cell("pcg_setup", "fp32")
    nvec = np.array([inp.N for inp in inps], np.float64)
    bytes_batch = float(np.sum(nvec * (S * per_vox + its.sum(axis=0) * algorithmic_bytes_per_cell("pcg_fused", "fp32"))))
    for f in films:
        f.close()
    sizes = {}
    for nx, nz, _, _ in specs:
        sizes[f"{nx}x{nx}x{nz}"] = sizes.get(f"{nx}x{nx}x{nz}", 0) + 1
    return {"workload": "%d films, nx=ny in {128,192,256,384}, nz in {16,24,32,48}, 50-300 grains, +-20%% Landau, "
                        "FP32-refined, sinusoid +-3 MV/cm / 1000 steps; %d warm-up + %d timed steps of the batch, "
                        "then the next %d steps film by film" % (args.batch_films, W, S, S),
            "films": args.batch_films, "cells_total": ntot, "film_shapes": sizes,
            "value": ntot * S / dt_batch / 1e6, "unit": "Mcell-steps/s",
            "ms_per_step": dt_batch / S * 1e3,
            "sequential_value": ntot * S / dt_seq / 1e6 if dt_seq else None,
            "speedup_vs_sequential": dt_seq / dt_batch if dt_seq else None,
            "pcg_iters_mean": float(its.mean()),
            "pcg_iters_mean_sequential": float(its_seq.mean()) if its_seq is not None else None,
            "algorithmic_GBs": bytes_batch / dt_batch / 1e9,
            "shape": shape, "lanes": args.batch_lanes,
            "autotune_ms_per_step": {f"kz{k}_group{m}MiB": v for (k, m), v in tune_ms.items()},
            "cg_launches_timed": cnt["cg_launches"],
            "unconverged_solves": cnt["unconverged"], "t_inputs_s": t_inputs}


def run_large(args):
    """Large-film workload, the pure-HBM case: 512x512x64 film (16.8 M
    cells), 1000 grains FE/AFE/DE 0.5/0.3/0.2, 1 % of the cells
    (grain-boundary cells) dielectric, +-20 % Landau coefficients, triangle
    +-4 MV/cm over 5000 steps; args.large_steps timed steps after 2 warm-up
    steps in each precision, then one step with kernel timing."""
    import cmsl_amd as H
    from cmsl_amd import _lib as L

    inp = H.make_film_inputs(512, 512, 64, 1000, fracs=(0.3, 0.5, 0.2), seed=41, jitter=0.2, label_device=0,
                             gb_dielectric=0.01)
    S, W = args.large_steps, 2
    Ez = inp.field_nd(H.triangle_wave(5000, 4.0, W + S + 1))
    out = {"workload": "512x512x64 film, 1000 grains FE/AFE/DE 0.5/0.3/0.2, 1%% dielectric grain-boundary cells, "
                       "+-20%% Landau, triangle +-4 MV/cm / 5000 steps, PCG ||r||<=1e-8||b||; "
                       "%d warm-up + %d timed steps" % (W, S)}
    for name, prec in (("fp32", L.PREC_FP32_REFINED), ("fp64", L.PREC_FP64)):
        film = H.Film(inp, precision=prec, pcg_rule=L.PCG_RULE_RELRES, tol_rel=1e-8, tol_abs=1e-30,
                      refine_every=20)
        film.run(Ez[:W])
        film.synchronize()
        t0 = time.perf_counter()
        its = film.run(Ez[W:W + S])
        film.synchronize()
        dt = time.perf_counter() - t0
        film.set_timing(True)
        its2 = film.run(Ez[W + S:])
        kt = film.kernel_times()
        cnt = film.cg_counters()
        form, kz = film.cg_form, film.tile_depth
        film.close()
        entry = {"value": inp.N * S / dt / 1e6, "unit": "Mcell-steps/s", "ms_per_step": dt / S * 1e3,
                 "pcg_iters_mean": float(np.mean(its)), "kernel_ms": {k: v[1] for k, v in kt.items()},
                 "tile_kz": kz, "cg_form": form}
        if name == "fp32":
            rl = cg_iteration_roofline(kt, its2, cnt, form, inp.N)
            entry["roofline_cg"] = dict(rl, peak=HBM_PEAK_GBS, unit="GB/s")
        out[name] = entry
    out["speedup_fp32_refined_vs_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--timing-steps", type=int, default=4)
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="steps of the 256x256x32 film timed on the oracle on all allotted cores")
    ap.add_argument("--cpu-single-steps", type=int, default=1, help="the same on one core")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-kz", type=int, default=0, help="z planes per workgroup (0: automatic)")
    ap.add_argument("--cg-fused", type=int, default=1,
                    help="FP32-refined solver: 1 automatic, 2 carried, 3 two-stencil, 4 resident, 0 two-kernel")
    ap.add_argument("--adj-steps", type=int, default=500, help="steps of the reverse-mode workload")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--batch-films", type=int, default=64, help="films of the batched workload (0: skip it)")
    ap.add_argument("--large-steps", type=int, default=4,
                    help="timed steps of the 512x512x64 large-film workload (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=8)
    ap.add_argument("--batch-warmup", type=int, default=2)
    ap.add_argument("--batch-lanes", type=int, default=2,
                    help="host threads (own streams) that run the batch's streaming groups side by side")
    ap.add_argument("--batch-shape", default="",
                    help="'KZ,MIB': tile depth and group budget of the batch instead of the autotuner's choice")
    ap.add_argument("--batch-no-seq", action="store_true", help="skip the film-by-film leg of the batched workload")
    ap.add_argument("--no-loop", action="store_true", help="skip the full-loop FP32-vs-FP64 error run")
    ap.add_argument("--stream-probe", action="store_true",
                    help="only time a plain kernel with the single-reduction CG's per-voxel traffic "
                         "(42 B) at 2^21 and 2^24 voxels: the attainable rate beside the CG kernel")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a device-to-device copy of 1 GiB three times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the film fits one MI355X); --gpus {args.gpus} "
              "is not supported", file=sys.stderr)
        sys.exit(2)
    if args.stream_probe or args.copy_probe:
        import ctypes as C
        probe = C.CDLL(os.path.join(ROOT, "cmsl_amd", "libhzo_probe.so"))
        probe.hzo_probe_stream.argtypes = [C.c_longlong, C.c_int, C.c_int, C.c_void_p]
        probe.hzo_probe_copy.argtypes = [C.c_longlong, C.c_int]
        probe.hzo_probe_last_error.restype = C.c_char_p
    if args.stream_probe:
        res = {}
        for n in (1 << 21, 1 << 24):
            for grid in (512, 1024, 2048, 4096):
                us = C.c_double(0.0)
                if probe.hzo_probe_stream(n, grid, 200, C.byref(us)) != 0:
                    raise RuntimeError(probe.hzo_probe_last_error().decode())
                res[f"n{n}_grid{grid}"] = {"us": us.value, "TBs": 42.0 * n / (us.value * 1e-6) / 1e12}
        print(json.dumps({"stream_probe_bytes_per_element": 42, "results": res}))
        return
    if args.copy_probe:
        n = 1 << 30  # larger than the 256 MiB last-level cache: every byte crosses HBM
        if probe.hzo_probe_copy(n, 3) != 0:
            raise RuntimeError(probe.hzo_probe_last_error().decode())
        print(json.dumps({"copy_probe_bytes_per_launch": n, "launches": 3}))
        return

    inp = build_inputs()
    N = inp.N
    res = {}
    for prec in ("fp64", "fp32"):
        dt, its, kt, form, its2, cnt = run_gpu(args, inp, prec)
        res[prec] = dict(dt=dt, its=its, kt=kt, form=form, value=N * args.steps / dt / 1e6, its2=its2, cnt=cnt)
    head = res["fp32"]
    value = head["value"]
    kt = head["kt"]
    rl = cg_iteration_roofline(kt, head["its2"], head["cnt"], head["form"], N)
    bpc = rl["bytes_per_cell"]
    achieved = rl["achieved"]
    traffic = measured_traffic(rl["kernel"], rl["iterations_per_launch"])
    # The same launch priced at the bytes of an unfused three-pass PCG
    # iteration (SpMV: read p and eps, write q; update: read x, p, r, q, D^-1,
    # write x, r; direction: read r, D^-1, p, write p: 14 FP32 words per
    # cell), and the timed steps at that PCG's bytes per cell-step (100 +
    # 56 N_it + 44 per FP64 residual refinement every 20 iterations).
    # Fractions above 1 mean the fused and resident kernels move fewer bytes
    # than the unfused iteration would, not that HBM ran past its peak.
    unf_bpc = 14.0 * 4.0
    unf_achieved = achieved * unf_bpc / bpc
    step_bytes = sum(100.0 + 56.0 * n + 44.0 * math.ceil(n / 20.0) for n in head["its"])
    unfused = {"bytes_per_cell_iteration": unf_bpc, "achieved": unf_achieved, "frac": unf_achieved / HBM_PEAK_GBS,
               "step_bytes_per_cell": step_bytes / len(head["its"]),
               "step_achieved": step_bytes * N / head["dt"] / 1e9,
               "step_frac": step_bytes * N / head["dt"] / 1e9 / HBM_PEAK_GBS}

    adj = None if args.no_adjoint else run_adjoint(args)
    loop = None if args.no_loop else run_loop_errors(inp)
    bat = run_batched(args) if args.batch_films > 0 else None
    large = run_large(args) if args.large_steps > 0 else None

    cpu = None if args.no_cpu else cpu_baselines(inp, args)

    line = {
        "metric": "Mcell-steps/s incl. PCG at 256x256x32: FP32-refined vs FP64; adjoint/fwd cost",
        "value": value,
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["dt"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Voronoi film, seed 1)",
        "config": {"workload": "256x256x32 film, 200 grains FE/AFE/DE 0.6/0.2/0.2, +-20% Landau, "
                               "triangle +-3 MV/cm / 2000 steps, PCG ||r||<=1e-8||b||",
                   "nx": NX, "ny": NY, "nz": NZ, "grains": NG,
                   "precision": "FP32 stencil+CG vectors, per-thread FP32 partial sums, FP64 global inner products, FP64 residual every 20 its, "
                                "single-reduction CG (cg_form 3: one register-resident launch per solve; "
                                "1/2: one streaming launch per iteration)"},
        "pcg_iters_mean": float(np.mean(head["its"])),
        "fp64_plain": {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["dt"] / args.steps * 1e3,
                       "pcg_iters_mean": float(np.mean(res["fp64"]["its"])),
                       "kernel_ms": {k: v[1] for k, v in res["fp64"]["kt"].items()}},
        "speedup_fp32_refined_vs_fp64": value / res["fp64"]["value"],
        "adjoint_fwd_cost": None if adj is None else adj["adjoint_fwd_cost"],
        "adjoint": adj,
        "fp32_vs_fp64_loop": loop,
        "batched": bat,
        "large_film": large,
        "roofline": {"bound": "hbm", "kernel": rl["kernel"], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_cell": bpc, "bytes_per_launch": rl["bytes_per_launch"],
                     "launch_ms": rl["launch_ms"], "iterations_per_launch": rl["iterations_per_launch"],
                     "ms_per_iteration": rl["ms_per_iteration"],
                     "launches_timed": rl["launches_timed"], "iterations_timed": rl["iterations_timed"],
                     "per": "launch; algorithmic bytes = bytes_per_cell x voxels x CG iterations of the "
                            "launch (replacement and re-formation passes amortized)",
                     "unfused_pcg": unfused},
        "cg_form": head["form"],
        "kernel_ms": {k: v[1] for k, v in kt.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
