"""Shared builders for the tests: small seeded films and their oracle twins."""
from __future__ import annotations

import numpy as np

import cmsl_amd as H
from cmsl_amd import _lib as L
from oracle import pyoracle as O


def film_inputs(nx, ny, nz, ng, fracs=(0.0, 1.0, 0.0), seed=3, jitter=0.0, del_h=0.5, del_t=0.01):
    return H.make_film_inputs(nx, ny, nz, ng, fracs=fracs, seed=seed, jitter=jitter, del_h=del_h,
                              del_t=del_t)


def oracle_film(inp, dep_flag=1, zbc=0, pcg_rule=0, tol_rel=1e-2, tol_abs=1e-3, max_iter=4000, g_bot=0.0,
                g_top=0.0):
    c = O.cfg(inp.nx, inp.ny, inp.nz, inp.del_h, inp.del_t, dep_flag=dep_flag, zbc=zbc,
              eps0_nd=inp.scales["eps0_nd"], pcg_rule=pcg_rule, tol_abs=tol_abs, tol_rel=tol_rel,
              max_iter=max_iter, g_bot=g_bot, g_top=g_top)
    return O.Film(c, inp.omega, inp.grain_phase, inp.TR, inp.phases, inp.iface)


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)


def uniform_grid_case(nx, ny, nz, phase=L.HZO_FE, ng=1, rows=None, TR=None, eps0=1.0):
    """Single-grain film with explicit coefficients (for closed-form checks)."""
    omega = np.zeros(nx * ny * nz, np.int32)
    gphase = np.full(ng, phase, np.int32)
    TR = np.tile(np.eye(3).reshape(-1), ng) if TR is None else np.asarray(TR, np.float64)
    return omega, gphase, TR
