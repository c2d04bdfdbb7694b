"""ctypes binding of the CPU FP64 oracle (oracle/libhzo_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (cmsl_amd/) never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhzo_oracle.so")


class OPhase(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("alpha", "beta", "gamma", "g", "a0", "kappa", "L", "eps_r")]


class OCfg(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("del_h", C.c_double), ("del_t", C.c_double),
                ("dep_flag", C.c_int), ("zbc", C.c_int), ("eps0_nd", C.c_double),
                ("g_bot", C.c_double), ("g_top", C.c_double), ("pcg_rule", C.c_int),
                ("tol_abs", C.c_double), ("tol_rel", C.c_double), ("max_iter", C.c_int)]


class OStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("rz0", C.c_double), ("rz", C.c_double),
                ("rr0", C.c_double), ("rr", C.c_double), ("bb", C.c_double)]


_P = C.c_void_p
_SIG = {
    "hzo_oracle_tdgl": [_P, _P, _P, _P, _P, C.c_int, _P, _P, _P, _P, _P, C.c_double],
    "hzo_oracle_divergence": [_P, _P, _P],
    "hzo_oracle_eps_map": [_P, _P, C.c_int, _P, _P, _P],
    "hzo_oracle_apply_A": [_P, _P, _P, _P],
    "hzo_oracle_diag_inv": [_P, _P, _P],
    "hzo_oracle_efield": [_P, _P, _P],
    "hzo_oracle_depol": [_P, _P, _P, _P, _P, _P, _P, _P, C.c_int],
    "hzo_oracle_step": [_P, _P, _P, _P, _P, C.c_int, _P, _P, _P, _P, _P, _P, _P, C.c_double, _P],
    "hzo_oracle_run": [_P, _P, _P, _P, C.c_int, _P, _P, _P, _P, _P, _P, _P, C.c_int, _P],
    "hzo_oracle_interface_flag": [_P, _P, _P, _P],
    "hzo_oracle_diagnostics": [_P, _P, _P, _P, _P, _P],
    "hzo_oracle_voronoi_layer": [C.c_int] * 5 + [_P, _P, _P, C.c_int, C.c_int, _P],
    "hzo_oracle_reorder_zf_to_c": [C.c_int, C.c_int, C.c_int, _P, _P],
    "hzo_oracle_set_threads": [C.c_int],
    "hzo_oracle_get_threads": [],
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        l = C.CDLL(LIB_PATH)
        for k, a in _SIG.items():
            getattr(l, k).restype = None
            getattr(l, k).argtypes = a
        l.hzo_oracle_get_threads.restype = C.c_int
        _lib = l
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def cfg(nx, ny, nz, del_h, del_t, dep_flag=1, zbc=0, eps0_nd=1.0, g_bot=0.0, g_top=0.0,
        pcg_rule=0, tol_abs=1e-3, tol_rel=1e-2, max_iter=4000) -> OCfg:
    return OCfg(nx, ny, nz, del_h, del_t, dep_flag, zbc, eps0_nd, g_bot, g_top, pcg_rule,
                tol_abs, tol_rel, max_iter)


def phases(rows) -> "C.Array":
    arr = (OPhase * 3)()
    names = ("alpha", "beta", "gamma", "g", "a0", "kappa", "L", "eps_r")
    for q, r in enumerate(rows):
        items = r.items() if isinstance(r, dict) else zip(names, r)
        for k, v in items:
            setattr(arr[q], k, float(v))
    return arr


class Film:
    """Oracle-side film: holds the same inputs the GPU context receives."""

    def __init__(self, c: OCfg, omega, grain_phase, TR, ph_rows, iface=None):
        self.c = c
        self.N = c.nx * c.ny * c.nz
        self.omega = np.ascontiguousarray(omega, dtype=np.int32)
        self.grain_phase = np.ascontiguousarray(grain_phase, dtype=np.int32)
        self.ng = int(self.grain_phase.size)
        self.TR = np.ascontiguousarray(TR, dtype=np.float64)
        self.ph = phases(ph_rows)
        if iface is None:
            iface = np.zeros(self.N, np.int32)
            lib().hzo_oracle_interface_flag(C.byref(c), _p(self.omega), _p(self.grain_phase), _p(iface))
        self.iface = np.ascontiguousarray(iface, dtype=np.int32)
        self.eps = np.zeros(self.N)
        lib().hzo_oracle_eps_map(C.byref(c), _p(self.omega), self.ng, _p(self.grain_phase), self.ph,
                                 _p(self.eps))

    def tdgl(self, P, E3, Ez_ext, want_local=False):
        Pn = np.empty_like(P)
        Pl = np.zeros_like(P) if want_local else None
        lib().hzo_oracle_tdgl(C.byref(self.c), _p(P), _p(Pn), _p(Pl), _p(self.omega), self.ng,
                              _p(self.grain_phase), _p(self.TR), self.ph, _p(self.iface), _p(E3), Ez_ext)
        return (Pn, Pl) if want_local else Pn

    def depol(self, P, phi, hist=0):
        E3 = np.zeros(3 * self.N)
        st = OStats()
        divP = np.zeros(self.N)
        h = np.zeros(max(hist, 1))
        lib().hzo_oracle_depol(C.byref(self.c), _p(self.eps), _p(P), _p(phi), _p(E3), _p(divP),
                               C.byref(st), _p(h), hist)
        return E3, st, divP, h

    def run(self, P, phi, E3, Ez, want_local=False):
        Ez = np.ascontiguousarray(Ez, dtype=np.float64)
        its = np.zeros(Ez.size, np.int32)
        Pl = np.zeros_like(P) if want_local else None
        lib().hzo_oracle_run(C.byref(self.c), _p(P), _p(Pl), _p(self.omega), self.ng, _p(self.grain_phase),
                             _p(self.TR), self.ph, _p(self.iface), _p(phi), _p(E3), _p(Ez), int(Ez.size),
                             _p(its))
        return its, Pl

    def diagnostics(self, P, Plocal):
        out = np.zeros(10)
        lib().hzo_oracle_diagnostics(C.byref(self.c), _p(P), _p(Plocal), _p(self.omega),
                                     _p(self.grain_phase), _p(out))
        return out


def set_threads(n: int) -> int:
    """Threads of the oracle's OpenMP loops; returns the count now in force."""
    lib().hzo_oracle_set_threads(int(n))
    return int(lib().hzo_oracle_get_threads())


def apply_A(c: OCfg, eps, x):
    y = np.zeros_like(x)
    lib().hzo_oracle_apply_A(C.byref(c), _p(eps), _p(x), _p(y))
    return y


def diag_inv(c: OCfg, eps):
    d = np.zeros_like(eps)
    lib().hzo_oracle_diag_inv(C.byref(c), _p(eps), _p(d))
    return d


def efield(c: OCfg, phi):
    E = np.zeros(3 * phi.size)
    lib().hzo_oracle_efield(C.byref(c), _p(phi), _p(E))
    return E


def divergence(c: OCfg, P):
    d = np.zeros(c.nx * c.ny * c.nz)
    lib().hzo_oracle_divergence(C.byref(c), _p(np.ascontiguousarray(P)), _p(d))
    return d


def voronoi_layer(nx, ny, nz, z0, span, cx, cy, cz, nL, goff, omega_zf):
    lib().hzo_oracle_voronoi_layer(nx, ny, nz, z0, span, _p(cx), _p(cy), _p(cz), nL, goff, _p(omega_zf))


def reorder_zf_to_c(nx, ny, nz, a):
    out = np.empty(nx * ny * nz, np.int32)
    lib().hzo_oracle_reorder_zf_to_c(nx, ny, nz, _p(np.ascontiguousarray(a, np.int32)), _p(out))
    return out
