"""ctypes binding of libhzo.so (the C ABI declared in include/hzo.h).

The shared library is built in-tree (``make -C cmsl_amd``).  Loading fails
loudly when it is missing: there is no CPU fallback for any compute call.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhzo.so")

HZO_OK = 0
HZO_ERR_ARG = -1
HZO_ERR_HIP = -2
HZO_ERR_IO = -3
HZO_ERR_STATE = -4
HZO_ERR_NOMEM = -5
HZO_ERR_DATA = -6
HZO_AFE, HZO_FE, HZO_DE = 0, 1, 2
ZBC_DIRICHLET, ZBC_NEUMANN = 0, 1
PREC_FP64, PREC_FP32_REFINED = 0, 1
PCG_RULE_REFERENCE, PCG_RULE_RELRES = 0, 1
# hzo_pcg_stats.status
HZO_PCG_CONVERGED, HZO_PCG_MAXITER = 0, 1
HZO_PCG_BAD_INITIAL, HZO_PCG_BREAKDOWN, HZO_PCG_BAD_RZ = -1, -2, -3
KCLASSES = 6
KCLASS_NAMES = ("tdgl", "pcg_setup", "pcg_dir_spmv", "pcg_update", "pcg_fused", "pcg_resident")


class PhaseND(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("alpha", "beta", "gamma", "g", "a0", "kappa", "L", "eps_r")]


class PhaseRaw(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double), ("g", C.c_double),
                ("a0", C.c_double), ("transition_type", C.c_int), ("L", C.c_double),
                ("kappa", C.c_double), ("eps_r", C.c_double)]


class Config(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("del_h", C.c_double), ("del_t", C.c_double),
                ("dep_flag", C.c_int), ("zbc", C.c_int),
                ("eps0_nd", C.c_double), ("g_bot", C.c_double), ("g_top", C.c_double),
                ("precision", C.c_int), ("pcg_rule", C.c_int),
                ("tol_abs", C.c_double), ("tol_rel", C.c_double),
                ("max_iter", C.c_int), ("refine_every", C.c_int), ("tile_kz", C.c_int),
                ("cg_fused", C.c_int)]


class PcgStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("refinements", C.c_int), ("reserved", C.c_int),
                ("rz0", C.c_double), ("rz", C.c_double), ("rr0", C.c_double), ("rr", C.c_double),
                ("bb", C.c_double)]


class SweepParams(C.Structure):
    _fields_ = [("maxEz", C.c_double), ("delEz", C.c_double), ("num_steps", C.c_int), ("max_levels", C.c_int),
                ("Ez_scale", C.c_double), ("P0_scale", C.c_double)]


class LandauRow(C.Structure):
    _fields_ = [("afe", C.c_double * 5), ("fe", C.c_double * 5), ("de", C.c_double * 2),
                ("fracs", C.c_double * 3), ("ref_type", C.c_int), ("ncols", C.c_int)]


class Deck(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("device_flag", "nx", "ny", "nz", "dep_flag", "num_steps", "out_flag",
                                      "dimx", "dimy", "dimz", "do_pre_relax", "n_layers")] + \
               [("layer_counts", C.c_int * 16)] + \
               [(n, C.c_double) for n in ("del_h", "del_t", "delEap", "Eap", "PsNoise", "G0", "omg",
                                         "kappa_fe", "kappa_afe", "kappa_de", "L_fe", "L_afe", "L_de",
                                         "eps_fe", "eps_afe", "eps_de")] + \
               [("crystal_type", C.c_char * 20), ("z_bc_type", C.c_char * 20)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double

# name -> (restype, argtypes); this table is also the list of exported symbols
# that include/hzo.h declares.
SIGNATURES = {
    "hzo_last_error": (C.c_char_p, []),
    "hzo_version": (_I, []),
    "hzo_device_count": (_I, [_P]),
    "hzo_config_default": (None, [_P]),
    "hzo_create": (_I, [_P, _P]),
    "hzo_destroy": (None, [_P]),
    "hzo_set_microstructure": (_I, [_P, _P, _I, _P, _P, _P]),
    "hzo_set_phase_table": (_I, [_P, _P]),
    "hzo_set_state": (_I, [_P, _P, _P]),
    "hzo_get_state": (_I, [_P, _P, _P, _P]),
    "hzo_tdgl_step": (_I, [_P, _D]),
    "hzo_depol_solve": (_I, [_P, _P]),
    "hzo_step": (_I, [_P, _D, _P]),
    "hzo_run": (_I, [_P, _P, _I, _P]),
    "hzo_diagnostics": (_I, [_P, _P]),
    "hzo_synchronize": (_I, [_P]),
    "hzo_grid_dims": (_I, [_P, _P]),
    "hzo_tile_depth": (_I, [_P]),
    "hzo_cg_form": (_I, [_P]),
    "hzo_sweep_schedule": (_I, [_P, _I, _P, _P, _P]),
    "hzo_sweep_run": (_I, [_P, _P, C.c_char_p, _I, _P, _P]),
    "hzo_set_timing": (_I, [_P, _I]),
    "hzo_kernel_times": (_I, [_P, _P, _P]),
    "hzo_cg_counters": (_I, [_P, _P]),
    "hzo_adjoint_reserve": (_I, [_P, _I, _I, _D]),
    "hzo_forward_record": (_I, [_P, _P, _I, _P]),
    "hzo_backward": (_I, [_P, _P, _P, _P]),
    "hzo_batch_create": (_I, [_P, _I, _I, _P]),
    "hzo_batch_destroy": (None, [_P]),
    "hzo_batch_config": (_I, [_P, _I, _I]),
    "hzo_batch_set_lanes": (_I, [_P, _I]),
    "hzo_batch_tile_depth": (_I, [_P]),
    "hzo_batch_film_group": (_I, [_P, _I, _P]),
    "hzo_batch_set_shape": (_I, [_P, _I, _D]),
    "hzo_batch_shape": (_I, [_P, _P]),
    "hzo_batch_run": (_I, [_P, _P, _I, _P, _P]),
    "hzo_batch_autotune": (_I, [_P, _P, _I, _P, _P]),
    "hzo_batch_counters": (_I, [_P, _P]),
    "hzo_save_snapshot": (_I, [_P, C.c_char_p, C.c_char_p, _D]),
    "hzo_checkpoint_save": (_I, [_P, C.c_char_p, _I, _I, _D, _D]),
    "hzo_checkpoint_load": (_I, [_P, C.c_char_p, _P, _P, _P, _P, _P]),
    "hzo_write_vector_field": (_I, [C.c_char_p, C.c_longlong, _P, _P, _P, _D]),
    "hzo_write_grain_map": (_I, [C.c_char_p, _I, _I, _I, _P, _D]),
    "hzo_write_midplane_slice": (_I, [C.c_char_p, _I, _I, _I, _P]),
    "hzo_write_checkpoint": (_I, [C.c_char_p, _I, _I, _I, _I, _I, _D, _D, _P, _P]),
    "hzo_read_checkpoint": (_I, [C.c_char_p, _I, _I, _I, _P, _P, _P, _P, _P, _P]),
    "hzo_write_raw_f64": (_I, [C.c_char_p, C.c_longlong, _P]),
    "hzo_read_deck": (_I, [C.c_char_p, _P]),
    "hzo_read_landau_table": (_I, [C.c_char_p, _I, _P, _P, _P]),
    "hzo_landau_to_raw": (_I, [_P, _P]),
    "hzo_read_grain_count": (_I, [C.c_char_p, _P]),
    "hzo_read_gdata": (_I, [C.c_char_p, _I, _I, _I, _P, _P]),
    "hzo_read_gchar": (_I, [C.c_char_p, _I, _P, _P]),
    "hzo_euler_to_TR": (_I, [_I, _P, _P]),
    "hzo_write_gdata": (_I, [C.c_char_p, _I, _I, _I, _P]),
    "hzo_count_structures": (_I, [C.c_char_p, _P]),
    "hzo_grain_boundary_cells": (_I, [_I, _I, _I, _P, _I, _D, C.c_uint64, _P, _P]),
    "hzo_write_gchar": (_I, [C.c_char_p, _I, _P, _P]),
    "hzo_voronoi": (_I, [_I, _I, _I, _I, _P, _I, C.c_uint64, _P, _P, _P]),
    "hzo_voronoi_gpu": (_I, [_I, _I, _I, _I, _I, _P, _I, C.c_uint64, _P, _P, _P]),
    "hzo_voronoi_gpu_noalloc": (_I, [_I, _I, _I, _I, _I, _P, _I, C.c_uint64, _P, _P, _P]),
    "hzo_nondim": (_I, [_P, _I, _P, _P]),
    "hzo_assign_phases": (_I, [_I, _I, _I, _P, _I, _P, _I, _P, _P, C.c_uint, _P]),
    "hzo_interface_flag": (_I, [_I, _I, _I, _P, _I, _P, _P]),
    "hzo_init_polarization": (_I, [_I, _I, _I, _P, _I, _P, _D, C.c_uint64, _P]),
    "hzo_splitmix64": (_I, [C.c_uint64, C.c_int64, _P]),
}

_lib = None


class HzoError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Loads libhzo.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C cmsl_amd` "
                              "(or __graft_entry__.build())")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to libhzo must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def check(fn: str, rc: int) -> None:
    if rc != HZO_OK:
        msg = lib().hzo_last_error()
        raise HzoError(fn, rc, msg.decode() if msg else "")


def call(name: str, *args):
    rc = getattr(lib(), name)(*args)
    check(name, rc)
    return rc


def device_count() -> int:
    n = C.c_int(0)
    call("hzo_device_count", C.byref(n))
    return n.value


def default_config(**kw) -> Config:
    c = Config()
    lib().hzo_config_default(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def phases_nd(rows) -> "C.Array":
    arr = (PhaseND * 3)()
    for q, r in enumerate(rows):
        for k, v in (r.items() if isinstance(r, dict) else zip(("alpha", "beta", "gamma", "g", "a0",
                                                                "kappa", "L", "eps_r"), r)):
            setattr(arr[q], k, float(v))
    return arr
