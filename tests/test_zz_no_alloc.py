"""The no-allocation GPU Voronoi labeller (hzo_voronoi_gpu_noalloc) against
the labeller that allocates (hzo_voronoi_gpu) and the CPU labeller: labels
equal bit for bit, on two calls in a row, the second smaller; sizes beyond
the scratch limit are refused before any GPU work."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[no-alloc] {msg}", flush=True)


def _alloc_labels(nx, ny, nz, layers, ng, seed):
    lc = np.array(layers, np.int32)
    om = np.zeros(nx * ny * nz, np.int32)
    cent = np.zeros(3 * ng)
    eul = np.zeros(3 * ng)
    L.call("hzo_voronoi_gpu", 0, nx, ny, nz, len(layers), L.ptr(lc), ng, C.c_uint64(seed), L.ptr(om),
           L.ptr(cent), L.ptr(eul))
    return om, cent, eul


def _hip():
    """The HIP runtime libhzo.so itself links (one runtime per process): the
    test allocates the caller's device array with it."""
    L.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _noalloc_labels(nx, ny, nz, layers, ng, seed):
    hip = _hip()
    N = nx * ny * nz
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), 4 * N) == 0
    try:
        assert hip.hipMemset(dev, 0xF9, 4 * N) == 0  # every label must be written
        assert hip.hipDeviceSynchronize() == 0
        lc = np.array(layers, np.int32)
        cent = np.zeros(3 * ng)
        eul = np.zeros(3 * ng)
        L.call("hzo_voronoi_gpu_noalloc", 0, nx, ny, nz, len(layers), L.ptr(lc), ng, C.c_uint64(seed), dev,
               L.ptr(cent), L.ptr(eul))
        om = np.zeros(N, np.int32)
        assert hip.hipMemcpy(L.ptr(om), dev, 4 * N, 2) == 0  # hipMemcpyDeviceToHost
    finally:
        hip.hipFree(dev)
    return om, cent, eul


def test_noalloc_labels_equal_the_allocating_labeller_twice():
    cases = [(96, 80, 24, [12, 12], 700, 21), (40, 36, 10, [10], 50, 5)]
    for nx, ny, nz, layers, ng, seed in cases:
        _say(f"case {nx}x{ny}x{nz} layers {layers} grains {ng} seed {seed}")
        om_a, cent_a, eul_a = _alloc_labels(nx, ny, nz, layers, ng, seed)
        om_b, cent_b, eul_b = _noalloc_labels(nx, ny, nz, layers, ng, seed)
        assert np.array_equal(om_a, om_b)
        om_c, _, _ = H.film.voronoi(nx, ny, nz, ng, seed, layers)  # CPU labeller
        assert np.array_equal(om_c, om_b)
        assert np.array_equal(cent_a, cent_b) and np.array_equal(eul_a, eul_b)
        assert om_b.min() >= 0 and om_b.max() < ng  # every voxel written


def test_noalloc_refuses_more_grains_than_its_scratch():
    ng = 65537  # HZO_VORONOI_NOALLOC_MAX_GRAINS + 1
    nx = ny = nz = 4
    lc = np.array([nz], np.int32)
    cent = np.zeros(3 * ng)
    eul = np.zeros(3 * ng)
    hip = _hip()
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), 4 * nx * ny * nz) == 0
    try:
        rc = L.lib().hzo_voronoi_gpu_noalloc(0, nx, ny, nz, 1, L.ptr(lc), ng, C.c_uint64(1), dev, L.ptr(cent),
                                             L.ptr(eul))
    finally:
        hip.hipFree(dev)
    assert rc == L.HZO_ERR_ARG
    assert b"scratch limit" in L.lib().hzo_last_error()
