"""Dump and checkpoint file formats (generate.cu:11-264), on the CPU.

The expected files are built here byte by byte from the layouts the
reference writes (struct packing and printf formats), independently of the
C++ writers: a swapped component, a wrong loop order or a missing header
field fails a comparison.
"""
import struct

import numpy as np
import pytest

from cmsl_amd import _lib as L
from cmsl_amd import film as F


def test_vector_field_is_interleaved_and_scaled(tmp_path):
    n = 5
    vx, vy, vz = np.arange(n) + 0.5, -np.arange(n) * 2.0, np.full(n, 3.25)
    p = tmp_path / "pm.bin"
    F.write_vector_field(p, vx, vy, vz, scale=100.0)
    raw = p.read_bytes()
    want = b"".join(struct.pack("<3d", vx[i] * 100.0, vy[i] * 100.0, vz[i] * 100.0) for i in range(n))
    assert raw == want


def test_grain_map_text_order_and_truncation(tmp_path):
    nx, ny, nz = 2, 2, 2
    g = np.array([0, 1, 2, 3, -3, 5, 7, -7], np.int32)  # id = i + nx*(j + ny*k)
    p = tmp_path / "microstructure.txt"
    F.write_grain_map(p, nx, ny, nz, g, scale=0.5)
    lines = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                v = g[i + nx * (j + ny * k)] * 0.5
                lines.append("%d\t%d\t%d\t%d\n" % (i, j, k, int(v)))  # int() truncates toward 0
            lines.append("\n")
        lines.append("\n")
    assert p.read_text() == "".join(lines)
    # -3 * 0.5 = -1.5 -> -1 and 7 * 0.5 = 3.5 -> 3 (truncation, not floor or rounding)
    assert "0\t0\t1\t-1\n" in p.read_text() and "0\t1\t1\t3\n" in p.read_text()


def test_midplane_slice_format(tmp_path):
    nx, ny, nz = 3, 2, 5
    v = np.arange(nx * ny * nz, dtype=np.float64) / 3.0
    p = tmp_path / "slice.dat"
    F.write_midplane_slice(p, nx, ny, nz, v)
    k = nz // 2
    want = []
    for j in range(ny):
        for i in range(nx):
            want.append("%d %d %.16e\n" % (i, j, v[i + nx * (j + ny * k)]))
        want.append("\n")
    assert p.read_text() == "".join(want)
    # plane k = 2 starts at voxel 12 (value 4), and 13/3 prints with 17 significant digits
    assert p.read_text().startswith("0 0 4.0000000000000000e+00\n1 0 4.3333333333333330e+00\n")


def _ckpt_bytes(nx, ny, nz, peloop, count, ez, inc, P, E):
    return struct.pack("<5i2d", nx, ny, nz, peloop, count, ez, inc) + P.astype("<f8").tobytes() + \
        E.astype("<f8").tobytes()


def test_checkpoint_layout_and_round_trip(tmp_path):
    nx, ny, nz = 3, 2, 2
    N = nx * ny * nz
    rng = np.random.default_rng(0)
    P = rng.normal(size=6 * N)
    E = rng.normal(size=3 * N)
    p = tmp_path / "ck.bin"
    F.write_checkpoint(p, nx, ny, nz, P, E, peloop=3, count=17, Ez_ext=-0.125, incEz=1e-3)
    assert p.read_bytes() == _ckpt_bytes(nx, ny, nz, 3, 17, -0.125, 1e-3, P, E)
    assert len(p.read_bytes()) == 36 + 9 * 8 * N  # 5 int32 + 2 f64 header, no padding
    hdr, P2, E2 = F.read_checkpoint(p, nx, ny, nz)
    assert hdr == {"peloop": 3, "count": 17, "Ez_ext": -0.125, "incEz": 1e-3}
    assert np.array_equal(P, P2) and np.array_equal(E, E2)


def test_checkpoint_read_rejects_grid_mismatch_and_short_files(tmp_path):
    nx, ny, nz = 2, 2, 2
    N = nx * ny * nz
    p = tmp_path / "ck.bin"
    p.write_bytes(_ckpt_bytes(nx, ny, nz, 0, 0, 0.0, 0.0, np.zeros(6 * N), np.zeros(3 * N)))
    with pytest.raises(L.HzoError) as e:
        F.read_checkpoint(p, nx, ny, nz + 1)
    assert e.value.code == L.HZO_ERR_DATA
    short = tmp_path / "short.bin"
    short.write_bytes(p.read_bytes()[:-8])
    with pytest.raises(L.HzoError) as e:
        F.read_checkpoint(short, nx, ny, nz)
    assert e.value.code == L.HZO_ERR_IO
    hdr_only = tmp_path / "hdr.bin"
    hdr_only.write_bytes(p.read_bytes()[:20])
    with pytest.raises(L.HzoError) as e:
        F.read_checkpoint(hdr_only, nx, ny, nz)
    assert e.value.code == L.HZO_ERR_IO
    with pytest.raises(L.HzoError):
        F.read_checkpoint(tmp_path / "missing.bin", nx, ny, nz)


def test_raw_f64_dump(tmp_path):
    v = np.linspace(-1.0, 1.0, 7)
    p = tmp_path / "phi.bin"
    F.write_raw_f64(p, v)
    assert np.array_equal(np.fromfile(p, dtype="<f8"), v)
