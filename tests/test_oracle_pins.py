"""Independent pins for the oracle's diagnostics, interface flag and
permittivity map.

Each rule is restated here in numpy from the reference text and checked on
hand-built grids whose answers a wrong slot, a wrong neighbour or a wrong
phase would change:
- diagnostics: evolve.cu:259-357 with build_mat_mask_and_p
  (kernels_evol.cu:23-44); the local-frame sums come from the grain-frame
  copies the TDGL kernel stores of its *input* polarisation
  (kernels_evol.cu:157-175);
- interface flag: build_interface_flag (main.cu:13-42);
- permittivity map: InitConfKernel (init_conf.cu:68-106), eps_m = eps_r of
  the voxel's phase (mat 2 -> DE, 1 -> FE, else AFE), void voxels keep the
  zero the allocation left (allocateVars.cu:24).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O

AFE, FE, DE = 0, 1, 2


def _cfg(nx, ny, nz):
    return O.cfg(nx, ny, nz, 0.5, 0.01)


def _rot(seed, ng):
    """ng proper rotations (QR of seeded Gaussian matrices, det +1)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ng):
        q, r = np.linalg.qr(rng.standard_normal((3, 3)))
        q = q * np.sign(np.diag(r))
        if np.linalg.det(q) < 0:
            q[:, 0] = -q[:, 0]
        out.append(q.reshape(-1))
    return np.ascontiguousarray(np.concatenate(out))


def _phase_rows(eps=(40.0, 30.0, 15.0)):
    rows = []
    for q in range(3):
        rows.append(dict(alpha=-1.0 + 0.1 * q, beta=0.3, gamma=0.2, g=0.05 * q, a0=0.7, kappa=0.5, L=1.0,
                         eps_r=eps[q]))
    return rows


# ------------------------------------------------------------------ diagnostics

def _dyadic(rng, n):
    """Values k/64 with |k| < 2^10: every sum of a few thousand is exact in
    FP64 whatever the order, so expected sums compare with ==."""
    return rng.integers(-1023, 1024, n).astype(np.float64) / 64.0


def _diag_numpy(P, Pl, omega, gphase):
    N = omega.size
    Pmz, Psz = P[2 * N:3 * N], P[5 * N:6 * N]
    mat = np.where(omega >= 0, gphase[np.maximum(omega, 0)], -1)
    out = [Pmz.sum(), Psz.sum(), Pl[2 * N:3 * N].sum(), Pl[5 * N:6 * N].sum()]
    for target in (FE, DE, AFE):  # evolve.cu:285, 313, 338
        m = (mat == target).astype(np.float64)
        out += [(m * Pmz).sum(), m.sum()]
    return np.array(out)


def test_diagnostics_slots_and_masks_on_a_hand_built_film():
    nx, ny, nz = 6, 5, 4
    N = nx * ny * nz
    rng = np.random.default_rng(11)
    # four grains of three phases plus void voxels; phase counts all differ
    gphase = np.array([FE, DE, AFE, FE], np.int32)
    omega = rng.choice(np.array([0, 1, 2, 3, -1], np.int32), size=N, p=[0.3, 0.15, 0.25, 0.2, 0.1])
    omega = omega.astype(np.int32)
    counts = {t: int(np.sum((omega >= 0) & (gphase[np.maximum(omega, 0)] == t))) for t in (FE, DE, AFE)}
    assert len(set(counts.values())) == 3
    P = _dyadic(rng, 6 * N)
    Pl = _dyadic(rng, 6 * N)
    out = np.zeros(10)
    c = _cfg(nx, ny, nz)
    O.lib().hzo_oracle_diagnostics(C.byref(c), O._p(P), O._p(Pl), O._p(omega), O._p(gphase), O._p(out))
    exp = _diag_numpy(P, Pl, omega, gphase)
    assert np.array_equal(out, exp)
    # the slot order is FE, DE, AFE with (sum of Pmz, count) each
    assert out[5] == counts[FE] and out[7] == counts[DE] and out[9] == counts[AFE]
    # the global sums use Pm_z and Ps_z (components 2 and 5), not x or y
    assert out[0] == P[2 * N:3 * N].sum() and out[1] == P[5 * N:6 * N].sum()
    assert out[0] != P[:N].sum() and out[1] != P[3 * N:4 * N].sum()


def test_diagnostics_mask_uses_the_grain_phase_not_the_grain_id():
    # grain ids chosen so that id == phase would give a different answer
    nx, ny, nz = 4, 3, 3
    N = nx * ny * nz
    gphase = np.array([DE, AFE, FE], np.int32)  # grain 0 is DE, grain 2 is FE
    omega = np.repeat(np.array([0, 1, 2], np.int32), N // 3)
    P = np.zeros(6 * N)
    P[2 * N:3 * N] = np.repeat([1.0, 10.0, 100.0], N // 3)
    out = np.zeros(10)
    c = _cfg(nx, ny, nz)
    O.lib().hzo_oracle_diagnostics(C.byref(c), O._p(P), O._p(np.zeros(6 * N)), O._p(omega), O._p(gphase),
                                   O._p(out))
    k = N // 3
    assert out[4] == 100.0 * k and out[6] == 1.0 * k and out[8] == 10.0 * k
    assert out[5] == out[7] == out[9] == k


def test_tdgl_local_frame_copies_are_the_rotated_input_state():
    # kernels_evol.cu:157-175: p_l = TR . P of the centre voxel before the
    # update, stored for every non-void voxel; the diagnostics read them one
    # step late (evolve.cu:259-357 after the next kernelCalc_antiferro)
    nx, ny, nz = 5, 4, 4
    N = nx * ny * nz
    ng = 3
    rng = np.random.default_rng(5)
    omega = rng.integers(-1, ng, N).astype(np.int32)
    gphase = np.array([FE, AFE, DE], np.int32)
    TR = _rot(9, ng)
    c = _cfg(nx, ny, nz)
    f = O.Film(c, omega, gphase, TR, _phase_rows())
    P = rng.uniform(-0.5, 0.5, 6 * N)
    Pn, Pl = f.tdgl(P, np.zeros(3 * N), 0.3, want_local=True)
    T = TR.reshape(ng, 3, 3)
    live = omega >= 0
    for base in (0, 3):  # Pm, then Ps
        v = np.stack([P[(base + a) * N:(base + a + 1) * N] for a in range(3)], axis=1)
        for row in range(3):
            t = T[np.maximum(omega, 0), row, :]
            # the reference's operand order: t1*x + t2*y + t3*z
            exp = (t[:, 0] * v[:, 0] + t[:, 1] * v[:, 1]) + t[:, 2] * v[:, 2]
            got = Pl[(base + row) * N:(base + row + 1) * N]
            assert np.array_equal(got[live], exp[live])
            assert np.all(got[~live] == 0.0)
    # and not the rotated output state
    t = T[np.maximum(omega, 0), 2, :]
    mz_new = (t[:, 0] * Pn[:N] + t[:, 1] * Pn[N:2 * N]) + t[:, 2] * Pn[2 * N:3 * N]
    assert not np.allclose(Pl[2 * N:3 * N][live], mz_new[live], rtol=1e-6, atol=0.0)


def test_run_leaves_the_local_copies_of_the_last_steps_input():
    nx, ny, nz = 6, 4, 5
    N = nx * ny * nz
    ng = 2
    rng = np.random.default_rng(8)
    omega = rng.integers(0, ng, N).astype(np.int32)
    gphase = np.array([FE, AFE], np.int32)
    TR = _rot(3, ng)
    c = _cfg(nx, ny, nz)
    f = O.Film(c, omega, gphase, TR, _phase_rows())
    P0 = rng.uniform(-0.3, 0.3, 6 * N)
    Ez = np.array([0.1, 0.2, -0.1])
    P2 = P0.copy()
    phi = np.zeros(N)
    E3 = np.zeros(3 * N)
    f.run(P2, phi, E3, Ez[:2])  # state after two steps
    P3 = P2.copy()
    _, Pl = f.run(P3, phi.copy(), E3.copy(), Ez[2:], want_local=True)
    T = TR.reshape(ng, 3, 3)[omega]
    mz = (T[:, 2, 0] * P2[:N] + T[:, 2, 1] * P2[N:2 * N]) + T[:, 2, 2] * P2[2 * N:3 * N]
    assert np.array_equal(Pl[2 * N:3 * N], mz)


# ------------------------------------------------------------------ interface flag

def _iface_numpy(omega, gphase, nx, ny, nz):
    """main.cu:13-42: an FE voxel off the first and last plane whose z-up or
    z-down neighbour is not FE (a void neighbour counts as not FE)."""
    om = omega.reshape(nz, ny, nx)
    mat = np.where(om >= 0, gphase[np.maximum(om, 0)], -1)
    out = np.zeros_like(om)
    fe = mat == FE
    inner = np.zeros_like(fe)
    inner[1:-1] = True
    up = np.zeros_like(fe)
    dn = np.zeros_like(fe)
    up[:-1] = mat[1:] != FE
    dn[1:] = mat[:-1] != FE
    out[fe & inner & (up | dn)] = 1
    return out.reshape(-1).astype(np.int32)


def _oracle_iface(omega, gphase, nx, ny, nz):
    c = _cfg(nx, ny, nz)
    out = np.zeros(nx * ny * nz, np.int32)
    O.lib().hzo_oracle_interface_flag(C.byref(c), O._p(omega), O._p(gphase), O._p(out))
    return out


def test_interface_flag_three_layer_stack():
    # planes FE / DE / FE / FE / FE: the FE voxels of plane 2 touch DE below,
    # plane 3 has FE on both sides, the first and last planes are never flagged
    nx, ny, nz = 3, 2, 5
    gphase = np.array([FE, DE], np.int32)
    layers = [0, 1, 0, 0, 0]
    omega = np.repeat(np.array(layers, np.int32), nx * ny)
    got = _oracle_iface(omega, gphase, nx, ny, nz).reshape(nz, -1)
    assert np.all(got[0] == 0) and np.all(got[4] == 0)  # faces
    assert np.all(got[1] == 0)  # DE voxels are never flagged
    assert np.all(got[2] == 1)  # only the lower neighbour is not FE
    assert np.all(got[3] == 0)


def test_interface_flag_only_looks_along_z():
    # an FE voxel whose x and y neighbours are DE but whose z neighbours are
    # FE stays 0; one whose upper neighbour alone is void gets 1
    nx, ny, nz = 3, 3, 4
    N = nx * ny * nz
    gphase = np.array([FE, DE], np.int32)
    omega = np.zeros(N, np.int32)
    om = omega.reshape(nz, ny, nx)
    om[1, 1, 0] = 1
    om[1, 1, 2] = 1
    om[1, 0, 1] = 1
    om[1, 2, 1] = 1  # centre (1,1,1) surrounded by DE in x and y
    om[2, 0, 0] = -1  # void above (1, 0, 0)
    got = _oracle_iface(omega, gphase, nx, ny, nz).reshape(nz, ny, nx)
    assert got[1, 1, 1] == 0
    assert got[1, 0, 0] == 1
    assert got[1, 1, 0] == 0  # DE
    # the voxels of plane 2 above the DE voxels of plane 1 have a non-FE
    # lower neighbour
    assert got[2, 1, 0] == 1 and got[2, 0, 1] == 1 and got[2, 1, 1] == 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_interface_flag_matches_numpy_restatement_on_random_films(seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = 7, 6, 6
    ng = 5
    gphase = rng.integers(0, 3, ng).astype(np.int32)
    omega = rng.integers(-1, ng, nx * ny * nz).astype(np.int32)
    got = _oracle_iface(omega, gphase, nx, ny, nz)
    assert np.array_equal(got, _iface_numpy(omega, gphase, nx, ny, nz))


# ------------------------------------------------------------------ permittivity map

def test_eps_map_by_phase_with_void_zero():
    nx, ny, nz = 4, 3, 2
    N = nx * ny * nz
    gphase = np.array([DE, FE, AFE, 7], np.int32)  # a phase code other than 1, 2 maps to AFE
    omega = np.array([0, 1, 2, 3, -1, 4] * (N // 6), np.int32)  # 4 >= ng is void too
    ph = O.phases(_phase_rows(eps=(40.0, 30.0, 15.0)))
    eps = np.full(N, -1.0)
    c = _cfg(nx, ny, nz)
    O.lib().hzo_oracle_eps_map(C.byref(c), O._p(omega), 4, O._p(gphase), ph, O._p(eps))
    exp = np.array([15.0, 30.0, 40.0, 40.0, 0.0, 0.0] * (N // 6))
    assert np.array_equal(eps, exp)
