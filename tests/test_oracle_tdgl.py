"""Pins the oracle's TDGL update (kernels_evol.cu:47-231) without re-typing it.

* Landau driving force: gradient of the closed-form free energy, derived
  symbolically by sympy, rotated by an independent scipy rotation.
* Laplacian: discrete Fourier modes are exact eigenvectors (periodic x/y,
  clamped z: cos(pi q (k+1/2)/nz)).
* Structural cases: Ps == 0 stays exactly 0 in FE films, void voxels give 0.
"""
import numpy as np
import pytest
import sympy as sp
from scipy.spatial.transform import Rotation

from oracle import pyoracle as O


def _energy_grad():
    px, py, pz, sx, sy, sz, a, g, b, c, a0 = sp.symbols("px py pz sx sy sz a g b c a0", real=True)
    f = ((a + g) * pz**2 + (a - g) * sz**2 + b * (pz**4 + 6 * pz**2 * sz**2 + sz**4)
         + c * (pz**6 + 15 * pz**4 * sz**2 + 15 * pz**2 * sz**4 + sz**6)
         + a0 * (px**2 + py**2 + sx**2 + sy**2))
    v = (px, py, pz, sx, sy, sz)
    return sp.lambdify(v + (a, g, b, c, a0), [sp.diff(f, t) for t in v], "numpy")


GRAD = _energy_grad()


def _nbsum(F, nx, ny, nz):
    """6-neighbour sum with periodic x,y and clamped z, F shaped (nz, ny, nx)."""
    up = np.concatenate([F[1:], F[-1:]], axis=0)
    dn = np.concatenate([F[:1], F[:-1]], axis=0)
    return np.roll(F, -1, 2) + np.roll(F, 1, 2) + np.roll(F, -1, 1) + np.roll(F, 1, 1) + up + dn


def _case(rng, nx=5, ny=4, nz=6, iface_frac=0.0):
    N = nx * ny * nz
    rows = [dict(alpha=rng.uniform(-2, 2), beta=rng.uniform(-1, 1), gamma=rng.uniform(0, 1),
                 g=rng.uniform(-0.5, 0.5), a0=rng.uniform(0, 3), kappa=rng.uniform(0.05, 0.5),
                 L=rng.uniform(0.1, 1.0), eps_r=30.0) for _ in range(3)]
    ng = 3
    omega = rng.integers(0, ng, N).astype(np.int32)
    gphase = np.array([0, 1, 2], np.int32)
    Rm = Rotation.random(ng, random_state=int(rng.integers(1 << 30))).as_matrix()
    TR = np.ascontiguousarray(Rm.reshape(-1))
    iface = (rng.random(N) < iface_frac).astype(np.int32)
    c = O.cfg(nx, ny, nz, 0.5, 0.01, dep_flag=1)
    f = O.Film(c, omega, gphase, TR, rows, iface)
    P = rng.uniform(-1, 1, 6 * N)
    E3 = rng.uniform(-0.5, 0.5, 3 * N)
    return f, rows, Rm, P, E3


@pytest.mark.parametrize("iface_frac", [0.0, 0.5])
def test_landau_gradient_matches_symbolic_energy(iface_frac):
    rng = np.random.default_rng(7)
    nx, ny, nz = 5, 4, 6
    N = nx * ny * nz
    f, rows, Rm, P, E3 = _case(rng, nx, ny, nz, iface_frac)
    Ez_ext = 0.37
    Pn, Pl = f.tdgl(P, E3, Ez_ext, want_local=True)
    dt, h = 0.01, 0.5
    Pm = P.reshape(6, N)
    exp = np.empty((6, N))
    S = np.stack([_nbsum(Pm[q].reshape(nz, ny, nx), nx, ny, nz).reshape(-1) for q in range(6)])
    for id_ in range(N):
        gid = f.omega[id_]
        r = rows[f.grain_phase[gid]]
        g, Lf = r["g"], r["L"]
        if f.iface[id_]:
            g, Lf = 0.4 * g, 11.0 * Lf
        T = Rm[gid]
        pl = T @ Pm[0:3, id_]
        sl = T @ Pm[3:6, id_]
        gr = np.array(GRAD(*pl, *sl, r["alpha"], g, r["beta"], r["gamma"], r["a0"]), dtype=float)
        dF = np.concatenate([T.T @ gr[0:3], T.T @ gr[3:6]])
        dF[0] -= E3[id_]
        dF[1] -= E3[N + id_]
        dF[2] -= Ez_ext + E3[2 * N + id_]
        A = r["kappa"] * dt / h**2
        exp[:, id_] = Pm[:, id_] * (1 - 6 * A) + A * S[:, id_] - Lf * dt * dF
        assert np.allclose(Pl.reshape(6, N)[:, id_], np.concatenate([pl, sl]), rtol=0, atol=1e-14)
    assert np.max(np.abs(Pn.reshape(6, N) - exp)) < 1e-12 * max(1.0, np.max(np.abs(exp)))


def _laplace_only(nx, ny, nz, kappa=0.3):
    N = nx * ny * nz
    rows = [dict(alpha=0, beta=0, gamma=0, g=0, a0=0, kappa=kappa, L=0.7, eps_r=1.0)] * 3
    c = O.cfg(nx, ny, nz, 0.5, 0.01, dep_flag=0)
    f = O.Film(c, np.zeros(N, np.int32), np.array([1], np.int32), np.eye(3).reshape(-1), rows,
               np.zeros(N, np.int32))
    return f, kappa * 0.01 / 0.25


@pytest.mark.parametrize("axis,m", [("x", 1), ("x", 3), ("y", 2), ("z", 1), ("z", 4)])
def test_laplacian_fourier_modes_are_eigenvectors(axis, m):
    nx, ny, nz = 16, 12, 9
    N = nx * ny * nz
    f, A = _laplace_only(nx, ny, nz)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    if axis == "x":
        mode = np.cos(2 * np.pi * m * i / nx + 0.3)
        lam = 2 * np.cos(2 * np.pi * m / nx) - 2
    elif axis == "y":
        mode = np.sin(2 * np.pi * m * j / ny + 0.1)
        lam = 2 * np.cos(2 * np.pi * m / ny) - 2
    else:  # clamped z: ghost = self  ->  Neumann cosine modes at half-integer points
        mode = np.cos(np.pi * m * (k + 0.5) / nz)
        lam = 2 * np.cos(np.pi * m / nz) - 2
    P = np.zeros(6 * N)
    for q in range(6):
        P[q * N:(q + 1) * N] = (q + 1) * mode.reshape(-1)
    Pn = f.tdgl(P, None, 0.0)
    assert np.allclose(Pn, P * (1 + A * lam), rtol=0, atol=1e-13)


def test_uniform_field_is_invariant_without_driving_force():
    f, A = _laplace_only(6, 5, 4)
    P = np.repeat(np.arange(1.0, 7.0), f.N)
    Pn = f.tdgl(P, None, 0.0)
    assert np.allclose(Pn, P, rtol=1e-15, atol=0)


def test_fe_film_keeps_zero_staggered_polarisation_exactly():
    rng = np.random.default_rng(3)
    f, rows, Rm, P, E3 = _case(rng)
    f.grain_phase[:] = 1  # all FE
    f = O.Film(f.c, f.omega, f.grain_phase, f.TR, rows, f.iface)
    N = f.N
    P[3 * N:] = 0.0
    Pn = f.tdgl(P, E3, 0.5)
    assert np.all(Pn[3 * N:] == 0.0)


def test_geometric_decay_with_quadratic_landau_only():
    # beta = gamma = 0, uniform P along the grain's polar axis, no field:
    # P_new = P (1 - 2 L dt (alpha + g)) up to the rounding of the flat Laplacian.
    nx, ny, nz = 4, 4, 4
    N = nx * ny * nz
    rows = [dict(alpha=-0.8, beta=0, gamma=0, g=0.3, a0=2.0, kappa=0.1, L=0.6, eps_r=1.0)] * 3
    T = Rotation.from_euler("ZXZ", [0.4, 1.1, 2.0]).as_matrix()
    c = O.cfg(nx, ny, nz, 0.5, 0.01, dep_flag=0)
    f = O.Film(c, np.zeros(N, np.int32), np.array([1], np.int32), T.reshape(-1), rows, np.zeros(N, np.int32))
    axis = T[2]  # local z in the lab frame
    P = np.concatenate([np.repeat(0.7 * axis, N), np.zeros(3 * N)])
    Pn = f.tdgl(P, None, 0.0)
    fac = 1 - 2 * 0.6 * 0.01 * (-0.8 + 0.3)
    assert np.allclose(Pn, P * fac, rtol=0, atol=1e-14)


def test_void_voxels_are_zeroed():
    rng = np.random.default_rng(11)
    f, rows, Rm, P, E3 = _case(rng)
    om = f.omega.copy()
    om[::5] = -1
    f2 = O.Film(f.c, om, f.grain_phase, f.TR, rows, f.iface)
    Pn = f2.tdgl(P, E3, 0.1)
    N = f.N
    for q in range(6):
        assert np.all(Pn[q * N:(q + 1) * N][om < 0] == 0.0)
