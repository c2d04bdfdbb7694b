"""Pins the oracle's depolarisation solve (depol_varEPS.cu) independently.

* divergence / gradient: exact sine-mode eigenvalues i*sin(2 pi m/n)/h in x,y;
  hand-computed clamped/Dirichlet boundary rows.
* operator: assembled column by column from apply_A on tiny grids, compared
  with a face-by-face finite-volume assembly (harmonic-mean face
  permittivities); symmetric positive definite on the interior; Jacobi
  inverse = 1/diag.
* PCG: converged solution equals a dense numpy solve; stopping rule and warm
  start follow depol_varEPS.cu:429-474.
"""
import numpy as np
import pytest

from oracle import pyoracle as O


def _cfg(nx, ny, nz, zbc=0, h=0.5, **kw):
    return O.cfg(nx, ny, nz, h, 0.01, dep_flag=1, zbc=zbc, **kw)


@pytest.mark.parametrize("axis,m", [("x", 1), ("x", 2), ("y", 3)])
def test_divergence_sine_modes(axis, m):
    nx, ny, nz, h, eps0 = 16, 12, 5, 0.5, 3.0
    c = _cfg(nx, ny, nz, h=h, eps0_nd=eps0)
    N = nx * ny * nz
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    P = np.zeros(6 * N)
    if axis == "x":
        th = 2 * np.pi * m * i / nx
        P[0:N] = np.sin(th).reshape(-1)
        exp = np.cos(th) * np.sin(2 * np.pi * m / nx) / h
    else:
        th = 2 * np.pi * m * j / ny
        P[N:2 * N] = np.sin(th).reshape(-1)
        exp = np.cos(th) * np.sin(2 * np.pi * m / ny) / h
    d = O.divergence(c, P)
    assert np.allclose(d, eps0 * exp.reshape(-1), atol=1e-13)


def test_divergence_clamped_z_rows():
    nx, ny, nz, h = 3, 2, 4, 0.5
    c = _cfg(nx, ny, nz, h=h, eps0_nd=1.0)
    N = nx * ny * nz
    P = np.zeros(6 * N)
    pz = np.array([1.0, 5.0, 2.0, 7.0])
    P[2 * N:3 * N] = np.repeat(pz, nx * ny)
    d = O.divergence(c, P).reshape(nz, ny, nx)
    # bottom uses d = k (ghost = self), top uses u = k (depol_varEPS.cu:299-300)
    assert np.allclose(d[:, 0, 0], [(5 - 1) / (2 * h), (2 - 1) / (2 * h), (7 - 5) / (2 * h), (7 - 2) / (2 * h)])


def test_efield_modes_and_electrode_ghosts():
    nx, ny, nz, h = 8, 6, 5, 0.5
    c = _cfg(nx, ny, nz, h=h, g_bot=0.0, g_top=0.0)
    N = nx * ny * nz
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    phi = (np.sin(2 * np.pi * i / nx) + 0.5 * np.cos(2 * np.pi * 2 * j / ny) + 0.1 * k * k).reshape(-1)
    E = O.efield(c, phi).reshape(3, nz, ny, nx)
    assert np.allclose(E[0], -np.cos(2 * np.pi * i / nx) * np.sin(2 * np.pi / nx) / h, atol=1e-13)
    assert np.allclose(E[1], 0.5 * np.sin(2 * np.pi * 2 * j / ny) * np.sin(2 * np.pi * 2 / ny) / h, atol=1e-13)
    ph = phi.reshape(nz, ny, nx)
    # interior: central difference; k=0 uses g_bot as the lower value, k=nz-1 uses g_top
    assert np.allclose(E[2, 2], -(ph[3] - ph[1]) / (2 * h))
    assert np.allclose(E[2, 0], -(ph[1] - 0.0) / (2 * h))
    assert np.allclose(E[2, nz - 1], -(0.0 - ph[nz - 2]) / (2 * h))
    cn = _cfg(nx, ny, nz, zbc=1, h=h)
    En = O.efield(cn, phi).reshape(3, nz, ny, nx)
    assert np.allclose(En[2, 0], 0.0) and np.allclose(En[2, nz - 1], 0.0)  # mirrored ghosts


def _harm(a, b):
    d = max(a + b, 1e-20)
    return 2 * a * b / d


def _assemble(c, eps):
    N = c.nx * c.ny * c.nz
    A = np.zeros((N, N))
    for col in range(N):
        e = np.zeros(N)
        e[col] = 1.0
        A[:, col] = O.apply_A(c, eps, e)
    return A


def _fv_matrix(nx, ny, nz, h, eps, zbc):
    """Face-by-face assembly of -div(eps grad) with harmonic face permittivity."""
    N = nx * ny * nz
    e3 = eps.reshape(nz, ny, nx)
    M = np.zeros((N, N))
    idx = lambda i, j, k: i + nx * (j + ny * k)
    ref = idx(nx // 2, ny // 2, nz // 2)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = idx(i, j, k)
                if (zbc == 0 and k in (0, nz - 1)) or (zbc == 1 and r == ref):
                    M[r, r] = 1.0
                    continue
                if zbc == 0:
                    kz = [k + 1, k - 1]
                else:
                    kz = [1, 1] if k == 0 else ([nz - 2, nz - 2] if k == nz - 1 else [k + 1, k - 1])
                nbs = [((i + 1) % nx, j, k), ((i - 1) % nx, j, k), (i, (j + 1) % ny, k), (i, (j - 1) % ny, k),
                       (i, j, kz[0]), (i, j, kz[1])]
                for (a, b, cc) in nbs:
                    w = _harm(e3[k, j, i], e3[cc, b, a]) / h**2
                    M[r, r] += w
                    M[r, idx(a, b, cc)] -= w
    return M


@pytest.mark.parametrize("zbc", [0, 1])
def test_operator_matches_finite_volume_assembly(zbc):
    rng = np.random.default_rng(5)
    nx, ny, nz, h = 4, 3, 5, 0.5
    c = _cfg(nx, ny, nz, zbc=zbc, h=h)
    eps = rng.choice([15.0, 30.0, 40.0], nx * ny * nz)
    A = _assemble(c, eps)
    M = _fv_matrix(nx, ny, nz, h, eps, zbc)
    assert np.allclose(A, M, rtol=1e-13, atol=1e-12)
    dinv = O.diag_inv(c, eps)
    assert np.allclose(dinv, 1.0 / np.diag(A), rtol=1e-15)


def test_operator_interior_block_is_spd_and_textbook_for_constant_eps():
    nx, ny, nz, h = 4, 4, 6, 0.5
    c = _cfg(nx, ny, nz, h=h)
    N = nx * ny * nz
    A = _assemble(c, np.full(N, 7.0))
    interior = np.arange(nx * ny, N - nx * ny)
    Ai = A[np.ix_(interior, interior)]
    assert np.allclose(Ai, Ai.T)
    assert np.linalg.eigvalsh(Ai).min() > 0
    assert np.allclose(np.diag(Ai), 6 * 7.0 / h**2)
    # Fourier mode in x/y times the Dirichlet sine mode in z is an eigenvector
    k, j, i = np.meshgrid(np.arange(1, nz - 1), np.arange(ny), np.arange(nx), indexing="ij")
    v = (np.cos(2 * np.pi * i / nx) * np.sin(np.pi * k / (nz - 1))).reshape(-1)
    lam = 7.0 / h**2 * (2 - 2 * np.cos(2 * np.pi / nx) + 2 - 2 * np.cos(np.pi / (nz - 1)))
    assert np.allclose(Ai @ v, lam * v)


def _solve_case(zbc, rule, tol, warm=None, nx=6, ny=5, nz=7, seed=2, max_iter=4000):
    rng = np.random.default_rng(seed)
    N = nx * ny * nz
    c = O.cfg(nx, ny, nz, 0.5, 0.01, dep_flag=1, zbc=zbc, eps0_nd=2.0, pcg_rule=rule, tol_abs=1e-30,
              tol_rel=tol, max_iter=max_iter)
    omega = rng.integers(0, 3, N).astype(np.int32)
    rows = [dict(alpha=0, beta=0, gamma=0, g=0, a0=0, kappa=0.1, L=0.5, eps_r=e) for e in (40.0, 30.0, 15.0)]
    f = O.Film(c, omega, np.array([0, 1, 2], np.int32), np.tile(np.eye(3).reshape(-1), 3), rows,
               np.zeros(N, np.int32))
    P = rng.uniform(-1, 1, 6 * N)
    phi = np.zeros(N) if warm is None else warm.copy()
    E3, st, divP, hist = f.depol(P, phi, hist=200)
    return c, f, P, phi, E3, st, divP, hist


@pytest.mark.parametrize("zbc", [0])
def test_pcg_converges_to_dense_solution(zbc):
    c, f, P, phi, E3, st, divP, hist = _solve_case(zbc, rule=1, tol=1e-13)
    N = f.N
    A = _assemble(c, f.eps)
    b = -divP
    if zbc == 0:
        b.reshape(c.nz, -1)[[0, c.nz - 1]] = 0.0
    else:
        b[c.nx // 2 + c.nx * (c.ny // 2 + c.ny * (c.nz // 2))] = 0.0
    x = np.linalg.solve(A, b)
    assert st.status == 0 and st.iters > 3
    assert np.allclose(phi, x, rtol=0, atol=1e-10 * np.abs(x).max())
    assert np.allclose(E3, O.efield(c, phi))
    assert np.isclose(st.bb, b @ b)


def test_neumann_gauge_row_breaks_symmetry_as_in_reference():
    # The gauge pin (depol_varEPS.cu:53-57) replaces row ref by the identity but
    # keeps column ref in the neighbouring rows, and the mirrored planes see
    # their one z neighbour twice, so A is not symmetric and the reference's
    # PCG stagnates at tight tolerances.  The oracle reproduces that behaviour
    # rather than fixing it.
    c, f, P, phi, E3, st, divP, hist = _solve_case(1, rule=1, tol=1e-12, nx=4, ny=4, nz=5)
    A = _assemble(c, f.eps)
    ref = c.nx // 2 + c.nx * (c.ny // 2 + c.ny * (c.nz // 2))
    assert np.all(A[ref] == np.eye(A.shape[0])[ref])
    assert np.count_nonzero(A[:, ref]) > 1
    assert st.status == 1 and st.rr < st.rr0


def _neumann_film(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    c = _cfg(nx, ny, nz, zbc=1, pcg_rule=1, tol_abs=1e-30, tol_rel=1e-8)
    eps = rng.choice([15.0, 30.0, 40.0], nx * ny * nz) * rng.uniform(0.8, 1.2, nx * ny * nz)
    S = np.ones(nx * ny * nz)
    S[: nx * ny] = 2.0
    S[-nx * ny:] = 2.0
    return c, eps, S, nx // 2 + nx * (ny // 2 + ny * (nz // 2))


def test_neumann_operator_is_symmetric_after_halving_the_mirrored_rows():
    # A mirrored plane (k = 0 or nz - 1) takes its ghost neighbour as the image
    # of plane 1 (nz - 2), so its row couples to that plane with twice the face
    # coefficient the neighbour's row couples back with: A = S A_s, S = 2 on
    # those planes.  With the pinned gauge unknown dropped, A_s = S^-1 A is
    # symmetric positive definite: the operator that CG with the mirrored
    # planes' inner-product terms halved runs on.
    nx, ny, nz = 4, 3, 6
    c, eps, S, ref = _neumann_film(nx, ny, nz, 11)
    A = _assemble(c, eps)
    keep = np.arange(A.shape[0]) != ref
    Ak = A[np.ix_(keep, keep)]
    As = (A / S[:, None])[np.ix_(keep, keep)]
    assert not np.allclose(Ak, Ak.T)
    assert np.allclose(As, As.T, rtol=1e-14, atol=1e-12 * np.abs(As).max())
    assert np.linalg.eigvalsh(As).min() > 0
    # the mirrored rows are exactly twice the symmetric ones
    k0 = np.arange(nx * ny)
    assert np.allclose(A[k0, nx * ny + k0], 2 * A[nx * ny + k0, k0])


def _single_reduction_cg(A, dinv, b, wt, iters):
    """Chronopoulos-Gear CG (one inner-product round per iteration) with the
    inner products r.u and w.u weighted by wt; returns the relative residual
    history of the recursively updated r."""
    x = np.zeros_like(b)
    r = b.copy()
    u = dinv * r
    w = A(u)
    g, d = (wt * r) @ u, (wt * w) @ u
    p = s = None
    g_old = a_old = None
    hist = []
    for i in range(iters):
        if i == 0:
            beta, alpha = 0.0, g / d
            p, s = u.copy(), w.copy()
        else:
            beta = g / g_old
            alpha = g / (d - beta * g / a_old)
            p, s = u + beta * p, w + beta * s
        x += alpha * p
        r -= alpha * s
        u = dinv * r
        w = A(u)
        g_old, a_old = g, alpha
        g, d = (wt * r) @ u, (wt * w) @ u
        hist.append(np.sqrt((r @ r) / (b @ b)))
        if hist[-1] <= 1e-8 or not np.isfinite(hist[-1]):
            break
    return x, np.array(hist)


def test_single_reduction_cg_needs_the_mirrored_plane_weights():
    # The single-reduction recurrence for p.Ap assumes u.Ap = p.Au.  On the
    # Neumann operator with plain inner products it stalls; with the mirrored
    # planes' terms halved it is CG on the symmetric A_s and converges to the
    # solution of A x = b (same iterates, same solution).
    nx, ny, nz = 16, 12, 8
    c, eps, S, ref = _neumann_film(nx, ny, nz, 12)
    A = lambda v: O.apply_A(c, eps, np.ascontiguousarray(v))
    dinv = O.diag_inv(c, eps)
    b = np.random.default_rng(3).standard_normal(nx * ny * nz)
    b[ref] = 0.0
    x_w, h_w = _single_reduction_cg(A, dinv, b, 1.0 / S, 2000)
    x_p, h_p = _single_reduction_cg(A, dinv, b, np.ones_like(b), 2000)
    assert h_w[-1] <= 1e-8 and len(h_w) < 500
    assert np.linalg.norm(b - A(x_w)) <= 1.05e-8 * np.linalg.norm(b)
    assert not (np.all(np.isfinite(h_p)) and h_p[-1] <= 1e-8 and len(h_p) <= len(h_w))
    assert np.nanmin(h_p[len(h_w):]) > 1e-7  # plain sums: no convergence in 4x the weighted count


def test_reference_stopping_rule_and_warm_start():
    c, f, P, phi, E3, st, divP, hist = _solve_case(0, rule=0, tol=1e-2)
    rz = hist[: st.iters + 1]
    # stops at the first iterate with rz/rz0 <= tol_rel^2 (tol_abs^2 = 1e-60 never binds)
    assert rz[-1] / rz[0] <= 1e-4
    assert np.all(rz[1:-1] / rz[0] > 1e-4)
    # warm start from the converged potential: the rule restarts from the new rz0
    c2, f2, P2, phi2, E32, st2, divP2, hist2 = _solve_case(0, rule=0, tol=1e-2, warm=phi)
    assert st2.rz0 < st.rz0 * 1e-3
    assert np.allclose(P2, P)


def test_pcg_breakdown_on_zero_permittivity():
    # eps == 0 everywhere: harm() gives 0 faces, p.Ap = 0 -> breakdown at iteration 1
    rng = np.random.default_rng(1)
    nx, ny, nz = 4, 4, 5
    N = nx * ny * nz
    c = O.cfg(nx, ny, nz, 0.5, 0.01, eps0_nd=1.0)
    rows = [dict(alpha=0, beta=0, gamma=0, g=0, a0=0, kappa=0.1, L=0.5, eps_r=0.0)] * 3
    f = O.Film(c, np.zeros(N, np.int32), np.array([1], np.int32), np.eye(3).reshape(-1), rows,
               np.zeros(N, np.int32))
    phi = np.zeros(N)
    E3, st, divP, hist = f.depol(rng.uniform(-1, 1, 6 * N), phi)
    assert st.status == -2 and st.iters == 1


def test_pcg_invalid_initial_residual_on_nan():
    nx, ny, nz = 4, 4, 5
    N = nx * ny * nz
    c = O.cfg(nx, ny, nz, 0.5, 0.01, eps0_nd=1.0)
    rows = [dict(alpha=0, beta=0, gamma=0, g=0, a0=0, kappa=0.1, L=0.5, eps_r=30.0)] * 3
    f = O.Film(c, np.zeros(N, np.int32), np.array([1], np.int32), np.eye(3).reshape(-1), rows,
               np.zeros(N, np.int32))
    P = np.zeros(6 * N)
    P[2 * N + 2 * nx * ny + 5] = np.nan
    E3, st, divP, hist = f.depol(P, np.zeros(N))
    assert st.status == -1 and st.iters == 0


def test_pcg_max_iter():
    # tolerance 1e-14 is not reached in 5 iterations: the loop stops at max_iter
    c, f, P, phi, E3, st, divP, hist = _solve_case(0, rule=1, tol=1e-14, max_iter=5)
    assert st.iters == 5 and st.status == 1
