"""Wiring of the reverse mode's torch.autograd.Function, on the CPU.

A stand-in film with an analytic loop m(theta) and its exact transpose
Jacobian replaces libhzo.so here, so torch.autograd.gradcheck verifies that
PELoop hands the coefficients, the loop seeds dL/dm and the returned
gradient through correctly (shapes, order AFE/FE/DE x alpha/beta/gamma/kappa,
no transposition).  The HIP adjoint itself is checked against finite
differences in test_gpu_adjoint.py.
"""
from types import SimpleNamespace

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from cmsl_amd.adjoint import COEFFS, coeff_matrix, pe_loop, rows_with  # noqa: E402


class AnalyticFilm:
    def __init__(self, nsteps, seed=0):
        rng = np.random.default_rng(seed)
        rows = [dict(alpha=-1.0, beta=0.5, gamma=0.2, g=0.0, a0=1.0, kappa=0.1, L=1.0, eps_r=30.0)
                for _ in range(3)]
        self.inp = SimpleNamespace(phases=rows, P0=np.zeros(6))
        self.W = rng.normal(size=(nsteps, 12)) * 0.3
        self._adj_steps = 0
        self.calls = []

    def adjoint_reserve(self, nsteps, ckpt_every, tol):
        self._adj_steps = nsteps
        self.calls.append(("reserve", nsteps, ckpt_every))

    def set_phase_table(self, rows):
        self.theta = coeff_matrix(rows).reshape(-1)

    def set_state(self, P0, phi0):
        self.calls.append(("state",))

    def forward_record(self, Ez):
        self.Ez = np.asarray(Ez)
        self.z = self.W @ self.theta
        return np.tanh(self.z) * self.Ez

    def backward(self, dL_dm):
        d = self.Ez * (1.0 - np.tanh(self.z) ** 2) * dL_dm
        return (self.W.T @ d).reshape(3, 4), 0


def test_rows_with_replaces_only_the_fitted_coefficients():
    rows = AnalyticFilm(3).inp.phases
    c = np.arange(12.0).reshape(3, 4)
    new = rows_with(rows, c)
    assert np.array_equal(coeff_matrix(new), c)
    for a, b in zip(rows, new):
        for k in a:
            if k not in COEFFS:
                assert a[k] == b[k]


def test_pe_loop_gradcheck_through_the_autograd_function():
    S = 7
    film = AnalyticFilm(S)
    Ez = np.linspace(-1.0, 1.0, S)
    c0 = torch.tensor(coeff_matrix(film.inp.phases) + 0.05 * np.arange(12).reshape(3, 4),
                      dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda c: pe_loop(c, film, Ez, ckpt_every=3), (c0,), eps=1e-6, atol=1e-8)
    assert film.calls[0] == ("reserve", S, 3)
