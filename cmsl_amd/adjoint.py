"""Differentiable P-E loop for Landau-coefficient fitting.

`pe_loop(coeffs, film, Ez)` runs the film through the applied-field sequence
Ez and returns the loop samples m_n = <Pmz> after every step as a torch
tensor.  Its backward pass is the hand-written adjoint of libhzo.so
(hzo_forward_record / hzo_backward): adjoint TDGL kernels, the same PCG for
the self-adjoint Poisson solve, checkpoints every `ckpt_every` steps with
recomputation of each segment.  Gradients are with respect to the
non-dimensional (alpha, beta, gamma, kappa) of each phase (AFE, FE, DE),
coeffs of shape (3, 4).

torch only carries the tensors across the autograd boundary; every forward
and backward step runs in the HIP library.
"""
from __future__ import annotations

import numpy as np
import torch

COEFFS = ("alpha", "beta", "gamma", "kappa")


def coeff_matrix(rows) -> np.ndarray:
    """(3, 4) array of (alpha, beta, gamma, kappa) from phase rows."""
    return np.array([[float(r[k]) for k in COEFFS] for r in rows])


def rows_with(rows, coeffs: np.ndarray) -> list:
    """Phase rows with alpha, beta, gamma, kappa replaced by coeffs (3, 4)."""
    out = []
    for q, r in enumerate(rows):
        d = dict(r)
        for t, k in enumerate(COEFFS):
            d[k] = float(coeffs[q, t])
        out.append(d)
    return out


class PELoop(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coeffs, film, Ez, P0, phi0, ckpt_every, tol):
        c = coeffs.detach().cpu().to(torch.float64).numpy()
        if getattr(film, "_adj_steps", 0) < Ez.size:
            film.adjoint_reserve(Ez.size, ckpt_every, tol)
        film.set_phase_table(rows_with(film.inp.phases, c))
        film.set_state(P0, phi0)
        m = film.forward_record(Ez)
        ctx.film = film
        return torch.from_numpy(m).to(coeffs.dtype)

    @staticmethod
    def backward(ctx, grad_m):
        g, _ = ctx.film.backward(grad_m.detach().cpu().to(torch.float64).numpy())
        return torch.from_numpy(g).to(grad_m.dtype), None, None, None, None, None, None


def pe_loop(coeffs: torch.Tensor, film, Ez, P0=None, phi0=None, ckpt_every: int = 50, tol: float = 1e-8):
    """Loop samples <Pmz>_n for the field sequence Ez (non-dimensional),
    differentiable in coeffs (3, 4)."""
    Ez = np.ascontiguousarray(Ez, np.float64)
    P0 = film.inp.P0 if P0 is None else P0
    return PELoop.apply(coeffs, film, Ez, P0, phi0, int(ckpt_every), float(tol))


def loop_mse(m: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """Mean squared error between a simulated and a target loop."""
    return torch.mean((m - target) ** 2)
