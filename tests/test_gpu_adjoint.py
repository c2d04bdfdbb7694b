"""Reverse mode (hzo_forward_record / hzo_backward) against finite differences.

The loss is L = sum_n (m_n - t_n)^2 over the loop samples m_n = <Pmz> after
each step.  The FP64 adjoint gradient with respect to the per-phase
(alpha, beta, gamma, kappa) is checked against central differences of the
FP64 forward run (step 1e-4 relative, PCG solved to 1e-13 so the solve error
stays far below the difference quotient).  The FP32 adjoint (FP32 vectors,
FP64 reductions) is checked against the FP64 adjoint, and the checkpoint
period must not change a single bit of the FP64 gradient.
"""
import numpy as np
import pytest

import cmsl_amd as H
from cmsl_amd import _lib as L
from cmsl_amd.adjoint import coeff_matrix, rows_with
from helpers import film_inputs

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu]", *a, flush=True)


def _film(inp, precision=L.PREC_FP64, tol=1e-13, cg_fused=1):
    return H.Film(inp, precision=precision, pcg_rule=L.PCG_RULE_RELRES, tol_rel=tol, tol_abs=1e-30,
                  max_iter=4000, refine_every=20, cg_fused=cg_fused)


def _inputs():
    inp = film_inputs(24, 20, 8, 6, fracs=(0.3, 0.5, 0.2), seed=4, jitter=0.2)
    Ez = inp.field_nd(H.triangle_wave(24, 3.0, 24))
    return inp, Ez


def _loss_and_seed(m, t):
    return float(np.sum((m - t) ** 2)), 2.0 * (m - t)


def _run(film, inp, Ez, coeffs):
    film.set_phase_table(rows_with(inp.phases, coeffs))
    film.set_state(inp.P0)
    return film.forward_record(Ez)


def test_adjoint_gradient_matches_central_differences():
    inp, Ez = _inputs()
    S = Ez.size
    _say("adjoint vs FD 24x20x8, 24 steps, FE/AFE/DE")
    f = _film(inp)
    f.adjoint_reserve(S, ckpt_every=7, tol=1e-13)
    base = coeff_matrix(inp.phases)
    m0 = _run(f, inp, Ez, base)
    t = 0.9 * m0 + 0.01
    _, seed = _loss_and_seed(m0, t)
    g, its = f.backward(seed)
    _say("adjoint grad", g.tolist(), "adjoint PCG iterations", its)
    fd = np.zeros((3, 4))
    scale = np.max(np.abs(base), axis=0)
    for p in range(3):
        if not np.any(inp.grain_phase == p):
            continue
        for q in range(4):
            h = 1e-4 * max(abs(base[p, q]), scale[q])
            cp, cm = base.copy(), base.copy()
            cp[p, q] += h
            cm[p, q] -= h
            lp, _ = _loss_and_seed(_run(f, inp, Ez, cp), t)
            lm, _ = _loss_and_seed(_run(f, inp, Ez, cm), t)
            fd[p, q] = (lp - lm) / (2.0 * h)
    err = np.abs(g - fd)
    _say("FD grad", fd.tolist(), "max abs err", err.max(), "max |g|", np.abs(fd).max())
    assert np.all(err <= 2e-5 * np.abs(fd).max() + 1e-4 * np.abs(fd)), (g, fd)
    f.close()


def test_checkpoint_period_leaves_gradient_bit_identical():
    inp, Ez = _inputs()
    S = Ez.size
    _say("checkpoint periods 1, 5, 24")
    grads = []
    for K in (1, 5, S):
        f = _film(inp, tol=1e-10)
        f.adjoint_reserve(S, ckpt_every=K, tol=1e-10)
        m = _run(f, inp, Ez, coeff_matrix(inp.phases))
        g, _ = f.backward(2.0 * (m - 0.5 * m))
        grads.append(g)
        f.close()
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


def test_backward_restores_the_recorded_final_state():
    inp, Ez = _inputs()
    f = _film(inp, tol=1e-10)
    f.adjoint_reserve(Ez.size, ckpt_every=5, tol=1e-10)
    m = _run(f, inp, Ez, coeff_matrix(inp.phases))
    P1, phi1, _ = f.get_state()
    f.backward(np.ones_like(m))
    P2, phi2, _ = f.get_state()
    assert np.array_equal(P1, P2) and np.array_equal(phi1, phi2)
    f.close()


@pytest.mark.parametrize("fused", [1, 3])
def test_fp32_adjoint_matches_fp64_adjoint(fused):
    # fused 3: the adjoint solves run the two-stencil single-reduction CG
    # (with the solve's own solution vector), the form large films take
    inp, Ez = _inputs()
    S = Ez.size
    _say(f"fp32 adjoint vs fp64 adjoint, cg_fused={fused}")
    out = {}
    for prec in (L.PREC_FP64, L.PREC_FP32_REFINED):
        f = _film(inp, precision=prec, tol=1e-8, cg_fused=fused)
        if prec == L.PREC_FP32_REFINED:
            assert f.cg_form == (2 if fused == 3 else 1)
        f.adjoint_reserve(S, ckpt_every=8, tol=1e-8)
        m = _run(f, inp, Ez, coeff_matrix(inp.phases))
        t = np.linspace(-0.01, 0.01, S)
        g, _ = f.backward(2.0 * (m - t))
        out[prec] = (m, g)
        f.close()
    m64, g64 = out[L.PREC_FP64]
    m32, g32 = out[L.PREC_FP32_REFINED]
    rel = np.abs(g32 - g64).max() / np.abs(g64).max()
    _say("loop max dev", np.abs(m32 - m64).max(), "grad rel dev", rel)
    assert rel < 2e-3


def test_autograd_function_drives_the_hip_adjoint():
    torch = pytest.importorskip("torch")
    from cmsl_amd.adjoint import loop_mse, pe_loop
    inp, Ez = _inputs()
    f = _film(inp, tol=1e-10)
    base = coeff_matrix(inp.phases)
    coeffs = torch.tensor(base, dtype=torch.float64, requires_grad=True)
    m = pe_loop(coeffs, f, Ez, ckpt_every=6, tol=1e-10)
    target = torch.linspace(-0.02, 0.02, Ez.size, dtype=torch.float64)
    loss = loop_mse(m, target)
    loss.backward()
    # the same gradient straight from the C ABI
    m_np = _run(f, inp, Ez, base)
    g, _ = f.backward(2.0 * (m_np - target.numpy()) / Ez.size)
    assert np.allclose(coeffs.grad.numpy(), g, rtol=1e-12, atol=0.0)
    f.close()
