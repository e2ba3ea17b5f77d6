"""The generated GRL1 adjoint (grl1_vjp in goss_amd/csrc/generated/*.h) on the CPU.

The same header the device adjoint kernels instantiate is compiled into the
oracle library (double, host libm).  Its vector-Jacobian product of one GRL1
step must equal
* torch.autograd through the plain-PyTorch GRL1 step (tests/torch_reference.py),
* central finite differences of the oracle's own GRL1 step, row by row.
"""
import numpy as np
import pytest
import torch

import oracle
from torch_reference import TorchModel

MODELS = ["tp06_m_fp6", "tp06_m", "fhn", "beeler_reuter_1977", "oscilator", "lorentz", "fitzhughnagumo"]


def _state(m, rng, N):
    return m.initial_conditions[None, :] * (1 + 0.05 * rng.uniform(-1, 1, (N, m.num_states)))


@pytest.mark.parametrize("name", MODELS)
@pytest.mark.parametrize("t", [5.0, 10.3])
def test_grl1_vjp_matches_torch_autograd(name, t):
    tm = TorchModel(name)
    m = oracle.Variant(name)
    rng = np.random.default_rng(7)
    N = 6
    y = _state(m, rng, N)
    fpi = m.field_parameter_indices
    fp = np.array([m.default_parameters[k] for k in fpi])[None, :].repeat(N, 0) * np.exp(
        0.1 * rng.normal(size=(N, len(fpi))))
    h = 0.02
    yt = torch.tensor(y, requires_grad=True)
    ft = torch.tensor(fp, requires_grad=True)
    lam = rng.normal(size=(N, m.num_states))
    (tm.grl1_step(yt, t, h, tm.params(ft)) * torch.tensor(lam)).sum().backward()
    for n in range(N):
        p = m.default_parameters.copy()
        for j, k in enumerate(fpi):
            p[k] = fp[n, j]
        lp, fb = m.grl1_vjp(y[n], t, h, lam[n], p)
        g = yt.grad[n].numpy()
        np.testing.assert_allclose(lp, g, rtol=1e-9, atol=1e-12 * np.abs(g).max())
        if fpi:
            np.testing.assert_allclose(fb, ft.grad[n].numpy(), rtol=1e-9, atol=1e-14)


@pytest.mark.parametrize("name", ["tp06_m", "beeler_reuter_1977", "lorentz"])
def test_grl1_vjp_rows_match_central_differences(name):
    m = oracle.Variant(name)
    rng = np.random.default_rng(8)
    y = _state(m, rng, 1)[0]
    p = m.default_parameters
    t, h, S = 10.3, 0.01, m.num_states
    J = np.zeros((S, S))

    def central(j, e):
        yp, ym = y.copy(), y.copy()
        yp[j] += e
        ym[j] -= e
        return (m.grl1_step(yp, t, h, p) - m.grl1_step(ym, t, h, p)) / (2 * e)

    for j in range(S):
        # Richardson-extrapolated central differences (error O(e^4)).
        e = 1e-3 * max(abs(y[j]), 1e-3)
        J[:, j] = (4 * central(j, e / 2) - central(j, e)) / 3
    for i in range(S):
        lam = np.zeros(S)
        lam[i] = 1.0
        lp, _ = m.grl1_vjp(y, t, h, lam, p)
        np.testing.assert_allclose(lp, J[i], rtol=1e-5, atol=1e-7 * max(np.abs(J[i]).max(), 1.0))


def test_phi1_series_and_closed_form_agree():
    # m_phi1 / m_dphi1 switch from the Taylor series to the closed form at
    # |x| = 1e-3 (double); the GRL1 adjoint of a step with a*h at that point
    # must be continuous.  Checked through linear_decay (y' = a y + b), whose
    # one-step map y' = y + (a y + b) h phi1(a h) has d y'/d y = exp(a h).
    m = oracle.Variant("linear_decay")
    for x in (-2e-3, -1.0001e-3, -0.9999e-3, -1e-5, 1e-5, 0.9999e-3, 1.0001e-3, 0.3):
        h = 0.5
        a = x / h
        p = m.default_parameters.copy()
        p[m.field_parameter_indices[0]] = a
        lp, fb = m.grl1_vjp(np.array([1.3]), 0.0, h, np.array([1.0]), p)
        np.testing.assert_allclose(lp[0], np.exp(a * h), rtol=1e-12)


@pytest.mark.parametrize("name", ["tp06_m_fp6", "beeler_reuter_1977", "oscilator"])
def test_grl1_vjp_tape_variant_agrees(name):
    # grl1_vjp_tape (the FP64 device kernels' form) runs the same reverse sweep
    # through memory, without the common-subexpression temporaries of grl1_vjp:
    # equal to rounding.
    m = oracle.Variant(name)
    rng = np.random.default_rng(9)
    for t in (5.0, 10.3, 11.7):
        y = _state(m, rng, 1)[0]
        lam = rng.normal(size=m.num_states)
        a = m.grl1_vjp(y, t, 0.02, lam)
        b = m.grl1_vjp(y, t, 0.02, lam, tape=True)
        np.testing.assert_allclose(b[0], a[0], rtol=1e-12, atol=1e-15 * np.abs(a[0]).max())
        np.testing.assert_allclose(b[1], a[1], rtol=1e-12, atol=1e-15 * max(np.abs(a[1]).max(initial=0), 1e-300))


@pytest.mark.parametrize("name", MODELS)
@pytest.mark.parametrize("t", [5.0, 10.3])
def test_rl1_and_eval_vjp_match_torch_autograd(name, t):
    # rl1_vjp: one RL1 step; eval_vjp: the right-hand side alone (the explicit schemes' building block).
    tm = TorchModel(name)
    m = oracle.Variant(name)
    rng = np.random.default_rng(11)
    N = 4
    y = _state(m, rng, N)
    fpi = m.field_parameter_indices
    fp = np.array([m.default_parameters[k] for k in fpi])[None, :].repeat(N, 0) * np.exp(
        0.1 * rng.normal(size=(N, len(fpi))))
    h = 0.02
    lam = rng.normal(size=(N, m.num_states))
    for kind in ("rl1", "eval"):
        if kind == "rl1":
            lin, _ = m.linearized_eval(y[0], t)
            if np.any(lin[np.array(m.linear_terms, dtype=bool)] == 0):
                continue  # RL1 divides by a zero linear coefficient here (RL1.h:68-77): the step is NaN
        yt = torch.tensor(y, requires_grad=True)
        ft = torch.tensor(fp, requires_grad=True)
        out = tm.rl1_step(yt, t, h, tm.params(ft)) if kind == "rl1" else tm.rhs(yt, t, tm.params(ft))
        (out * torch.tensor(lam)).sum().backward()
        for n in range(N):
            p = m.default_parameters.copy()
            for j, k in enumerate(fpi):
                p[k] = fp[n, j]
            for tape in (False, True):
                if kind == "rl1":
                    gy, gp = m.rl1_vjp(y[n], t, h, lam[n], p, tape=tape)
                else:
                    gy, gp = m.eval_vjp(y[n], t, lam[n], p, tape=tape)
                g = yt.grad[n].numpy()
                np.testing.assert_allclose(gy, g, rtol=1e-9, atol=1e-12 * np.abs(g).max(), err_msg=f"{kind} {tape}")
                if fpi:
                    np.testing.assert_allclose(gp, ft.grad[n].numpy(), rtol=1e-9, atol=1e-14)
