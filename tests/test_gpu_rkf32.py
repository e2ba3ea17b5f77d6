"""RKF32 on the device against the oracle.

The reference keeps RKF32's controller state (the initial step guess _ldt
and facmax, AdaptiveExplicitSolver.cpp:216-241) in one solver instance per
thread and carries it from node to node; the device keeps it per node, as
the oracle's per-node mode does (the oracle's carried mode is pinned bit for
bit to the reference build in test_oracle.py).

Tolerance (FP64).  The controller accepts or rejects steps on an error norm;
the device and the host differ in the last bit of pow/exp, so the bound is
calibrated per case as for the implicit solvers: the oracle is run from y0
and from y0 moved by one ulp, and with s the spread of that pair,
    |x_gpu - x_cpu| <= max(1e-10, 10 s) * max(|x_cpu|, floor_s).
Where the one-ulp pair takes the same steps (equal counters) the device must
take them too.

FP32 runs the states, stages and error vector in FP32 (step sizes and time
in FP64) and is checked against the FP64 oracle with stated bounds.
"""
import numpy as np
import pytest

import oracle
from test_gpu_explicit import assert_parity, dev_system, perturbed
from test_gpu_implicit import state_scale, ulp_sensitivity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant,dt,steps,params,spread", [
    ("lorentz", 0.01, 100, {}, 0.05),
    ("lorentz", 0.05, 40, {"atol": 1e-7, "rtol": 1e-9}, 0.05),
    ("fitzhughnagumo", 0.5, 100, {}, 0.01),
    ("tp06_m", 0.1, 150, {}, 0.01),
    ("oscilator", 0.1, 100, {"atol": 1e-8}, 0.01),
])
def test_rkf32_fp64_matches_oracle_per_node(variant, dt, steps, params, spread):
    N = 65
    o = oracle.OracleSystem(variant, "RKF32", N, params, per_node_state=True)
    dev = dev_system(variant, "RKF32", N, "fp64", params)
    y0 = perturbed(o.states(), 51, spread)
    t = np.arange(steps + 1) * dt
    sens, c_a, c_b = ulp_sensitivity(variant, "RKF32", params, y0, t)
    o.set_states(y0)
    dev.set_states(y0)
    o.forward_times(t)
    dev.forward_times(t)
    same_decisions = list(c_a[1:]) == list(c_b[1:])
    tol = max(1e-10, 10 * sens)
    if same_decisions:
        tol = min(tol, 1e-8)
    err = assert_parity(dev.states(), o.states(), rtol=tol, what=f"{variant}/RKF32",
                        scale=state_scale(y0, o.states()))
    oc, dc = o.counters(), dev.solver_counters()
    print(f"{variant}/RKF32: max rel err {err:.3e} (one-ulp spread {sens:.3e}, bound {tol:.1e}), "
          f"counters oracle {oc} device {dc}")
    if same_decisions:
        assert list(oc[1:]) == list(dc[1:])


def test_rkf32_field_parameters_and_solve_rows():
    # Per-node field parameters reach the RHS, and solve() writes one row of
    # field states per outer step (rows 1..n-1; ODESystemSolver.cpp:81-99).
    N = 40
    o = oracle.OracleSystem("oscilator", "RKF32", N, per_node_state=True)
    dev = dev_system("oscilator", "RKF32", N, "fp64")
    rng = np.random.Generator(np.random.PCG64(52))
    fp = o.field_parameters() * (1.0 + rng.uniform(-0.2, 0.2, (N, 2)))
    o.set_field_parameters(fp)
    dev.set_field_parameters(fp)
    t = np.arange(31) * 0.1
    want = o.solve(t)
    got = np.zeros_like(want)
    dev.solve(t, got)
    assert_parity(got[1:].reshape(len(t) - 1, -1), want[1:].reshape(len(t) - 1, -1), rtol=1e-9, what="rows")


def test_rkf32_fp32_accuracy():
    # FitzHugh-Nagumo over 50 ms and Lorenz over 0.5 time units (before the
    # chaotic divergence of nearby trajectories grows past FP32 rounding).
    for variant, dt, steps, bound in (("fitzhughnagumo", 0.5, 100, 2e-3), ("lorentz", 0.01, 50, 1e-3)):
        N = 64
        o = oracle.OracleSystem(variant, "RKF32", N, per_node_state=True)
        dev = dev_system(variant, "RKF32", N, "fp32")
        y0 = perturbed(o.states(), 53, 0.05)
        o.set_states(y0)
        dev.set_states(y0)
        t = np.arange(steps + 1) * dt
        o.forward_times(t)
        dev.forward_times(t)
        a, b = dev.states(), o.states()
        scale = np.max(np.abs(b), axis=0, keepdims=True)
        err = np.max(np.abs(a - b) / scale)
        print(f"{variant} FP32 RKF32: max error relative to the state's range {err:.3e}")
        assert err <= bound


def test_rkf32_stuck_node_reports_error():
    # A NaN state makes every error norm non-finite (taken as 2000,
    # AdaptiveExplicitSolver.cpp:195-196): the step size is halved until it
    # underflows and the interval end is never reached.  The device and the
    # oracle stop that node after the same sub-step cap and report it.
    y0 = np.array([[1.0, 0.0, 0.0], [np.nan, 0.0, 0.0]])
    o = oracle.OracleSystem("lorentz", "RKF32", 2, per_node_state=True)
    dev = dev_system("lorentz", "RKF32", 2, "fp64")
    o.set_states(y0)
    dev.set_states(y0)
    with pytest.raises(RuntimeError) as a:
        o.forward(0.0, 0.01)
    from goss_amd._lib import GossHipError

    with pytest.raises(GossHipError) as b:
        dev.forward(0.0, 0.01)
    assert "RKF32 did not reach the end of the interval" in str(a.value)
    assert "RKF32 did not reach the end of the interval" in str(b.value)
