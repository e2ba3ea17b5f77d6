"""ImplicitEuler and ESDIRK23a on the device against the oracle.

The reference carries Newton/Jacobian/step-size state from one node to the
next inside one solver instance (ImplicitEuler.cpp:81-82,
ESDIRK23a.cpp:280-316), so its results depend on node order and thread
count.  The device keeps that state per node; the oracle's per-node mode is
the same algorithm on the CPU and is the parity target here.  Both use the
finite-difference Jacobian (ODE.cpp:47-67) and the no-pivot LU (ODE.cpp:69-100).

Tolerance.  The reference algorithms make discrete decisions (Jacobian
reuse, step rejection, step halving) on quantities near rounding level: the
ImplicitEuler Newton iteration runs until eta * relative_residual < 1e-13
(ImplicitODESolver.cpp:244).  The device and the host differ in the last
bit of exp/log/pow, so the bound is calibrated per case on the CPU: the
oracle is run twice, once from initial states moved by one ulp, and the
spread s of that pair is the algorithm's own resolution.  Then
    |x_gpu - x_cpu| <= max(1e-10, 10 s) * max(|x_cpu|, floor_s)
per state (floor as in test_gpu_explicit).  Where the one-ulp pair makes
the same decisions (equal counters), the device must make them too.
"""
import numpy as np
import pytest

import oracle
from test_gpu_explicit import assert_parity, dev_system, perturbed

pytestmark = pytest.mark.gpu


def ulp_sensitivity(variant, solver, params, y0, t):
    """Spread of the oracle itself between y0 and y0 moved by one ulp."""
    N = y0.shape[0]
    a = oracle.OracleSystem(variant, solver, N, params, per_node_state=True)
    b = oracle.OracleSystem(variant, solver, N, params, per_node_state=True)
    a.set_states(y0)
    b.set_states(np.nextafter(y0, np.inf))
    a.forward_times(t)
    b.forward_times(t)
    x, z = a.states(), b.states()
    floor = 1e-6 * np.max(np.abs(x), axis=0, keepdims=True)
    return float((np.abs(x - z) / np.maximum(np.abs(x), floor)).max()), a.counters(), b.counters()


@pytest.mark.parametrize("solver", ["ImplicitEuler", "ESDIRK23a"])
@pytest.mark.parametrize("variant,dt,steps,ldt", [
    ("tp06_epi", 0.1, 300, -1.0),
    ("tp06_epi", 0.1, 200, 0.05),
    ("beeler_reuter_1977", 0.5, 100, -1.0),
    ("fhn", 0.5, 100, -1.0),
    ("tp06_m", 0.05, 300, -1.0),
])
def test_implicit_matches_oracle_per_node(solver, variant, dt, steps, ldt):
    N = 65
    params = {"internal_time_step": ldt}
    o = oracle.OracleSystem(variant, solver, N, params, per_node_state=True)
    dev = dev_system(variant, solver, N, "fp64", params)
    y0 = perturbed(o.states(), 41)
    t = np.arange(steps + 1) * dt
    sens, c_a, c_b = ulp_sensitivity(variant, solver, params, y0, t)
    o.set_states(y0)
    dev.set_states(y0)
    o.forward_times(t)
    dev.forward_times(t)
    tol = max(1e-10, 10 * sens)
    err = assert_parity(dev.states(), o.states(), rtol=tol, what=f"{variant}/{solver}")
    oc, dc = o.counters(), dev.solver_counters()
    print(f"{variant}/{solver}: max rel err {err:.3e} (one-ulp spread {sens:.3e}), counters oracle {oc} device {dc}")
    if list(c_a[1:]) == list(c_b[1:]):
        assert list(oc[1:]) == list(dc[1:])


def test_config3_epi_subset_and_full_size_equality():
    """Config 3 shape: TP06 epi, dt 0.1 ms, init +-1%.  A 1,024-node subset
    for 100 ms against the oracle, and the full 262,144 nodes from identical
    initial states for 20 ms: all nodes stay bit-identical."""
    for solver in ("ESDIRK23a", "ImplicitEuler"):
        N = 1024
        o = oracle.OracleSystem("tp06_epi", solver, N, per_node_state=True)
        o.set_num_threads(8)
        dev = dev_system("tp06_epi", solver, N, "fp64")
        y0 = perturbed(o.states(), 43)
        o.set_states(y0)
        dev.set_states(y0)
        t = np.arange(1001) * 0.1
        sens, _, _ = ulp_sensitivity("tp06_epi", solver, {}, y0[:128], t)
        o.forward_times(t)
        dev.forward_times(t)
        err = assert_parity(dev.states(), o.states(), rtol=max(1e-10, 10 * sens), what=f"cfg3 subset {solver}")
        print(f"cfg3 subset {solver}: max rel err {err:.3e} (one-ulp spread on 128 nodes {sens:.3e})")
        big = dev_system("tp06_epi", solver, 262144, "fp64")
        big.set_states(np.tile(y0[0], (262144, 1)))
        big.forward_times(np.arange(201) * 0.1)
        s = big.states()
        assert np.all(s == s[0])
        np.testing.assert_allclose(s[0], dev_system_single(y0[0], solver), rtol=1e-12)


def dev_system_single(y, solver):
    o = oracle.OracleSystem("tp06_epi", solver, 1, per_node_state=True)
    o.set_states(y[None, :])
    o.forward_times(np.arange(201) * 0.1)
    return o.states()[0]


def test_implicit_euler_failure_reports_reference_message():
    """Newton cannot converge (max_iterations = 0) and the step would have to
    shrink below min_dt: the call fails with the reference's message
    (ImplicitEuler.cpp:132-141); every kernel still returns normally."""
    from goss_amd._lib import GossHipError

    params = {"max_iterations": 0, "min_dt": 0.5}
    o = oracle.OracleSystem("tp06_m", "ImplicitEuler", 4, params, per_node_state=True)
    with pytest.raises(RuntimeError) as cpu_err:
        o.forward(0.0, 0.1)
    dev = dev_system("tp06_m", "ImplicitEuler", 4, "fp64", params)
    with pytest.raises(GossHipError) as gpu_err:
        dev.forward(0.0, 0.1)
    assert str(gpu_err.value) == str(cpu_err.value)
    assert "Newtons solver failed to converge as dt become smaller than 'min_dt' 0.5" in str(gpu_err.value)
