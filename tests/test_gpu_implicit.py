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
per state, with floor_s = 1e-6 of the state's largest magnitude over the
nodes at the start and the end of the run.  Where the one-ulp pair makes
the same decisions (equal counters), the device must make them too and the
bound is capped at 1e-8.  Bit-level parity of these solvers is pinned
separately: the parity build (host-identical FP64 elementary functions)
equals the host restatement bit for bit (test_gpu_parity_build.py).
"""
import numpy as np
import pytest

import oracle
from test_gpu_explicit import assert_parity, dev_system, perturbed

pytestmark = pytest.mark.gpu


def state_scale(*arrays):
    """Per-state magnitude over the run: the largest |value| over nodes of the given states."""
    return np.max([np.max(np.abs(a), axis=0) for a in arrays], axis=0)[None, :]


def ulp_sensitivity(variant, solver, params, y0, t, threads=1):
    """Spread of the oracle itself between y0 and y0 moved by one ulp."""
    N = y0.shape[0]
    a = oracle.OracleSystem(variant, solver, N, params, per_node_state=True)
    b = oracle.OracleSystem(variant, solver, N, params, per_node_state=True)
    a.set_num_threads(threads)
    b.set_num_threads(threads)
    a.set_states(y0)
    b.set_states(np.nextafter(y0, np.inf))
    a.forward_times(t)
    b.forward_times(t)
    x, z = a.states(), b.states()
    floor = 1e-6 * state_scale(y0, x)
    return float((np.abs(x - z) / np.maximum(np.abs(x), floor)).max()), a.counters(), b.counters()


@pytest.mark.parametrize("solver", ["ImplicitEuler", "ESDIRK23a", "ThetaSolver", "BasicImplicitEuler"])
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
    same_decisions = list(c_a[1:]) == list(c_b[1:])
    tol = max(1e-10, 10 * sens)
    if same_decisions:
        tol = min(tol, 1e-8)
    err = assert_parity(dev.states(), o.states(), rtol=tol, what=f"{variant}/{solver}",
                        scale=state_scale(y0, o.states()))
    oc, dc = o.counters(), dev.solver_counters()
    print(f"{variant}/{solver}: max rel err {err:.3e} (one-ulp spread {sens:.3e}, bound {tol:.1e}), "
          f"counters oracle {oc} device {dc}")
    if same_decisions:
        assert list(oc[1:]) == list(dc[1:])


def test_esdirk4o32_matches_oracle_per_node():
    """ESDIRK4O32 as the reference computes it (1e4-1e5 steps per interval on FHN), 64 nodes,
    three intervals of 0.05 ms: the oracle's one-ulp calibration of the other implicit solvers."""
    N = 64
    o = oracle.OracleSystem("fhn", "ESDIRK4O32", N, per_node_state=True)
    o.set_num_threads(8)
    dev = dev_system("fhn", "ESDIRK4O32", N, "fp64")
    y0 = perturbed(o.states(), 44)
    t = np.arange(4) * 0.05
    sens, c_a, c_b = ulp_sensitivity("fhn", "ESDIRK4O32", {}, y0, t, threads=8)
    o.set_states(y0)
    dev.set_states(y0)
    o.forward_times(t)
    dev.forward_times(t)
    same_decisions = list(c_a[1:]) == list(c_b[1:])
    tol = max(1e-10, 10 * sens)
    if same_decisions:
        tol = min(tol, 1e-8)
    err = assert_parity(dev.states(), o.states(), rtol=tol, what="fhn/ESDIRK4O32", scale=state_scale(y0, o.states()))
    oc, dc = o.counters(), dev.solver_counters()
    print(f"fhn/ESDIRK4O32: max rel err {err:.3e} (one-ulp spread {sens:.3e}, bound {tol:.1e}), "
          f"counters oracle {oc} device {dc}")
    assert dc[3] > 0
    if same_decisions:
        assert list(oc[1:]) == list(dc[1:])


def test_tp06_epi_implicit_subset_and_full_size_equality():
    """TP06 epi, dt 0.1 ms, init +-1%.  A 1,024-node subset
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
        sens, c_a, c_b = ulp_sensitivity("tp06_epi", solver, {}, y0, t, threads=8)
        o.forward_times(t)
        dev.forward_times(t)
        tol = max(1e-10, 10 * sens)
        if list(c_a[1:]) == list(c_b[1:]):
            tol = min(tol, 1e-8)
        err = assert_parity(dev.states(), o.states(), rtol=tol, what=f"cfg3 subset {solver}",
                            scale=state_scale(y0, o.states()))
        print(f"cfg3 subset {solver}: max rel err {err:.3e} (one-ulp spread {sens:.3e}, bound {tol:.1e})")
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


def _epi_run(kind, solver, params, y0, t, jacobian="fd"):
    """TP06 epi with the M-cell's 52 pA/pF stimulus (the epi .ode ships with
    stim_amplitude 0), so that the window holds an upstroke."""
    if kind == "oracle":
        s = oracle.OracleSystem("tp06_epi", solver, y0.shape[0], params, per_node_state=True)
        s.set_num_threads(8)
    else:
        s = dev_system("tp06_epi", solver, y0.shape[0], "fp64", {**params, "jacobian": jacobian})
    s.set_parameter("stim_amplitude", 52.0)
    s.set_states(y0)
    s.forward_times(t)
    return s.states(), (s.counters() if kind == "oracle" else s.solver_counters())


def _diff(a, b):
    floor = 1e-6 * np.max(np.abs(b), axis=0, keepdims=True)
    return float(np.abs(a[:, 17] - b[:, 17]).max()), float((np.abs(a - b) / np.maximum(np.abs(b), floor)).max())


@pytest.mark.parametrize("solver,rtol", [("ImplicitEuler", 1e-6), ("ImplicitEuler", None), ("ESDIRK23a", None)])
def test_fp32_jacobian_precision(solver, rtol):
    """FP32 right-hand side in the Jacobian, FP64 Newton: the Jacobian from forward differences of the
    FP32 right-hand side, the Newton iteration (residual, LU, substitutions)
    in FP64.  TP06 epi, dt 0.1 ms, 30 ms through the stimulus and upstroke,
    against the FP64 oracle with the reference's Jacobian; ImplicitEuler at
    the config's relative_tolerance 1e-6 and at the default (ESDIRK23a's
    Newton tolerance is fixed, ESDIRK23a.cpp:34,189).

    A different Jacobian changes the Newton iterates within what the stopping
    rule allows, and with them the solvers' discrete decisions: where
    ImplicitEuler's Newton iteration is declared divergent and the step is
    halved (ImplicitEuler.cpp:108-150).  A different halving pattern is a
    different, equally valid discretisation, so ImplicitEuler is bounded by
    its own discretisation error: the FP32-Jacobian result must lie within
    the distance between the oracle at dt 0.1 and at dt 0.01, in every state
    (measured: 0.083 mV in V against 0.21 mV, worst state ratio 0.79; a
    one-ulp change of y0 flips no halving decision here, a different
    Jacobian flips a few).  ESDIRK23a controls its local error (tol 1e-5): within
    1e-3 mV in V and 1e-3 relative in every state (floor 1e-6 of the state's
    maximum).  The figures, and the oracle's own one-ulp spread, are printed."""
    N = 256
    y0 = perturbed(oracle.OracleSystem("tp06_epi", solver, N, {}, per_node_state=True).states(), 47)
    t = np.arange(301) * 0.1
    params = {} if rtol is None else {"relative_tolerance": rtol}
    ref, oc = _epi_run("oracle", solver, params, y0, t)
    got, dc = _epi_run("device", solver, params, y0, t, "fd_fp32")
    ulp, _ = _epi_run("oracle", solver, params, np.nextafter(y0, np.inf), t)
    dv, rel = _diff(got, ref)
    dv_ulp, rel_ulp = _diff(ulp, ref)
    msg = (f"{solver} rtol {rtol}: FP32 Jacobian vs oracle max |dV| {dv:.3e} mV, max rel {rel:.3e}; oracle one-ulp "
           f"spread |dV| {dv_ulp:.3e} rel {rel_ulp:.3e}; V {ref[:, 17].min():.1f}..{ref[:, 17].max():.1f} mV; "
           f"counters oracle {oc} device {dc}")
    assert ref[:, 17].max() > 0  # the upstroke is in the window
    if solver == "ImplicitEuler":
        fine, _ = _epi_run("oracle", solver, params, y0, np.arange(3001) * 0.01)
        floor = 1e-6 * np.max(np.abs(fine), axis=0, keepdims=True)
        disc = np.abs(ref - fine).max(axis=0)
        err = np.abs(got - ref).max(axis=0)
        print(msg + f"; ImplicitEuler dt 0.1 vs dt 0.01 |dV| {disc[17]:.3e} mV, worst ratio "
                    f"{(err / np.maximum(disc, floor[0])).max():.3e}")
        assert np.all(err <= np.maximum(disc, floor[0])), msg
    else:
        print(msg)
        assert dv <= 1e-3 and rel <= 1e-3, msg


def test_unknown_jacobian_mode_is_rejected():
    from goss_amd._lib import GossHipError

    with pytest.raises(GossHipError, match="Jacobian"):
        dev_system("tp06_epi", "ImplicitEuler", 4, "fp64", {"jacobian": 7})
