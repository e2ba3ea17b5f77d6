"""Parity of the HIP explicit-solver path with the reference and the oracle.

FP64 tolerance (stated here, applied per state after every run):
    |x_gpu - x_ref| <= 1e-10 * max(|x_ref|, floor_s),
    floor_s = 1e-6 * max over nodes of |x_ref[:, s]|
The floor only matters for states that pass near zero (gates such as r, d,
j); the GPU and CPU differ only through the last-bit differences of the
device and host exp/log/pow routines.

FP32 / FP32C are approximations of the FP64 path and are checked against
stated accuracy bounds (max |dV| in mV and per-state relative error).
"""
import numpy as np
import pytest

import oracle
from conftest import golden

pytestmark = pytest.mark.gpu

EXPLICIT = ["ExplicitEuler", "RK2", "RK4", "RL1", "RL2", "GRL1", "GRL2"]
RTOL = 1e-10


def dev_system(*a, **k):
    from goss_amd.system import DeviceSystem

    return DeviceSystem(*a, **k)


def assert_parity(a, b, rtol=RTOL, what="", scale=None):
    """max over nodes and states of |a - b| / max(|b|, 1e-6 * scale), scale per state: the
    largest |b| over the nodes unless given (a state's magnitude over the run)."""
    assert a.shape == b.shape
    assert np.all(np.isfinite(a)) == np.all(np.isfinite(b))
    if scale is None:
        scale = np.max(np.abs(b), axis=0, keepdims=True)
    floor = 1e-6 * np.asarray(scale)
    err = np.abs(a - b) / np.maximum(np.abs(b), np.maximum(floor, 1e-300))
    worst = np.unravel_index(np.argmax(err), err.shape)
    assert err.max() <= rtol, f"{what}: max rel err {err.max():.3e} at {worst}: {a[worst]} vs {b[worst]}"
    return err.max()


def perturbed(sys_states, seed, spread=0.01):
    rng = np.random.Generator(np.random.PCG64(seed))
    return sys_states * (1.0 + rng.uniform(-spread, spread, sys_states.shape))


def test_tp06_grl1_fp64_1024_nodes_matches_reference():
    """TP06 M-cell GRL1 FP64: 1,024 nodes, dt 0.01 ms, 10,000 steps (through a full AP)."""
    d = np.load(golden("ref_tp06_grl1_1024x10000.npz"))
    dev = dev_system("tp06_m", "GRL1", 1024, "fp64")
    dev.set_states(d["y0"])
    t = d["t"]
    trace = [dev.states()[:8, 17]]
    for k in range(0, len(t) - 1, 100):
        dev.forward_times(t[k:k + 101])
        trace.append(dev.states()[:8, 17])
    err = assert_parity(dev.states(), d["final"], what="1,024-node final states")
    np.testing.assert_allclose(np.array(trace), d["V_trace"], rtol=0, atol=1e-8)
    print(f"1,024 nodes x 10,000 steps: max rel err {err:.3e}")


@pytest.mark.parametrize("solver", ["GRL1", "RL2"])
def test_field_parameter_window_matches_reference(solver):
    d = np.load(golden("ref_tp06_fp3_window.npz"))
    N = d["y0"].shape[0]
    dev = dev_system("tp06_m_fp3", solver, N, "fp64")
    dev.set_states(d["y0"])
    dev.set_field_parameters(d["field_parameters"])
    dev.forward_times(np.arange(int(d["steps"]) + 1) * float(d["dt"]))
    assert_parity(dev.states(), d[f"final_{solver}"], what=solver)


@pytest.mark.parametrize("solver", EXPLICIT)
@pytest.mark.parametrize("variant,ldt,dt,steps", [
    ("tp06_m", -1.0, 0.01, 1500),
    ("tp06_m_fp3", 0.005, 0.02, 600),
    ("beeler_reuter_1977", 0.01, 0.1, 300),
    ("fhn", -1.0, 0.1, 300),
    ("fitzhughnagumo", 0.01, 0.05, 300),
    ("lorentz", 0.001, 0.01, 200),
])
def test_explicit_solver_matches_oracle(solver, variant, ldt, dt, steps):
    N = 65  # ragged: one full wavefront plus one lane
    o = oracle.OracleSystem(variant, solver, N, {"internal_time_step": ldt})
    dev = dev_system(variant, solver, N, "fp64", {"internal_time_step": ldt})
    y0 = perturbed(o.states(), 21)
    o.set_states(y0)
    dev.set_states(y0)
    if o.variant.num_field_parameters:
        rng = np.random.Generator(np.random.PCG64(22))
        fp = o.field_parameters() * np.exp(rng.normal(0, 0.2, o.field_parameters().shape))
        o.set_field_parameters(fp)
        dev.set_field_parameters(fp)
    t = np.arange(steps + 1) * dt
    o.forward_times(t)
    dev.forward_times(t)
    a, b = dev.states(), o.states()
    if not np.all(np.isfinite(b)):
        # An unstable explicit scheme (RK4 on TP06 at dt 0.01) diverges on both.
        assert not np.all(np.isfinite(a))
        return
    assert_parity(a, b, what=f"{variant}/{solver}")


@pytest.mark.parametrize("N", [1, 63, 64, 65, 1000])
def test_node_counts(N):
    o = oracle.OracleSystem("tp06_m", "GRL1", N)
    dev = dev_system("tp06_m", "GRL1", N, "fp64")
    y0 = perturbed(o.states(), N)
    o.set_states(y0)
    dev.set_states(y0)
    t = np.arange(1201) * 0.01
    o.forward_times(t)
    dev.forward_times(t)
    assert_parity(dev.states(), o.states(), what=f"N={N}")


def test_forward_single_steps_equal_fused():
    N = 300
    a = dev_system("tp06_m", "GRL1", N, "fp64")
    b = dev_system("tp06_m", "GRL1", N, "fp64")
    y0 = perturbed(a.states(), 5)
    a.set_states(y0)
    b.set_states(y0)
    t = np.arange(1101) * 0.01
    for k in range(1100):
        a.forward(t[k], t[k + 1] - t[k])
    b.set_steps_per_launch(7)
    b.forward_times(t)
    assert np.array_equal(a.states(), b.states())


def test_solve_matches_oracle_rows():
    N = 64
    o = oracle.OracleSystem("tp06_m", "GRL1", N)
    dev = dev_system("tp06_m", "GRL1", N, "fp64")
    y0 = perturbed(o.states(), 8)
    o.set_states(y0)
    dev.set_states(y0)
    t = np.linspace(0, 20, 2001)
    for tangled in (True, False):
        ref = o.solve(t, tangled=tangled) if tangled else None
        out = np.zeros((len(t), N, 1)) if tangled else np.zeros((len(t), 1, N))
        dev.set_states(y0)
        dev.solve(t, out, tangled=tangled)
        if tangled:
            np.testing.assert_allclose(out[1:], ref[1:], rtol=1e-10, atol=1e-9)
            assert np.all(out[0] == 0)  # row 0 is the caller's
        else:
            o2 = oracle.OracleSystem("tp06_m", "GRL1", N)
            o2.set_states(y0)
            ref_u = o2.solve(t, tangled=False)
            np.testing.assert_allclose(out[1:], ref_u[1:], rtol=1e-10, atol=1e-9)


def test_layout_roundtrips_fp64_and_fp32c():
    N = 77
    rng = np.random.Generator(np.random.PCG64(3))
    for prec in ("fp64", "fp32c", "fp32"):
        dev = dev_system("tp06_m_fp3", "GRL1", N, prec)
        y = perturbed(dev.states(), 4, 0.3)
        dev.set_states(y)
        back = dev.states()
        if prec == "fp64":
            assert np.array_equal(back, y)
        elif prec == "fp32c":
            # hi + lo holds ~48 bits for compensated states, FP32 for gates.
            comp = np.array(dev.model.linear_terms) == 0
            np.testing.assert_allclose(back[:, comp], y[:, comp], rtol=1e-13)
            np.testing.assert_allclose(back[:, ~comp], y[:, ~comp], rtol=1e-7)
        else:
            np.testing.assert_allclose(back, y, rtol=1e-7)
        # field states, tangled and untangled
        fs = rng.uniform(-90, 20, (N, 1))
        dev.set_field_states(fs)
        out = np.zeros((N, 1))
        dev.get_field_states(out)
        np.testing.assert_allclose(out, fs, rtol=1e-13 if prec != "fp32" else 1e-7)
        out_u = np.zeros((1, N))
        dev.get_field_states(out_u, tangled=False)
        np.testing.assert_allclose(out_u[0], fs[:, 0], rtol=1e-13 if prec != "fp32" else 1e-7)
        # components
        comp_out = np.full((N, 1), -1.0)
        dev.get_field_state_components(comp_out, [0])
        np.testing.assert_allclose(comp_out, fs, rtol=1e-13 if prec != "fp32" else 1e-7)
        # field parameters
        fp = rng.uniform(0.5, 2.0, (N, 3)) * np.array([0.153, 0.098, 3.98e-5])
        dev.set_field_parameters(fp)
        fpo = np.zeros((N, 3))
        dev.get_field_parameters(fpo)
        np.testing.assert_allclose(fpo, fp, rtol=0 if prec == "fp64" else 1e-7)
        fpu = np.zeros((3, N))
        dev.get_field_parameters(fpu, tangled=False)
        np.testing.assert_allclose(fpu, fp.T, rtol=0 if prec == "fp64" else 1e-7)
        # reset_default restores initial conditions and default field parameters
        dev.reset_default()
        np.testing.assert_allclose(dev.states(), np.tile(dev.model.initial_conditions, (N, 1)), rtol=1e-7)
        dev.get_field_parameters(fpo)
        np.testing.assert_allclose(fpo, np.tile([0.153, 0.098, 3.98e-5], (N, 1)), rtol=1e-7)


def test_parameters_and_errors():
    from goss_amd._lib import GossHipError

    dev = dev_system("tp06_m", "GRL1", 32, "fp64")
    assert dev.get_parameter("stim_amplitude") == 52.0
    dev.set_parameter("stim_amplitude", 0.0)
    dev.forward_times(np.arange(2001) * 0.01)
    assert dev.states()[:, 17].max() < -80.0  # no stimulus, no upstroke
    with pytest.raises(GossHipError, match='"g_nope" is not a parameter'):
        dev.set_parameter("g_nope", 1.0)
    dev.set_num_threads(4)
    assert dev.get_num_threads() == 4


@pytest.mark.parametrize("prec,max_dv", [("fp32c", 0.05), ("fp32", 0.05)])
def test_fp32_accuracy_against_fp64(prec, max_dv):
    """100 ms of TP06 M-cell GRL1 (one AP) with lognormal conductances:
    stated bound on max |dV| against the FP64 device path."""
    N = 2048
    rng = np.random.Generator(np.random.PCG64(31))
    ref = dev_system("tp06_m_fp3", "GRL1", N, "fp64")
    low = dev_system("tp06_m_fp3", "GRL1", N, prec)
    y0 = perturbed(ref.states(), 32)
    fp = ref_fp = np.tile([0.153, 0.098, 3.98e-5], (N, 1)) * np.exp(rng.normal(0, 0.2, (N, 3)))
    for s in (ref, low):
        s.set_states(y0)
        s.set_field_parameters(fp)
    t = np.arange(5001) * 0.02
    ref.forward_times(t)
    low.forward_times(t)
    a, b = low.states(), ref.states()
    dv = np.abs(a[:, 17] - b[:, 17]).max()
    print(f"{prec}: max |dV| after 100 ms = {dv:.3e} mV")
    assert dv <= max_dv


def test_full_size_node_equality_and_permutation():
    """1,048,576 nodes (the per-GPU share of the headline run): identical
    nodes stay bit-identical (cpp/tests/odesystem.cpp:67-83 checks all nodes
    equal), and permuting nodes permutes results exactly."""
    N = 1 << 20
    dev = dev_system("tp06_m_fp3", "GRL1", N, "fp64")
    t = np.arange(601) * 0.02  # 12 ms: through the stimulus (10-11 ms) and upstroke
    dev.forward_times(t)
    s = dev.states()
    assert np.all(s == s[0])
    assert s[0, 17] > 0.0  # upstroke reached
    rng = np.random.Generator(np.random.PCG64(2))
    y0 = perturbed(np.tile(dev.model.initial_conditions, (N, 1)), 6)
    perm = rng.permutation(N)
    dev.set_states(y0)
    dev.forward_times(t)
    a = dev.states()
    dev.set_states(y0[perm])
    dev.forward_times(t)
    assert np.array_equal(dev.states(), a[perm])


def test_full_size_closed_form_linear_decay():
    """y' = a y + b per node with random (a, b): GRL1/RL1 are exact, so the
    1,048,576-node result equals the closed form to rounding."""
    N = 1 << 20
    rng = np.random.Generator(np.random.PCG64(12))
    a = -rng.uniform(0.05, 4.0, N)
    b = rng.uniform(-2, 2, N)
    y0 = rng.uniform(-1, 1, N)
    for solver in ("GRL1", "RL1"):
        dev = dev_system("linear_decay", solver, N, "fp64")
        dev.set_states(y0[:, None])
        dev.set_field_parameters(np.column_stack([a, b]))
        dev.forward_times(np.linspace(0, 3.0, 301))
        exact = (y0 + b / a) * np.exp(a * 3.0) - b / a
        np.testing.assert_allclose(dev.states()[:, 0], exact, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("prec", ["fp32c", "fp32"])
@pytest.mark.parametrize("solver,variant,dt,t_end,N,max_dv", [
    ("RL2", "tp06_m_fp3", 0.02, 100.0, 512, 0.05),
    ("RK4", "tp06_m", 0.001, 30.0, 64, 0.05),
    ("ExplicitEuler", "tp06_m", 0.001, 30.0, 64, 0.05),
])
def test_fp32_solvers_against_the_fp64_oracle(prec, solver, variant, dt, t_end, N, max_dv):
    """FP32 and compensated FP32 for the explicit schemes besides GRL1, against the FP64 oracle:
    RL2 over 100 ms (one action potential) at dt 0.02 ms, RK4 and ExplicitEuler through the
    stimulus and upstroke at dt 0.001 ms (within their stability limit for TP06's fastest
    gate).  Bound on max |dV| at the end, in mV; every state within 1e-3 (compensated) or
    2e-3 (plain FP32) relative of its largest magnitude.  Measured: |dV| 3.8e-5 / 3.2e-4 mV
    (RL2), 2.7e-3 / 3.9e-3 mV (RK4), 2.7e-3 / 3.8e-3 mV (ExplicitEuler); states 7.5e-4
    (compensated) and 1.01e-3 (plain) for RK4 and ExplicitEuler."""
    o = oracle.OracleSystem(variant, solver, N)
    o.set_num_threads(8)
    dev = dev_system(variant, solver, N, prec)
    y0 = perturbed(o.states(), 61)
    o.set_states(y0)
    dev.set_states(y0)
    t = np.arange(int(round(t_end / dt)) + 1) * dt
    o.forward_times(t)
    dev.forward_times(t)
    a, b = dev.states(), o.states()
    if t_end < 50:
        assert b[:, 17].min() > -20.0  # the 30 ms runs end on the plateau, after the upstroke
    dv = np.abs(a[:, 17] - b[:, 17]).max()
    rel = (np.abs(a - b) / np.max(np.abs(b), axis=0, keepdims=True)).max()
    print(f"{solver} {prec}: max |dV| {dv:.3e} mV, max state error {rel:.3e} of the state's magnitude")
    assert dv <= max_dv
    assert rel <= (1e-3 if prec == "fp32c" else 2e-3)
