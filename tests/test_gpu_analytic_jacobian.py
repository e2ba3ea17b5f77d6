"""The implicit solvers with the model's analytic Jacobian on the device
(solver parameter jacobian = GOSS_HIP_JACOBIAN_ANALYTIC: a plugin generated
with generate_jacobian, python/goss/codegeneration.py:186, 622-641).

* Parity build (host-identical FP64 elementary functions): the device equals
  the host build of the same kernel logic bit for bit, states and counters.
  That host build equals the oracle's analytic mode bit for bit, and the
  oracle's analytic mode equals the reference's solvers running the plugin
  with its compute_jacobian override (tests/test_analytic_jacobian.py).
* Product build: the device against the oracle's analytic mode at the
  one-ulp-calibrated bound of the implicit solvers (tests/test_gpu_implicit.py
  explains it).
* The Python API: ParameterizedODE(..., code_params={"generate_jacobian":
  True}) makes compute_jacobian analytic and the implicit solvers use it.
"""
import ctypes
import json

import numpy as np
import pytest

import oracle
from test_gpu_explicit import assert_parity, dev_system, perturbed
from test_gpu_implicit import state_scale, ulp_sensitivity
from test_gpu_parity_build import DEV_LIB, META, SOLVER_IDS, _libs, _p

pytestmark = pytest.mark.gpu


def _host_parity(variant, solver, y0, t, params):
    L = _libs()
    dp = ctypes.POINTER(ctypes.c_double)
    L.kernel_logic_run_analytic_jacobian.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint, dp, dp,
                                                     ctypes.c_ulong, dp, dp, ctypes.POINTER(ctypes.c_long)]
    out = np.zeros_like(y0)
    counters = (ctypes.c_long * 4)()
    rc = L.kernel_logic_run_analytic_jacobian(variant.encode(), SOLVER_IDS[solver], y0.shape[0], _p(y0), _p(t),
                                              len(t), _p(params), _p(out), counters)
    assert rc == 0, rc
    return out, np.array(list(counters))


@pytest.mark.parametrize("variant,solver,N,dt,steps,stim", [
    ("tp06_epi", "ImplicitEuler", 1024, 0.1, 3000, 52.0),   # config-3 shape through the action potential
    ("tp06_epi", "ESDIRK23a", 1024, 0.1, 3000, 52.0),
    ("tp06_m", "ThetaSolver", 256, 0.1, 400, None),
    ("tp06_m", "BasicImplicitEuler", 256, 0.1, 400, None),
    ("fhn", "ESDIRK4O32", 16, 0.05, 3, None),
])
def test_analytic_device_equals_host_bit_for_bit(variant, solver, N, dt, steps, stim):
    from goss_amd.system import DeviceSystem

    meta = json.load(open(META))[variant]
    params = np.ascontiguousarray(meta["default_parameters"], dtype=np.float64)
    if stim is not None:
        params[meta["parameter_names"].index("stim_amplitude")] = stim
    rng = np.random.Generator(np.random.PCG64(2028))
    ic = np.array(meta["initial_conditions"])
    y0 = np.ascontiguousarray(ic * (1.0 + rng.uniform(-0.01, 0.01, (N, ic.size))))
    t = np.arange(steps + 1) * dt
    print(f"{variant} {solver} N={N} steps={steps}", flush=True)
    host, host_c = _host_parity(variant, solver, y0, t, params)
    s = DeviceSystem(variant, solver, N, "fp64", {"jacobian": "analytic"}, library=DEV_LIB)
    for k, v in enumerate(params):
        if v != s.model.default_parameters[k]:
            s.set_parameter(meta["parameter_names"][k], float(v))
    s.set_steps_per_launch(100)
    s.set_states(y0)
    s.forward_times(t)
    dev, dev_c = s.states(), s.solver_counters()
    s.close()
    assert np.all(np.isfinite(host))
    np.testing.assert_array_equal(dev, host)
    np.testing.assert_array_equal(dev_c[1:], host_c[1:])


@pytest.mark.parametrize("solver", ["ImplicitEuler", "ESDIRK23a", "ThetaSolver", "BasicImplicitEuler"])
@pytest.mark.parametrize("variant,dt,steps", [("tp06_epi", 0.1, 300), ("beeler_reuter_1977", 0.5, 100),
                                              ("tp06_m", 0.05, 300), ("van_der_pol_dae", 0.01, 60)])
def test_analytic_matches_oracle(solver, variant, dt, steps):
    N = 65
    params = {"jacobian": 2}
    o = oracle.OracleSystem(variant, solver, N, params, per_node_state=True)
    dev = dev_system(variant, solver, N, "fp64", {"jacobian": "analytic"})
    if variant == "van_der_pol_dae":
        z = np.random.Generator(np.random.PCG64(31)).uniform(1.95, 2.2, N)
        y0 = np.column_stack([(z * z / 3.0 - 1.0) * z, z])
    else:
        y0 = perturbed(o.states(), 42)
    t = np.arange(steps + 1) * dt
    sens, c_a, c_b = ulp_sensitivity(variant, solver, params, y0, t)
    o.set_states(y0)
    dev.set_states(y0)
    o.forward_times(t)
    dev.forward_times(t)
    same = list(c_a[1:]) == list(c_b[1:])
    tol = max(1e-10, 10 * sens)
    if same:
        tol = min(tol, 1e-8)
    err = assert_parity(dev.states(), o.states(), rtol=tol, what=f"{variant}/{solver} analytic",
                        scale=state_scale(y0, o.states()))
    oc, dc = o.counters(), dev.solver_counters()
    print(f"{variant}/{solver} analytic: max rel err {err:.3e} (one-ulp spread {sens:.3e}, bound {tol:.1e}), "
          f"counters oracle {oc} device {dc}")
    if same:
        assert list(oc[1:]) == list(dc[1:])


def test_python_api_generate_jacobian():
    import goss_amd as goss

    ode = goss.ParameterizedODE("tp06_m", code_params={"generate_jacobian": True})
    assert ode.has_analytic_jacobian and not goss.ParameterizedODE("tp06_m").has_analytic_jacobian
    y = ode.get_ic() * 1.001
    J = ode.compute_jacobian(y, 0.0)
    ref = oracle.Variant("tp06_m").analytic_jacobian(y, 0.0)
    assert np.abs(J - ref).max() <= 1e-12 * np.abs(ref).max()

    N, dt, steps = 64, 0.05, 200
    solver = goss.solvers.ImplicitEuler()
    system = goss.ODESystemSolver(N, solver, ode)
    y0 = perturbed(np.tile(ode.get_ic(), (N, 1)), 43)
    system.states = y0
    t = np.arange(steps + 1) * dt
    for k in range(steps):
        system.forward(t[k], dt)
    o = oracle.OracleSystem("tp06_m", "ImplicitEuler", N, {"jacobian": 2}, per_node_state=True)
    o.set_states(y0)
    o.forward_times(t)
    err = assert_parity(system.states, o.states(), rtol=1e-6, what="pygoss analytic",
                        scale=state_scale(y0, o.states()))
    print(f"pygoss ImplicitEuler with the analytic Jacobian: max rel err {err:.2e}")
