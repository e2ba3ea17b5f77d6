"""The implicit solvers with the model's analytic Jacobian (a plugin generated
with generate_jacobian, python/goss/codegeneration.py:186, 622-641), on the CPU.

* The generated analytic Jacobian equals central differences of the generated
  right-hand side (mathematics), and for van der Pol the reference's
  hand-written class (tests/test_dsl_dae_markov.py); its structured form (the
  entries the implicit kernels store) holds exactly the dense form's entries.
* The oracle's solvers in analytic mode equal the reference's solvers
  (cpp/goss/*.cpp compiled from their sources) running the generated plugin
  with its compute_jacobian override, bit for bit: ImplicitEuler, ESDIRK23a,
  ThetaSolver, BasicImplicitEuler and ESDIRK4O32.  The oracle's Jacobian
  dispatch is the one that also runs the reference's own analytic-Jacobian
  class, van_der_Pol.h, bit for bit equal to the reference (test_oracle_pinned.py).
* The host build of the kernels' per-node logic (goss_implicit.h with
  compute_jacobian_structured) equals the oracle's per-node mode bit for bit.
* The analytic and the forward-difference Jacobian solve the same equations:
  the states after a run agree to the Newton tolerance.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle
from test_kernel_logic_cpu import HARNESS, META, _p

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build oracle/_ref not present")

SOLVER_IDS = {"ImplicitEuler": 7, "ESDIRK23a": 8, "ThetaSolver": 10, "BasicImplicitEuler": 11, "ESDIRK4O32": 12}
IMPLICIT = ["ImplicitEuler", "ESDIRK23a", "ThetaSolver", "BasicImplicitEuler"]


def _y0(variant, N, seed, spread=0.01):
    base = oracle.Variant(variant).initial_conditions
    rng = np.random.Generator(np.random.PCG64(seed))
    return base[None, :] * (1.0 + rng.uniform(-spread, spread, (N, base.size)))


@pytest.mark.parametrize("variant", ["tp06_m", "beeler_reuter_1977", "fhn", "van_der_pol_dae", "markov_three_state"])
def test_analytic_jacobian_against_central_differences(variant):
    v = oracle.Variant(variant)
    rng = np.random.Generator(np.random.PCG64(3))
    S = v.num_states
    for _ in range(5):
        y = v.initial_conditions * (1.0 + rng.uniform(-0.02, 0.02, S))
        if variant == "van_der_pol_dae":
            y = rng.uniform(1.5, 2.5, 2)
        J = v.analytic_jacobian(y, 0.3)
        C = np.zeros((S, S))
        for i in range(S):
            h = 1e-6 * max(1e-3, abs(y[i]))
            yp, ym = y.copy(), y.copy()
            yp[i] += h
            ym[i] -= h
            C[:, i] = (v.eval(yp, 0.3) - v.eval(ym, 0.3)) / (2 * h)
        scale = np.maximum(np.abs(C).max(axis=1, keepdims=True), 1e-12)
        assert (np.abs(J - C) / scale).max() < 1e-5, variant


@needs_ref
@pytest.mark.parametrize("solver", IMPLICIT + ["ESDIRK4O32"])
@pytest.mark.parametrize("variant,dt,steps", [("tp06_m", 0.05, 200), ("beeler_reuter_1977", 0.1, 100),
                                              ("fhn", 0.5, 40), ("van_der_pol_dae", 0.01, 60)])
def test_oracle_analytic_equals_reference_analytic(variant, dt, steps, solver):
    if solver == "ESDIRK4O32":
        if variant in ("tp06_m", "van_der_pol_dae", "beeler_reuter_1977"):
            pytest.skip("ESDIRK4O32 takes 1e4-1e5 steps per interval (and refuses a DAE): fhn only")
        dt, steps = 0.05, 3
    N = 4
    y0 = _y0(variant, N, 5)
    if variant == "van_der_pol_dae":
        z = np.random.Generator(np.random.PCG64(6)).uniform(1.95, 2.2, N)
        y0 = np.column_stack([(z * z / 3.0 - 1.0) * z, z])
    t = np.arange(steps + 1) * dt
    o = oracle.OracleSystem(variant, solver, N, {"jacobian": 2})
    r = oracle.RefSystem(variant, solver, N, analytic_jacobian=True)
    fd = oracle.OracleSystem(variant, solver, N)
    for s in (o, r, fd):
        s.set_states(y0)
        s.forward_times(t)
    assert np.array_equal(o.states(), r.states()), (variant, solver)
    # the analytic Jacobian is the one that ran: on TP06 the forward-difference run differs in the last bits
    # (on the small models Newton can land on the same bits with either Jacobian)
    if variant == "tp06_m":
        assert not np.array_equal(o.states(), fd.states())
    if solver in ("ImplicitEuler", "ThetaSolver", "BasicImplicitEuler") and variant != "tp06_m":
        np.testing.assert_allclose(o.states(), fd.states(), rtol=1e-8, atol=1e-10)


def _harness_run(variant, solver, y0, t):
    if not os.path.exists(HARNESS):
        pytest.skip("host build of the kernel logic is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(HARNESS)
    dp = ctypes.POINTER(ctypes.c_double)
    L.kernel_logic_run_analytic_jacobian.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint, dp, dp,
                                                     ctypes.c_ulong, dp, dp, ctypes.POINTER(ctypes.c_long)]
    meta = json.load(open(META))[variant]
    params = np.ascontiguousarray(meta["default_parameters"], dtype=np.float64)
    y0 = np.ascontiguousarray(y0)
    out = np.zeros_like(y0)
    counters = (ctypes.c_long * 4)()
    rc = L.kernel_logic_run_analytic_jacobian(variant.encode(), SOLVER_IDS[solver], y0.shape[0], _p(y0), _p(t),
                                              len(t), _p(params), _p(out), counters)
    assert rc in (0, 1)
    return out, np.array(list(counters))


@pytest.mark.parametrize("solver", IMPLICIT)
@pytest.mark.parametrize("variant,dt,steps", [("tp06_m", 0.05, 300), ("tp06_epi", 0.1, 120),
                                              ("beeler_reuter_1977", 0.1, 100), ("fhn", 0.5, 60)])
def test_kernel_logic_analytic_matches_oracle_bit_exact(variant, dt, steps, solver):
    N = 9
    y0 = _y0(variant, N, 41)
    t = np.arange(steps + 1) * dt
    o = oracle.OracleSystem(variant, solver, N, {"jacobian": 2}, per_node_state=True)
    o.set_states(y0)
    for k in range(len(t) - 1):
        o.forward(t[k], t[k + 1] - t[k])
    got, got_c = _harness_run(variant, solver, y0, t)
    np.testing.assert_array_equal(got, o.states())
    np.testing.assert_array_equal(got_c[1:], o.counters()[1:])


def test_oracle_refuses_unknown_jacobian_mode():
    with pytest.raises(RuntimeError, match="jacobian must be 0"):
        oracle.OracleSystem("fhn", "ImplicitEuler", 2, {"jacobian": 1})
