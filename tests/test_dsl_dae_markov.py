"""The .ode language's DAE and Markov-chain forms: frontend, generated code and
the oracle's solvers on them.

* DAE (``alg_<X>_0 = expr``, ``codegeneration.py:452-459``): the van der Pol
  oscillator of the reference's DAE test model (cpp/tests/van_der_Pol.h) as a
  .ode file of this project (goss_amd/odes/van_der_pol_dae.ode).  Its generated
  right-hand side and analytic Jacobian match the reference's hand-written
  class; the reference's solvers running the generated plugin equal the
  oracle bit for bit; and the mathematics pins the result: on the constraint
  y = (z^2/3 - 1) z the DAE reduces to z' = -z / (z^2 - 1), which backward
  Euler and the theta method approach at orders 1 and 2.
* Markov chains (``rates[A, B] = expr``): a three-state channel whose opening
  rate is modulated in time.  The chain is the master equation
  x' = Q(t)^T x: with constant rates its solution is a matrix exponential, and
  otherwise a tightly integrated reference solution; RK4 reaches both, the
  occupancies stay summed to 1, and the reference's solvers running the
  generated plugin equal the oracle bit for bit.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.linalg import expm

import oracle
from goss_amd.codegen.ode_model import ODEParseError, parse_ode_text

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build oracle/_ref not present")
needs_refpin = pytest.mark.skipif(not oracle.refpin_available(),
                                  reason="reference build oracle/_ref/libgoss_refpin.so not present")


def test_frontend_dae_and_markov_forms():
    m = parse_ode_text("states(y=1.0, z=2.0)\ndy_dt = -z\nalg_z_0 = y - z\n", "d")
    assert m.is_dae and m.differential_states == [1, 0] and m.derivatives == {"y": "dy_dt", "z": "alg_z_0"}
    m = parse_ode_text("parameters(a=2.0, b=0.5)\nstates(C=1.0, O=0.0)\nrates[C, O] = a\nrates[O, C] = b\n", "m")
    assert not m.is_dae and m.state_names == ["C", "O"]
    bd = m.body_dict()
    C, O, a, b = (m.symbols[k] for k in ("C", "O", "rates_C_O", "rates_O_C"))
    assert (bd["dC_dt"] - (b * O - a * C)).simplify() == 0
    assert (bd["dO_dt"] - (a * C - b * O)).simplify() == 0
    # the derivatives of chain states come after the explicitly defined ones
    m = parse_ode_text("parameters(a=1.0)\nstates(C=1.0, O=0.0, V=0.0)\nrates[C, O] = a\ndV_dt = -V\n", "m2")
    assert m.state_names == ["V", "C", "O"]
    for text in ("states(C=1.0)\nparameters(a=1.0)\nrates[C, X] = a\n",
                 "states(C=1.0, O=0.0)\nparameters(a=1.0)\nrates[C, O] = a\ndC_dt = -C\n",
                 "states(C=1.0, O=0.0)\nparameters(a=1.0)\nrates[C, O] = a\nrates[C, O] = a\n",
                 "states(C=1.0)\nparameters(a=1.0)\nrates[C, C] = a\n",
                 "states(y=1.0)\ndy_dt = -y\nalg_y_0 = y\n"):
        with pytest.raises(ODEParseError):
            parse_ode_text(text, "bad")


def test_dae_metadata_through_the_python_api():
    """ODE.is_dae (python/goss/ode.py:74-77) and the mass matrix's diagonal, from the library's metadata."""
    import goss_amd as goss

    ode = goss.ParameterizedODE("van_der_pol_dae")
    assert ode.is_dae and ode.differential_states.tolist() == [1, 0]
    assert not goss.ParameterizedODE("markov_three_state").is_dae


def test_generated_metadata():
    v = oracle.Variant("van_der_pol_dae")
    assert v.differential_states == [1, 0]
    assert v.initial_conditions.tolist() == [(2.0 * 2.0 / 3.0 - 1.0) * 2.0, 2.0]
    m = oracle.Variant("markov_three_state")
    assert m.differential_states == [1, 1, 1] and m.linear_terms == [1, 1, 1]


@needs_refpin
def test_van_der_pol_rhs_and_jacobian_match_the_reference_class():
    """Generated eval and analytic Jacobian vs cpp/tests/van_der_Pol.h (eval, compute_jacobian).  The
    generator writes z*z/3 as z**2 * (1/3), so the two agree to rounding, not bit for bit."""
    oracle.register_reference_models()
    ours, ref = oracle.Variant("van_der_pol_dae"), oracle.Variant("ref_van_der_pol")
    rng = np.random.Generator(np.random.PCG64(5))
    for _ in range(200):
        y = rng.uniform(-3, 3, 2)
        a, b = ours.eval(y, 0.0), ref.eval(y, 0.0)
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-14 * np.abs(y).max() ** 3)
        np.testing.assert_allclose(ours.analytic_jacobian(y, 0.0).ravel(),
                                   np.array([0.0, -1.0, 1.0, 1.0 - y[1] ** 2]), rtol=1e-14, atol=1e-14)


@needs_ref
@pytest.mark.parametrize("threads", [0, 3])
@pytest.mark.parametrize("variant,solvers,dt,steps", [
    ("van_der_pol_dae", ["ThetaSolver", "ImplicitEuler", "BasicImplicitEuler", "ESDIRK23a"], 0.01, 60),
    ("markov_three_state", ["ExplicitEuler", "RK2", "RK4", "RL1", "RL2", "GRL1", "GRL2", "ImplicitEuler",
                            "ESDIRK23a", "RKF32", "ThetaSolver", "BasicImplicitEuler"], 0.05, 40),
])
def test_generated_plugins_oracle_equals_reference(variant, solvers, dt, steps, threads):
    """The reference's solvers (cpp/goss/*.cpp) with the generated plugin, against the oracle with the same
    model: bit for bit, node after node (the reference carries solver state from node to node)."""
    v = oracle.Variant(variant)
    N = 6
    rng = np.random.Generator(np.random.PCG64(8))
    y0 = np.tile(v.initial_conditions, (N, 1))
    if variant == "markov_three_state":
        y0 = rng.dirichlet(np.ones(3), N)
    else:
        # Consistent initial values (on the constraint).  From z0 the solution reaches the fold at z = 1,
        # where the DAE has no continuation, at t = z0^2/2 - ln z0 - 1/2 (0.73 for z0 = 1.95): the run
        # (0.6) stays before it.
        z = rng.uniform(1.95, 2.2, N)
        y0 = np.column_stack([(z * z / 3.0 - 1.0) * z, z])
    t = np.arange(steps + 1) * dt
    for solver in solvers:
        if solver == "RKF32" and threads:
            continue  # the reference's thread copies of RKF32 start from an uninitialised `first`
        o = oracle.OracleSystem(variant, solver, N)
        r = oracle.RefSystem(variant, solver, N)
        for s in (o, r):
            s.set_num_threads(threads)
            s.set_states(y0)
            if v.num_field_parameters:
                s.set_field_parameters(np.full((N, v.num_field_parameters), 0.3) * rng.uniform(0.5, 1.5, (N, 1)))
        o.set_field_parameters(r.field_parameters()) if v.num_field_parameters else None
        o.forward_times(t)
        r.forward_times(t)
        assert np.array_equal(o.states(), r.states()), (variant, solver)


def _vdp_exact(z0, t_end):
    sol = solve_ivp(lambda t, z: -z / (z * z - 1.0), (0.0, t_end), [z0], rtol=1e-13, atol=1e-15, method="DOP853")
    z = sol.y[0, -1]
    return np.array([(z * z / 3.0 - 1.0) * z, z])


@pytest.mark.parametrize("solver,order", [("ImplicitEuler", 1), ("ThetaSolver", 2), ("BasicImplicitEuler", 1)])
def test_dae_solvers_converge_to_the_reduced_ode(solver, order):
    """On the constraint the DAE is z' = -z / (z^2 - 1) (differentiate y = (z^2/3 - 1) z and use y' = -z).
    The errors at t = 0.5 fall by 10^order for a tenfold smaller step, and the states stay on the
    constraint."""
    t_end, z0 = 0.5, 2.0
    exact = _vdp_exact(z0, t_end)
    errs = []
    for dt in (1e-2, 1e-3):
        n = int(round(t_end / dt))
        o = oracle.OracleSystem("van_der_pol_dae", solver, 1,
                                {"theta": 0.5} if solver == "ThetaSolver" else None)
        y0 = oracle.Variant("van_der_pol_dae").initial_conditions
        o.set_states(y0[None, :])
        o.forward_times(np.arange(n + 1) * dt)
        y = o.states()[0]
        assert abs(y[0] - (y[1] ** 2 / 3 - 1) * y[1]) < 1e-10
        errs.append(abs(y[1] - exact[1]))
    ratio = errs[0] / errs[1]
    print(f"{solver}: errors {errs[0]:.3e}, {errs[1]:.3e}, ratio {ratio:.1f}")
    assert 10 ** order / 2 < ratio < 10 ** order * 2


def _markov_params():
    import json
    import os
    meta = json.load(open(os.path.join(os.path.dirname(oracle.__file__), "..", "goss_amd", "csrc", "generated",
                                       "variants.json")))["markov_three_state"]
    return meta["parameter_names"], dict(zip(meta["parameter_names"], meta["default_parameters"]))


def _generator(p, t):
    """The chain's generator matrix: Q[a, b] the rate from a to b, rows summing to zero."""
    q = np.zeros((3, 3))
    q[0, 1] = p["k_co"] * (1.0 + p["a_co"] * np.sin(p["w_co"] * t))
    q[1, 0], q[1, 2], q[2, 1] = p["k_oc"], p["k_oi"], p["k_io"]
    return q - np.diag(q.sum(axis=1))


@pytest.mark.parametrize("modulated", [False, True])
def test_markov_chain_is_the_master_equation(modulated):
    """x' = Q(t)^T x (the master equation of a continuous-time Markov chain).  With constant rates
    (a_co = 0) its solution is expm(Q^T t) x0; with the opening rate modulated in time it is integrated
    with DOP853 at rtol 1e-13.  RK4 at dt 1e-3 to t = 2 is within 1e-11 of it, and C + O + I stays 1 to
    rounding."""
    names, p = _markov_params()
    if not modulated:
        p["a_co"] = 0.0
    x0 = np.array([0.7, 0.2, 0.1])
    dt, t_end = 1e-3, 2.0
    o = oracle.OracleSystem("markov_three_state", "RK4", 1)
    o.set_parameter("a_co", p["a_co"])
    o.set_states(x0[None, :])
    o.set_field_parameters(np.array([[p["k_oi"]]]))
    o.forward_times(np.arange(int(round(t_end / dt)) + 1) * dt)
    x = o.states()[0]
    if modulated:
        sol = solve_ivp(lambda t, y: _generator(p, t).T @ y, (0.0, t_end), x0, rtol=1e-13, atol=1e-15,
                        method="DOP853")
        exact = sol.y[:, -1]
    else:
        exact = expm(_generator(p, 0.0).T * t_end) @ x0
    print("RK4 vs exact:", np.abs(x - exact).max())
    assert np.abs(x - exact).max() < 1e-11
    assert abs(x.sum() - 1.0) < 1e-14
