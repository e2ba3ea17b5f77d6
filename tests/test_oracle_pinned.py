"""The oracle's solvers pinned against the reference with no code of ours in
the model, and against what the mathematics fixes.

* The reference's own gotran-generated test models (cpp/tests/fhn.h,
  beeler_reuter_1977.h, Winslow.h, Panfilov.h) are compiled with the
  reference's solvers (cpp/goss/*.cpp) into oracle/_ref/libgoss_refpin.so.
  The same model classes are registered with the oracle, whose restated
  solvers then run on them: states after every solver must equal the
  reference's bit for bit, serially and with OpenMP threads, and the oracle's
  per-node solver state must equal one reference system per node.  The FD
  Jacobian (ODE.cpp:47-67) must equal ODE::compute_jacobian bit for bit.
* The reference's DAE test model (cpp/tests/van_der_Pol.h: one algebraic
  state, its own analytic compute_jacobian) advanced by the reference's
  implicit solvers through ODESolver::forward equals the oracle's restatement
  on the same class bit for bit; the explicit solvers refuse it with the
  reference's messages.
* Closed forms on y' = a*y + b: one backward-Euler step is
  (y0 + h*b) / (1 - h*a); ESDIRK23a's single-step amplification R(z) has a
  local error of order z^4 (a third-order method) and tends to 0 as z -> -inf
  (L-stability); the FD Jacobian of the linear oscillator is its matrix to
  forward-difference rounding.
"""
import numpy as np
import pytest

import oracle

EXPLICIT = ["ExplicitEuler", "RK2", "RK4", "RL1", "RL2", "GRL1", "GRL2"]
ALL = EXPLICIT + ["ImplicitEuler", "ESDIRK23a", "RKF32", "ThetaSolver", "BasicImplicitEuler"]

needs_refpin = pytest.mark.skipif(not oracle.refpin_available(),
                                  reason="reference build oracle/_ref/libgoss_refpin.so not present")

# (model, solvers, dt, steps).  Winslow and Panfilov have no linearized_eval:
# the Rush-Larsen schemes do not apply to them.
CASES = [
    ("ref_fhn", ALL, 0.1, 400),
    ("ref_beeler_reuter_1977", ALL, 0.05, 400),
    ("ref_winslow", ["ExplicitEuler", "RK2", "RK4", "ImplicitEuler", "ESDIRK23a", "ThetaSolver",
                     "BasicImplicitEuler"], 0.01, 200),
    ("ref_winslow", ["RKF32"], 0.01, 20),
    ("ref_panfilov", ["ExplicitEuler", "RK2", "RK4", "ImplicitEuler", "ESDIRK23a", "RKF32", "ThetaSolver",
                      "BasicImplicitEuler"], 0.1, 200),
    # ESDIRK4O32 as the reference computes it takes 1e4-1e6 steps per interval (its stage-2
    # evaluation at the increment and its two-buffer update, ESDIRK4O32.cpp:163-297): short runs.
    ("ref_fhn", ["ESDIRK4O32"], 0.05, 3),
    ("ref_panfilov", ["ESDIRK4O32"], 0.1, 4),
]


def _inputs(v, N, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    y0 = v.initial_conditions * (1 + rng.uniform(-0.01, 0.01, (N, v.num_states)))
    fp = None
    if v.num_field_parameters:
        g = v.default_parameters[v.field_parameter_indices]
        fp = g * np.exp(rng.normal(0.0, 0.1, (N, v.num_field_parameters)))
    return y0, fp


def _params(solver, dt):
    # Fixed-step solvers with sub-steps, so the internal time step is exercised too.
    return {"internal_time_step": dt / 4} if solver in EXPLICIT and solver != "GRL2" else {}


@needs_refpin
@pytest.mark.parametrize("threads", [0, 3])
@pytest.mark.parametrize("model,solvers,dt,steps", CASES)
def test_oracle_solvers_equal_reference_on_reference_models(model, solvers, dt, steps, threads):
    oracle.register_reference_models()
    v = oracle.Variant(model)
    N = 7
    y0, fp = _inputs(v, N, 5)
    t = np.arange(steps + 1) * dt
    for solver in solvers:
        if threads and solver == "RKF32":
            # The threaded reference runs copies of the solver (ODESystemSolver.cpp:236-241), and RKF32's copy
            # constructor (RKF32.cpp:52-61) leaves `first` uninitialised: a copy may skip the k1 evaluation
            # (RKF32.cpp:127-131) and start a node from the previous node's derivative.  That read is undefined;
            # the oracle keeps the serial semantics (k1 evaluated on every call), pinned with threads = 0.
            continue
        o = oracle.OracleSystem(model, solver, N, _params(solver, dt))
        r = oracle.RefSystem(model, solver, N, _params(solver, dt))
        for s in (o, r):
            if threads:
                s.set_num_threads(threads)
            s.set_states(y0)
            if fp is not None:
                s.set_field_parameters(fp)
            s.forward_times(t)
        a, b = o.states(), r.states()
        assert np.all(np.isfinite(b)), (model, solver)
        assert np.array_equal(a, b), (model, solver, np.abs(a - b).max())
        if fp is not None:
            assert np.array_equal(o.field_states(), r.field_states())


@needs_refpin
@pytest.mark.parametrize("theta", [1.0, 0.7])
@pytest.mark.parametrize("model", ["ref_fhn", "ref_beeler_reuter_1977", "ref_winslow"])
def test_theta_solver_tunable_theta(model, theta):
    """ThetaSolver with theta != 0.5, serially.  (The threaded reference runs
    copies of the solver, and ThetaSolver's copy constructor (ThetaSolver.cpp:
    27-31) does not copy the public theta, min_dt and refinement count, nor
    ImplicitODESolver's tunables: the copies run with the defaults.)"""
    oracle.register_reference_models()
    v = oracle.Variant(model)
    y0, fp = _inputs(v, 6, 3)
    t = np.arange(61) * 0.1
    o = oracle.OracleSystem(model, "ThetaSolver", 6, {"theta": theta})
    r = oracle.RefSystem(model, "ThetaSolver", 6, {"theta": theta})
    for s in (o, r):
        s.set_states(y0)
        if fp is not None:
            s.set_field_parameters(fp)
        s.forward_times(t)
    assert np.array_equal(o.states(), r.states())


@needs_refpin
@pytest.mark.parametrize("model,solver,dt,steps", [
    ("ref_fhn", "ThetaSolver", 0.5, 60), ("ref_winslow", "ThetaSolver", 0.05, 40),
    ("ref_fhn", "ImplicitEuler", 0.5, 60), ("ref_fhn", "ESDIRK23a", 0.5, 60), ("ref_fhn", "RKF32", 0.5, 60),
    ("ref_beeler_reuter_1977", "ImplicitEuler", 0.2, 60), ("ref_beeler_reuter_1977", "ESDIRK23a", 0.2, 60),
    ("ref_winslow", "ImplicitEuler", 0.05, 40), ("ref_winslow", "ESDIRK23a", 0.05, 40),
])
def test_oracle_per_node_state_equals_one_reference_system_per_node(model, solver, dt, steps):
    """The device keeps solver state per node; the oracle's per-node mode is
    that semantics, and equals a separate reference system for each node."""
    oracle.register_reference_models()
    v = oracle.Variant(model)
    N = 5
    y0, fp = _inputs(v, N, 9)
    t = np.arange(steps + 1) * dt
    o = oracle.OracleSystem(model, solver, N, {}, per_node_state=True)
    o.set_states(y0)
    if fp is not None:
        o.set_field_parameters(fp)
    o.forward_times(t)
    got = o.states()
    for n in range(N):
        r = oracle.RefSystem(model, solver, 1)
        r.set_states(y0[n:n + 1])
        if fp is not None:
            r.set_field_parameters(fp[n:n + 1])
        r.forward_times(t)
        assert np.array_equal(got[n], r.states()[0]), (model, solver, n)


@needs_refpin
@pytest.mark.parametrize("model", [m for m in oracle.REFERENCE_MODELS if m != "ref_van_der_pol"])
def test_fd_jacobian_equals_reference_compute_jacobian(model):
    oracle.register_reference_models()
    v = oracle.Variant(model)
    m = oracle.RefModel(model)
    rng = np.random.Generator(np.random.PCG64(2))
    for k in range(5):
        y = v.initial_conditions * (1 + rng.uniform(-0.05, 0.05, v.num_states))
        t = float(rng.uniform(0, 20))
        assert np.array_equal(v.eval(y, t), m.eval(y, t))
        assert np.array_equal(v.fd_jacobian(y, t), m.compute_jacobian(y, t))


def test_backward_euler_step_is_the_closed_form():
    # y' = a*y + b: one ImplicitEuler step of size h solves y1 = y0 + h*(a*y1 + b).
    a = np.array([-2.0, -50.0, -0.3, 1.5, -1e3])
    b = np.array([3.0, 1.0, -0.7, 0.2, 5.0])
    h = 0.01
    N = len(a)
    o = oracle.OracleSystem("linear_decay", "ImplicitEuler", N, {}, per_node_state=True)
    y0 = np.array([1.0, -0.5, 2.0, 0.3, 4.0])
    o.set_states(y0[:, None])
    o.set_field_parameters(np.stack([a, b], 1))
    o.forward_times(np.array([0.0, h]))
    expect = (y0 + h * b) / (1 - h * a)
    np.testing.assert_allclose(o.states()[:, 0], expect, rtol=1e-12, atol=0)
    # One step each: no refinement, no rejection.
    assert o.counters()[2] == 0


def _esdirk_single_step(z, h=1e-3, y0=1e-6):
    """Amplification y1/y0 of one ESDIRK23a step of size h on y' = (z/h) y.
    The first step of ESDIRK23a is at least min_dt = 1e-3 (ESDIRK23a.h:38), so
    an interval of h = 1e-3 is one step; y0 is small enough that the step's
    absolute error estimate is below tol = 1e-5 and it is accepted."""
    z = np.asarray(z, dtype=float)
    N = z.size
    o = oracle.OracleSystem("linear_decay", "ESDIRK23a", N, {}, per_node_state=True)
    o.set_states(np.full((N, 1), y0))
    o.set_field_parameters(np.stack([z / h, np.zeros(N)], 1))
    o.forward_times(np.array([0.0, h]))
    c = o.counters()
    assert c[2] == 0 and c[3] == N, c  # no rejection, one accepted step per node
    return o.states()[:, 0] / y0


def test_esdirk23a_is_third_order():
    z = -np.array([0.0125, 0.025, 0.05, 0.1, 0.2])
    err = _esdirk_single_step(z) - np.exp(z)
    # Local error C z^4 of a third-order method: halving z divides it by 16.
    ratios = err[:-1] / err[1:]
    np.testing.assert_allclose(ratios, 1 / 16, rtol=0.2)
    assert np.all(np.abs(err) < 1e-3)


def test_esdirk23a_is_l_stable():
    z = np.array([-1e2, -1e3, -1e4, -1e5])
    R = _esdirk_single_step(z)
    # R(z) -> 0 like c/z (stiffly accurate, A-stable with R(inf) = 0).
    assert np.all(np.abs(R) < 0.05)
    zr = z * R
    np.testing.assert_allclose(zr, zr[-1], rtol=0.1)


def test_fd_jacobian_of_linear_model_is_its_matrix():
    v = oracle.Variant("oscilator")
    a, b = 1.7, -0.6
    p = v.default_parameters.copy()
    p[0], p[1] = a, b
    for y in ([1.0, 0.0], [0.3, -2.0], [1e-7, 5.0]):
        J = v.fd_jacobian(np.array(y), 0.0, p)
        exact = np.array([[0.0, -a], [b, 0.0]])
        # delta_i = sqrt(1e-15 max(1e-5, |y_i|)); the rounding error of the
        # difference quotient is about eps * |f| / delta.
        delta = np.sqrt(1e-15 * np.maximum(1e-5, np.abs(y)))
        f = np.abs(v.eval(np.array(y), 0.0, p)).max() + 1.0
        assert np.all(np.abs(J - exact) <= 4 * 2.2e-16 * f / delta.min())


@needs_refpin
@pytest.mark.parametrize("model,dt,steps", [("ref_fhn", 0.05, 3), ("ref_panfilov", 0.1, 4)])
def test_esdirk4o32_counters_equal_reference(model, dt, steps):
    """ESDIRK4O32's accepted and rejected steps and right-hand-side evaluations
    (AdaptiveImplicitSolver::num_accepted / num_rejected, ESDIRK4O32::nfevals)
    equal the reference's, serially."""
    oracle.register_reference_models()
    v = oracle.Variant(model)
    y0, _ = _inputs(v, 3, 8)
    t = np.arange(steps + 1) * dt
    o = oracle.OracleSystem(model, "ESDIRK4O32", 3, {"atol": 1e-6, "rtol": 1e-7})
    r = oracle.RefSystem(model, "ESDIRK4O32", 3, {"atol": 1e-6, "rtol": 1e-7})
    for s in (o, r):
        s.set_states(y0)
        s.forward_times(t)
    assert np.array_equal(o.states(), r.states())
    assert list(o.esdirk4o32_counters()) == list(r.esdirk4o32_counters())
    assert o.esdirk4o32_counters()[0] > 0


@needs_refpin
def test_basic_implicit_euler_internal_time_step_equals_reference():
    """BasicImplicitEuler with sub-steps: the reference's residual and clock use the whole
    interval while the Jacobian uses the sub-step (BasicImplicitEuler.cpp:123,168); the
    oracle follows it bit for bit."""
    oracle.register_reference_models()
    for model, dt in (("ref_panfilov", 0.1), ("ref_beeler_reuter_1977", 0.02)):
        v = oracle.Variant(model)
        y0, _ = _inputs(v, 4, 9)
        t = np.arange(41) * dt
        params = {"internal_time_step": dt / 2}
        o = oracle.OracleSystem(model, "BasicImplicitEuler", 4, params)
        r = oracle.RefSystem(model, "BasicImplicitEuler", 4, params)
        for s in (o, r):
            s.set_states(y0)
            s.forward_times(t)
        assert np.all(np.isfinite(r.states()))
        assert np.array_equal(o.states(), r.states()), model


@needs_refpin
@pytest.mark.parametrize("solver,dt,steps", [("ThetaSolver", 0.01, 80), ("ThetaSolver", 0.001, 800),
                                             ("ImplicitEuler", 0.01, 80), ("BasicImplicitEuler", 0.01, 80),
                                             ("ESDIRK23a", 0.01, 20)])
def test_dae_van_der_pol_oracle_equals_reference(solver, dt, steps):
    """The reference's DAE test (cpp/tests/ode.cpp:171-182 runs ThetaSolver to t = 0.8 with dt 0.01 and
    0.001): the mass matrix enters the Newton residual and the iteration matrix
    (ImplicitODESolver.cpp:140,258-262; BasicImplicitEuler.cpp:53,123; ESDIRK23a.cpp:112).  ESDIRK23a's
    stage equations leave the mass matrix out (ESDIRK23a.cpp:223-224), and on this model the reference
    rejects every step and stays at the initial state; the oracle does the same."""
    oracle.register_reference_models()
    v = oracle.Variant("ref_van_der_pol")
    assert v.differential_states == [1, 0]
    y_ref = oracle.refpin_ode_run(solver, v.initial_conditions, 0.0, dt, steps)
    o = oracle.OracleSystem("ref_van_der_pol", solver, 1)
    o.set_states(v.initial_conditions[None, :])
    t = 0.0
    for _ in range(steps):
        o.forward(t, dt)
        t += dt
    y = o.states()[0]
    assert np.array_equal(y, y_ref), (y, y_ref)
    if solver != "ESDIRK23a":
        assert y[1] < 1.2  # the run has moved along the slow branch of the limit cycle (z from 2 down)
        assert abs(y[0] - (y[1] ** 2 / 3 - 1) * y[1]) < 1e-12  # on the constraint
    else:
        assert np.array_equal(y, v.initial_conditions)


@needs_refpin
@pytest.mark.parametrize("solver,message", [
    ("ExplicitEuler", "an explicit solver"), ("RK2", "an explicit solver"), ("RK4", "an explicit solver"),
    ("RL1", "Rush Larsen method"), ("RL2", "an explicit solver"), ("GRL1", "an explicit solver"),
    ("GRL2", "an explicit solver"), ("RKF32", "an explicit solver"), ("ESDIRK4O32", "an explicit solver")])
def test_dae_refused_by_explicit_solvers_with_reference_messages(solver, message):
    oracle.register_reference_models()
    with pytest.raises(RuntimeError, match=f"cannot integrate a DAE ode with {message}") as ref:
        oracle.refpin_ode_run(solver, [0.5, 2.0], 0.0, 0.01, 1)
    with pytest.raises(RuntimeError) as orc:
        oracle.OracleSystem("ref_van_der_pol", solver, 1)
    assert str(orc.value) == str(ref.value)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build oracle/_ref not present")
def test_rkf32_initial_step_from_dtinit_equals_reference():
    """RKF32 calls that start from AdaptiveExplicitSolver::dtinit (AdaptiveExplicitSolver.cpp:67-166): after
    an overshoot of the interval end, an accepted backward step leaves a negative step guess
    (tests/markov_cases.py).  The oracle's per-node state equals one reference system per node, bit for
    bit, on the generated Markov chain (the reference's solvers with the generated plugin)."""
    from markov_cases import DT, STEPS, finishing_nodes, markov_inputs

    y0, fp = markov_inputs(64, 32)
    t = np.arange(STEPS + 1) * DT
    good = finishing_nodes(y0, fp, t)[:12]
    assert len(good) >= 6
    o = oracle.OracleSystem("markov_three_state", "RKF32", len(good), per_node_state=True)
    o.set_states(y0[good])
    o.set_field_parameters(fp[good])
    o.forward_times(t)
    for k, n in enumerate(good):
        r = oracle.RefSystem("markov_three_state", "RKF32", 1)
        r.set_states(y0[n:n + 1])
        r.set_field_parameters(fp[n:n + 1])
        for i in range(STEPS):
            r.forward(t[i], t[i + 1] - t[i])
        assert np.array_equal(o.states()[k], r.states()[0]), n
