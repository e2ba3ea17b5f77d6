"""The CPU oracle (oracle/goss_oracle.cpp) pinned before it is trusted.

* bit-for-bit against the golden trajectories written by the reference's own
  ODESystemSolver (tests/golden/ref_*.npz, tests/golden/make_ref_trajectories.py);
* bit-for-bit against the live reference build (oracle/_ref) for every
  explicit solver, when that build is present;
* closed forms: Rush-Larsen schemes are exact for y' = a*y + b, Euler/RK
  convergence orders, the reference's oscillator expectations
  (tests/test_ode_solver_bindings.py:57-66,125-136);
* dense helpers: no-pivot LU and substitution against numpy.
"""
import numpy as np
import pytest

import oracle
from conftest import golden

EXPLICIT = ["ExplicitEuler", "RK2", "RK4", "RL1", "RL2", "GRL1", "GRL2"]


def test_oracle_matches_reference_tp06_1024_node_fixture():
    d = np.load(golden("ref_tp06_grl1_1024x10000.npz"))
    o = oracle.OracleSystem("tp06_m", "GRL1", 1024)
    o.set_num_threads(8)
    o.set_states(d["y0"])
    t = d["t"]
    trace = [o.states()[:8, 17].copy()]
    for k in range(0, len(t) - 1, 100):
        o.forward_times(t[k:k + 101])
        trace.append(o.states()[:8, 17].copy())
    assert np.array_equal(o.states(), d["final"])
    assert np.array_equal(np.array(trace), d["V_trace"])
    # The run goes through a full upstroke.
    assert d["V_trace"].max() > 20 and d["V_trace"][-1].max() < 40


@pytest.mark.parametrize("solver", ["GRL1", "RL2"])
def test_oracle_matches_reference_field_parameter_window(solver):
    d = np.load(golden("ref_tp06_fp3_window.npz"))
    N = d["y0"].shape[0]
    o = oracle.OracleSystem("tp06_m_fp3", solver, N)
    o.set_states(d["y0"])
    o.set_field_parameters(d["field_parameters"])
    o.forward_times(np.arange(int(d["steps"]) + 1) * float(d["dt"]))
    assert np.array_equal(o.states(), d[f"final_{solver}"])


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("solver", EXPLICIT)
@pytest.mark.parametrize("variant,ldt,dt", [("tp06_m", -1.0, 0.01), ("tp06_m_fp3", 0.005, 0.02),
                                             ("beeler_reuter_1977", 0.01, 0.1), ("fhn", -1.0, 0.1)])
def test_oracle_matches_reference_build(solver, variant, ldt, dt):
    N = 17
    rng = np.random.Generator(np.random.PCG64(11))
    o = oracle.OracleSystem(variant, solver, N, {"internal_time_step": ldt})
    r = oracle.RefSystem(variant, solver, N, {"internal_time_step": ldt})
    y0 = o.states() * (1 + rng.uniform(-0.01, 0.01, o.states().shape))
    o.set_states(y0)
    r.set_states(y0)
    if o.variant.num_field_parameters:
        fp = o.field_parameters() * np.exp(rng.normal(0, 0.2, o.field_parameters().shape))
        o.set_field_parameters(fp)
        r.set_field_parameters(fp)
    t = np.arange(301) * dt
    o.forward_times(t)
    r.forward_times(t)
    a, b = o.states(), r.states()
    if solver == "GRL2":
        # GRL2.cpp:67 reads an uninitialised `delta`; the oracle uses the
        # member value 1e-8, so only agreement to rounding is expected.
        finite = np.isfinite(b)
        np.testing.assert_allclose(a[finite], b[finite], rtol=1e-8, atol=1e-12)
    else:
        assert np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("solver", ["RL1", "GRL1"])
def test_rush_larsen_exact_for_affine(solver):
    # y' = a y + b has y(t) = (y0 + b/a) e^{a t} - b/a; the one-stage
    # Rush-Larsen schemes integrate it exactly up to rounding.
    N = 64
    rng = np.random.Generator(np.random.PCG64(5))
    a = -rng.uniform(0.1, 5.0, N)
    b = rng.uniform(-3, 3, N)
    y0 = rng.uniform(-2, 2, N)
    o = oracle.OracleSystem("linear_decay", solver, N)
    o.set_states(y0[:, None])
    o.set_field_parameters(np.column_stack([a, b]))
    t = np.linspace(0.0, 2.0, 41)
    o.forward_times(t)
    exact = (y0 + b / a) * np.exp(a * 2.0) - b / a
    np.testing.assert_allclose(o.states()[:, 0], exact, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("solver,order", [("ExplicitEuler", 1), ("RK2", 2), ("RK4", 4), ("RL2", 1), ("GRL2", 1)])
def test_explicit_convergence_order(solver, order):
    # RL2/GRL2 as written in the reference (RL2.cpp:75-85) apply the
    # half-step rate f(y2)/a from y0 without the a*(y0 - y2) correction of
    # the linearisation, so on y' = a y + b the local error is
    # (y2 - y0)(e^{ah} - 1) = O(h^2): first order globally.
    a, b, y0, T = -1.5, 2.0, 1.0, 1.0
    exact = (y0 + b / a) * np.exp(a * T) - b / a
    errs = []
    for n in (20, 40):
        o = oracle.OracleSystem("linear_decay", solver, 1)
        o.set_states(np.array([[y0]]))
        o.set_field_parameters(np.array([[a, b]]))
        o.forward_times(np.linspace(0.0, T, n + 1))
        errs.append(abs(o.states()[0, 0] - exact))
    rate = np.log2(errs[0] / errs[1])
    assert abs(rate - order) < 0.2, rate


def test_oscillator_expectations():
    # tests/test_ode_solver_bindings.py:57-66: one ExplicitEuler forward step
    # of size 0.1 from (1, 0) gives (1, 0.1).
    o = oracle.OracleSystem("oscilator", "ExplicitEuler", 1)
    o.forward(0.0, 0.1)
    np.testing.assert_allclose(o.states()[0], [1.0, 0.1])
    # tests/test_ode_solver_bindings.py:125-136: solve against cos/sin.
    t = np.linspace(0, 10, 101)
    for solver, rtol in (("ExplicitEuler", 0.14), ("RK4", 0.02), ("GRL1", 0.14)):
        o = oracle.OracleSystem("oscilator", solver, 4, {"internal_time_step": 0.001})
        out = o.solve(t)
        np.testing.assert_allclose(out[:, 0, 0], np.cos(t), rtol=rtol, atol=rtol)
        np.testing.assert_allclose(out[:, 0, 1], np.sin(t), rtol=rtol, atol=rtol)


def test_lu_and_substitution_against_numpy():
    rng = np.random.Generator(np.random.PCG64(9))
    L = oracle.lib()
    for n in (1, 2, 5, 19):
        A = rng.normal(size=(n, n)) + n * np.eye(n)  # diagonally dominant: no pivoting needed
        M = np.ascontiguousarray(A.copy())
        L.oracle_lu_factorize(n, oracle._ptr(M))
        lower = np.tril(M, -1) + np.eye(n)
        upper = np.triu(M)
        np.testing.assert_allclose(lower @ upper, A, rtol=1e-12, atol=1e-12)
        b = rng.normal(size=n)
        x = np.zeros(n)
        L.oracle_forward_backward_subst(n, oracle._ptr(M), oracle._ptr(np.ascontiguousarray(b)), oracle._ptr(x))
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_fd_jacobian_matches_reference_base_class():
    if not oracle.ref_available():
        pytest.skip("reference build not present")
    v = oracle.Variant("tp06_m")
    r = oracle.RefModel("tp06_m")
    y = v.initial_conditions * 1.01
    np.testing.assert_array_equal(v.fd_jacobian(y, 10.5), r.compute_jacobian(y, 10.5))


def test_field_storage_roundtrip():
    N = 5
    o = oracle.OracleSystem("tp06_m_fp3", "GRL1", N)
    fp = np.arange(N * 3, dtype=float).reshape(N, 3) + 1
    o.set_field_parameters(fp)
    np.testing.assert_array_equal(o.field_parameters(), fp)
    o.set_field_parameters(np.ascontiguousarray(fp.T), tangled=False)
    np.testing.assert_array_equal(o.field_parameters(), fp)
    fs = np.linspace(-80, -70, N)[:, None]
    o.set_field_states(fs)
    np.testing.assert_array_equal(o.states()[:, 17], fs[:, 0])


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("solver", ["ImplicitEuler", "ESDIRK23a"])
@pytest.mark.parametrize("variant,dt,ldt", [("tp06_epi", 0.1, -1.0), ("tp06_m", 0.1, -1.0), ("tp06_m", 0.1, 0.05),
                                            ("beeler_reuter_1977", 0.5, -1.0), ("fhn", 0.5, -1.0)])
@pytest.mark.parametrize("threads", [0, 3])
def test_oracle_implicit_matches_reference_build(solver, variant, dt, ldt, threads):
    # The reference's carried solver state (one instance per thread, static
    # partition) is reproduced exactly by the oracle's default mode.
    N = 17
    rng = np.random.Generator(np.random.PCG64(13))
    o = oracle.OracleSystem(variant, solver, N, {"internal_time_step": ldt})
    r = oracle.RefSystem(variant, solver, N, {"internal_time_step": ldt})
    if threads:
        o.set_num_threads(threads)
        r.set_num_threads(threads)
    y0 = o.states() * (1 + rng.uniform(-0.01, 0.01, o.states().shape))
    o.set_states(y0)
    r.set_states(y0)
    t = np.arange(151) * dt
    o.forward_times(t)
    r.forward_times(t)
    assert np.array_equal(o.states(), r.states(), equal_nan=True)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build oracle/_ref not present")
def test_oracle_implicit_failure_message_matches_reference():
    params = {"max_iterations": 0, "min_dt": 0.5}
    o = oracle.OracleSystem("tp06_m", "ImplicitEuler", 2, params)
    r = oracle.RefSystem("tp06_m", "ImplicitEuler", 2, params)
    with pytest.raises(RuntimeError) as a:
        o.forward(0.0, 0.1)
    with pytest.raises(RuntimeError) as b:
        r.forward(0.0, 0.1)
    assert str(a.value) == str(b.value)


def test_esdirk_per_node_state_is_order_independent():
    # With per-node solver state a node's result does not depend on the
    # other nodes: permuting the nodes permutes the results exactly.
    N = 9
    rng = np.random.Generator(np.random.PCG64(14))
    a = oracle.OracleSystem("tp06_epi", "ESDIRK23a", N, per_node_state=True)
    y0 = a.states() * (1 + rng.uniform(-0.01, 0.01, a.states().shape))
    perm = rng.permutation(N)
    b = oracle.OracleSystem("tp06_epi", "ESDIRK23a", N, per_node_state=True)
    a.set_states(y0)
    b.set_states(y0[perm])
    t = np.arange(101) * 0.1
    a.forward_times(t)
    b.forward_times(t)
    assert np.array_equal(a.states()[perm], b.states())


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("variant,dt,nsteps,params", [
    ("lorentz", 0.01, 300, {}),
    ("lorentz", 0.05, 80, {"atol": 1e-7, "rtol": 1e-9}),
    ("fitzhughnagumo", 0.5, 100, {}),
    ("tp06_m", 0.1, 150, {}),            # through the stimulus at 10 ms
    ("oscilator", 0.1, 100, {"atol": 1e-8}),
])
def test_oracle_rkf32_matches_reference_build(variant, dt, nsteps, params):
    # Serial mode: one RKF32 instance carries its step-size guess (_ldt) and
    # facmax from node to node (AdaptiveExplicitSolver.cpp:216-241); the
    # oracle's default mode does the same.  Threaded reference runs are not
    # compared: the thread copies leave RKF32::first uninitialised (RKF32.h:95
    # is not copied by RKF32.cpp:52-61).
    N = 11
    rng = np.random.Generator(np.random.PCG64(21))
    o = oracle.OracleSystem(variant, "RKF32", N, params)
    r = oracle.RefSystem(variant, "RKF32", N, params)
    y0 = o.states() * (1 + rng.uniform(-0.01, 0.01, o.states().shape))
    o.set_states(y0)
    r.set_states(y0)
    t = np.arange(nsteps + 1) * dt
    o.forward_times(t)
    r.forward_times(t)
    assert np.array_equal(o.states(), r.states())
    nfevals, ndtsa, ndtsr = r.rkf32_counters()
    oc = o.counters()
    assert (oc[1], oc[3], oc[2]) == (nfevals, ndtsa, ndtsr)


def test_rkf32_linear_decay_tolerance_and_order():
    # y' = a*y + b on one node: the global error of RKF32 at t = 2 stays
    # within a small multiple of the tolerance, and tightening atol/rtol
    # by 1e3 cuts the error by at least 10x (a 3(2) pair, error ~ tol^(3/4)).
    a, b = -1.3, 0.7
    exact = lambda t, y0: (y0 + b / a) * np.exp(a * t) - b / a
    errs = []
    for tol in (1e-6, 1e-9):
        o = oracle.OracleSystem("linear_decay", "RKF32", 1, {"atol": tol, "rtol": tol})
        o.set_field_parameters(np.array([[a, b]]))
        o.set_states(np.array([[1.0]]))
        o.forward_times(np.linspace(0.0, 2.0, 21))
        errs.append(abs(o.states()[0, 0] - exact(2.0, 1.0)))
        assert errs[-1] < 200 * tol
    assert errs[1] < errs[0] / 10


def test_rkf32_per_node_state_is_order_independent():
    N = 7
    rng = np.random.Generator(np.random.PCG64(22))
    a = oracle.OracleSystem("lorentz", "RKF32", N, per_node_state=True)
    y0 = a.states() * (1 + rng.uniform(-0.05, 0.05, a.states().shape))
    perm = rng.permutation(N)
    b = oracle.OracleSystem("lorentz", "RKF32", N, per_node_state=True)
    a.set_states(y0)
    b.set_states(y0[perm])
    t = np.arange(101) * 0.01
    a.forward_times(t)
    b.forward_times(t)
    assert np.array_equal(a.states()[perm], b.states())


def test_rkf32_stuck_node_raises():
    # A NaN state gives a non-finite error norm on every step (taken as 2000,
    # AdaptiveExplicitSolver.cpp:195-196); the step size underflows and the
    # reference would loop forever.  The oracle stops after its sub-step cap.
    o = oracle.OracleSystem("lorentz", "RKF32", 1)
    o.set_states(np.array([[np.nan, 0.0, 0.0]]))
    with pytest.raises(RuntimeError, match="RKF32 did not reach the end of the interval"):
        o.forward(0.0, 0.01)
