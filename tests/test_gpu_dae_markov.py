"""The DAE and Markov-chain forms of the .ode language on the device, against
the oracle (whose solvers equal the reference's bit for bit on the same
generated models, tests/test_dsl_dae_markov.py).

* van der Pol as a DAE (one algebraic state): the implicit solvers put the
  mass matrix into the Newton residual and the iteration matrix; the explicit
  solvers, RKF32 and ESDIRK4O32 refuse the model with the reference's
  messages.
* The three-state Markov chain with a time-modulated opening rate and a
  field parameter: every solver in FP64, and FP32 RK4 against FP64.

The device keeps the implicit and adaptive solvers' state per node, as the
oracle's per-node mode does.  These models use only +, -, *, / and sin, so the
tolerance is 1e-10 relative per state (floor 1e-6 of the state's largest
magnitude).  RKF32 accepts or rejects steps on an error norm, so the last bit
of the device's sin and pow moves its step sequence: its bound is calibrated
as in test_gpu_rkf32.py, from the oracle's own spread under a one-ulp change of
the initial states, capped at 1e-8 where that pair takes the same steps.
"""
import numpy as np
import pytest

import oracle
from test_gpu_explicit import assert_parity, dev_system
from test_gpu_implicit import state_scale

pytestmark = pytest.mark.gpu

IMPLICIT = ["ImplicitEuler", "ThetaSolver", "BasicImplicitEuler", "ESDIRK23a"]


def _vdp_y0(N, seed):
    # Consistent initial values on the constraint y = (z^2/3 - 1) z; from z0 >= 1.95 the fold at z = 1
    # is reached after t = 0.73, beyond the runs below.
    z = np.random.Generator(np.random.PCG64(seed)).uniform(1.95, 2.2, N)
    return np.column_stack([(z * z / 3.0 - 1.0) * z, z])


@pytest.mark.parametrize("solver", IMPLICIT)
def test_van_der_pol_dae_matches_oracle(solver):
    N, dt, steps = 1000, 0.01, 60
    y0 = _vdp_y0(N, 31)
    t = np.arange(steps + 1) * dt
    dev = dev_system("van_der_pol_dae", solver, N, "fp64")
    o = oracle.OracleSystem("van_der_pol_dae", solver, N, per_node_state=True)
    dev.set_states(y0)
    o.set_states(y0)
    dev.forward_times(t)
    o.forward_times(t)
    a, b = dev.states(), o.states()
    if solver != "ESDIRK23a":
        # the run moved along the slow branch and stays on the constraint
        assert b[:, 1].max() < 1.95 and np.abs(a[:, 0] - (a[:, 1] ** 2 / 3 - 1) * a[:, 1]).max() < 1e-10
    assert_parity(a, b, 1e-10, f"van_der_pol_dae {solver}")
    assert np.array_equal(dev.solver_counters(), o.counters())


@pytest.mark.parametrize("solver,message", [
    ("ExplicitEuler", "an explicit solver"), ("RK4", "an explicit solver"), ("RL1", "Rush Larsen method"),
    ("GRL1", "an explicit solver"), ("RKF32", "an explicit solver"), ("ESDIRK4O32", "an explicit solver")])
def test_dae_refused_by_explicit_solvers(solver, message):
    from goss_amd._lib import GossHipError

    with pytest.raises(GossHipError, match=f"cannot integrate a DAE ode with {message}"):
        dev_system("van_der_pol_dae", solver, 4, "fp64")


def _rkf32_bound(y0, fp, t):
    """The oracle's spread between y0 and y0 moved by one ulp, as a bound (see the module docstring)."""
    runs = []
    for y in (y0, np.nextafter(y0, np.inf)):
        o = oracle.OracleSystem("markov_three_state", "RKF32", y0.shape[0], per_node_state=True)
        o.set_states(y)
        o.set_field_parameters(fp)
        o.forward_times(t)
        runs.append((o.states(), o.counters()))
    (x, ca), (z, cb) = runs
    floor = 1e-6 * state_scale(y0, x)
    spread = float((np.abs(x - z) / np.maximum(np.abs(x), floor)).max())
    same = list(ca[1:]) == list(cb[1:])
    tol = max(1e-10, 10 * spread)
    return (min(tol, 1e-8) if same else tol), same


MARKOV_SOLVERS = ["ExplicitEuler", "RK2", "RK4", "RL1", "RL2", "GRL1", "GRL2", "ImplicitEuler", "ESDIRK23a",
                  "RKF32", "ThetaSolver", "BasicImplicitEuler"]


def _markov_inputs(N, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.dirichlet(np.ones(3), N), 0.3 * np.exp(rng.normal(0.0, 0.3, (N, 1)))


@pytest.mark.parametrize("solver", MARKOV_SOLVERS)
def test_markov_chain_matches_oracle(solver):
    # RKF32 at intervals of 0.1, its first step guess: no call overshoots an interval end (the next test
    # covers calls that do).
    N, dt, steps = 4096, (0.1 if solver == "RKF32" else 0.05), 40
    y0, fp = _markov_inputs(N, 32)
    t = np.arange(steps + 1) * dt
    dev = dev_system("markov_three_state", solver, N, "fp64")
    o = oracle.OracleSystem("markov_three_state", solver, N, per_node_state=True)
    for s in (dev, o):
        s.set_states(y0)
        s.set_field_parameters(fp)
    dev.forward_times(t)
    o.forward_times(t)
    a, b = dev.states(), o.states()
    if solver not in ("RL1", "RL2", "GRL1", "GRL2"):
        # Runge-Kutta and the implicit schemes keep linear invariants; the Rush-Larsen schemes, which
        # update each state with its own exponential, do not.
        assert np.abs(b.sum(axis=1) - 1.0).max() < 1e-12
    tol = _rkf32_bound(y0, fp, t)[0] if solver == "RKF32" else 1e-10
    err = assert_parity(a, b, tol, f"markov_three_state {solver}", scale=state_scale(y0, b))
    print(f"markov_three_state {solver}: max rel err {err:.2e} (bound {tol:.1e})")


def test_markov_rkf32_initial_step_from_dtinit():
    """RKF32 calls that start from AdaptiveExplicitSolver::dtinit after an overshoot (tests/markov_cases.py),
    FP64 device against the oracle, counters included."""
    from markov_cases import DT, STEPS, finishing_nodes, markov_inputs

    y0, fp = markov_inputs(1024, 34)
    t = np.arange(STEPS + 1) * DT
    good = finishing_nodes(y0, fp, t)
    assert len(good) >= 64
    y0, fp = y0[good], fp[good]
    dev = dev_system("markov_three_state", "RKF32", len(good), "fp64")
    o = oracle.OracleSystem("markov_three_state", "RKF32", len(good), per_node_state=True)
    for s in (dev, o):
        s.set_states(y0)
        s.set_field_parameters(fp)
    dev.forward_times(t)
    o.forward_times(t)
    tol, same = _rkf32_bound(y0, fp, t)
    err = assert_parity(dev.states(), o.states(), tol, "markov_three_state RKF32 from dtinit",
                        scale=state_scale(y0, o.states()))
    print(f"markov_three_state RKF32 from dtinit: max rel err {err:.2e} (bound {tol:.1e}), "
          f"counters oracle {o.counters()} device {dev.solver_counters()}")
    if same:
        assert np.array_equal(dev.solver_counters()[1:], o.counters()[1:])


def test_markov_chain_fp32_rk4_against_fp64():
    N, dt, steps = 4096, 0.01, 200
    y0, fp = _markov_inputs(N, 33)
    t = np.arange(steps + 1) * dt
    o = oracle.OracleSystem("markov_three_state", "RK4", N)
    dev = dev_system("markov_three_state", "RK4", N, "fp32")
    for s in (dev, o):
        s.set_states(y0)
        s.set_field_parameters(fp)
    dev.forward_times(t)
    o.forward_times(t)
    err = np.abs(dev.states() - o.states()).max()
    print(f"markov FP32 RK4 vs FP64 oracle after 2 time units: max |error| {err:.2e}")
    assert err < 1e-5
