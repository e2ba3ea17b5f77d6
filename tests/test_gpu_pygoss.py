"""The goss Python API on the GPU: the expectations of the reference's
tests/test_ode_system_solver_bindings.py:11-127, test_ode_solver_bindings.py
and test_ode_bindings.py restated through goss_amd, plus the device model
methods against the oracle (eval, linearized_eval, FD Jacobian, LU and
substitution, monitored values)."""
import numpy as np
import pytest

import goss_amd as goss
import oracle
from goss_amd import solvers

pytestmark = pytest.mark.gpu

OSC = "oscilator.ode"


@pytest.fixture(scope="module")
def oscilator():
    return goss.ParameterizedODE(OSC)


@pytest.fixture(scope="module")
def oscilator_fields():
    # The reference's test_solve_system_oscilator model: field states x, y; field parameters a, b.
    return goss.ParameterizedODE(OSC, field_states=["x", "y"], field_parameters=["a", "b"])


@pytest.fixture(scope="module")
def br_fields():
    # A cell model with two field states and two field parameters (compiled at run time).
    return goss.ParameterizedODE("beeler_reuter_1977.ode", field_states=["Cai", "V"], field_parameters=["g_Na", "g_s"])


def test_construct_ode_system_solver(oscilator):
    solver = solvers.ExplicitEuler()
    system = goss.ODESystemSolver(1, solver, oscilator)
    assert system.num_nodes == 1
    assert solver.get_ode().num_states == oscilator.num_states
    assert system.solver is solver and system.ode is oscilator


def test_num_threads(oscilator):
    system = goss.ODESystemSolver(1, solvers.ExplicitEuler(), oscilator)
    assert system.num_threads == 0
    system.num_threads = 1
    assert system.num_threads == 1
    system.reset_default()


def test_states(oscilator):
    system = goss.ODESystemSolver(3, solvers.ExplicitEuler(), oscilator)
    states = system.states
    assert states.shape == (3, oscilator.num_states)
    for node in range(3):
        assert np.allclose(states[node, :], oscilator.get_ic())


def test_field_states(br_fields):
    system = goss.ODESystemSolver(3, solvers.ExplicitEuler(), br_fields)
    ic = br_fields.get_ic()
    values = ic[br_fields.field_state_indices]
    fs = system.field_states
    assert fs.shape == (3, br_fields.num_field_states)
    assert (fs[:, 0] == values[0]).all() and (fs[:, 1] == values[1]).all()
    fs[:] = 0
    system.field_states = fs
    assert (system.field_states == 0).all()
    with pytest.raises(ValueError, match="Expected shape of field states"):
        system.field_states = np.zeros((2, 2))


def test_field_parameters(br_fields):
    system = goss.ODESystemSolver(4, solvers.ExplicitEuler(), br_fields)
    fp = system.field_parameters
    assert fp.shape == (4, br_fields.num_field_parameters)
    for i, name in enumerate(br_fields.field_parameter_names):
        assert np.allclose(fp[:, i], br_fields.parameters[name])
    fp[:] = 0
    system.field_parameters = fp
    assert (system.field_parameters == 0).all()
    with pytest.raises(ValueError, match="Expected shape of field parameters"):
        system.field_parameters = np.zeros((3, 2))


@pytest.mark.parametrize("Solver", solvers.GOSSSolvers)
def test_solve_system_oscilator(Solver, oscilator_fields):
    solver = solvers.solver_mapper[Solver.name]()
    system = goss.ODESystemSolver(4, solver, oscilator_fields)
    solver.internal_time_step = 0.0001
    dt = 0.1
    t = np.arange(0, 10.0 + dt, dt)
    fs = system.solve(t)
    assert fs.shape == (t.size, 4, 2)
    for i in range(4):
        assert np.isclose(fs[:, i, 0], np.cos(t), rtol=0.14).all()
        assert np.isclose(fs[:, i, 1], np.sin(t), rtol=0.02).all()


@pytest.mark.parametrize("Solver", solvers.GOSSExplicitSolvers)
def test_forward(Solver, oscilator):
    solver = solvers.solver_mapper[Solver.name](oscilator)
    y = oscilator.get_ic()
    solver.forward(y, 0, 0.1)
    assert np.allclose(y, [1, 0.1], rtol=0.01)


@pytest.mark.parametrize("Solver", solvers.GOSSSolvers)
def test_solve_oscilator(Solver, oscilator):
    solver = solvers.solver_mapper[Solver.name](oscilator)
    solver.internal_time_step = 0.0001
    dt = 0.1
    t = np.arange(0, 10.0 + dt, dt)
    u = solver.solve(t)
    assert np.isclose(u[:, 0], np.cos(t), rtol=0.14).all()
    assert np.isclose(u[:, 1], np.sin(t), rtol=0.02).all()


@pytest.mark.parametrize("Solver", solvers.GOSSImplicitSolvers)
def test_num_jac_comp(Solver, oscilator):
    solver = solvers.solver_mapper[Solver.name](oscilator)
    assert solver.num_jac_comp() == 0
    solver.forward(oscilator.get_ic(), 0.0, 0.1)
    assert solver.num_jac_comp() > 0


def test_solver_tunables_apply_after_creation():
    """A tunable or a model parameter changed between two calls applies to the
    next call, as with the reference's solver and model objects; the result
    equals a solver created with those values from the start."""
    ode = goss.ParameterizedODE("tp06_m")
    y0 = ode.get_ic()
    a = solvers.GRL1(ode)
    ya = y0.copy()
    a.forward(ya, 0.0, 1.0)
    a.internal_time_step = 0.01
    ode.set_parameter("stim_amplitude", 40.0)
    a.forward(ya, 1.0, 10.0)
    ode2 = goss.ParameterizedODE("tp06_m")
    b = solvers.GRL1(ode2)
    yb = y0.copy()
    b.forward(yb, 0.0, 1.0)
    c = solvers.GRL1(ode2)
    c.internal_time_step = 0.01
    ode2.set_parameter("stim_amplitude", 40.0)
    c.forward(yb, 1.0, 10.0)
    np.testing.assert_array_equal(ya, yb)


def test_ode_eval_and_components(oscilator):
    values = oscilator.eval(np.array([0.1, 0.2]), 0)
    assert np.allclose(values, [-0.2, 0.1])
    assert np.allclose(oscilator.eval([0.1, 0.2], 0), [-0.2, 0.1])
    with pytest.raises(ValueError):
        oscilator.eval(np.array([0.1, 0.2, 0.3]), 0)
    assert np.isclose(oscilator.eval_component(0, np.array([0.1, 0.2]), 0), -0.2)
    assert np.isclose(oscilator.eval_component(1, np.array([0.1, 0.2]), 0), 0.1)
    with pytest.raises(RuntimeError):
        oscilator.eval_component(3, np.array([0.1, 0.2]), 0)
    with pytest.raises(TypeError):
        oscilator.eval_component(-1, np.array([0.1, 0.2]), 0)


def test_monitored_energy():
    ode = goss.ParameterizedODE(OSC, monitored=["energy"])
    states = np.array([3.0, 5.0])
    assert np.isclose(ode.eval_monitored(states, 0), states.sum())
    trace = np.array([[1.0, 2.0], [0.5, -1.0], [3.0, 5.0]])
    np.testing.assert_allclose(ode.monitored_values(trace, np.zeros(3))[:, 0], trace.sum(axis=1))


def test_model_methods_match_the_oracle_tp06():
    ode = goss.ParameterizedODE("tp06_m")
    v = oracle.Variant("tp06_m")
    rng = np.random.Generator(np.random.PCG64(5))
    y = v.initial_conditions * (1 + rng.uniform(-0.05, 0.05, 19))
    for t in (0.0, 10.5, 37.0):
        f_dev, f_cpu = ode.eval(y, t), v.eval(y, t)
        np.testing.assert_allclose(f_dev, f_cpu, rtol=1e-12, atol=1e-14 * np.abs(f_cpu).max())
        lin_dev = ode.linearized_eval(y, t)
        lin_cpu, rhs_cpu = v.linearized_eval(y, t)
        np.testing.assert_allclose(lin_dev.rhs, rhs_cpu, rtol=1e-12, atol=1e-14 * np.abs(rhs_cpu).max())
        np.testing.assert_allclose(lin_dev.linear, lin_cpu, rtol=1e-12, atol=1e-14 * np.abs(lin_cpu).max())
        J_dev, J_cpu = ode.compute_jacobian(y, t), v.fd_jacobian(y, t)
        np.testing.assert_allclose(J_dev, J_cpu, rtol=1e-6, atol=1e-6 * np.abs(J_cpu).max())
    # LU and substitution are plain arithmetic: bit for bit with the oracle's restatement.
    M = np.eye(19) - 0.1 * J_cpu
    A = M.copy()
    ode.lu_factorize(A)
    B = M.copy().reshape(-1)
    oracle.lib().oracle_lu_factorize(19, B.ctypes.data_as(oracle._dp))
    np.testing.assert_array_equal(A.reshape(-1), B)
    b = rng.uniform(-1, 1, 19)
    x = np.zeros(19)
    ode.forward_backward_subst(A, b, x)
    xo = np.zeros(19)
    oracle.lib().oracle_forward_backward_subst(19, B.ctypes.data_as(oracle._dp), b.ctypes.data_as(oracle._dp),
                                               xo.ctypes.data_as(oracle._dp))
    np.testing.assert_array_equal(x, xo)
    np.testing.assert_allclose(M @ x, b, rtol=1e-10, atol=1e-12)


def test_ode_system_solver_forward_matches_oracle_tp06():
    """ODESystemSolver.forward through the Python API on the TP06 M-cell with
    per-node field parameters, against the oracle (FP64, 1e-10)."""
    ode = goss.ParameterizedODE("tentusscher_panfilov_2006_M_cell.ode", field_states=["V"],
                                field_parameters=["g_Kr", "g_Ks", "g_CaL"])
    assert ode.variant == "tp06_m_fp3"
    N = 300
    system = goss.ODESystemSolver(N, solvers.GRL1(), ode)
    o = oracle.OracleSystem("tp06_m_fp3", "GRL1", N)
    rng = np.random.Generator(np.random.PCG64(9))
    fp = system.field_parameters * np.exp(rng.normal(0, 0.2, (N, 3)))
    system.field_parameters = fp
    o.set_field_parameters(fp)
    t = np.arange(1001) * 0.02
    for k in range(len(t) - 1):  # through the stimulus at 10 ms
        system.forward(t[k], t[k + 1] - t[k])
        o.forward(t[k], t[k + 1] - t[k])
    a, b = system.states, o.states()
    assert b[:, 17].max() > 0  # upstroke
    np.testing.assert_allclose(a, b, rtol=1e-10, atol=0)
