"""Host build of the implicit and RKF32 kernels' per-node logic against the oracle.

tests/harness/kernel_logic_cpu.hip compiles goss_amd/csrc/goss_implicit.h and
goss_adaptive.h for the host (the same source the GPU kernels instantiate) and runs it node by
node with the host math library.  The oracle's per-node-state mode
(oracle/goss_oracle.cpp) gives each node its own solver state, which is what
the device kernels do; with the same libm the two must agree bit for bit, so
any difference is a logic difference of the kernel code (an operation order,
a dropped reset, a wrong branch) and not floating-point noise.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HARNESS = os.path.join(HERE, "harness", "libkernel_logic_cpu.so")
META = os.path.join(ROOT, "goss_amd", "csrc", "generated", "variants.json")

SOLVER_IDS = {"ImplicitEuler": 7, "ESDIRK23a": 8, "BasicImplicitEuler": 11, "ESDIRK4O32": 12}
# ImplicitODESolver / ImplicitEuler / ESDIRK23a defaults
# (ImplicitODESolver.cpp:31-45, ImplicitEuler.h:36-37, ESDIRK23a.cpp:46-60).
DEFAULTS = dict(eta_0=1.0, kappa=0.1, relative_tolerance=1e-12, max_relative_previous_residual=0.01,
                max_iterations=30, always=0, num_refinements=2)


def _harness():
    if not os.path.exists(HARNESS):
        pytest.skip("host build of the kernel logic is not built (run __graft_entry__.build())")
    L = ctypes.CDLL(HARNESS)
    dp = ctypes.POINTER(ctypes.c_double)
    L.kernel_logic_run.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint, dp, dp, ctypes.c_ulong, dp,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp,
                                   ctypes.POINTER(ctypes.c_long)]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def run_harness(variant, solver, y0, t, ldt=-1.0, min_dt=None):
    L = _harness()
    meta = json.load(open(META))[variant]
    params = np.ascontiguousarray(meta["default_parameters"], dtype=np.float64)
    y0 = np.ascontiguousarray(y0, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    N, S = y0.shape
    out = np.zeros_like(y0)
    counters = (ctypes.c_long * 4)()
    if min_dt is None:
        min_dt = 1e-3 if solver == "ESDIRK23a" else 1e-4
    d = DEFAULTS
    rc = L.kernel_logic_run(variant.encode(), SOLVER_IDS[solver], N, _p(y0), _p(t), len(t), _p(params), ldt,
                            d["eta_0"], d["kappa"], d["relative_tolerance"], d["max_relative_previous_residual"],
                            min_dt, d["max_iterations"], d["always"], d["num_refinements"], _p(out), counters)
    assert rc in (0, 1)
    return out, np.array(list(counters)), rc


def run_oracle(variant, solver, y0, t, ldt=-1.0):
    N = y0.shape[0]
    params = {"internal_time_step": ldt}
    o = oracle.OracleSystem(variant, solver, N, solver_params=params, per_node_state=True)
    o.set_states(y0)
    for k in range(len(t) - 1):
        o.forward(t[k], t[k + 1] - t[k])
    return o.states(), o.counters()


CASES = [
    ("tp06_m", 0.05, 300, -1.0),      # crosses the V < -40 switch of the h/j gates after the stimulus
    ("tp06_epi", 0.1, 120, -1.0),
    ("beeler_reuter_1977", 0.1, 100, -1.0),
    ("fhn", 0.5, 60, -1.0),
    ("tp06_epi", 0.1, 60, 0.05),
]


@pytest.mark.parametrize("solver", ["ImplicitEuler", "ESDIRK23a", "BasicImplicitEuler"])
@pytest.mark.parametrize("variant,dt,nsteps,ldt", CASES)
def test_kernel_logic_matches_oracle_bit_exact(variant, dt, nsteps, ldt, solver):
    S = len(json.load(open(META))[variant]["state_names"])
    o = oracle.OracleSystem(variant, solver, 1, per_node_state=True)
    base = o.states()[0]
    rng = np.random.Generator(np.random.PCG64(41))
    N = 9
    y0 = base[None, :] * (1.0 + rng.uniform(-0.01, 0.01, (N, S)))
    t = np.arange(nsteps + 1) * dt
    want, want_c = run_oracle(variant, solver, y0, t, ldt)
    got, got_c, _ = run_harness(variant, solver, y0, t, ldt)
    np.testing.assert_array_equal(got, want)
    # counters: Newton iterations of the last solve, Jacobians, rejects, accepted
    np.testing.assert_array_equal(got_c[1:], want_c[1:])


@pytest.mark.parametrize("variant,dt,nsteps", [("fhn", 0.05, 3), ("fhn", 0.02, 5)])
def test_esdirk4o32_kernel_logic_matches_oracle_bit_exact(variant, dt, nsteps):
    """ESDIRK4O32 as the reference computes it (its stage-2 evaluation at the increment, its
    two-buffer update): 1e4-1e5 steps per interval, so short runs on a small model."""
    S = len(json.load(open(META))[variant]["state_names"])
    o = oracle.OracleSystem(variant, "ESDIRK4O32", 1, per_node_state=True)
    base = o.states()[0]
    rng = np.random.Generator(np.random.PCG64(42))
    y0 = base[None, :] * (1.0 + rng.uniform(-0.01, 0.01, (3, S)))
    t = np.arange(nsteps + 1) * dt
    want, want_c = run_oracle(variant, "ESDIRK4O32", y0, t)
    got, got_c, _ = run_harness(variant, "ESDIRK4O32", y0, t)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got_c[1:], want_c[1:])
    assert want_c[3] > 0


@pytest.mark.parametrize("variant,dt,nsteps,tol", [
    ("tp06_m", 0.1, 150, (1e-5, 1e-8)),
    ("lorentz", 0.01, 100, (1e-5, 1e-8)),
    ("lorentz", 0.05, 60, (1e-7, 1e-9)),
    ("fitzhughnagumo", 0.5, 100, (1e-5, 1e-8)),
])
def test_rkf32_kernel_logic_matches_oracle_bit_exact(variant, dt, nsteps, tol):
    L = _harness()
    o = oracle.OracleSystem(variant, "RKF32", 7, {"atol": tol[0], "rtol": tol[1]}, per_node_state=True)
    rng = np.random.Generator(np.random.PCG64(42))
    y0 = o.states() * (1.0 + rng.uniform(-0.01, 0.01, o.states().shape))
    t = np.arange(nsteps + 1) * dt
    o.set_states(y0)
    o.forward_times(t)
    params = np.ascontiguousarray(json.load(open(META))[variant]["default_parameters"], dtype=np.float64)
    out = np.zeros_like(y0)
    counters = (ctypes.c_long * 4)()
    # solver 9 = RKF32; atol / rtol travel in the eta_0 / kappa slots of the harness call.
    rc = L.kernel_logic_run(variant.encode(), 9, y0.shape[0], _p(y0), _p(t), len(t), _p(params), -1.0, tol[0],
                            tol[1], 0.0, 0.0, 0.0, 0, 0, 0, _p(out), counters)
    assert rc == 0
    np.testing.assert_array_equal(out, o.states())
    np.testing.assert_array_equal(np.array(list(counters))[1:], o.counters()[1:])


def test_rkf32_kernel_logic_initial_step_from_dtinit_matches_oracle():
    """The Markov chain at intervals of 0.05 (tests/markov_cases.py): calls that start from
    AdaptiveExplicitSolver::dtinit after an overshoot, with a field parameter per node."""
    from markov_cases import DT, STEPS, finishing_nodes, markov_inputs

    L = _harness()
    dp = ctypes.POINTER(ctypes.c_double)
    L.kernel_logic_run_rkf32_fp.argtypes = [ctypes.c_char_p, ctypes.c_uint, dp, dp, dp, ctypes.c_ulong, dp,
                                            ctypes.c_double, ctypes.c_double, dp, ctypes.POINTER(ctypes.c_long)]
    y0, fp = markov_inputs(256, 32)
    t = np.arange(STEPS + 1) * DT
    good = finishing_nodes(y0, fp, t)
    assert len(good) >= 20
    y0, fp = np.ascontiguousarray(y0[good]), np.ascontiguousarray(fp[good])
    o = oracle.OracleSystem("markov_three_state", "RKF32", len(good), per_node_state=True)
    o.set_states(y0)
    o.set_field_parameters(fp)
    o.forward_times(t)
    params = np.ascontiguousarray(json.load(open(META))["markov_three_state"]["default_parameters"], dtype=np.float64)
    out = np.zeros_like(y0)
    counters = (ctypes.c_long * 4)()
    rc = L.kernel_logic_run_rkf32_fp(b"markov_three_state", len(good), _p(y0), _p(fp), _p(t), len(t), _p(params),
                                     1e-5, 1e-8, _p(out), counters)
    assert rc == 0
    np.testing.assert_array_equal(out, o.states())
    np.testing.assert_array_equal(np.array(list(counters))[1:], o.counters()[1:])


STEP_VJP_SOLVERS = {"GRL1": 5, "RL1": 3, "ExplicitEuler": 0, "RK4": 2, "RK2": 1, "RL2": 4, "GRL2": 6}


@pytest.mark.parametrize("solver", list(STEP_VJP_SOLVERS))
@pytest.mark.parametrize("fp64,tol", [(1, 1e-12), (0, 2e-4)])
def test_step_adjoint_logic_matches_torch_autograd(solver, fp64, tol):
    """The reverse-sweep kernel's step adjoint (StepVjp in goss_adjoint.h, built for the host) against
    torch autograd of the same step (tests/torch_reference.py, FP64) for tp06_m_fp6: lp = (dy'/dy)^T lam
    and pb = (dy'/d field parameters)^T lam, at states on both sides of the V = -40 mV switch of the
    sodium gates and with the stimulus on."""
    import torch

    from torch_reference import TorchModel

    L = _harness()
    dp = ctypes.POINTER(ctypes.c_double)
    L.kernel_logic_step_vjp.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, dp, dp,
                                        ctypes.c_double, dp, dp]
    meta = json.load(open(META))["tp06_m_fp6"]
    tm = TorchModel("tp06_m_fp6")
    S = tm.S
    fpi = tm.field_parameter_indices
    rng = np.random.Generator(np.random.PCG64(91))
    ic = np.asarray(meta["initial_conditions"], dtype=np.float64)
    vi = meta["state_names"].index("V")
    for V in (-85.0, -41.0, -39.0, 15.0):
        y = ic * (1.0 + rng.uniform(-0.02, 0.02, S))
        y[vi] = V
        prm = np.asarray(meta["default_parameters"], dtype=np.float64) * np.exp(rng.normal(0.0, 0.1, tm.P))
        lam = rng.normal(size=S)
        t, h = 10.3, 0.01
        lp = np.zeros(S)
        pb = np.zeros(len(fpi))
        rc = L.kernel_logic_step_vjp(STEP_VJP_SOLVERS[solver], fp64, _p(y), t, h, _p(lam), _p(prm), 1e-8, _p(lp),
                                     _p(pb))
        assert rc == 0
        yt = torch.tensor(y[None], requires_grad=True)
        pt = torch.tensor(prm[None])
        ft = torch.tensor(prm[fpi][None], requires_grad=True)
        cols = [pt[:, k] for k in range(tm.P)]
        for j, k in enumerate(fpi):
            cols[k] = ft[:, j]
        y1 = tm.step(solver, yt, t, h, torch.stack(cols, dim=1))
        (y1[0] * torch.tensor(lam)).sum().backward()
        ref_y, ref_p = yt.grad.numpy()[0], ft.grad.numpy()[0]
        err_y = np.abs(lp - ref_y).max() / np.abs(ref_y).max()
        err_p = np.abs(pb - ref_p).max() / np.abs(ref_p).max()
        assert err_y <= tol and err_p <= tol, (solver, fp64, V, err_y, err_p)


@pytest.mark.parametrize("theta", [0.5, 1.0, 0.7])
@pytest.mark.parametrize("variant,dt,nsteps,ldt", [("tp06_m", 0.05, 300, -1.0), ("beeler_reuter_1977", 0.1, 100, -1.0),
                                                   ("fhn", 0.5, 60, -1.0), ("tp06_epi", 0.1, 60, 0.05)])
def test_theta_kernel_logic_matches_oracle_bit_exact(variant, dt, nsteps, ldt, theta):
    """ThetaSolver (ThetaSolver.cpp:48-155): the kernel's per-node logic on the host against the oracle's
    per-node mode."""
    L = _harness()
    dp = ctypes.POINTER(ctypes.c_double)
    L.kernel_logic_run_theta.argtypes = [ctypes.c_char_p, ctypes.c_uint, dp, dp, ctypes.c_ulong, dp, ctypes.c_double,
                                         ctypes.c_double, dp, ctypes.POINTER(ctypes.c_long)]
    S = len(json.load(open(META))[variant]["state_names"])
    o = oracle.OracleSystem(variant, "ThetaSolver", 7, {"internal_time_step": ldt, "theta": theta},
                            per_node_state=True)
    rng = np.random.Generator(np.random.PCG64(44))
    y0 = o.states() * (1.0 + rng.uniform(-0.01, 0.01, (7, S)))
    t = np.arange(nsteps + 1) * dt
    o.set_states(y0)
    o.forward_times(t)
    params = np.ascontiguousarray(json.load(open(META))[variant]["default_parameters"], dtype=np.float64)
    out = np.zeros_like(y0)
    counters = (ctypes.c_long * 4)()
    rc = L.kernel_logic_run_theta(variant.encode(), 7, _p(y0), _p(t), len(t), _p(params), ldt, theta, _p(out),
                                  counters)
    assert rc == 0
    np.testing.assert_array_equal(out, o.states())
    np.testing.assert_array_equal(np.array(list(counters))[1:3], o.counters()[1:3])
