"""C++ host API (include/goss_hip.hpp): goss_hip::ODESystemSolver, the solver
classes with their public tunables and ParameterizedODE over the C ABI, driven
by tests/cpp/test_cpp_api.cpp.

CPU: the model metadata and solver objects work without a GPU, and building
a system without one fails loudly (no CPU fallback).  GPU: a TP06 GRL1 run
through the C++ classes (set_parameter on the model before construction,
tangled field parameters, forward, solve, states) equals the same run through
the Python binding bit for bit, and the FP64 oracle to 1e-10 relative; solve_async into pinned memory
(goss_hip_host_alloc) equals solve bit for bit, and ImplicitEuler with the analytic Jacobian agrees
with the forward differences to the Newton tolerance.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_cpp_api")


def _exe():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built; run __graft_entry__.build()")
    return EXE


def test_cpp_model_and_solver_objects():
    r = subprocess.run([_exe(), "info"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "states 19 parameters 53 field_states 1 field_parameters 3 first_state Xr1 g_Kr 0.153" in r.stdout
    assert '"no_such_parameter" is not a parameter in  this ODE.' in r.stdout  # ParameterizedODE.h:185-195
    assert "adaptive 01 delta 1e-08" in r.stdout
    # GRL1.h:39, ThetaSolver.h:37, ImplicitEuler.h:37, ImplicitODESolver.h:36-41, ESDIRK23a.h:38,
    # AdaptiveExplicitSolver.cpp:52-53, RKF32 iord 3.
    assert "tunables 1e-07 0.5 0.0001 1 0.1 30 0.001 1e-05 3" in r.stdout
    assert "block 0.75 12 1 0.02 1e-06 1e-09" in r.stdout
    assert "solver 1" in r.stdout


def test_cpp_system_without_gpu_fails_loudly():
    from goss_amd._lib import lib

    import ctypes

    n = ctypes.c_int(0)
    lib().goss_hip_device_count(ctypes.byref(n))
    if n.value > 0:
        pytest.skip("a GPU is visible")
    r = subprocess.run([_exe(), "create"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 4 and "error" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_run_equals_python_binding_and_oracle(tmp_path):
    import oracle
    from goss_amd.system import DeviceSystem

    out = tmp_path / "cpp.bin"
    r = subprocess.run([_exe(), "run", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # Tunables and model parameters changed between calls apply to the next call.
    assert "tunables and parameters changed between calls apply" in r.stdout
    assert "solve_async into pinned memory equals solve" in r.stdout
    assert "analytic Jacobian: max |dV| vs forward differences" in r.stdout
    raw = np.fromfile(out, dtype=np.float64)
    N, S, F = (int(x) for x in raw[:3])
    NFP = 3
    off = 3
    fp = raw[off:off + N * NFP].reshape(N, NFP)
    off += N * NFP
    states = raw[off:off + N * S].reshape(N, S)
    off += N * S
    rows = raw[off:].reshape(51, N * F)

    dt = 0.01
    dev = DeviceSystem("tp06_m_fp3", "GRL1", N, "fp64")
    dev.set_parameter("stim_start", 0.5)
    dev.set_field_parameters(fp, tangled=True)
    dev.reset_default()
    t = 0.0
    for _ in range(100):
        dev.forward(t, dt)
        t += dt
    times = t + np.arange(51) * dt
    prow = np.zeros((51, N * F))
    dev.solve(times, prow, tangled=True)
    assert np.array_equal(dev.states(), states)
    assert np.array_equal(prow[1:], rows[1:])

    o = oracle.OracleSystem("tp06_m_fp3", "GRL1", N)
    o.set_parameter("stim_start", 0.5)
    o.set_field_parameters(fp)
    o.reset_default()
    t = 0.0
    for _ in range(100):
        o.forward(t, dt)
        t += dt
    o.forward_times(times)
    ref = o.states()
    floor = 1e-6 * np.max(np.abs(ref), axis=0, keepdims=True)
    assert float((np.abs(states - ref) / np.maximum(np.abs(ref), floor)).max()) <= 1e-10
    assert states[:, 17].max() > 0  # the 0.5 ms stimulus start fired the upstroke
