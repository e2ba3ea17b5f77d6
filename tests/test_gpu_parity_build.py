"""Implicit solvers and RKF32 on the GPU, bit for bit against the host.

The product's FP64 kernels call the device library's exp/log/pow/sqrt, whose
last bit differs from glibc's; the implicit and adaptive solvers turn such
differences into different Newton and step-size decisions, so the product is
compared with the oracle at a tolerance (tests/test_gpu_implicit.py).  Here
the same solver source is built twice with GOSS_PARITY_MATH, whose FP64
elementary functions (goss_amd/csrc/goss_parity_math.h) round identically on
the host and the device:

* tests/harness/libgoss_hip_parity.so -- the device library (C ABI);
* tests/harness/libkernel_logic_cpu_parity.so -- the kernels' per-node logic
  built for the host.

Their states and counters must be equal bit for bit.  The host build with
glibc equals the oracle bit for bit (tests/test_kernel_logic_cpu.py), and the
oracle equals the reference (tests/test_oracle_pinned.py), so the device runs
exactly the reference's solver logic; only the elementary functions differ.
"""
import ctypes
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DEV_LIB = os.path.join(HERE, "harness", "libgoss_hip_parity.so")
HOST_LIB = os.path.join(HERE, "harness", "libkernel_logic_cpu_parity.so")
META = os.path.join(ROOT, "goss_amd", "csrc", "generated", "variants.json")
SOLVER_IDS = {"ImplicitEuler": 7, "ESDIRK23a": 8, "RKF32": 9, "ThetaSolver": 10, "BasicImplicitEuler": 11,
              "ESDIRK4O32": 12}


def _libs():
    from goss_amd._lib import _share_the_hsa_runtime_with_torch

    for p in (DEV_LIB, HOST_LIB):
        if not os.path.exists(p):
            pytest.fail(f"{p} is not built (run __graft_entry__.build())")
    # The host build links the HIP runtime too: torch's runtime must come first, or torch (which the
    # tests after this one use) finds no device (goss_amd._lib explains why).
    _share_the_hsa_runtime_with_torch()
    L = ctypes.CDLL(HOST_LIB)
    dp = ctypes.POINTER(ctypes.c_double)
    L.kernel_logic_run.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint, dp, dp, ctypes.c_ulong, dp,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp,
                                   ctypes.POINTER(ctypes.c_long)]
    L.kernel_logic_run_theta.argtypes = [ctypes.c_char_p, ctypes.c_uint, dp, dp, ctypes.c_ulong, dp, ctypes.c_double,
                                         ctypes.c_double, dp, ctypes.POINTER(ctypes.c_long)]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _run_host(variant, solver, y0, t, params, relative_tolerance, atol=1e-5, rtol=1e-8):
    L = _libs()
    out = np.zeros_like(y0)
    counters = (ctypes.c_long * 4)()
    if solver == "ThetaSolver":
        rc = L.kernel_logic_run_theta(variant.encode(), y0.shape[0], _p(y0), _p(t), len(t), _p(params), -1.0, 0.5,
                                      _p(out), counters)
    elif solver == "RKF32":  # atol / rtol travel in the eta_0 / kappa slots of the harness call
        rc = L.kernel_logic_run(variant.encode(), 9, y0.shape[0], _p(y0), _p(t), len(t), _p(params), -1.0, atol, rtol,
                                0.0, 0.0, 0.0, 0, 0, 0, _p(out), counters)
    else:
        min_dt = 1e-3 if solver == "ESDIRK23a" else 1e-4
        rc = L.kernel_logic_run(variant.encode(), SOLVER_IDS[solver], y0.shape[0], _p(y0), _p(t), len(t),
                                _p(params), -1.0, 1.0, 0.1, relative_tolerance, 0.01, min_dt, 30, 0, 2, _p(out),
                                counters)
    assert rc == 0, rc
    return out, np.array(list(counters))


def _run_device(variant, solver, y0, t, params, relative_tolerance):
    from goss_amd.system import DeviceSystem

    sp = {"relative_tolerance": relative_tolerance} if solver == "ImplicitEuler" else {}
    s = DeviceSystem(variant, solver, y0.shape[0], "fp64", sp, library=DEV_LIB)
    names = s.model.parameter_names
    for k, v in enumerate(params):
        if v != s.model.default_parameters[k]:
            s.set_parameter(names[k], float(v))
    s.set_steps_per_launch(100)
    s.set_states(y0)
    s.forward_times(t)
    out, c = s.states(), s.solver_counters()
    s.close()
    return out, c


@pytest.mark.parametrize("variant,solver,N,dt,steps,stim,reltol", [
    # Config-3 shape: TP06 epi-cell, dt 0.1 ms, stimulated at 10 ms, through the action potential (300 ms).
    ("tp06_epi", "ImplicitEuler", 1024, 0.1, 3000, 52.0, 1e-6),
    ("tp06_epi", "ESDIRK23a", 1024, 0.1, 3000, 52.0, None),
    # TP06 M-cell with its default stimulus, through the upstroke.
    ("tp06_m", "ImplicitEuler", 256, 0.1, 400, None, 1e-12),
    ("tp06_m", "ESDIRK23a", 256, 0.1, 400, None, None),
    ("tp06_m", "ThetaSolver", 256, 0.1, 400, None, None),
    ("tp06_epi", "ThetaSolver", 1024, 0.1, 1000, 52.0, None),
    ("tp06_m", "RKF32", 128, 0.1, 200, None, None),
    ("lorentz", "RKF32", 1024, 0.02, 500, None, None),
    ("tp06_epi", "BasicImplicitEuler", 256, 0.1, 300, 52.0, None),
    # ESDIRK4O32 as the reference computes it takes 1e4-1e5 steps per interval on FHN.
    ("fhn", "ESDIRK4O32", 16, 0.05, 3, None, None),
])
def test_device_equals_host_bit_for_bit(variant, solver, N, dt, steps, stim, reltol):
    meta = json.load(open(META))[variant]
    params = np.ascontiguousarray(meta["default_parameters"], dtype=np.float64)
    if stim is not None:
        params[meta["parameter_names"].index("stim_amplitude")] = stim
    rng = np.random.Generator(np.random.PCG64(2027))
    ic = np.array(meta["initial_conditions"])
    y0 = np.ascontiguousarray(ic * (1.0 + rng.uniform(-0.01, 0.01, (N, ic.size))))
    t = np.arange(steps + 1) * dt
    reltol = 1e-12 if reltol is None else reltol
    host, host_c = _run_host(variant, solver, y0, t, params, reltol)
    dev, dev_c = _run_device(variant, solver, y0, t, params, reltol)
    assert np.all(np.isfinite(host))
    np.testing.assert_array_equal(dev, host)
    # Counters: implicit -- Jacobians, rejections, accepted steps (the first entry holds only the last
    # solve's Newton iterations); RKF32 -- nfevals, rejected, accepted steps.
    np.testing.assert_array_equal(dev_c[1:], host_c[1:])


def test_rkf32_initial_step_from_dtinit_device_equals_host():
    """RKF32 calls that start from AdaptiveExplicitSolver::dtinit (tests/markov_cases.py) on the Markov chain
    with one field parameter per node.  The opening rate's time modulation is switched off (a_co = 0: the
    product with sin is an exact zero on both sides, so sin's last bit, which the parity build does not
    control, cannot enter)."""
    from goss_amd.system import DeviceSystem
    from markov_cases import DT, STEPS, finishing_nodes, markov_inputs

    L = _libs()
    dp = ctypes.POINTER(ctypes.c_double)
    L.kernel_logic_run_rkf32_fp.argtypes = [ctypes.c_char_p, ctypes.c_uint, dp, dp, dp, ctypes.c_ulong, dp,
                                            ctypes.c_double, ctypes.c_double, dp, ctypes.POINTER(ctypes.c_long)]
    meta = json.load(open(META))["markov_three_state"]
    names = meta["parameter_names"]
    params = np.ascontiguousarray(meta["default_parameters"], dtype=np.float64)
    params[names.index("a_co")] = 0.0
    y0, fp = markov_inputs(1024, 35)
    t = np.arange(STEPS + 1) * DT
    good = finishing_nodes(y0, fp, t, {"a_co": 0.0})
    assert len(good) >= 64
    y0, fp = np.ascontiguousarray(y0[good]), np.ascontiguousarray(fp[good])
    host = np.zeros_like(y0)
    host_c = (ctypes.c_long * 4)()
    assert L.kernel_logic_run_rkf32_fp(b"markov_three_state", len(good), _p(y0), _p(fp), _p(t), len(t), _p(params),
                                       1e-5, 1e-8, _p(host), host_c) == 0
    s = DeviceSystem("markov_three_state", "RKF32", len(good), "fp64", {}, library=DEV_LIB)
    s.set_parameter("a_co", 0.0)
    s.set_states(y0)
    s.set_field_parameters(fp)
    s.forward_times(t)
    dev, dev_c = s.states(), s.solver_counters()
    s.close()
    np.testing.assert_array_equal(dev, host)
    np.testing.assert_array_equal(dev_c[1:], np.array(list(host_c))[1:])
