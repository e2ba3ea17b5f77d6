#!/usr/bin/env python3
"""Benchmark of the ODESystemSolver::forward hot path on one MI355X.

Workload (the headline configuration): the ten Tusscher-Panfilov 2006 M-cell
(19 states) advanced with GRL1 over 1,048,576 nodes at dt = 0.02 ms; per-node
g_Kr, g_Ks, g_CaL = default * exp(N(0, 0.2^2)) and initial states
default * (1 + U(-0.01, 0.01)), both seeded; default stimulus (-52 pA/pF,
1 ms, every 1000 ms).  A step is one ODESystemSolver::forward(t, dt) over all
nodes.  `--steps` steps are timed after `--warmup` untimed ones.

Reported: FP32C node-steps/s (FP32 arithmetic with compensated V, Ca_i,
Ca_SR, Ca_ss, Na_i, K_i) as `value`, the plain FP64 kernel beside it, and the
roofline of the dominant kernel.  The accuracy gate does not depend on
`--steps`: FP32C and FP64 advance all nodes over the metric's 1000 ms
(50,000 steps) and V is compared on every node every 1 ms; the FP64 states are
checked against the CPU oracle on a node subset.  Optional legs (adjoint,
implicit solvers, mixed batch, CPU baseline) record an error in the JSON line
instead of ending the run.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import platform
import sys
import time
import traceback

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("node-steps/s TP06 GRL1 1M nodes, FP32c vs FP64 @max|ΔV|≤0.05mV; adjoint/fwd")
VARIANT = "tp06_m_fp3"
FP_DEFAULTS = np.array([0.153, 0.098, 3.98e-5])
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# VALU issue: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 VALU instruction per SIMD every 4 cycles.
PEAK_VALU_WAVE_INSTS = 256 * 4 * 2.4e9 / 4
# The accuracy bound of the metric: max |dV| <= 0.05 mV over 1000 ms.
GATE_HORIZON_MS = 1000.0
GATE_SAMPLE_MS = 1.0
GATE_DV_MV = 0.05


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on an arbitrary stream (the system's own stream)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return ms.value

    def device_synchronize(self):
        assert self.hip.hipDeviceSynchronize() == 0


def inputs(N, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    from goss_amd.system import ModelVariant

    ic = ModelVariant(VARIANT).initial_conditions
    y0 = ic * (1.0 + rng.uniform(-0.01, 0.01, (N, ic.size)))
    fp = FP_DEFAULTS * np.exp(rng.normal(0.0, 0.2, (N, 3)))
    return y0, fp


def gate_schedule(horizon_ms, sample_ms, dt):
    """Outer steps per sample and number of samples of the accuracy gate.

    V is compared every `sample_ms` (or once, at the horizon, when the horizon
    is shorter than one sample interval), so there is always at least one
    sample; the steps must tile the horizon exactly."""
    total = horizon_ms / dt
    if total < 1 or abs(total - round(total)) > 1e-9 * max(1.0, total):
        raise ValueError(f"horizon {horizon_ms} ms is not a whole number of {dt} ms steps")
    total = int(round(total))
    per = max(1, min(total, int(round(sample_ms / dt))))
    if total % per:
        raise ValueError(f"{total} steps do not split into samples of {per} steps")
    return per, total // per


def flops_per_node_step():
    """State-dependent operations of one GRL1 node-step: the generated
    linearized_eval after CSE (each add, mul, div, exp, log, sqrt, pow is one
    operation) plus the GRL1 update of each state (a*h, exp, -1, b/a, *, +)."""
    with open(os.path.join(ROOT, "goss_amd", "csrc", "generated", "variants.json")) as f:
        meta = json.load(f)[VARIANT]
    c = meta["census"]["linearized_eval_state"]
    ops = sum(c[k] for k in ("add", "mul", "div", "exp", "log", "sqrt", "pow"))
    return ops + 6 * meta["num_states"], ops


def timed_run(prec, N, y0, fp, t_warm, t_timed, spl, ev):
    from goss_amd.system import DeviceSystem

    dev = DeviceSystem(VARIANT, "GRL1", N, prec)
    dev.set_steps_per_launch(spl)
    dev.set_states(y0)
    dev.set_field_parameters(fp)
    if len(t_warm) > 1:
        dev.forward_times(t_warm)
    dev.synchronize()
    ev.device_synchronize()
    stream = dev.stream()
    dev.set_kernel_timing(True)
    e0, e1 = ev.create(), ev.create()
    w0 = time.perf_counter()
    ev.record(e0, stream)
    dev.forward_times(t_timed)
    ev.record(e1, stream)
    dev.synchronize()
    ev.device_synchronize()
    wall = time.perf_counter() - w0
    call_ms = ev.elapsed_ms(e0, e1)
    # HIP events the library records on the same stream right before the first
    # and right after the last kernel launch of the call.
    kernel_ms, launches = dev.kernel_timing()
    dev.close()
    return wall, call_ms, kernel_ms, launches


def one_step_calls(N, y0, fp, dt, calls=100):
    """One outer step per forward call, as an operator-splitting caller advances the ODE nodes once per PDE
    step (dolfinutils.py:750-780): every call reads and writes all states, so the launch is bound by that
    HBM traffic (212 B per node in FP32C, 328 B in FP64) as much as by the arithmetic."""
    ev = HipEvents()
    out = {"workload": f"TP06 M-cell GRL1, {N} nodes, one step of {dt} ms per forward call, {calls} calls"}
    for prec, nbytes in (("fp32c", 212), ("fp64", 328)):
        t_warm = np.arange(11) * dt
        t_timed = t_warm[-1] + np.arange(calls + 1) * dt
        wall, call_ms, kernel_ms, launches = timed_run(prec, N, y0, fp, t_warm, t_timed, 1, ev)
        per_call_s = kernel_ms * 1e-3 / launches
        gbs = nbytes * N / per_call_s / 1e9
        out[prec] = {"us_per_call": round(per_call_s * 1e6, 2), "node_steps_per_s": N * calls / (call_ms * 1e-3),
                     "algorithmic_bytes_per_call": nbytes * N, "achieved_GBps": round(gbs, 1),
                     "hbm_frac": round(gbs / PEAK_HBM_GBS, 3)}
    return out


def solve_rows(N, y0, fp, dt, steps=250):
    """ODESystemSolver::solve (ODESystemSolver.cpp:81-99) on the same nodes: V of every node after each step
    (8 bytes per node and step) into pageable host memory (goss_hip_system_solve: the rows of a chunk of steps
    are copied after the chunk), into pinned host memory (goss_hip_system_solve_async: the copy of one chunk
    runs while the next chunk steps) and into device memory, beside forward_times, which writes no rows.
    Wall time of each call, from its start to the end of its last copy."""
    from goss_amd.system import DeviceSystem, pinned_array
    import torch

    t = np.arange(steps + 1) * dt
    nbytes = steps * N * 8
    dests = {"pageable": np.zeros((steps + 1, N)), "pinned": pinned_array((steps + 1, N)),
             "device": torch.zeros((steps + 1, N), dtype=torch.float64, device="cuda")}
    dests["pinned"][:] = 0.0
    out = {"workload": f"TP06 M-cell GRL1 FP32C, {N} nodes, {steps} steps of {dt} ms, V after every step "
                       f"({nbytes / 1e9:.2f} GB of rows)"}

    def once(kind):
        dev = DeviceSystem(VARIANT, "GRL1", N, "fp32c")
        dev.set_states(y0)
        dev.set_field_parameters(fp)
        dev.forward_times(t[:3])  # first launch of the kernel outside the timing
        dev.synchronize()
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        if kind == "forward_times":
            dev.forward_times(t)
        elif kind == "pageable":
            dev.solve(t, dests[kind])
        else:
            dev.solve_async(t, dests[kind])
        dev.synchronize()
        ms = (time.perf_counter() - w0) * 1e3
        dev.close()
        return ms

    for kind in ("forward_times", "pageable", "pinned", "device"):
        once(kind)  # the destination's pages touched, the copy paths set up
        ms = once(kind)
        out[f"{kind}_ms"] = round(ms, 2)
        if kind != "forward_times":
            out[f"{kind}_rows_GBps"] = round(nbytes / (ms * 1e-3) / 1e9, 2)
    out["pinned_over_forward_times"] = round(out["pinned_ms"] / out["forward_times_ms"], 3)
    out["pageable_over_forward_times"] = round(out["pageable_ms"] / out["forward_times_ms"], 3)
    return out


def accuracy_gate(y0, fp, dt, horizon_ms=GATE_HORIZON_MS, sample_ms=GATE_SAMPLE_MS, oracle_nodes=64):
    """FP32C against FP64 on every node over [0, horizon_ms]: V of all nodes is
    compared on the device every `sample_ms` (through the device field-state
    accessor, no host copy), the end states per state, and the FP64 end states
    of the first `oracle_nodes` nodes against the CPU oracle."""
    import torch

    from goss_amd.system import DeviceSystem

    N = y0.shape[0]
    per, n_samples = gate_schedule(horizon_ms, sample_ms, dt)
    t = np.arange(per * n_samples + 1) * dt
    sys32 = DeviceSystem(VARIANT, "GRL1", N, "fp32c")
    sys64 = DeviceSystem(VARIANT, "GRL1", N, "fp64")
    for s in (sys32, sys64):
        s.set_steps_per_launch(per)
        s.set_states(y0)
        s.set_field_parameters(fp)
    v32 = torch.empty(N, dtype=torch.float64, device="cuda")
    v64 = torch.empty(N, dtype=torch.float64, device="cuda")
    dv = torch.empty(n_samples, dtype=torch.float64, device="cuda")
    vmax = torch.empty(n_samples, dtype=torch.float64, device="cuda")
    w = time.perf_counter()
    for k in range(n_samples):
        seg = t[k * per:(k + 1) * per + 1]
        sys32.forward_times(seg)
        sys64.forward_times(seg)
        # Both accessors finish on their system's stream before returning.
        sys32.get_field_states_device(v32.data_ptr())
        sys64.get_field_states_device(v64.data_ptr())
        dv[k] = (v32 - v64).abs().max()
        vmax[k] = v64.max()
        torch.cuda.current_stream().synchronize()  # v32/v64 are overwritten next sample
    elapsed = time.perf_counter() - w
    dv_h = dv.cpu().numpy()
    vmax_h = vmax.cpu().numpy()
    a, b = sys32.states(), sys64.states()
    sys32.close()
    floor = 1e-6 * np.max(np.abs(b), axis=0, keepdims=True)
    rel = (np.abs(a - b) / np.maximum(np.abs(b), floor)).max(axis=0)
    worst = int(np.argmax(dv_h))
    out = {
        "nodes": N, "horizon_ms": horizon_ms, "sample_every_ms": sample_ms, "samples": n_samples,
        "max_abs_dV_mV": float(dv_h.max()), "t_of_max_ms": float((worst + 1) * per * dt),
        "bound_mV": GATE_DV_MV, "pass": bool(dv_h.max() <= GATE_DV_MV),
        "max_abs_dV_end_mV": float(np.abs(a[:, 17] - b[:, 17]).max()),
        "max_rel_err_per_state_end": [float(x) for x in rel],
        "upstroke_seen": bool(vmax_h.max() > 0.0), "seconds": round(elapsed, 2),
    }
    del a
    sys64.close()
    try:
        out["fp64_vs_oracle_max_rel_err"] = oracle_parity(y0, fp, t, b, oracle_nodes)
        out["fp64_vs_oracle_nodes"] = oracle_nodes
    except Exception as e:  # the oracle is a checker; its absence must not hide the figures above
        out["fp64_vs_oracle_error"] = f"{type(e).__name__}: {e}"
    return out


def oracle_parity(y0, fp, t_all, gpu_final, nodes=64):
    import oracle

    o = oracle.OracleSystem(VARIANT, "GRL1", nodes)
    o.set_num_threads(min(8, os.cpu_count() or 1))
    o.set_states(y0[:nodes])
    o.set_field_parameters(fp[:nodes])
    o.forward_times(t_all)
    b = o.states()
    a = gpu_final[:nodes]
    floor = 1e-6 * np.max(np.abs(b), axis=0, keepdims=True)
    return float((np.abs(a - b) / np.maximum(np.abs(b), floor)).max())


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_threads():
    """Threads for the CPU baseline: every core in this process's affinity
    mask, capped by OMP_NUM_THREADS when the host sets it (the CPU share of
    one GPU on a shared node)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = cores
    if env.isdigit() and int(env) > 0:
        threads = min(cores, int(env))
    return threads, cores


def cpu_baseline(N, y0, fp, dt, budget_s=12.0):
    """The reference ODESystemSolver (cpp/goss compiled from its sources,
    OpenMP over nodes) on this host's cores, on a window of the workload."""
    import oracle

    threads, cores = host_threads()
    if oracle.ref_available():
        sysc = oracle.RefSystem(VARIANT, "GRL1", N)
        kind = "reference"
    else:
        sysc = oracle.OracleSystem(VARIANT, "GRL1", N)
        kind = "port"
    sysc.set_num_threads(threads)
    sysc.set_states(y0)
    sysc.set_field_parameters(fp)
    # calibrate on 2 steps, then time a window of about budget_s seconds
    w = time.perf_counter()
    sysc.forward_times(np.arange(3) * dt)
    per_step = (time.perf_counter() - w) / 2
    k = int(max(2, min(1000, budget_s / max(per_step, 1e-6))))
    t = (np.arange(k + 1) + 2) * dt
    w = time.perf_counter()
    sysc.forward_times(t)
    el = time.perf_counter() - w
    return {"value": N * k / el, "unit": "node-steps/s", "cores": threads, "kind": kind,
            "host_cores_in_affinity_mask": cores, "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
            "cpu_model": cpu_model(),
            "sample": f"all {N:,} nodes x {k} GRL1 steps (FP64, dt {dt} ms) of the same workload, "
                      f"{el:.1f} s, OpenMP {threads} threads"}


def adjoint_case(N, steps, dt, sample_every, checkpoint_every, seed, ev):
    """One gradient evaluation of the adjoint workload; HIP-event times of the
    record pass (forward with checkpoints and samples), the backward pass
    (recompute + reverse sweep) and a plain forward over the same steps."""
    import torch

    from goss_amd.adjoint import voltage_trace
    from goss_amd.system import DeviceSystem, ModelVariant

    variant = "tp06_m_fp6"
    mv = ModelVariant(variant)
    rng = np.random.Generator(np.random.PCG64(seed))
    y0 = mv.initial_conditions[None, :] * (1.0 + rng.uniform(-0.01, 0.01, (N, mv.num_states)))
    g0 = np.array([mv.default_parameters[mv.parameter_names.index(p)] for p in mv.field_parameter_names])
    fp = g0[None, :] * np.exp(rng.normal(0.0, 0.2, (N, g0.size)))
    target_fp = fp * np.exp(rng.normal(0.0, 0.1, fp.shape))
    t = np.arange(steps + 1) * dt
    sys_ = DeviceSystem(variant, "GRL1", N, "fp32")
    sys_.set_steps_per_launch(250)
    with torch.no_grad():
        target = voltage_trace(sys_, torch.tensor(y0), torch.tensor(target_fp), t, sample_every, 200)
    # warm-up of both kernels on a short run
    yw = torch.tensor(y0, requires_grad=True)
    fw = torch.tensor(fp, requires_grad=True)
    vw = voltage_trace(sys_, yw, fw, t[:401], sample_every, checkpoint_every)
    (vw ** 2).mean().backward()
    torch.cuda.synchronize()

    stream = sys_.stream()
    marks = [ev.create() for _ in range(4)]
    yt = torch.tensor(y0, requires_grad=True)
    ft = torch.tensor(fp, requires_grad=True)
    ev.record(marks[0], stream)
    v = voltage_trace(sys_, yt, ft, t, sample_every, checkpoint_every)
    ev.record(marks[1], stream)
    loss = ((v - target) ** 2).mean()
    torch.cuda.synchronize()
    ev.record(marks[2], stream)
    loss.backward()
    ev.record(marks[3], stream)
    sys_.synchronize()
    rec_ms = ev.elapsed_ms(marks[0], marks[1])
    bwd_ms = ev.elapsed_ms(marks[2], marks[3])
    gy, gf = yt.grad.numpy(), ft.grad.numpy()
    loss_value = float(loss)
    ckpt_bytes = sys_.adjoint_buffer_sizes(len(t), sample_every, checkpoint_every)[0]
    del v, loss
    sys_.set_states(y0)
    sys_.set_field_parameters(fp)
    torch.cuda.synchronize()
    plain = [ev.create() for _ in range(2)]
    ev.record(plain[0], stream)
    sys_.forward_times(t)
    ev.record(plain[1], stream)
    sys_.synchronize()
    plain_ms = ev.elapsed_ms(plain[0], plain[1])
    sys_.close()
    return {
        "checkpoint_every": checkpoint_every, "record_ms": rec_ms, "backward_ms": bwd_ms,
        "plain_forward_ms": plain_ms,
        # adjoint cost: the backward pass (recompute of each segment + reverse sweep) over a plain forward
        "adjoint_over_forward": bwd_ms / plain_ms,
        # one gradient evaluation: record pass + backward pass, over a plain forward
        "gradient_over_forward": (rec_ms + bwd_ms) / plain_ms,
        "loss": loss_value, "grads_finite": bool(np.all(np.isfinite(gy)) and np.all(np.isfinite(gf))),
        "checkpoint_bytes": ckpt_bytes,
    }


def adjoint_run(N=65536, steps=20000, dt=0.02, sample_every=50, seed=2025, every_state=True):
    """Adjoint workload: TP06 M-cell GRL1 in FP32 with six per-node conductances
    (g_Kr, g_Ks, g_CaL, g_Na, g_K1, g_to) and perturbed initial states; loss =
    MSE of V sampled every 1 ms against a target trace from the conductances x
    lognormal(0, 0.1); gradients with respect to the conductances and the 19
    initial states.  The headline figure uses checkpoints every 200 steps (the
    workload's stated interval); the every-state mode (checkpoint interval 1,
    no recomputation) is reported beside it when its states fit in half of the
    free device memory."""
    import torch

    ev = HipEvents()
    main = adjoint_case(N, steps, dt, sample_every, 200, seed, ev)
    out = {
        "workload": f"TP06 M-cell GRL1 FP32, {N} nodes x {steps} steps (dt {dt} ms), V sampled every "
                    f"{sample_every} steps, checkpoints every 200 steps, grads wrt 6 conductances + 19 initial "
                    f"states",
        "adjoint_over_forward_definition": "backward pass (segment recompute + reverse sweep) / plain forward; "
                                           "gradient_over_forward adds the record pass",
        **main,
    }
    if every_state:
        free, _ = torch.cuda.mem_get_info()
        if steps * N * 19 * 4 <= free // 2:
            e1 = adjoint_case(N, steps, dt, sample_every, 1, seed, ev)
            out["every_state"] = {k: e1[k] for k in ("record_ms", "backward_ms", "plain_forward_ms",
                                                     "adjoint_over_forward", "gradient_over_forward",
                                                     "checkpoint_bytes", "grads_finite")}
    return out


def implicit_run(N=262144, steps=10000, dt=0.1, seed=2027, fd64_steps=500):
    """Stiff workload: TP06 epi-cell, ImplicitEuler (relative_tolerance 1e-6)
    and ESDIRK23a (its fixed tol 1e-5), N nodes, dt 0.1 ms, initial states x
    (1 + U(-0.01, 0.01)), with a -52 pA/pF, 1 ms stimulus at t = 10 ms (the
    .ode's default amplitude is 0).  The Jacobian is formed from forward
    differences of the FP32 right-hand side, with the Newton iteration, LU and
    substitutions in FP64, over the full `steps`; the reference's FP64
    forward-difference Jacobian and the model's analytic Jacobian (a plugin
    generated with generate_jacobian) run over the first `fd64_steps` steps
    beside it.  Wall time around the call and a synchronize, after a warm-up
    step."""
    from goss_amd.system import DeviceSystem, ModelVariant

    mv = ModelVariant("tp06_epi")
    rng = np.random.Generator(np.random.PCG64(seed))
    y0 = mv.initial_conditions[None, :] * (1.0 + rng.uniform(-0.01, 0.01, (N, mv.num_states)))
    out = {"workload": f"TP06 epi, stimulus -52 pA/pF 1 ms at 10 ms, {N} nodes, dt {dt} ms"}
    for solver in ("ImplicitEuler", "ESDIRK23a"):
        base = {"relative_tolerance": 1e-6} if solver == "ImplicitEuler" else {}
        for jac, n_steps in (("fd_fp32", steps), ("fd", fd64_steps), ("analytic", fd64_steps)):
            t = np.arange(n_steps + 1) * dt
            s = DeviceSystem("tp06_epi", solver, N, "fp64", {**base, "jacobian": jac})
            s.set_parameter("stim_amplitude", 52.0)
            s.set_steps_per_launch(50)
            s.set_states(y0)
            s.forward_times(t[:2])  # warm-up
            s.synchronize()
            s.close()
            s = DeviceSystem("tp06_epi", solver, N, "fp64", {**base, "jacobian": jac})
            s.set_parameter("stim_amplitude", 52.0)
            s.set_steps_per_launch(50)
            s.set_states(y0)
            s.synchronize()
            w = time.perf_counter()
            s.forward_times(t)
            s.synchronize()
            el = time.perf_counter() - w
            fin = s.states()
            out[f"{solver}_{jac}"] = {"steps": n_steps, "s": el, "node_steps_per_s": N * n_steps / el,
                                      "counters": [int(c) for c in s.solver_counters()],
                                      "V_range_end_mV": [float(fin[:, 17].min()), float(fin[:, 17].max())]}
            s.close()
    return out


MIXED_POPS = [("tp06_m", "GRL1", 300017, 0.02), ("tp06_epi", "GRL1", 211003, 0.02),
              ("fitzhughnagumo", "RK4", 700001, 0.05), ("lorentz", "RKF32", 9973, 0.02)]


def mixed_run(sim_ms=100.0, tune_ms=1.0, seed=2026):
    """Mixed persistent batch: TP06 M 300,017 + TP06 epi 211,003 (GRL1, dt 0.02 ms) + FitzHugh-Nagumo
    700,001 (RK4, dt 0.05) + Lorenz 9,973 (RKF32, outer interval 0.02), FP32, initial states
    x (1 + U(-0.01, 0.01)).  The batch's launch shape is autotuned on [0, tune_ms]; then [tune_ms,
    tune_ms + sim_ms] is timed (wall clock around the call and a synchronize).  The same populations
    advanced one after another, each with its own kernel, are timed beside it."""
    from goss_amd.batch import DeviceBatch
    from goss_amd.system import DeviceSystem, ModelVariant

    rng = np.random.Generator(np.random.PCG64(seed))
    y0 = []
    for v, _, n, _ in MIXED_POPS:
        mv = ModelVariant(v)
        y0.append(mv.initial_conditions[None, :] * (1.0 + rng.uniform(-0.01, 0.01, (n, mv.num_states))))
    dts = [dt for *_, dt in MIXED_POPS]

    def make():
        out = []
        for (v, solver, n, _), y in zip(MIXED_POPS, y0):
            s = DeviceSystem(v, solver, n, "fp32")
            s.set_states(y)
            out.append(s)
        return out

    systems = make()
    batch = DeviceBatch(systems)
    cfg = batch.autotune(0.0, tune_ms, dts)
    batch.forward(0.0, tune_ms, dts)  # warm-up, also captures the graph shape once
    batch.synchronize()
    t0, t1 = tune_ms, tune_ms + sim_ms
    w = time.perf_counter()
    batch.forward(t0, t1, dts)
    batch.synchronize()
    fused = time.perf_counter() - w
    node_steps = sum(n * int(round(sim_ms / dt)) for (_, _, n, dt) in MIXED_POPS)

    def times(a, b, dt):  # the batch's times: t_k = a + k dt, ending exactly at b
        t = a + np.arange(int(round((b - a) / dt)) + 1) * dt
        t[-1] = b
        return t

    alone = make()
    for s, dt in zip(alone, dts):
        s.set_steps_per_launch(250)
        s.forward_times(times(0.0, tune_ms, dt))
        s.synchronize()
    w = time.perf_counter()
    for s, dt in zip(alone, dts):
        s.forward_times(times(t0, t1, dt))
    for s in alone:
        s.synchronize()
    separate = time.perf_counter() - w
    same = all(np.array_equal(a.states(), b.states()) for a, b in zip(systems, alone))
    batch.close()
    for s in systems + alone:
        s.close()
    return {
        "workload": "TP06 M 300,017 + TP06 epi 211,003 (GRL1, dt 0.02 ms) + FitzHugh-Nagumo 700,001 (RK4, dt 0.05) "
                    f"+ Lorenz 9,973 (RKF32, interval 0.02), FP32, {sim_ms:g} time units",
        "node_steps": node_steps, "fused_s": fused, "node_steps_per_s": node_steps / fused,
        "separate_kernels_s": separate, "fused_speedup": separate / fused,
        "autotuned": {"block_size": cfg.block_size, "nodes_per_lane": cfg.nodes_per_lane, "span": cfg.span,
                      "dispatch_order": [MIXED_POPS[p][0] for p in cfg.order]},
        "states_equal_to_separate": bool(same),
    }


def load_traffic():
    p = os.path.join(ROOT, "profiles", "grl1_tp06_fp32c_1M_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def leg(name, fn, *a, **kw):
    """Runs an optional leg; a failure is recorded instead of ending the run."""
    w = time.perf_counter()
    try:
        r = fn(*a, **kw)
    except Exception as e:
        log(f"{name} failed: {type(e).__name__}: {e}")
        traceback.print_exc(file=sys.stderr)
        return {"error": f"{type(e).__name__}: {e}"}
    log(f"{name} ({time.perf_counter() - w:.1f} s) {json.dumps(r)}")
    return r


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--nodes", type=int, default=1 << 20)
    ap.add_argument("--dt", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--steps-per-launch", type=int, default=250)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-checks", action="store_true", help="skip the accuracy gate and the oracle parity")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint workload")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the adjoint workload and print its figures as one JSON line")
    ap.add_argument("--adjoint-steps", type=int, default=20000, help="steps of the adjoint workload")
    ap.add_argument("--adjoint-checkpoint-every", type=int, default=None,
                    help="with --adjoint-only: one checkpoint interval instead of 200 and 1")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-population batch workload")
    ap.add_argument("--no-implicit", action="store_true", help="skip the implicit-solver workload")
    ap.add_argument("--implicit-steps", type=int, default=10000, help="steps of the implicit workload")
    ap.add_argument("--implicit-only", action="store_true",
                    help="only run the implicit-solver workload and print its figures as one JSON line")
    ap.add_argument("--solve-only", action="store_true",
                    help="only run the solve workload (rows into pageable, pinned and device memory) and print "
                         "its figures as one JSON line")
    ap.add_argument("--mixed-only", action="store_true",
                    help="only run the mixed-population batch workload and print its figures as one JSON line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run one device-to-device copy of 256 MiB (read 256 MiB + write 256 MiB), the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE readings")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the ODE system path shards by node ranges with no "
              "communication, and this build measures the single-GPU share (--gpus 1).", file=sys.stderr)
        return 2

    if args.calibrate_copy:
        import torch

        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4}))
        return 0
    if args.adjoint_only:
        if args.adjoint_checkpoint_every:
            r = adjoint_case(65536, args.adjoint_steps, 0.02, 50, args.adjoint_checkpoint_every, 2025, HipEvents())
        else:
            r = adjoint_run(steps=args.adjoint_steps)
        print(json.dumps(r), flush=True)
        return 0
    if args.mixed_only:
        print(json.dumps(mixed_run()), flush=True)
        return 0
    if args.implicit_only:
        print(json.dumps(implicit_run(steps=args.implicit_steps)), flush=True)
        return 0

    N, K, W, dt = args.nodes, args.steps, args.warmup, args.dt
    if K < 1 or W < 0:
        raise SystemExit("--steps must be >= 1 and --warmup >= 0")
    t_all = np.arange(W + K + 1) * dt
    t_warm, t_timed = t_all[: W + 1], t_all[W:]
    y0, fp = inputs(N, args.seed)
    if args.solve_only:
        print(json.dumps(solve_rows(N, y0, fp, dt)), flush=True)
        return 0
    ev = HipEvents()
    log(f"{N} nodes, warmup {W}, steps {K}, dt {dt}")

    # The accuracy gate runs first: besides checking the metric's bound it brings the GPU to its
    # working clocks, so that a short timed region (a few steps) is not measured from idle.
    accuracy = None if args.no_checks else leg("accuracy", accuracy_gate, y0, fp, dt)
    res = {}
    for prec in ("fp32c", "fp64"):
        wall, call_ms, kms, launches = timed_run(prec, N, y0, fp, t_warm, t_timed, args.steps_per_launch, ev)
        assert launches == math.ceil(K / args.steps_per_launch)
        res[prec] = {"wall": wall, "call_ms": call_ms, "kernel_ms": kms, "launches": launches, "value": N * K / wall}
        log(f"{prec}: {wall:.3f} s wall, {call_ms:.3f} ms call events, {kms:.3f} ms kernel events over {launches} "
            f"launches, {N * K / wall:.4e} node-steps/s")

    flops, rhs_ops = flops_per_node_step()
    spl = min(args.steps_per_launch, K)
    avg_launch_s = res["fp32c"]["kernel_ms"] * 1e-3 / res["fp32c"]["launches"]
    achieved = flops * N * spl / avg_launch_s / 1e12
    traffic = load_traffic()
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": traffic.get("bytes_per_launch") if traffic else None,
                "kernel": "goss_hip::explicit_kernel<tp06_m_fp3, GRL1, float, COMP=true>",
                "flops_per_node_step": flops, "node_steps_per_launch": N * spl,
                "avg_launch_ms": round(avg_launch_s * 1e3, 4), "launches": res["fp32c"]["launches"],
                "algorithmic_bytes_per_launch": 212 * N}
    valu = (traffic or {}).get("valu", {}).get("f32c")
    if valu and valu.get("insts_per_node_step"):
        # Wave-level VALU instructions per launch (rocprofv3 SQ_INSTS_VALU per node-step, scaled to
        # this launch) over the live launch time, against the issue peak.
        insts = valu["insts_per_node_step"] * N * spl / 64
        roofline["valu_issue"] = {"wave_insts_per_launch": insts, "insts_per_node_step": valu["insts_per_node_step"],
                                  "achieved": insts / avg_launch_s, "peak": PEAK_VALU_WAVE_INSTS,
                                  "frac": round(insts / avg_launch_s / PEAK_VALU_WAVE_INSTS, 4)}
    avg64 = res["fp64"]["kernel_ms"] * 1e-3 / res["fp64"]["launches"]
    fp64_achieved = flops * N * spl / avg64 / 1e12

    out = {
        "metric": METRIC,
        "value": res["fp32c"]["value"],
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": res["fp32c"]["wall"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "TP06 M-cell GRL1 ODESystemSolver::forward, 1,048,576 nodes, dt 0.02 ms, "
                               "lognormal(0,0.2) g_Kr/g_Ks/g_CaL, stimulus -52 pA/pF every 1000 ms",
                   "nodes": N, "dt_ms": dt, "solver": "GRL1", "precision": "fp32c",
                   "steps_per_launch": args.steps_per_launch},
        "fp64": {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["wall"] / K * 1e3,
                 "achieved_tflops": round(fp64_achieved, 3), "peak": PEAK_FP64_TFLOPS,
                 "frac": round(fp64_achieved / PEAK_FP64_TFLOPS, 4)},
        "fp32c_over_fp64": res["fp32c"]["value"] / res["fp64"]["value"],
        "adjoint_over_forward": None,
        "roofline": roofline,
    }
    if accuracy is not None:
        out["accuracy"] = accuracy
    out["one_step_calls"] = leg("one_step_calls", one_step_calls, N, y0, fp, dt)
    out["solve"] = leg("solve", solve_rows, N, y0, fp, dt)
    if not args.no_adjoint:
        adj = leg("adjoint", adjoint_run, steps=args.adjoint_steps)
        out["adjoint"] = adj
        out["adjoint_over_forward"] = adj.get("adjoint_over_forward")
    if not args.no_implicit:
        out["implicit"] = leg("implicit", implicit_run, steps=args.implicit_steps)
    if not args.no_mixed:
        out["mixed"] = leg("mixed", mixed_run)
    if not args.no_cpu:
        out["cpu_baseline"] = leg("cpu", cpu_baseline, N, y0, fp, dt)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
