"""solve_async (goss_hip_system_solve_async): ODESystemSolver::solve
(cpp/goss/ODESystemSolver.cpp:81-99) into pinned host or device memory,
returning before the rows are there.

The rows must equal, bit for bit, those of the synchronous solve into pageable
host memory (the same kernels over the same steps; only where the rows travel
differs), for destinations made by the library (pinned_array), by torch
(pin_memory, a CUDA tensor), for tangled and untangled storage, and with
enough rows that the pinned path's two staging buffers are each reused
(chunks of 64 MiB of rows).
"""
import numpy as np
import pytest
import torch

from test_gpu_explicit import dev_system, perturbed

pytestmark = pytest.mark.gpu


def _pageable_rows(variant, solver, N, precision, y0, t, tangled):
    s = dev_system(variant, solver, N, precision)
    s.set_states(y0)
    nfs = s.model.num_field_states
    out = np.zeros((len(t), N * nfs))
    s.solve(t, out, tangled=tangled)
    return out, s.states()


@pytest.mark.parametrize("variant,solver,precision,N,steps", [
    ("tp06_m", "GRL1", "fp32", 1 << 20, 40),      # 8 rows per 64 MiB chunk: 5 chunks
    ("tp06_m", "GRL1", "fp64", 65536, 300),       # 128 rows per chunk: 3 chunks
    ("oscilator", "RK4", "fp64", 300000, 60),     # two field states: 13 rows per chunk
    ("beeler_reuter_1977", "ImplicitEuler", "fp64", 32768, 500),
])
def test_solve_async_equals_solve(variant, solver, precision, N, steps):
    from goss_amd.system import pinned_array

    dt = 0.05 if variant != "oscilator" else 0.01
    t = np.arange(steps + 1) * dt
    probe = dev_system(variant, solver, N, precision)
    y0 = perturbed(probe.states(), 11)
    nfs = probe.model.num_field_states
    del probe
    for tangled in ((True, False) if nfs > 1 else (True,)):
        print(f"{variant} {solver} {precision} N={N} steps={steps} tangled={tangled}", flush=True)
        ref, ref_states = _pageable_rows(variant, solver, N, precision, y0, t, tangled)
        assert np.any(ref[1:] != 0)
        dests = {
            "pinned_array": pinned_array((len(t), N * nfs)),
            "torch_pinned": torch.zeros((len(t), N * nfs), dtype=torch.float64).pin_memory(),
            "cuda": torch.zeros((len(t), N * nfs), dtype=torch.float64, device="cuda"),
        }
        dests["pinned_array"][:] = 0.0
        for name, out in dests.items():
            s = dev_system(variant, solver, N, precision)
            s.set_states(y0)
            torch.cuda.synchronize()
            s.solve_async(t, out, tangled=tangled)
            s.synchronize()
            got = out.cpu().numpy() if isinstance(out, torch.Tensor) else out
            assert np.array_equal(got[1:], ref[1:]), (name, tangled)
            assert np.all(got[0] == 0), name  # row 0 is the caller's
            assert np.array_equal(s.states(), ref_states), name


def test_solve_into_pinned_memory_takes_the_pipelined_path():
    """goss_hip_system_solve (the synchronous entry point) with a pinned destination: the same rows, and the
    call has waited (the rows are there when it returns)."""
    from goss_amd.system import pinned_array

    N, t = 1 << 20, np.arange(41) * 0.05
    probe = dev_system("tp06_m", "GRL1", N, "fp32")
    y0 = perturbed(probe.states(), 12)
    del probe
    ref, _ = _pageable_rows("tp06_m", "GRL1", N, "fp32", y0, t, True)
    out = pinned_array((len(t), N))
    out[:] = 0.0
    s = dev_system("tp06_m", "GRL1", N, "fp32")
    s.set_states(y0)
    s.solve(t, out)
    assert np.array_equal(out[1:], ref[1:])


def test_two_async_solves_in_a_row_then_forward():
    """Calls on one system are ordered: a second solve_async into another buffer, then forward_times, read
    after one synchronize, equal the same calls made synchronously."""
    from goss_amd.system import pinned_array

    N = 1 << 19
    t1, t2 = np.arange(31) * 0.05, 1.5 + np.arange(21) * 0.05
    t3 = 2.5 + np.arange(11) * 0.05
    probe = dev_system("tp06_m", "GRL1", N, "fp32")
    y0 = perturbed(probe.states(), 13)
    del probe
    a = dev_system("tp06_m", "GRL1", N, "fp32")
    a.set_states(y0)
    r1, r2 = np.zeros((len(t1), N)), np.zeros((len(t2), N))
    a.solve(t1, r1)
    a.solve(t2, r2)
    a.forward_times(t3)
    b = dev_system("tp06_m", "GRL1", N, "fp32")
    b.set_states(y0)
    p1, p2 = pinned_array((len(t1), N)), pinned_array((len(t2), N))
    b.solve_async(t1, p1)
    b.solve_async(t2, p2)
    b.forward_times(t3)
    b.synchronize()
    assert np.array_equal(p1[1:], r1[1:]) and np.array_equal(p2[1:], r2[1:])
    assert np.array_equal(b.states(), a.states())


def test_solve_async_refuses_pageable_memory():
    from goss_amd._lib import GossHipError

    s = dev_system("tp06_m", "GRL1", 64, "fp64")
    with pytest.raises(GossHipError, match="pinned host memory"):
        s.solve_async(np.arange(3) * 0.1, np.zeros((3, 64)))
    # the system still works afterwards
    out = np.zeros((3, 64))
    s.solve(np.arange(3) * 0.1, out)
    assert np.all(out[1:] != 0)
