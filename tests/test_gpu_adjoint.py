"""Discrete adjoint of the explicit steps on the device (goss_amd.adjoint, goss_adjoint.h).

References:
* plain-PyTorch FP64 autograd through the same scheme (tests/torch_reference.py,
  run on the CPU): the FP64 device gradients must agree to 1e-8 relative per entry
  (floor: 1e-9 of the entry's column maximum);
* central finite differences of the device FP64 loss;
* the FP64 device adjoint itself for FP32 (stated bound) and for other checkpoint
  intervals (bit-identical: the backward sweep recomputes the forward states
  with the same code).
"""
import numpy as np
import pytest
import torch

from test_gpu_explicit import dev_system

pytestmark = pytest.mark.gpu

VARIANT = "tp06_m_fp6"


def _inputs(N, seed, spread=0.01):
    from goss_amd.system import ModelVariant

    mv = ModelVariant(VARIANT)
    rng = np.random.Generator(np.random.PCG64(seed))
    y0 = mv.initial_conditions[None, :] * (1.0 + rng.uniform(-spread, spread, (N, mv.num_states)))
    g0 = np.array([mv.default_parameters[mv.parameter_names.index(p)] for p in mv.field_parameter_names])
    fp = g0[None, :] * np.exp(rng.normal(0.0, 0.1, (N, g0.size)))
    return y0, fp


def _times(nsteps, dt=0.02, t0=9.0):
    # t0 = 9 ms: the run crosses the stimulus (10-11 ms) and the upstroke.
    return t0 + np.arange(nsteps + 1) * dt


def _device_grads(prec, y0, fp, t, sample_every, ckpt_every, weights, solver="GRL1"):
    from goss_amd.adjoint import voltage_trace

    sys_ = dev_system(VARIANT, solver, y0.shape[0], prec)
    yt = torch.tensor(y0, requires_grad=True)
    ft = torch.tensor(fp, requires_grad=True)
    v = voltage_trace(sys_, yt, ft, t, sample_every, ckpt_every)
    loss = (v * torch.as_tensor(weights, device=v.device)).sum()
    loss.backward()
    return float(loss), v.detach().cpu().numpy(), yt.grad.numpy(), ft.grad.numpy()


def _rel(a, b):
    floor = 1e-9 * np.max(np.abs(b), axis=0, keepdims=True)
    return float((np.abs(a - b) / np.maximum(np.abs(b), np.maximum(floor, 1e-300))).max())


def test_adjoint_fp64_matches_torch_autograd():
    from torch_reference import TorchModel, sampled_trace

    N, nsteps, se, ce = 6, 300, 25, 40
    y0, fp = _inputs(N, 61)
    t = _times(nsteps)
    rng = np.random.Generator(np.random.PCG64(62))
    w = rng.normal(size=(nsteps // se, N))
    loss, v, gy, gp = _device_grads("fp64", y0, fp, t, se, ce, w)

    tm = TorchModel(VARIANT)
    yt = torch.tensor(y0, requires_grad=True)
    ft = torch.tensor(fp, requires_grad=True)
    vr = sampled_trace(tm, yt, ft, t, se, 17)
    (vr * torch.tensor(w)).sum().backward()
    print(f"V range {v.min():.2f} .. {v.max():.2f} mV; samples rel err {_rel(v, vr.detach().numpy()):.2e}; "
          f"grad y0 {_rel(gy, yt.grad.numpy()):.2e}; grad g {_rel(gp, ft.grad.numpy()):.2e}")
    assert v.max() > 0.0  # the run includes the upstroke
    assert _rel(v, vr.detach().numpy()) <= 1e-10
    assert _rel(gy, yt.grad.numpy()) <= 1e-8
    assert _rel(gp, ft.grad.numpy()) <= 1e-8


def test_adjoint_fp64_matches_central_differences():
    from goss_amd.adjoint import voltage_trace

    N, nsteps, se, ce = 2, 200, 20, 50
    y0, fp = _inputs(N, 63)
    t = _times(nsteps)
    rng = np.random.Generator(np.random.PCG64(64))
    w = rng.normal(size=(nsteps // se, N))
    _, _, gy, gp = _device_grads("fp64", y0, fp, t, se, ce, w)
    sys_ = dev_system(VARIANT, "GRL1", N, "fp64")

    def loss(y, f):
        with torch.no_grad():
            v = voltage_trace(sys_, torch.tensor(y), torch.tensor(f), t, se, ce)
        return (v.cpu().numpy() * w).sum(axis=0)  # per node (nodes are independent)

    for j in range(fp.shape[1]):
        e = 1e-5 * fp[:, j]
        fpp, fpm = fp.copy(), fp.copy()
        fpp[:, j] += e
        fpm[:, j] -= e
        fd = (loss(y0, fpp) - loss(y0, fpm)) / (2 * e)
        np.testing.assert_allclose(gp[:, j], fd, rtol=2e-4, atol=1e-9 * np.abs(gp).max())
    # V, Ca_i, h.  (Fast gates such as m forget their initial value within a few steps:
    # d loss / d m0 is at the rounding level of the loss and has no usable difference quotient.)
    for i in (17, 12, 4):
        e = 1e-4 * np.abs(y0[:, i])
        yp, ym = y0.copy(), y0.copy()
        yp[:, i] += e
        ym[:, i] -= e
        fd = (loss(yp, fp) - loss(ym, fp)) / (2 * e)
        np.testing.assert_allclose(gy[:, i], fd, rtol=2e-4, atol=1e-7 * np.abs(gy[:, i]).max())


def test_checkpoint_interval_does_not_change_gradients():
    N, nsteps, se = 64, 240, 30
    y0, fp = _inputs(N, 65)
    t = _times(nsteps)
    w = np.random.Generator(np.random.PCG64(66)).normal(size=(nsteps // se, N))
    ref = _device_grads("fp64", y0, fp, t, se, 240, w)
    for ce in (1, 37, 1000):
        got = _device_grads("fp64", y0, fp, t, se, ce, w)
        assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3]), ce


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_record_samples_equal_forward(prec):
    from goss_amd.adjoint import voltage_trace

    N, nsteps, se = 100, 150, 10
    y0, fp = _inputs(N, 67)
    t = _times(nsteps)
    a = dev_system(VARIANT, "GRL1", N, prec)
    with torch.no_grad():
        v = voltage_trace(a, torch.tensor(y0), torch.tensor(fp), t, se, 50).cpu().numpy()
    b = dev_system(VARIANT, "GRL1", N, prec)
    b.set_states(y0)
    b.set_field_parameters(fp)
    rows = np.zeros((len(t), N, 1))
    b.solve(t, rows)
    if prec == "fp64":
        # FP64 is compiled without contraction: the record kernel's step is the forward kernel's, bit for bit.
        assert np.array_equal(v, rows[se::se, :, 0])
        assert np.array_equal(a.states(), b.states())
    else:
        # FP32 allows FMA contraction, which the compiler may form differently in the two kernels
        # (measured: 2 of 100 nodes differ, by 1.2e-4 mV; FP32 vs FP64 is 3.8e-4 mV here).
        assert np.abs(v - rows[se::se, :, 0]).max() <= 2e-3
        sa, sb = a.states(), b.states()
        assert (np.abs(sa - sb) <= 1e-4 * np.abs(sb) + 1e-12).all()


def test_adjoint_fp32_close_to_fp64():
    # The calibration workload's shape on a subset: 20 ms through the stimulus and upstroke,
    # MSE of V sampled every 1 ms against a target from conductances x lognormal(0, 0.1).
    N, nsteps, se, ce = 256, 1000, 50, 200
    y0, fp = _inputs(N, 68)
    t = _times(nsteps)
    rng = np.random.Generator(np.random.PCG64(69))
    target_fp = fp * np.exp(rng.normal(0.0, 0.1, fp.shape))
    from goss_amd.adjoint import voltage_trace

    with torch.no_grad():
        tgt = voltage_trace(dev_system(VARIANT, "GRL1", N, "fp64"), torch.tensor(y0), torch.tensor(target_fp), t,
                            se, ce).cpu().numpy()
    grads = {}
    for prec in ("fp64", "fp32"):
        sys_ = dev_system(VARIANT, "GRL1", N, prec)
        yt = torch.tensor(y0, requires_grad=True)
        ft = torch.tensor(fp, requires_grad=True)
        v = voltage_trace(sys_, yt, ft, t, se, ce)
        loss = ((v - torch.as_tensor(tgt, device=v.device)) ** 2).mean()
        loss.backward()
        grads[prec] = (yt.grad.numpy(), ft.grad.numpy())
    for k, name in ((0, "y0"), (1, "conductances")):
        a, b = grads["fp32"][k], grads["fp64"][k]
        rel_l2 = np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(b, axis=0), 1e-300)
        print(f"FP32 vs FP64 adjoint, d loss / d {name}: relative L2 error per column, max {rel_l2.max():.2e}")
        assert np.all(np.isfinite(a))
        # States whose gradient is exactly zero in FP64 (none influence V here) are skipped.
        cols = np.linalg.norm(b, axis=0) > 0
        assert rel_l2[cols].max() <= 5e-3  # measured 3.9e-4


def test_calibration_workload_shape_65536_nodes():
    """The calibration workload's shape on 2,000 of its 20,000 steps: TP06 M-cell GRL1 FP32,
    65,536 nodes, dt 0.02 ms from t = 0 (stimulus at 10 ms), V sampled every 1 ms, MSE against a
    target from conductances x lognormal(0, 0.1), checkpoints every 200 steps (10 segments, so the
    pipelined backward alternates its two segment buffers).  Checks:
    * interval 200 and interval 1 (every state stored, no recompute) give the same gradients bit
      for bit at all 65,536 nodes;
    * on 64 of the nodes, the FP64 adjoint of the same per-node losses agrees with central
      differences of the FP64 loss (6 conductances, V, Ca_i), and the FP32 gradients agree with the FP64 ones.  Nodes are independent, so a 64-node system gives
      those nodes' gradients."""
    from goss_amd.adjoint import voltage_trace

    N, nsteps, se, dt = 65536, 2000, 50, 0.02
    y0, fp = _inputs(N, 71)
    t = np.arange(nsteps + 1) * dt
    rng = np.random.Generator(np.random.PCG64(72))
    target_fp = fp * np.exp(rng.normal(0.0, 0.1, fp.shape))
    with torch.no_grad():
        tgt = voltage_trace(dev_system(VARIANT, "GRL1", N, "fp32"), torch.tensor(y0), torch.tensor(target_fp), t,
                            se, 200).cpu().numpy()
    tgt_t = torch.as_tensor(tgt)

    def grads(prec, y, f, ce, target):
        sys_ = dev_system(VARIANT, "GRL1", y.shape[0], prec)
        yt = torch.tensor(y, requires_grad=True)
        ft = torch.tensor(f, requires_grad=True)
        v = voltage_trace(sys_, yt, ft, t, se, ce)
        # sum over nodes of the per-node MSE over samples: a node's gradient does not depend on N
        loss = ((v - target.to(v.device)) ** 2).mean(dim=0).sum()
        loss.backward()
        return v.detach().cpu().numpy(), yt.grad.numpy(), ft.grad.numpy()

    v200, gy200, gp200 = grads("fp32", y0, fp, 200, tgt_t)
    assert v200.max() > 0.0 and np.all(np.isfinite(gy200)) and np.all(np.isfinite(gp200))
    _, gy1, gp1 = grads("fp32", y0, fp, 1, tgt_t)
    assert np.array_equal(gy200, gy1) and np.array_equal(gp200, gp1)

    n = 64
    sub = np.arange(0, N, N // n)  # spread over the node range
    tsub = tgt_t[:, sub]
    _, gy64, gp64 = grads("fp64", y0[sub], fp[sub], 200, tsub)
    fd_sys = dev_system(VARIANT, "GRL1", n, "fp64")

    def loss(y, f):
        with torch.no_grad():
            v = voltage_trace(fd_sys, torch.tensor(y), torch.tensor(f), t, se, 200).cpu()
        return ((v - tsub) ** 2).mean(dim=0).numpy()  # per node

    # Central differences with a relative perturbation of 1e-6: at 1e-5 some nodes' perturbed runs move
    # a step across the -40 mV switch of the gate rates, where the discrete loss jumps (measured: 1.3e-2
    # of the column's largest gradient for g_Kr); at 1e-7 the quotient is rounding-limited (7e-4).  At
    # 1e-6 the largest deviation is 7.8e-5 of the column's largest gradient.
    def check(adj, perturb, what):
        fd = perturb(1e-6)
        scale = np.abs(adj).max()
        print(f"{what}: max |adjoint - central difference| {np.abs(adj - fd).max() / scale:.2e} of the largest")
        np.testing.assert_allclose(adj, fd, rtol=1e-3, atol=1e-3 * scale)

    for j in range(fp.shape[1]):
        def fd_p(rel, j=j):
            e = rel * fp[sub, j]
            fpp, fpm = fp[sub].copy(), fp[sub].copy()
            fpp[:, j] += e
            fpm[:, j] -= e
            return (loss(y0[sub], fpp) - loss(y0[sub], fpm)) / (2 * e)
        check(gp64[:, j], fd_p, f"field parameter {j}")
    for i in (17, 12):
        def fd_y(rel, i=i):
            e = rel * np.abs(y0[sub, i])
            yp, ym = y0[sub].copy(), y0[sub].copy()
            yp[:, i] += e
            ym[:, i] -= e
            return (loss(yp, fp[sub]) - loss(ym, fp[sub])) / (2 * e)
        check(gy64[:, i], fd_y, f"initial state {i}")
    for a, b, name in ((gy200[sub], gy64, "y0"), (gp200[sub], gp64, "conductances")):
        rel_l2 = np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(b, axis=0), 1e-300)
        cols = np.linalg.norm(b, axis=0) > 0
        print(f"65,536 nodes x 2,000 steps: FP32 vs FP64 d loss / d {name} on 64 nodes, max relative L2 per column "
              f"{rel_l2[cols].max():.2e}")
        assert rel_l2[cols].max() <= 5e-3


def test_adjoint_requires_a_supported_solver():
    from goss_amd._lib import GossHipError

    s = dev_system(VARIANT, "ImplicitEuler", 4, "fp64")
    with pytest.raises(GossHipError, match="explicit fixed-step solvers only"):
        s.adjoint_buffer_sizes(11, 5, 5)


# The explicit schemes: RL1 at the GRL1 step; RK4 at 2 us and ExplicitEuler at 1 us (TP06's fast gates
# make 5 us unstable for both, and 2 us for Euler with these perturbed nodes), from 9.8 ms through the
# stimulus and the upstroke.
EXPLICIT_CASES = [("RL1", 0.02, 300, 25, 40), ("ExplicitEuler", 0.001, 1200, 100, 128), ("RK4", 0.002, 600, 50, 64),
                  ("RK2", 0.001, 1200, 100, 128), ("RL2", 0.02, 300, 25, 40), ("GRL2", 0.02, 300, 25, 40)]


@pytest.mark.parametrize("solver,dt,nsteps,se,ce", EXPLICIT_CASES)
def test_explicit_adjoint_fp64_matches_torch_autograd(solver, dt, nsteps, se, ce):
    from torch_reference import TorchModel, sampled_trace

    N = 5
    y0, fp = _inputs(N, 71)
    t = 9.8 + np.arange(nsteps + 1) * dt
    w = np.random.Generator(np.random.PCG64(72)).normal(size=(nsteps // se, N))
    loss, v, gy, gp = _device_grads("fp64", y0, fp, t, se, ce, w, solver)
    tm = TorchModel(VARIANT)
    yt = torch.tensor(y0, requires_grad=True)
    ft = torch.tensor(fp, requires_grad=True)
    vr = sampled_trace(tm, yt, ft, t, se, 17, solver=solver)
    (vr * torch.tensor(w)).sum().backward()
    print(f"{solver}: V {v.min():.2f} .. {v.max():.2f} mV; samples rel err {_rel(v, vr.detach().numpy()):.2e}; "
          f"grad y0 {_rel(gy, yt.grad.numpy()):.2e}; grad g {_rel(gp, ft.grad.numpy()):.2e}")
    assert np.all(np.isfinite(v))
    assert _rel(v, vr.detach().numpy()) <= 1e-10
    assert _rel(gy, yt.grad.numpy()) <= 1e-8
    assert _rel(gp, ft.grad.numpy()) <= 1e-8


@pytest.mark.parametrize("solver,dt,nsteps,se,ce", EXPLICIT_CASES)
def test_explicit_adjoint_fp32_close_to_fp64_and_checkpoint_invariant(solver, dt, nsteps, se, ce):
    N = 128
    y0, fp = _inputs(N, 73)
    t = 9.8 + np.arange(nsteps + 1) * dt
    w = np.random.Generator(np.random.PCG64(74)).normal(size=(nsteps // se, N))
    ref = _device_grads("fp64", y0, fp, t, se, ce, w, solver)
    for every in (7, 1):  # 1: every state stored, the reverse sweep alone
        other = _device_grads("fp64", y0, fp, t, se, every, w, solver)
        assert np.array_equal(ref[2], other[2]) and np.array_equal(ref[3], other[3]), every
    got = _device_grads("fp32", y0, fp, t, se, ce, w, solver)
    got1 = _device_grads("fp32", y0, fp, t, se, 1, w, solver)
    assert np.array_equal(got[2], got1[2]) and np.array_equal(got[3], got1[3])
    for k, name in ((2, "y0"), (3, "conductances")):
        a, b = got[k], ref[k]
        rel_l2 = np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(b, axis=0), 1e-300)
        cols = np.linalg.norm(b, axis=0) > 0
        print(f"{solver} FP32 vs FP64, d loss / d {name}: max relative L2 per column {rel_l2[cols].max():.2e}")
        assert np.all(np.isfinite(a)) and rel_l2[cols].max() <= 5e-3


@pytest.mark.parametrize("prec,elem", [("fp64", 8), ("fp32", 4)])
def test_every_state_checkpoint_sizes(prec, elem):
    """Checkpoint interval 1: one state copy before every step, (n - 1) x S x ld values; the segment
    buffer still holds one state row block, the work rows, lam and the FP64 parameter gradients."""
    N, n = 100, 41  # ld = 128 (N rounded up to a multiple of 64)
    s = dev_system(VARIANT, "GRL1", N, prec)
    ckpt, seg, ns = s.adjoint_buffer_sizes(n, 10, 1)
    S = s.model.num_states
    assert ckpt == (n - 1) * S * 128 * elem
    assert ns == (n - 1) // 10
    ckpt200, seg200, _ = s.adjoint_buffer_sizes(n, 10, 200)
    assert ckpt200 == S * 128 * elem  # one checkpoint: the 40 steps fit in one segment
    # Interval 200: two buffers of a segment's recomputed states (the recompute of one segment overlaps
    # the reverse sweep of the next), 2 x 200 row blocks against 1.
    assert seg200 - seg == 399 * S * 128 * elem


# RK4 at dt 0.002 ms is unstable for some of these 4,096 nodes (as the torch reference shows); 0.001 is not.
MANY_NODE_CASES = [c if c[0] != "RK4" else ("RK4", 0.001, 1200, 100, 128) for c in EXPLICIT_CASES] + \
    [("GRL1", 0.02, 300, 25, 40)]


@pytest.mark.parametrize("solver,dt,nsteps,se,ce", MANY_NODE_CASES)
def test_explicit_adjoint_many_divergent_nodes(solver, dt, nsteps, se, ce):
    """4,096 nodes (64 wavefronts) with initial states spread by +-3 %: the nodes reach the upstroke at
    different times, so the lanes of every wavefront sit on both sides of the V = -40 mV gate switches.
    The FP64 and FP32 adjoints must stay finite and agree (relative L2 per column)."""
    N = 4096
    y0, fp = _inputs(N, 75, spread=0.03)
    t = 9.8 + np.arange(nsteps + 1) * dt
    w = np.random.Generator(np.random.PCG64(77)).normal(size=(nsteps // se, N))
    ref = _device_grads("fp64", y0, fp, t, se, ce, w, solver)
    got = _device_grads("fp32", y0, fp, t, se, ce, w, solver)
    assert np.all(np.isfinite(ref[1])) and np.all(np.isfinite(ref[2])) and np.all(np.isfinite(ref[3]))
    v = ref[1].reshape(ref[1].shape[0], N // 64, 64)
    straddle = ((v.max(axis=2) > -40.0) & (v.min(axis=2) < -40.0)).any(axis=0)
    assert straddle.all()  # every wavefront diverges at some sample
    for k, name in ((2, "y0"), (3, "conductances")):
        a, b = got[k], ref[k]
        rel_l2 = np.linalg.norm(a - b, axis=0) / np.maximum(np.linalg.norm(b, axis=0), 1e-300)
        cols = np.linalg.norm(b, axis=0) > 0
        print(f"{solver} {N} nodes: FP32 vs FP64 d loss / d {name}: max relative L2 per column {rel_l2[cols].max():.2e}")
        assert np.all(np.isfinite(a)) and rel_l2[cols].max() <= 5e-3
