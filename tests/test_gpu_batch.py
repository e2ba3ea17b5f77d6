"""Mixed-population batch (goss_hip_batch_*, goss_amd.batch) on the device.

The fused kernel runs each node with its own population's per-node code, so the
reference for a batch is each system advanced alone with its own kernel
(goss_hip_system_forward_times) over the same times: the states must be
identical, bit for bit, for every launch shape.  The single-system kernels are
themselves checked against the oracle elsewhere (FP32 GRL1/RK4 accuracy,
RKF32 in FP32).
"""
import numpy as np
import pytest

from test_gpu_explicit import dev_system

pytestmark = pytest.mark.gpu

POPS = [("tp06_m", "GRL1", 0.02), ("tp06_epi", "GRL1", 0.02), ("fitzhughnagumo", "RK4", 0.05),
        ("lorentz", "RKF32", 0.02)]


def _make(sizes, seed):
    from goss_amd.system import ModelVariant

    rng = np.random.Generator(np.random.PCG64(seed))
    systems, y0 = [], []
    for (v, solver, _), n in zip(POPS, sizes):
        mv = ModelVariant(v)
        y = mv.initial_conditions[None, :] * (1.0 + rng.uniform(-0.01, 0.01, (n, mv.num_states)))
        s = dev_system(v, solver, n, "fp32")
        s.set_states(y)
        systems.append(s)
        y0.append(y)
    return systems, y0


def _alone(sizes, y0, t0, t1):
    out = []
    for (v, solver, dt), n, y in zip(POPS, sizes, y0):
        s = dev_system(v, solver, n, "fp32")
        s.set_states(y)
        k = int(round((t1 - t0) / dt))
        t = t0 + np.arange(k + 1) * dt
        t[-1] = t1
        s.forward_times(t)
        out.append(s.states())
        s.close()
    return out


def _dts():
    return [dt for _, _, dt in POPS]


@pytest.mark.parametrize("block,npl,span,order", [(256, 1, 0.5, ()), (64, 3, 3.0, (3, 0, 1, 2)),
                                                  (128, 2, 0.1, (2, 3, 1, 0)), (192, 1, 1.5, ())])
def test_batch_equals_each_population_alone(block, npl, span, order):
    from goss_amd.batch import Config, DeviceBatch

    sizes = [1000, 701, 3001, 97]
    t0, t1 = 9.0, 12.0  # through the TP06 stimulus edges at 10 and 11 ms
    systems, y0 = _make(sizes, 11)
    b = DeviceBatch(systems)
    b.config = Config(block, npl, span, order)
    b.forward(t0, t1, _dts())
    b.synchronize()
    ref = _alone(sizes, y0, t0, t1)
    for (v, _, _), s, r, y in zip(POPS, systems, ref, y0):
        got = s.states()
        assert np.all(np.isfinite(got)), v
        assert np.abs(got - y.astype(np.float32)).max() > 0, v  # the population moved
        assert np.array_equal(got, r), f"{v}: max abs diff {np.abs(got - r).max():.3e}"
    b.close()


def test_batch_second_call_reuses_and_continues():
    from goss_amd.batch import DeviceBatch

    sizes = [300, 200, 500, 64]
    systems, y0 = _make(sizes, 12)
    b = DeviceBatch(systems)
    b.forward(0.0, 1.0, _dts())
    b.forward(1.0, 2.0, _dts())  # same shape: the captured graph is replayed with new times
    b.synchronize()
    ref = _alone(sizes, y0, 0.0, 2.0)
    for s, r in zip(systems, ref):
        assert np.array_equal(s.states(), r)


def test_autotune_keeps_states_and_picks_a_valid_config():
    from goss_amd.batch import DeviceBatch

    sizes = [2000, 1500, 4000, 128]
    systems, y0 = _make(sizes, 13)
    b = DeviceBatch(systems)
    before = [s.states() for s in systems]
    cfg = b.autotune(0.0, 1.0, _dts())
    after = [s.states() for s in systems]
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    assert cfg.block_size in (64, 128, 256) and cfg.nodes_per_lane in (1, 2, 4) and cfg.span > 0
    assert sorted(cfg.order) == [0, 1, 2, 3]
    assert b.config == cfg
    b.forward(0.0, 1.0, _dts())
    b.synchronize()
    ref = _alone(sizes, y0, 0.0, 1.0)
    for s, r in zip(systems, ref):
        assert np.array_equal(s.states(), r)


def test_config_4_shape_one_millisecond():
    # TP06 M 300,017 + TP06 epi 211,003 (GRL1, dt 0.02) + FitzHugh-Nagumo 700,001 (RK4, dt 0.05)
    # + Lorenz 9,973 (RKF32, interval 0.02), FP32.
    from goss_amd.batch import DeviceBatch

    sizes = [300017, 211003, 700001, 9973]
    systems, y0 = _make(sizes, 14)
    b = DeviceBatch(systems)
    b.forward(0.0, 1.0, _dts())
    b.synchronize()
    ref = _alone(sizes, y0, 0.0, 1.0)
    for (v, _, _), s, r in zip(POPS, systems, ref):
        assert np.array_equal(s.states(), r), v


def test_batch_rejects_what_it_cannot_run():
    from goss_amd._lib import GossHipError
    from goss_amd.batch import Config, DeviceBatch

    a = dev_system("tp06_m", "GRL1", 64, "fp64")
    with pytest.raises(GossHipError, match="FP32"):
        DeviceBatch([a])
    c = dev_system("tp06_m", "RK4", 64, "fp32")
    with pytest.raises(GossHipError, match="no fused batch kernel"):
        DeviceBatch([c])
    d = dev_system("fitzhughnagumo", "RK4", 64, "fp32")
    with pytest.raises(GossHipError, match="twice"):
        DeviceBatch([d, d])
    b = DeviceBatch([d])
    with pytest.raises(GossHipError, match="whole number"):
        b.forward(0.0, 1.0, [0.03])
    with pytest.raises(GossHipError, match="block size"):
        b.config = Config(100, 1, 1.0)
    e = dev_system("lorentz", "RKF32", 64, "fp32")
    b2 = DeviceBatch([d, e])
    with pytest.raises(GossHipError, match="permutation"):
        b2.config = Config(256, 1, 1.0, (1, 1))


def test_batch_fp32_populations_against_the_fp64_oracle():
    """The batch's FitzHugh-Nagumo (RK4) and TP06 epi (GRL1) populations after 20 ms in FP32,
    against the FP64 oracle: V within 1e-3 of its range for FitzHugh-Nagumo and within
    0.05 mV for TP06 epi."""
    import oracle
    from goss_amd.batch import DeviceBatch

    sizes = [256, 256, 512, 32]
    systems, y0 = _make(sizes, 21)
    b = DeviceBatch(systems)
    b.forward(0.0, 20.0, _dts())
    b.synchronize()
    for k in (1, 2):
        v, solver, dt = POPS[k]
        o = oracle.OracleSystem(v, solver, sizes[k])
        o.set_num_threads(8)
        o.set_states(y0[k])
        o.forward_times(np.arange(int(round(20.0 / dt)) + 1) * dt)
        got, ref = systems[k].states(), o.states()
        vi = 17 if v == "tp06_epi" else 0
        dv = np.abs(got[:, vi] - ref[:, vi]).max()
        span = np.ptp(np.concatenate([y0[k][:, vi], ref[:, vi]]))
        print(f"batch {v} {solver} FP32 vs FP64 oracle: max |dV| {dv:.3e} (V range {span:.3e})")
        assert dv <= (0.05 if v == "tp06_epi" else 1e-3 * span)
    b.close()
