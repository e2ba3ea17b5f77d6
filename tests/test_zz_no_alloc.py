"""goss_hip_system_forward_times_noalloc against goss_hip_system_forward_times.

The allocation-free entry point stages the steps in the library's fixed
device array; its results must equal the allocating entry point's bit for
bit.  Two calls in a row, the second with fewer time points than the first,
check that stale contents of the shared array never leak into a call."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from goss_amd._lib import GossHipError  # noqa: E402
from goss_amd.system import DeviceSystem  # noqa: E402

MAX_POINTS = 32768


def _pair(variant, solver, N, precision, params=None):
    a = DeviceSystem(variant, solver, N, precision, params)
    b = DeviceSystem(variant, solver, N, precision, params)
    return a, b


def _inputs(a, N, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    ic = a.model.initial_conditions
    y0 = ic * (1.0 + rng.uniform(-0.01, 0.01, (N, ic.size)))
    fp = None
    if a.model.num_field_parameters:
        g = np.array([a.model.default_parameters[i] for i in a.model.field_parameter_indices])
        fp = g * np.exp(rng.normal(0.0, 0.2, (N, g.size)))
    return y0, fp


@pytest.mark.parametrize("variant,solver,precision,dt", [
    ("tp06_m_fp3", "GRL1", "fp64", 0.02),
    ("tp06_m_fp3", "GRL1", "fp32c", 0.02),
    ("fitzhughnagumo", "RK4", "fp32", 0.05),
    ("tp06_epi", "ImplicitEuler", "fp64", 0.1),
    ("lorentz", "RKF32", "fp64", 0.02),
])
def test_noalloc_equals_forward_times_two_calls(variant, solver, precision, dt):
    N = 1000  # not a multiple of 64: the padded lanes of every row stay untouched
    a, b = _pair(variant, solver, N, precision)
    y0, fp = _inputs(a, N, 11)
    for s in (a, b):
        s.set_states(y0)
        if fp is not None:
            s.set_field_parameters(fp)
    # First call: a long run through the stimulus (TP06) with the full-size staging.
    n1 = 900 if solver in ("ImplicitEuler", "RKF32") else 1200
    t1 = np.arange(n1 + 1) * dt
    a.forward_times(t1)
    b.forward_times_noalloc(t1)
    np.testing.assert_array_equal(a.states(), b.states())
    # Second call: fewer points, continuing in time.
    t2 = t1[-1] + np.arange(37) * dt
    a.forward_times(t2)
    b.forward_times_noalloc(t2)
    sa, sb = a.states(), b.states()
    assert np.all(np.isfinite(sa))
    np.testing.assert_array_equal(sa, sb)
    np.testing.assert_array_equal(a.solver_counters(), b.solver_counters())
    a.close()
    b.close()


def test_noalloc_rejects_more_points_than_its_array():
    s = DeviceSystem("tp06_m_fp3", "GRL1", 64, "fp64")
    before = s.states()
    with pytest.raises(GossHipError, match="at most 32768 time points"):
        s.forward_times_noalloc(np.arange(MAX_POINTS + 1) * 0.02)
    np.testing.assert_array_equal(s.states(), before)
    # The maximum itself is accepted and equals the allocating path.
    r = DeviceSystem("tp06_m_fp3", "GRL1", 64, "fp64")
    t = np.arange(MAX_POINTS) * 1e-4
    s.forward_times_noalloc(t)
    r.forward_times(t)
    np.testing.assert_array_equal(s.states(), r.states())
    with pytest.raises(GossHipError, match="finite"):
        s.forward_times_noalloc(np.array([0.0, np.nan]))
    s.close()
    r.close()
