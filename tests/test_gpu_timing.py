"""Kernel timing of a stepping call (goss_hip_system_set/get_kernel_timing)."""
import numpy as np
import pytest

from goss_amd._lib import GossHipError
from goss_amd.system import DeviceSystem

pytestmark = pytest.mark.gpu


def test_kernel_timing_counts_launches_and_time():
    s = DeviceSystem("tp06_m_fp3", "GRL1", 65536, "fp32c")
    with pytest.raises(GossHipError, match="not enabled"):
        s.kernel_timing()
    s.set_kernel_timing(True)
    s.set_steps_per_launch(100)
    s.forward_times(np.arange(251) * 0.02)
    ms, launches = s.kernel_timing()
    assert launches == 3 and ms > 0
    s.forward_times(np.arange(51) * 0.02)
    ms2, launches2 = s.kernel_timing()
    assert launches2 == 1 and 0 < ms2 < ms
    s.set_kernel_timing(False)
    with pytest.raises(GossHipError, match="not enabled"):
        s.kernel_timing()
    s.close()
