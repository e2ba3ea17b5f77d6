"""Host logic of bench.py: the accuracy gate's schedule and the CPU thread
count.  The gate must cover its horizon whatever --steps the timed region
uses, and must never reduce an empty set of samples."""
import os

import numpy as np
import pytest

import bench


def test_gate_schedule_headline():
    # 1000 ms at dt 0.02 ms, V every 1 ms: 1000 samples of 50 steps.
    assert bench.gate_schedule(1000.0, 1.0, 0.02) == (50, 1000)


@pytest.mark.parametrize("horizon", [0.02, 0.1, 0.5, 0.98])
def test_gate_schedule_shorter_than_one_sample(horizon):
    # A horizon below the sample interval still yields one sample, at the horizon.
    per, n = bench.gate_schedule(horizon, 1.0, 0.02)
    assert n == 1 and per == int(round(horizon / 0.02))
    t = np.arange(per * n + 1) * 0.02
    assert abs(t[-1] - horizon) < 1e-12


def test_gate_schedule_rejects_ragged_horizons():
    with pytest.raises(ValueError):
        bench.gate_schedule(0.01, 1.0, 0.02)  # less than one step
    with pytest.raises(ValueError):
        bench.gate_schedule(1.03, 1.0, 0.02)  # 51.5 steps


def test_gate_schedule_does_not_depend_on_timed_steps():
    # The gate's constants are the metric's (1000 ms, every 1 ms, 0.05 mV), not argparse values.
    assert (bench.GATE_HORIZON_MS, bench.GATE_SAMPLE_MS, bench.GATE_DV_MV) == (1000.0, 1.0, 0.05)


def test_host_threads_respects_affinity(monkeypatch):
    monkeypatch.delenv("OMP_NUM_THREADS", raising=False)
    threads, cores = bench.host_threads()
    assert threads == cores == len(os.sched_getaffinity(0))
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    threads, cores = bench.host_threads()
    assert threads == 1 and cores >= 1


def test_leg_records_failures():
    def boom():
        raise RuntimeError("no device")

    r = bench.leg("x", boom)
    assert r == {"error": "RuntimeError: no device"}
    assert bench.leg("y", lambda: {"ok": 1}) == {"ok": 1}


def test_gpus_above_one_exits_with_message(capsys, monkeypatch):
    monkeypatch.setattr("sys.argv", ["bench.py", "--gpus", "2"])
    assert bench.main() == 2
    assert "one GPU" in capsys.readouterr().err
