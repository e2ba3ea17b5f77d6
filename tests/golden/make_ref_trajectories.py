"""Golden trajectories from the reference's own ODESystemSolver.

Runs in a container where oracle/_ref/libgoss_ref.so was built from the
reference sources (cpp/goss/*.cpp) and writes the fixtures that the tests
compare the oracle and the HIP path against:

* ref_tp06_grl1_1024x10000.npz -- TP06 M-cell, GRL1, FP64, 1,024 nodes,
  dt = 0.01 ms, 10,000 steps; y0 = defaults * (1 + U(-0.01, 0.01)), seeded;
  default stimulus (-52 pA/pF on [10, 11] ms).  Final states and V of nodes
  0..7 every 100 steps.
* ref_tp06_fp3_window.npz -- TP06 M-cell with field parameters g_Kr, g_Ks,
  g_CaL = default * exp(N(0, 0.2^2)), GRL1 and RL2, dt = 0.02 ms, 128 nodes,
  5,000 steps (100 ms); final states.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

SEEDS = {"tp06_1024": 1001, "fp3_window": 1002}


def init_states(sys_, rng):
    y = sys_.states()
    return y * (1.0 + rng.uniform(-0.01, 0.01, y.shape))


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libgoss_ref.so is not built")
    # TP06 M-cell GRL1: 1,024 nodes, dt 0.01 ms, 10,000 steps.
    rng = np.random.Generator(np.random.PCG64(SEEDS["tp06_1024"]))
    N, dt, steps = 1024, 0.01, 10000
    r = oracle.RefSystem("tp06_m", "GRL1", N)
    r.set_num_threads(8)
    y0 = init_states(r, rng)
    r.set_states(y0)
    t = np.arange(steps + 1) * dt
    trace = [r.states()[:8, 17].copy()]
    for k in range(0, steps, 100):
        r.forward_times(t[k:k + 101])
        trace.append(r.states()[:8, 17].copy())
    np.savez_compressed(os.path.join(HERE, "ref_tp06_grl1_1024x10000.npz"), y0=y0, t=t, final=r.states(),
                        V_trace=np.array(trace))
    # TP06 M-cell with 3 field parameters, a 5,000-step window of 128 nodes at dt 0.02 ms.
    rng = np.random.Generator(np.random.PCG64(SEEDS["fp3_window"]))
    N, dt, steps = 128, 0.02, 5000
    out = {}
    y0 = None
    for solver in ("GRL1", "RL2"):
        r = oracle.RefSystem("tp06_m_fp3", solver, N)
        if y0 is None:
            y0 = init_states(r, rng)
            fp = r.field_parameters() * np.exp(rng.normal(0.0, 0.2, (N, 3)))
        r.set_states(y0)
        r.set_field_parameters(fp)
        r.forward_times(np.arange(steps + 1) * dt)
        out[solver] = r.states()
    np.savez_compressed(os.path.join(HERE, "ref_tp06_fp3_window.npz"), y0=y0, field_parameters=fp,
                        dt=dt, steps=steps, final_GRL1=out["GRL1"], final_RL2=out["RL2"])
    with open(os.path.join(HERE, "seeds.json"), "w") as f:
        json.dump(SEEDS, f, indent=1)
    print("written")


if __name__ == "__main__":
    main()
