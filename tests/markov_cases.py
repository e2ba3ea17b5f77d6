"""Inputs of the three-state Markov chain that make RKF32 start a call from
AdaptiveExplicitSolver::dtinit (test infrastructure).

With intervals of 0.05 shorter than RKF32's first step guess (_ldt = 0.1), the
first step of a call overshoots the end of the interval and the next one
integrates backwards to it (AdaptiveExplicitSolver.cpp:230-241); when that
backward step is accepted, the negative step is stored as the guess and the
next call takes its first step from dtinit (RKF32.cpp:134-141).  When it is
rejected, the reference loops for ever (the step is reset to the same
backward step after every rejection); the oracle and the device stop such a
node after 10^6 sub-steps.  finishing_nodes() keeps the nodes that finish.
"""
import numpy as np

import oracle

DT, STEPS = 0.05, 40


def markov_inputs(N, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.dirichlet(np.ones(3), N), 0.3 * np.exp(rng.normal(0.0, 0.3, (N, 1)))


def finishing_nodes(y0, fp, t, model_params=None):
    good = []
    for n in range(y0.shape[0]):
        o = oracle.OracleSystem("markov_three_state", "RKF32", 1, per_node_state=True)
        for k, v in (model_params or {}).items():
            o.set_parameter(k, v)
        o.set_states(y0[n:n + 1])
        o.set_field_parameters(fp[n:n + 1])
        try:
            o.forward_times(t)
        except RuntimeError:
            continue
        good.append(n)
    return np.array(good, dtype=int)
