"""Known-answer vectors for the TP06 M-cell right-hand side.

The only independent statement of TP06 in the reference tree is the
gotran-generated Python function ``rhs(t, states, parameters)`` in
benchmarks/tentusscher_panfilov_2006_M_cell.py:7-356 (default parameters at
:400-455).  Its module imports goss/numba, which are absent, so this script
parses the file, takes only the ``rhs`` function definition and evaluates it
at seeded state points.  It runs in a container that has the reference tree
and writes tp06_rhs_vectors.npz (inputs and outputs only).
"""
import ast
import math
import os
import types

import numpy as np

REF = os.environ.get("GOSS_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(REF, "benchmarks", "tentusscher_panfilov_2006_M_cell.py")


def load_rhs_and_defaults():
    tree = ast.parse(open(SRC).read())
    fn = next(n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name == "rhs")
    # The function was written for numpy < 2 (np.float_); give it a numpy
    # namespace that still has that alias.
    np_compat = types.SimpleNamespace(**{k: getattr(np, k) for k in ("zeros", "exp", "log", "sqrt", "floor")})
    np_compat.float_ = np.float64
    ns = {"math": math, "np": np_compat, "numpy": np_compat}
    exec(compile(ast.Module(body=[fn], type_ignores=[]), SRC, "exec"), ns)
    main = next(n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name == "main")
    call = next(n for n in ast.walk(main) if isinstance(n, ast.Call) and getattr(n.func, "attr", "") == "array")
    params = np.array([float(ast.literal_eval(e)) for e in call.args[0].elts])
    return ns["rhs"], params


def main():
    rhs, params = load_rhs_and_defaults()
    rng = np.random.Generator(np.random.PCG64(20241204))
    ic = np.array([0.0165, 0.473, 0.0174, 0.00165, 0.749, 0.6788, 3.288e-05, 0.7026, 0.9526, 0.9942, 0.999998,
                   2.347e-08, 0.000153, 0.8978, 4.272, 0.00042, 10.132, -85.423, 138.52])
    n = 400
    states = ic * (1.0 + rng.uniform(-0.3, 0.3, (n, ic.size)))
    # Voltages across the whole range, including both sides of the V < -40
    # switches and the removable point of V_eff at V = 15.
    states[:, 17] = rng.uniform(-95.0, 50.0, n)
    states[:20, 17] = 15.0 + rng.uniform(-0.02, 0.02, 20)
    states[20:40, 17] = -40.0 + rng.uniform(-0.05, 0.05, 20)
    t = rng.uniform(0.0, 3000.0, n)
    t[:60] = rng.uniform(9.9, 11.1, 60)  # stimulus edges
    t[60:70] = [10.0, 11.0, 1010.0, 1011.0, 2010.0, 9.999999, 11.000001, 0.0, 1000.0, 2999.0]
    out = np.array([rhs(t[k], states[k], params) for k in range(n)])
    np.savez_compressed(os.path.join(HERE, "tp06_rhs_vectors.npz"), time=t, states=states, parameters=params,
                        values=out)
    print("written", n, "points")


if __name__ == "__main__":
    main()
