"""Write known-answer vectors from the gotran-generated model classes that the
reference ships (cpp/tests/fhn.h, cpp/tests/beeler_reuter_1977.h).

Runs in a container that has the reference tree: compiles
gotran_header_kat.cpp against those headers with g++, evaluates eval and
linearized_eval at seeded random points and stores inputs and outputs in
gotran_kat_<model>.npz.  The tests compare the code generated by
goss_amd.codegen from the matching .ode files against these vectors.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

REF = os.environ.get("GOSS_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def sample(model, n, rng):
    if model == "fhn":
        t = rng.uniform(0.0, 4.0, n)
        v = rng.uniform(-90.0, 45.0, n)
        s = rng.uniform(-5.0, 40.0, n)
        return np.column_stack([t, v, s])
    # beeler_reuter_1977 states: m, h, j, Cai, d, f, x1, V
    t = rng.uniform(0.0, 2100.0, n)
    t[: n // 4] = rng.uniform(9.5, 11.5, n // 4)  # around the stimulus edges
    gates = rng.uniform(0.0, 1.0, (n, 3))
    cai = 10.0 ** rng.uniform(-5.0, -2.5, n)
    gates2 = rng.uniform(0.0, 1.0, (n, 3))
    V = rng.uniform(-95.0, 50.0, n)
    V[: 16] = -47.0 + rng.uniform(-0.02, 0.02, 16)  # V_eff_0 switch
    V[16:32] = -23.0 + rng.uniform(-0.02, 0.02, 16)  # V_eff_1 switch
    return np.column_stack([t, gates, cai, gates2, V])


def main():
    rng = np.random.Generator(np.random.PCG64(20241203))
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "kat")
        subprocess.check_call(["g++", "-O0", "-std=c++17", f"-I{REF}/cpp", f"-I{REF}/cpp/tests",
                               os.path.join(HERE, "gotran_header_kat.cpp"), f"{REF}/cpp/goss/ODE.cpp",
                               "-o", exe])
        for model, S in (("fhn", 2), ("beeler_reuter_1977", 8)):
            n = 512
            x = sample(model, n, rng)
            fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
            np.concatenate([[float(n)], x.ravel()]).astype(np.float64).tofile(fin)
            subprocess.check_call([exe, model, fin, fout])
            out = np.fromfile(fout, dtype=np.float64)
            body, flags = out[: n * 3 * S].reshape(n, 3, S), out[n * 3 * S:]
            np.savez_compressed(os.path.join(HERE, f"gotran_kat_{model}.npz"), time=x[:, 0], states=x[:, 1:],
                                eval=body[:, 0], linear=body[:, 1], rhs=body[:, 2],
                                linear_terms=flags.astype(np.int8))
            print(model, "written", n, "points")


if __name__ == "__main__":
    sys.exit(main())
