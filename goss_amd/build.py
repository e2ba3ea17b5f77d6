"""Build libgoss_hip.so: regenerate the model code, then compile every
variant for gfx950 (goss_amd/csrc/Makefile)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 8, force: bool = True) -> str:
    from .codegen.generate import generate

    generate()
    csrc = os.path.join(HERE, "csrc")
    if force:
        # Always recompile (the build check compiles every time it is asked).
        subprocess.check_call(["make", "-s", "-C", csrc, "clean"])
    subprocess.check_call(["make", "-s", f"-j{jobs}", "-C", csrc])
    return os.path.join(HERE, "libgoss_hip.so")


if __name__ == "__main__":
    print(build())
