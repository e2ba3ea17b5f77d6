"""Model libraries built at run time from a gotran ``.ode`` file.

The reference compiles a ``goss::ParameterizedODE`` subclass for each
(model, field states, field parameters, monitored) combination with gotran and
cppyy at run time (``python/goss/compilemodule.py:19-58``,
``python/goss/codegeneration.py:226-370``).  Here the project's own code
generator writes the variant's expression templates and hipcc compiles them
with the library's kernel templates for gfx950 into a small model library;
``goss_hip_model_load`` registers it with the main library, after which the
variant is used like the compiled-in ones.

Libraries are cached under ``goss_amd/_jit/<key>/`` by a key over the model
text, the field selection and the library's sources, so each combination is
compiled once.  Monitored values need no compilation: every named
intermediate is evaluated by the generated ``eval_intermediates``.
"""
from __future__ import annotations

import fcntl
import hashlib
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor
from typing import Dict, Optional, Sequence, Tuple

from ._lib import check, lib

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
CACHE = os.path.join(HERE, "_jit")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wno-unused-command-line-argument",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-DGOSS_PLUGIN"]
# The two arithmetic families of goss_amd/csrc/Makefile.
FAMILIES = {
    "f64": ["-ffp-contract=off", "-DGOSS_PREC_F64"],
    "f32": ["-ffp-contract=fast", "-fgpu-flush-denormals-to-zero", "-DGOSS_PREC_F32"],
}
SOURCES = ("variant_kernels.hip", "goss_kernels.h", "goss_implicit.h", "goss_adaptive.h", "goss_adjoint.h",
           "goss_model_wrap.h", "goss_model_support.h", "goss_registry.h", "goss_layout.h")

_loaded: Dict[str, str] = {}


def _key(ode_text: str, field_states: Sequence[str], field_parameters: Sequence[str]) -> str:
    h = hashlib.sha256()
    h.update(ode_text.encode())
    h.update(repr((list(field_states), list(field_parameters))).encode())
    for f in SOURCES:
        with open(os.path.join(CSRC, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def build(ode_path: str, field_states: Sequence[str] = (), field_parameters: Sequence[str] = (),
          jobs: int = 2) -> Tuple[str, str]:
    """(variant name, library path) of the model library for this selection,
    compiled unless it is cached."""
    from .codegen.emit import Variant, emit_model_header
    from .codegen.ode_model import load_ode

    with open(ode_path) as f:
        text = f.read()
    name = "jit_" + _key(text, field_states, field_parameters)
    out_dir = os.path.join(CACHE, name)
    so = os.path.join(out_dir, f"libgoss_model_{name}.so")
    if os.path.exists(so):
        return name, so
    os.makedirs(os.path.join(out_dir, "generated"), exist_ok=True)
    with open(os.path.join(out_dir, ".lock"), "w") as lockf:
        fcntl.flock(lockf, fcntl.LOCK_EX)
        if os.path.exists(so):
            return name, so
        model = load_ode(ode_path)
        var = Variant(model=model, name=name, field_states=list(field_states),
                      field_parameters=list(field_parameters))
        em = emit_model_header(var)
        with open(os.path.join(out_dir, "generated", f"{name}.h"), "w") as f:
            f.write(em.header)
        incs = [f"-I{CSRC}", f"-I{out_dir}", f"-I{os.path.join(HERE, '..', 'include')}"]

        def compile_family(fam):
            obj = os.path.join(out_dir, f"{name}_{fam}.o")
            cmd = [HIPCC] + COMMON + FAMILIES[fam] + incs + [f"-DGOSS_VARIANT={name}", "-c",
                                                             os.path.join(CSRC, "variant_kernels.hip"), "-o", obj]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"compiling {ode_path} ({fam}) failed:\n{r.stderr[-4000:]}")
            return obj

        with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
            objs = list(ex.map(compile_family, FAMILIES))
        tmp = so + ".tmp"
        r = subprocess.run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs,
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"linking the model library of {ode_path} failed:\n{r.stderr[-4000:]}")
        os.replace(tmp, so)
    return name, so


def load(ode_path: str, field_states: Sequence[str] = (), field_parameters: Sequence[str] = (),
         library: Optional[str] = None) -> str:
    """Builds (or finds) the model library and registers its variant with the
    main library; returns the variant name."""
    name, so = build(ode_path, field_states, field_parameters)
    if name not in _loaded:
        check(lib().goss_hip_model_load(so.encode(), name.encode()))
        _loaded[name] = so
    return name
