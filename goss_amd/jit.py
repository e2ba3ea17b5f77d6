"""Model libraries built at run time from a gotran ``.ode`` file.

The reference compiles a ``goss::ParameterizedODE`` subclass for each
(model, field states, field parameters, monitored) combination with gotran and
cppyy at run time (``python/goss/compilemodule.py:19-58``,
``python/goss/codegeneration.py:226-370``).  Here the project's own code
generator writes the variant's expression templates and hipcc compiles them
with the library's kernel templates for gfx950 into a small model library;
``goss_hip_model_load`` registers it with the main library, after which the
variant is used like the compiled-in ones.

A variant's metadata (sizes, names, defaults, linear terms, intermediates)
comes from the code generator without compiling (:func:`metadata`), so a
model can be inspected before its first device use, which compiles it.

Libraries are cached under ``goss_amd/_jit/<key>/`` by a key over the model
text, the field selection, the code generator and the library's sources, so
each combination is compiled once.  Monitored values need no compilation: every named
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
    codegen = os.path.join(HERE, "codegen")
    for f in sorted(os.listdir(codegen)):
        if f.endswith(".py"):
            with open(os.path.join(codegen, f), "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def variant_name(ode_path: str, field_states: Sequence[str] = (), field_parameters: Sequence[str] = ()) -> str:
    """Name of the run-time variant for this selection (no compilation)."""
    with open(ode_path) as f:
        text = f.read()
    return "jit_" + _key(text, field_states, field_parameters)


class ModelMeta:
    """The attributes of :class:`goss_amd.system.ModelVariant`, computed by the
    code generator for a variant that is not compiled yet."""

    def __init__(self, ode_path: str, field_states: Sequence[str], field_parameters: Sequence[str]):
        import numpy as np

        from .codegen.analysis import linearize
        from .codegen.emit import Variant
        from .codegen.ode_model import load_ode

        model = load_ode(ode_path)
        var = Variant(model=model, name=variant_name(ode_path, field_states, field_parameters),
                      field_states=list(field_states), field_parameters=list(field_parameters))
        self.name = var.name
        self.num_states = model.num_states
        self.num_parameters = model.num_parameters
        self.num_field_states = len(var.field_states)
        self.num_field_parameters = len(var.field_parameters)
        self.state_names = model.state_names
        self.parameter_names = model.parameter_names
        self.field_state_names = list(var.field_states)
        self.field_parameter_names = list(var.field_parameters)
        self.field_state_indices = list(var.field_state_indices)
        self.field_parameter_indices = list(var.field_parameter_indices)
        self.default_parameters = np.array([float(p.value) for p in model.parameters])
        self.initial_conditions = np.array([float(st.init) for st in model.states])
        self.linear_terms = list(linearize(model).linear_terms)
        self.differential_states = list(model.differential_states)
        self.is_dae = model.is_dae
        derivs = set(model.derivatives.values())
        # The order of the generated eval_intermediates (intermediate_names).
        self.intermediate_names = [n for n, _ in model.body if n not in derivs]


_meta: Dict[str, ModelMeta] = {}


def metadata(ode_path: str, field_states: Sequence[str] = (), field_parameters: Sequence[str] = ()) -> ModelMeta:
    name = variant_name(ode_path, field_states, field_parameters)
    if name not in _meta:
        _meta[name] = ModelMeta(ode_path, field_states, field_parameters)
    return _meta[name]


def build(ode_path: str, field_states: Sequence[str] = (), field_parameters: Sequence[str] = (),
          jobs: int = 2) -> Tuple[str, str]:
    """(variant name, library path) of the model library for this selection,
    compiled unless it is cached."""
    from .codegen.emit import Variant, emit_model_header
    from .codegen.ode_model import load_ode

    name = variant_name(ode_path, field_states, field_parameters)
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
