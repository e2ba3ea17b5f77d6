"""The C-ABI library loads and exports what include/goss_hip.h declares.

No compute call is made here (no GPU is needed)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from goss_amd import _lib
from goss_amd.codegen.generate import VARIANTS
from goss_amd.system import ModelVariant, model_names, solver_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "goss_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(goss_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "libgoss_hip.so is not built"
    out = subprocess.check_output(["nm", "-D", "--defined-only", _lib.LIB_PATH], text=True)
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    decl = declared_functions()
    assert len(decl) >= 30
    missing = [d for d in decl if d not in exported]
    assert not missing, missing
    # ...and the ctypes binding covers exactly the declared set.
    assert sorted(_lib.EXPORTS) == decl
    L = _lib.lib()
    for d in decl:
        assert hasattr(L, d)


def test_model_registry_matches_generated_variants():
    names = model_names()
    assert sorted(names) == sorted(VARIANTS)
    m = ModelVariant("tp06_m_fp3")
    assert m.num_states == 19 and m.num_parameters == 53
    assert m.field_parameter_names == ["g_Kr", "g_Ks", "g_CaL"]
    assert m.field_state_names == ["V"] and m.field_state_indices == [17]
    assert m.default_parameters[m.parameter_names.index("g_Kr")] == 0.153
    assert m.initial_conditions[17] == -85.423
    # linear_terms: the 12 gates and R_prime are affine in themselves.
    assert m.linear_terms == [1] * 12 + [0, 1, 0, 0, 0, 0, 0]
    with pytest.raises(_lib.GossHipError):
        ModelVariant("no_such_model")


def test_solver_param_defaults_are_the_reference_defaults():
    p = solver_params("GRL1")
    assert p.internal_time_step == -1.0  # ODESolver.h:137
    assert p.delta == 1e-8  # GRL1.h:39
    q = solver_params("ImplicitEuler")
    # ImplicitODESolver.h:36-41, ImplicitEuler.h:36-37
    assert (q.eta_0, q.kappa, q.relative_tolerance, q.max_iterations, q.max_relative_previous_residual) == (
        1.0, 0.1, 1e-12, 30, 0.01)
    assert q.min_dt == 0.0001 and q.num_refinements_without_always_recomputing_jacobian == 2
    assert solver_params("ESDIRK23a").min_dt == 0.001  # ESDIRK23a.h:38
    with pytest.raises(KeyError):
        solver_params("GRL1", no_such=1.0)
    L = _lib.lib()
    assert L.goss_hip_solver_is_adaptive(_lib.SOLVERS["RKF32"]) == 1
    assert L.goss_hip_solver_is_adaptive(_lib.SOLVERS["GRL1"]) == 0


def test_create_without_device_fails_cleanly():
    L = _lib.lib()
    n = ctypes.c_int(-1)
    assert L.goss_hip_device_count(ctypes.byref(n)) == 0
    if n.value > 0:
        pytest.skip("a GPU is visible")
    from goss_amd.system import DeviceSystem

    with pytest.raises(_lib.GossHipError):
        DeviceSystem("tp06_m", "GRL1", 16)


def test_unknown_names_are_rejected_before_any_device_work():
    from goss_amd.system import DeviceSystem

    with pytest.raises(_lib.GossHipError, match="unknown model variant"):
        DeviceSystem("nope", "GRL1", 4)
