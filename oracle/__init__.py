"""ctypes bindings of the CPU checkers.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of bench.py, never by the product package.

* :class:`OracleSystem` -- our restatement of goss's ODESystemSolver and
  solvers (``goss_oracle.cpp``), built into ``oracle/liboracle.so``;
* :class:`RefSystem` -- the reference's own C++ (``cpp/goss/*.cpp``) built into
  ``oracle/_ref/libgoss_ref.so`` when the reference tree was available at build
  time (with our generated model plugins), and into
  ``oracle/_ref/libgoss_refpin.so`` with the reference's own gotran-generated
  test models (``cpp/tests/*.h``; variants ``ref_fhn``,
  ``ref_beeler_reuter_1977``, ``ref_winslow``, ``ref_panfilov``);
* :func:`register_reference_models` -- makes those reference classes oracle
  variants of the same names, so the oracle's solvers run on them.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libgoss_ref.so")
REFPIN_SO = os.path.join(HERE, "_ref", "libgoss_refpin.so")
REFERENCE_MODELS = ("ref_fhn", "ref_beeler_reuter_1977", "ref_winslow", "ref_panfilov", "ref_van_der_pol")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint)


def build(ref: Optional[str] = None, jobs: int = 8) -> None:
    """Build liboracle.so, and _ref/libgoss_ref.so when the reference exists."""
    ref = ref or os.environ.get("GOSS_REFERENCE", "/root/reference")
    targets = ["oracle"]
    if os.path.isdir(os.path.join(ref, "cpp", "goss")):
        targets.append("ref")
    subprocess.check_call(["make", "-s", f"-j{jobs}", "-C", HERE, f"REF={ref}"] + targets)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


_lib = None
_ref = {}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is not built (run oracle.build())")
        L = ctypes.CDLL(ORACLE_SO)
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_variant_name.restype = ctypes.c_char_p
        L.oracle_variant_differential_states.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_ubyte)]
        L.oracle_system_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_char_p), _dp, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_system_forward.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.oracle_system_forward_times.argtypes = [ctypes.c_void_p, _dp, ctypes.c_ulong]
        L.oracle_system_solve.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_ulong, ctypes.c_int]
        L.oracle_system_set_parameter.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]
        for fn in ("get_states", "set_states"):
            getattr(L, f"oracle_system_{fn}").argtypes = [ctypes.c_void_p, _dp]
        for fn in ("get_field_states", "set_field_states", "get_field_parameters", "set_field_parameters"):
            getattr(L, f"oracle_system_{fn}").argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        for fn in ("get_field_state_components", "set_field_state_components"):
            getattr(L, f"oracle_system_{fn}").argtypes = [ctypes.c_void_p, _dp, ctypes.c_uint, _up, ctypes.c_int]
        L.oracle_system_set_num_threads.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        L.oracle_system_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_system_reset_default.argtypes = [ctypes.c_void_p]
        L.oracle_system_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        L.oracle_system_esdirk4o32_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        L.oracle_model_eval.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp, _dp]
        L.oracle_model_linearized_eval.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp, _dp, _dp, ctypes.c_int]
        L.oracle_model_analytic_jacobian.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp, _dp]
        L.oracle_model_fd_jacobian.argtypes = [ctypes.c_int, _dp, ctypes.c_double, _dp, _dp]
        L.oracle_model_step_vjp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, _dp,
                                            ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.oracle_register_external_model.argtypes = [ctypes.c_void_p]
        L.oracle_lu_factorize.argtypes = [ctypes.c_int, _dp]
        L.oracle_forward_backward_subst.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        _lib = L
    return _lib


def ref_available(path: str = REF_SO) -> bool:
    return os.path.exists(path)


def refpin_available() -> bool:
    return os.path.exists(REFPIN_SO)


def ref_lib(path: str = REF_SO):
    if path not in _ref:
        if not ref_available(path):
            raise RuntimeError(f"{path} is not built")
        L = ctypes.CDLL(path)
        L.ref_last_error.restype = ctypes.c_char_p
        L.ref_system_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint,
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _dp,
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _dp,
                                        ctypes.POINTER(ctypes.c_void_p)]
        L.ref_system_forward.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.ref_system_forward_times.argtypes = [ctypes.c_void_p, _dp, ctypes.c_ulong]
        L.ref_system_solve.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_ulong, ctypes.c_int]
        for fn in ("get_states", "set_states"):
            getattr(L, f"ref_system_{fn}").argtypes = [ctypes.c_void_p, _dp]
        for fn in ("get_field_states", "set_field_states", "get_field_parameters", "set_field_parameters"):
            getattr(L, f"ref_system_{fn}").argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        for fn in ("get_field_state_components", "set_field_state_components"):
            getattr(L, f"ref_system_{fn}").argtypes = [ctypes.c_void_p, _dp, ctypes.c_uint, _up, ctypes.c_int]
        L.ref_system_set_num_threads.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        L.ref_system_destroy.argtypes = [ctypes.c_void_p]
        L.ref_system_reset_default.argtypes = [ctypes.c_void_p]
        L.ref_system_rkf32_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        L.ref_system_esdirk4o32_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_long)]
        L.ref_model_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)]
        L.ref_model_eval.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp]
        L.ref_model_linearized_eval.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp, _dp, ctypes.c_int]
        L.ref_model_compute_jacobian.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp]
        L.ref_model_lu_factorize.argtypes = [ctypes.c_void_p, _dp]
        L.ref_model_forward_backward_subst.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.ref_model_set_parameter.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]
        L.ref_model_set_field_parameters.argtypes = [ctypes.c_void_p, _dp]
        L.ref_model_linear_term.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        L.ref_model_destroy.argtypes = [ctypes.c_void_p]
        if path == REFPIN_SO:
            L.refpin_model_ops.restype = ctypes.c_void_p
            L.refpin_model_ops.argtypes = [ctypes.c_char_p]
            L.refpin_ode_run.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), _dp, _dp,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int]
        _ref[path] = L
    return _ref[path]


def refpin_ode_run(solver: str, y0, t0: float, dt: float, nsteps: int, solver_params=None) -> np.ndarray:
    """The reference's solver advancing the reference's van der Pol DAE test model
    (cpp/tests/van_der_Pol.h) from y0: forward(y, t, dt); t += dt, nsteps times."""
    R = ref_lib(REFPIN_SO)
    y = _f64(y0).copy()
    n, names, vals = _names(solver_params or {})
    if R.refpin_ode_run(solver.encode(), n, names, _ptr(vals), _ptr(y), float(t0), float(dt), int(nsteps)) != 0:
        raise RuntimeError(R.ref_last_error().decode())
    return y


def register_reference_models() -> None:
    """Registers the reference's gotran-generated test models (built into
    _ref/libgoss_refpin.so) as oracle variants of the same names."""
    R = ref_lib(REFPIN_SO)
    L = lib()
    for name in REFERENCE_MODELS:
        ops = R.refpin_model_ops(name.encode())
        if not ops:
            raise RuntimeError(R.ref_last_error().decode())
        if L.oracle_register_external_model(ops) < 0:
            raise RuntimeError(L.oracle_last_error().decode())


class Variant:
    """Sizes, defaults and field selection of a generated model variant."""

    def __init__(self, name: str):
        L = lib()
        self.index = L.oracle_variant_find(name.encode())
        if self.index < 0:
            raise KeyError(name)
        self.name = name
        sz = (ctypes.c_int * 4)()
        L.oracle_variant_sizes(self.index, sz)
        self.num_states, self.num_parameters, self.num_field_states, self.num_field_parameters = list(sz)
        self.default_parameters = np.zeros(self.num_parameters)
        self.initial_conditions = np.zeros(self.num_states)
        L.oracle_variant_defaults(self.index, _ptr(self.default_parameters), _ptr(self.initial_conditions))
        fsi = (ctypes.c_int * max(self.num_field_states, 1))()
        fpi = (ctypes.c_int * max(self.num_field_parameters, 1))()
        L.oracle_variant_field_indices(self.index, fsi, fpi)
        self.field_state_indices = list(fsi)[: self.num_field_states]
        self.field_parameter_indices = list(fpi)[: self.num_field_parameters]
        lt = (ctypes.c_ubyte * self.num_states)()
        L.oracle_variant_linear_terms(self.index, lt)
        self.linear_terms = list(lt)
        ds = (ctypes.c_ubyte * self.num_states)()
        L.oracle_variant_differential_states(self.index, ds)
        self.differential_states = list(ds)

    def eval(self, y, t, params=None):
        p = _f64(self.default_parameters if params is None else params)
        y = _f64(y)
        out = np.zeros(self.num_states)
        lib().oracle_model_eval(self.index, _ptr(y), float(t), _ptr(p), _ptr(out))
        return out

    def linearized_eval(self, y, t, params=None, only_linear=False):
        p = _f64(self.default_parameters if params is None else params)
        y = _f64(y)
        lin, rhs = np.zeros(self.num_states), np.zeros(self.num_states)
        lib().oracle_model_linearized_eval(self.index, _ptr(y), float(t), _ptr(p), _ptr(lin), _ptr(rhs),
                                           int(only_linear))
        return lin, rhs

    def analytic_jacobian(self, y, t, params=None):
        p = _f64(self.default_parameters if params is None else params)
        y = _f64(y)
        jac = np.zeros(self.num_states * self.num_states)
        lib().oracle_model_analytic_jacobian(self.index, _ptr(y), float(t), _ptr(p), _ptr(jac))
        return jac.reshape(self.num_states, self.num_states)

    def _step_vjp(self, kind, y, t, h, lam, params, delta, tape, out):
        p = _f64(self.default_parameters if params is None else params)
        y, lam = _f64(y), _f64(lam)
        fp_bar = np.zeros(max(self.num_field_parameters, 1))
        lib().oracle_model_step_vjp(self.index, kind, int(tape), _ptr(y), float(t), _ptr(p), float(h), float(delta),
                                    _ptr(lam), _ptr(out), _ptr(fp_bar))
        return out, fp_bar[: self.num_field_parameters]

    def grl1_vjp(self, y, t, h, lam, params=None, delta=1e-8, tape=False):
        """((dy'/dy)^T lam, (dy'/dp_field)^T lam) for one GRL1 step y -> y' (generated code;
        tape=True: the variant that keeps intermediates and adjoints in a tape)."""
        return self._step_vjp(0, y, t, h, lam, params, delta, tape, np.zeros(self.num_states))

    def rl1_vjp(self, y, t, h, lam, params=None, tape=False):
        """The same for one RL1 step (RL1.h:68-77)."""
        return self._step_vjp(1, y, t, h, lam, params, 1e-8, tape, np.zeros(self.num_states))

    def eval_vjp(self, y, t, seed, params=None, tape=False):
        """((d rhs/dy)^T seed, (d rhs/dp_field)^T seed) (generated code)."""
        return self._step_vjp(2, y, t, 0.0, seed, params, 0.0, tape, np.zeros(self.num_states))

    def grl1_step(self, y, t, h, params=None, delta=1e-8):
        """One GRL1 step (GRL1.h:72-84) in FP64 from linearized_eval."""
        lin, rhs = self.linearized_eval(y, t, params)
        y = _f64(y)
        with np.errstate(all="ignore"):
            inc = np.where(np.abs(lin) > delta, rhs / lin * np.expm1(lin * h), rhs * h)
        return y + inc

    def fd_jacobian(self, y, t, params=None):
        p = _f64(self.default_parameters if params is None else params)
        y = _f64(y).copy()
        jac = np.zeros(self.num_states * self.num_states)
        lib().oracle_model_fd_jacobian(self.index, _ptr(y), float(t), _ptr(p), _ptr(jac))
        return jac.reshape(self.num_states, self.num_states)


def _names(d: Dict[str, float]):
    keys = list(d)
    arr = (ctypes.c_char_p * max(len(keys), 1))(*[k.encode() for k in keys])
    vals = np.array([float(d[k]) for k in keys] or [0.0])
    return len(keys), arr, vals


class _SystemBase:
    _prefix = ""

    def __init__(self, variant: str, solver: str, num_nodes: int):
        self.variant = Variant(variant)
        self.num_nodes = int(num_nodes)
        self.solver = solver

    @property
    def S(self):
        return self.variant.num_states

    def states(self) -> np.ndarray:
        out = np.zeros((self.num_nodes, self.S))
        self._call("get_states", _ptr(out))
        return out

    def set_states(self, states) -> None:
        a = _f64(states).reshape(self.num_nodes, self.S)
        self._call("set_states", _ptr(a))

    def field_parameters(self, tangled=True) -> np.ndarray:
        F = self.variant.num_field_parameters
        out = np.zeros((self.num_nodes, F) if tangled else (F, self.num_nodes))
        self._call("get_field_parameters", _ptr(out), int(tangled))
        return out

    def set_field_parameters(self, fp, tangled=True) -> None:
        a = _f64(fp)
        self._call("set_field_parameters", _ptr(a), int(tangled))

    def field_states(self, tangled=True) -> np.ndarray:
        F = self.variant.num_field_states
        out = np.zeros((self.num_nodes, F) if tangled else (F, self.num_nodes))
        self._call("get_field_states", _ptr(out), int(tangled))
        return out

    def set_field_states(self, fs, tangled=True) -> None:
        a = _f64(fs)
        self._call("set_field_states", _ptr(a), int(tangled))

    def set_num_threads(self, n: int) -> None:
        self._call("set_num_threads", int(n))

    def forward_times(self, t) -> None:
        t = _f64(t)
        rc = self._call("forward_times", _ptr(t), len(t))
        self._check(rc)

    def forward(self, t, interval) -> None:
        self._check(self._call("forward", float(t), float(interval)))

    def solve(self, t, tangled=True) -> np.ndarray:
        t = _f64(t)
        F = self.variant.num_field_states
        out = np.zeros((len(t), self.num_nodes, F))
        out[0] = self.field_states(True)
        if not tangled:
            out = np.zeros((len(t), F, self.num_nodes))
            out[0] = self.field_states(False)
        self._check(self._call("solve", _ptr(out), _ptr(t), len(t), int(tangled)))
        return out

    def reset_default(self):
        self._call("reset_default")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._call("destroy")
            self._h = None


class OracleSystem(_SystemBase):
    """Restated ODESystemSolver (oracle/goss_oracle.cpp)."""

    def __init__(self, variant, solver, num_nodes, solver_params=None, per_node_state=False):
        super().__init__(variant, solver, num_nodes)
        self._L = lib()
        n, names, vals = _names(solver_params or {})
        self._h = ctypes.c_void_p()
        rc = self._L.oracle_system_create(variant.encode(), solver.encode(), self.num_nodes, n, names, _ptr(vals),
                                          int(per_node_state), ctypes.byref(self._h))
        if rc != 0:
            raise RuntimeError(self._L.oracle_last_error().decode())

    def _call(self, fn, *args):
        return getattr(self._L, f"oracle_system_{fn}")(self._h, *args)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self._L.oracle_last_error().decode())

    def set_parameter(self, name: str, value: float) -> None:
        self._check(self._call("set_parameter", name.encode(), float(value)))

    def counters(self) -> np.ndarray:
        out = (ctypes.c_long * 4)()
        self._call("counters", out)
        return np.array(list(out))

    def esdirk4o32_counters(self) -> np.ndarray:
        """Accepted steps, error-test rejections and right-hand-side evaluations of ESDIRK4O32."""
        out = (ctypes.c_long * 3)()
        self._call("esdirk4o32_counters", out)
        return np.array(list(out))


class RefSystem(_SystemBase):
    """The reference's ODESystemSolver compiled from its own sources."""

    def __init__(self, variant, solver, num_nodes, solver_params=None, model_params=None, analytic_jacobian=False):
        """analytic_jacobian: the generated plugin overrides ODE::compute_jacobian with its analytic Jacobian,
        as gotran's plugin does with generate_jacobian (codegeneration.py:186, 622-641)."""
        if variant in REFERENCE_MODELS:
            register_reference_models()  # the Variant metadata comes from the oracle's registry
        super().__init__(variant, solver, num_nodes)
        self._L = ref_lib(REFPIN_SO if variant in REFERENCE_MODELS else REF_SO)
        n, names, vals = _names(solver_params or {})
        m, mnames, mvals = _names(model_params or {})
        self._h = ctypes.c_void_p()
        name = variant + (":analytic" if analytic_jacobian else "")
        rc = self._L.ref_system_create(name.encode(), solver.encode(), self.num_nodes, n, names, _ptr(vals),
                                       m, mnames, _ptr(mvals), ctypes.byref(self._h))
        if rc != 0:
            raise RuntimeError(self._L.ref_last_error().decode())

    def _call(self, fn, *args):
        return getattr(self._L, f"ref_system_{fn}")(self._h, *args)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self._L.ref_last_error().decode())


    def rkf32_counters(self) -> np.ndarray:
        """nfevals, ndtsa, ndtsr of the system's RKF32 solver (serial mode)."""
        out = (ctypes.c_long * 3)()
        self._check(self._L.ref_system_rkf32_counters(self._h, out))
        return np.array(list(out))

    def esdirk4o32_counters(self) -> np.ndarray:
        """num_accepted, num_rejected, nfevals of the system's ESDIRK4O32 solver (serial mode)."""
        out = (ctypes.c_long * 3)()
        self._check(self._L.ref_system_esdirk4o32_counters(self._h, out))
        return np.array(list(out))


class RefModel:
    """A reference-side goss::ParameterizedODE instance (generated plugin)."""

    def __init__(self, variant: str, analytic_jacobian: bool = False):
        if variant in REFERENCE_MODELS:
            register_reference_models()
        self._L = ref_lib(REFPIN_SO if variant in REFERENCE_MODELS else REF_SO)
        self._h = ctypes.c_void_p()
        name = variant + (":analytic" if analytic_jacobian else "")
        if self._L.ref_model_create(name.encode(), ctypes.byref(self._h)) != 0:
            raise RuntimeError(self._L.ref_last_error().decode())
        self.variant = Variant(variant)

    def eval(self, y, t):
        y = _f64(y)
        out = np.zeros(len(y))
        self._L.ref_model_eval(self._h, _ptr(y), float(t), _ptr(out))
        return out

    def linearized_eval(self, y, t, only_linear=False):
        y = _f64(y)
        lin, rhs = np.zeros(len(y)), np.zeros(len(y))
        self._L.ref_model_linearized_eval(self._h, _ptr(y), float(t), _ptr(lin), _ptr(rhs), int(only_linear))
        return lin, rhs

    def compute_jacobian(self, y, t):
        y = _f64(y).copy()
        jac = np.zeros(len(y) * len(y))
        self._L.ref_model_compute_jacobian(self._h, _ptr(y), float(t), _ptr(jac))
        return jac.reshape(len(y), len(y))

    def set_parameter(self, name, value):
        self._L.ref_model_set_parameter(self._h, name.encode(), float(value))

    def linear_term(self, i):
        return self._L.ref_model_linear_term(self._h, i)

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.ref_model_destroy(self._h)
            self._h = None
