"""goss's ``ODE`` and ``ParameterizedODE`` (``python/goss/ode.py:40-231``).

A ``ParameterizedODE`` names a model variant of the device library: a gotran
``.ode`` file (or the name of a compiled-in variant) with a choice of field
states and field parameters.  A compiled-in variant of the same model text
and field selection is used when there is one; otherwise the variant is
compiled at run time into a model library (:mod:`goss_amd.jit`), as the
reference compiles its model class with gotran and cppyy.  Monitored values
are selected at run time among the model's named intermediates.

The model methods (``eval``, ``linearized_eval``, ``compute_jacobian``,
``eval_monitored``, ``lu_factorize``, ``forward_backward_subst``) are
evaluated on the GPU.  Shared parameter values live in this object and are
handed to the systems and solvers that use it.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
from pathlib import Path
from typing import Dict, List, NamedTuple, Optional, Sequence

import numpy as np

from ._lib import check, lib
from .system import ModelVariant

HERE = os.path.dirname(os.path.abspath(__file__))
ODE_DIR = os.path.join(HERE, "odes")
META = os.path.join(HERE, "csrc", "generated", "variants.json")

_dp = ctypes.POINTER(ctypes.c_double)


def _p(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class LinearizedEval(NamedTuple):
    rhs: np.ndarray
    linear: Optional[np.ndarray] = None


def _compiled() -> Dict[str, dict]:
    with open(META) as f:
        return json.load(f)


def _resolve_path(path) -> Path:
    p = Path(path)
    for cand in (p, p.with_suffix(".ode"), Path(ODE_DIR) / p, (Path(ODE_DIR) / p).with_suffix(".ode")):
        if cand.is_file():
            return cand
    raise FileNotFoundError(f"File {path} does not exist")


def _sha(text: str) -> str:
    return hashlib.sha256(text.encode()).hexdigest()


def resolve_variant(ode, field_states: Optional[Sequence[str]] = None,
                    field_parameters: Optional[Sequence[str]] = None) -> str:
    """The device variant for a model and a field selection (see module doc),
    compiled and registered with the library if it is a run-time variant."""
    name, pending = _find_variant(ode, field_states, field_parameters)
    if pending is not None:
        from . import jit

        jit.load(*pending)
    return name


def _find_variant(ode, field_states, field_parameters):
    """(variant name, None) for a compiled-in variant, or (name, (path,
    field states, field parameters)) for one compiled at its first use."""
    meta = _compiled()
    if isinstance(ode, str) and ode in meta and field_states is None and field_parameters is None:
        return ode, None
    path = _resolve_path(ode)
    text = path.read_text()
    fs, fp = list(field_states or []), list(field_parameters or [])
    from .codegen.ode_model import load_ode

    model = load_ode(str(path))
    bad = [s for s in fs if s not in model.state_names]
    if bad:
        raise ValueError(f"field states {bad} are not states of {model.name}")
    bad = [q for q in fp if q not in model.parameter_names]
    if bad:
        raise ValueError(f"field parameters {bad} are not parameters of {model.name}")
    digest = _sha(text)
    for name, m in meta.items():
        if m.get("source_sha256") == digest and m["field_states"] == fs and m["field_parameters"] == fp:
            return name, None
    from . import jit

    return jit.variant_name(str(path), fs, fp), (str(path), fs, fp)


class ODE:
    """An ODE model on the device library (``python/goss/ode.py:40-138``)."""

    def __init__(self, *args, **kwargs):
        if not args:
            raise ValueError("an ODE needs a model: a .ode file, a variant name or another ODE")
        src = args[0]
        if isinstance(src, ODE):
            self._copy_from(src)
            return
        if not isinstance(src, (str, os.PathLike)):
            raise ValueError(f"Unknown argument of type {type(src)}")
        self._variant, self._pending = _find_variant(src, kwargs.get("field_states"), kwargs.get("field_parameters"))
        if self._pending is None:
            self._info = ModelVariant(self._variant)
        else:
            from . import jit

            self._info = jit.metadata(*self._pending)
        self._params = self._info.default_parameters.copy()
        # code_params={"generate_jacobian": True} (codegeneration.py:186): the model's compute_jacobian is
        # its analytic Jacobian, and the implicit solvers use it.
        self._analytic_jacobian = bool((kwargs.get("code_params") or {}).get("generate_jacobian", False))
        self._monitored: List[int] = []
        monitored = kwargs.get("monitored") or []
        names = self._intermediate_names()
        bad = [m for m in monitored if m not in names]
        if bad:
            raise ValueError(f"monitored values {bad} are not intermediates of the model")
        self._monitored = [names.index(m) for m in monitored]

    def _copy_from(self, other: "ODE") -> None:
        self._variant = other._variant
        self._pending = other._pending
        self._info = other._info
        self._params = other._params.copy()
        self._analytic_jacobian = getattr(other, "_analytic_jacobian", False)
        self._monitored = list(other._monitored)

    def __repr__(self) -> str:
        return f"{self.__class__.__name__}({self._variant!r})"

    def __str__(self) -> str:
        return f"{self.__class__.__name__} with {self.num_states} states"

    def copy(self):
        return self.__class__(self)

    @property
    def variant(self) -> str:
        """Name of the device model variant."""
        return self._variant

    def device_variant(self) -> str:
        """The variant name, after compiling and registering a run-time
        variant at its first device use."""
        if self._pending is not None:
            from . import jit

            jit.load(*self._pending)
            self._pending = None
        return self._variant

    @property
    def num_states(self) -> int:
        return self._info.num_states

    @property
    def is_dae(self) -> bool:
        """ODE::is_dae (python/goss/ode.py:74-77): the model has algebraic states
        (``alg_<X>_0 = expr`` in the .ode file)."""
        return bool(self._info.is_dae)

    @property
    def differential_states(self) -> np.ndarray:
        """ODE::differential_states (ODE.h:85-89): the diagonal of the mass
        matrix, 0 for the algebraic states."""
        return np.array(self._info.differential_states, dtype=np.uint8)

    def get_ic(self) -> np.ndarray:
        return self._info.initial_conditions.copy()

    def _check_state_input(self, states) -> np.ndarray:
        if len(states) != self.num_states:
            raise ValueError("State array has to be of same size as the number"
                             f"of states, got {len(states)}, expected {self.num_states}")
        return np.ascontiguousarray(states, dtype=np.float64)

    def _points(self, states, time):
        y = np.ascontiguousarray(np.atleast_2d(np.asarray(states, dtype=np.float64)))
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(time, dtype=np.float64), (y.shape[0],)))
        return y, t

    def eval(self, states, time: float) -> np.ndarray:
        """Right-hand side at one state (ODE.h:55)."""
        y = self._check_state_input(states)
        return self.eval_many(y[None, :], np.array([time]))[0]

    def eval_many(self, states: np.ndarray, times) -> np.ndarray:
        """Right-hand sides at n states [n, S] and times [n] in one device launch."""
        y, t = self._points(states, times)
        out = np.zeros_like(y)
        check(lib().goss_hip_model_eval(self.device_variant().encode(), _p(self._params), _p(y), _p(t), y.shape[0], _p(out)))
        return out

    def eval_component(self, component: int, states, time: float) -> float:
        """Component `component` of the right-hand side (ODE.h:58)."""
        if not isinstance(component, (int, np.integer)) or component < 0:
            raise TypeError("component must be a non-negative integer")
        y = self._check_state_input(states)
        if component >= self.num_states:
            raise RuntimeError("throw - Index out of range")
        return float(self.eval(y, time)[component])

    @property
    def has_analytic_jacobian(self) -> bool:
        """Whether the model was made with code_params={"generate_jacobian": True}."""
        return getattr(self, "_analytic_jacobian", False)

    def compute_jacobian(self, states, time: float) -> np.ndarray:
        """The Jacobian, [S, S]: forward differences (ODE.cpp:47-67), or the generated analytic
        Jacobian for a model made with code_params={"generate_jacobian": True} (codegeneration.py:622-641)."""
        y = self._check_state_input(states)
        S = self.num_states
        jac = np.zeros((1, S * S))
        t = np.array([time], dtype=np.float64)
        fn = (lib().goss_hip_model_compute_analytic_jacobian if self.has_analytic_jacobian
              else lib().goss_hip_model_compute_jacobian)
        check(fn(self.device_variant().encode(), _p(self._params), _p(y[None, :].copy()), _p(t), 1, _p(jac)))
        return jac.reshape(S, S)

    def linearized_eval(self, states, time: float, only_linear: bool = False) -> LinearizedEval:
        """Right-hand side and the diagonal of the Jacobian (ODE.h:61)."""
        y = self._check_state_input(states)
        S = self.num_states
        linear = np.zeros((1, S))
        rhs = np.zeros((1, S))
        t = np.array([time], dtype=np.float64)
        check(lib().goss_hip_model_linearized_eval(self.device_variant().encode(), _p(self._params), _p(y[None, :].copy()),
                                                   _p(t), 1, int(only_linear), _p(linear), _p(rhs)))
        if only_linear:
            return LinearizedEval(rhs=rhs[0])
        return LinearizedEval(linear=linear[0], rhs=rhs[0])

    def lu_factorize(self, mat: np.ndarray) -> None:
        """In-place no-pivot LU (ODE.cpp:69-100)."""
        S = self.num_states
        a = np.ascontiguousarray(mat, dtype=np.float64).reshape(1, S * S)
        check(lib().goss_hip_lu_factorize(S, 1, _p(a)))
        mat[...] = a.reshape(mat.shape)

    def forward_backward_subst(self, mat: np.ndarray, b: np.ndarray, x: np.ndarray) -> None:
        """Solve LU x = b with a factorised matrix (ODE.cpp:102-132)."""
        S = self.num_states
        a = np.ascontiguousarray(mat, dtype=np.float64).reshape(1, S * S)
        bb = np.ascontiguousarray(b, dtype=np.float64).reshape(1, S)
        xx = np.zeros((1, S))
        check(lib().goss_hip_forward_backward_subst(S, 1, _p(a), _p(bb), _p(xx)))
        x[...] = xx.reshape(x.shape)

    def _intermediate_names(self) -> List[str]:
        if self._pending is not None:
            return list(self._info.intermediate_names)
        n = ctypes.c_int()
        names = ctypes.POINTER(ctypes.c_char_p)()
        check(lib().goss_hip_model_intermediates(self._variant.encode(), ctypes.byref(n), ctypes.byref(names)))
        return [names[i].decode() for i in range(n.value)]


class ParameterizedODE(ODE):
    """``goss.ParameterizedODE`` (``python/goss/ode.py:141-231``).

    ``ParameterizedODE(ode, field_states=None, field_parameters=None,
    monitored=None)`` with ``ode`` a gotran ``.ode`` file or a compiled-in
    variant name; ``ParameterizedODE(num_states=.., num_parameters=..)`` gives
    an empty model with those sizes (no right-hand side)."""

    def __init__(self, *args, **kwargs):
        if not args and "num_states" in kwargs and "num_parameters" in kwargs:
            self._variant = None
            self._pending = None
            self._info = None
            self._empty = dict(num_states=int(kwargs["num_states"]), num_parameters=int(kwargs["num_parameters"]),
                               num_field_states=int(kwargs.get("num_field_states", 0)),
                               num_field_parameters=int(kwargs.get("num_field_parameters", 0)),
                               num_monitored=int(kwargs.get("num_monitored", 0)))
            self._params = np.zeros(self._empty["num_parameters"])
            self._monitored = []
            return
        self._empty = None
        super().__init__(*args, **kwargs)

    def _copy_from(self, other: "ODE") -> None:
        super()._copy_from(other)
        self._empty = getattr(other, "_empty", None)

    def _size(self, key: str, value):
        return self._empty[key] if self._empty is not None else value()

    @property
    def num_states(self) -> int:
        return self._size("num_states", lambda: self._info.num_states)

    @property
    def num_parameters(self) -> int:
        return self._size("num_parameters", lambda: self._info.num_parameters)

    @property
    def num_field_parameters(self) -> int:
        return self._size("num_field_parameters", lambda: self._info.num_field_parameters)

    @property
    def num_field_states(self) -> int:
        return self._size("num_field_states", lambda: self._info.num_field_states)

    @property
    def num_monitored(self) -> int:
        return self._size("num_monitored", lambda: len(self._monitored))

    def eval_monitored(self, states, time: float) -> np.ndarray:
        """The monitored intermediates at one state (ParameterizedODE.h:90)."""
        y = self._check_state_input(states)
        return self.monitored_values(y[None, :], np.array([time]))[0]

    def monitored_values(self, states: np.ndarray, time) -> np.ndarray:
        """Monitored intermediates along a trace: states [n_t, S], time [n_t] (ParameterizedODE.h:93-105)."""
        y, t = self._points(states, time)
        if y.shape[1] != self.num_states:
            raise ValueError(f"states must have {self.num_states} columns")
        out = np.zeros((y.shape[0], len(self._monitored)))
        if not self._monitored:
            return out
        idx = (ctypes.c_int * len(self._monitored))(*self._monitored)
        check(lib().goss_hip_model_eval_monitored(self.device_variant().encode(), _p(self._params), len(self._monitored), idx,
                                                  _p(y), _p(t), y.shape[0], _p(out)))
        return out

    def _pindex(self, name: str) -> int:
        try:
            return self.parameter_names.index(name)
        except ValueError:
            # ParameterizedODE.h:185-195
            raise RuntimeError(f'"{name}" is not a parameter in  {self._variant}') from None

    def set_parameter(self, name: str, value: float) -> None:
        self._params[self._pindex(name)] = float(value)

    def get_parameter(self, name: str) -> float:
        return float(self._params[self._pindex(name)])

    @property
    def parameters(self) -> Dict[str, float]:
        return {name: self.get_parameter(name) for name in self.parameter_names}

    @property
    def parameter_values(self) -> np.ndarray:
        """All parameter values in declaration order (a copy)."""
        return self._params.copy()

    @property
    def state_names(self) -> List[str]:
        return list(self._info.state_names) if self._info else []

    @property
    def field_state_names(self) -> List[str]:
        return list(self._info.field_state_names) if self._info else []

    @property
    def parameter_names(self) -> List[str]:
        return list(self._info.parameter_names) if self._info else []

    @property
    def field_parameter_names(self) -> List[str]:
        return list(self._info.field_parameter_names) if self._info else []

    @property
    def field_state_indices(self) -> List[int]:
        return list(self._info.field_state_indices) if self._info else []

    @property
    def monitored_names(self) -> List[str]:
        if not self._monitored:
            return []
        names = self._intermediate_names()
        return [names[i] for i in self._monitored]
