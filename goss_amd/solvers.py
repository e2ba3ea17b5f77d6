"""goss's solver classes (``python/goss/solvers.py:14-491``) over the device library.

Each class carries the reference solver's public tunables
(``default_parameters``, ``GRL1.h:39``, ``ImplicitODESolver.h:36-41``,
``ImplicitEuler.h:36-37``, ``ThetaSolver.h:35-37``, ``ESDIRK23a.h:37-38``)
and its internal time step (``ODESolver.h:119-128``).  Attached to a
:class:`~goss_amd.ode.ParameterizedODE`, ``forward(y, t, interval)`` and
``solve(t, y0)`` advance one state vector on the GPU through a one-node
system; the solver's per-node state (Newton eta, step-size history, RKF32's
step guess) carries from call to call as the reference solver object's does,
and tunables or model parameters changed between calls apply to the next one.
An :class:`~goss_amd.systemsolver.ODESystemSolver` reads the same tunables.
"""
from __future__ import annotations

import abc
from enum import Enum
from typing import Any, Dict, Optional

import numpy as np

from .ode import ODE

_DEVICE_FIELDS = ("delta", "eta_0", "kappa", "relative_tolerance", "max_iterations",
                  "max_relative_previous_residual", "always_recompute_jacobian",
                  "num_refinements_without_always_recomputing_jacobian", "min_dt", "theta")


class ODESolver(abc.ABC):
    """Base class (``python/goss/solvers.py:14-214``): ``Solver()``,
    ``Solver(ode)`` or ``solver.attach(ode)``."""

    adaptive = False

    def __init__(self, *args, **kwargs):
        self._params: Dict[str, Any] = self.__class__.default_parameters()
        self._ldt = -1.0  # ODESolver::_ldt (ODESolver.h:137)
        self._atol, self._rtol, self._iord = 1.0e-5, 1.0e-8, 1  # AdaptiveExplicitSolver.cpp:28-31
        self._ode: Optional[ODE] = None
        self._system = None
        self._synced = None
        if args:
            if isinstance(args[0], ODE):
                self.attach(args[0])
            elif isinstance(args[0], ODESolver):
                self._copy_settings(args[0])
            else:
                raise ValueError(f"Unknown argument of type {type(args[0])}")

    @property
    def name(self) -> str:
        return self.__class__.__name__

    def __str__(self) -> str:
        return self.name

    def __repr__(self) -> str:
        return f"{self.name}()"

    def _copy_settings(self, other: "ODESolver") -> None:
        self._params = dict(other._params)
        self._ldt = other._ldt
        self._atol, self._rtol, self._iord = other._atol, other._rtol, other._iord
        if other._ode is not None:
            self.attach(other._ode.copy())

    def reset(self) -> None:
        """Fresh solver state (ODESolver::reset): the next forward starts over."""
        if self._system is not None:
            self._system.close()
        self._system = None
        self._synced = None

    def attach(self, ode: ODE) -> None:
        self._ode = ode
        self.reset()

    @property
    def is_adaptive(self) -> bool:
        return self.adaptive

    @staticmethod
    def default_parameters() -> Dict[str, Any]:
        return {}

    def update_parameters(self, parameters: Dict[str, Any]) -> None:
        for k, v in parameters.items():
            self.set_parameter(k, v)

    @property
    def internal_time_step(self) -> float:
        return self._ldt

    @internal_time_step.setter
    def internal_time_step(self, time_step: float) -> None:
        self._ldt = float(time_step)

    @property
    def parameters(self) -> Dict[str, Any]:
        return {name: self.get_parameter(name) for name in self.__class__.default_parameters()}

    def get_parameter(self, name: str) -> Any:
        if name not in self._params:
            raise AttributeError(name)
        return self._params[name]

    def set_parameter(self, name: str, value: Any) -> None:
        parameters = self.__class__.default_parameters()
        if name not in parameters:
            raise KeyError(f"Invalid parameter {name}, expected one of {parameters.keys()}")
        if not isinstance(value, type(parameters[name])):
            raise TypeError(f"Expected parameter {name} to be of type{type(parameters[name])}, got {type(value)}")
        self._params[name] = value

    def get_ode(self) -> ODE:
        if self._ode is None:
            raise RuntimeError("no ODE is attached to the solver")
        return self._ode

    def copy(self) -> "ODESolver":
        return self.__class__(self)

    @property
    def num_states(self) -> int:
        return 0 if self._ode is None else self._ode.num_states

    # -- device settings ------------------------------------------------------
    def device_params(self) -> Dict[str, Any]:
        """The tunables as goss_hip_solver_params fields."""
        d: Dict[str, Any] = {"internal_time_step": self._ldt}
        for k in _DEVICE_FIELDS:
            if k in self._params:
                d[k] = self._params[k]
        if self.adaptive:
            d["atol"], d["rtol"] = self._atol, self._rtol
        if isinstance(self, ImplicitODESolver) and self._ode is not None and self._ode.has_analytic_jacobian:
            d["jacobian"] = "analytic"  # the model's ODE::compute_jacobian is its analytic Jacobian
        return d

    def _one_node(self):
        """The one-node device system of forward/solve, with the current
        tunables and model parameters pushed into it."""
        from .system import DeviceSystem

        ode = self.get_ode()
        if self._system is None:
            self._system = DeviceSystem(ode.device_variant(), self.name, 1, "fp64", self.device_params())
            self._synced = (dict(self.device_params()), np.array(ode.parameter_values), )
            _push_parameters(self._system, ode, None)
        else:
            params, values = self._synced
            now = self.device_params()
            if now != params:
                self._system.set_solver_params(**now)
            _push_parameters(self._system, ode, values)
            self._synced = (dict(now), np.array(ode.parameter_values))
        return self._system

    def forward(self, y: np.ndarray, t: float, interval: float) -> None:
        """Advance y (mutated in place) from t by interval (ODESolver.h:75)."""
        s = self._one_node()
        s.set_states(np.asarray(y, dtype=np.float64).reshape(1, -1))
        s.forward(float(t), float(interval))
        y[...] = s.states().reshape(np.shape(y))

    def solve(self, t: np.ndarray, y0: Optional[np.ndarray] = None) -> np.ndarray:
        """States at each time point, [n_t, S] (ODESolver.h:78-98; row 0 = y0)."""
        if y0 is None:
            y0 = self.get_ode().get_ic()
        t = np.asarray(t, dtype=np.float64)
        y = np.zeros((t.size, self.num_states))
        y[0, :] = y0
        cur = np.array(y0, dtype=np.float64)
        for it in range(1, t.size):
            self.forward(cur, t[it - 1], t[it] - t[it - 1])
            y[it] = cur
        return y

    def _counters(self) -> np.ndarray:
        if self._system is None:
            return np.zeros(4, dtype=np.int64)
        return self._system.solver_counters()


def _push_parameters(system, ode: ODE, previous: Optional[np.ndarray]) -> None:
    """Model parameter values that differ from what the system holds."""
    values = ode.parameter_values
    base = previous if previous is not None else ode._info.default_parameters
    names = ode.parameter_names
    for k in np.nonzero(values != base)[0]:
        system.set_parameter(names[k], float(values[k]))


class ExplicitEuler(ODESolver):
    pass


class RL1(ODESolver):
    pass


class RL2(ODESolver):
    pass


class RK2(ODESolver):
    pass


class RK4(ODESolver):
    pass


class GRL1(ODESolver):
    @staticmethod
    def default_parameters() -> Dict[str, Any]:
        names = ODESolver.default_parameters()
        names.update({"delta": 1e-8})
        return names


class GRL2(GRL1):
    pass


class ImplicitODESolver(ODESolver, abc.ABC):
    @staticmethod
    def default_parameters() -> Dict[str, Any]:
        names = ODESolver.default_parameters()
        names.update({"eta_0": 1.0, "kappa": 0.1, "relative_tolerance": 1e-12, "max_iterations": 30,
                      "max_relative_previous_residual": 0.01, "always_recompute_jacobian": False})
        return names

    def num_jac_comp(self) -> int:
        """Number of Jacobian evaluations so far (ImplicitODESolver.h:65-68)."""
        return int(self._counters()[1])


class ImplicitEuler(ImplicitODESolver):
    @staticmethod
    def default_parameters() -> Dict[str, Any]:
        names = ImplicitODESolver.default_parameters()
        names.update({"num_refinements_without_always_recomputing_jacobian": 2, "min_dt": 0.0001})
        return names


class BasicImplicitEuler(ImplicitODESolver):
    """BasicImplicitEuler (BasicImplicitEuler.cpp:62-181): full Newton with a new
    Jacobian at every iteration.  The reference's Python layer marks it as not
    working as expected (python/goss/solvers.py:279-282) and leaves it out of
    its solver lists; so does this package."""


class ThetaSolver(ImplicitODESolver):
    @staticmethod
    def default_parameters() -> Dict[str, Any]:
        names = ImplicitODESolver.default_parameters()
        names.update({"num_refinements_without_always_recomputing_jacobian": 2, "min_dt": 0.0001, "theta": 0.5})
        return names


class AdaptiveSolver(ODESolver, abc.ABC):
    adaptive = True

    def get_num_accepted(self) -> int:
        c = self._counters()
        return int(c[3])

    def get_num_rejected(self) -> int:
        return int(self._counters()[2])

    def set_tol(self, atol: float, rtol: float = 1e-8) -> None:
        self._atol, self._rtol = float(atol), float(rtol)

    def set_iord(self, iord: int) -> None:
        self._iord = int(iord)

    @property
    def atol(self) -> float:
        return self._atol

    @property
    def rtol(self) -> float:
        return self._rtol

    @property
    def iord(self) -> int:
        return self._iord


class AdaptiveExplicitSolver(AdaptiveSolver, abc.ABC):
    pass


class RKF32(AdaptiveExplicitSolver):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._iord = 3  # RKF32.cpp:39

    @property
    def nfevals(self) -> int:
        return int(self._counters()[1])

    @property
    def ndtsa(self) -> int:
        return int(self._counters()[3])

    @property
    def ndtsr(self) -> int:
        return int(self._counters()[2])


class AdaptiveImplicitSolver(ImplicitODESolver, AdaptiveSolver, abc.ABC):
    adaptive = True


class ESDIRK23a(AdaptiveImplicitSolver):
    """ESDIRK23a keeps its own tolerance (tol = 1e-5, ESDIRK23a.cpp:34); set_tol is recorded only."""

    @staticmethod
    def default_parameters() -> Dict[str, Any]:
        names = AdaptiveImplicitSolver.default_parameters()
        names.update({"num_refinements_without_always_recomputing_jacobian": 2, "min_dt": 0.001})
        return names

    @property
    def ndtsa(self) -> int:
        return int(self._counters()[3])

    @property
    def ndtsr(self) -> int:
        return int(self._counters()[2])


class ESDIRK4O32(AdaptiveImplicitSolver):
    """ESDIRK4O32 (ESDIRK4O32.cpp:92-298) as the reference computes it, error
    test with atol / rtol (set_tol) and its two-buffer solution update.  The
    reference's Python layer marks it as not working as expected
    (python/goss/solvers.py:389-392) and leaves it out of its solver lists; so
    does this package."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._iord = 3  # ESDIRK4O32.cpp:38

    @property
    def ndtsa(self) -> int:
        return int(self._counters()[3])

    @property
    def ndtsr(self) -> int:
        return int(self._counters()[2])


class GOSSSolvers(Enum):
    ExplicitEuler = "ExplicitEuler"
    RK2 = "RK2"
    RK4 = "RK4"
    RL1 = "RL1"
    RL2 = "RL2"
    GRL1 = "GRL1"
    GRL2 = "GRL2"
    ImplicitEuler = "ImplicitEuler"
    ThetaSolver = "ThetaSolver"
    RKF32 = "RKF32"
    ESDIRK23a = "ESDIRK23a"


class GOSSImplicitSolvers(Enum):
    ImplicitEuler = "ImplicitEuler"
    ThetaSolver = "ThetaSolver"
    ESDIRK23a = "ESDIRK23a"


class GOSSExplicitSolvers(Enum):
    ExplicitEuler = "ExplicitEuler"
    RK2 = "RK2"
    RK4 = "RK4"
    RL1 = "RL1"
    RL2 = "RL2"
    GRL1 = "GRL1"
    GRL2 = "GRL2"
    RKF32 = "RKF32"


class GOSSNonAdaptiveSolvers(Enum):
    ExplicitEuler = "ExplicitEuler"
    RK2 = "RK2"
    RK4 = "RK4"
    RL1 = "RL1"
    RL2 = "RL2"
    GRL1 = "GRL1"
    GRL2 = "GRL2"
    ImplicitEuler = "ImplicitEuler"
    ThetaSolver = "ThetaSolver"


class GOSSAdaptiveSolvers(Enum):
    RKF32 = "RKF32"
    ESDIRK23a = "ESDIRK23a"


solver_mapper = {
    "ExplicitEuler": ExplicitEuler,
    "RK2": RK2,
    "RK4": RK4,
    "RL1": RL1,
    "RL2": RL2,
    "GRL1": GRL1,
    "GRL2": GRL2,
    "ImplicitEuler": ImplicitEuler,
    "ThetaSolver": ThetaSolver,
    "RKF32": RKF32,
    "ESDIRK23a": ESDIRK23a,
}
