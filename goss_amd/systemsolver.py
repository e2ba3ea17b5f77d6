"""goss's ``ODESystemSolver`` (``python/goss/systemsolver.py:15-134``) on the GPU.

``ODESystemSolver(num_nodes, solver, ode)`` advances ``ode`` over
``num_nodes`` nodes with ``solver``'s scheme and tunables, in the device
library's kernels (the reference runs ``cpp/goss/ODESystemSolver.cpp`` on the
host).  As in the reference, the system holds the solver and the model: their
tunables and shared parameter values are read again at every ``forward`` and
``solve``.  ``precision`` ("fp64", the reference's arithmetic, "fp32" or
"fp32c") is an extension of the reference's constructor.
"""
from __future__ import annotations

import numpy as np

from .ode import ParameterizedODE
from .solvers import ODESolver, _push_parameters
from .system import DeviceSystem


def has_shape(y: np.ndarray, xdim: int, ydim: int) -> bool:
    return y.shape == (xdim, ydim) or y.size == xdim * ydim


class ODESystemSolver:
    def __init__(self, num_nodes: int, solver: ODESolver, ode: ParameterizedODE, precision: str = "fp64"):
        solver.attach(ode)  # ODESystemSolver.cpp:55
        self._solver = solver
        self._ode = ode
        self._tangled_storage = True
        self._system = DeviceSystem(ode.device_variant(), solver.name, int(num_nodes), precision, solver.device_params())
        _push_parameters(self._system, ode, None)
        self._synced = (dict(solver.device_params()), ode.parameter_values)
        self._num_threads = 0
        # ODESystemSolver::reset_default (ODESystemSolver.cpp:192-216) with the model's current
        # parameter values for the field parameters.
        self._system.reset_default()

    def _sync(self) -> None:
        params, values = self._synced
        now = self._solver.device_params()
        if now != params:
            self._system.set_solver_params(**now)
        _push_parameters(self._system, self._ode, values)
        self._synced = (dict(now), self._ode.parameter_values)

    @property
    def system(self) -> DeviceSystem:
        """The underlying device system (goss_hip_system)."""
        return self._system

    @property
    def num_nodes(self) -> int:
        return self._system.num_nodes

    @property
    def solver(self) -> ODESolver:
        return self._solver

    @property
    def ode(self) -> ParameterizedODE:
        return self._ode

    @property
    def num_threads(self) -> int:
        return self._system.get_num_threads()

    @num_threads.setter
    def num_threads(self, num_threads: int) -> None:
        self._system.set_num_threads(num_threads)

    def reset_default(self) -> None:
        """Initial conditions and the current field-parameter values on every node."""
        self._sync()
        self._system.reset_default()

    @property
    def states(self) -> np.ndarray:
        """All states, node-major [num_nodes, num_states] (a copy)."""
        return self._system.states()

    @states.setter
    def states(self, states: np.ndarray) -> None:
        states = np.ascontiguousarray(states, dtype=np.float64)
        if not has_shape(states, self.num_nodes, self.ode.num_states):
            raise ValueError(f"Expected shape of states to be {(self.num_nodes, self.ode.num_states)}, "
                             f"got {states.shape}.")
        self._system.set_states(states)

    @property
    def field_parameters(self) -> np.ndarray:
        out = np.zeros((self.num_nodes, self.ode.num_field_parameters))
        if out.size:
            self._system.get_field_parameters(out, self._tangled_storage)
        return out

    @field_parameters.setter
    def field_parameters(self, field_parameters: np.ndarray) -> None:
        field_parameters = np.ascontiguousarray(field_parameters, dtype=np.float64)
        if not has_shape(field_parameters, self.num_nodes, self.ode.num_field_parameters):
            raise ValueError(("Expected shape of field parameters to be "
                              f"{(self.num_nodes, self.ode.num_field_parameters)}, "
                              f"got {field_parameters.shape}."))
        self._system.set_field_parameters(field_parameters, self._tangled_storage)

    @property
    def field_states(self) -> np.ndarray:
        out = np.zeros((self.num_nodes, self.ode.num_field_states))
        if out.size:
            self._system.get_field_states(out, self._tangled_storage)
        return out

    @field_states.setter
    def field_states(self, field_states: np.ndarray) -> None:
        field_states = np.ascontiguousarray(field_states, dtype=np.float64)
        if not has_shape(field_states, self.num_nodes, self.ode.num_field_states):
            raise ValueError(("Expected shape of field states to be "
                              f"{(self.num_nodes, self.ode.num_field_states)}, "
                              f"got {field_states.shape}."))
        self._system.set_field_states(field_states, self._tangled_storage)

    def forward(self, t: float, interval: float) -> None:
        self._sync()
        self._system.forward(float(t), float(interval))

    def solve(self, t: np.ndarray) -> np.ndarray:
        """Field states at each time point, [n_t, num_nodes, num_field_states]; row 0 holds the
        field states before the first step (systemsolver.py:107-134)."""
        t = np.asarray(t, dtype=np.float64)
        self._sync()
        field_states = np.zeros((t.size, self.num_nodes, self.ode.num_field_states))
        field_states[0, :, :] = self.field_states
        self._system.solve(t, field_states, self._tangled_storage)
        return field_states
