"""Device-resident operator-splitting caller: goss's ``DOLFINODESystemSolver``
(``python/goss/dolfinutils.py:394-810``) without DOLFIN.

In an operator-splitting PDE solver (monodomain, bidomain) the ODE step
reads the field states (the membrane potential V) from the PDE's solution
vector, advances every node's cell model, and writes the field states back.
The reference does both exchanges on the host through numpy index arrays
(``to_field_states`` / ``from_field_states``, ``dolfinutils.py:670-748``).
Here the solution vectors ``vs`` and ``vs_`` live in device memory (torch
tensors, float64) and the exchange runs in the library's gather and scatter
kernels (``goss_hip_system_gather_field_states_device`` /
``goss_hip_system_scatter_field_states_device``), so a step moves no data
over PCIe.

The mesh is described by what the reference derives from it:

* ``num_entities``: the mesh entities the ODE nodes live on (vertices for
  P1; ``dolfinutils.py:269-276``);
* ``domains``: the label of every entity (the reference's MeshFunction), one
  ODE per distinct label (``check_domains``, ``dolfinutils.py:359-391``);
* ``entity_to_dof``: the dof of field ``f`` on entity ``e`` at
  ``e * num_field_states + f`` (the reference's ``entity_to_dofs(V)``,
  ``dolfinutils.py:46-74``); the identity when omitted.

The dof maps are built as ``setup_dofs`` builds them
(``dolfinutils.py:213-346``): ``dof_maps[label][field] =
entity_to_dof[mesh_entities[label] * F + f]``.
"""
from __future__ import annotations

from typing import Dict, Iterator, Optional, Tuple, Union

import numpy as np

from . import solvers
from .ode import ParameterizedODE
from .systemsolver import ODESystemSolver


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("the device-resident splitting solver needs a GPU (no CPU fallback)")
    return torch


class DeviceODESystemSolver:
    """``DOLFINODESystemSolver(mesh, odes, domains, params)`` with the mesh
    replaced by ``num_entities``, ``domains`` and ``entity_to_dof``."""

    @staticmethod
    def default_parameters() -> dict:
        # dolfinutils.py:401-407; "precision" is this library's ("fp64", "fp32c", "fp32").
        return {"solver": "GRL1", "num_threads": 0, "precision": "fp64", "solver_parameters": {}}

    def __init__(self, num_entities: int, odes: Union[ParameterizedODE, Dict[int, ParameterizedODE]],
                 domains: Optional[np.ndarray] = None, entity_to_dof: Optional[np.ndarray] = None,
                 params: Optional[dict] = None):
        torch = _torch()
        self.parameters = self.default_parameters()
        self.parameters.update(params or {})
        if isinstance(odes, ParameterizedODE):
            odes = {0: odes}
        if not isinstance(odes, dict) or not odes:
            raise TypeError("expected a ParameterizedODE or a dict of them for the odes argument")
        num_entities = int(num_entities)
        if len(odes) > 1:
            # check_domains (dolfinutils.py:359-391).
            if domains is None:
                raise ValueError("expected domains when more than 1 ODE is given")
            domains = np.asarray(domains)
            if domains.shape != (num_entities,):
                raise ValueError(f"expected {num_entities} domain labels, got shape {domains.shape}")
            if sorted(set(domains.tolist())) != sorted(odes):
                raise ValueError("expected the distinct domains to be the labels of the ODEs")
        first = next(iter(odes.values()))
        for ode in odes.values():
            if ode.field_state_names != first.field_state_names:
                raise ValueError("expected all odes to have the same name and order of the field states")
        F = first.num_field_states
        if F == 0:
            raise ValueError("the ODEs have no field states to exchange")
        if entity_to_dof is None:
            entity_to_dof = np.arange(num_entities * F, dtype=np.int64)
        entity_to_dof = np.asarray(entity_to_dof, dtype=np.int64)
        if entity_to_dof.shape != (num_entities * F,):
            raise ValueError(f"expected {num_entities * F} entity dofs, got shape {entity_to_dof.shape}")
        solver_cls = solvers.solver_mapper[self.parameters["solver"]]
        solver = solver_cls()
        solver.update_parameters(self.parameters.get("solver_parameters", {}))

        self._odes = odes
        self._field_names = list(first.field_state_names)
        self._distinct_domains = sorted(odes)
        self.num_values = num_entities * F
        self._num_entities = num_entities
        self.mesh_entities: Dict[int, np.ndarray] = {}
        self.dof_maps: Dict[int, Dict[str, np.ndarray]] = {}
        self._ode_system_solvers: Dict[int, ODESystemSolver] = {}
        for label in self._distinct_domains:
            ents = (np.arange(num_entities) if domains is None
                    else np.nonzero(np.asarray(domains) == label)[0])
            self.mesh_entities[label] = ents
            self.dof_maps[label] = {name: entity_to_dof[ents * F + f] for f, name in enumerate(self._field_names)}
            system = ODESystemSolver(len(ents), solver.copy(), odes[label], self.parameters["precision"])
            system.num_threads = self.parameters["num_threads"]
            system.system.set_field_dofs(np.stack([self.dof_maps[label][n] for n in self._field_names]),
                                         self.num_values)
            self._ode_system_solvers[label] = system
        self._changed_field_parameters: Dict[str, np.ndarray] = {}
        self._saved_states: Dict[int, object] = {}

        # Current and previous solution (dolfinutils.py:516-518).
        self.vs = torch.zeros(self.num_values, dtype=torch.float64, device="cuda")
        self.vs_ = torch.zeros_like(self.vs)
        self.initial_conditions_to_field_states()
        self.from_field_states()
        self.vs_.copy_(self.vs)

    # -- properties (dolfinutils.py:617-668) -----------------------------------
    def solution_fields(self) -> Tuple[object, object]:
        """The previous and current solution vectors (device tensors)."""
        return self.vs_, self.vs

    @property
    def field_names(self):
        return self._field_names

    @property
    def num_field_states(self) -> int:
        return len(self._field_names)

    @property
    def num_distinct_domains(self) -> int:
        return len(self._distinct_domains)

    @property
    def ode_system_solvers(self) -> Dict[int, ODESystemSolver]:
        return self._ode_system_solvers

    # -- field parameters (dolfinutils.py:530-650) ------------------------------
    def set_field_parameter(self, name: str, values) -> None:
        """Per-entity values [num_entities] of a field parameter (the reference's
        ``ode.field_params[name]`` Function); applied at the next ``step`` or
        ``update_parameters``."""
        values = np.asarray(values.cpu() if hasattr(values, "cpu") else values, dtype=np.float64)
        if values.shape != (self._num_entities,):
            raise ValueError(f"expected {self._num_entities} values, got shape {values.shape}")
        if not any(name in ode.field_parameter_names for ode in self._odes.values()):
            raise ValueError(f"{name!r} is not a field parameter of the ODEs")
        self._changed_field_parameters[name] = values

    def update_parameters(self) -> None:
        """Apply changed field parameters to the domains whose ODE has them."""
        if not self._changed_field_parameters:
            return
        for label, system in self._ode_system_solvers.items():
            ode = system.ode
            names = [n for n in ode.field_parameter_names if n in self._changed_field_parameters]
            if not names:
                continue
            fp = system.field_parameters
            for n in names:
                fp[:, ode.field_parameter_names.index(n)] = self._changed_field_parameters[n][self.mesh_entities[label]]
            system.field_parameters = fp
        self._changed_field_parameters = {}

    # -- exchange (dolfinutils.py:670-748) --------------------------------------
    def _ready(self) -> None:
        # torch writes vs on its current stream; the library's calls run on the system's stream.
        _torch().cuda.current_stream().synchronize()

    def from_field_states(self) -> None:
        """vs[dofs] <- field states of every domain."""
        self._ready()
        for system in self._ode_system_solvers.values():
            system.system.scatter_field_states_device(self.vs.data_ptr(), self.num_values)

    def to_field_states(self) -> None:
        """Field states of every domain <- vs[dofs]."""
        self._ready()
        for system in self._ode_system_solvers.values():
            system.system.gather_field_states_device(self.vs.data_ptr(), self.num_values)

    def initial_conditions_to_field_states(self) -> None:
        """Every node's field states <- its ODE's initial conditions (dolfinutils.py:698-721)."""
        torch = _torch()
        values = np.zeros(self.num_values)
        for label, system in self._ode_system_solvers.items():
            ic = system.ode.get_ic()[system.ode.field_state_indices]
            for f, name in enumerate(self._field_names):
                values[self.dof_maps[label][name]] = ic[f]
        staged = torch.from_numpy(values).to("cuda")
        self._ready()
        for system in self._ode_system_solvers.values():
            system.system.gather_field_states_device(staged.data_ptr(), self.num_values)

    # -- stepping (dolfinutils.py:750-780) --------------------------------------
    def step(self, interval: Tuple[float, float]) -> None:
        """Solve on the time step (t0, t1): vs <- vs_, field states <- vs,
        changed field parameters, forward(t0, t1 - t0) on every domain,
        vs <- field states."""
        t0, t1 = interval
        self.vs.copy_(self.vs_)
        self.to_field_states()
        self.update_parameters()
        for system in self._ode_system_solvers.values():
            system.forward(t0, t1 - t0)
        self.from_field_states()

    def solve(self, interval: Tuple[float, float], dt: float) -> Iterator[Tuple[Tuple[float, float], object]]:
        """Steps of at most dt over interval, yielding ((t0, t1), vs) after each;
        vs_ takes vs before the next step."""
        T0, T1 = interval
        t0 = T0
        while t0 < T1 - 1e-12:
            t1 = min(t0 + dt, T1)
            self.step((t0, t1))
            yield (t0, t1), self.vs
            self.vs_.copy_(self.vs)
            t0 = t1

    # -- save / restore (dolfinutils.py:782-810) ---------------------------------
    def save_states(self) -> None:
        """Copy every domain's full states into device memory."""
        torch = _torch()
        for label, system in self._ode_system_solvers.items():
            if label not in self._saved_states:
                self._saved_states[label] = torch.empty((system.num_nodes, system.ode.num_states),
                                                        dtype=torch.float64, device="cuda")
            self._ready()
            system.system.get_states_device(self._saved_states[label].data_ptr())

    def restore_states(self) -> None:
        if not self._saved_states:
            raise RuntimeError("Cannot restore any states when none are saved.")
        self._ready()
        for label, system in self._ode_system_solvers.items():
            system.system.set_states_device(self._saved_states[label].data_ptr())
