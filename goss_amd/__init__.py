"""MI355X implementation of goss's ODESystemSolver hot path.

The Python API mirrors goss's (``python/goss/__init__.py``):

* :class:`ODE`, :class:`ParameterizedODE` (``goss_amd.ode``) -- a gotran
  ``.ode`` model with its field states, field parameters and monitored values;
* ``goss_amd.solvers`` -- the solver classes with their public tunables;
* :class:`ODESystemSolver` (``goss_amd.systemsolver``) -- one model advanced
  over many nodes.

The work runs in the HIP kernels of ``goss_amd/libgoss_hip.so`` (C ABI:
``include/goss_hip.h``; lower-level handle: :class:`goss_amd.system.DeviceSystem`).
The code generator (:mod:`goss_amd.codegen`) turns ``.ode`` models into the
sources compiled into the library, at build time or at run time
(:mod:`goss_amd.jit`).  There is no CPU fallback: without a GPU the calls
that compute fail.
"""
from . import solvers
from .ode import ODE, LinearizedEval, ParameterizedODE
from .systemsolver import ODESystemSolver

__version__ = "0.2.0"


def has_openmp() -> bool:
    """Whether ``num_threads`` changes how a system runs: it does not here --
    the device path uses no host threads (the value is recorded only)."""
    return False


__all__ = ["ODE", "ParameterizedODE", "LinearizedEval", "ODESystemSolver", "solvers", "has_openmp",
           "__version__"]
