"""MI355X implementation of goss's ODESystemSolver hot path.

The package mirrors the Python API of goss (python/goss/__init__.py):
ParameterizedODE, the solver classes and ODESystemSolver, backed by
libgoss_hip.so.  The code generator (goss_amd.codegen) turns .ode models
into the C++/HIP sources compiled into the library.
"""
__version__ = "0.1.0"
