"""ctypes binding of libgoss_hip.so (the C ABI declared in include/goss_hip.h).

The library is loaded from the package directory, next to this file.  There
is no fallback: if the library is missing, or no GPU is visible when a
system is created, the call raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgoss_hip.so")

SOLVERS = {
    "ExplicitEuler": 0,
    "RK2": 1,
    "RK4": 2,
    "RL1": 3,
    "RL2": 4,
    "GRL1": 5,
    "GRL2": 6,
    "ImplicitEuler": 7,
    "ESDIRK23a": 8,
    "RKF32": 9,
    "ThetaSolver": 10,
    "BasicImplicitEuler": 11,
    "ESDIRK4O32": 12,
}
PRECISIONS = {"fp64": 0, "fp32": 1, "fp32c": 2}
JACOBIANS = {"fd": 0, "fd_fp32": 1, "analytic": 2}

# Functions whose C signatures this module binds (checked by the tests against
# the exported symbols).
EXPORTS = (
    "goss_hip_last_error", "goss_hip_device_count", "goss_hip_model_count", "goss_hip_model_name",
    "goss_hip_model_info_get", "goss_hip_model_load", "goss_hip_model_eval", "goss_hip_model_linearized_eval",
    "goss_hip_model_compute_jacobian", "goss_hip_model_compute_analytic_jacobian", "goss_hip_model_intermediates", "goss_hip_model_eval_monitored",
    "goss_hip_lu_factorize", "goss_hip_forward_backward_subst",
    "goss_hip_solver_params_default", "goss_hip_solver_is_adaptive",
    "goss_hip_system_create", "goss_hip_system_destroy", "goss_hip_system_forward",
    "goss_hip_system_forward_times", "goss_hip_system_forward_times_noalloc", "goss_hip_system_solve", "goss_hip_system_solve_async", "goss_hip_host_alloc",
    "goss_hip_host_free", "goss_hip_system_get_field_states",
    "goss_hip_system_set_field_states", "goss_hip_system_get_field_state_components",
    "goss_hip_system_set_field_state_components", "goss_hip_system_get_field_parameters",
    "goss_hip_system_set_field_parameters", "goss_hip_system_get_field_states_device",
    "goss_hip_system_set_field_states_device", "goss_hip_system_get_states", "goss_hip_system_set_states",
    "goss_hip_system_get_states_device", "goss_hip_system_set_states_device", "goss_hip_system_set_field_dofs",
    "goss_hip_system_gather_field_states_device", "goss_hip_system_scatter_field_states_device",
    "goss_hip_system_reset_default", "goss_hip_system_set_parameter", "goss_hip_system_get_parameter",
    "goss_hip_system_set_num_threads", "goss_hip_system_get_num_threads", "goss_hip_system_num_nodes",
    "goss_hip_system_set_steps_per_launch", "goss_hip_system_synchronize", "goss_hip_system_stream",
    "goss_hip_system_set_kernel_timing", "goss_hip_system_get_kernel_timing",
    "goss_hip_system_set_solver_params", "goss_hip_system_get_solver_params",
    "goss_hip_system_get_solver_counters", "goss_hip_adjoint_buffer_sizes", "goss_hip_adjoint_record",
    "goss_hip_adjoint_backward", "goss_hip_batch_create", "goss_hip_batch_destroy", "goss_hip_batch_set_config",
    "goss_hip_batch_get_config", "goss_hip_batch_forward", "goss_hip_batch_autotune",
)


class SolverParams(ctypes.Structure):
    _fields_ = [
        ("internal_time_step", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("eta_0", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("relative_tolerance", ctypes.c_double),
        ("max_relative_previous_residual", ctypes.c_double),
        ("min_dt", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("rtol", ctypes.c_double),
        ("max_iterations", ctypes.c_int),
        ("always_recompute_jacobian", ctypes.c_int),
        ("num_refinements_without_always_recomputing_jacobian", ctypes.c_int),
        ("jacobian", ctypes.c_int),
        ("theta", ctypes.c_double),
    ]


class ModelInfo(ctypes.Structure):
    _fields_ = [
        ("num_states", ctypes.c_int),
        ("num_parameters", ctypes.c_int),
        ("num_field_states", ctypes.c_int),
        ("num_field_parameters", ctypes.c_int),
        ("state_names", ctypes.POINTER(ctypes.c_char_p)),
        ("parameter_names", ctypes.POINTER(ctypes.c_char_p)),
        ("field_state_names", ctypes.POINTER(ctypes.c_char_p)),
        ("field_parameter_names", ctypes.POINTER(ctypes.c_char_p)),
        ("field_state_indices", ctypes.POINTER(ctypes.c_int)),
        ("field_parameter_indices", ctypes.POINTER(ctypes.c_int)),
        ("default_parameters", ctypes.POINTER(ctypes.c_double)),
        ("initial_conditions", ctypes.POINTER(ctypes.c_double)),
        ("linear_terms", ctypes.POINTER(ctypes.c_ubyte)),
        ("differential_states", ctypes.POINTER(ctypes.c_ubyte)),
        ("is_dae", ctypes.c_int),
    ]


BATCH_MAX_POPULATIONS = 16


class BatchConfig(ctypes.Structure):
    _fields_ = [
        ("block_size", ctypes.c_uint32),
        ("nodes_per_lane", ctypes.c_uint32),
        ("span", ctypes.c_double),
        ("order", ctypes.c_uint32 * BATCH_MAX_POPULATIONS),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_lib: Optional[ctypes.CDLL] = None


class GossHipError(RuntimeError):
    """A failure reported by libgoss_hip.so (message from goss_hip_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def lib() -> ctypes.CDLL:
    """The product library, goss_amd/libgoss_hip.so."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C goss_amd/csrc")
    _lib = load(LIB_PATH)
    return _lib


_loaded = {}


def _share_the_hsa_runtime_with_torch() -> None:
    """Load torch's ROCm libraries before this library when torch is installed.

    The torch wheel carries its own HSA runtime, with the soname of ROCm's
    (libhsa-runtime64.so.1) but required by torch under another name.  Loaded
    after this library, torch would bring up a second HSA runtime in the
    process, which cannot open the GPU again (torch then reports no device).
    Loaded first, this library's HIP runtime binds to that same HSA runtime,
    so torch tensors (device buffers of the Python API) and the library's
    kernels use one device context."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(path: str) -> ctypes.CDLL:
    """A library with the C ABI of include/goss_hip.h at `path`, with its
    signatures bound (the product library, or a test build of it)."""
    if path in _loaded:
        return _loaded[path]
    _share_the_hsa_runtime_with_torch()
    L = ctypes.CDLL(path)
    L.goss_hip_last_error.restype = ctypes.c_char_p
    L.goss_hip_model_name.restype = ctypes.c_char_p
    L.goss_hip_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    L.goss_hip_model_info_get.argtypes = [ctypes.c_char_p, ctypes.POINTER(ModelInfo)]
    if hasattr(L, "goss_hip_model_load"):
        L.goss_hip_model_load.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
        L.goss_hip_model_eval.argtypes = [ctypes.c_char_p, _dp, _dp, _dp, ctypes.c_uint64, _dp]
        L.goss_hip_model_linearized_eval.argtypes = [ctypes.c_char_p, _dp, _dp, _dp, ctypes.c_uint64, ctypes.c_int,
                                                     _dp, _dp]
        L.goss_hip_model_compute_jacobian.argtypes = [ctypes.c_char_p, _dp, _dp, _dp, ctypes.c_uint64, _dp]
        L.goss_hip_model_compute_analytic_jacobian.argtypes = [ctypes.c_char_p, _dp, _dp, _dp, ctypes.c_uint64, _dp]
        L.goss_hip_model_intermediates.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                                   ctypes.POINTER(ctypes.POINTER(ctypes.c_char_p))]
        L.goss_hip_model_eval_monitored.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                    _dp, _dp, ctypes.c_uint64, _dp]
        L.goss_hip_lu_factorize.argtypes = [ctypes.c_int, ctypes.c_uint64, _dp]
        L.goss_hip_forward_backward_subst.argtypes = [ctypes.c_int, ctypes.c_uint64, _dp, _dp, _dp]
    L.goss_hip_solver_params_default.argtypes = [ctypes.c_int, ctypes.POINTER(SolverParams)]
    L.goss_hip_system_create.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(SolverParams), ctypes.c_int,
                                         ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]
    V = ctypes.c_void_p
    L.goss_hip_system_destroy.argtypes = [V]
    L.goss_hip_system_forward.argtypes = [V, ctypes.c_double, ctypes.c_double]
    L.goss_hip_system_forward_times.argtypes = [V, _dp, ctypes.c_uint64]
    L.goss_hip_system_forward_times_noalloc.argtypes = [V, _dp, ctypes.c_uint64]
    L.goss_hip_system_solve.argtypes = [V, _dp, _dp, ctypes.c_uint64, ctypes.c_int]
    L.goss_hip_system_solve_async.argtypes = [V, ctypes.c_void_p, _dp, ctypes.c_uint64, ctypes.c_int]
    L.goss_hip_host_alloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.goss_hip_host_free.argtypes = [V]
    for fn in ("get_field_states", "set_field_states", "get_field_parameters", "set_field_parameters",
               "get_field_states_device", "set_field_states_device"):
        getattr(L, f"goss_hip_system_{fn}").argtypes = [V, ctypes.c_void_p, ctypes.c_int]
    for fn in ("get_field_state_components", "set_field_state_components"):
        getattr(L, f"goss_hip_system_{fn}").argtypes = [V, _dp, ctypes.c_uint32, _u32p, ctypes.c_int]
    L.goss_hip_system_get_states.argtypes = [V, _dp]
    L.goss_hip_system_set_states.argtypes = [V, _dp]
    L.goss_hip_system_get_states_device.argtypes = [V, ctypes.c_void_p]
    L.goss_hip_system_set_states_device.argtypes = [V, ctypes.c_void_p]
    L.goss_hip_system_set_field_dofs.argtypes = [V, ctypes.POINTER(ctypes.c_int32), ctypes.c_uint64]
    L.goss_hip_system_gather_field_states_device.argtypes = [V, ctypes.c_void_p, ctypes.c_uint64]
    L.goss_hip_system_scatter_field_states_device.argtypes = [V, ctypes.c_void_p, ctypes.c_uint64]
    L.goss_hip_system_reset_default.argtypes = [V]
    L.goss_hip_system_set_parameter.argtypes = [V, ctypes.c_char_p, ctypes.c_double]
    L.goss_hip_system_get_parameter.argtypes = [V, ctypes.c_char_p, _dp]
    L.goss_hip_system_set_num_threads.argtypes = [V, ctypes.c_uint32]
    L.goss_hip_system_get_num_threads.argtypes = [V, _u32p]
    L.goss_hip_system_num_nodes.argtypes = [V, _u32p]
    L.goss_hip_system_set_steps_per_launch.argtypes = [V, ctypes.c_uint32]
    L.goss_hip_system_synchronize.argtypes = [V]
    L.goss_hip_system_stream.argtypes = [V, ctypes.POINTER(ctypes.c_void_p)]
    L.goss_hip_system_set_kernel_timing.argtypes = [V, ctypes.c_int]
    L.goss_hip_system_set_solver_params.argtypes = [V, ctypes.POINTER(SolverParams)]
    L.goss_hip_system_get_solver_params.argtypes = [V, ctypes.POINTER(SolverParams)]
    L.goss_hip_system_get_kernel_timing.argtypes = [V, _dp, ctypes.POINTER(ctypes.c_uint64)]
    L.goss_hip_system_get_solver_counters.argtypes = [V, ctypes.POINTER(ctypes.c_int64)]
    u64p = ctypes.POINTER(ctypes.c_uint64)
    L.goss_hip_adjoint_buffer_sizes.argtypes = [V, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, u64p, u64p, u64p]
    L.goss_hip_adjoint_record.argtypes = [V, _dp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, V, V]
    L.goss_hip_adjoint_backward.argtypes = [V, _dp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, V, V, V, V, V]
    if hasattr(L, "goss_hip_batch_create"):
        L.goss_hip_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint32,
                                            ctypes.POINTER(ctypes.c_void_p)]
        L.goss_hip_batch_destroy.argtypes = [V]
        L.goss_hip_batch_set_config.argtypes = [V, ctypes.POINTER(BatchConfig)]
        L.goss_hip_batch_get_config.argtypes = [V, ctypes.POINTER(BatchConfig)]
        L.goss_hip_batch_forward.argtypes = [V, ctypes.c_double, ctypes.c_double, _dp]
        L.goss_hip_batch_autotune.argtypes = [V, ctypes.c_double, ctypes.c_double, _dp,
                                              ctypes.POINTER(BatchConfig)]
    _loaded[path] = L
    return L


def check(rc: int, L: Optional[ctypes.CDLL] = None) -> None:
    if rc != 0:
        raise GossHipError(rc, (L or lib()).goss_hip_last_error().decode())
