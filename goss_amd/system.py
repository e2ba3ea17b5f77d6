"""Low-level Python handle of one device system (``goss_hip_system``)."""
from __future__ import annotations

import ctypes
import weakref
from typing import Dict, List, Optional, Sequence

import numpy as np

from ._lib import JACOBIANS, PRECISIONS, SOLVERS, GossHipError, ModelInfo, SolverParams, check, lib, load

_dp = ctypes.POINTER(ctypes.c_double)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def pinned_array(shape, library: Optional[str] = None) -> np.ndarray:
    """A float64 array in pinned host memory (goss_hip_host_alloc), freed when
    the array and its views are gone: a destination for
    DeviceSystem.solve_async."""
    L = load(library) if library else lib()
    n = int(np.prod(shape, dtype=np.int64))
    p = ctypes.c_void_p()
    check(L.goss_hip_host_alloc(ctypes.byref(p), max(n, 1) * 8), L)
    buf = (ctypes.c_double * max(n, 1)).from_address(p.value)
    weakref.finalize(buf, L.goss_hip_host_free, p.value)
    return np.frombuffer(buf, dtype=np.float64)[:n].reshape(shape)


def model_names() -> List[str]:
    L = lib()
    return [L.goss_hip_model_name(i).decode() for i in range(L.goss_hip_model_count())]


class ModelVariant:
    """Metadata of a compiled model variant (sizes, names, defaults)."""

    def __init__(self, name: str, library: Optional[str] = None):
        L = load(library) if library else lib()
        info = ModelInfo()
        check(L.goss_hip_model_info_get(name.encode(), ctypes.byref(info)), L)
        self.name = name
        self.num_states = info.num_states
        self.num_parameters = info.num_parameters
        self.num_field_states = info.num_field_states
        self.num_field_parameters = info.num_field_parameters
        self.state_names = [info.state_names[i].decode() for i in range(self.num_states)]
        self.parameter_names = [info.parameter_names[i].decode() for i in range(self.num_parameters)]
        self.field_state_names = [info.field_state_names[i].decode() for i in range(self.num_field_states)]
        self.field_parameter_names = [info.field_parameter_names[i].decode()
                                      for i in range(self.num_field_parameters)]
        self.field_state_indices = [info.field_state_indices[i] for i in range(self.num_field_states)]
        self.field_parameter_indices = [info.field_parameter_indices[i] for i in range(self.num_field_parameters)]
        self.default_parameters = np.array([info.default_parameters[i] for i in range(self.num_parameters)])
        self.initial_conditions = np.array([info.initial_conditions[i] for i in range(self.num_states)])
        self.linear_terms = [int(info.linear_terms[i]) for i in range(self.num_states)]
        self.differential_states = [int(info.differential_states[i]) for i in range(self.num_states)]
        self.is_dae = bool(info.is_dae)


def solver_params(solver: str, library: Optional[str] = None, **overrides) -> SolverParams:
    L = load(library) if library else lib()
    sp = SolverParams()
    check(L.goss_hip_solver_params_default(SOLVERS[solver], ctypes.byref(sp)), L)
    for k, v in overrides.items():
        if v is None:
            continue
        if not hasattr(sp, k):
            raise KeyError(f"unknown solver parameter {k!r}")
        if k == "jacobian" and isinstance(v, str):
            v = JACOBIANS[v]  # "fd", "fd_fp32" or "analytic"
        setattr(sp, k, type(getattr(sp, k))(v))
    return sp


class DeviceSystem:
    """One ODESystemSolver on the GPU (wraps goss_hip_system_*)."""

    def __init__(self, variant: str, solver: str, num_nodes: int, precision: str = "fp64",
                 params: Optional[Dict[str, float]] = None, library: Optional[str] = None):
        """`library`: path of another build of the C ABI (tests); default the product library."""
        self._L = load(library) if library else lib()
        self.model = ModelVariant(variant, library)
        self.solver = solver
        self.precision = precision
        self.num_nodes = int(num_nodes)
        self._sp = solver_params(solver, library, **(params or {}))
        self._h = ctypes.c_void_p()
        check(self._L.goss_hip_system_create(variant.encode(), SOLVERS[solver], ctypes.byref(self._sp),
                                             PRECISIONS[precision], self.num_nodes, ctypes.byref(self._h)), self._L)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.goss_hip_system_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping -----------------------------------------------------------
    def forward(self, t: float, interval: float) -> None:
        check(self._L.goss_hip_system_forward(self._h, float(t), float(interval)), self._L)

    def forward_times(self, t) -> None:
        t = _f64(t)
        check(self._L.goss_hip_system_forward_times(self._h, _ptr(t), len(t)), self._L)

    def forward_times_noalloc(self, t) -> None:
        """forward_times with the steps staged in the library's fixed device
        array (no allocation; at most 32,768 time points per call)."""
        t = _f64(t)
        check(self._L.goss_hip_system_forward_times_noalloc(self._h, _ptr(t), len(t)), self._L)

    def solve(self, t, out: np.ndarray, tangled: bool = True) -> None:
        t = _f64(t)
        if not (out.flags.c_contiguous and out.dtype == np.float64):
            raise ValueError("out must be a C-contiguous float64 array")
        check(self._L.goss_hip_system_solve(self._h, _ptr(out), _ptr(t), len(t), int(tangled)), self._L)

    def solve_async(self, t, out, tangled: bool = True) -> None:
        """solve into pinned host memory (pinned_array, a pinned torch tensor)
        or device memory (a CUDA tensor, or a device address), returning once
        the steps and the row copies are enqueued on the system's streams;
        synchronize() waits for them.  Rows 1..len(t)-1 are written."""
        t = _f64(t)
        need = len(t) * self.num_nodes * self.model.num_field_states
        if isinstance(out, int):
            addr = out
        elif isinstance(out, np.ndarray):
            if not (out.flags.c_contiguous and out.dtype == np.float64 and out.size >= need):
                raise ValueError(f"out must be a C-contiguous float64 array of at least {need} values")
            addr = out.ctypes.data
        elif hasattr(out, "data_ptr"):
            if not (str(out.dtype) == "torch.float64" and out.is_contiguous() and out.numel() >= need):
                raise ValueError(f"out must be a contiguous float64 tensor of at least {need} values")
            addr = out.data_ptr()
        else:
            raise TypeError("out must be a pinned or device buffer (numpy array, torch tensor or address)")
        check(self._L.goss_hip_system_solve_async(self._h, addr, _ptr(t), len(t), int(tangled)), self._L)

    def synchronize(self) -> None:
        check(self._L.goss_hip_system_synchronize(self._h), self._L)

    def set_steps_per_launch(self, n: int) -> None:
        check(self._L.goss_hip_system_set_steps_per_launch(self._h, int(n)), self._L)

    # -- data ----------------------------------------------------------------
    def states(self) -> np.ndarray:
        out = np.zeros((self.num_nodes, self.model.num_states))
        check(self._L.goss_hip_system_get_states(self._h, _ptr(out)), self._L)
        return out

    def set_states(self, states) -> None:
        a = _f64(states).reshape(self.num_nodes, self.model.num_states)
        check(self._L.goss_hip_system_set_states(self._h, _ptr(a)), self._L)

    def get_field_states(self, out: np.ndarray, tangled: bool = True) -> None:
        check(self._L.goss_hip_system_get_field_states(self._h, out.ctypes.data, int(tangled)), self._L)

    def set_field_states(self, a: np.ndarray, tangled: bool = True) -> None:
        a = _f64(a)
        check(self._L.goss_hip_system_set_field_states(self._h, a.ctypes.data, int(tangled)), self._L)

    def get_field_state_components(self, out: np.ndarray, components: Sequence[int], tangled: bool = True) -> None:
        c = np.ascontiguousarray(components, dtype=np.uint32)
        check(self._L.goss_hip_system_get_field_state_components(
            self._h, _ptr(out), len(c), c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), int(tangled)), self._L)

    def set_field_state_components(self, a: np.ndarray, components: Sequence[int], tangled: bool = True) -> None:
        a = _f64(a)
        c = np.ascontiguousarray(components, dtype=np.uint32)
        check(self._L.goss_hip_system_set_field_state_components(
            self._h, _ptr(a), len(c), c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), int(tangled)), self._L)

    def get_field_parameters(self, out: np.ndarray, tangled: bool = True) -> None:
        check(self._L.goss_hip_system_get_field_parameters(self._h, out.ctypes.data, int(tangled)), self._L)

    def set_field_parameters(self, a: np.ndarray, tangled: bool = True) -> None:
        a = _f64(a)
        check(self._L.goss_hip_system_set_field_parameters(self._h, a.ctypes.data, int(tangled)), self._L)

    def get_field_states_device(self, dev_ptr: int, tangled: bool = True) -> None:
        check(self._L.goss_hip_system_get_field_states_device(self._h, ctypes.c_void_p(dev_ptr), int(tangled)), self._L)

    def set_field_states_device(self, dev_ptr: int, tangled: bool = True) -> None:
        check(self._L.goss_hip_system_set_field_states_device(self._h, ctypes.c_void_p(dev_ptr), int(tangled)), self._L)

    def get_states_device(self, dev_ptr: int) -> None:
        """All states, node-major [N][S] FP64, into device memory at dev_ptr."""
        check(self._L.goss_hip_system_get_states_device(self._h, ctypes.c_void_p(dev_ptr)), self._L)

    def set_states_device(self, dev_ptr: int) -> None:
        check(self._L.goss_hip_system_set_states_device(self._h, ctypes.c_void_p(dev_ptr)), self._L)

    def set_field_dofs(self, dofs: np.ndarray, num_values: int) -> None:
        """dofs [F][N]: index in a field vector of num_values doubles of field state c of node n."""
        d = np.ascontiguousarray(dofs, dtype=np.int32)
        if d.size != self.num_nodes * self.model.num_field_states:
            raise ValueError(f"expected {self.model.num_field_states} x {self.num_nodes} dofs, got {d.shape}")
        check(self._L.goss_hip_system_set_field_dofs(self._h, d.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                     int(num_values)), self._L)

    def gather_field_states_device(self, dev_ptr: int, num_values: int) -> None:
        """Field states <- values[dofs] of the device field vector at dev_ptr."""
        check(self._L.goss_hip_system_gather_field_states_device(self._h, ctypes.c_void_p(dev_ptr),
                                                                 int(num_values)), self._L)

    def scatter_field_states_device(self, dev_ptr: int, num_values: int) -> None:
        """values[dofs] <- field states, into the device field vector at dev_ptr."""
        check(self._L.goss_hip_system_scatter_field_states_device(self._h, ctypes.c_void_p(dev_ptr),
                                                                  int(num_values)), self._L)

    def reset_default(self) -> None:
        check(self._L.goss_hip_system_reset_default(self._h), self._L)

    def set_parameter(self, name: str, value: float) -> None:
        check(self._L.goss_hip_system_set_parameter(self._h, name.encode(), float(value)), self._L)

    def get_parameter(self, name: str) -> float:
        v = ctypes.c_double()
        check(self._L.goss_hip_system_get_parameter(self._h, name.encode(), ctypes.byref(v)), self._L)
        return v.value

    def set_num_threads(self, n: int) -> None:
        check(self._L.goss_hip_system_set_num_threads(self._h, int(n)), self._L)

    def get_num_threads(self) -> int:
        v = ctypes.c_uint32()
        check(self._L.goss_hip_system_get_num_threads(self._h, ctypes.byref(v)), self._L)
        return v.value

    def solver_counters(self) -> np.ndarray:
        out = (ctypes.c_int64 * 4)()
        check(self._L.goss_hip_system_get_solver_counters(self._h, out), self._L)
        return np.array(list(out))

    # -- adjoint of GRL1 stepping (device buffers passed as integer addresses) --
    def adjoint_buffer_sizes(self, n: int, sample_every: int, checkpoint_every: int):
        """(checkpoint bytes, segment bytes, number of samples) for n time points."""
        c, g, m = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(self._L.goss_hip_adjoint_buffer_sizes(self._h, int(n), int(sample_every), int(checkpoint_every),
                                                  ctypes.byref(c), ctypes.byref(g), ctypes.byref(m)), self._L)
        return c.value, g.value, m.value

    def adjoint_record(self, t, sample_every: int, checkpoint_every: int, ckpt_ptr: int, samples_ptr: int) -> None:
        t = _f64(t)
        check(self._L.goss_hip_adjoint_record(self._h, _ptr(t), len(t), int(sample_every), int(checkpoint_every),
                                            ctypes.c_void_p(ckpt_ptr), ctypes.c_void_p(samples_ptr)), self._L)

    def adjoint_backward(self, t, sample_every: int, checkpoint_every: int, ckpt_ptr: int, grad_samples_ptr: int,
                         segment_ptr: int, grad_states_ptr: int, grad_fp_ptr: int) -> None:
        t = _f64(t)
        check(self._L.goss_hip_adjoint_backward(self._h, _ptr(t), len(t), int(sample_every), int(checkpoint_every),
                                              ctypes.c_void_p(ckpt_ptr), ctypes.c_void_p(grad_samples_ptr),
                                              ctypes.c_void_p(segment_ptr), ctypes.c_void_p(grad_states_ptr),
                                              ctypes.c_void_p(grad_fp_ptr)), self._L)

    def set_solver_params(self, **overrides) -> None:
        """Changes solver tunables (names of goss_hip_solver_params) from the next launch on."""
        sp = SolverParams()
        check(self._L.goss_hip_system_get_solver_params(self._h, ctypes.byref(sp)), self._L)
        for k, v in overrides.items():
            if not hasattr(sp, k):
                raise KeyError(f"unknown solver parameter {k!r}")
            if k == "jacobian" and isinstance(v, str):
                v = JACOBIANS[v]
            setattr(sp, k, type(getattr(sp, k))(v))
        check(self._L.goss_hip_system_set_solver_params(self._h, ctypes.byref(sp)), self._L)
        self._sp = sp

    def solver_params(self) -> SolverParams:
        sp = SolverParams()
        check(self._L.goss_hip_system_get_solver_params(self._h, ctypes.byref(sp)), self._L)
        return sp

    def set_kernel_timing(self, enable: bool = True) -> None:
        """Bracket the solver kernels of each stepping call with HIP events."""
        check(self._L.goss_hip_system_set_kernel_timing(self._h, int(enable)), self._L)

    def kernel_timing(self):
        """(milliseconds, launches) of the solver kernels of the last stepping call."""
        ms, n = ctypes.c_double(), ctypes.c_uint64()
        check(self._L.goss_hip_system_get_kernel_timing(self._h, ctypes.byref(ms), ctypes.byref(n)), self._L)
        return ms.value, n.value

    def stream(self) -> int:
        v = ctypes.c_void_p()
        check(self._L.goss_hip_system_stream(self._h, ctypes.byref(v)), self._L)
        return v.value or 0
