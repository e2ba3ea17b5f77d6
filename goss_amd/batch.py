"""Mixed-population batches: several device systems advanced together.

A :class:`DeviceBatch` groups :class:`~goss_amd.system.DeviceSystem` objects of
different cell models and solvers (FP32; the (model, solver) pairs compiled into
the fused kernel) and advances them over a common time window with one fused
persistent kernel launch per block of time steps, the launches captured in a
hipGraph (``goss_hip_batch_*`` in ``include/goss_hip.h``).  Each population
steps with its own ``dt``; a node's result equals that of its own system's
``forward_times`` over the same times.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Sequence, Tuple

import numpy as np

from ._lib import BatchConfig, check, lib
from .system import DeviceSystem


@dataclass
class Config:
    block_size: int = 256      # threads per block (64..256, multiple of 64)
    nodes_per_lane: int = 1    # block_size-node chunks per block and launch
    span: float = 1.0          # simulated time per launch
    order: Tuple[int, ...] = ()  # dispatch order of the populations (empty: as given)

    def _c(self, npop: int) -> BatchConfig:
        c = BatchConfig()
        c.block_size, c.nodes_per_lane, c.span = int(self.block_size), int(self.nodes_per_lane), float(self.span)
        order = list(self.order) if self.order else list(range(npop))
        for i, p in enumerate(order):
            c.order[i] = int(p)
        return c

    @staticmethod
    def _from(c: BatchConfig, npop: int) -> "Config":
        return Config(c.block_size, c.nodes_per_lane, c.span, tuple(int(c.order[i]) for i in range(npop)))


class DeviceBatch:
    def __init__(self, systems: Sequence[DeviceSystem]):
        self.systems = list(systems)
        arr = (ctypes.c_void_p * len(self.systems))(*[s._h.value for s in self.systems])
        self._h = ctypes.c_void_p()
        check(lib().goss_hip_batch_create(arr, len(self.systems), ctypes.byref(self._h)))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().goss_hip_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def config(self) -> Config:
        c = BatchConfig()
        check(lib().goss_hip_batch_get_config(self._h, ctypes.byref(c)))
        return Config._from(c, len(self.systems))

    @config.setter
    def config(self, cfg: Config) -> None:
        c = cfg._c(len(self.systems))
        check(lib().goss_hip_batch_set_config(self._h, ctypes.byref(c)))

    def _dt(self, dt) -> np.ndarray:
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (len(self.systems),)))
        return d

    def forward(self, t0: float, t1: float, dt) -> None:
        """Advance population p from t0 to t1 in steps of dt[p] (enqueued; see synchronize)."""
        d = self._dt(dt)
        check(lib().goss_hip_batch_forward(self._h, float(t0), float(t1),
                                           d.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def autotune(self, t0: float, t1: float, dt) -> Config:
        """Time candidate configs on [t0, t1], keep the fastest; states are left unchanged."""
        d = self._dt(dt)
        c = BatchConfig()
        check(lib().goss_hip_batch_autotune(self._h, float(t0), float(t1),
                                            d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(c)))
        return Config._from(c, len(self.systems))

    def synchronize(self) -> None:
        for s in self.systems:
            s.synchronize()
