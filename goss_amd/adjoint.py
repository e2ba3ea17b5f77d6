"""Gradients of sampled field-state traces through explicit time stepping.

:class:`SampledTrace` is a ``torch.autograd.Function``: its forward advances a
:class:`~goss_amd.system.DeviceSystem` (solver GRL1, RL1, ExplicitEuler or
RK4, FP64 or FP32, one step per interval) over the given times and returns
the first field state sampled
every ``sample_every`` steps, as a ``[n_samples, N]`` FP64 tensor on the GPU.
Its backward runs the device discrete adjoint (``goss_hip_adjoint_backward``):
the states are checkpointed every ``checkpoint_every`` steps in the forward
and recomputed segment by segment in the backward sweep (``checkpoint_every=1``
stores every state and skips the recomputation), and the gradients
with respect to the initial states ``[N, S]`` and the per-node field
parameters ``[N, NFP]`` come back as FP64 tensors.

PyTorch provides the device buffers and the autograd plumbing; the forward
and reverse sweeps are the library's HIP kernels.  The system's field
parameters must not change between forward and backward.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .system import DeviceSystem


class SampledTrace(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y0: torch.Tensor, fp: torch.Tensor, system: DeviceSystem, t: np.ndarray, sample_every: int,
                checkpoint_every: int):
        N, S = system.num_nodes, system.model.num_states
        NFP = system.model.num_field_parameters
        if tuple(y0.shape) != (N, S):
            raise ValueError(f"y0 must have shape ({N}, {S})")
        if NFP and tuple(fp.shape) != (N, NFP):
            raise ValueError(f"fp must have shape ({N}, {NFP})")
        t = np.ascontiguousarray(t, dtype=np.float64)
        system.set_states(y0.detach().to("cpu", torch.float64).numpy())
        if NFP:
            system.set_field_parameters(fp.detach().to("cpu", torch.float64).numpy())
        ckpt_bytes, seg_bytes, M = system.adjoint_buffer_sizes(len(t), sample_every, checkpoint_every)
        dev = torch.device("cuda")
        ckpt = torch.empty(max(ckpt_bytes, 1), dtype=torch.uint8, device=dev)
        samples = torch.empty((M, N), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        system.adjoint_record(t, sample_every, checkpoint_every, ckpt.data_ptr(), samples.data_ptr())
        system.synchronize()
        ctx.system, ctx.t, ctx.sample_every, ctx.checkpoint_every = system, t, sample_every, checkpoint_every
        ctx.ckpt, ctx.seg_bytes = ckpt, seg_bytes
        ctx.y0_device, ctx.fp_device = y0.device, fp.device
        return samples

    @staticmethod
    def backward(ctx, grad_samples: torch.Tensor):
        system: DeviceSystem = ctx.system
        N, S = system.num_nodes, system.model.num_states
        NFP = system.model.num_field_parameters
        dev = torch.device("cuda")
        g = grad_samples.detach().to(dev, torch.float64).contiguous()
        seg = torch.empty(max(ctx.seg_bytes, 1), dtype=torch.uint8, device=dev)
        grad_y0 = torch.empty((N, S), dtype=torch.float64, device=dev)
        grad_fp = torch.empty((N, max(NFP, 1)), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        system.adjoint_backward(ctx.t, ctx.sample_every, ctx.checkpoint_every, ctx.ckpt.data_ptr(), g.data_ptr(),
                                seg.data_ptr(), grad_y0.data_ptr(), grad_fp.data_ptr())
        system.synchronize()
        gfp = grad_fp[:, :NFP].to(ctx.fp_device) if NFP else None
        return grad_y0.to(ctx.y0_device), gfp, None, None, None, None


def voltage_trace(system: DeviceSystem, y0: torch.Tensor, fp: torch.Tensor, t, sample_every: int,
                  checkpoint_every: int = 200) -> torch.Tensor:
    """Differentiable ``[n_samples, N]`` trace of the first field state (see :class:`SampledTrace`)."""
    return SampledTrace.apply(y0, fp, system, np.asarray(t, dtype=np.float64), int(sample_every), int(checkpoint_every))
