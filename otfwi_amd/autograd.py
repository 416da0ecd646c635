"""torch.autograd.Function around the propagator: J(c^2) = mean over shots
of the misfit between simulated and observed traces, differentiable in c^2.

mode="exact"      the hand-written exact discrete adjoint (the transpose of
                  the forward recursion) with snapshot checkpointing every
                  ``snap_every`` steps (0: sqrt(T-1)); J is the L2 value
                  sum (sim - obs)^2 or the reference W2 value, and the
                  gradient is their exact derivative.
mode="reference"  the reference's gradient (regression_PML_single.cu:704-846):
                  boundary-saving reconstruction of the forward field and the
                  imaging condition with the reference adjoint sources
                  (adjoint_functions.py:8-11, 25-67).  Not the exact derivative
                  of J; provided so that the reference's optimizer flow can sit
                  behind autograd.

The gradient is formed in the forward call (the adjoint needs that run's
wavefields) and scaled by the incoming gradient in backward.

A float32 CUDA tensor on an FP32 model stays on the device: the library
reads c^2 from the tensor's memory and writes the gradient into a tensor on
the same device (after torch's stream has finished producing c^2); other
inputs (CPU tensors, float64, the FP64 build) go through host arrays.
"""
from __future__ import annotations

import numpy as np
import torch

from .model import PMLModel


class PropagatorMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cquad, model: PMLModel, shots, misfit, mode, snap_every):
        if tuple(cquad.shape) != (model.Z, model.X):
            raise ValueError(f"cquad must be ({model.Z}, {model.X})")
        if mode not in ("exact", "reference"):
            raise ValueError(f"unknown mode {mode!r}")
        shots = np.asarray(shots, np.int32)
        if cquad.is_cuda and cquad.dtype == torch.float32 and model.precision == 4:
            cq = cquad.detach().contiguous()
            g = torch.empty_like(cq)
            # the library's stream reads c^2: torch's work producing it must be done
            torch.cuda.current_stream(cq.device).synchronize()
            model.set_cquad_device(cq.data_ptr())
            if mode == "exact":
                mis = model.gradient_exact_device(shots, misfit, snap_every, g.data_ptr())
            else:
                mis = model.gradient_device(shots, misfit, g.data_ptr())
            ctx.save_for_backward(g)
            return torch.tensor(float(np.mean(mis)), dtype=cquad.dtype, device=cquad.device)
        cq = cquad.detach().to("cpu", torch.float64).numpy()
        if mode == "exact":
            if model.precision == 8:
                model.set_cquad64(cq)
            else:
                model.set_cquad(cq.astype(np.float32))
            g, mis = model.gradient_exact(shots, misfit, snap_every, fp64=True)
        elif mode == "reference":
            model.set_cquad(cq.astype(np.float32))
            g, mis = model.gradient(shots, misfit)
        ctx.save_for_backward(torch.as_tensor(np.asarray(g), dtype=cquad.dtype,
                                              device=cquad.device))
        return torch.tensor(float(np.mean(mis)), dtype=cquad.dtype, device=cquad.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None, None, None, None


def propagator_misfit(cquad, model: PMLModel, shots, misfit="l2", mode="exact", snap_every=0):
    """Differentiable J(cquad); see the module docstring.  `model` must have
    damping, sources, receivers and observed traces set (and the boundary
    ring for mode="reference")."""
    return PropagatorMisfit.apply(cquad, model, shots, misfit, mode, snap_every)
