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
"""
from __future__ import annotations

import numpy as np
import torch

from .model import PMLModel


class PropagatorMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cquad, model: PMLModel, shots, misfit, mode, snap_every):
        cq = cquad.detach().to("cpu", torch.float64).numpy()
        if cq.shape != (model.Z, model.X):
            raise ValueError(f"cquad must be ({model.Z}, {model.X})")
        shots = np.asarray(shots, np.int32)
        if mode == "exact":
            if model.precision == 8:
                model.set_cquad64(cq)
            else:
                model.set_cquad(cq.astype(np.float32))
            g, mis = model.gradient_exact(shots, misfit, snap_every, fp64=True)
        elif mode == "reference":
            model.set_cquad(cq.astype(np.float32))
            g, mis = model.gradient(shots, misfit)
        else:
            raise ValueError(f"unknown mode {mode!r}")
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
