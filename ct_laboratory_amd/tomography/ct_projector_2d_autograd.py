"""Autograd Function of the 2D projector (ct_laboratory/tomography/ct_projector_2d_autograd.py:16-72).

The forward pass projects, the backward pass back-projects ``grad_output``
with the same crossing table (or the same on-the-fly traversal), so the
gradient is the exact transpose of the forward: both kernels enumerate the
same segments with the same voxel and length.
"""
from __future__ import annotations

import torch

from ._backend import check_backend
from .ct_projector_2d_hip import (back_project_2d_hip, back_project_2d_on_the_fly_hip, forward_project_2d_hip,
                                  forward_project_2d_on_the_fly_hip)

__all__ = ["CTProjector2DFunction"]


class CTProjector2DFunction(torch.autograd.Function):
    """forward(image [R,C] or [B,R,C], tvals, M, b, src, dst, backend='cuda', semantics='torch',
    zero_rule=None) -> sinogram [n_ray] or [B, n_ray].  ``tvals=None`` uses the
    on-the-fly traversal with the given segment rule (and zero rule, None: the
    rule's own)."""

    @staticmethod
    def forward(ctx, image, tvals, M, b, src, dst, backend="cuda", semantics="torch", zero_rule=None):
        check_backend(backend)
        ctx.image_shape = image.shape
        ctx.semantics = semantics
        ctx.zero_rule = zero_rule
        ctx.on_the_fly = tvals is None
        if ctx.on_the_fly:
            ctx.save_for_backward(M, b, src, dst)
            return forward_project_2d_on_the_fly_hip(image.contiguous(), M, b, src, dst, semantics=semantics,
                                                     zero_rule=zero_rule)
        ctx.save_for_backward(tvals, M, b, src, dst)
        return forward_project_2d_hip(image.contiguous(), tvals, M, b, src, dst)

    @staticmethod
    def backward(ctx, grad_output):
        grad_output = grad_output.contiguous()
        n_row, n_col = ctx.image_shape[-2:]
        if ctx.on_the_fly:
            M, b, src, dst = ctx.saved_tensors
            grad = back_project_2d_on_the_fly_hip(grad_output, M, b, src, dst, n_row, n_col,
                                                  semantics=ctx.semantics, zero_rule=ctx.zero_rule)
        else:
            tvals, M, b, src, dst = ctx.saved_tensors
            grad = back_project_2d_hip(grad_output, tvals, M, b, src, dst, n_row, n_col)
        return grad, None, None, None, None, None, None, None, None
