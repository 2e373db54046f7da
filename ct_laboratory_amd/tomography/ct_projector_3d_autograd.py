"""Autograd Function of the 3D projector (ct_laboratory/tomography/ct_projector_3d_autograd.py:16-170).

The backward pass is the hand-written adjoint kernel: it back-projects
``grad_output`` over exactly the segments the forward pass summed (same crossing
table, compressed table or on-the-fly traversal), so gradients are exact
transposes.  Compressed tables run in one launch each way (the reference loops
over 65,536-ray chunks in Python, ct_projector_3d_autograd.py:60-95, 130-160).
"""
from __future__ import annotations

import torch

from ._backend import check_backend
from .ct_projector_3d_hip import back_project_3d_hip, forward_project_3d_hip

__all__ = ["CTProjector3DFunction", "CTBackProjector3DFunction", "OnTheFlyOptions"]


class OnTheFlyOptions:
    """How the on-the-fly traversal runs for one geometry: the segment rule
    (``"torch"`` or ``"otf"``), the detector layout (rows, cols) of rays ordered
    [view][rows][cols] if known, and the brick plan of that layout if built."""

    def __init__(self, semantics: str = "torch", det_shape=None, plan=None):
        self.semantics = semantics
        self.det_shape = det_shape
        self.plan = plan


class CTProjector3DFunction(torch.autograd.Function):
    """forward(volume [X,Y,Z] or [B,X,Y,Z], t_main, t_start, t_scale, M, b, src, dst,
    backend='cuda', options=None) -> sinogram [n_ray] or [B, n_ray].

    ``t_main`` is a float crossing table, or the uint16 table when ``t_start``
    and ``t_scale`` are given, or None for the on-the-fly traversal (configured
    by ``options``)."""

    @staticmethod
    def forward(ctx, volume, t_main, t_start, t_scale, M, b, src, dst, backend="cuda", options=None):
        check_backend(backend)
        opts = options if options is not None else OnTheFlyOptions()
        ctx.opts = opts
        ctx.volume_shape = volume.shape
        if t_start is not None and t_scale is not None:
            tv = (t_main, t_start, t_scale)
            ctx.kind = "u16"
            ctx.save_for_backward(t_main, t_start, t_scale, M, b, src, dst)
        elif t_main is not None:
            tv = t_main
            ctx.kind = "f32"
            ctx.save_for_backward(t_main, M, b, src, dst)
        else:
            tv = None
            ctx.kind = "otf"
            ctx.save_for_backward(M, b, src, dst)
        return forward_project_3d_hip(volume.contiguous(), tv, src, dst, M, b, semantics=opts.semantics,
                                      det_shape=opts.det_shape)

    @staticmethod
    def backward(ctx, grad_output):
        grad_output = grad_output.contiguous()
        saved = ctx.saved_tensors
        if ctx.kind == "u16":
            t_main, t_start, t_scale, M, b, src, dst = saved
            tv = (t_main, t_start, t_scale)
        elif ctx.kind == "f32":
            t_main, M, b, src, dst = saved
            tv = t_main
        else:
            M, b, src, dst = saved
            tv = None
        n_x, n_y, n_z = ctx.volume_shape[-3:]
        opts = ctx.opts
        grad = back_project_3d_hip(grad_output, tv, src, dst, M, b, n_x, n_y, n_z, semantics=opts.semantics,
                                   det_shape=opts.det_shape, brick_plan=opts.plan)
        return grad, None, None, None, None, None, None, None, None, None


class CTBackProjector3DFunction(torch.autograd.Function):
    """apply(sinogram, projector): the back projection of a 3D projector module
    as a differentiable operation, whose backward is the module's forward
    projection.  The reference declares this class with an instance
    constructor and static methods that read ``self.base_forward_projector``
    (ct_laboratory/tomography/ct_projector_3d_autograd.py:176-186), so it cannot
    run; here the projector is an argument."""

    @staticmethod
    def forward(ctx, sinogram, projector):
        ctx.projector = projector
        return projector.back_project(sinogram.contiguous())

    @staticmethod
    def backward(ctx, grad_volume):
        return ctx.projector.forward_project(grad_volume.contiguous()), None
