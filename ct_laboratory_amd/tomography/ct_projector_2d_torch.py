"""The names of the reference's pure-torch 2D projector
(ct_laboratory/tomography/ct_projector_2d_torch.py:1-310), computed by the HIP
kernels with the torch projector's segment rule (sorted crossings with the
exact ``dir == 0`` test, world-space lengths, midpoints rounded half to even;
ct_projector_2d_torch.py:4-250).

* ``compute_intersections_2d_torch``: [n_ray, n_row+n_col+2] sorted crossings,
  INF padded (:4-65).
* ``forward_project_2d_torch`` / ``back_project_2d_torch``: the table
  consumers (:68-250).

Deliberate differences: the inputs must be HIP tensors (there is no CPU
projector), a [B, R, C] image or [B, n_ray] sinogram is projected as a batch,
and a table of ``None`` projects on the fly with the same rule.  The
reference's ``CTProjector2DFunction`` passes M where the table belongs
(:268 and :285) and cannot run; here it projects on the fly with the torch
rule, so ``CTProjector2D`` works as its docstrings describe.
"""
from __future__ import annotations

import torch

from .. import _native as N
from .ct_projector_2d_hip import (_M_inv, _b, back_project_2d_hip, back_project_2d_on_the_fly_hip,
                                  forward_project_2d_hip, forward_project_2d_on_the_fly_hip)
from .. import _C

__all__ = ["compute_intersections_2d_torch", "forward_project_2d_torch", "back_project_2d_torch",
           "CTProjector2DFunction", "CTProjector2D"]


def compute_intersections_2d_torch(n_row: int, n_col: int, M, b, src, dst) -> torch.Tensor:
    """Sorted crossings t in [0, 1] of every row and column line, INF padded,
    with the torch projector's exact zero test (ct_projector_2d_torch.py:4-65)."""
    return _C.compute_intersections_2d(n_row, n_col, src.contiguous(), dst.contiguous(), _M_inv(M), _b(b),
                                       zero_rule=N.ZERO_EXACT)


def forward_project_2d_torch(image, t_sorted, M, b, src, dst) -> torch.Tensor:
    """Line integrals [n_ray] ([B, n_ray] for a [B, R, C] image)
    (ct_projector_2d_torch.py:68-163)."""
    if t_sorted is None or t_sorted.numel() == 0:
        return forward_project_2d_on_the_fly_hip(image.contiguous(), M, b, src.contiguous(), dst.contiguous(),
                                                 semantics="torch")
    return forward_project_2d_hip(image.contiguous(), t_sorted.contiguous(), M, b, src.contiguous(), dst.contiguous())


def back_project_2d_torch(sinogram, t_sorted, M, b, src, dst, n_row: int, n_col: int) -> torch.Tensor:
    """Back projection [n_row, n_col] ([B, ...] for a [B, n_ray] sinogram)
    (ct_projector_2d_torch.py:166-250)."""
    if t_sorted is None or t_sorted.numel() == 0:
        return back_project_2d_on_the_fly_hip(sinogram.contiguous(), M, b, src.contiguous(), dst.contiguous(), n_row,
                                              n_col, semantics="torch")
    return back_project_2d_hip(sinogram.contiguous(), t_sorted.contiguous(), M, b, src.contiguous(),
                               dst.contiguous(), n_row, n_col)


class CTProjector2DFunction(torch.autograd.Function):
    """forward(image, M, b, src, dst): the on-the-fly torch-rule projection;
    backward: its back projection (ct_projector_2d_torch.py:255-290)."""

    @staticmethod
    def forward(ctx, image, M, b, src, dst):
        ctx.save_for_backward(M, b, src, dst)
        ctx.image_shape = tuple(image.shape[-2:])
        return forward_project_2d_torch(image, None, M, b, src, dst)

    @staticmethod
    def backward(ctx, grad_output):
        M, b, src, dst = ctx.saved_tensors
        n_row, n_col = ctx.image_shape
        return back_project_2d_torch(grad_output, None, M, b, src, dst, n_row, n_col), None, None, None, None


class CTProjector2D(torch.nn.Module):
    """nn.Module over CTProjector2DFunction with M, b, src, dst as buffers
    (ct_projector_2d_torch.py:293-310)."""

    def __init__(self, M, b, src, dst):
        super().__init__()
        self.register_buffer("M", M)
        self.register_buffer("b", b)
        self.register_buffer("src", src)
        self.register_buffer("dst", dst)

    def forward(self, image):
        return CTProjector2DFunction.apply(image, self.M, self.b, self.src, self.dst)
