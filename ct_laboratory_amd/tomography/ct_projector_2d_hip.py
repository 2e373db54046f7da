"""2D projector functions over the HIP kernels.

Same names, argument order and squeeze rules as the reference's
tomography/ct_projector_2d_cuda.py:1-46 (``*_hip`` plus the reference
``*_cuda`` names), and on-the-fly functions the reference exposes only
through ``_C`` (src/ct_projector_2d.cu:246-465, bindings.cpp:75-80).

The on-the-fly functions take ``semantics``: ``"torch"`` is the segment rule
of the reference's torch projector (ct_projector_2d_torch.py:68-250: sorted
crossings with exact ``dir == 0`` tests, world-space lengths, midpoint
rounded half to even); ``"otf"`` is the rule of the reference's 3D on-the-fly
kernel applied to 2D rays ((t1-t0)*|dst-src| lengths with 0 and 1 inserted and
the ``|dir| < 1e-12`` test); ``"march"`` is the reference's own 2D on-the-fly
kernel (src/ct_projector_2d.cu:246-413: an incremental march from a
floor-indexed entry voxel, every in-grid step visited, zero and negative
lengths included), which ``_C.forward/back_project_2d_on_the_fly_cuda`` use by
default, as the reference's ``_C`` does.  No reference Python code calls the 2D
on-the-fly kernels; the default here stays ``"torch"``.
"""
from __future__ import annotations

import torch

from .. import _C
from .. import _native as N
from .ct_projector_3d_hip import semantics_mode

__all__ = [
    "compute_intersections_2d_hip", "forward_project_2d_hip", "back_project_2d_hip",
    "forward_project_2d_on_the_fly_hip", "back_project_2d_on_the_fly_hip",
    "compute_intersections_2d_cuda", "forward_project_2d_cuda", "back_project_2d_cuda",
]


def _M_inv(M: torch.Tensor) -> torch.Tensor:
    """Contiguous float32 inverse on M's device (ct_projector_2d_cuda.py:6-8)."""
    if M.dim() > 2:
        M = M[0]
    return _C.cached_inverse(M)


def _b(b: torch.Tensor) -> torch.Tensor:
    if b.dim() > 1:
        b = b[0]
    return b.to(torch.float32).contiguous()


def compute_intersections_2d_hip(n_row, n_col, M, b, src, dst, *, K=None):
    """[n_ray, (n_row+1)+(n_col+1)] sorted crossings (ct_projector_2d_cuda.py:11-20)."""
    return _C.compute_intersections_2d(n_row, n_col, src.contiguous(), dst.contiguous(), _M_inv(M), _b(b), K=K)


def forward_project_2d_hip(image, tvals, M, b, src, dst):
    """[B, n_ray] or [n_ray] line integrals from a crossing table (ct_projector_2d_cuda.py:23-33)."""
    out = _C.forward_project_2d_cuda(image, tvals, src, dst, _M_inv(M), _b(b))
    return out.squeeze(0) if image.dim() == 2 else out


def back_project_2d_hip(sinogram, tvals, M, b, src, dst, n_row, n_col):
    """[B, R, C] or [R, C] back projection from a crossing table (ct_projector_2d_cuda.py:36-46)."""
    out = _C.back_project_2d_cuda(sinogram, tvals, src, dst, _M_inv(M), _b(b), n_row, n_col)
    return out.squeeze(0) if sinogram.dim() == 1 else out


def _zero_rule(semantics: str) -> int:
    return N.ZERO_EXACT if semantics == "torch" else N.ZERO_THRESH


def forward_project_2d_on_the_fly_hip(image, M, b, src, dst, *, semantics: str = "torch", zero_rule=None):
    """On-the-fly 2D forward projection: [B, n_ray] or [n_ray].  zero_rule
    (None: the rule's own) selects which axes have no crossings."""
    zr = _zero_rule(semantics) if zero_rule is None else int(zero_rule)
    fn = _C.forward_project_2d_on_the_fly_f64 if image.dtype == torch.float64 else _C.forward_project_2d_on_the_fly_cuda
    out = fn(image, src, dst, _M_inv(M), _b(b), mode=semantics_mode(semantics), zero_rule=zr)
    return out.squeeze(0) if image.dim() == 2 else out


def back_project_2d_on_the_fly_hip(sinogram, M, b, src, dst, n_row, n_col, *, semantics: str = "torch",
                                   zero_rule=None):
    """On-the-fly 2D back projection, the exact transpose of the forward."""
    zr = _zero_rule(semantics) if zero_rule is None else int(zero_rule)
    fn = _C.back_project_2d_on_the_fly_f64 if sinogram.dtype == torch.float64 else _C.back_project_2d_on_the_fly_cuda
    out = fn(sinogram, src, dst, _M_inv(M), _b(b), n_row, n_col, mode=semantics_mode(semantics), zero_rule=zr)
    return out.squeeze(0) if sinogram.dim() == 1 else out


# Reference names (ct_projector_2d_cuda.py)
compute_intersections_2d_cuda = compute_intersections_2d_hip
forward_project_2d_cuda = forward_project_2d_hip
back_project_2d_cuda = back_project_2d_hip
