"""2D projector nn.Module (ct_laboratory/tomography/ct_projector_2d_module.py:11-65).

Buffers M, b, src, dst and, by default, the float crossing table computed at
construction on the rays' device, as in the reference.  ``forward`` is
differentiable (CTProjector2DFunction); ``forward_project`` / ``back_project``
are the raw operators.  ``precomputed_intersections=False`` uses the on-the-fly
traversal instead of a table (no table memory; same segments under the
``"torch"`` rule).
"""
from __future__ import annotations

import torch

from ._backend import check_backend
from .ct_projector_2d_autograd import CTProjector2DFunction
from .ct_projector_2d_hip import (back_project_2d_hip, back_project_2d_on_the_fly_hip,
                                  compute_intersections_2d_hip, forward_project_2d_hip,
                                  forward_project_2d_on_the_fly_hip)

__all__ = ["CTProjector2DModule"]


class CTProjector2DModule(torch.nn.Module):
    def __init__(self, n_row, n_col, M, b, src, dst, backend="cuda", *, precomputed_intersections: bool = True,
                 semantics: str = "torch"):
        super().__init__()
        self.backend = check_backend(backend)
        self.n_row = int(n_row)
        self.n_col = int(n_col)
        self.semantics = semantics
        self.register_buffer("M", M)
        self.register_buffer("b", b)
        self.register_buffer("src", src.contiguous())
        self.register_buffer("dst", dst.contiguous())
        tvals = None
        if precomputed_intersections:
            tvals = compute_intersections_2d_hip(self.n_row, self.n_col, self.M, self.b, self.src, self.dst)
        self.register_buffer("tvals", tvals)

    def forward_project(self, image):
        if self.tvals is None:
            return forward_project_2d_on_the_fly_hip(image.contiguous(), self.M, self.b, self.src, self.dst,
                                                     semantics=self.semantics)
        return forward_project_2d_hip(image.contiguous(), self.tvals, self.M, self.b, self.src, self.dst)

    def back_project(self, sinogram):
        if self.tvals is None:
            return back_project_2d_on_the_fly_hip(sinogram.contiguous(), self.M, self.b, self.src, self.dst,
                                                  self.n_row, self.n_col, semantics=self.semantics)
        return back_project_2d_hip(sinogram.contiguous(), self.tvals, self.M, self.b, self.src, self.dst,
                                   self.n_row, self.n_col)

    def forward(self, image):
        """image [R,C] or [B,R,C] -> sinogram [n_ray] or [B,n_ray] (differentiable)."""
        return CTProjector2DFunction.apply(image, self.tvals, self.M, self.b, self.src, self.dst, self.backend,
                                           self.semantics)
