"""2D projector nn.Module (ct_laboratory/tomography/ct_projector_2d_module.py:11-65).

Buffers M, b, src, dst and, by default, the float crossing table computed at
construction on the rays' device, as in the reference.  ``forward`` is
differentiable (CTProjector2DFunction); ``forward_project`` / ``back_project``
are the raw operators.  ``precomputed_intersections=False`` uses the on-the-fly
traversal instead of a table (no table memory; same segments under the
``"torch"`` rule).

While the table is the one the module computed and nothing has written to it,
the projections run on the fly with the table's own segment rule (torch
crossings, midpoints and lengths; axes with |dir| < 1e-12 dropped, as
compute_intersections_2d drops them): the traversal yields exactly the table's
segments, in the same order (the forward is bit-identical), without reading
4 (n_row + n_col + 2) bytes of table per ray, and the back projection sums LDS
tiles instead of adding each segment with a global atomic.  A table that is
replaced or written in place, or whose rays or map (src, dst, M, b) are
replaced or written in place, is consumed as given, as the reference would
apply it to the current rays.
"""
from __future__ import annotations

import weakref

import torch

from .. import _native as N

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
        self._own_table = self._record(tvals)

    def _record(self, t):
        """The table and the geometry it was computed for: weak references to the
        table, M, b, src and dst buffers with their version counters."""
        if t is None:
            return None
        return tuple((weakref.ref(x), x._version) for x in (t, self.M, self.b, self.src, self.dst))

    def _own_untouched(self) -> bool:
        """Whether the table is the one this module computed and neither it nor
        the geometry it was computed for (M, b, src, dst: same tensors, unwritten)
        has changed since; otherwise the table is consumed as given."""
        t = self.tvals
        if t is None or self._own_table is None:
            return False
        return all(ref() is x and x._version == ver
                   for (ref, ver), x in zip(self._own_table, (t, self.M, self.b, self.src, self.dst)))

    def __getstate__(self):
        # a weak reference does not pickle (torch.save, copy.deepcopy): keep whether the
        # table is still the module's own, untouched one, and point at the copy on load
        state = self.__dict__.copy()
        state["_own_table"] = self._own_untouched()
        return state

    def __setstate__(self, state):
        own = state.pop("_own_table", False)
        self.__dict__.update(state)
        t = self._buffers.get("tvals")
        self._own_table = self._record(t) if own is True else None

    def _rule(self):
        """(tvals, semantics, zero_rule) to project with: no table and the
        table's rule while the table is the module's own, untouched one."""
        t = self.tvals
        if t is None:
            return None, self.semantics, None
        if self._own_untouched():
            return None, "torch", N.ZERO_THRESH
        return t, self.semantics, None

    def forward_project(self, image):
        tvals, sem, zr = self._rule()
        if tvals is None:
            return forward_project_2d_on_the_fly_hip(image.contiguous(), self.M, self.b, self.src, self.dst,
                                                     semantics=sem, zero_rule=zr)
        return forward_project_2d_hip(image.contiguous(), tvals, self.M, self.b, self.src, self.dst)

    def back_project(self, sinogram):
        tvals, sem, zr = self._rule()
        if tvals is None:
            return back_project_2d_on_the_fly_hip(sinogram.contiguous(), self.M, self.b, self.src, self.dst,
                                                  self.n_row, self.n_col, semantics=sem, zero_rule=zr)
        return back_project_2d_hip(sinogram.contiguous(), tvals, self.M, self.b, self.src, self.dst,
                                   self.n_row, self.n_col)

    def forward(self, image):
        """image [R,C] or [B,R,C] -> sinogram [n_ray] or [B,n_ray] (differentiable)."""
        tvals, sem, zr = self._rule()
        return CTProjector2DFunction.apply(image, tvals, self.M, self.b, self.src, self.dst, self.backend, sem, zr)
