"""2D fan-beam geometry and projector (ct_laboratory/tomography/fanbeam_projector_2d.py:6-173).

``build_fanbeam_2d_geometry`` reproduces the reference's arithmetic: the view
angles are a float32 ``torch.linspace`` (plus the offset, in float32), every
other quantity is evaluated in float64 from those angles with ``math``
functions, and the result is rounded to float32 once.  The detector loop is
vectorised (numpy float64 elementwise operations, the same operations in the
same order as the reference's per-element Python arithmetic).
"""
import math

import numpy as np
import torch

from .ct_projector_2d_module import CTProjector2DModule
from .standard_image_transform import standard_image_transform_2d

__all__ = ["build_fanbeam_2d_geometry", "FanBeam2DProjector"]


def build_fanbeam_2d_geometry(n_view: int, n_det: int, sid: float, sdd: float, det_spacing: float = 1.0,
                              angle_offset: float = 0.0, image_center_xy=(0.0, 0.0)):
    """src, dst [n_view * n_det, 2] float32, ray order view -> detector element."""
    cx, cy = image_center_xy
    angles = torch.linspace(0, 2 * math.pi - 2 * math.pi / n_view, n_view) + angle_offset
    mid = (n_det - 1) / 2.0
    offsets = (np.arange(n_det, dtype=np.float64) - mid) * det_spacing
    src_rows, dst_rows = [], []
    for theta in angles.tolist():
        sx = cx + sid * math.cos(theta)
        sy = cy + sid * math.sin(theta)
        dxc = cx - (sdd - sid) * math.cos(theta)
        dyc = cy - (sdd - sid) * math.sin(theta)
        px, py = -(dyc - sy), (dxc - sx)
        nrm = math.sqrt(px ** 2 + py ** 2)
        if nrm < 1e-12:
            continue  # degenerate view, skipped as in the reference
        px /= nrm
        py /= nrm
        cell = np.stack([dxc + offsets * px, dyc + offsets * py], axis=1)
        src_rows.append(np.broadcast_to(np.array([sx, sy]), cell.shape))
        dst_rows.append(cell)
    if not src_rows:
        return torch.empty((0, 2), dtype=torch.float32), torch.empty((0, 2), dtype=torch.float32)
    src = torch.from_numpy(np.concatenate(src_rows).astype(np.float32))
    dst = torch.from_numpy(np.concatenate(dst_rows).astype(np.float32))
    return src, dst


class FanBeam2DProjector(CTProjector2DModule):
    """CTProjector2DModule for a fan-beam scan (fanbeam_projector_2d.py:85-173)."""

    def __init__(self, n_row, n_col, n_view, n_det, sid, sdd, det_spacing=1.0, pixel_size=1.0, angle_offset=0.0,
                 backend="cuda", device=None, *, precomputed_intersections: bool = True, semantics: str = "torch"):
        src, dst = build_fanbeam_2d_geometry(n_view=n_view, n_det=n_det, sid=sid, sdd=sdd, det_spacing=det_spacing,
                                             angle_offset=angle_offset, image_center_xy=(0.0, 0.0))
        M, b = standard_image_transform_2d(n_x=n_row, n_y=n_col, s_x=pixel_size, s_y=pixel_size)
        if device is None:
            device = torch.device("cuda")
        super().__init__(n_row=n_row, n_col=n_col, M=M.to(device), b=b.to(device), src=src.to(device),
                         dst=dst.to(device), backend=backend, precomputed_intersections=precomputed_intersections,
                         semantics=semantics)
        self.n_view, self.n_det = int(n_view), int(n_det)
        self.sid, self.sdd = sid, sdd
        self.det_spacing, self.pixel_size, self.angle_offset = det_spacing, pixel_size, angle_offset
        self.device = device
