"""Circular flat-panel cone-beam geometry and projector
(ct_laboratory/tomography/conebeam_projector_3d.py:6-172).

``build_conebeam_3d_geometry`` uses the same rays, order and float32
arithmetic as the reference: angles from a float32 linspace over
[0, 2pi - 2pi/n_view], source at radius sid, detector centre at -(sdd - sid),
detector u along the in-plane perpendicular, v along +z, ray order
view -> detector row (v) -> detector column (u).
"""
import math

import torch

from .ct_projector_3d_module import CTProjector3DModule
from .standard_image_transform import standard_image_transform_3d

__all__ = ["build_conebeam_3d_geometry", "ConeBeam3DProjector"]


def build_conebeam_3d_geometry(n_view, det_nx, det_ny, sid, sdd, det_spacing=1.0, angle_offset=0.0,
                               volume_center_xyz=(0.0, 0.0, 0.0), device=None):
    """Returns src, dst of shape [n_view * det_ny * det_nx, 3] (float32).

    The per-view quantities (angles, their cosines and sines, source, detector
    centre and in-plane direction: n_view values each) are computed on the CPU;
    with ``device`` given, the expansion over the detector pixels (float32
    multiplies and adds, correctly rounded on either processor) runs there, so
    the rays equal the CPU-built ones bit for bit without a host copy of
    n_view * det_ny * det_nx rays."""
    cx, cy, cz = volume_center_xyz
    ang = torch.linspace(0, 2 * math.pi - 2 * math.pi / n_view, n_view, dtype=torch.float32) + angle_offset
    c, s = torch.cos(ang), torch.sin(ang)
    sx, sy = cx + sid * c, cy + sid * s
    sz = torch.full_like(sx, cz)
    dxc, dyc = cx - (sdd - sid) * c, cy - (sdd - sid) * s
    dzc = torch.full_like(dxc, cz)
    px, py = -(dyc - sy), (dxc - sx)
    nrm = torch.sqrt(px * px + py * py)
    px, py = px / nrm, py / nrm
    pz = torch.zeros_like(px)
    u = (torch.arange(det_nx, dtype=torch.float32) - (det_nx - 1) / 2.0) * det_spacing
    v = (torch.arange(det_ny, dtype=torch.float32) - (det_ny - 1) / 2.0) * det_spacing
    if device is not None:
        u, v = u.to(device), v.to(device)
        sx, sy, sz, dxc, dyc, dzc, px, py, pz = (t.to(device) for t in (sx, sy, sz, dxc, dyc, dzc, px, py, pz))
    U = u.unsqueeze(0).expand(det_ny, det_nx)
    V = v.unsqueeze(1).expand(det_ny, det_nx)
    sh = (n_view, 1, 1)
    sx, sy, sz, dxc, dyc, dzc, px, py, pz = (t.view(sh) for t in (sx, sy, sz, dxc, dyc, dzc, px, py, pz))
    cell_x = dxc + px * U + 0.0 * V
    cell_y = dyc + py * U + 0.0 * V
    cell_z = dzc + pz * U + 1.0 * V
    src = torch.stack([sx.expand(-1, det_ny, det_nx), sy.expand(-1, det_ny, det_nx),
                       sz.expand(-1, det_ny, det_nx)], dim=-1).reshape(-1, 3)
    dst = torch.stack([cell_x, cell_y, cell_z], dim=-1).reshape(-1, 3)
    return src, dst


class ConeBeam3DProjector(CTProjector3DModule):
    """CTProjector3DModule for a circular cone-beam scan (conebeam_projector_3d.py:101-172).

    The rays come as 720-style views of det_ny x det_nx pixels with one source
    each, so the module declares that layout: the forward runs the
    detector-tiled schedule and the backward the LDS-brick back projector."""

    def __init__(self, n_x, n_y, n_z, n_view, det_nx, det_ny, sid, sdd, det_spacing=1.0, voxel_size=1.0,
                 angle_offset=0.0, backend="cuda", device=None, precomputed_intersections=False, *,
                 semantics: str = "torch"):
        if device is None:
            device = torch.device("cuda")
        src, dst = build_conebeam_3d_geometry(n_view=n_view, det_nx=det_nx, det_ny=det_ny, sid=sid, sdd=sdd,
                                              det_spacing=det_spacing, angle_offset=angle_offset,
                                              volume_center_xyz=(0.0, 0.0, 0.0), device=device)
        M, b = standard_image_transform_3d(n_x=n_x, n_y=n_y, n_z=n_z, s_x=voxel_size, s_y=voxel_size,
                                           s_z=voxel_size)
        self.n_view, self.det_nx, self.det_ny = int(n_view), int(det_nx), int(det_ny)
        super().__init__(n_x=n_x, n_y=n_y, n_z=n_z, M=M.to(device), b=b.to(device), src=src.to(device),
                         dst=dst.to(device), backend=backend, device=device,
                         precomputed_intersections=precomputed_intersections, semantics=semantics,
                         det_shape=(det_ny, det_nx))
