"""3D static-gantry (StaticCT) geometry, frame sequences and projectors
(ct_laboratory/tomography/staticct_projector_3d.py:8-539).

Rays: for every frame, every active source, every module the source
illuminates, every module pixel (u outer, v inner), in that order
(staticct_projector_3d.py:113-160).  Each (frame, source, module) block is a
flat panel of det_nx_per_module x det_ny_per_module pixels lit by one source,
so the projector declares that layout to the on-the-fly kernels (detector-tiled
forward, LDS-brick backward when the plan fits in memory).
"""
import math
from typing import Optional

import torch

from .ct_projector_3d_module import CTProjector3DModule

__all__ = ["build_circular_sequence", "build_helical_sequence", "build_static_3d_geometry", "StaticCTProjector3D",
           "build_uniform_static_3d_geometry", "UniformStaticCTProjector3D"]


def build_circular_sequence(n_frame: int, n_source: int, device: Optional[torch.device] = None):
    """Identity gantry transforms; frame i activates source i mod n_source
    (staticct_projector_3d.py:8-30)."""
    device = torch.device("cpu") if device is None else device
    M_gantry = torch.eye(3, dtype=torch.float32, device=device).unsqueeze(0).repeat(n_frame, 1, 1)
    b_gantry = torch.zeros((n_frame, 3), dtype=torch.float32, device=device)
    active = torch.zeros(n_frame, n_source, dtype=torch.bool, device=device)
    active[torch.arange(n_frame), torch.arange(n_frame) % n_source] = True
    return M_gantry, b_gantry, active


def build_helical_sequence(n_frame: int, n_source: int, helical_travel_per_rot: float,
                           device: Optional[torch.device] = None):
    """Circular sequence plus a z translation of travel/n_source per frame, centred
    on the middle frame (staticct_projector_3d.py:33-63)."""
    M_gantry, b_gantry, active = build_circular_sequence(n_frame, n_source, torch.device("cpu"))
    step = helical_travel_per_rot / n_source
    # float64 products rounded once to float32, as the reference's per-element assignment
    b_gantry[:, 2] = torch.tensor([(i - (n_frame - 1) / 2.0) * step for i in range(n_frame)], dtype=torch.float32)
    if device is not None:
        M_gantry, b_gantry, active = M_gantry.to(device), b_gantry.to(device), active.to(device)
    return M_gantry, b_gantry, active


def _module_pixels_3d(module_centers, module_orientations, det_nx, det_ny, sx, sy):
    """Gantry coordinates [n_module, det_nx*det_ny, 3]: Rm (u, v, 0) + centre
    (staticct_projector_3d.py:100-110, 143-145)."""
    mid_u, mid_v = (det_nx - 1) / 2.0, (det_ny - 1) / 2.0
    local = torch.tensor([[(iu - mid_u) * sx, (iv - mid_v) * sy, 0.0] for iu in range(det_nx) for iv in range(det_ny)],
                         dtype=torch.float32)
    out = []
    for m in range(module_centers.shape[0]):
        p = local @ module_orientations[m].transpose(0, 1)
        p += module_centers[m]
        out.append(p)
    return torch.stack(out)


def build_static_3d_geometry(source_positions, module_centers, module_orientations, det_nx_per_module: int,
                             det_ny_per_module: int, det_spacing_x: float, det_spacing_y: float, source_module_mask,
                             M_gantry, b_gantry, active_sources, device: Optional[torch.device] = None):
    """src, dst [N, 3] float32 over all active (frame, source, module, pixel), built on
    the CPU and moved to ``device``."""
    device = torch.device("cpu") if device is None else device
    sp, mc, mo = source_positions.cpu(), module_centers.cpu(), module_orientations.cpu()
    Mg, bg = M_gantry.cpu(), b_gantry.cpu()
    mask, active = source_module_mask.cpu().to(torch.bool), active_sources.cpu().to(torch.bool)
    pix = _module_pixels_3d(mc, mo, det_nx_per_module, det_ny_per_module, det_spacing_x, det_spacing_y)
    all_src, all_dst = [], []
    for f in range(Mg.shape[0]):
        Mf, bf = Mg[f], bg[f]
        for s in range(sp.shape[0]):
            if not bool(active[f, s]):
                continue
            mods = torch.nonzero(mask[s], as_tuple=False).flatten()
            if mods.numel() == 0:
                continue
            src_image = Mf @ sp[s] + bf
            pixel_image = pix[mods].reshape(-1, 3) @ Mf.transpose(0, 1) + bf
            all_src.append(src_image.unsqueeze(0).expand_as(pixel_image))
            all_dst.append(pixel_image)
    if not all_src:
        e = torch.empty((0, 3), dtype=torch.float32, device=device)
        return e, e.clone()
    return torch.cat(all_src, dim=0).to(device), torch.cat(all_dst, dim=0).to(device)


class StaticCTProjector3D(CTProjector3DModule):
    """3D static-gantry projector (staticct_projector_3d.py:166-270)."""

    def __init__(self, n_x, n_y, n_z, M_gantry, b_gantry, source_positions, module_centers, module_orientations,
                 det_nx_per_module, det_ny_per_module, det_spacing_x, det_spacing_y, source_module_mask,
                 active_sources, M, b, backend="cuda", device: Optional[torch.device] = None,
                 precomputed_intersections: bool = False, tvals=None, use_compression: bool = True, *,
                 semantics: str = "torch"):
        if device is None:
            device = torch.device("cuda")
        src_all, dst_all = build_static_3d_geometry(source_positions=source_positions, module_centers=module_centers,
                                                    module_orientations=module_orientations,
                                                    det_nx_per_module=det_nx_per_module,
                                                    det_ny_per_module=det_ny_per_module, det_spacing_x=det_spacing_x,
                                                    det_spacing_y=det_spacing_y,
                                                    source_module_mask=source_module_mask, M_gantry=M_gantry,
                                                    b_gantry=b_gantry, active_sources=active_sources, device=device)
        super().__init__(n_x=n_x, n_y=n_y, n_z=n_z, M=M.to(device), b=b.to(device), src=src_all, dst=dst_all,
                         backend=backend, device=device, precomputed_intersections=precomputed_intersections,
                         tvals=tvals, use_compression=use_compression, semantics=semantics,
                         det_shape=(det_nx_per_module, det_ny_per_module))
        self.M_gantry, self.b_gantry = M_gantry, b_gantry
        self.source_positions, self.module_centers = source_positions, module_centers
        self.module_orientations = module_orientations
        self.det_nx_per_module, self.det_ny_per_module = det_nx_per_module, det_ny_per_module
        self.det_spacing_x, self.det_spacing_y = det_spacing_x, det_spacing_y
        self.source_module_mask, self.active_sources = source_module_mask, active_sources


def build_uniform_static_3d_geometry(n_source: int, source_radius: float, source_z_offset: float, n_module: int,
                                     module_radius: float, module_z_offset: float, modules_per_source: int,
                                     device: Optional[torch.device] = None):
    """Source ring at z = source_z_offset, inward-facing module ring at
    z = module_z_offset, each source illuminating its ``modules_per_source``
    farthest modules (staticct_projector_3d.py:378-459).  The selection uses
    ``torch.topk`` as the reference does; equal distances resolve as topk
    resolves them."""
    device = torch.device("cpu") if device is None else device
    angles_source = torch.linspace(0, 2 * math.pi - 2 * math.pi / n_source, n_source)
    source_positions = torch.tensor([[source_radius * math.cos(t), source_radius * math.sin(t), source_z_offset]
                                     for t in angles_source.tolist()], dtype=torch.float32)
    angles_module = torch.linspace(0, 2 * math.pi - 2 * math.pi / n_module, n_module)
    centers, orients = [], []
    up = torch.tensor([0.0, 0.0, 1.0], dtype=torch.float32)
    for t in angles_module.tolist():
        cx, cy = module_radius * math.cos(t), module_radius * math.sin(t)
        centers.append([cx, cy, module_z_offset])
        nn = math.sqrt(cx * cx + cy * cy)
        nx, ny, nz = (cx / nn, cy / nn, 0.0) if nn >= 1e-12 else (1.0, 0.0, 0.0)
        normal = torch.tensor([nx, ny, nz], dtype=torch.float32)
        side = torch.linalg.cross(up, normal)
        side_len = side.norm()
        if side_len < 1e-12:
            side, side_len = torch.tensor([1.0, 0.0, 0.0], dtype=torch.float32), 1.0
        side = side / side_len
        up_ortho = torch.linalg.cross(normal, side)
        orients.append(torch.stack([side, up_ortho, normal], dim=1))
    module_centers = torch.tensor(centers, dtype=torch.float32)
    module_orientations = torch.stack(orients, dim=0)
    mask = torch.zeros((n_source, n_module), dtype=torch.bool)
    for i in range(n_source):
        dist = torch.norm(module_centers - source_positions[i], dim=1)
        _, far = torch.topk(dist, modules_per_source, largest=True)
        mask[i, far] = True
    return (source_positions.to(device), module_centers.to(device), module_orientations.to(device),
            mask.to(device))


class UniformStaticCTProjector3D(StaticCTProjector3D):
    """Ring StaticCT projector in 3D (staticct_projector_3d.py:462-539)."""

    def __init__(self, n_x, n_y, n_z, n_source, source_radius, source_z_offset, n_module, module_radius,
                 det_nx_per_module, det_ny_per_module, det_spacing_x, det_spacing_y, module_z_offset, M_gantry,
                 b_gantry, active_sources, modules_per_source, M, b, backend="cuda", precomputed_intersections=False,
                 tvals=None, use_compression=True, device: Optional[torch.device] = None, *,
                 semantics: str = "torch"):
        sp, mc, mo, mask = build_uniform_static_3d_geometry(n_source=n_source, source_radius=source_radius,
                                                            source_z_offset=source_z_offset, n_module=n_module,
                                                            module_radius=module_radius,
                                                            module_z_offset=module_z_offset,
                                                            modules_per_source=modules_per_source)
        super().__init__(n_x=n_x, n_y=n_y, n_z=n_z, M_gantry=M_gantry, b_gantry=b_gantry, source_positions=sp,
                         module_centers=mc, module_orientations=mo, det_nx_per_module=det_nx_per_module,
                         det_ny_per_module=det_ny_per_module, det_spacing_x=det_spacing_x,
                         det_spacing_y=det_spacing_y, source_module_mask=mask, active_sources=active_sources, M=M,
                         b=b, backend=backend, device=device, precomputed_intersections=precomputed_intersections,
                         tvals=tvals, use_compression=use_compression, semantics=semantics)
        self.n_source, self.source_radius, self.source_z_offset = n_source, source_radius, source_z_offset
        self.n_module, self.module_radius, self.module_z_offset = n_module, module_radius, module_z_offset
        self.modules_per_source = modules_per_source
