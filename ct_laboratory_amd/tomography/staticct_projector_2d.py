"""2D static-gantry (StaticCT) geometry and projectors
(ct_laboratory/tomography/staticct_projector_2d.py:9-299).

Rays: for every frame, every active source, every module the source
illuminates, every detector column of the module, in that order
(staticct_projector_2d.py:50-80).  The float32 tensor operations are the
reference's (the same matrix products per frame and module); the loop over
modules is batched per (frame, source).
"""
import math
from typing import Optional

import torch

from .ct_projector_2d_module import CTProjector2DModule

__all__ = ["build_static_2d_geometry", "StaticCTProjector2D", "build_uniform_static_2d_geometry",
           "UniformStaticCTProjector2D"]


def _module_pixels_2d(module_center_positions, module_orientations, det_n_col, det_spacing):
    """Gantry coordinates [n_module, det_n_col, 2] of every module pixel:
    Rm (offset, 0) + centre (staticct_projector_2d.py:40-48, 66)."""
    mid_col = (det_n_col - 1) / 2.0
    local = torch.tensor([[(ic - mid_col) * det_spacing, 0.0] for ic in range(det_n_col)], dtype=torch.float32,
                         device=module_center_positions.device)
    return torch.stack([local @ module_orientations[m].transpose(0, 1) + module_center_positions[m]
                        for m in range(module_center_positions.shape[0])])


def build_static_2d_geometry(source_positions, module_center_positions, module_orientations, det_n_col: int,
                             det_spacing: float, source_module_mask, M_gantry, b_gantry, active_sources):
    """src, dst [N, 2] float32 over all active (frame, source, module, column)."""
    device = source_positions.device
    n_frame = M_gantry.shape[0]
    n_source = source_positions.shape[0]
    pix = _module_pixels_2d(module_center_positions, module_orientations, det_n_col, det_spacing)
    mask = source_module_mask.to(torch.bool)
    active = active_sources.to(torch.bool).cpu()
    all_src, all_dst = [], []
    for f in range(n_frame):
        Mf, bf = M_gantry[f], b_gantry[f]
        for s in range(n_source):
            if not bool(active[f, s]):
                continue
            src_image = Mf @ source_positions[s] + bf
            mods = torch.nonzero(mask[s], as_tuple=False).flatten()
            if mods.numel() == 0:
                continue
            pixel_image = pix[mods].reshape(-1, 2) @ Mf.transpose(0, 1) + bf
            all_src.append(src_image.unsqueeze(0).expand_as(pixel_image))
            all_dst.append(pixel_image)
    if not all_src:
        e = torch.empty((0, 2), dtype=torch.float32, device=device)
        return e, e.clone()
    return torch.cat(all_src, dim=0), torch.cat(all_dst, dim=0)


class StaticCTProjector2D(CTProjector2DModule):
    """2D static-gantry projector (staticct_projector_2d.py:86-148)."""

    def __init__(self, n_row, n_col, M_gantry, b_gantry, source_positions, module_centers, module_orientations,
                 det_n_col, det_spacing, source_module_mask, active_sources, M, b, backend="cuda",
                 device: Optional[torch.device] = None, *, precomputed_intersections: bool = True,
                 semantics: str = "torch"):
        src_all, dst_all = build_static_2d_geometry(source_positions=source_positions,
                                                    module_center_positions=module_centers,
                                                    module_orientations=module_orientations, det_n_col=det_n_col,
                                                    det_spacing=det_spacing, source_module_mask=source_module_mask,
                                                    M_gantry=M_gantry, b_gantry=b_gantry,
                                                    active_sources=active_sources)
        if device is None:
            device = torch.device("cuda")
        super().__init__(n_row=n_row, n_col=n_col, M=M.to(device), b=b.to(device), src=src_all.to(device),
                         dst=dst_all.to(device), backend=backend, precomputed_intersections=precomputed_intersections,
                         semantics=semantics)


def build_uniform_static_2d_geometry(n_source, source_radius, n_module, module_radius, det_n_col, det_spacing,
                                     M_gantry, b_gantry, active_sources):
    """Sources and inward-facing modules on two circles, every source illuminating
    every module (staticct_projector_2d.py:154-214).  Returns source_positions
    [n_source,2], module_centers [n_module,2], module_orientations [n_module,2,2],
    source_module_mask [n_source,n_module]."""
    device = M_gantry.device
    angles_src = torch.linspace(0, 2 * math.pi - 2 * math.pi / n_source, n_source)
    source_positions = torch.tensor([[source_radius * math.cos(t), source_radius * math.sin(t)]
                                     for t in angles_src.tolist()], dtype=torch.float32, device=device)
    angles_mod = torch.linspace(0, 2 * math.pi - 2 * math.pi / n_module, n_module)
    centers, orients = [], []
    for t in angles_mod.tolist():
        cx, cy = module_radius * math.cos(t), module_radius * math.sin(t)
        centers.append([cx, cy])
        nrm = math.sqrt(cx * cx + cy * cy)
        nx, ny = (cx / nrm, cy / nrm) if nrm >= 1e-12 else (1.0, 0.0)
        orients.append([[-ny, nx], [nx, ny]])  # columns: local x = perpendicular, local y = normal
    module_centers = torch.tensor(centers, dtype=torch.float32, device=device)
    module_orientations = torch.tensor(orients, dtype=torch.float32, device=device)
    source_module_mask = torch.ones((n_source, n_module), dtype=torch.bool, device=device)
    return source_positions, module_centers, module_orientations, source_module_mask


class UniformStaticCTProjector2D(StaticCTProjector2D):
    """Ring StaticCT projector (staticct_projector_2d.py:217-299)."""

    def __init__(self, n_row, n_col, M, b, n_source, source_radius, n_module, module_radius, det_n_col, det_spacing,
                 M_gantry, b_gantry, active_sources, backend="cuda", device: Optional[torch.device] = None, *,
                 precomputed_intersections: bool = True, semantics: str = "torch"):
        if device is None:
            device = torch.device("cuda")
        sp, mc, mo, mask = build_uniform_static_2d_geometry(n_source=n_source, source_radius=source_radius,
                                                            n_module=n_module, module_radius=module_radius,
                                                            det_n_col=det_n_col, det_spacing=det_spacing,
                                                            M_gantry=M_gantry.cpu(), b_gantry=b_gantry.cpu(),
                                                            active_sources=active_sources.cpu())
        self_attrs = dict(n_source=n_source, source_radius=source_radius, n_module=n_module,
                          module_radius=module_radius, det_n_col=det_n_col, det_spacing=det_spacing)
        super().__init__(n_row=n_row, n_col=n_col, M_gantry=M_gantry.cpu(), b_gantry=b_gantry.cpu(),
                         source_positions=sp, module_centers=mc, module_orientations=mo, det_n_col=det_n_col,
                         det_spacing=det_spacing, source_module_mask=mask, active_sources=active_sources.cpu(),
                         M=M, b=b, backend=backend, device=device,
                         precomputed_intersections=precomputed_intersections, semantics=semantics)
        for k, v in self_attrs.items():
            setattr(self, k, v)
