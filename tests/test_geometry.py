"""CPU tests of the geometry builders and the host-side projector layer.

The builders are checked through what the geometry fixes (distances, detector
planes, pixel pitch, ray order and counts), not by retyping their formulas.
"""
import math

import numpy as np
import pytest
import torch

from ct_laboratory_amd.tomography import (build_circular_sequence, build_conebeam_3d_geometry,
                                          build_fanbeam_2d_geometry, build_helical_sequence,
                                          build_static_2d_geometry, build_static_3d_geometry,
                                          build_uniform_static_2d_geometry, build_uniform_static_3d_geometry,
                                          standard_image_transform_2d, standard_image_transform_3d)
from ct_laboratory_amd.tomography._backend import check_backend


def test_standard_transforms_map_centre_to_origin():
    M, b = standard_image_transform_3d(5, 8, 3, 0.5, 2.0, 1.25)
    centre = torch.tensor([2.0, 3.5, 1.0])
    assert torch.allclose(M @ centre + b, torch.zeros(3))
    assert torch.equal(torch.diag(M), torch.tensor([0.5, 2.0, 1.25]))
    M2, b2 = standard_image_transform_2d(7, 4, 1.0, 3.0)
    assert torch.allclose(M2 @ torch.tensor([3.0, 1.5]) + b2, torch.zeros(2))


def test_fanbeam_geometry():
    n_view, n_det, sid, sdd, sp = 36, 17, 1000.0, 1500.0, 1.25
    src, dst = build_fanbeam_2d_geometry(n_view, n_det, sid, sdd, det_spacing=sp, angle_offset=0.1)
    assert src.shape == dst.shape == (n_view * n_det, 2) and src.dtype == torch.float32
    s = src.double().view(n_view, n_det, 2)
    d = dst.double().view(n_view, n_det, 2)
    assert torch.allclose(s.norm(dim=-1), torch.full((n_view, n_det), sid, dtype=torch.float64), rtol=1e-6)
    assert (s == s[:, :1]).all(), "one source per view"
    # pixels equally spaced along a line perpendicular to the central ray, at distance sdd
    step = d[:, 1:] - d[:, :-1]
    assert torch.allclose(step.norm(dim=-1), torch.full_like(step[..., 0], sp), rtol=1e-4)
    centre = d[:, n_det // 2]
    axis = (centre - s[:, 0]) / (centre - s[:, 0]).norm(dim=-1, keepdim=True)
    assert torch.allclose((centre - s[:, 0]).norm(dim=-1), torch.full((n_view,), sdd, dtype=torch.float64),
                          rtol=1e-6)
    assert (step * axis[:, None]).sum(-1).abs().max() < 1e-3
    # the central ray passes through the isocentre
    assert (s[:, 0] + axis * sid).norm(dim=-1).max() < 1e-3
    # views in increasing angle from the offset
    ang = torch.atan2(s[:, 0, 1], s[:, 0, 0])
    expect = torch.remainder(0.1 + 2 * math.pi * torch.arange(n_view, dtype=torch.float64) / n_view + math.pi,
                             2 * math.pi) - math.pi
    assert torch.allclose(ang, expect, atol=1e-5)


def test_conebeam_geometry():
    n_view, nx, ny, sid, sdd, sp = 8, 6, 5, 500.0, 800.0, 1.5
    src, dst = build_conebeam_3d_geometry(n_view, nx, ny, sid, sdd, det_spacing=sp)
    assert src.shape == (n_view * ny * nx, 3)
    d = dst.double().view(n_view, ny, nx, 3)
    s = src.double().view(n_view, ny, nx, 3)
    assert (s == s[:, :1, :1]).all()
    assert torch.allclose(s[..., :2].norm(dim=-1), torch.full(s.shape[:3], sid, dtype=torch.float64), rtol=1e-6)
    # v (rows) along +z, u (columns) in-plane, both at the pitch
    assert torch.allclose(d[:, 1:, :, 2] - d[:, :-1, :, 2], torch.full_like(d[:, 1:, :, 2], sp), atol=1e-4)
    du = d[:, :, 1:] - d[:, :, :-1]
    assert torch.allclose(du.norm(dim=-1), torch.full_like(du[..., 0], sp), atol=1e-4)
    assert du[..., 2].abs().max() < 1e-6


def test_static_2d_and_uniform_ring():
    n_frame, n_source, n_module, n_col = 5, 4, 12, 3
    Mg = torch.eye(2).repeat(n_frame, 1, 1)
    bg = torch.zeros(n_frame, 2)
    active = torch.zeros(n_frame, n_source, dtype=torch.bool)
    active[torch.arange(n_frame), torch.arange(n_frame) % n_source] = True
    sp, mc, mo, mask = build_uniform_static_2d_geometry(n_source, 100.0, n_module, 90.0, n_col, 2.0, Mg, bg, active)
    assert torch.allclose(sp.norm(dim=1), torch.full((n_source,), 100.0))
    assert torch.allclose(mc.norm(dim=1), torch.full((n_module,), 90.0))
    # orthonormal orientations whose second column (module normal) points along the centre
    eye = torch.eye(2).expand(n_module, 2, 2)
    assert torch.allclose(mo.transpose(1, 2) @ mo, eye, atol=1e-6)
    assert torch.allclose(mo[:, :, 1], mc / mc.norm(dim=1, keepdim=True), atol=1e-6)
    src, dst = build_static_2d_geometry(sp, mc, mo, n_col, 2.0, mask, Mg, bg, active)
    assert src.shape == (n_frame * n_module * n_col, 2)
    # order: frame -> module -> column; frame f uses source f mod n_source
    s = src.view(n_frame, n_module, n_col, 2)
    for f in range(n_frame):
        assert torch.equal(s[f], sp[f % n_source].expand(n_module, n_col, 2))
    d = dst.view(n_frame, n_module, n_col, 2)
    assert torch.allclose((d[:, :, 1] - d[:, :, 0]).norm(dim=-1), torch.full((n_frame, n_module), 2.0), atol=1e-4)
    # a per-frame gantry translation moves every point of that frame
    bg2 = bg.clone()
    bg2[2] = torch.tensor([5.0, -3.0])
    src2, dst2 = build_static_2d_geometry(sp, mc, mo, n_col, 2.0, mask, Mg, bg2, active)
    assert torch.allclose(dst2.view(n_frame, -1, 2)[2] - d.view(n_frame, -1, 2)[2], torch.tensor([5.0, -3.0]))
    # a source that illuminates no module contributes no rays
    mask2 = mask.clone()
    mask2[0] = False
    src3, _ = build_static_2d_geometry(sp, mc, mo, n_col, 2.0, mask2, Mg, bg, active)
    assert src3.shape[0] == (n_frame - 2) * n_module * n_col  # frames 0 and 4 use source 0


def test_sequences_and_uniform_3d():
    Mg, bg, act = build_circular_sequence(7, 3)
    assert torch.equal(act.int().argmax(1), torch.arange(7) % 3) and (act.sum(1) == 1).all()
    assert (Mg == torch.eye(3)).all() and (bg == 0).all()
    Mh, bh, acth = build_helical_sequence(5, 2, 10.0)
    assert torch.allclose(bh[:, 2], torch.tensor([-10.0, -5.0, 0.0, 5.0, 10.0]))
    sp, mc, mo, mask = build_uniform_static_3d_geometry(6, 50.0, 2.0, 20, 45.0, -1.0, 7)
    assert (mask.sum(1) == 7).all()
    # every selected module is at least as far from its source as every unselected one
    for i in range(6):
        dist = (mc - sp[i]).norm(dim=1)
        assert dist[mask[i]].min() >= dist[~mask[i]].max() - 1e-4
    assert torch.allclose(mo.transpose(1, 2) @ mo, torch.eye(3).expand(20, 3, 3), atol=1e-6)
    assert torch.allclose(sp[:, 2], torch.full((6,), 2.0)) and torch.allclose(mc[:, 2], torch.full((20,), -1.0))
    src, dst = build_static_3d_geometry(sp, mc, mo, 3, 2, 1.0, 1.5, mask, Mg[:6], bg[:6],
                                        build_circular_sequence(6, 6)[2])
    assert src.shape == (6 * 7 * 3 * 2, 3)
    # pixel order inside a module: u outer (spacing_x along the module's first axis), v inner
    d = dst.view(6, 7, 3, 2, 3)
    assert torch.allclose((d[:, :, :, 1] - d[:, :, :, 0]).norm(dim=-1), torch.full((6, 7, 3), 1.5), atol=1e-4)
    assert torch.allclose((d[:, :, 1, :] - d[:, :, 0, :]).norm(dim=-1), torch.full((6, 7, 2), 1.0), atol=1e-4)


def test_backend_names():
    assert check_backend("cuda") == "hip" and check_backend("hip") == "hip"
    with pytest.raises(ValueError, match="not provided"):
        check_backend("torch")
    with pytest.raises(ValueError):
        check_backend("cpu")


def test_projector_layer_requires_hip_tensors():
    """The product path has no CPU fallback: CPU tensors are refused loudly."""
    from ct_laboratory_amd.tomography import forward_project_3d_hip

    vol = torch.zeros(4, 4, 4)
    src = torch.zeros(2, 3)
    with pytest.raises((RuntimeError, TypeError)):
        forward_project_3d_hip(vol, None, src, src.clone(), torch.eye(3), torch.zeros(3))
