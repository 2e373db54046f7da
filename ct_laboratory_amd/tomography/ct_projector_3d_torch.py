"""The names of the reference's pure-torch 3D projector
(ct_laboratory/tomography/ct_projector_3d_torch.py:1-420), computed by the HIP
kernels with the torch projector's segment rule (sorted crossings, world-space
lengths, midpoints rounded half to even; ct_projector_3d_torch.py:78-386).

The arguments and results are the reference's:

* ``compress_tvals_to_uint16`` / ``decompress_tvals_from_uint16``: the torch
  compression rule (round half to even, INF -> 10.0) and its float decode
  (cumsum, times scale, plus start; :4-75).
* ``compute_intersections_3d_torch``: [n_ray, n_x+n_y+n_z+3] sorted crossings,
  INF padded (:78-158).
* ``forward_project_3d_torch`` / ``back_project_3d_torch``: a float table or a
  compressed (uint16, start, scale) tuple, decompressed chunk by chunk as the
  reference does (:161-386), then the table consumers.

Deliberate differences: the inputs must be HIP tensors (there is no CPU
projector), a 4D volume or 2D sinogram is projected as a batch, and a table of
``None`` projects on the fly with the same rule, where the reference fails on
``None.shape`` (:186-187, :303-304; ``CTProjector3DFunction`` passes ``None``,
:396 and :404).
"""
from __future__ import annotations

import torch

from .. import _C
from .ct_projector_3d_hip import (_collapse, _f32c, _prepare_M_inv, back_project_3d_on_the_fly_hip,
                                  compress_tvals_3d_hip, compute_intersections_3d_hip,
                                  forward_project_3d_on_the_fly_hip)
from .ct_projector_3d_module import decompress_tvals_from_uint16

__all__ = ["compress_tvals_to_uint16", "decompress_tvals_from_uint16", "compute_intersections_3d_torch",
           "forward_project_3d_torch", "back_project_3d_torch", "CTProjector3DFunction", "CTProjector3D"]

# Table bytes decompressed at a time (the reference decompresses 65,536 rays at a time).
_CHUNK_BYTES = 1 << 30


def compress_tvals_to_uint16(tvals: torch.Tensor, M: torch.Tensor, src: torch.Tensor, dst: torch.Tensor):
    """(int16 deltas [N, K] holding uint16 values, starts [N], scales [N]) of a
    sorted crossing table, with the torch rule (ct_projector_3d_torch.py:4-65)."""
    if tvals.numel() == 0:
        return tvals.to(torch.int16), torch.empty(0, device=tvals.device), torch.empty(0, device=tvals.device)
    return compress_tvals_3d_hip(tvals, M, src, dst, semantics="torch")


def compute_intersections_3d_torch(n_x: int, n_y: int, n_z: int, M, b, src, dst) -> torch.Tensor:
    """Sorted crossings t in [0, 1] of every plane, INF padded
    (ct_projector_3d_torch.py:78-158)."""
    return compute_intersections_3d_hip(n_x, n_y, n_z, M, b, src, dst)


def _chunks(t_sorted_arg, n_ray):
    """(r0, r1, float table of rays r0..r1) over the rays."""
    if isinstance(t_sorted_arg, (tuple, list)):
        t_u16, t_start, t_scale = t_sorted_arg
        step = max(1, _CHUNK_BYTES // max(1, 4 * int(t_u16.shape[1])))
        for r0 in range(0, n_ray, step):
            r1 = min(n_ray, r0 + step)
            yield r0, r1, decompress_tvals_from_uint16(t_u16[r0:r1], t_start[r0:r1], t_scale[r0:r1]).contiguous()
    else:
        yield 0, n_ray, _f32c(t_sorted_arg)


def _no_table(t) -> bool:
    return t is None or (isinstance(t, (tuple, list)) and len(t) == 0) or (
        isinstance(t, torch.Tensor) and t.numel() == 0)


def forward_project_3d_torch(volume, t_sorted_arg, M, b, src, dst) -> torch.Tensor:
    """Line integrals [n_ray] ([B, n_ray] for a [B, X, Y, Z] volume)
    (ct_projector_3d_torch.py:161-274)."""
    M, b = _collapse(M, b)
    if _no_table(t_sorted_arg):
        return forward_project_3d_on_the_fly_hip(volume, M, b, src, dst, semantics="torch")
    vol = _f32c(volume)
    src, dst, bb = _f32c(src), _f32c(dst), _f32c(b)
    M_inv = _prepare_M_inv(M)
    batch = vol.shape[0] if vol.dim() == 4 else 1
    n_ray = int(src.shape[0])
    out = torch.empty((batch, n_ray), device=vol.device, dtype=torch.float32)
    for r0, r1, tv in _chunks(t_sorted_arg, n_ray):
        if (r0, r1) == (0, n_ray):
            _C.forward_project_3d_cuda(vol, tv, src, dst, M_inv, bb, out=out)
        else:
            out[:, r0:r1] = _C.forward_project_3d_cuda(vol, tv, src[r0:r1], dst[r0:r1], M_inv, bb)
    return out.squeeze(0) if volume.dim() == 3 else out


def back_project_3d_torch(sinogram, t_sorted_arg, M, b, src, dst, n_x: int, n_y: int, n_z: int) -> torch.Tensor:
    """Back projection [n_x, n_y, n_z] ([B, ...] for a [B, n_ray] sinogram)
    (ct_projector_3d_torch.py:277-386)."""
    M, b = _collapse(M, b)
    if _no_table(t_sorted_arg):
        return back_project_3d_on_the_fly_hip(sinogram, src, dst, M, b, n_x, n_y, n_z, semantics="torch")
    sino = _f32c(sinogram)
    sino2 = sino.reshape(1, -1) if sino.dim() == 1 else sino
    src, dst, bb = _f32c(src), _f32c(dst), _f32c(b)
    M_inv = _prepare_M_inv(M)
    out = torch.zeros((sino2.shape[0], int(n_x), int(n_y), int(n_z)), device=sino.device, dtype=torch.float32)
    for r0, r1, tv in _chunks(t_sorted_arg, int(src.shape[0])):
        _C.back_project_3d_cuda(sino2[:, r0:r1].contiguous(), tv, src[r0:r1], dst[r0:r1], M_inv, bb, n_x, n_y, n_z,
                                out=out, accumulate=True)
    return out.squeeze(0) if sinogram.dim() == 1 else out


class CTProjector3DFunction(torch.autograd.Function):
    """forward(volume, M, b, src, dst): the on-the-fly torch-rule projection;
    backward: its back projection (ct_projector_3d_torch.py:389-407)."""

    @staticmethod
    def forward(ctx, volume, M, b, src, dst):
        ctx.save_for_backward(M, b, src, dst)
        ctx.volume_shape = tuple(volume.shape[-3:])
        return forward_project_3d_torch(volume, None, M, b, src, dst)

    @staticmethod
    def backward(ctx, grad_output):
        M, b, src, dst = ctx.saved_tensors
        n_x, n_y, n_z = ctx.volume_shape
        return back_project_3d_torch(grad_output.contiguous(), None, M, b, src, dst, n_x, n_y, n_z), None, None, None, \
            None


class CTProjector3D(torch.nn.Module):
    """nn.Module over CTProjector3DFunction with M, b, src, dst as buffers
    (ct_projector_3d_torch.py:410-418)."""

    def __init__(self, M, b, src, dst):
        super().__init__()
        self.register_buffer("M", M)
        self.register_buffer("b", b)
        self.register_buffer("src", src)
        self.register_buffer("dst", dst)

    def forward(self, volume):
        return CTProjector3DFunction.apply(volume, self.M, self.b, self.src, self.dst)
