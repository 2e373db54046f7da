"""3D projector functions over the HIP kernels.

Same public names, argument order and squeeze rules as the reference's
tomography/ct_projector_3d_cuda.py:1-323 (each ``*_hip`` function also has
its reference ``*_cuda`` name), with these deliberate differences:

* ``back_project_3d_hip(sinogram, None, ...)`` passes ``inverse(M)`` to the
  on-the-fly kernel once.  The reference passes the already inverted matrix to
  a function that inverts it again (ct_projector_3d_cuda.py:210 and :319), so
  its default on-the-fly backward is not the adjoint when M != I.
* ``semantics`` selects the segment rule of the on-the-fly kernels:
  ``"torch"`` (world-space segment lengths, the rule of the reference's torch
  projector, ct_projector_3d_torch.py:161-274) or ``"otf"`` (the reference's
  on-the-fly CUDA kernel: (t1-t0)*|dst-src| with 0 and 1 inserted,
  src/ct_projector_3d.cu:834-1012).  The default here is ``"otf"``, as in the
  reference's ``_cuda`` functions; the nn.Modules default to ``"torch"``.
* ``det_shape=(rows, cols)`` declares rays ordered [view][rows][cols] (one
  source per view).  It only changes the launch schedule; results are the same.

There is no CPU path: every function needs HIP tensors and the built library.
"""
from __future__ import annotations

import torch

from .. import _C
from .. import _native as N

__all__ = [
    "compute_intersections_3d_hip", "compress_tvals_3d_hip",
    "forward_project_3d_hip", "forward_project_3d_compressed_hip",
    "back_project_3d_hip", "back_project_3d_compressed_hip",
    "forward_project_3d_on_the_fly_hip", "back_project_3d_on_the_fly_hip",
    "voxel_diagonal_length",
    # reference names
    "compute_intersections_3d_cuda", "compress_tvals_3d_cuda",
    "forward_project_3d_cuda", "forward_project_3d_compressed_cuda",
    "back_project_3d_cuda", "back_project_3d_compressed_cuda",
    "forward_project_3d_on_the_fly_cuda", "back_project_3d_on_the_fly_cuda",
]

_MODES = {"torch": N.MODE_TORCH, "otf": N.MODE_OTF, "march": N.MODE_MARCH}


def semantics_mode(semantics: str) -> int:
    """Kernel mode of a segment rule name ("torch", "otf", or "march": the
    reference's 2D on-the-fly kernel, 2D rays only)."""
    if semantics not in _MODES:
        raise ValueError(f"semantics must be 'torch', 'otf' or 'march', got {semantics!r}")
    return _MODES[semantics]


def _collapse(M: torch.Tensor, b: torch.Tensor, n_ray: int | None = None):
    """Per-ray affine stacks [n_ray,3,3] collapse to their first entry
    (ct_projector_3d_cuda.py:75-80, 130-132)."""
    if M.dim() > 2:
        M = M[0]
    if b.dim() > 1:
        b = b[0]
    return M, b


def _prepare_M_inv(M: torch.Tensor) -> torch.Tensor:
    """Row-major float32 inverse of M on M's device (ct_projector_3d_cuda.py:42-48)."""
    if not M.is_cuda:
        raise TypeError("M must be a HIP (cuda) tensor")
    return _C.cached_inverse(M)


def _f32c(t: torch.Tensor) -> torch.Tensor:
    return t.to(torch.float32).contiguous()


def _empty_or_none(t) -> bool:
    if t is None:
        return True
    if isinstance(t, (tuple, list)):
        return len(t) == 0
    return t.numel() == 0


def voxel_diagonal_length(M: torch.Tensor) -> float:
    """Largest body diagonal |M d| of a voxel, d in {(1,1,1), (1,1,-1), (1,-1,1),
    (-1,1,1)} (ct_projector_3d_cuda.py:92-104; ct_projector_3d_torch.py:27-35),
    evaluated with the same float32 torch operations on the CPU."""
    M, _ = _collapse(M, M.new_zeros(3))
    M = M.detach().to("cpu", torch.float32)
    d = torch.tensor([[1, 1, 1], [1, 1, -1], [1, -1, 1], [-1, 1, 1]], dtype=torch.float32).t()
    return float(torch.norm(M @ d, dim=0).max().item())


# ---------------------------------------------------------------------------
# Crossing tables
# ---------------------------------------------------------------------------


def compute_intersections_3d_hip(n_x, n_y, n_z, M, b, src, dst, *, K=None):
    """Sorted crossings t in [0,1] per ray, INF padded: [n_ray, (n_x+1)+(n_y+1)+(n_z+1)]
    (ct_projector_3d_cuda.py:62-82).  ``K`` trims the table width (it must be at
    least the largest crossing count, see ``_C.count_crossings``)."""
    M, b = _collapse(M, b)
    return _C.compute_intersections_3d(n_x, n_y, n_z, _f32c(src), _f32c(dst), _prepare_M_inv(M), _f32c(b), K=K)


def compress_tvals_3d_hip(tvals, M, src, dst, *, semantics: str = "cuda"):
    """uint16 delta compression of a crossing table (ct_projector_3d_cuda.py:84-106).

    ``semantics="cuda"`` reproduces the reference's GPU kernel (roundf, INF ->
    delta 0, src/ct_projector_3d.cu:160-212); ``"torch"`` reproduces
    ``compress_tvals_to_uint16`` (round half to even, INF -> 10.0,
    ct_projector_3d_torch.py:4-65)."""
    sem = {"cuda": N.COMPRESS_CUDA, "torch": N.COMPRESS_TORCH}[semantics]
    return _C.compress_tvals_3d_cuda(_f32c(tvals), _f32c(src), _f32c(dst), voxel_diagonal_length(M), semantics=sem)


# ---------------------------------------------------------------------------
# Forward / back with a crossing table, compressed table or none
# ---------------------------------------------------------------------------


def _check_inputs(**tensors):
    for name, t in tensors.items():
        assert t.is_contiguous(), f"{name} must be C-contiguous"
        assert t.dtype == torch.float32, f"{name} must be float32"


def forward_project_3d_hip(volume, tvals, src, dst, M, b, *, semantics: str = "otf", det_shape=None):
    """Line integrals [B, n_ray] (or [n_ray] for a 3D volume) (ct_projector_3d_cuda.py:112-160).

    ``tvals``: float table, (t_uint16, t_start, t_scale) tuple, or None/empty for
    the on-the-fly traversal.  A float64 volume (on-the-fly only) is projected
    with double values over the same segments."""
    M, b = _collapse(M, b)
    if _empty_or_none(tvals):
        return forward_project_3d_on_the_fly_hip(volume, M, b, src, dst, semantics=semantics, det_shape=det_shape)
    _check_inputs(volume=volume, src=src, dst=dst, M=M, b=b)
    M_inv = _prepare_M_inv(M)
    if isinstance(tvals, (tuple, list)):
        t_u16, t_start, t_scale = tvals
        out = _C.forward_project_3d_compressed_cuda(volume, t_u16, t_start, t_scale, src, dst, M_inv, b)
    else:
        assert tvals.dtype == torch.float32, "tvals must be float32"
        out = _C.forward_project_3d_cuda(volume, tvals, src, dst, M_inv, b)
    return out.squeeze(0) if volume.dim() == 3 else out


def forward_project_3d_compressed_hip(volume, t_uint16, t_start, t_scale, M, b, src, dst):
    """Forward projection from a compressed table (ct_projector_3d_cuda.py:162-179)."""
    M, b = _collapse(M, b)
    out = _C.forward_project_3d_compressed_cuda(volume, t_uint16, t_start, t_scale, src, dst, _prepare_M_inv(M), b)
    return out.squeeze(0) if volume.dim() == 3 else out


def back_project_3d_hip(sinogram, tvals, src, dst, M, b, n_x, n_y, n_z, *, semantics: str = "otf",
                        det_shape=None, brick_plan=None, table_runs=None):
    """Adjoint of ``forward_project_3d_hip`` with the same ``tvals``: [B, n_x, n_y, n_z]
    (or [n_x, n_y, n_z] for a 1D sinogram) (ct_projector_3d_cuda.py:186-225).

    With no table, ``brick_plan`` (from ``ConeBeamPlan``) selects the LDS-brick
    back projector; otherwise the atomic one runs.  With a table,
    ``table_runs`` (``_C.table_runs`` of that table, rays and map) selects the
    table brick back projector; without it, large tables get cached run lists
    and small ones the per-segment atomic kernels."""
    M, b = _collapse(M, b)
    if _empty_or_none(tvals):
        return back_project_3d_on_the_fly_hip(sinogram, src, dst, M, b, n_x, n_y, n_z, semantics=semantics,
                                              det_shape=det_shape, brick_plan=brick_plan)
    _check_inputs(sinogram=sinogram)
    M_inv = _prepare_M_inv(M)
    if table_runs is not None:
        out = _C.back_project_3d_table_brick(sinogram, tvals, src, dst, table_runs, M_inv, b, n_x, n_y, n_z)
    elif isinstance(tvals, (tuple, list)):
        t_u16, t_start, t_scale = tvals
        out = _C.back_project_3d_compressed_cuda(sinogram, t_u16, t_start, t_scale, src, dst, M_inv, b,
                                                 n_x, n_y, n_z)
    else:
        assert tvals.dtype == torch.float32, "tvals must be float32"
        out = _C.back_project_3d_cuda(sinogram, tvals, src, dst, M_inv, b, n_x, n_y, n_z)
    return out.squeeze(0) if sinogram.dim() == 1 else out


def back_project_3d_compressed_hip(sinogram, t_uint16, t_start, t_scale, M, b, src, dst, n_x, n_y, n_z):
    """Back projection from a compressed table (ct_projector_3d_cuda.py:228-247)."""
    M, b = _collapse(M, b)
    out = _C.back_project_3d_compressed_cuda(sinogram, t_uint16, t_start, t_scale, src, dst, _prepare_M_inv(M), b,
                                             n_x, n_y, n_z)
    return out.squeeze(0) if sinogram.dim() == 1 else out


# ---------------------------------------------------------------------------
# On-the-fly traversal
# ---------------------------------------------------------------------------


def forward_project_3d_on_the_fly_hip(volume, M, b, src, dst, *, semantics: str = "otf", det_shape=None):
    """On-the-fly forward projection (ct_projector_3d_cuda.py:253-287).  A
    float64 volume gives float64 line integrals over the same segments."""
    M, b = _collapse(M, b)
    M, b = _f32c(M), _f32c(b)
    _check_inputs(src=src, dst=dst)
    if volume.dtype == torch.float64:
        out = _C.forward_project_3d_on_the_fly_f64(volume.contiguous(), src, dst, _prepare_M_inv(M), b,
                                                   mode=semantics_mode(semantics))
    else:
        _check_inputs(volume=volume)
        out = _C.forward_project_3d_on_the_fly_cuda(volume, src, dst, _prepare_M_inv(M), b,
                                                    mode=semantics_mode(semantics), det_shape=det_shape)
    return out.squeeze(0) if volume.dim() == 3 else out


def back_project_3d_on_the_fly_hip(sinogram, src, dst, M, b, n_x, n_y, n_z, *, semantics: str = "otf",
                                   det_shape=None, brick_plan=None):
    """On-the-fly back projection, the exact transpose of the forward
    (ct_projector_3d_cuda.py:290-323; ``M`` is the index->world matrix, inverted
    here once)."""
    M, b = _collapse(M, b)
    M, b = _f32c(M), _f32c(b)
    _check_inputs(src=src, dst=dst)
    mode = semantics_mode(semantics)
    M_inv = _prepare_M_inv(M)
    if sinogram.dtype == torch.float64:  # double values over the same segments
        out = _C.back_project_3d_on_the_fly_f64(sinogram.contiguous(), src, dst, M_inv, b, n_x, n_y, n_z, mode=mode)
        return out.squeeze(0) if sinogram.dim() == 1 else out
    _check_inputs(sinogram=sinogram)
    if brick_plan is not None and det_shape is not None:
        out = _C.back_project_3d_brick(sinogram, src, dst, det_shape, brick_plan, M_inv, b, n_x, n_y, n_z, mode=mode)
    else:
        out = _C.back_project_3d_on_the_fly_cuda(sinogram, src, dst, M_inv, b, n_x, n_y, n_z, mode=mode,
                                                 det_shape=det_shape)
    return out.squeeze(0) if sinogram.dim() == 1 else out


class ConeBeamPlan:
    """Geometry-only data of the brick back projector for rays ordered
    [view][rows][cols] with one source per view: the per-view detector frames,
    the brick plan and its filed unit lists.  Built once per geometry on the
    rays' device."""

    # Plans above this size fall back to the atomic back projector.
    max_plan_bytes = 1 << 30

    def __init__(self, src, dst, M, b, shape, det_shape):
        M, b = _collapse(M, b)
        self.det_shape = (int(det_shape[0]), int(det_shape[1]))
        self.shape = tuple(int(s) for s in shape)
        rows, cols = self.det_shape
        n_ray = int(src.shape[0])
        if rows * cols == 0 or n_ray % (rows * cols) != 0:
            raise ValueError("ConeBeamPlan: n_ray must be a multiple of rows*cols")
        self.n_view = n_ray // (rows * cols)
        ints = N.load().ctlab_brick_plan_ints_3d(*self.shape, self.n_view)
        self.plan = None
        self.frames = None
        self.lists = None
        if ints <= 0 or 4 * ints > self.max_plan_bytes or n_ray > 0xFFFFFFFF - 4096:
            return
        M_inv = _prepare_M_inv(M)
        self.frames = _C.cone_frames(_f32c(src), _f32c(dst), self.det_shape, M_inv, _f32c(b), self.shape)
        self.plan = _C.brick_plan(self.frames, self.det_shape, self.shape)
        # the plan's filed unit lists (about 23 bytes per ray at 256^3), registered for the plan, unless
        # they would take more than half of the free device memory
        free, _ = torch.cuda.mem_get_info(src.device)
        self.lists = _C.brick_filed_lists(_f32c(src), _f32c(dst), self.det_shape, self.plan, M_inv, _f32c(b),
                                          self.shape, max_bytes=free // 2)

    @property
    def usable(self) -> bool:
        return self.plan is not None


# Reference names (ct_projector_3d_cuda.py:29-38)
compute_intersections_3d_cuda = compute_intersections_3d_hip
compress_tvals_3d_cuda = compress_tvals_3d_hip
forward_project_3d_cuda = forward_project_3d_hip
forward_project_3d_compressed_cuda = forward_project_3d_compressed_hip
back_project_3d_cuda = back_project_3d_hip
back_project_3d_compressed_cuda = back_project_3d_compressed_hip
forward_project_3d_on_the_fly_cuda = forward_project_3d_on_the_fly_hip
back_project_3d_on_the_fly_cuda = back_project_3d_on_the_fly_hip
