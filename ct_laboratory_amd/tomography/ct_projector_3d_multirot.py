"""Rolling-window multi-rotation projector (ct_laboratory/tomography/ct_projector_3d_multirot.py:13-393).

A helical-like scan is modelled as ``n_rotations`` copies of one rotation's
geometry, each seeing a z-window of ``n_z_window`` slices of the full volume,
the window centre moving by ``shift_offset`` slices per rotation.  Windows
reaching past the volume are padded as the reference pads them:

- ``'repeat'``: the nearest edge slice, with the padded copies detached from
  the graph (no gradient reaches the edge slice through them);
- ``'reflect'``: reflection about the edge slice centres (index -i -> i,
  n-1+i -> n-1-i, repeated until in range), with slices farther than
  ``reflection_padding_slices`` outside set to zero when that is given;
- anything else: zeros.

forward(volume [X, Y, Z]) returns [n_rays, n_rotations]; back_project(sino
[n_rays, n_rotations], Z) returns [X, Y, Z].  Lists of volumes are projected
together (the reference's multi-volume subset path).

MI355X-first: instead of one projection per rotation (and per volume), all
windows are gathered with one index into a [n_rotations, X, Y, n_z_window]
batch and projected in one batched launch of the base projector each way;
the back projection's windows are folded back into the volume with one
index_add_.  With an OrderedSubsetProjector base this is one batched launch
per subset.
"""
from __future__ import annotations

import torch
from torch import nn

__all__ = ["MultiRotationProjector"]


class MultiRotationProjector(nn.Module):
    def __init__(self, base_projector, n_rotations, shift_offset, n_z_window, scan_center_slice=None,
                 padding_mode="repeat", reflection_padding_slices=None, verbose=False):
        super().__init__()
        self.base_projector = base_projector
        self.n_rotations = int(n_rotations)
        self.shift_offset = float(shift_offset)
        self.n_z_window = int(n_z_window)
        self.scan_center_slice = scan_center_slice
        self.padding_mode = padding_mode
        self.reflection_padding_slices = reflection_padding_slices
        self.verbose = verbose

    # -- geometry ------------------------------------------------------------
    def _get_rotation_z_bounds(self, full_volume_z_size, rot_idx):
        """(start_z, end_z) of rotation rot_idx's window (multirot.py:51-61)."""
        if self.scan_center_slice is not None:
            centre = float(self.scan_center_slice) - 0.5
        else:
            centre = (full_volume_z_size - 1) / 2.0
        first = centre - (self.n_rotations - 1) / 2.0 * self.shift_offset
        start = int(round(first + rot_idx * self.shift_offset - (self.n_z_window - 1) / 2.0))
        return start, start + self.n_z_window

    def _window_index(self, z_size, device):
        """Source slice of every window slot, [n_rotations, n_z_window] int64,
        and a mask of the slots that take that slice (False: zero) and of the
        slots whose copy is detached (repeat padding)."""
        starts = torch.tensor([self._get_rotation_z_bounds(z_size, i)[0] for i in range(self.n_rotations)],
                              dtype=torch.int64, device=device)
        raw = starts[:, None] + torch.arange(self.n_z_window, device=device)[None, :]
        outside = (raw < 0) | (raw >= z_size)
        detach = torch.zeros_like(outside)
        if self.padding_mode == "reflect":
            idx = raw.clone()
            while True:  # reflection about the edge slices' centres
                lo, hi = idx < 0, idx >= z_size
                if not bool((lo | hi).any()):
                    break
                idx = torch.where(lo, -idx, idx)
                idx = torch.where(hi, 2 * z_size - 2 - idx, idx)
                if z_size == 1:
                    idx = torch.zeros_like(idx)
            keep = torch.ones_like(outside)
            if self.reflection_padding_slices is not None:
                lim = int(self.reflection_padding_slices)
                keep = (raw >= -lim) & (raw < z_size + lim)
        elif self.padding_mode == "repeat":
            idx = raw.clamp(0, z_size - 1)
            keep = torch.ones_like(outside)
            detach = outside
        else:
            idx = raw.clamp(0, z_size - 1)
            keep = ~outside
        return idx, keep, detach

    # -- forward ---------------------------------------------------------------
    def _windows(self, volume):
        """[n_rotations, X, Y, n_z_window] windows of one volume [X, Y, Z]."""
        idx, keep, detach = self._window_index(volume.shape[-1], volume.device)
        w = volume.index_select(-1, idx.reshape(-1)).reshape(*volume.shape[:-1], self.n_rotations,
                                                              self.n_z_window)
        w = w.movedim(-2, 0)  # [R, X, Y, W]
        if bool(detach.any()):
            d = detach[:, None, None, :]
            w = torch.where(d, w.detach(), w)
        if not bool(keep.all()):
            w = torch.where(keep[:, None, None, :], w, torch.zeros((), dtype=w.dtype, device=w.device))
        return w.contiguous()

    def forward(self, full_volume):
        """Forward projection of every rotation: [n_rays, n_rotations] (a list
        of them for a list of volumes)."""
        if isinstance(full_volume, (list, tuple)):
            vols = list(full_volume)
            batch = torch.cat([self._windows(v) for v in vols], dim=0)
            sino = self.base_projector(batch)  # [V*R, n_rays]
            return [s.t().contiguous() for s in sino.split(self.n_rotations, dim=0)]
        if full_volume.dim() != 3:
            raise ValueError("MultiRotationProjector.forward takes a [X, Y, Z] volume (or a list of them)")
        sino = self.base_projector(self._windows(full_volume))  # [R, n_rays]
        return sino.t().contiguous()

    def back_project(self, sino, full_volume_z_size):
        """Back projection of [n_rays, n_rotations] into [X, Y, full_volume_z_size]."""
        z = int(full_volume_z_size)
        windows = self.base_projector.back_project(sino.t().contiguous())  # [R, X, Y, W]
        idx, keep, _ = self._window_index(z, windows.device)
        w = windows.movedim(0, -2)  # [X, Y, R, W]
        w = w.reshape(*w.shape[:-2], -1)
        flat_idx, flat_keep = idx.reshape(-1), keep.reshape(-1)
        if not bool(flat_keep.all()):
            w = w[..., flat_keep]
            flat_idx = flat_idx[flat_keep]
        out = torch.zeros((*w.shape[:-1], z), device=w.device, dtype=torch.float32)
        out.index_add_(-1, flat_idx, w.to(torch.float32))
        return out.contiguous()
