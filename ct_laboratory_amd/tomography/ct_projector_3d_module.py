"""3D projector nn.Module (ct_laboratory/tomography/ct_projector_3d_module.py:28-253) and the
chunked crossing-table precompute (precompute_tvals_stitched, :256-315).

Crossing tables, when used, are computed on the GPU (the reference computes
them on the CPU with torch, ct_projector_3d_module.py:84-109), trimmed to the
largest crossing count and compressed with the reference module's uint16
rule (compress_tvals_to_uint16 semantics).  Without tables the on-the-fly
traversal runs; when the rays are ordered [view][rows][cols] with one source
per view (``det_shape`` given, as the cone-beam and StaticCT builders do) the
forward uses the detector-tiled schedule and the backward the LDS-brick back
projector, whose geometry-only plan is built once and cached.
"""
from __future__ import annotations

import math

import torch

from .. import _C
from ._backend import check_backend
from .ct_projector_3d_autograd import CTProjector3DFunction, OnTheFlyOptions
from .ct_projector_3d_hip import (ConeBeamPlan, _collapse, _prepare_M_inv, back_project_3d_hip,
                                  compress_tvals_3d_hip, compute_intersections_3d_hip)

__all__ = ["CTProjector3DModule", "precompute_tvals_stitched", "decompress_tvals_from_uint16"]


def decompress_tvals_from_uint16(t_uint16, t_start, t_scale):
    """t = cumsum(uint16 deltas) * scale + start, in float32
    (ct_projector_3d_torch.py:68-75)."""
    deltas = t_uint16.to(torch.int32) & 0xFFFF
    k = torch.cumsum(deltas.to(torch.float32), dim=1)
    return k * t_scale.unsqueeze(1) + t_start.unsqueeze(1)


def _table_on_device(n_x, n_y, n_z, M, b, src, dst, compress: bool, chunk_bytes: int = 1 << 30):
    """Crossing table of all rays on the rays' device, trimmed to the largest
    crossing count; compressed chunk by chunk when ``compress``."""
    Mc, bc = _collapse(M, b)
    M_inv = _prepare_M_inv(Mc)
    bc = bc.to(torch.float32).contiguous()
    counts = _C.count_crossings((n_x, n_y, n_z), src, dst, M_inv, bc)
    K = max(1, int(counts.max().item()) if counts.numel() else 1)
    n_ray = int(src.shape[0])
    if not compress:
        return compute_intersections_3d_hip(n_x, n_y, n_z, Mc, bc, src, dst, K=K)
    t16 = torch.empty((n_ray, K), dtype=torch.int16, device=src.device)
    ts = torch.empty((n_ray,), dtype=torch.float32, device=src.device)
    tsc = torch.empty((n_ray,), dtype=torch.float32, device=src.device)
    step = max(1, chunk_bytes // (4 * K))
    for r0 in range(0, n_ray, step):
        r1 = min(n_ray, r0 + step)
        s, d = src[r0:r1].contiguous(), dst[r0:r1].contiguous()
        t = compute_intersections_3d_hip(n_x, n_y, n_z, Mc, bc, s, d, K=K)
        a, st, sc = compress_tvals_3d_hip(t, Mc, s, d, semantics="torch")
        t16[r0:r1], ts[r0:r1], tsc[r0:r1] = a, st, sc
        del t
    return t16, ts, tsc


class CTProjector3DModule(torch.nn.Module):
    """CT projector with crossing tables (float or uint16) or on-the-fly traversal.

    Arguments as in the reference (n_x, n_y, n_z, M, b, src, dst, backend,
    device, precomputed_intersections, tvals, use_compression), plus
    ``semantics`` ("torch" or "otf": the segment rule of the on-the-fly
    traversal) and ``det_shape`` (rows, cols) of rays ordered [view][rows][cols]."""

    def __init__(self, n_x, n_y, n_z, M, b, src, dst, backend="cuda", device=None,
                 precomputed_intersections=False, tvals=None, use_compression=True, *, semantics: str = "torch",
                 det_shape=None):
        super().__init__()
        self.backend = check_backend(backend)
        if device is None:
            device = src.device if src.is_cuda else torch.device("cuda")
        self.n_x, self.n_y, self.n_z = int(n_x), int(n_y), int(n_z)
        self.use_compression = use_compression
        self.semantics = semantics
        self.det_shape = None if det_shape is None else (int(det_shape[0]), int(det_shape[1]))
        self.register_buffer("M", M.to(device))
        self.register_buffer("b", b.to(device))
        self.register_buffer("src", src.to(device).contiguous())
        self.register_buffer("dst", dst.to(device).contiguous())
        self.tvals_uint16 = self.tvals_start = self.tvals_scale = self.tvals_full = None
        self._plan = None
        self._runs = None
        self.precomputed_intersections = False
        if tvals is not None:
            self.set_tvals(tvals, device=device)
        elif precomputed_intersections:
            tab = _table_on_device(self.n_x, self.n_y, self.n_z, self.M, self.b, self.src, self.dst,
                                   compress=use_compression)
            if use_compression:
                self.tvals_uint16, self.tvals_start, self.tvals_scale = tab
            else:
                self.tvals_full = tab
            self.precomputed_intersections = True
            self._build_table_runs()

    # -- device moves -------------------------------------------------------
    def to(self, *args, **kwargs):
        """Moves buffers and crossing tables (ct_projector_3d_module.py:119-140)."""
        device = torch._C._nn._parse_to(*args, **kwargs)[0]
        super().to(*args, **kwargs)
        if device is not None:
            for name in ("tvals_uint16", "tvals_start", "tvals_scale", "tvals_full"):
                t = getattr(self, name)
                if t is not None:
                    setattr(self, name, t.to(device))
            self._plan = None
            self._build_table_runs()
        return self

    def _build_table_runs(self):
        """Run lists of the crossing table for the brick back projection, built
        once per table (a geometry-only pass over it) when the table is large
        enough for them to pay; None otherwise."""
        self._runs = None
        tv = self._table()
        if tv is None or not self.src.is_cuda:
            return
        t0 = tv[0] if isinstance(tv, tuple) else tv
        if t0.numel() < _C._MIN_TABLE_ENTRIES:
            return
        M, b = _collapse(self.M, self.b)
        free, _ = torch.cuda.mem_get_info(self.src.device)
        # None when the lists would take more than half of the free device memory
        self._runs = _C.table_runs(tv, self.src, self.dst, _prepare_M_inv(M), b.to(torch.float32).contiguous(),
                                   (self.n_x, self.n_y, self.n_z), max_bytes=free // 2)

    # -- on-the-fly configuration -------------------------------------------
    def _options(self) -> OnTheFlyOptions:
        plan = None
        if self.det_shape is not None:
            if self._plan is None or self._plan_device != self.src.device:
                self._plan = ConeBeamPlan(self.src, self.dst, self.M, self.b, (self.n_x, self.n_y, self.n_z),
                                          self.det_shape)
                self._plan_device = self.src.device
            plan = self._plan.plan if self._plan.usable else None
        return OnTheFlyOptions(self.semantics, self.det_shape, plan)

    def _table(self):
        if self.tvals_uint16 is not None:
            return (self.tvals_uint16, self.tvals_start, self.tvals_scale)
        return self.tvals_full

    # -- operators ----------------------------------------------------------
    def forward(self, volume):
        """volume [X,Y,Z] or [B,X,Y,Z] -> sinogram [n_ray] or [B,n_ray] (differentiable)."""
        t16, ts, tsc, tf = self.tvals_uint16, self.tvals_start, self.tvals_scale, self.tvals_full
        opts = self._options() if (t16 is None and tf is None) else None
        return CTProjector3DFunction.apply(volume, t16 if t16 is not None else tf, ts, tsc, self.M, self.b,
                                           self.src, self.dst, self.backend, opts)

    def forward_project(self, volume):
        return self.forward(volume)

    def back_project(self, sinogram):
        tv = self._table()
        opts = self._options() if tv is None else OnTheFlyOptions(self.semantics)
        runs = None
        if tv is not None and self._runs is not None:
            M, b = _collapse(self.M, self.b)
            if self._runs.matches(tv, self.src, self.dst, _prepare_M_inv(M), b):
                runs = self._runs
        return back_project_3d_hip(sinogram.contiguous(), tv, self.src, self.dst, self.M, self.b, self.n_x,
                                   self.n_y, self.n_z, semantics=opts.semantics, det_shape=opts.det_shape,
                                   brick_plan=opts.plan, table_runs=runs)

    # -- crossing tables ----------------------------------------------------
    def save_tvals(self, filePath):
        """Saves the tables as the reference's dict (ct_projector_3d_module.py:190-204)."""
        if self.tvals_uint16 is None and self.tvals_full is None:
            raise ValueError("No precomputed t-values to save.")

        def cpu(t):
            return None if t is None else t.cpu()

        torch.save({"tvals_uint16": cpu(self.tvals_uint16), "tvals_start": cpu(self.tvals_start),
                    "tvals_scale": cpu(self.tvals_scale), "tvals_full": cpu(self.tvals_full),
                    "n_x": self.n_x, "n_y": self.n_y, "n_z": self.n_z, "n_ray": int(self.src.shape[0])}, filePath)

    def load_tvals(self, filePath, device=None):
        """Loads a dict written by save_tvals (ct_projector_3d_module.py:206-216), with a
        loader that executes nothing from the file."""
        device = self.src.device if device is None else device
        d = torch.load(filePath, map_location="cpu", weights_only=True)
        if d["n_x"] != self.n_x or d["n_y"] != self.n_y or d["n_z"] != self.n_z:
            raise ValueError("Dimension mismatch")
        for name in ("tvals_uint16", "tvals_start", "tvals_scale", "tvals_full"):
            t = d.get(name)
            setattr(self, name, None if t is None else t.to(device))
        self.precomputed_intersections = True
        self._build_table_runs()

    def set_tvals(self, tvals, device=None):
        """Float table, (uint16, start, scale) tuple, or None for on-the-fly
        (ct_projector_3d_module.py:218-243)."""
        device = self.src.device if device is None else device
        if tvals is None:
            self.tvals_uint16 = self.tvals_start = self.tvals_scale = self.tvals_full = None
            self.precomputed_intersections = False
            self._runs = None
            return
        if isinstance(tvals, (tuple, list)):
            tu16, tstart, tscale = tvals
            self.tvals_uint16, self.tvals_start, self.tvals_scale = (tu16.to(device).contiguous(),
                                                                     tstart.to(device).contiguous(),
                                                                     tscale.to(device).contiguous())
            self.tvals_full = None
        elif self.use_compression:
            tf = tvals.to(device).to(torch.float32).contiguous()
            self.tvals_uint16, self.tvals_start, self.tvals_scale = compress_tvals_3d_hip(tf, self.M, self.src,
                                                                                          self.dst,
                                                                                          semantics="torch")
            self.tvals_full = None
        else:
            self.tvals_full = tvals.to(device).to(torch.float32).contiguous()
            self.tvals_uint16 = self.tvals_start = self.tvals_scale = None
        self.precomputed_intersections = True
        self._build_table_runs()

    @property
    def tvals(self):
        if self.tvals_uint16 is not None:
            return decompress_tvals_from_uint16(self.tvals_uint16, self.tvals_start, self.tvals_scale)
        return self.tvals_full

    @tvals.setter
    def tvals(self, value):
        self.set_tvals(value)


def precompute_tvals_stitched(n_x, n_y, n_z, M, b, src_all, dst_all, chunk_size=1000000, backend="cuda",
                              device=None, verbose=True, use_compression=True):
    """Crossing tables of many rays in chunks on the GPU, stitched into host
    tensors of width (n_x+1)+(n_y+1)+(n_z+1) (ct_projector_3d_module.py:256-315):
    (int16 [n_ray, W], start [n_ray], scale [n_ray]) with the reference GPU
    compression rule, or float32 [n_ray, W]."""
    check_backend(backend)
    device = torch.device("cuda") if device is None else torch.device(device)
    n_ray = int(src_all.shape[0])
    W = (n_x + 1) + (n_y + 1) + (n_z + 1)
    if verbose:
        print(f"Precomputing {n_ray} rays in chunks of {chunk_size} on {device}...")
    if use_compression:
        main = torch.zeros((n_ray, W), dtype=torch.int16)
        starts = torch.zeros(n_ray, dtype=torch.float32)
        scales = torch.zeros(n_ray, dtype=torch.float32)
    else:
        full = torch.zeros((n_ray, W), dtype=torch.float32)
    Md, bd = M.to(device), b.to(device)
    for i in range(math.ceil(n_ray / chunk_size)):
        r0, r1 = i * chunk_size, min((i + 1) * chunk_size, n_ray)
        s = src_all[r0:r1].to(device).to(torch.float32).contiguous()
        d = dst_all[r0:r1].to(device).to(torch.float32).contiguous()
        t = compute_intersections_3d_hip(n_x, n_y, n_z, Md, bd, s, d)
        if use_compression:
            t16, ts, tsc = compress_tvals_3d_hip(t, Md, s, d, semantics="cuda")
            main[r0:r1] = t16.cpu()
            starts[r0:r1] = ts.cpu()
            scales[r0:r1] = tsc.cpu()
        else:
            full[r0:r1] = t.cpu()
        del t, s, d
    return (main, starts, scales) if use_compression else full
