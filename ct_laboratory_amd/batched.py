"""Many 2D scans of uneven geometry and size in one persistent launch.

``BatchedProjector2D`` holds the rays of every scan on the device and a
table of ``ctlab_scan2d`` records.  A forward or back projection of the whole
batch is one persistent launch:

- forward (``ctlab_forward_project_2d_batch``): workgroups take work items of
  ``rays_per_item`` consecutive rays from a per-batch queue, one ray per lane;
- back (``ctlab_back_project_2d_batch_tiled``): work units are (``tile_rows``
  x 64 image tile, ``rays_per_part`` consecutive rays); a unit sums into LDS
  and writes its tile once, so no segment costs a global atomic.  ``tiled_back=False``
  selects the one-ray-per-lane atomic launch (``ctlab_back_project_2d_batch``)
  instead.

``autotune`` chooses ``rays_per_item`` and the number of persistent
workgroups per CU for the forward, and the tile size (``tile_rows``) and
``rays_per_part`` for the back projection, by timing the batch's own
projections; results are cached per geometry-class signature of the batch.

Per scan the segments are those of the single-scan on-the-fly kernels
(``_C.forward/back_project_2d_on_the_fly_cuda``), so the forward projections
are bit-identical to them.
"""
from __future__ import annotations

import itertools
import json
import os

import numpy as np
import torch

from . import _native as N
from .tomography.ct_projector_3d_hip import semantics_mode

__all__ = ["Scan2D", "BatchedProjector2D", "SCAN_DTYPE", "c4_workload"]

SCAN_DTYPE = np.dtype([("image", "<u8"), ("sino", "<u8"), ("src", "<u8"), ("dst", "<u8"), ("minv", "<f4", (4,)),
                       ("b", "<f4", (2,)), ("n_row", "<i4"), ("n_col", "<i4"), ("n_ray", "<i8"), ("mode", "<i4"),
                       ("zero_rule", "<i4"), ("item_begin", "<i8")])
assert SCAN_DTYPE.itemsize == 88

N_CU = 256  # MI355X; the projector reads the device's own count
# Tiled back projection: workgroups of 1024 lanes with up to 75 KB of LDS, two per CU.
TILED_GROUPS_PER_CU = 2


class Scan2D:
    """One scan: image shape, affine map (M, b), rays and segment rule.
    ``geometry_class`` labels it for the tuner (e.g. "fan" or "ring")."""

    def __init__(self, n_row, n_col, M, b, src, dst, semantics="torch", geometry_class="fan"):
        self.n_row, self.n_col = int(n_row), int(n_col)
        if self.n_row * self.n_col >= 2 ** 31 or self.n_col >= 2 ** 24 or self.n_row >= 2 ** 24:
            raise ValueError("image too large for 32-bit voxel indices")
        M = M.detach().to("cpu", torch.float32).reshape(2, 2)
        self.minv = torch.inverse(M).contiguous()
        self.b = b.detach().to("cpu", torch.float32).reshape(2).contiguous()
        self.src, self.dst = src.contiguous(), dst.contiguous()
        if self.src.dtype != torch.float32 or self.src.shape != self.dst.shape or self.src.shape[1] != 2:
            raise ValueError("src/dst must be float32 [n_ray, 2] of equal shape")
        self.n_ray = int(self.src.shape[0])
        if semantics not in ("torch", "otf"):
            raise ValueError(f"the multi-scan launches take the 'torch' or 'otf' rule, not {semantics!r}")
        self.mode = semantics_mode(semantics)
        self.zero_rule = N.ZERO_EXACT if semantics == "torch" else N.ZERO_THRESH
        self.geometry_class = geometry_class

    @classmethod
    def from_projector(cls, proj, geometry_class="fan"):
        return cls(proj.n_row, proj.n_col, proj.M, proj.b, proj.src, proj.dst, semantics=proj.semantics,
                   geometry_class=geometry_class)


class BatchedProjector2D:
    _tuned: dict = {}

    def __init__(self, scans, device=None, rays_per_item=1024, groups_per_cu=4, tuner_file=None, rays_per_part=32768,
                 tiled_back=True, tile_rows=128):
        self.scans = list(scans)
        self.device = torch.device("cuda") if device is None else torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:  # tensors always carry the index
            self.device = torch.device("cuda", torch.cuda.current_device())
        for sc in self.scans:
            if sc.src.device != self.device:
                sc.src, sc.dst = sc.src.to(self.device), sc.dst.to(self.device)
        self.n_cu = torch.cuda.get_device_properties(self.device).multi_processor_count or N_CU
        self.counter = torch.zeros(1, dtype=torch.int64, device=self.device)
        self.sinos = [torch.empty(sc.n_ray, device=self.device) for sc in self.scans]
        self.images = [torch.empty(sc.n_row, sc.n_col, device=self.device) for sc in self.scans]
        self.tuner_file = tuner_file
        self.rays_per_item, self.groups_per_cu = int(rays_per_item), int(groups_per_cu)
        self.rays_per_part, self.tiled_back, self.tile_rows = int(rays_per_part), bool(tiled_back), int(tile_rows)
        if self.tile_rows not in (64, 128):
            raise ValueError("tile_rows must be 64 or 128")
        key = self.class_signature()
        cached = self._load_tuned().get(key)
        if cached:
            self.rays_per_item, self.groups_per_cu = cached[:2]
            if len(cached) > 3:
                self.rays_per_part, self.tile_rows = cached[2], cached[3]
        self._tables = {}

    # -- tuner cache ----------------------------------------------------------
    def class_signature(self) -> str:
        counts = {}
        for sc in self.scans:
            counts[sc.geometry_class] = counts.get(sc.geometry_class, 0) + 1
        return ",".join(f"{k}:{counts[k]}" for k in sorted(counts))

    def _load_tuned(self):
        if self.tuner_file and os.path.exists(self.tuner_file):
            with open(self.tuner_file) as f:
                BatchedProjector2D._tuned.update(json.load(f))
        return BatchedProjector2D._tuned

    # -- descriptor table -------------------------------------------------------
    def _build(self, images, sinos, rays_per_item, tiled=False, tile_rows=128):
        """Descriptor table and total work count: items of rays_per_item rays,
        or (tiled) units of tile_rows x 64 tiles x rays_per_item rays."""
        lib = N.load() if tiled else None
        rec = np.zeros(len(self.scans), SCAN_DTYPE)
        begin = 0
        for k, sc in enumerate(self.scans):
            rec[k]["image"] = images[k].data_ptr()
            rec[k]["sino"] = sinos[k].data_ptr()
            rec[k]["src"] = sc.src.data_ptr()
            rec[k]["dst"] = sc.dst.data_ptr()
            rec[k]["minv"] = sc.minv.reshape(-1).numpy()
            rec[k]["b"] = sc.b.numpy()
            rec[k]["n_row"], rec[k]["n_col"], rec[k]["n_ray"] = sc.n_row, sc.n_col, sc.n_ray
            rec[k]["mode"], rec[k]["zero_rule"] = sc.mode, sc.zero_rule
            rec[k]["item_begin"] = begin
            if tiled:
                begin += int(lib.ctlab_back_project_2d_tile_units(sc.n_row, sc.n_col, sc.n_ray, rays_per_item,
                                                                  tile_rows))
            else:
                begin += -(-sc.n_ray // rays_per_item)
        table = torch.from_numpy(rec.view(np.uint8).copy()).to(self.device)
        return table, begin

    def _launch(self, back, images, sinos, rays_per_item, groups_per_cu, accumulate=False, tiled=False,
                tile_rows=128):
        """One persistent launch; for a tiled back projection rays_per_item is
        the rays per part."""
        for t in list(images) + list(sinos):
            if t.dtype != torch.float32 or not t.is_contiguous() or t.device != self.device:
                raise RuntimeError("batched projector: images and sinograms must be contiguous float32 on the device")
        tiled = bool(back and tiled)
        key = (tuple(t.data_ptr() for t in images), tuple(t.data_ptr() for t in sinos), rays_per_item, tiled,
               tile_rows)
        if key not in self._tables:
            table, items = self._build(images, sinos, rays_per_item, tiled, tile_rows)
            # the pointed-to tensors stay referenced as long as the table
            self._tables[key] = (table, items, list(images), list(sinos))
            while len(self._tables) > 8:
                self._tables.pop(next(iter(self._tables)))
        table, items = self._tables[key][:2]
        stream = N.stream_of(self.counter)
        with torch.cuda.device(self.device):
            if tiled:
                N.call("ctlab_back_project_2d_batch_tiled", N.ptr(table), len(self.scans), items, rays_per_item,
                       tile_rows, TILED_GROUPS_PER_CU * self.n_cu, N.ptr(self.counter), int(bool(accumulate)),
                       stream)
            elif back:
                N.call("ctlab_back_project_2d_batch", N.ptr(table), len(self.scans), items, rays_per_item,
                       groups_per_cu * self.n_cu, N.ptr(self.counter), int(bool(accumulate)), stream)
            else:
                N.call("ctlab_forward_project_2d_batch", N.ptr(table), len(self.scans), items,
                       rays_per_item, groups_per_cu * self.n_cu, N.ptr(self.counter), stream)

    # -- projections ------------------------------------------------------------
    def forward(self, images):
        """Line integrals of every scan: list of [n_ray] tensors (resident buffers)."""
        for sc, im in zip(self.scans, images):
            if tuple(im.shape) != (sc.n_row, sc.n_col):
                raise RuntimeError("batched projector: image shape does not match its scan")
        self._launch(False, images, self.sinos, self.rays_per_item, self.groups_per_cu)
        return self.sinos

    def back(self, sinos):
        """Back projections of every scan: list of [n_row, n_col] tensors (resident buffers)."""
        for sc, s in zip(self.scans, sinos):
            if tuple(s.shape) != (sc.n_ray,):
                raise RuntimeError("batched projector: sinogram shape does not match its scan")
        if self.tiled_back:
            self._launch(True, self.images, sinos, self.rays_per_part, 0, tiled=True, tile_rows=self.tile_rows)
        else:
            self._launch(True, self.images, sinos, self.rays_per_item, self.groups_per_cu)
        return self.images

    def _time(self, fn, reps):
        fn()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            fn()
        ev[1].record()
        ev[1].synchronize()
        return ev[0].elapsed_time(ev[1]) / reps

    def autotune(self, images, rays_per_item=(256, 512, 1024, 2048, 4096), groups_per_cu=(1, 2, 4, 8, 16),
                 rays_per_part=(8192, 16384, 32768, 65536, 131072), tile_rows=(64, 128), reps=3):
        """Times this batch's forward for every (rays_per_item, groups_per_cu),
        and its back projection for every (tile_rows, rays_per_part) (tiled) or
        every (rays_per_item, groups_per_cu) (atomic), and keeps the fastest.
        Returns ((rays_per_item, groups_per_cu, rays_per_part, tile_rows),
        {"forward": {config: ms}, "back": {config: ms}})."""
        sinos = [torch.empty(sc.n_ray, device=self.device) for sc in self.scans]
        fwd = {(rpi, gpc): self._time(lambda: self._launch(False, images, sinos, rpi, gpc), reps)
               for rpi, gpc in itertools.product(rays_per_item, groups_per_cu)}
        if self.tiled_back:
            back = {(tr, rpp): self._time(lambda: self._launch(True, self.images, sinos, rpp, 0, tiled=True,
                                                               tile_rows=tr), reps)
                    for tr, rpp in itertools.product(tile_rows, rays_per_part)}
            best_fwd = min(fwd, key=fwd.get)
            best_tr, best_rpp = min(back, key=back.get)
        else:
            back = {(rpi, gpc): self._time(lambda: self._launch(True, self.images, sinos, rpi, gpc), reps)
                    for rpi, gpc in itertools.product(rays_per_item, groups_per_cu)}
            # one (rays_per_item, groups_per_cu) serves both launches
            best_fwd = min(fwd, key=lambda k: fwd[k] + back[k])
            best_tr, best_rpp = self.tile_rows, self.rays_per_part
        best = (best_fwd[0], best_fwd[1], best_rpp, best_tr)
        self.rays_per_item, self.groups_per_cu, self.rays_per_part, self.tile_rows = best
        BatchedProjector2D._tuned[self.class_signature()] = list(best)
        if self.tuner_file:
            with open(self.tuner_file, "w") as f:
                json.dump(BatchedProjector2D._tuned, f)
        return best, {"forward": fwd, "back": back}


def c4_workload(n_scan=64, device=None):
    """64 scans: scan i draws from default_rng(100 + i) an image size in
    {128, 192, 256, 384, 512}^2 at 256/n mm per pixel, 90-720 views, and a
    geometry: fan beam (512-1024 bins at 1 mm, SID 1000 / SDD 1500) or a StaticCT
    ring (1024 detectors as 64 modules x 16 columns on a 500 mm circle, 8-64
    sources on a 518 mm circle, frame f lit by source f mod n_source); plus a
    seeded ellipse phantom per scan.  Returns (scans, images)."""
    from .phantoms import ellipse_phantom_2d
    from .tomography import (build_fanbeam_2d_geometry, build_static_2d_geometry, build_uniform_static_2d_geometry,
                             standard_image_transform_2d)

    device = torch.device("cuda") if device is None else torch.device(device)
    scans, images = [], []
    for i in range(n_scan):
        rng = np.random.default_rng(100 + i)
        n = int(rng.choice([128, 192, 256, 384, 512]))
        px = 256.0 / n
        n_view = int(rng.integers(90, 721))
        M, b = standard_image_transform_2d(n, n, px, px)
        if rng.uniform() < 0.5:
            bins = int(rng.integers(512, 1025))
            src, dst = build_fanbeam_2d_geometry(n_view, bins, 1000.0, 1500.0, det_spacing=1.0)
            cls = "fan"
        else:
            n_source = int(rng.integers(8, 65))
            Mg = torch.eye(2).repeat(n_view, 1, 1)
            bg = torch.zeros(n_view, 2)
            act = torch.zeros(n_view, n_source, dtype=torch.bool)
            act[torch.arange(n_view), torch.arange(n_view) % n_source] = True
            sp, mc, mo, mask = build_uniform_static_2d_geometry(n_source, 518.0, 64, 500.0, 16, 500.0 * 2 * np.pi
                                                                / 1024, Mg, bg, act)
            src, dst = build_static_2d_geometry(sp, mc, mo, 16, 500.0 * 2 * np.pi / 1024, mask, Mg, bg, act)
            cls = "ring"
        scans.append(Scan2D(n, n, M, b, src.to(device), dst.to(device), geometry_class=cls))
        images.append(torch.from_numpy(ellipse_phantom_2d(n, n, px, seed=100 + i)).to(device))
    return scans, images
