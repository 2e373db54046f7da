"""Ray-subset projectors (ct_laboratory/tomography/ct_projector_3d_subsets.py:9-276).

``OrderedSubsetProjector`` splits the core module's rays into contiguous
ranges of at most ``max_subset_size`` rays and projects range by range;
``RandomSubsetProjector`` projects through a random subset of the rays.  The
names, arguments, attributes (``n_subsets``, ``subset_boundaries``,
``last_precompute_time``, ``last_projection_time``) and the per-subset cache
file names are the reference's.

What differs, MI355X-first:

- A subset's crossing table is a slice (a view, no copy) of the core module's
  table when it has one; otherwise it is computed on the GPU by the
  traversal (``precompute_tvals_stitched``), optionally loaded from / saved to
  ``cache_dir`` with the reference's file names (loaded with
  ``weights_only=True``).
- Subset modules stay resident once built (``resident=True``, the default):
  288 GB of HBM holds the tables of every subset of the configurations here,
  so nothing is rebuilt per call.  ``resident=False`` rebuilds per call and
  frees as the reference does.
- Forward results are written straight into slices of one output tensor and
  back projections accumulate in place.
"""
from __future__ import annotations

import os
import time

import numpy as np
import torch
from torch import nn

from .ct_projector_3d_module import CTProjector3DModule, precompute_tvals_stitched

__all__ = ["SubsetProjector", "OrderedSubsetProjector", "RandomSubsetProjector"]


def _core_module(projector) -> CTProjector3DModule:
    core = projector
    while not isinstance(core, CTProjector3DModule) and hasattr(core, "base_projector"):
        core = core.base_projector
    if not isinstance(core, CTProjector3DModule):
        raise ValueError("base_projector must eventually be a CTProjector3DModule")
    return core


class SubsetProjector(nn.Module):
    """Base of the subset projectors: resolves the innermost
    ``CTProjector3DModule`` and builds modules over index sets of its rays."""

    def __init__(self, base_projector, subset_size, precomputed_intersections=False, verbose=False):
        super().__init__()
        self.base_projector = base_projector
        self.subset_size = int(subset_size)
        self.precomputed_intersections = precomputed_intersections
        self.verbose = verbose
        core = _core_module(base_projector)
        self.core_module = core
        self.n_total_rays_core = int(core.src.shape[0])
        self.device = core.src.device if core.src.is_cuda else torch.device("cuda")
        self.n_x, self.n_y, self.n_z = core.n_x, core.n_y, core.n_z
        self.M, self.b = core.M, core.b
        self.backend = core.backend

    def _get_subset_module(self, indices):
        """On-the-fly module over the core rays ``indices``."""
        core = self.core_module
        idx = torch.as_tensor(indices, device=core.src.device)
        return CTProjector3DModule(self.n_x, self.n_y, self.n_z, self.M.to(self.device), self.b.to(self.device),
                                   core.src[idx].to(self.device), core.dst[idx].to(self.device),
                                   backend=self.backend, device=self.device,
                                   precomputed_intersections=self.precomputed_intersections,
                                   semantics=core.semantics)

    def _run_forward(self, module, volume):
        return module.forward(volume)

    def _run_backward(self, module, sinogram):
        return module.back_project(sinogram)


class OrderedSubsetProjector(SubsetProjector):
    """Projects the core rays range by range (at most ``max_subset_size`` rays
    each), with per-subset crossing tables as in the reference."""

    def __init__(self, base_projector, max_subset_size=1_000_000, cache_dir=None, verbose=False, *,
                 resident: bool = True):
        super().__init__(base_projector, max_subset_size, precomputed_intersections=True, verbose=verbose)
        self.cache_dir = cache_dir
        self.resident = resident
        self.n_subsets = max(1, int(np.ceil(self.n_total_rays_core / self.subset_size)))
        self.subset_boundaries = np.linspace(0, self.n_total_rays_core, self.n_subsets + 1, dtype=int)
        self.last_precompute_time = 0.0
        self.last_projection_time = 0.0
        self._subset_cache: dict[int, CTProjector3DModule] = {}
        self.persistent_projector = None
        if self.n_subsets == 1:
            t0 = time.time()
            self.persistent_projector = self._get_cached_subset_projector(0)
            self.last_precompute_time = time.time() - t0

    def _range(self, j):
        return int(self.subset_boundaries[j]), int(self.subset_boundaries[j + 1])

    def _get_subset_cache_path(self, j):
        """The reference's per-subset cache file name (subsets.py:84-94), or None."""
        if self.cache_dir is None:
            return None
        a, b = self._range(j)
        return os.path.join(self.cache_dir, f"tvals_subset{j:04d}_rays{a}-{b}_nx{self.n_x}_ny{self.n_y}_nz{self.n_z}.pt")

    def _subset_tvals(self, j):
        """Slice of the core tables, else the cache file, else a GPU precompute
        (saved to the cache when a cache_dir is set)."""
        a, b = self._range(j)
        core = self.core_module
        if core.tvals_uint16 is not None:
            return core.tvals_uint16[a:b], core.tvals_start[a:b], core.tvals_scale[a:b]
        if core.tvals_full is not None:
            return core.tvals_full[a:b]
        path = self._get_subset_cache_path(j)
        if path is not None and os.path.exists(path):
            if self.verbose:
                print(f"  Loading subset {j} tvals: {path}", flush=True)
            t = torch.load(path, map_location="cpu", weights_only=True)
            return tuple(t) if isinstance(t, (tuple, list)) else t
        if self.verbose:
            print(f"  Computing subset {j} tvals ({b - a} rays)...", flush=True)
        t = precompute_tvals_stitched(self.n_x, self.n_y, self.n_z, self.M, self.b, core.src[a:b], core.dst[a:b],
                                      verbose=self.verbose)
        if path is not None:
            os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
            torch.save(tuple(x.cpu() for x in t) if isinstance(t, tuple) else t.cpu(), path)
        return t

    def _get_cached_subset_projector(self, j):
        if j in self._subset_cache:
            return self._subset_cache[j]
        a, b = self._range(j)
        core = self.core_module
        m = CTProjector3DModule(self.n_x, self.n_y, self.n_z, self.M, self.b, core.src[a:b], core.dst[a:b],
                                backend=self.backend, device=self.device, tvals=self._subset_tvals(j),
                                semantics=core.semantics)
        if self.resident:
            self._subset_cache[j] = m
        return m

    def forward(self, volume):
        self.last_precompute_time = self.last_projection_time = 0.0
        if self.persistent_projector is not None:
            t0 = time.time()
            out = self._run_forward(self.persistent_projector, volume)
            self.last_projection_time = time.time() - t0
            return out
        shape = (volume.shape[0], self.n_total_rays_core) if volume.dim() == 4 else (self.n_total_rays_core,)
        out = torch.zeros(shape, device=volume.device, dtype=volume.dtype)
        for j in range(self.n_subsets):
            a, b = self._range(j)
            t0 = time.time()
            m = self._get_cached_subset_projector(j)
            t1 = time.time()
            out[..., a:b] = self._run_forward(m, volume)
            t2 = time.time()
            self.last_precompute_time += t1 - t0
            self.last_projection_time += t2 - t1
            if self.verbose:
                print(f"  [Subset {j + 1}/{self.n_subsets}] rays {a}:{b} {t2 - t0:.2f}s", flush=True)
        return out

    def back_project(self, sinogram):
        self.last_precompute_time = self.last_projection_time = 0.0
        if self.persistent_projector is not None:
            t0 = time.time()
            out = self._run_backward(self.persistent_projector, sinogram)
            self.last_projection_time = time.time() - t0
            return out
        out = None
        for j in range(self.n_subsets):
            a, b = self._range(j)
            t0 = time.time()
            m = self._get_cached_subset_projector(j)
            t1 = time.time()
            part = self._run_backward(m, sinogram[..., a:b].contiguous())
            out = part if out is None else out.add_(part)
            t2 = time.time()
            self.last_precompute_time += t1 - t0
            self.last_projection_time += t2 - t1
            if self.verbose:
                print(f"  [Subset {j + 1}/{self.n_subsets}] rays {a}:{b} (backward) {t2 - t0:.2f}s", flush=True)
        return out


class RandomSubsetProjector(SubsetProjector):
    """Projects through ``subset_size`` rays drawn at random (torch.randperm
    on the CPU, seeded by ``seed`` when given), optionally redrawn on every
    forward or back projection (subsets.py:225-276)."""

    def __init__(self, base_projector, subset_size=10000, precomputed_intersections=False,
                 randomize_on_forward=False, randomize_on_backward=False, seed=None, verbose=False):
        super().__init__(base_projector, subset_size, precomputed_intersections, verbose)
        self.randomize_on_forward = randomize_on_forward
        self.randomize_on_backward = randomize_on_backward
        self.seed = seed
        self.current_subset_module = None
        self.indices = None
        self.randomize_projector()

    def randomize_projector(self, subset_size=None, seed=None):
        if subset_size is not None:
            self.subset_size = int(subset_size)
        if seed is not None:
            self.seed = seed
        if self.seed is not None:
            torch.manual_seed(self.seed)
        self.indices = torch.randperm(self.n_total_rays_core, device="cpu")[:self.subset_size]
        self.current_subset_module = self._get_subset_module(self.indices)

    def forward(self, volume):
        if self.randomize_on_forward:
            self.randomize_projector()
        return self._run_forward(self.current_subset_module, volume)

    def back_project(self, sinogram):
        if self.randomize_on_backward:
            self.randomize_projector()
        return self._run_backward(self.current_subset_module, sinogram)

    def forward_project(self, volume):
        return self.forward(volume)
