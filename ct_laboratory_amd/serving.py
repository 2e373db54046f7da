"""Continuous-batching server for MAP reconstruction requests.

Requests arrive over time.  The server admits them while it has a free slot and
its HBM budget allows the request's resident bytes, which are:

- the geometry (src/dst);
- the measurements;
- two iterates, the projection and residual buffers;
- the back projection and the preconditioner;
- the brick plan or crossing table.

Each admitted request gets a slot, that is, a HIP stream at default priority. On
that stream it builds its projector, computes its SQS preconditioner, captures
two MAP iterations in a HIP graph, and queues all its replays. Requests on
different slots overlap on the device. The host only polls completion events.
A finished request frees its slot and budget, and the next queued request is
admitted at once. Every iterate stays resident until the request completes.

``plain_loop`` is the comparison: one request at a time, the reference-style
autograd loop with one host synchronisation per iteration, run over the same
projector, preconditioner and iterations, so the final images agree.  With
``reference_kernels=True`` it runs the same loop on the reference's kernel
structure (table kernels and storage-order on-the-fly kernels with per-segment
global atomics), a second baseline.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .map_estimator import (DiagonalPreconditioner, LinearGaussianRandomVariable, MaximumAPosterioriEstimator,
                            ProjectorOps, QuadraticMAP, QuadraticSmoothnessRandomVariable, sqs_preconditioner)

__all__ = ["MAPRequest", "ReconServer", "plain_loop", "poisson_workload", "build_problem", "latency_stats"]


@dataclass
class MAPRequest:
    """One reconstruction: ``kind`` "fan2d" (n x n image, n_view views, ``det``
    detector bins) or "cone3d" (n^3 volume, n_view views, det x det panel)."""

    rid: int
    kind: str
    n: int
    n_view: int
    det: int
    iters: int = 20
    beta: float = 1e-2
    seed: int = 0
    arrival: float = 0.0
    y: torch.Tensor | None = field(default=None, repr=False)
    var: float = 1.0
    x_true: torch.Tensor | None = field(default=None, repr=False)


def _projector(req: MAPRequest, device):
    from .tomography import ConeBeam3DProjector, FanBeam2DProjector

    if req.kind == "fan2d":
        # 1 mm pixels, 1 mm bins, SID 1000 / SDD 1500 (the fan-beam configuration).  On the
        # fly: no crossing table to hold (4 (2n + 2) bytes per ray), and the back
        # projection sums LDS tiles instead of streaming the table into global atomics.
        return FanBeam2DProjector(req.n, req.n, req.n_view, req.det, 1000.0, 1500.0, det_spacing=1.0,
                                  pixel_size=1.0, device=device, precomputed_intersections=False)
    # 1 mm voxels, 1.5 mm pixels, SID 1000 / SDD 1500 (the cone-beam configuration)
    return ConeBeam3DProjector(req.n, req.n, req.n, req.n_view, req.det, req.det, 1000.0, 1500.0, det_spacing=1.5,
                               voxel_size=1.0, device=device)


def _reference_projector(req: MAPRequest, device):
    """The request's projector with the reference's kernel structure, for the
    baseline loop: a fan-beam module holding its crossing table, consumed by the
    table kernels (the back projection adds every segment with a global atomic,
    2d.cu:117-155), and a cone-beam module without a detector layout, traced one
    ray per lane in storage order with per-segment global atomics (the
    structure of 3d.cu:821-1079).  The segments are those of ``_projector``'s
    (torch rule).  Use with ``_C.AUTO_LAYOUT`` off, so that the shim does not
    swap in its own schedules."""
    from .tomography import CTProjector3DModule, FanBeam2DProjector
    from .tomography.conebeam_projector_3d import build_conebeam_3d_geometry
    from .tomography.standard_image_transform import standard_image_transform_3d

    if req.kind == "fan2d":
        proj = FanBeam2DProjector(req.n, req.n, req.n_view, req.det, 1000.0, 1500.0, det_spacing=1.0,
                                  pixel_size=1.0, device=device, precomputed_intersections=True)
        # a copy of the table (as loaded from disk) is consumed as given, by the table kernels
        proj.tvals = proj.tvals.clone()
        return proj
    src, dst = build_conebeam_3d_geometry(req.n_view, req.det, req.det, 1000.0, 1500.0, det_spacing=1.5,
                                          device=device)
    M, b = standard_image_transform_3d(n_x=req.n, n_y=req.n, n_z=req.n, s_x=1.0, s_y=1.0, s_z=1.0)
    return CTProjector3DModule(req.n, req.n, req.n, M.to(device), b.to(device), src, dst, device=device)


def build_problem(req: MAPRequest, device) -> None:
    """Synthetic measurements: seeded phantom, forward projection, Gaussian noise of
    standard deviation 0.01 * max line integral (fills req.y, req.var, req.x_true)."""
    from .phantoms import ellipse_phantom_2d, ellipsoid_phantom_3d

    proj = _projector(req, device)
    if req.kind == "fan2d":
        x = torch.from_numpy(ellipse_phantom_2d(req.n, req.n, 1.0, seed=req.seed)).to(device)
    else:
        x = ellipsoid_phantom_3d(req.n, req.n, req.n, 1.0, seed=req.seed, device=device)
    y = proj.forward_project(x)
    sigma = 0.01 * float(y.max())
    g = torch.Generator().manual_seed(10_000 + req.seed)
    req.y = (y + sigma * torch.randn(y.shape, generator=g).to(device)).contiguous()
    req.var = sigma * sigma
    req.x_true = x


def resident_bytes(req: MAPRequest) -> int:
    """Device bytes a request holds while active: rays (src + dst), measurements,
    projection and residual buffers, two iterates, back projection, preconditioner
    and (3D) the brick plan."""
    if req.kind == "fan2d":
        n_ray, vox = req.n_view * req.det, req.n * req.n
        geometry = 16 * n_ray
    else:
        n_ray, vox = req.n_view * req.det * req.det, req.n ** 3
        bricks = math.ceil(req.n / 32) ** 2 * math.ceil(req.n / 8)  # 32 x 32 x 8 bricks
        geometry = 24 * n_ray + 4 * (bricks * (4 * req.n_view + 1) + req.n_view) + 64 * req.n_view
    return int(geometry + 3 * 4 * n_ray + 5 * 4 * vox)


class _Active:
    def __init__(self, req, stream, device):
        self.req = req
        self.stream = stream
        # the request's measurements were written on the caller's stream
        stream.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(stream):
            self.proj = _projector(req, device)
            self.map = QuadraticMAP(self.proj, req.y, req.var, req.beta, torch.zeros_like(req.x_true))
            self.map.capture()
            self.map.run(req.iters)
            self.done = torch.cuda.Event()
            self.done.record(stream)


class ReconServer:
    """Admits requests by HBM budget and slot count, overlaps them on one GPU."""

    def __init__(self, device=None, hbm_budget=200e9, slots=8):
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self.hbm_budget = float(hbm_budget)
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(slots)]

    def run(self, requests):
        """Serves ``requests`` (arrival times in seconds from the start); returns
        {rid: (estimate, latency_s)} and the makespan in seconds."""
        pending = sorted(requests, key=lambda r: r.arrival)
        free = list(range(len(self.streams)))
        active = {}
        used = 0.0
        results = {}
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            # admit in arrival order while a slot and the budget allow
            while pending and pending[0].arrival <= now and free:
                need = resident_bytes(pending[0])
                if used + need > self.hbm_budget and active:
                    break
                req = pending.pop(0)
                slot = free.pop(0)
                used += need
                active[req.rid] = (_Active(req, self.streams[slot], self.device), slot, need)
            # retire finished requests
            for rid in list(active):
                a, slot, need = active[rid]
                if a.done.query():
                    # the copy is queued on the slot's stream, so the next request admitted to
                    # this slot (whose buffers may reuse this request's blocks) runs after it
                    with torch.cuda.stream(a.stream):
                        est = a.map.estimate.clone()
                    results[rid] = (est, time.perf_counter() - t0 - a.req.arrival)
                    free.append(slot)
                    used -= need
                    del active[rid]
            if not active and pending:
                wait = pending[0].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.01))
            elif active:
                time.sleep(0.0002)
        torch.cuda.synchronize(self.device)
        return results, time.perf_counter() - t0


def plain_loop(requests, device=None, reference_kernels: bool = False):
    """One request at a time with the reference-style autograd MAP loop (one host
    synchronisation per iteration), same preconditioner and iterations as the
    server.  The projectors are the server's, or with ``reference_kernels``
    those of ``_reference_projector`` (the reference's kernel structure, with
    the shim's schedules switched off for the loop's duration)."""
    from . import _C

    device = torch.device("cuda") if device is None else torch.device(device)
    saved = _C.AUTO_LAYOUT
    if reference_kernels:
        _C.AUTO_LAYOUT = False
    try:
        return _plain_loop(requests, device, _reference_projector if reference_kernels else _projector)
    finally:
        _C.AUTO_LAYOUT = saved


def _plain_loop(requests, device, make_projector):
    results = {}
    t0 = time.perf_counter()
    for req in sorted(requests, key=lambda r: r.arrival):
        now = time.perf_counter() - t0
        if req.arrival > now:
            time.sleep(req.arrival - now)
        proj = make_projector(req, device)
        d2 = sqs_preconditioner(ProjectorOps(proj), req.var, req.beta)
        est = MaximumAPosterioriEstimator(LinearGaussianRandomVariable(proj, req.var),
                                          QuadraticSmoothnessRandomVariable(1.0), req.y,
                                          torch.zeros_like(req.x_true), preconditioner=DiagonalPreconditioner(
                                              d2.sqrt().clamp_min(1e-30)), lr=1.0, prior_weight=req.beta,
                                          fused=False)
        for _ in range(req.iters):
            est.step()
        torch.cuda.synchronize(device)
        results[req.rid] = (est.x.clone(), time.perf_counter() - t0 - req.arrival)
    return results, time.perf_counter() - t0


def poisson_workload(n_req, qps, seed=7, frac_2d=0.7, sizes_2d=(256, 320, 384, 448, 512), sizes_3d=(128, 160, 192,
                                                                                                      224, 256),
                     views_2d=(360, 720), views_3d=(180, 720), dets_3d=(256, 384), iters=20, beta=1e-2):
    """Requests with Poisson arrivals at ``qps``: 70% fan-beam (n^2 image, 2n bins),
    30% cone-beam (n^3 volume, det^2 panel)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n_req):
        t += rng.exponential(1.0 / qps)
        if rng.uniform() < frac_2d:
            n = int(rng.choice(sizes_2d))
            reqs.append(MAPRequest(i, "fan2d", n, int(rng.integers(views_2d[0], views_2d[1] + 1)), 2 * n, iters, beta,
                                   seed=i, arrival=t))
        else:
            n = int(rng.choice(sizes_3d))
            det = int(rng.integers(dets_3d[0] // 32, dets_3d[1] // 32 + 1)) * 32
            reqs.append(MAPRequest(i, "cone3d", n, int(rng.integers(views_3d[0], views_3d[1] + 1)), det, iters, beta,
                                   seed=i, arrival=t))
    return reqs


def latency_stats(results):
    lat = np.array([v[1] for v in results.values()])
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean())}
