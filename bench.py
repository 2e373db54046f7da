#!/usr/bin/env python
"""Headline benchmark: 3D circular cone-beam Siddon forward + back projection.

Workload: 256^3 float32 volume at 1.0 mm (20 seeded ellipsoids); 720 views over
360 deg; 384 x 384 flat detector at 1.5 mm pitch; SID 1000 mm, SDD 1500 mm
(106,168,320 rays).  One step = one forward projection of the whole volume plus
one back projection of a dense sinogram (y ~ N(0,1), seed 2: every ray nonzero,
as a MAP residual is), on one GPU, inputs resident in HBM.  The back projectors
skip rays whose value is exactly 0, so a forward projection of the phantom
(65% zero rays at this size) would time a sparse workload; it is reported
beside the dense time, never as it.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--mode torch|otf] [--no-cpu]

Roofline accounting (dominant kernel): algorithmic bytes per launch
    forward: 4*S (volume gathers) + 4*N (sinogram write) + 24*N (src/dst)
    back:    4*S (volume scatter adds) + 4*N (sinogram read) + 24*N (src/dst)
with S = number of positive-length in-grid segments (counted exactly by the
segment-count kernel), N = number of rays; peak HBM 8.0 TB/s.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM = 8.0e12
METRIC = "fwd+back rays/s at 256³ cone-beam (% HBM peak); MAP p50/p99 latency vs plain"

CFG = dict(n=256, voxel=1.0, n_view=720, det_nx=384, det_ny=384, det_spacing=1.5, sid=1000.0, sdd=1500.0)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def build_inputs(dev):
    import torch

    from ct_laboratory_amd.phantoms import ellipsoid_phantom_3d
    from ct_laboratory_amd.tomography.conebeam_projector_3d import build_conebeam_3d_geometry

    c = CFG
    src, dst = build_conebeam_3d_geometry(c["n_view"], c["det_nx"], c["det_ny"], c["sid"], c["sdd"],
                                          det_spacing=c["det_spacing"])
    vol = ellipsoid_phantom_3d(c["n"], c["n"], c["n"], voxel_size=c["voxel"], seed=1)
    n = c["n"]
    b = np.full(3, -(n - 1) / 2.0 * c["voxel"], np.float32)
    minv = np.eye(3, dtype=np.float32) / np.float32(c["voxel"])
    return (src, dst, vol, minv, b,
            src.to(dev), dst.to(dev), torch.from_numpy(vol).to(dev), torch.from_numpy(minv).to(dev),
            torch.from_numpy(b).to(dev))


def cpu_baseline(src, dst, vol, minv, b, mode_name, n_sample_views=8):
    """Oracle (sort-based restatement of the reference torch path) on a bounded
    sample: n_sample_views evenly spaced views, forward + back, all host threads."""
    import oracle as O

    c = CFG
    per_view = c["det_nx"] * c["det_ny"]
    views = np.linspace(0, c["n_view"], n_sample_views, endpoint=False).astype(int)
    sel = np.concatenate([np.arange(v * per_view, (v + 1) * per_view) for v in views])
    s, d = src.numpy()[sel], dst.numpy()[sel]
    threads = len(os.sched_getaffinity(0))
    O.set_threads(threads)
    t0 = time.perf_counter()
    sino = O.forward(vol, minv, b, s, d, mode=mode_name)
    t1 = time.perf_counter()
    O.back(sino.astype(np.float32), vol.shape, minv, b, s, d, mode=mode_name)
    t2 = time.perf_counter()
    n = len(sel)
    return {
        "value": n / (t2 - t0),
        "unit": "fwd+back rays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n_sample_views} evenly spaced views of the workload ({n} rays), oracle forward "
                  f"{t1 - t0:.2f} s + back {t2 - t1:.2f} s, {threads} threads of {cpu_model()}; the oracle is "
                  "a C restatement of the reference's torch rule (OpenMP, -O2), several times faster per "
                  "ray-plane than the reference's own torch CPU path, so the baseline is conservative",
    }


def fan2d_bench(a, dev) -> int:
    """The 2D fan-beam configuration: 256x256 image at 1 mm, 360 views, 512 bins at
    1 mm, SID 1000 / SDD 1500 (184,320 rays), seeded 10-ellipse phantom with a
    0.02/mm background; FanBeam2DProjector with its crossing table as the
    reference builds it (projected on the fly with the table's rule).  Forward +
    back per step on the GPU, and the oracle (C restatement of the torch path) on
    the full problem on the host cores (one warm-up, median of three).  Prints one
    JSON line."""
    import torch

    import oracle as O
    from ct_laboratory_amd.phantoms import ellipse_phantom_2d
    from ct_laboratory_amd.tomography import FanBeam2DProjector

    n, n_view, n_det = 256, 360, 512
    proj = FanBeam2DProjector(n, n, n_view, n_det, 1000.0, 1500.0, det_spacing=1.0, pixel_size=1.0, device=dev)
    img = torch.from_numpy(ellipse_phantom_2d(n, n, 1.0, seed=0)).to(dev)
    n_ray = int(proj.src.shape[0])

    def step():
        proj.back_project(proj.forward_project(img))

    for _ in range(max(1, a.warmup)):
        step()
    torch.cuda.synchronize()
    steps = max(a.steps, 20)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t_gpu = (time.perf_counter() - t0) / steps

    minv = torch.inverse(proj.M).to(torch.float32).cpu().numpy()
    b, src, dst = proj.b.cpu().numpy(), proj.src.cpu().numpy(), proj.dst.cpu().numpy()
    x = img.cpu().numpy()
    threads = len(os.sched_getaffinity(0))
    O.set_threads(threads)

    def cpu_step():
        y = O.forward(x, minv, b, src, dst, mode="torch", zero_rule="thresh")
        O.back(y.astype(np.float32), (n, n), minv, b, src, dst, mode="torch", zero_rule="thresh")

    cpu_step()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        cpu_step()
        times.append(time.perf_counter() - t0)
    t_cpu = float(np.median(times))
    res = {
        "metric": "2D fan-beam forward + back: rays/s (256x256, 360 views x 512 bins)",
        "n_ray": n_ray, "steps": steps,
        "gpu": {"ms_per_step": 1e3 * t_gpu, "rays_per_s": n_ray / t_gpu},
        "cpu_baseline": {"ms_per_step": 1e3 * t_cpu, "rays_per_s": n_ray / t_cpu, "cores": threads, "kind": "port",
                         "sample": f"full problem, median of 3 after a warm-up, {threads} threads of {cpu_model()}"},
        "data": "synthetic (seeded ellipse phantom)",
    }
    print(json.dumps(res), flush=True)
    return 0


def refcalls_bench(a, dev) -> int:
    """The 256^3 cone-beam workload through the _C functions called as the
    reference's ct_projector_3d_cuda.py calls them (no detector layout, otf
    rule): with the shim's layout recognition (tiled forward, brick back
    projection) and without it (one ray per lane in storage order, global
    atomics).  Prints one JSON line."""
    import torch

    from ct_laboratory_amd import _C

    _, _, _, _, _, ts, td, tv, tm, tb = build_inputs(dev)
    ts, td = ts.contiguous(), td.contiguous()
    n = CFG["n"]

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return 1e3 * (time.perf_counter() - t0) / reps

    res = {"metric": "256^3 cone-beam through the reference-style _C calls: ms per call",
           "workload": "720 views, 384x384 detector, otf rule, no det_shape passed"}
    y = _C.forward_project_3d_on_the_fly_cuda(tv, ts, td, tm, tb)
    for auto in (True, False):
        _C.AUTO_LAYOUT = auto
        try:
            res["recognised" if auto else "plain"] = {
                "forward_ms": timed(lambda: _C.forward_project_3d_on_the_fly_cuda(tv, ts, td, tm, tb), 3),
                "back_ms": timed(lambda: _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, n, n, n), 3 if auto else 1),
            }
        finally:
            _C.AUTO_LAYOUT = True
    res["layout"] = _C._layout_of(ts, td)
    print(json.dumps(res), flush=True)
    return 0


def generic_bench(a, dev) -> int:
    """The 256^3 cone-beam workload's rays in a seeded random order, so that no
    detector layout describes them: the brick lists' build time, the back
    projection of a dense sinogram over brick lists and with one ray per lane
    (global atomics), the forward in storage order, the adjoint of the list
    pair, and the _C path that sorts the rays into spatial order once.  Prints
    one JSON line."""
    import torch

    from ct_laboratory_amd import _C
    from ct_laboratory_amd import _native as N

    _, _, _, _, _, ts, td, tv, tm, tb = build_inputs(dev)
    n = CFG["n"]
    n_ray = int(ts.shape[0])
    perm = torch.randperm(n_ray, generator=torch.Generator().manual_seed(5)).to(dev)
    ts, td = ts[perm].contiguous(), td[perm].contiguous()
    del perm
    y = torch.from_numpy(np.random.default_rng(2).standard_normal(n_ray, dtype=np.float32)).to(dev)[None]
    mode = N.MODE_TORCH

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            out = fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / reps, out

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bins = _C.brick_bins(ts, td, tm, tb, (n, n, n), mode=mode)
    torch.cuda.synchronize()
    t_bins = time.perf_counter() - t0
    ws = torch.empty(1, device=dev)
    back_ms, bk = timed(lambda: _C.back_project_3d_brick_bins(y, ts, td, bins, tm, tb, n, n, n, mode=mode,
                                                                 workspace=ws), a.steps)
    fwd_ms, sino = timed(lambda: _C._forward_otf(tv, ts, td, tm, tb, mode, None, None), max(1, a.steps // 2))
    # the _C path: spatial order built once (sort + reordered copies + lists), then
    # the sinogram gathered into that order for each back projection
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    order = _C._order_of(ts, td)
    _C._bins_of(order.src, order.dst, tm, tb, (n, n, n), mode)
    torch.cuda.synchronize()
    t_order = time.perf_counter() - t0
    back_o_ms, bk_o = timed(lambda: _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, n, n, n, mode=mode), a.steps)
    fwd_o_ms, sino_o = timed(lambda: _C.forward_project_3d_on_the_fly_cuda(tv, ts, td, tm, tb, mode=mode),
                             max(1, a.steps // 2))
    _C.AUTO_LAYOUT = False
    try:
        atomic_ms, _ = timed(lambda: _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, n, n, n, mode=mode), 1)
    finally:
        _C.AUTO_LAYOUT = True
    lhs = float(torch.dot(sino[0].double(), y[0].double()))
    rhs = float(torch.dot(tv.reshape(-1).double(), bk.reshape(-1).double()))
    res = {
        "metric": "256^3 cone-beam rays in random order (no detector layout): back projection over brick lists",
        "workload": "720 views x 384x384 rays at 1.5 mm, SID 1000, SDD 1500, shuffled (seed 5); dense y ~ N(0,1)",
        "n_ray": n_ray, "pairs_listed": int(bins.ids.numel()), "pairs_per_ray": bins.ids.numel() / n_ray,
        "bins_build_s": t_bins, "back_bins_ms": back_ms, "back_atomic_ms": atomic_ms, "forward_ray_order_ms": fwd_ms,
        "adjoint_rel_err": abs(lhs - rhs) / abs(lhs),
        "ordered": {"build_s": t_order, "back_ms": back_o_ms, "forward_ms": fwd_o_ms,
                    "back_max_abs_diff_vs_unordered_lists": float((bk_o - bk).abs().max()),
                    "back_max_abs": float(bk.abs().max()),
                    "forward_equal_to_storage_order": bool(torch.equal(sino_o, sino))},
    }
    print(json.dumps(res), flush=True)
    return 0


# The reference's projector timing benchmark (benchmark/projector_timing_benchmark.py)
# runs a calibrated 80-source, 3-rotation StaticCT ring that is not available here.
# The synthetic ring below has its published shape (benchmark/README.md:36-40): 80
# sources x 3 gantry rotations = 240 views, each lighting the 16 farthest of 48
# inward-facing modules of 48 x 16 pixels (12,288 rays per view, 2,949,120 rays);
# radii and pixel pitch are those of the reference's StaticCT example
# (python_examples/test_staticct_projector_3d.py:43-48).  Rotation r turns the
# gantry by r/3 of the source pitch.
RING = dict(n_source=80, n_rotation=3, source_radius=380.0, n_module=48, module_radius=366.7, modules_per_source=16,
            det_nx=48, det_ny=16, det_sx=1.0, det_sy=2.0)
# res: (n_z, voxel pitch mm); 400 mm transaxial FOV, +5 mm z shift
# (projector_timing_benchmark.py:43-51, 106-118)
RING_GRIDS = {64: (8, 6.4), 128: (16, 3.2), 256: (32, 1.6), 512: (64, 0.8)}
RING_B_SHIFT_Z = 5.0
# H100 80GB HBM3, res512, CUDA kernels with uint16 tables: forward / back ms
# (benchmark/README.md:102)
RING_H100_U16_MS = (46.2, 36.1)


def staticct_ring_rays(dev):
    """src, dst [2,949,120, 3] of the synthetic 240-view ring, in the StaticCT
    builders' order (frame, module, u, v)."""
    import math

    import torch

    from ct_laboratory_amd.tomography.staticct_projector_3d import (build_static_3d_geometry,
                                                                    build_uniform_static_3d_geometry)

    r = RING
    sp, mc, mo, mask = build_uniform_static_3d_geometry(r["n_source"], r["source_radius"], 0.0, r["n_module"],
                                                        r["module_radius"], 0.0, r["modules_per_source"])
    n_frame = r["n_source"] * r["n_rotation"]
    Mg = torch.zeros(n_frame, 3, 3)
    active = torch.zeros(n_frame, r["n_source"], dtype=torch.bool)
    for f in range(n_frame):
        ang = (f // r["n_source"]) * 2.0 * math.pi / (r["n_source"] * r["n_rotation"])
        Mg[f] = torch.tensor([[math.cos(ang), -math.sin(ang), 0.0], [math.sin(ang), math.cos(ang), 0.0],
                              [0.0, 0.0, 1.0]])
        active[f, f % r["n_source"]] = True
    return build_static_3d_geometry(sp, mc, mo, r["det_nx"], r["det_ny"], r["det_sx"], r["det_sy"], mask, Mg,
                                    torch.zeros(n_frame, 3), active, device=dev)


def ring_grid(res: int, dev):
    """(M, b, n_z) of the 400 mm FOV grid (projector_timing_benchmark.py:106-118)."""
    import torch

    nz, sp = RING_GRIDS[res]
    M = torch.eye(3, device=dev) * sp
    b = torch.tensor([-(res - 1) / 2.0 * sp, -(res - 1) / 2.0 * sp, -(nz - 1) / 2.0 * sp + RING_B_SHIFT_Z], device=dev)
    return M, b, nz


def staticct_bench(a, dev) -> int:
    """The reference's projector timing benchmark on the synthetic ring: for each
    grid, CTProjector3DModule built as that benchmark builds it (uint16 tables
    from precompute_tvals_stitched in chunks of 32,768 rays, float tables, or
    none: on the fly), plus the on-the-fly module told the rays' module layout
    (48 x 16 pixels per source and module); forward_project of a randn volume
    and back_project of its sinogram, median wall ms over the repeats with a
    synchronize either side of each call, as the reference times them.  Prints
    one JSON line."""
    import torch

    from ct_laboratory_amd import _C
    from ct_laboratory_amd import _native as N
    from ct_laboratory_amd.tomography.ct_projector_3d_module import CTProjector3DModule, precompute_tvals_stitched

    src, dst = staticct_ring_rays(dev)
    n_ray = int(src.shape[0])
    print(f"[staticct] ring built: {n_ray} rays", file=sys.stderr, flush=True)
    repeats = max(3, a.steps)

    def median_ms(fn):
        vals, out = [], None
        for _ in range(repeats + 1):  # the first call is a warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            vals.append(1e3 * (time.perf_counter() - t0))
        return float(np.median(vals[1:])), out

    rows, segments = [], {}
    for res in a.ring_res:
        M, b, nz = ring_grid(res, dev)
        vol = torch.randn(res, res, nz, device=dev, generator=torch.Generator(device=dev).manual_seed(res))
        # positive-length in-grid segments of the torch rule (the on-the-fly segments)
        _, _, cnt = _C.trace_segments((res, res, nz), src, dst, _C.cached_inverse(M), b, mode=N.MODE_TORCH, cap=0)
        segments[res] = int(cnt.to(torch.int64).sum().item())
        del cnt
        ref = {}
        for label in ("uint16", "float", "none", "none+layout"):
            print(f"[staticct] res{res} {label}", file=sys.stderr, flush=True)
            torch.cuda.empty_cache()
            torch.cuda.reset_peak_memory_stats(dev)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if label in ("uint16", "float"):
                comp = label == "uint16"
                tv = precompute_tvals_stitched(res, res, nz, M.cpu(), b.cpu(), src.cpu(), dst.cpu(), chunk_size=32768,
                                               device=dev, verbose=False, use_compression=comp)
                proj = CTProjector3DModule(res, res, nz, M, b, src, dst, device=dev, tvals=tv, use_compression=comp)
                del tv
            else:
                det = (RING["det_nx"], RING["det_ny"]) if label == "none+layout" else None
                proj = CTProjector3DModule(res, res, nz, M, b, src, dst, device=dev, det_shape=det)
            torch.cuda.synchronize()
            build_ms = 1e3 * (time.perf_counter() - t0)
            fwd_ms, sino = median_ms(lambda: proj.forward_project(vol))
            back_ms, bk = median_ms(lambda: proj.back_project(sino))
            lhs = float(torch.dot(sino.reshape(-1).double(), sino.reshape(-1).double()))
            rhs = float(torch.dot(vol.reshape(-1).double(), bk.reshape(-1).double()))
            row = {"res": res, "n_z": nz, "voxel_mm": RING_GRIDS[res][1], "tvals": label, "build_ms": build_ms,
                   "forward_ms": fwd_ms, "back_ms": back_ms,
                   "pair_rays_per_s": n_ray / (1e-3 * (fwd_ms + back_ms)),
                   "peak_gpu_gb": torch.cuda.max_memory_allocated(dev) / 1e9, "adjoint_rel_err": abs(lhs - rhs) / lhs}
            if getattr(proj, "_runs", None) is not None:
                row["table_runs"] = proj._runs.n_runs
                row["table_units"] = proj._runs.n_units
            if label in ("uint16", "float", "none"):
                ref[label] = sino.reshape(-1)
            if label != "uint16" and "uint16" in ref:
                d = sino.reshape(-1) - ref["uint16"]
                row["forward_rel_diff_vs_uint16"] = float(d.norm() / ref["uint16"].norm())
            if label == "none+layout":
                row["forward_equal_to_none"] = bool(torch.equal(sino.reshape(-1), ref["none"]))
            if res == 512 and label == "uint16":
                row["h100_forward_back_ms"] = list(RING_H100_U16_MS)
                row["speedup_vs_h100"] = [RING_H100_U16_MS[0] / fwd_ms, RING_H100_U16_MS[1] / back_ms]
            print(f"[staticct] {json.dumps(row)}", file=sys.stderr, flush=True)
            rows.append(row)
            del proj, sino, bk
        del ref, vol
    res = {"metric": "StaticCT ring (240 views x 12,288 rays) CTProjector3DModule forward / back ms, median",
           "workload": "synthetic 80-source x 3-rotation ring, 16 of 48 modules of 48x16 px per view, "
                       "2,949,120 rays; 400 mm FOV grids with +5 mm z shift; randn volume",
           "n_ray": n_ray, "segments": segments, "repeats": repeats, "rows": rows}
    print(json.dumps(res), flush=True)
    return 0


def batched_bench(a, dev) -> int:
    """64 independent 2D scans (sizes 128^2-512^2, 90-720 views, half fan beam, half
    StaticCT ring; scan i seeded by default_rng(100 + i)): forward + back of all
    scans as one persistent launch each way (tuned) against 64 sequential
    single-scan launches each way.  Both back projections use the LDS tile
    kernel; the persistent one-ray-per-lane back projection with global float
    atomics is timed beside them.  Prints one JSON line."""
    import torch

    from ct_laboratory_amd import _C
    from ct_laboratory_amd.batched import BatchedProjector2D, c4_workload

    scans, images = c4_workload(64, dev)
    n_rays = sum(sc.n_ray for sc in scans)
    minvs = [sc.minv.to(dev) for sc in scans]
    bs = [sc.b.to(dev) for sc in scans]
    sinos = [torch.empty((1, sc.n_ray), device=dev) for sc in scans]
    backs = [torch.empty((1, sc.n_row, sc.n_col), device=dev) for sc in scans]

    def sequential():
        for k, sc in enumerate(scans):
            _C.forward_project_2d_on_the_fly_cuda(images[k], sc.src, sc.dst, minvs[k], bs[k], mode=sc.mode,
                                                  zero_rule=sc.zero_rule, out=sinos[k])
        for k, sc in enumerate(scans):
            _C.back_project_2d_on_the_fly_cuda(sinos[k], sc.src, sc.dst, minvs[k], bs[k], sc.n_row, sc.n_col,
                                               mode=sc.mode, zero_rule=sc.zero_rule, out=backs[k])

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    t_seq = timed(sequential, 5)
    bp = BatchedProjector2D(scans, dev)
    best, timing = bp.autotune(images)
    ys = [s[0] for s in sinos]

    def batched():
        out = bp.forward(images)
        bp.back(out)

    t_bat = timed(batched, 5)
    out = bp.forward(images)
    same = all(torch.equal(out[k], sinos[k][0]) for k in range(len(scans)))
    bp_atomic = BatchedProjector2D(scans, dev, rays_per_item=bp.rays_per_item, groups_per_cu=bp.groups_per_cu,
                                   tiled_back=False)

    def batched_atomic():
        bp_atomic.back(bp_atomic.forward(images))

    t_atomic = timed(batched_atomic, 3)
    res = {
        "metric": "64 uneven 2D scans, forward + back: aggregate rays/s, one persistent launch vs 64 launches",
        "n_scans": len(scans), "n_rays": n_rays,
        "classes": {c: sum(sc.geometry_class == c for sc in scans) for c in ("fan", "ring")},
        "sequential": {"ms": 1e3 * t_seq, "rays_per_s": n_rays / t_seq},
        "batched": {"ms": 1e3 * t_bat, "rays_per_s": n_rays / t_bat, "rays_per_item": best[0],
                    "groups_per_cu": best[1], "rays_per_part": best[2], "tile": f"{best[3]}x64"},
        "speedup": t_seq / t_bat,
        "batched_atomic_back": {"ms": 1e3 * t_atomic, "rays_per_s": n_rays / t_atomic},
        "forward_bit_identical": same,
        "tuner_ms": {"forward": {f"{k[0]}x{k[1]}": v for k, v in sorted(timing["forward"].items())},
                     "back": {f"{k[0]}x64 tiles, {k[1]} rays per part": v
                              for k, v in sorted(timing["back"].items())}},
    }
    print(json.dumps(res), flush=True)
    return 0


def serve_bench(a, dev) -> int:
    """Served MAP stream: 70% fan-beam (256^2-512^2, 360-720 views, 2n bins), 30%
    cone-beam (128^3-256^3, 180-720 views, 256^2-384^2 panel), 20 iterations each
    with the quadratic prior (weight 1e-2), Poisson arrivals; admission by a 200 GB
    HBM budget.  The server runs first, then the plain loop; each phase is
    bracketed by a short spin kernel so that a kernel trace of this run splits
    into the two phases.  Prints one JSON line."""
    import torch

    from ct_laboratory_amd.serving import ReconServer, build_problem, latency_stats, plain_loop, poisson_workload

    reqs = poisson_workload(a.requests, a.qps)
    print(f"[serve] building {len(reqs)} problems", file=sys.stderr, flush=True)
    for i, r in enumerate(reqs):
        build_problem(r, dev)
        if i % 20 == 19:
            print(f"[serve] {i + 1} problems built", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    # warm the library and torch kernels once on a small request
    from ct_laboratory_amd.serving import MAPRequest
    warm = [MAPRequest(-1, "fan2d", 64, 30, 128, iters=2), MAPRequest(-2, "cone3d", 32, 8, 32, iters=2)]
    for r in warm:
        build_problem(r, dev)
    ReconServer(dev, slots=2).run(warm)
    plain_loop(warm, dev)
    def mark():
        # a spin kernel of ~1000 cycles marks a phase boundary in a kernel trace
        torch.cuda.synchronize()
        torch.cuda._sleep(1000)
        torch.cuda.synchronize()

    print("[serve] server", file=sys.stderr, flush=True)
    mark()
    served, span_s = ReconServer(dev, hbm_budget=200e9, slots=a.slots).run(reqs)
    mark()
    print("[serve] plain loop", file=sys.stderr, flush=True)
    plain, span_p = plain_loop(reqs, dev)
    mark()

    def rmse(res):
        e = [float(((res[r.rid][0] - r.x_true) ** 2).mean().sqrt()) for r in reqs]
        return float(np.mean(e))

    def by_kind(res):
        out = {}
        for kind in ("fan2d", "cone3d"):
            sub = {r.rid: res[r.rid] for r in reqs if r.kind == kind}
            if sub:
                out[kind] = latency_stats(sub)
        return out

    worst = max(float((served[r.rid][0] - plain[r.rid][0]).abs().max() / (plain[r.rid][0].abs().max() + 1e-30))
                for r in reqs)
    ref = None
    if a.ref_plain:
        # the first requests as one burst (all arrived at 0): the server, the plain loop, and the
        # plain loop on the reference's kernel structure (table kernels with per-segment atomics
        # for fan-beam, storage-order on-the-fly kernels with atomics for cone-beam)
        import dataclasses
        sub = [dataclasses.replace(r, arrival=0.0) for r in sorted(reqs, key=lambda r: r.arrival)[:a.ref_plain]]
        plain_loop(warm, dev, reference_kernels=True)
        print(f"[serve] burst of {len(sub)}: server", file=sys.stderr, flush=True)
        mark()
        b_served, b_span_s = ReconServer(dev, hbm_budget=200e9, slots=a.slots).run(sub)
        mark()
        print(f"[serve] burst of {len(sub)}: plain loop", file=sys.stderr, flush=True)
        b_plain, b_span_p = plain_loop(sub, dev)
        mark()
        print(f"[serve] burst of {len(sub)}: plain loop, reference kernel structure", file=sys.stderr, flush=True)
        b_ref, b_span_r = plain_loop(sub, dev, reference_kernels=True)
        mark()
        ref = {
            "n_requests": len(sub),
            "mix": {"fan2d": sum(r.kind == "fan2d" for r in sub), "cone3d": sum(r.kind == "cone3d" for r in sub)},
            "server_s": b_span_s, "plain_s": b_span_p, "plain_reference_kernels_s": b_span_r,
            "throughput_ratio": b_span_p / b_span_s,
            "throughput_ratio_vs_reference_kernels": b_span_r / b_span_s,
            "max_rel_diff_reference_kernels_vs_plain": max(
                float((b_ref[r.rid][0] - b_plain[r.rid][0]).abs().max() / (b_plain[r.rid][0].abs().max() + 1e-30))
                for r in sub),
        }
    out = {
        "metric": "served MAP stream: requests/s and p50/p99 latency, continuous batching vs plain per-request loop",
        "n_requests": len(reqs), "qps": a.qps, "slots": a.slots, "hbm_budget_gb": 200,
        "mix": {"fan2d": sum(r.kind == "fan2d" for r in reqs), "cone3d": sum(r.kind == "cone3d" for r in reqs)},
        "server": {"requests_per_s": len(reqs) / span_s, "makespan_s": span_s, **latency_stats(served),
                   "mean_rmse": rmse(served), "by_kind": by_kind(served)},
        "plain": {"requests_per_s": len(reqs) / span_p, "makespan_s": span_p, **latency_stats(plain),
                  "mean_rmse": rmse(plain), "by_kind": by_kind(plain)},
        "throughput_ratio": span_p / span_s,
        "max_rel_diff_server_vs_plain": worst,
    }
    if ref is not None:
        out["burst"] = ref
    print(json.dumps(out), flush=True)
    return 0


def map_bench(a, dev) -> int:
    """Full-size MAP reconstruction: quadratic smoothness prior (weight
    1e-2, non-periodic) on a 512x512x256 volume at 0.75 mm, 600 views, 512x384
    detector at 1.125 mm, SID 1000 / SDD 1500 (117,964,800 rays).  Data = A
    x_true + sigma N(0,1) (default_rng(3)), sigma = 0.01 max(A x_true).  Times
    the fused, graph-captured iteration (SQS preconditioner) and, for a few
    iterations, the reference-style autograd loop with the same projector and
    preconditioner.  Prints one JSON line."""
    import torch

    from ct_laboratory_amd.map_estimator import (DiagonalPreconditioner, LinearGaussianRandomVariable,
                                                 MaximumAPosterioriEstimator, ProjectorOps, QuadraticMAP,
                                                 QuadraticSmoothnessRandomVariable, sqs_preconditioner)
    from ct_laboratory_amd.phantoms import ellipsoid_phantom_3d
    from ct_laboratory_amd.tomography.conebeam_projector_3d import ConeBeam3DProjector

    nx, ny, nz, vox = 512, 512, 256, 0.75
    proj = ConeBeam3DProjector(nx, ny, nz, 600, 512, 384, 1000.0, 1500.0, det_spacing=1.125, voxel_size=vox,
                               device=dev)
    n_ray = int(proj.src.shape[0])
    x_true = torch.from_numpy(ellipsoid_phantom_3d(nx, ny, nz, vox, seed=1)).to(dev)
    with torch.no_grad():
        y0 = proj.forward_project(x_true)
    sigma = 0.01 * float(y0.max())
    noise = torch.from_numpy(np.random.default_rng(3).standard_normal(n_ray).astype(np.float32)).to(dev)
    y = (y0 + sigma * noise).contiguous()
    beta, var = 1e-2, sigma * sigma
    print(f"[map] n_ray={n_ray} sigma={sigma:.4g}", file=sys.stderr, flush=True)
    ops = ProjectorOps(proj)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d2 = sqs_preconditioner(ops, var, beta)
    torch.cuda.synchronize()
    t_sqs = time.perf_counter() - t0
    m = QuadraticMAP(ops, y, var, beta, torch.zeros_like(x_true), d2=d2)
    m.run(2)  # capture (and warm) the graph
    torch.cuda.synchronize()
    e_start = m.energies()
    iters = max(2, a.map_iters - a.map_iters % 2)
    t0 = time.perf_counter()
    m.run(iters)
    torch.cuda.synchronize()
    t_fused = (time.perf_counter() - t0) / iters
    m.step()
    e_end = m.energies()
    rmse = float(((m.estimate - x_true) ** 2).mean().sqrt())
    # reference-style loop: autograd gradient, preconditioner and update in torch ops
    est = MaximumAPosterioriEstimator(LinearGaussianRandomVariable(proj, var), QuadraticSmoothnessRandomVariable(1.0),
                                      y, torch.zeros_like(x_true),
                                      preconditioner=DiagonalPreconditioner(d2.sqrt().clamp_min(1e-30)), lr=1.0,
                                      prior_weight=beta, fused=False)
    est.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        est.step()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / 3
    res = {
        "metric": "MAP iteration, 512x512x256 cone-beam at 0.75 mm, 600 views (fused, graph-captured) vs the reference-style autograd loop",
        "workload": "512x512x256 at 0.75 mm, 600 views, 512x384 detector at 1.125 mm, SID 1000, SDD 1500; quadratic "
                    "prior 1e-2; SQS preconditioner",
        "n_ray": n_ray, "iterations_timed": iters,
        "fused_ms_per_iter": 1e3 * t_fused, "fused_rays_per_s": n_ray / t_fused,
        "plain_ms_per_iter": 1e3 * t_plain, "speedup": t_plain / t_fused,
        "projected_50_iter_s": 50 * t_fused + t_sqs, "sqs_setup_s": t_sqs,
        "energy_start": [float(e) for e in e_start], "energy_end": [float(e) for e in e_end],
        "rmse_after": rmse, "data": "synthetic (seeded ellipsoid phantom)",
    }
    print(json.dumps(res), flush=True)
    return 0


def trace_summary(path: str) -> int:
    """Device-busy breakdown of a rocprofv3 kernel trace (its *_results.db) of
    a ``--serve`` run: the run's phases are delimited by the spin kernels it
    launches between them; per phase, the wall time from marker to marker and
    the time with at least one kernel running (kernel intervals merged), with
    the busy time split by kernel name.  Prints one JSON line."""
    import sqlite3
    from collections import defaultdict

    con = sqlite3.connect(path)
    rows = con.execute("select name, start, end from kernels order by start").fetchall()
    marks = [r for r in rows if "spin" in r[0].lower() or "sleep" in r[0].lower()]
    if len(marks) < 3:
        print(json.dumps({"error": f"expected 3 phase markers, found {len(marks)}"}))
        return 1
    phases = {"server": (marks[0][2], marks[1][1]), "plain": (marks[1][2], marks[2][1])}
    if len(marks) >= 7:  # a --ref-plain burst: server, plain loop, plain loop on the reference's kernel structure
        for i, name in enumerate(("burst_server", "burst_plain", "burst_plain_reference_kernels")):
            phases[name] = (marks[3 + i][2], marks[4 + i][1])
    out = {}
    for name, (t0, t1) in phases.items():
        ks = [r for r in rows if r[1] >= t0 and r[2] <= t1]
        busy, cur_s, cur_e = 0, None, None
        by_kernel = defaultdict(int)
        for kname, s0, s1 in ks:
            by_kernel[kname.split("(")[0].replace("void ", "")[:60]] += s1 - s0
            if cur_e is None or s0 > cur_e:
                if cur_e is not None:
                    busy += cur_e - cur_s
                cur_s, cur_e = s0, s1
            else:
                cur_e = max(cur_e, s1)
        if cur_e is not None:
            busy += cur_e - cur_s
        top = sorted(by_kernel.items(), key=lambda kv: -kv[1])[:8]
        out[name] = {"wall_s": (t1 - t0) * 1e-9, "busy_s": busy * 1e-9, "busy_frac": busy / max(t1 - t0, 1),
                     "kernels": len(ks), "kernel_time_s": sum(by_kernel.values()) * 1e-9,
                     "top_kernels_s": {k: v * 1e-9 for k, v in top}}
    print(json.dumps(out), flush=True)
    return 0


def measured_traffic(kernel: str, a) -> float | None:
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of this workload (scaled by the same counters
    on a copy of known size).  None when no summary matches the configuration."""
    path = os.path.join(HERE, "profiles", "conebeam256_hbm_traffic.md")
    if a.mode != "torch" or a.back != "brick" or a.plain_order or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("bench-readable:"):
                d = json.loads(line.split(":", 1)[1])
                return d.get(kernel)
    return None


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["torch", "otf"], default="torch")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-views", type=int, default=8)
    ap.add_argument("--plain-order", action="store_true", help="one ray per lane in storage order (no tiling)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the full-size adjoint check")
    ap.add_argument("--no-sparse", action="store_true",
                    help="skip the extra back projection of the forward output (profiling runs)")
    ap.add_argument("--back", choices=["brick", "atomic"], default="brick",
                    help="back projector: LDS bricks (no global atomics) or per-ray atomics")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 device copies of 1 GiB (the reference kernel for scaling HBM counters)")
    ap.add_argument("--serve", action="store_true",
                    help="served MAP stream instead of the projector pair: Poisson arrivals, continuous batching "
                         "vs the plain per-request loop (p50/p99 latency, throughput)")
    ap.add_argument("--fan2d", action="store_true",
                    help="the 2D fan-beam configuration (256^2, 360 views x 512 bins) with its CPU baseline")
    ap.add_argument("--refcalls", action="store_true",
                    help="the 256^3 cone-beam workload through the _C calls as the reference's Python makes them")
    ap.add_argument("--generic", action="store_true",
                    help="the 256^3 workload's rays in random order: back projection over brick lists vs atomics")
    ap.add_argument("--staticct", action="store_true",
                    help="the reference's projector timing benchmark on a synthetic 240-view StaticCT ring "
                         "(uint16 / float tables / on the fly through CTProjector3DModule)")
    ap.add_argument("--ring-res", type=int, nargs="+", default=[256, 512], choices=sorted(RING_GRIDS))
    ap.add_argument("--batched", action="store_true",
                    help="64 uneven 2D scans: one persistent launch (autotuned) vs 64 sequential launches")
    ap.add_argument("--map", action="store_true",
                    help="MAP iterations on the 512x512x256 / 600-view cone-beam problem (fused graph vs autograd loop); one JSON line")
    ap.add_argument("--map-iters", type=int, default=6)
    ap.add_argument("--trace-summary", default=None, metavar="DB",
                    help="device-busy breakdown of a rocprofv3 kernel-trace database of a --serve run; one JSON line")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--qps", type=float, default=40.0)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--ref-plain", type=int, default=0, metavar="N",
                    help="with --serve: also run the first N requests as one burst through the server, the plain "
                         "loop and the plain loop on the reference's kernel structure")
    a = ap.parse_args(argv)
    if a.trace_summary:
        return trace_summary(a.trace_summary)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the projector shards over rays "
              "across GPUs by design but no multi-GPU run is part of this build.", file=sys.stderr)
        return 2

    import torch

    from ct_laboratory_amd import _C
    from ct_laboratory_amd import _native as N

    dev = torch.device("cuda:0")
    if a.serve:
        return serve_bench(a, dev)
    if a.map:
        return map_bench(a, dev)
    if a.batched:
        return batched_bench(a, dev)
    if a.generic:
        return generic_bench(a, dev)
    if a.staticct:
        return staticct_bench(a, dev)
    if a.fan2d:
        return fan2d_bench(a, dev)
    if a.refcalls:
        return refcalls_bench(a, dev)
    if a.copy_calibration:
        x = torch.ones(1 << 28, device=dev)  # 1 GiB of float32
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": x.numel() * 4, "copies": 3}), flush=True)
        return 0
    mode = N.MODES[a.mode]
    print(f"[bench] building inputs (cone-beam {CFG})", file=sys.stderr, flush=True)
    src_c, dst_c, vol_c, minv_c, b_c, src, dst, vol, minv, b = build_inputs(dev)
    n = CFG["n"]
    n_ray = int(src.shape[0])

    # Exact segment count S (positive-length, in-grid) per ray.
    _, _, cnt = _C.trace_segments((n, n, n), src, dst, minv, b, mode=mode, cap=0)
    S = int(cnt.to(torch.int64).sum().item())
    del cnt
    print(f"[bench] n_ray={n_ray} segments={S}", file=sys.stderr, flush=True)

    det = None if a.plain_order else (CFG["det_ny"], CFG["det_nx"])
    frames = None
    if a.back == "brick":
        frames = _C.cone_frames(src, dst, (CFG["det_ny"], CFG["det_nx"]), minv, b, (n, n, n))
        plan = _C.brick_plan(frames, (CFG["det_ny"], CFG["det_nx"]), (n, n, n))
        # geometry-only filed unit lists (registered on the plan; built once, outside the timed steps)
        torch.cuda.synchronize()
        t_l = time.perf_counter()
        lists = _C.brick_filed_lists(src, dst, (CFG["det_ny"], CFG["det_nx"]), plan, minv, b, (n, n, n))
        torch.cuda.synchronize()
        lists_s = time.perf_counter() - t_l
        print(f"[bench] filed unit lists: {lists.ids.numel()} (unit, ray) pairs in {lists_s:.3f} s",
              file=sys.stderr, flush=True)
        ws = torch.empty(CFG["n_view"], device=dev)
        n_valid = int((frames[:, 15] == 1).sum().item())
        print(f"[bench] detector frames: {n_valid}/{frames.shape[0]} views fit the flat-panel model",
              file=sys.stderr, flush=True)

    def back(y):
        if a.back == "brick":
            return _C.back_project_3d_brick(y, src, dst, (CFG["det_ny"], CFG["det_nx"]), plan, minv, b, n, n, n,
                                            workspace=ws,
                                            mode=mode)
        return _C.back_project_3d_on_the_fly_cuda(y, src, dst, minv, b, n, n, n, mode=mode, det_shape=det)

    # dense sinogram for the back projection: every ray carries a nonzero value
    y_dense = torch.from_numpy(np.random.default_rng(2).standard_normal(n_ray, dtype=np.float32)).to(dev)
    n_zero = int((y_dense == 0).sum().item())

    def step(ev=None):
        if ev is not None:
            ev[0].record()
        sino = _C.forward_project_3d_on_the_fly_cuda(vol, src, dst, minv, b, mode=mode, det_shape=det)
        if ev is not None:
            ev[1].record()
        bk = back(y_dense)
        if ev is not None:
            ev[2].record()
        return sino, bk

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(a.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        sino, bk = step(evs[k])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    fwd_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    back_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    total_s = t1 - t0
    ms_per_step = 1e3 * total_s / a.steps
    value = n_ray * a.steps / total_s

    adjoint = None
    if not a.no_adjoint:
        # Config check: <A x, y> vs <x, A^T y> with y ~ N(0, 1) (seed 2, the timed back
        # projection's input), FP64 dot products.
        lhs = float(torch.dot(sino[0].double(), y_dense.double()))
        rhs = float(torch.dot(vol.reshape(-1).double(), bk.reshape(-1).double()))
        adjoint = abs(lhs - rhs) / abs(lhs)

    # Beside the headline: the back projection of the forward projection itself,
    # whose zero rays the back projectors skip (not the timed workload).
    back_ms_sparse = zero_rays_sparse = None
    if not a.no_sparse:
        sparse_ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        sparse_ev[0].record()
        back(sino)
        sparse_ev[1].record()
        torch.cuda.synchronize()
        back_ms_sparse = sparse_ev[0].elapsed_time(sparse_ev[1])
        zero_rays_sparse = int((sino == 0).sum().item())

    fwd_bytes = 4 * S + 4 * n_ray + 24 * n_ray
    back_bytes = 4 * S + 4 * n_ray + 24 * n_ray
    dom, dom_bytes, dom_ms = ("back_project_3d", back_bytes, back_ms) if back_ms >= fwd_ms else (
        "forward_project_3d", fwd_bytes, fwd_ms)
    achieved = dom_bytes / (dom_ms * 1e-3)
    pair_bytes = fwd_bytes + back_bytes
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "fwd+back rays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid phantom, cone-beam rays built in-process)",
        "config": {
            "workload": "3D circular cone-beam 256^3 f32, 1.0 mm voxels, 720 views, 384x384 detector at 1.5 mm, "
                        "SID 1000 mm, SDD 1500 mm; forward + back projection per step",
            "n_ray": n_ray,
            "segments": S,
            "semantics": a.mode,
            "schedule": "plain" if a.plain_order else "detector-tiled (64 rows x 4 cols per block)",
            "back_projector": ("brick over filed unit lists (geometry-only, built once per plan in %.2f s: %d "
                               "(unit, ray) pairs)" % (lists_s, lists.ids.numel())) if a.back == "brick" else a.back,
            "back_input": "dense y ~ N(0,1), seed 2 (rays skipped for a zero value: %d of %d)" % (n_zero, n_ray),
            "segments_traced_back": S,
            "adjoint_rel_err": adjoint,
            "fwd_ms": fwd_ms,
            "back_ms": back_ms,
            "voxel_updates_per_s": 2 * S * a.steps / total_s,
            "pair_hbm_frac": pair_bytes / (1e-3 * (fwd_ms + back_ms)) / PEAK_HBM,
            "back_ms_of_forward_output": back_ms_sparse,
            "zero_rays_of_forward_output": zero_rays_sparse,
        },
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": PEAK_HBM / 1e9,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM,
            "traffic": measured_traffic(dom, a),
        },
    }
    if not a.no_cpu:
        print("[bench] cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(src_c, dst_c, vol_c, minv_c, b_c, a.mode, a.cpu_views)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
