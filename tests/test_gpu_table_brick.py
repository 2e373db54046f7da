"""Table back projection in LDS bricks (ctlab_table_runs_count_3d / _fill_3d,
ctlab_back_project_3d_table_brick) against the oracle's table consumers
(back_tvals, back_u16: restatements of src/ct_projector_3d.cu:215-269, 325-376)
and the per-segment atomic kernels.

* Random grids that are not multiples of the 32 x 32 x 16 brick, diagonal and
  rotated maps, float tables with INF padding and uint16 tables with the CUDA
  rule, batches with zero values, bricks split into many work units.
* Run records: units tile the run list once, each ray's segment ranges are
  disjoint and inside [1, K).  (A segment filed under the wrong brick would be
  dropped by the kernel's brick bounds check and show as missing mass in the
  comparisons with the oracle.)
* The 2,949,120-ray StaticCT ring (240 views of 16 modules x 48 x 16 pixels)
  on the 512 x 512 x 64 grid at 0.8 mm with its uint16 table: the brick result
  against the atomic kernel on a dense sinogram, and against the oracle on a
  sample of rays.

Bar: back projections of signed data within 1e-5 of the back projection of
their magnitudes, voxel by voxel (the brick path groups the float sums
differently from the atomic one, and both differ from the oracle's float64
sums only by rounding).
"""
import numpy as np
import pytest
import torch

import oracle as O
from ct_laboratory_amd import _C
from helpers import random_rays

pytestmark = pytest.mark.gpu


def _t(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def _within_bound(got, ref, bound, tol=1e-5):
    err = np.abs(np.asarray(got, np.float64) - ref)
    ok = err <= tol * bound + 1e-30
    return bool(ok.all()), float((err / np.maximum(bound, 1e-30)).max())


def _case(rng, rotated, n_ray=2500):
    shape = (70, 45, 37)
    s = rng.uniform(0.6, 1.4, 3).astype(np.float32)
    M = np.diag(s).astype(np.float64)
    if rotated:
        th = 0.4
        R = np.eye(3)
        R[0, 0] = R[1, 1] = np.cos(th)
        R[0, 1], R[1, 0] = -np.sin(th), np.sin(th)
        M = R @ M
    M = M.astype(np.float32)
    b = (-(np.array(shape) - 1) / 2 * s + rng.uniform(-2, 2, 3)).astype(np.float32)
    minv = np.linalg.inv(M.astype(np.float64)).astype(np.float32)
    src, dst = random_rays(rng, shape, s, n_ray, "far")
    s2, d2 = random_rays(rng, shape, s, n_ray // 5, "inside")
    return shape, M, minv, b, np.concatenate([src, s2]), np.concatenate([dst, d2])


@pytest.mark.parametrize("rotated", [False, True])
@pytest.mark.parametrize("kind", ["f32", "u16"])
def test_table_brick_vs_oracle_and_atomic(dev, kind, rotated):
    rng = np.random.default_rng(700 + 2 * rotated + (kind == "u16"))
    shape, M, minv, b, src, dst = _case(rng, rotated)
    n_ray = src.shape[0]
    tab = O.intersections(shape, minv, b, src, dst)
    ys = rng.standard_normal((2, n_ray)).astype(np.float32)
    ys[:, rng.random(n_ray) < 0.2] = 0.0
    ts, td, tm, tb = _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev)
    if kind == "u16":
        u, st, sc = O.compress_cuda(tab, src, dst, O.voxel_diag_len(M))
        table = (_t(u.view(np.int16), dev), _t(st, dev), _t(sc, dev))
        ref = O.back_u16(ys, u, st, sc, shape, minv, b, src, dst)
        bound = O.back_u16(np.abs(ys), u, st, sc, shape, minv, b, src, dst)
    else:
        table = _t(tab, dev)
        ref = O.back_tvals(ys, tab, shape, minv, b, src, dst)
        bound = O.back_tvals(np.abs(ys), tab, shape, minv, b, src, dst)
    print(f"[table_brick] {kind} rotated={rotated} shape={shape} n_ray={n_ray}", flush=True)
    for unit_runs in (_C.TABLE_UNIT_RUNS, 7):
        runs = _C.table_runs(table, ts, td, tm, tb, shape, unit_runs=unit_runs)
        rays = runs.runs[:runs.n_runs, 0].cpu().numpy()
        assert runs.n_runs > 2 * len(np.unique(rays)) > 0  # rays cross several bricks
        got = _C.back_project_3d_table_brick(_t(ys, dev), table, ts, td, runs, tm, tb, *shape)
        ok, worst = _within_bound(got.cpu().numpy(), ref, bound)
        assert ok, (unit_runs, worst)
    # a table that does not start on a 16-byte boundary (a view 3 entries into a buffer)
    t_dev = table[0] if kind == "u16" else table
    buf = torch.empty(t_dev.numel() + 3, dtype=t_dev.dtype, device=dev)
    buf[3:] = t_dev.reshape(-1)
    view = buf[3:].view(t_dev.shape)
    assert view.data_ptr() % 16 != 0
    shifted = (view, table[1], table[2]) if kind == "u16" else view
    for _ in range(1):
        runs = _C.table_runs(shifted, ts, td, tm, tb, shape)
        got = _C.back_project_3d_table_brick(_t(ys, dev), shifted, ts, td, runs, tm, tb, *shape)
        ok, worst = _within_bound(got.cpu().numpy(), ref, bound)
        assert ok, ("shifted", worst)
    # the per-segment atomic kernels on the same table
    if kind == "u16":
        atomic = _C.back_project_3d_compressed_cuda(_t(ys, dev), *table, ts, td, tm, tb, *shape)
    else:
        atomic = _C.back_project_3d_cuda(_t(ys, dev), table, ts, td, tm, tb, *shape)
    ok, worst = _within_bound(atomic.cpu().numpy(), ref, bound)
    assert ok, worst


def test_table_runs_records(dev):
    rng = np.random.default_rng(710)
    shape, M, minv, b, src, dst = _case(rng, False, n_ray=1500)
    tab = O.intersections(shape, minv, b, src, dst)
    u, st, sc = O.compress_cuda(tab, src, dst, O.voxel_diag_len(M))
    table = (_t(u.view(np.int16), dev), _t(st, dev), _t(sc, dev))
    runs = _C.table_runs(table, _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev), shape, unit_runs=50)
    rec = runs.runs[:runs.n_runs].cpu().numpy().view(np.uint32).astype(np.int64)
    units = runs.units.cpu().numpy()
    brick, n, first = units[:, 0] & 0xFFFFFFFF, units[:, 0] >> 32, units[:, 1]
    K = tab.shape[1]
    assert (rec[:, 1] >= 1).all() and (rec[:, 1] + rec[:, 2] <= K).all() and (rec[:, 2] >= 1).all()
    # units tile the run list exactly once
    cover = np.zeros(runs.n_runs, np.int64)
    for f, m in zip(first, n):
        cover[f:f + m] += 1
    assert (cover == 1).all() and (n <= 50).all() and (np.diff(n) <= 0).all()
    n_brick = ((shape[0] + 31) // 32) * ((shape[1] + 31) // 32) * ((shape[2] + 15) // 16)
    assert (brick < n_brick).all()
    # a ray's runs never overlap
    order = np.lexsort((rec[:, 1], rec[:, 0]))
    r_, s0, ns = rec[order, 0], rec[order, 1], rec[order, 2]
    same = r_[1:] == r_[:-1]
    assert (s0[1:][same] >= (s0 + ns)[:-1][same]).all()


def test_table_brick_accumulate_and_cache(dev):
    rng = np.random.default_rng(720)
    shape, M, minv, b, src, dst = _case(rng, False, n_ray=1200)
    tab = O.intersections(shape, minv, b, src, dst)
    table = _t(tab, dev)
    ts, td, tm, tb = _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev)
    y = _t(rng.standard_normal((1, src.shape[0])).astype(np.float32), dev)
    runs = _C.table_runs(table, ts, td, tm, tb, shape)
    base = torch.full((1,) + shape, 2.5, device=dev)
    got = _C.back_project_3d_table_brick(y, table, ts, td, runs, tm, tb, *shape, out=base.clone(), accumulate=True)
    plain = _C.back_project_3d_table_brick(y, table, ts, td, runs, tm, tb, *shape)
    assert torch.allclose(got, plain + 2.5, rtol=0, atol=1e-5)
    # the cache: same object for the same tensors; rebuilt after an in-place write to the table
    old = _C._MIN_TABLE_ENTRIES
    _C._MIN_TABLE_ENTRIES = 1
    try:
        r1 = _C._truns_of(table, ts, td, tm, tb, shape)
        assert r1 is _C._truns_of(table, ts, td, tm, tb, shape)
        table.add_(0.0)
        # a lookup alone (as the table forwards make it) builds nothing for the written table
        assert _C._truns_of(table, ts, td, tm, tb, shape, build=False) is None
        r2 = _C._truns_of(table, ts, td, tm, tb, shape)
        assert r2 is not r1 and _C._truns_of(table, ts, td, tm, tb, shape, build=False) is r2
    finally:
        _C._MIN_TABLE_ENTRIES = old
    with pytest.raises(RuntimeError):
        _C.back_project_3d_table_brick(y, table, ts, td, runs, tm, tb, shape[0] + 1, shape[1], shape[2])
    # lists that would not fit the given memory are not built (the callers then use the atomic kernels)
    assert _C.table_runs(table, ts, td, tm, tb, shape, max_bytes=48 * runs.n_runs - 1) is None
    assert _C.table_runs(table, ts, td, tm, tb, shape, max_bytes=48 * runs.n_runs) is not None


@pytest.mark.parametrize("rotated", [False, True])
@pytest.mark.parametrize("kind", ["f32", "u16"])
def test_ranged_forward_equals_full_walk(dev, kind, rotated):
    """The forward over each ray's visited range (from the run lists) equals the
    forward that walks whole rows bit for bit, and the oracle's table forward."""
    rng = np.random.default_rng(740 + 2 * rotated + (kind == "u16"))
    shape, M, minv, b, src, dst = _case(rng, rotated)
    tab = O.intersections(shape, minv, b, src, dst)
    ts, td, tm, tb = _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev)
    vol = rng.standard_normal((2,) + shape).astype(np.float32)
    tv = _t(vol, dev)
    if kind == "u16":
        u, st, sc = O.compress_cuda(tab, src, dst, O.voxel_diag_len(M))
        table = (_t(u.view(np.int16), dev), _t(st, dev), _t(sc, dev))
        ref = O.forward_u16(vol, u, st, sc, minv, b, src, dst)
        bound = O.forward_u16(np.abs(vol), u, st, sc, minv, b, src, dst)
    else:
        table = _t(tab, dev)
        ref = O.forward_tvals(vol, tab, minv, b, src, dst)
        bound = O.forward_tvals(np.abs(vol), tab, minv, b, src, dst)
    runs = _C.table_runs(table, ts, td, tm, tb, shape)
    first, end, _ = runs.ray_ranges()
    assert int((end > first).sum()) > 0 and int(end.max()) <= tab.shape[1]
    got = _C.forward_project_3d_table_ranged(tv, table, ts, td, runs, tm, tb)
    _C.AUTO_LAYOUT = False  # the whole-row kernels
    try:
        if kind == "u16":
            full = _C.forward_project_3d_compressed_cuda(tv, *table, ts, td, tm, tb)
        else:
            full = _C.forward_project_3d_cuda(tv, table, ts, td, tm, tb)
    finally:
        _C.AUTO_LAYOUT = True
    assert torch.equal(got, full)
    # signed values cancel along a ray: the rounding bound scales with the integral of |vol|
    ok, worst = _within_bound(got.cpu().numpy(), ref, bound)
    assert ok, worst


@pytest.fixture(scope="module")
def ring512(dev):
    import bench
    from ct_laboratory_amd.tomography.ct_projector_3d_hip import compress_tvals_3d_hip

    print("[table_brick] building the 2,949,120-ray ring and its res512 uint16 table", flush=True)
    src, dst = bench.staticct_ring_rays(dev)
    M, b, nz = bench.ring_grid(512, dev)
    shape = (512, 512, nz)
    minv = _C.cached_inverse(M)
    n_ray = int(src.shape[0])
    K = 512 + 512 + nz + 3
    u = torch.empty((n_ray, K), dtype=torch.int16, device=dev)
    st = torch.empty(n_ray, device=dev)
    sc = torch.empty(n_ray, device=dev)
    step = 262144
    for r0 in range(0, n_ray, step):
        r1 = min(n_ray, r0 + step)
        t = _C.compute_intersections_3d(512, 512, nz, src[r0:r1].contiguous(), dst[r0:r1].contiguous(), minv, b,
                                        K=K)
        a, s_, c_ = compress_tvals_3d_hip(t, M, src[r0:r1].contiguous(), dst[r0:r1].contiguous(), semantics="cuda")
        u[r0:r1], st[r0:r1], sc[r0:r1] = a, s_, c_
        del t
    g = dict(src=src, dst=dst, minv=minv, b=b, shape=shape, table=(u, st, sc))
    yield g
    del g
    torch.cuda.empty_cache()


def test_ring512_brick_vs_atomic_dense(dev, ring512):
    g = ring512
    n_ray = int(g["src"].shape[0])
    y = torch.randn((1, n_ray), device=dev, generator=torch.Generator(device=dev).manual_seed(3))
    print("[table_brick] ring res512: brick vs atomic on a dense sinogram", flush=True)
    runs = _C.table_runs(g["table"], g["src"], g["dst"], g["minv"], g["b"], g["shape"])
    got = _C.back_project_3d_table_brick(y, g["table"], g["src"], g["dst"], runs, g["minv"], g["b"], *g["shape"])
    _C.AUTO_LAYOUT = False
    try:
        atomic = _C.back_project_3d_compressed_cuda(y, *g["table"], g["src"], g["dst"], g["minv"], g["b"], *g["shape"])
        bound = _C.back_project_3d_compressed_cuda(y.abs(), *g["table"], g["src"], g["dst"], g["minv"], g["b"],
                                                   *g["shape"])
    finally:
        _C.AUTO_LAYOUT = True
    # both sides are float32 sums: allow the rounding of each (2e-5 of the magnitude sum)
    err = (got.double() - atomic.double()).abs()
    assert bool((err <= 2e-5 * bound.double() + 1e-30).all()), float((err / bound.double().clamp_min(1e-30)).max())


def test_ring512_ranged_forward_equals_full_walk(dev, ring512):
    g = ring512
    vol = torch.randn(g["shape"], device=dev, generator=torch.Generator(device=dev).manual_seed(4))
    print("[table_brick] ring res512: ranged forward vs whole-row forward", flush=True)
    runs = _C.table_runs(g["table"], g["src"], g["dst"], g["minv"], g["b"], g["shape"])
    got = _C.forward_project_3d_table_ranged(vol, g["table"], g["src"], g["dst"], runs, g["minv"], g["b"])
    _C.AUTO_LAYOUT = False
    try:
        full = _C.forward_project_3d_compressed_cuda(vol, *g["table"], g["src"], g["dst"], g["minv"], g["b"])
    finally:
        _C.AUTO_LAYOUT = True
    assert torch.equal(got, full)


def test_ring512_brick_vs_oracle_sample(dev, ring512):
    g = ring512
    n_ray = int(g["src"].shape[0])
    rng = np.random.default_rng(33)
    pick = np.sort(rng.choice(n_ray, 3000, replace=False))
    ys = rng.standard_normal(3000).astype(np.float32)
    y = torch.zeros((1, n_ray), device=dev)
    y[0, torch.from_numpy(pick).to(dev)] = torch.from_numpy(ys).to(dev)
    print("[table_brick] ring res512: brick vs oracle on 3000 sampled rays", flush=True)
    got = _C.back_project_3d_compressed_cuda(y, *g["table"], g["src"], g["dst"], g["minv"], g["b"], *g["shape"])
    tp = torch.from_numpy(pick).to(dev)
    u, st, sc = (t[tp].cpu().numpy() for t in g["table"])
    s, d = g["src"].cpu().numpy()[pick], g["dst"].cpu().numpy()[pick]
    minv, b = g["minv"].cpu().numpy(), g["b"].cpu().numpy()
    ref = O.back_u16(ys[None], u, st, sc, g["shape"], minv, b, s, d)
    bound = O.back_u16(np.abs(ys)[None], u, st, sc, g["shape"], minv, b, s, d)
    ok, worst = _within_bound(got.cpu().numpy(), ref, bound)
    assert ok, worst


@pytest.mark.parametrize("compress", [True, False])
def test_module_back_project_uses_table_runs(dev, compress, monkeypatch):
    """CTProjector3DModule with a table set builds the run lists once and its
    back_project equals the oracle's table back projection; after the map is
    written in place the stale lists are not used."""
    from ct_laboratory_amd.tomography.ct_projector_3d_module import CTProjector3DModule, precompute_tvals_stitched

    monkeypatch.setattr(_C, "_MIN_TABLE_ENTRIES", 1)
    rng = np.random.default_rng(730 + compress)
    shape, M, minv, b, src, dst = _case(rng, False, n_ray=1500)
    tm, tb, ts, td = _t(M, dev), _t(b, dev), _t(src, dev), _t(dst, dev)
    tv = precompute_tvals_stitched(*shape, tm.cpu(), tb.cpu(), ts.cpu(), td.cpu(), chunk_size=700, device=dev,
                                   verbose=False, use_compression=compress)
    proj = CTProjector3DModule(*shape, tm, tb, ts, td, device=dev, tvals=tv, use_compression=compress)
    assert proj._runs is not None
    y = rng.standard_normal(src.shape[0]).astype(np.float32)
    got = proj.back_project(_t(y, dev)).cpu().numpy()
    minv_dev = _C.cached_inverse(tm).cpu().numpy()
    if compress:
        u, st, sc = (t.numpy() for t in tv)
        ref = O.back_u16(y, u, st, sc, shape, minv_dev, b, src, dst)
        bound = O.back_u16(np.abs(y), u, st, sc, shape, minv_dev, b, src, dst)
    else:
        tab = tv.numpy()
        ref = O.back_tvals(y, tab, shape, minv_dev, b, src, dst)
        bound = O.back_tvals(np.abs(y), tab, shape, minv_dev, b, src, dst)
    ok, worst = _within_bound(got, ref, bound)
    assert ok, worst
    runs = proj._runs
    proj.M.mul_(1.01)  # the map changes in place: the module's lists no longer match
    minv2 = _C.cached_inverse(proj.M)
    assert not runs.matches(proj._table(), proj.src, proj.dst, minv2, proj.b)
    again = proj.back_project(_t(y, dev)).cpu().numpy()
    m2 = minv2.cpu().numpy()
    if compress:
        ref = O.back_u16(y, u, st, sc, shape, m2, b, src, dst)
        bound = O.back_u16(np.abs(y), u, st, sc, shape, m2, b, src, dst)
    else:
        ref = O.back_tvals(y, tab, shape, m2, b, src, dst)
        bound = O.back_tvals(np.abs(y), tab, shape, m2, b, src, dst)
    ok, worst = _within_bound(again, ref, bound)
    assert ok, worst
