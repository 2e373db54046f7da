"""GPU parity of the brick back projection of rays in any order (per-brick ray
lists, ctlab_brick_bins_* and ctlab_back_project_3d_brick_bins) against the
oracle and against the one-ray-per-lane kernel: random rays (far, inside,
half-integer lattice, distant) on multi-brick grids with rotated maps, both
rules, batches, accumulation, INF/NaN values, a grid with more bricks than the
binning's LDS counters hold, and the _C path for a ray set without a
flat-panel layout.

Bar: within 1e-5 of the oracle relative to the back projection of |y|, voxel by
voxel (signed data cancel)."""
import numpy as np
import pytest
import torch

import oracle as O
from ct_laboratory_amd import _C
from ct_laboratory_amd import _native as N
from helpers import random_rays

pytestmark = pytest.mark.gpu

MODES = [("torch", N.MODE_TORCH), ("otf", N.MODE_OTF)]


def _t(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)


def _geometry(rng, shape, rotated):
    s = rng.uniform(0.5, 1.5, 3).astype(np.float32)
    M = np.diag(s).astype(np.float64)
    if rotated:
        th = rng.uniform(0, np.pi)
        R = np.eye(3)
        R[0, 0] = R[1, 1] = np.cos(th)
        R[0, 1], R[1, 0] = -np.sin(th), np.sin(th)
        M = R @ M
    M = M.astype(np.float32)
    b = (-(np.array(shape) - 1) / 2 * s + rng.uniform(-2, 2, 3)).astype(np.float32)
    minv = np.linalg.inv(M.astype(np.float64)).astype(np.float32)
    return M, minv, b, s


def _check(dev, shape, minv, b, src, dst, y, mname, mode, **kw):
    ts, td = _t(src, dev), _t(dst, dev)
    tm, tb = _t(minv, dev), _t(b, dev)
    bins = _C.brick_bins(ts, td, tm, tb, shape, mode=mode)
    got = _C.back_project_3d_brick_bins(_t(y, dev), ts, td, bins, tm, tb, *shape, mode=mode, **kw).cpu().numpy()
    ref = O.back(y, shape, minv, b, src, dst, mode=mname)
    bound = O.back(np.abs(y), shape, minv, b, src, dst, mode=mname)
    return got, ref, bound


@pytest.mark.parametrize("mname,mode", MODES)
@pytest.mark.parametrize("kind", ["far", "inside", "lattice", "distant"])
def test_bins_back_projection_matches_oracle(dev, mname, mode, kind):
    rng = np.random.default_rng(200 + 10 * mode + ["far", "inside", "lattice", "distant"].index(kind))
    for trial, shape in enumerate([(70, 45, 37), (33, 64, 20), (9, 8, 100)]):
        M, minv, b, s = _geometry(rng, shape, rotated=(trial == 1))
        src, dst = random_rays(rng, shape, s, 4000, kind)
        y = rng.standard_normal((1, src.shape[0])).astype(np.float32)
        got, ref, bound = _check(dev, shape, minv, b, src, dst, y, mname, mode)
        err = np.abs(got.astype(np.float64) - ref)
        assert np.all(err <= 1e-5 * bound + 1e-30), (trial, float((err / np.maximum(bound, 1e-30)).max()))
        assert bound.max() > 0


@pytest.mark.parametrize("mname,mode", MODES)
def test_bins_batch_accumulate_and_nonfinite(dev, mname, mode):
    rng = np.random.default_rng(230 + mode)
    shape = (40, 50, 30)
    M, minv, b, s = _geometry(rng, shape, rotated=True)
    src, dst = random_rays(rng, shape, s, 6000, "far")
    y = rng.uniform(0, 1, (2, src.shape[0])).astype(np.float32)
    ts, td, tm, tb = _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev)
    bins = _C.brick_bins(ts, td, tm, tb, shape, mode=mode)
    ref = O.back(y, shape, minv, b, src, dst, mode=mname)
    got = _C.back_project_3d_brick_bins(_t(y, dev), ts, td, bins, tm, tb, *shape, mode=mode).cpu().numpy()
    scale = np.abs(ref).max()
    assert np.all(np.abs(got - ref) <= 1e-5 * (np.abs(ref) + 1e-3 * scale))
    base = rng.uniform(0, 1, (2,) + shape).astype(np.float32)
    acc = _C.back_project_3d_brick_bins(_t(y, dev), ts, td, bins, tm, tb, *shape, mode=mode, out=_t(base, dev),
                                        accumulate=True).cpu().numpy()
    assert np.all(np.abs(acc - (ref + base)) <= 1e-5 * (np.abs(ref + base) + 1e-3 * scale))
    # an INF and a NaN ray in batch element 0: their voxels carry INF / NaN, the rest is unchanged
    yn = y.copy()
    k_inf, k_nan = 100, 200
    yn[0, k_inf], yn[0, k_nan] = np.inf, np.nan
    gn = _C.back_project_3d_brick_bins(_t(yn, dev), ts, td, bins, tm, tb, *shape, mode=mode).cpu().numpy()
    one = np.zeros_like(y[:1])
    one[0, k_inf] = 1.0
    hit_inf = O.back(one, shape, minv, b, src, dst, mode=mname)[0] > 0
    one[0, k_inf], one[0, k_nan] = 0.0, 1.0
    hit_nan = O.back(one, shape, minv, b, src, dst, mode=mname)[0] > 0
    assert np.all(np.isnan(gn[0][hit_nan])) and np.all(np.isposinf(gn[0][hit_inf & ~hit_nan]))
    clean = ~(hit_inf | hit_nan)
    assert np.all(np.abs(gn[0][clean] - ref[0][clean]) <= 1e-5 * (np.abs(ref[0][clean]) + 1e-3 * scale))
    assert np.all(np.abs(gn[1] - ref[1]) <= 1e-5 * (np.abs(ref[1]) + 1e-3 * scale))


def test_bins_grid_beyond_lds_counters(dev):
    """16,388 bricks (more than the 16,384 the binning counts in LDS): the
    global-counter binning path."""
    rng = np.random.default_rng(240)
    shape = (33, 33, 32776)
    minv, b = np.eye(3, dtype=np.float32), np.array([-16.0, -16.0, -16387.5], np.float32)
    n = 3000
    src = np.stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n), rng.uniform(-17000, 17000, n)], 1)
    dst = np.stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n), rng.uniform(-17000, 17000, n)], 1)
    src, dst = src.astype(np.float32), dst.astype(np.float32)
    assert _C.N.load().ctlab_brick_count_3d(*shape) > 16384
    y = rng.uniform(0, 1, (1, n)).astype(np.float32)
    got, ref, bound = _check(dev, shape, minv, b, src, dst, y, "torch", N.MODE_TORCH)
    assert np.all(np.abs(got.astype(np.float64) - ref) <= 1e-5 * bound + 1e-30)
    assert bound.max() > 0


@pytest.mark.parametrize("mname,mode", MODES)
def test_c_path_lists_unstructured_rays(dev, mname, mode):
    """_C.back_project_3d_on_the_fly_cuda on a ray set with no flat-panel layout
    (a cone-beam scan in shuffled order) builds brick lists once and matches
    the one-ray-per-lane kernel and the oracle."""
    from ct_laboratory_amd.tomography.conebeam_projector_3d import build_conebeam_3d_geometry

    shape = (64, 48, 40)
    src, dst = build_conebeam_3d_geometry(12, 48, 32, 600.0, 900.0, det_spacing=2.0)
    perm = np.random.default_rng(250).permutation(src.shape[0])
    src, dst = src.numpy()[perm], dst.numpy()[perm]
    b = np.array([-(n - 1) / 2 for n in shape], np.float32)
    ts, td, tm, tb = _t(src, dev), _t(dst, dev), _t(np.eye(3), dev), _t(b, dev)
    y = _t(np.random.default_rng(251).uniform(0, 1, (1, src.shape[0])), dev)
    got = _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, *shape, mode=mode)
    assert _C._layout_of(ts, td) is None
    key = [k for k in _C._BINS if k[0] == _C._ray_key(ts, td)]
    assert key, "no brick lists were built for the ray set"
    _C.AUTO_LAYOUT = False
    try:
        atomic = _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, *shape, mode=mode)
    finally:
        _C.AUTO_LAYOUT = True
    ref = O.back(y.cpu().numpy(), shape, np.eye(3, dtype=np.float32), b, src, dst, mode=mname)
    scale = np.abs(ref).max()
    for arr in (got.cpu().numpy(), atomic.cpu().numpy()):
        assert np.all(np.abs(arr - ref) <= 1e-5 * (np.abs(ref) + 1e-3 * scale))


@pytest.mark.parametrize("mname,mode", MODES)
def test_c_path_orders_large_unstructured_sets(dev, mname, mode):
    """A cone-beam scan of 1,228,800 rays in shuffled order: the _C path sorts it
    once into spatial order (ray_order); the forward equals the storage-order
    forward ray for ray (bit for bit), the back projection equals the brick
    lists over the shuffled order to float rounding (each brick sums in fixed
    point; a brick split over several blocks adds its parts' float sums), and
    the oracle on a sample of rays."""
    from ct_laboratory_amd.tomography.conebeam_projector_3d import build_conebeam_3d_geometry

    shape = (96, 96, 64)
    src, dst = build_conebeam_3d_geometry(120, 128, 80, 600.0, 900.0, det_spacing=1.2)
    n = src.shape[0]
    assert n >= _C._MIN_ORDER_RAYS
    perm = torch.from_numpy(np.random.default_rng(260).permutation(n))
    ts, td = src[perm].contiguous().to(dev), dst[perm].contiguous().to(dev)
    b = np.array([-(k - 1) / 2 for k in shape], np.float32)
    tm, tb = _t(np.eye(3), dev), _t(b, dev)
    print(f"[bins] ordered path, {mname}: {n} shuffled rays on {shape}", flush=True)
    order = _C.ray_order(ts, td)
    assert torch.equal(torch.sort(order.perm).values, torch.arange(n, device=dev))
    assert torch.equal(order.src, ts[order.perm]) and torch.equal(order.dst, td[order.perm])
    vol = torch.rand(shape, device=dev, generator=torch.Generator(device=dev).manual_seed(261))
    fwd = _C.forward_project_3d_on_the_fly_cuda(vol, ts, td, tm, tb, mode=mode)
    assert _C._layout_of(ts, td) is None and _C._order_of(ts, td) is not None
    plain = _C._forward_otf(vol, ts, td, tm, tb, mode, None, None)
    assert torch.equal(fwd, plain)
    y = torch.randn((1, n), device=dev, generator=torch.Generator(device=dev).manual_seed(262))
    got = _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, *shape, mode=mode)
    bins = _C.brick_bins(ts, td, tm, tb, shape, mode=mode)
    unordered = _C.back_project_3d_brick_bins(y, ts, td, bins, tm, tb, *shape, mode=mode)
    mag = _C.back_project_3d_brick_bins(y.abs(), ts, td, bins, tm, tb, *shape, mode=mode)
    diff = (got.double() - unordered.double()).abs()
    assert bool((diff <= 1e-6 * mag.double() + 1e-30).all()), float((diff / mag.double().clamp_min(1e-30)).max())
    pick = torch.from_numpy(np.sort(np.random.default_rng(263).choice(n, 4000, replace=False))).to(dev)
    ys = torch.zeros_like(y)
    ys[0, pick] = y[0, pick]
    got_s = _C.back_project_3d_on_the_fly_cuda(ys, ts, td, tm, tb, *shape, mode=mode).cpu().numpy()
    s_np, d_np, y_np = ts[pick].cpu().numpy(), td[pick].cpu().numpy(), ys[0, pick].cpu().numpy()[None]
    ref = O.back(y_np, shape, np.eye(3, dtype=np.float32), b, s_np, d_np, mode=mname)
    bound = O.back(np.abs(y_np), shape, np.eye(3, dtype=np.float32), b, s_np, d_np, mode=mname)
    err = np.abs(got_s.astype(np.float64) - ref)
    assert np.all(err <= 1e-5 * bound + 1e-30), float((err / np.maximum(bound, 1e-30)).max())
    ref_f = O.forward(vol.cpu().numpy(), np.eye(3, dtype=np.float32), b, s_np, d_np, mode=mname)
    f_np = fwd[0, pick].cpu().numpy().astype(np.float64)
    assert np.all(np.abs(f_np - ref_f) <= 1e-5 * np.maximum(np.abs(ref_f), 1e-3 * np.abs(ref_f).max()))
