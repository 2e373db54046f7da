"""GPU parity of the brick back projection over filed unit lists.

The filed lists hold, per (brick, half of the views) unit, the rays of the
unit's shadow rectangles that pass the shadow kernel's slab test.  The traced
rays, their segments and the per-brick fixed-point scale are those of the
shadow kernel, so the two results must be equal bit for bit whenever the
fixed-point path runs; both are checked against the oracle within 1e-5.
"""
import numpy as np
import pytest
import torch

import oracle as O
from ct_laboratory_amd import _C
from ct_laboratory_amd import _native as N
from ct_laboratory_amd.tomography.conebeam_projector_3d import build_conebeam_3d_geometry

pytestmark = pytest.mark.gpu

MODES = [("torch", N.MODE_TORCH), ("otf", N.MODE_OTF)]


def _t(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)


def _setup(dev, shape, n_view, cols, rows, sid, sdd, spacing, scale=1.0, rotate=False):
    src, dst = build_conebeam_3d_geometry(n_view, cols, rows, sid, sdd, det_spacing=spacing)
    M = np.diag([scale] * 3).astype(np.float64)
    if rotate:
        c, s = np.cos(0.3), np.sin(0.3)
        M = M @ np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    minv = np.linalg.inv(M).astype(np.float32)
    b = np.array([-(n - 1) / 2 * scale for n in shape], np.float32)
    ts, td, tm, tb = src.to(dev).contiguous(), dst.to(dev).contiguous(), _t(minv, dev), _t(b, dev)
    det = (rows, cols)
    plan = _C.brick_plan(_C.cone_frames(ts, td, det, tm, tb, shape), det, shape)
    lists = _C.brick_filed_lists(ts, td, det, plan, tm, tb, shape)
    return src, dst, minv, b, ts, td, tm, tb, det, plan, lists


@pytest.mark.parametrize("mname,mode", MODES)
@pytest.mark.parametrize("rotate", [False, True])
def test_filed_equals_shadow_and_oracle(dev, mname, mode, rotate):
    """Ragged brick grid (70 x 45 x 37), two batch members, diagonal and
    rotated maps: bit-identical to the shadow kernel, within 1e-5 of the
    oracle; accumulate mode (one unit per brick) bit-identical too."""
    shape = (70, 45, 37)
    src, dst, minv, b, ts, td, tm, tb, det, plan, lists = _setup(dev, shape, 11, 40, 52, 600.0, 900.0, 2.0,
                                                                 scale=1.1, rotate=rotate)
    y = np.random.default_rng(21).standard_normal((2, src.shape[0])).astype(np.float32)
    ty = _t(y, dev)
    old = _C.back_project_3d_brick(ty, ts, td, det, plan, tm, tb, *shape, mode=mode, lists=False)
    new = _C.back_project_3d_brick_filed(ty, ts, td, det, plan, lists, tm, tb, *shape, mode=mode)
    assert torch.equal(old, new)
    ref = O.back(y, shape, minv, b, src.numpy(), dst.numpy(), mode=mname)
    bound = O.back(np.abs(y), shape, minv, b, src.numpy(), dst.numpy(), mode=mname)
    assert np.all(np.abs(new.cpu().numpy() - ref) <= 1e-5 * bound + 1e-10 * float(np.abs(y).max()))
    base = torch.randn((2,) + shape, device=dev)
    acc_old = _C.back_project_3d_brick(ty, ts, td, det, plan, tm, tb, *shape, mode=mode, out=base.clone(),
                                       accumulate=True, lists=False)
    acc_new = _C.back_project_3d_brick_filed(ty, ts, td, det, plan, lists, tm, tb, *shape, mode=mode,
                                             out=base.clone(), accumulate=True)
    assert torch.equal(acc_old, acc_new)


@pytest.mark.parametrize("mname,mode", MODES)
def test_filed_sparse_invalid_frames_and_nonfinite(dev, mname, mode):
    """Sinograms that are mostly zero (the per-wave queue skips zero rays),
    views without a valid frame (whole detectors per view), repeat
    determinism, and INF / NaN rays (float fallback: INF and NaN land in the
    voxels those rays cross, the rest matches the oracle)."""
    shape = (40, 33, 36)
    src, dst, minv, b, ts, td, tm, tb, det, plan, lists = _setup(dev, shape, 7, 30, 34, 500.0, 800.0, 2.0)
    rng = np.random.default_rng(22)
    y = rng.standard_normal((1, src.shape[0])).astype(np.float32)
    y[rng.uniform(size=y.shape) < 0.8] = 0.0
    ty = _t(y, dev)
    old = _C.back_project_3d_brick(ty, ts, td, det, plan, tm, tb, *shape, mode=mode, lists=False)
    new = _C.back_project_3d_brick_filed(ty, ts, td, det, plan, lists, tm, tb, *shape, mode=mode)
    assert torch.equal(old, new)
    assert torch.equal(new, _C.back_project_3d_brick_filed(ty, ts, td, det, plan, lists, tm, tb, *shape, mode=mode))
    ref = O.back(y, shape, minv, b, src.numpy(), dst.numpy(), mode=mname)
    bound = O.back(np.abs(y), shape, minv, b, src.numpy(), dst.numpy(), mode=mname)
    floor = 1e-10 * float(np.abs(y).max())
    # signed data: the error is bounded against A|y| (cancellation leaves max|ref| no scale)
    assert bool(np.all(np.abs(new.cpu().numpy() - ref) <= 1e-5 * bound + floor))
    # invalid frames: the plan and the lists fall back to whole detectors
    fr = _C.cone_frames(ts, td, det, tm, tb, shape)
    fr[:, 15] = 0
    plan_bad = _C.brick_plan(fr, det, shape)
    lists_bad = _C.brick_filed_lists(ts, td, det, plan_bad, tm, tb, shape)
    got = _C.back_project_3d_brick_filed(ty, ts, td, det, plan_bad, lists_bad, tm, tb, *shape, mode=mode)
    assert torch.equal(got, _C.back_project_3d_brick(ty, ts, td, det, plan_bad, tm, tb, *shape, mode=mode,
                                                     lists=False))
    assert bool(np.all(np.abs(got.cpu().numpy() - ref) <= 1e-5 * bound + floor))
    # an INF and a NaN ray
    yn = y.copy()
    k_inf, k_nan = src.shape[0] // 3, 2 * src.shape[0] // 3
    yn[0, k_inf], yn[0, k_nan] = np.inf, np.nan
    gn = _C.back_project_3d_brick_filed(_t(yn, dev), ts, td, det, plan, lists, tm, tb, *shape,
                                        mode=mode).cpu().numpy()
    one = np.zeros_like(y)
    one[0, k_inf] = 1.0
    hit_inf = O.back(one, shape, minv, b, src.numpy(), dst.numpy(), mode=mname)[0] > 0
    one[0, k_inf], one[0, k_nan] = 0.0, 1.0
    hit_nan = O.back(one, shape, minv, b, src.numpy(), dst.numpy(), mode=mname)[0] > 0
    assert hit_inf.any() and hit_nan.any()
    assert np.all(np.isnan(gn[0][hit_nan]))
    assert np.all(np.isposinf(gn[0][hit_inf & ~hit_nan]))
    clean = ~(hit_inf | hit_nan)
    ref_n = O.back(np.where(np.isfinite(yn), yn, 0.0).astype(np.float32), shape, minv, b, src.numpy(), dst.numpy(),
                   mode=mname)
    bound_n = O.back(np.abs(np.where(np.isfinite(yn), yn, 0.0)).astype(np.float32), shape, minv, b, src.numpy(),
                     dst.numpy(), mode=mname)
    assert bool(np.all(np.abs(gn[0][clean] - ref_n[0][clean]) <= 1e-5 * bound_n[0][clean] + 1e-6))


def test_filed_lists_cover_the_traced_rays(dev):
    """The lists hold, per unit, distinct rays of that unit's views; every ray
    with a segment in a brick is listed in one of the brick's two units; zero
    values everywhere give a zero volume; lists of another plan are refused."""
    shape = (64, 64, 24)
    src, dst, minv, b, ts, td, tm, tb, det, plan, lists = _setup(dev, shape, 8, 48, 20, 400.0, 700.0, 2.0)
    n_ray = src.shape[0]
    per_view = det[0] * det[1]
    off = lists.offsets.cpu().numpy()
    ids = lists.ids.cpu().numpy().astype(np.int64) & 0xFFFFFFFF
    nbk = int(N.load().ctlab_brick_count_3d(*shape))
    n_view = n_ray // per_view
    S = int(N.load().ctlab_brick_filed_slots_3d(*shape, n_view)) // (2 * nbk)  # slots per unit
    K = 64  # length classes per view chunk
    assert S % K == 0 and off.shape[0] == nbk * 2 * S + 1 and off[-1] <= ids.shape[0]
    covered = {}
    for bk in range(nbk):
        for part in range(2):
            u = (bk * 2 + part) * S
            rays = ids[off[u]:off[u + S]]
            assert len(np.unique(rays)) == len(rays)
            views = rays // per_view
            assert np.all((views >= part * n_view // 2) & (views < (part + 1) * n_view // 2))
            # view chunks in order: the views of a unit's list never decrease from chunk to chunk
            chunk_first = [ids[off[u + c]:off[u + c + K]] // per_view for c in range(0, S, K)]
            chunk_views = [(v.min(), v.max()) for v in chunk_first if v.size]
            assert all(a[1] < b[0] for a, b in zip(chunk_views, chunk_views[1:]))
            covered.setdefault(bk, set()).update(rays.tolist())
    # every ray with a segment in brick bk is in bk's lists (brick of each segment from the oracle)
    idx, ln, cnt = O.trace(shape, minv, b, src.numpy(), dst.numpy(), mode="torch")
    keep = (np.arange(idx.shape[1])[None, :] < cnt[:, None]) & (ln > 0)
    i = idx[keep]
    r = np.broadcast_to(np.arange(n_ray)[:, None], idx.shape)[keep]
    x, yy, z = i // (shape[1] * shape[2]), (i // shape[2]) % shape[1], i % shape[2]
    nb = [(shape[0] + 31) // 32, (shape[1] + 31) // 32, (shape[2] + 7) // 8]
    brick = ((x // 32) * nb[1] + yy // 32) * nb[2] + z // 8
    for bk in np.unique(brick):
        need = set(np.unique(r[brick == bk]).tolist())
        assert need <= covered[int(bk)]
    zero = _C.back_project_3d_brick_filed(torch.zeros((1, n_ray), device=dev), ts, td, det, plan, lists, tm, tb,
                                          *shape)
    assert float(zero.abs().max()) == 0.0
    plan2 = plan.clone()
    with pytest.raises(RuntimeError):
        _C.back_project_3d_brick_filed(torch.zeros((1, n_ray), device=dev), ts, td, det, plan2, lists, tm, tb, *shape)


def test_registered_lists_are_used_and_dropped_on_change(dev):
    """back_project_3d_brick takes the lists registered for its plan (same
    result, bit for bit), and stops using them once the rays change in place
    (the shadow kernel then traces the new rays with the plan as given)."""
    shape = (48, 40, 30)
    src, dst, minv, b, ts, td, tm, tb, det, plan, lists = _setup(dev, shape, 6, 36, 24, 500.0, 800.0, 2.0)
    assert _C.filed_lists_of(plan, ts, td, det, shape) is lists
    y = _t(np.random.default_rng(23).standard_normal((1, src.shape[0])), dev)
    auto = _C.back_project_3d_brick(y, ts, td, det, plan, tm, tb, *shape)
    assert torch.equal(auto, _C.back_project_3d_brick_filed(y, ts, td, det, plan, lists, tm, tb, *shape))
    assert torch.equal(auto, _C.back_project_3d_brick(y, ts, td, det, plan, tm, tb, *shape, lists=False))
    td.add_(0.0)  # an in-place write (same values, new version)
    assert _C.filed_lists_of(plan, ts, td, det, shape) is None
    with pytest.raises(RuntimeError):
        _C.back_project_3d_brick_filed(y, ts, td, det, plan, lists, tm, tb, *shape)
    assert torch.equal(auto, _C.back_project_3d_brick(y, ts, td, det, plan, tm, tb, *shape))
