"""GPU parity of the on-the-fly projectors (2D and 3D) against the oracle.

Bar: per-ray segment lists (voxel index of every segment with length > 0)
bit-exact, lengths within 3 ulp; line integrals and back projections within 1e-5
relative (floor 1e-3 of the max); adjoint |<Ax,y>-<x,A^T y>|/|<Ax,y>| < 1e-5.
"""
import math

import numpy as np
import pytest
import torch

import oracle as O
from ct_laboratory_amd import _C
from ct_laboratory_amd import _native as N
from ct_laboratory_amd.tomography.conebeam_projector_3d import build_conebeam_3d_geometry
from helpers import lattice_rays, positive_lists, random_geometry, random_rays, rel_err

pytestmark = pytest.mark.gpu

MODES = [("torch", N.MODE_TORCH), ("otf", N.MODE_OTF)]


def _t(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)


def _gpu_lists(shape, minv, b, src, dst, mode, dev, zero_rule=None):
    idx, ln, cnt = _C.trace_segments(shape, _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev), mode=mode,
                                     zero_rule=zero_rule)
    return positive_lists(idx.cpu().numpy(), ln.cpu().numpy(), cnt.cpu().numpy())


def _assert_lists_equal(shape, minv, b, src, dst, mname, mode, dev, zero_rule=None):
    nd = len(shape)
    zr = zero_rule if zero_rule is not None else (N.ZERO_EXACT if (nd == 2 and mname == "torch") else N.ZERO_THRESH)
    g = _gpu_lists(shape, minv, b, src, dst, mode, dev, zero_rule=zr)
    oi, ol, oc = O.trace(shape, minv, b, src, dst, mode=mname, zero_rule="exact" if zr == N.ZERO_EXACT else "thresh")
    o = positive_lists(oi, ol, oc)
    # Voxel indices bit-exact; lengths within 3 ulp: the device sums the squared
    # differences with fmas (up to 2 ulp of the length from the torch rule's
    # mul-then-add order) and takes v_sqrt_f32 (1 ulp); the oracle keeps the
    # torch order and a correctly rounded sqrtf.
    bad = []
    for r in range(len(o)):
        gi, gl = g[r]
        ri, rl = o[r]
        if len(gi) != len(ri) or (gi != ri).any() or (np.abs(gl.view(np.int32) - rl.view(np.int32)) > 3).any():
            bad.append(r)
    assert not bad, f"{len(bad)} rays differ, first {bad[:5]}: gpu {g[bad[0]][0][:8]} oracle {o[bad[0]][0][:8]}"


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("mname,mode", MODES)
@pytest.mark.parametrize("kind", ["far", "inside", "lattice", "distant"])
def test_segment_lists_bit_exact_random(dev, nd, mname, mode, kind):
    rng = np.random.default_rng(1000 + nd * 10 + mode)
    for trial in range(4):
        shape, M, minv, b, s = random_geometry(rng, nd, rotated=(trial == 3))
        src, dst = random_rays(rng, shape, s, 3000, kind)
        _assert_lists_equal(shape, minv, b, src, dst, mname, mode, dev)


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("mname,mode", MODES)
def test_segment_lists_bit_exact_degenerate(dev, nd, mname, mode):
    """Rays on voxel planes, through corners and axis-parallel (H6/H7 ties)."""
    rng = np.random.default_rng(7 + nd)
    shape = (9, 9, 9) if nd == 3 else (13, 11)
    src, dst = lattice_rays(rng, nd, 20000)
    eye = np.eye(nd, dtype=np.float32)
    _assert_lists_equal(shape, eye, np.zeros(nd, np.float32), src, dst, mname, mode, dev)


@pytest.mark.parametrize("mname,mode", MODES)
def test_segment_lists_conebeam(dev, mname, mode):
    n = 64
    src, dst = build_conebeam_3d_geometry(24, 48, 40, 1000.0, 1500.0, det_spacing=6.0)
    b = np.full(3, -(n - 1) / 2 * 1.0, np.float32)
    _assert_lists_equal((n, n, n), np.eye(3, dtype=np.float32), b, src.numpy(), dst.numpy(), mname, mode, dev)


def test_closed_forms(dev):
    """Known answers from the reference's python_examples (as asserts)."""
    img = np.zeros((32, 32), np.float32)
    img[16, :] = 1
    e2, z2 = _t(np.eye(2), dev), _t(np.zeros(2), dev)
    for mode in (N.MODE_TORCH, N.MODE_OTF, N.MODE_MARCH):
        out = _C.forward_project_2d_on_the_fly_cuda(_t(img, dev), _t([[16.0, -5.0]], dev), _t([[16.0, 37.0]], dev),
                                                    e2, z2, mode=mode).cpu().numpy()
        assert abs(out[0, 0] - 32.0) < 1e-4  # test_ct_projector_2d_line_integral.py:31-43,70
    img = np.zeros((16, 16), np.float32)
    img[8, :] = 1
    src = [[8.0, -5.0], [0, 0], [15, 15]]
    dst = [[8.0, 25.0], [15, 15], [0, 8]]
    out = _C.forward_project_2d_on_the_fly_cuda(_t(img, dev), _t(src, dev), _t(dst, dev), e2, z2,
                                                mode=N.MODE_TORCH, zero_rule=N.ZERO_EXACT).cpu().numpy()[0]
    np.testing.assert_allclose(out, [16.0, math.sqrt(2), math.sqrt(15**2 + 7**2) / 15], rtol=1e-6)
    vol = np.zeros((32, 32, 32), np.float32)
    vol[:, 16, 16] = 1
    e3, z3 = _t(np.eye(3), dev), _t(np.zeros(3), dev)
    out = _C.forward_project_3d_on_the_fly_cuda(_t(vol, dev), _t([[200.0, 16, 16]], dev), _t([[-200.0, 16, 16]], dev),
                                                e3, z3, mode=N.MODE_TORCH).cpu().numpy()
    assert abs(out[0, 0] - 32.0) < 1e-4  # test_ct_projector_3d_line_integral.py:33-45,74
    vol = np.zeros((8, 8, 8), np.float32)
    vol[4] = 1
    src = [[4.0, -5, 4], [0, 0, 0]]
    dst = [[4.0, 15, 4], [7, 7, 7]]
    for mode in (N.MODE_TORCH, N.MODE_OTF):
        out = _C.forward_project_3d_on_the_fly_cuda(_t(vol, dev), _t(src, dev), _t(dst, dev), e3, z3,
                                                    mode=mode).cpu().numpy()[0]
        np.testing.assert_allclose(out, [8.0, math.sqrt(3)], rtol=1e-6)  # test_ct_projector_3d_module.py:17-32


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("mname,mode", MODES)
def test_forward_back_vs_oracle(dev, nd, mname, mode):
    rng = np.random.default_rng(50 + nd + 3 * mode)
    shape, M, minv, b, s = random_geometry(rng, nd, max_n=40)
    src, dst = random_rays(rng, shape, s, 5000, "far")
    zr = N.ZERO_EXACT if (nd == 2 and mname == "torch") else N.ZERO_THRESH
    ozr = "exact" if zr == N.ZERO_EXACT else "thresh"
    B = 2
    ts, td, tm, tb = _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev)

    def run(vol, sino):
        tv = _t(vol, dev)
        if nd == 3:
            fwd = _C.forward_project_3d_on_the_fly_cuda(tv, ts, td, tm, tb, mode=mode)
            bk = _C.back_project_3d_on_the_fly_cuda(_t(sino, dev), ts, td, tm, tb, *shape, mode=mode)
        else:
            fwd = _C.forward_project_2d_on_the_fly_cuda(tv, ts, td, tm, tb, mode=mode, zero_rule=zr)
            bk = _C.back_project_2d_on_the_fly_cuda(_t(sino, dev), ts, td, tm, tb, *shape, mode=mode,
                                                    zero_rule=zr)
        return fwd.cpu().numpy(), bk.cpu().numpy()

    # Nonnegative data (attenuation, line integrals): 1e-5 relative.
    vol = rng.uniform(0, 1, size=(B,) + shape).astype(np.float32)
    sino = rng.uniform(0, 1, size=(B, src.shape[0])).astype(np.float32)
    fwd, bk = run(vol, sino)
    assert rel_err(fwd, O.forward(vol, minv, b, src, dst, mode=mname, zero_rule=ozr)) < 1e-5
    assert rel_err(bk, O.back(sino, shape, minv, b, src, dst, mode=mname, zero_rule=ozr)) < 1e-5
    # Signed data: float summation error scales with A|x| (resp. A^T|y|), not |Ax|.
    vol = rng.normal(size=(B,) + shape).astype(np.float32)
    sino = rng.normal(size=(B, src.shape[0])).astype(np.float32)
    fwd, bk = run(vol, sino)
    ref_f = O.forward(vol, minv, b, src, dst, mode=mname, zero_rule=ozr)
    ref_b = O.back(sino, shape, minv, b, src, dst, mode=mname, zero_rule=ozr)
    cond_f = O.forward(np.abs(vol), minv, b, src, dst, mode=mname, zero_rule=ozr)
    cond_b = O.back(np.abs(sino), shape, minv, b, src, dst, mode=mname, zero_rule=ozr)
    assert (np.abs(fwd - ref_f) <= 1e-5 * cond_f + 1e-30).all()
    assert (np.abs(bk - ref_b) <= 1e-5 * cond_b + 1e-30).all()


@pytest.mark.parametrize("mname,mode", MODES)
def test_adjoint_conebeam(dev, mname, mode):
    n = 48
    src, dst = build_conebeam_3d_geometry(30, 40, 36, 1000.0, 1500.0, det_spacing=2.5)
    b = np.full(3, -(n - 1) / 2, np.float32)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(n, n, n)).astype(np.float32)
    y = rng.normal(size=src.shape[0]).astype(np.float32)
    tm, tb = _t(np.eye(3), dev), _t(b, dev)
    ts, td = src.to(dev), dst.to(dev)
    Ax = _C.forward_project_3d_on_the_fly_cuda(_t(x, dev), ts, td, tm, tb, mode=mode)[0].cpu().numpy()
    ATy = _C.back_project_3d_on_the_fly_cuda(_t(y, dev), ts, td, tm, tb, n, n, n, mode=mode)[0].cpu().numpy()
    # Dot products in FP64 on the host: the adjoint identity of the GPU pair.
    lhs = float(np.dot(Ax.astype(np.float64), y.astype(np.float64)))
    rhs = float(np.dot(x.astype(np.float64).ravel(), ATy.astype(np.float64).ravel()))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    # Same operator as the oracle's FP32-semantics projector (only summation order differs).
    ref = O.forward(x, np.eye(3), b, src.numpy(), dst.numpy(), mode=mname)
    lhs_o = float(np.dot(ref, y.astype(np.float64)))
    assert abs(lhs - lhs_o) / abs(lhs_o) < 1e-5


@pytest.mark.parametrize("mname,mode", MODES)
def test_detector_tiled_schedule(dev, mname, mode):
    """The detector-tiled lane schedule changes only which lane traces a ray."""
    n = 40
    src, dst = build_conebeam_3d_geometry(7, 37, 70, 1000.0, 1500.0, det_spacing=3.0)  # ragged tiles
    b = _t(np.full(3, -(n - 1) / 2, np.float32), dev)
    e3 = _t(np.eye(3), dev)
    rng = np.random.default_rng(5)
    vol = _t(rng.uniform(0, 1, (2, n, n, n)), dev)
    ts, td = src.to(dev), dst.to(dev)
    f0 = _C.forward_project_3d_on_the_fly_cuda(vol, ts, td, e3, b, mode=mode)
    f1 = _C.forward_project_3d_on_the_fly_cuda(vol, ts, td, e3, b, mode=mode, det_shape=(70, 37))
    assert torch.equal(f0, f1)
    y = _t(rng.uniform(0, 1, (2, src.shape[0])), dev)
    b0 = _C.back_project_3d_on_the_fly_cuda(y, ts, td, e3, b, n, n, n, mode=mode).cpu().numpy()
    b1 = _C.back_project_3d_on_the_fly_cuda(y, ts, td, e3, b, n, n, n, mode=mode, det_shape=(70, 37)).cpu().numpy()
    assert rel_err(b1, b0) < 1e-5
    with pytest.raises(RuntimeError, match="multiple of the detector size"):
        _C.forward_project_3d_on_the_fly_cuda(vol, ts, td, e3, b, det_shape=(71, 37))


@pytest.mark.parametrize("mname,mode", MODES)
def test_forward_batch_groups(dev, mname, mode):
    """A forward launch traces each ray once for a group of up to four batch
    members: every member's line integrals equal its own batch-1 projection bit
    for bit, for batches of 2, 3 and 5 (ragged last group), with and without a
    detector layout, in 3D and 2D."""
    n = 24
    src, dst = build_conebeam_3d_geometry(5, 21, 18, 200.0, 300.0, det_spacing=2.5)
    b = _t(np.full(3, -(n - 1) / 2, np.float32), dev)
    e3 = _t(np.eye(3), dev)
    ts, td = src.to(dev).contiguous(), dst.to(dev).contiguous()
    rng = np.random.default_rng(12)
    vol = _t(rng.uniform(-1, 1, (5, n, n, n)), dev)
    for det in (None, (18, 21)):
        one = torch.cat([_C.forward_project_3d_on_the_fly_cuda(vol[i], ts, td, e3, b, mode=mode, det_shape=det)
                         for i in range(5)])
        for nb in (2, 3, 5):
            got = _C.forward_project_3d_on_the_fly_cuda(vol[:nb].contiguous(), ts, td, e3, b, mode=mode,
                                                        det_shape=det)
            assert torch.equal(got, one[:nb])
    shape, M, minv, b2, s = random_geometry(rng, 2, max_n=30)
    s2, d2 = random_rays(rng, shape, s, 3000, "far")
    img = _t(rng.uniform(-1, 1, (5,) + shape), dev)
    args = (_t(s2, dev), _t(d2, dev), _t(minv, dev), _t(b2, dev))
    one = torch.cat([_C.forward_project_2d_on_the_fly_cuda(img[i], *args, mode=mode) for i in range(5)])
    for nb in (2, 3, 5):
        assert torch.equal(_C.forward_project_2d_on_the_fly_cuda(img[:nb].contiguous(), *args, mode=mode), one[:nb])


def test_empty_and_zero(dev):
    z = torch.zeros((0, 3), device=dev)
    vol = torch.ones((4, 4, 4), device=dev)
    e3, z3 = torch.eye(3, device=dev), torch.zeros(3, device=dev)
    out = _C.forward_project_3d_on_the_fly_cuda(vol, z, z, e3, z3)
    assert out.shape == (1, 0)
    bk = _C.back_project_3d_on_the_fly_cuda(torch.zeros((0,), device=dev), z, z, e3, z3, 4, 4, 4)
    assert bk.shape == (1, 4, 4, 4) and float(bk.abs().sum()) == 0.0
    # zero-length ray
    p = torch.tensor([[1.0, 1.0, 1.0]], device=dev)
    assert float(_C.forward_project_3d_on_the_fly_cuda(vol, p, p, e3, z3)[0, 0]) == 0.0


@pytest.mark.parametrize("mname,mode", MODES)
def test_brick_back_projection(dev, mname, mode):
    """The LDS-brick back projector adds the same segments as the atomic one:
    within 1e-5 of the oracle, bricks with ragged edges (n not a multiple of 32)."""
    shape = (70, 45, 37)
    src, dst = build_conebeam_3d_geometry(11, 40, 52, 600.0, 900.0, det_spacing=2.0)
    b = np.array([-(n - 1) / 2 * 1.1 for n in shape], np.float32)
    M = np.diag([1.1, 1.1, 1.1]).astype(np.float32)
    minv = np.linalg.inv(M.astype(np.float64)).astype(np.float32)
    rng = np.random.default_rng(12)
    y = rng.uniform(0, 1, (2, src.shape[0])).astype(np.float32)
    ts, td, tm, tb = src.to(dev), dst.to(dev), _t(minv, dev), _t(b, dev)
    fr = _C.cone_frames(ts, td, (52, 40), tm, tb, shape)
    assert bool((fr[:, 15] == 1).all()), "cone-beam views should fit the flat-panel model"
    plan = _C.brick_plan(fr, (52, 40), shape)
    got = _C.back_project_3d_brick(_t(y, dev), ts, td, (52, 40), plan, tm, tb, *shape, mode=mode).cpu().numpy()
    ref = O.back(y, shape, minv, b, src.numpy(), dst.numpy(), mode=mname)
    assert rel_err(got, ref) < 1e-5
    # invalid frames (whole detector per view) give the same result
    fr_bad = fr.clone()
    fr_bad[:, 15] = 0
    plan_bad = _C.brick_plan(fr_bad, (52, 40), shape)
    got2 = _C.back_project_3d_brick(_t(y, dev), ts, td, (52, 40), plan_bad, tm, tb, *shape, mode=mode).cpu().numpy()
    assert rel_err(got2, ref) < 1e-5
    # accumulate mode adds into the given volume
    base = torch.ones((2,) + shape, device=dev)
    acc = _C.back_project_3d_brick(_t(y, dev), ts, td, (52, 40), plan, tm, tb, *shape, mode=mode, out=base,
                                   accumulate=True).cpu().numpy()
    assert rel_err(acc - 1.0, ref) < 1e-4


@pytest.mark.parametrize("mname,mode", MODES)
def test_brick_signed_nonfinite_deterministic(dev, mname, mode):
    """Fixed-point brick accumulation: signed data (checked with the A|y| bound
    plus the fixed-point quantum: a voxel's absolute error is at most half a
    quantum per contribution, the quantum being below the larger of
    4 * max|y| * diag * rays / 2^62 and 4 * max|y| * diag / 2^50 of the brick,
    far under 1e-10 * max|y| here), bit-identical repeats,
    value scales near both ends of the float range, and the float fallback for
    bricks whose rays carry INF/NaN."""
    shape = (40, 33, 36)
    src, dst = build_conebeam_3d_geometry(7, 30, 34, 500.0, 800.0, det_spacing=2.0)
    b = np.array([-(n - 1) / 2 for n in shape], np.float32)
    minv = np.eye(3, dtype=np.float32)
    rng = np.random.default_rng(13)
    y = rng.standard_normal((1, src.shape[0])).astype(np.float32)
    ts, td, tm, tb = src.to(dev), dst.to(dev), _t(minv, dev), _t(b, dev)
    det = (34, 30)
    plan = _C.brick_plan(_C.cone_frames(ts, td, det, tm, tb, shape), det, shape)

    def brick(v):
        return _C.back_project_3d_brick(_t(v, dev), ts, td, det, plan, tm, tb, *shape, mode=mode).cpu().numpy()

    got = brick(y)
    ref = O.back(y, shape, minv, b, src.numpy(), dst.numpy(), mode=mname)
    bound = O.back(np.abs(y), shape, minv, b, src.numpy(), dst.numpy(), mode=mname)
    floor = 1e-10 * float(np.abs(y).max())
    assert np.all(np.abs(got - ref) <= 1e-5 * bound + floor)
    assert np.array_equal(got, brick(y)), "fixed-point accumulation must be order independent"
    for s in (1e-30, 1e30):  # power-of-two-free scales near both ends of the float range
        ys = (y * np.float32(s)).astype(np.float32)
        gs = brick(ys).astype(np.float64)
        assert np.all(np.abs(gs - ref * s) <= (1e-5 * bound + floor) * s)
    # an INF and a NaN ray: the voxels they cross carry INF / NaN, the rest is unchanged
    yn = y.copy()
    k_inf, k_nan = src.shape[0] // 3, 2 * src.shape[0] // 3
    yn[0, k_inf] = np.inf
    yn[0, k_nan] = np.nan
    gn = brick(yn)
    one = np.zeros_like(y)
    one[0, k_inf] = 1.0
    hit_inf = O.back(one, shape, minv, b, src.numpy(), dst.numpy(), mode=mname)[0] > 0
    one[0, k_inf], one[0, k_nan] = 0.0, 1.0
    hit_nan = O.back(one, shape, minv, b, src.numpy(), dst.numpy(), mode=mname)[0] > 0
    assert hit_inf.any() and hit_nan.any()
    assert np.all(np.isnan(gn[0][hit_nan]))
    assert np.all(np.isposinf(gn[0][hit_inf & ~hit_nan]))
    clean = ~(hit_inf | hit_nan)
    assert np.all(np.abs(gn[0][clean] - ref[0][clean]) <= 1e-5 * bound[0][clean] + floor)


@pytest.mark.parametrize("mname,mode", MODES)
def test_auto_layout_uses_bricks(dev, mname, mode):
    """The _C 3D on-the-fly functions called as the reference calls them (no
    detector layout) recognise a cone-beam ray stack: the back projection is the
    brick kernel's (bit-identical to an explicit brick call, within 1e-5 of the
    oracle and of the atomic kernel), the forward equals the explicitly tiled
    one, and an in-place change of the rays invalidates the cached layout."""
    shape = (50, 42, 36)
    src, dst = build_conebeam_3d_geometry(9, 44, 30, 600.0, 900.0, det_spacing=2.0)
    ts, td = src.to(dev).contiguous(), dst.to(dev).contiguous()
    b = np.array([-(n - 1) / 2 for n in shape], np.float32)
    tm, tb = _t(np.eye(3), dev), _t(b, dev)
    rng = np.random.default_rng(21)
    y = _t(rng.uniform(0, 1, (1, src.shape[0])), dev)
    auto = _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, *shape, mode=mode)
    assert _C._layout_of(ts, td) == (30, 44)
    det = (30, 44)
    plan = _C.brick_plan(_C.cone_frames(ts, td, det, tm, tb, shape), det, shape)
    assert torch.equal(auto, _C.back_project_3d_brick(y, ts, td, det, plan, tm, tb, *shape, mode=mode))
    ref = O.back(y.cpu().numpy(), shape, np.eye(3, dtype=np.float32), b, src.numpy(), dst.numpy(), mode=mname)
    assert rel_err(auto.cpu().numpy(), ref) < 1e-5
    _C.AUTO_LAYOUT = False
    try:
        atomic = _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, *shape, mode=mode)
    finally:
        _C.AUTO_LAYOUT = True
    assert rel_err(atomic.cpu().numpy(), auto.cpu().numpy()) < 1e-5
    vol = _t(rng.uniform(0, 1, (1,) + shape), dev)
    assert torch.equal(_C.forward_project_3d_on_the_fly_cuda(vol, ts, td, tm, tb, mode=mode),
                       _C.forward_project_3d_on_the_fly_cuda(vol, ts, td, tm, tb, mode=mode, det_shape=det))
    with torch.no_grad():
        ts[5] += 1.0  # the first view's source now moves: no layout, atomic kernel
    assert _C._layout_of(ts, td) is None
    moved = _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, *shape, mode=mode)
    ref2 = O.back(y.cpu().numpy(), shape, np.eye(3, dtype=np.float32), b, ts.cpu().numpy(), dst.numpy(), mode=mname)
    assert rel_err(moved.cpu().numpy(), ref2) < 1e-5


@pytest.mark.parametrize("mname,mode", MODES)
def test_auto_layout_new_rays_in_freed_block(dev, mname, mode):
    """A ray set freed after its layout and brick plan were cached, followed by a
    new ray set of the same size (another angle offset) that the caching
    allocator may place in the same blocks with the same version counters: the
    new set gets its own plan, and its back projection matches the oracle.  A
    second call with the same map tensors reads no map values from the device."""
    shape = (40, 36, 28)
    b = np.array([-(n - 1) / 2 for n in shape], np.float32)
    tm, tb = _t(np.eye(3), dev), _t(b, dev)
    rng = np.random.default_rng(23)
    src_a, dst_a = build_conebeam_3d_geometry(8, 40, 24, 600.0, 900.0, det_spacing=2.0)
    ts, td = src_a.to(dev).contiguous(), dst_a.to(dev).contiguous()
    y = _t(rng.uniform(0, 1, (1, src_a.shape[0])), dev)
    _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, *shape, mode=mode)
    assert _C._layout_of(ts, td) == (24, 40)
    ptrs_a = (ts.data_ptr(), td.data_ptr())
    del ts, td
    src_b, dst_b = build_conebeam_3d_geometry(8, 40, 24, 600.0, 900.0, det_spacing=2.0, angle_offset=0.37)
    ts, td = src_b.to(dev).contiguous(), dst_b.to(dev).contiguous()
    print("same blocks as the freed ray set:", (ts.data_ptr(), td.data_ptr()) == ptrs_a)
    got = _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, *shape, mode=mode)
    ref = O.back(y.cpu().numpy(), shape, np.eye(3, dtype=np.float32), b, src_b.numpy(), dst_b.numpy(), mode=mname)
    assert rel_err(got.cpu().numpy(), ref) < 1e-5
    # map values come from the per-tensor cache on a repeated call
    key_m = _C._MAP_VALUES.get(id(tm))
    assert key_m is not None and key_m[0]() is tm
    again = _C.back_project_3d_on_the_fly_cuda(y, ts, td, tm, tb, *shape, mode=mode)
    assert torch.equal(again, got)


def test_march_2d_matches_oracle(dev):
    """The reference's 2D on-the-fly march (the _C default for 2D on-the-fly
    calls, as in the reference's _C): every in-grid step bit-exact against the
    oracle (index and length, zero and negative lengths included), forward and
    back projections within 1e-5, and the hand-derived known answers."""
    rng = np.random.default_rng(77)
    for trial in range(8):
        shape, M, minv, b, s = random_geometry(rng, 2, rotated=(trial % 4 == 3))
        kind = ["far", "inside", "lattice", "distant"][trial % 4]
        src, dst = random_rays(rng, shape, s, 3000, kind)
        idx, ln, cnt = _C.trace_segments(shape, _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev),
                                         mode=N.MODE_MARCH)
        gi, gl, gc = idx.cpu().numpy(), ln.cpu().numpy(), cnt.cpu().numpy()
        oi, ol, oc = O.trace(shape, minv, b, src, dst, mode="march")
        assert np.array_equal(gc, oc)
        for r in range(src.shape[0]):
            k = int(oc[r])
            assert np.array_equal(gi[r, :k], oi[r, :k]), (trial, r)
            assert np.array_equal(gl[r, :k].view(np.int32), ol[r, :k].view(np.int32)), (trial, r)
        img = rng.uniform(0, 1, shape).astype(np.float32)
        got = _C.forward_project_2d_on_the_fly_cuda(_t(img, dev), _t(src, dev), _t(dst, dev), _t(minv, dev),
                                                    _t(b, dev)).cpu().numpy()[0]
        ref = O.forward(img, minv, b, src, dst, mode="march")
        assert rel_err(got, ref) < 1e-5
        y = rng.uniform(0, 1, (1, src.shape[0])).astype(np.float32)
        back = _C.back_project_2d_on_the_fly_cuda(_t(y, dev), _t(src, dev), _t(dst, dev), _t(minv, dev),
                                                  _t(b, dev), *shape).cpu().numpy()
        refb = O.back(y, shape, minv, b, src, dst, mode="march")
        assert np.all(np.abs(back - refb) <= 1e-5 * (np.abs(refb) + np.abs(refb).max()))
    e2, z2 = _t(np.eye(2), dev), _t(np.zeros(2), dev)
    one = np.zeros((8, 8), np.float32)
    one[3, 2] = 1
    got = _C.forward_project_2d_on_the_fly_cuda(_t(one, dev), _t([[3.7, 2.0]], dev), _t([[9.7, 2.0]], dev), e2, z2)
    assert abs(float(got[0, 0]) + 0.2) < 1e-6  # the negative first step of a ray starting past a voxel centre
    col5 = np.zeros((8, 8), np.float32)
    col5[:, 5] = 1
    got = _C.forward_project_2d_on_the_fly_cuda(_t(col5, dev), _t([[-2.0, 4.5]], dev), _t([[10.0, 4.5]], dev), e2, z2)
    assert abs(float(got[0, 0]) - 8.0) < 1e-5  # column coordinate 4.5 walks column 5
