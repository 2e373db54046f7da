"""GPU tests of the projector layer: autograd Functions, nn.Modules and the
fan-beam, cone-beam and StaticCT projectors, against the oracle on the same rays.

The modules invert M with torch on the device; the oracle gets that same
float32 inverse, so both sides trace identical voxel coordinates.  Bars as in
the kernel tests: 1e-5 relative for line integrals and back projections,
adjoint < 1e-5 in FP64 dot products; gradients are checked against the
back projector and by central finite differences of an FP64 loss.
"""
import math

import numpy as np
import pytest
import torch

import oracle as O
from ct_laboratory_amd.phantoms import ellipse_phantom_2d, ellipsoid_phantom_3d
from ct_laboratory_amd.tomography import (CTProjector2DFunction, CTProjector2DModule, CTProjector3DFunction,
                                          CTProjector3DModule, ConeBeam3DProjector, FanBeam2DProjector,
                                          StaticCTProjector3D, UniformStaticCTProjector2D,
                                          UniformStaticCTProjector3D, build_circular_sequence,
                                          build_conebeam_3d_geometry, build_uniform_static_3d_geometry,
                                          decompress_tvals_from_uint16, precompute_tvals_stitched,
                                          standard_image_transform_2d, standard_image_transform_3d)
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _minv(M):
    return _np(torch.inverse(M.float()))


def _fd_check(loss_fn, grad, x, seed):
    """Central differences of an FP64 loss along random directions against <grad, v>."""
    g = torch.Generator().manual_seed(seed)
    for _ in range(2):
        v = torch.randn(x.shape, generator=g).to(x.device)
        h = 1e-2 * float(x.norm()) / float(v.norm())
        fd = (loss_fn(x + h * v) - loss_fn(x - h * v)) / (2 * h)
        an = float((grad.double() * v.double()).sum())
        assert abs(fd - an) <= 2e-4 * abs(an) + 1e-9, (fd, an)


# ---------------------------------------------------------------------------
# 2D
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("tables", [True, False])
def test_fanbeam_module_matches_oracle(dev, tables):
    n = 64
    proj = FanBeam2DProjector(n, n, 60, 96, 300.0, 450.0, det_spacing=1.0, pixel_size=1.0, device=dev,
                              precomputed_intersections=tables)
    img = torch.from_numpy(ellipse_phantom_2d(n, n, 1.0, seed=0)).to(dev)
    minv, b = _minv(proj.M), _np(proj.b)
    src, dst = _np(proj.src), _np(proj.dst)
    # tables follow the reference GPU crossing rule (|dir| < 1e-12); on-the-fly
    # with the torch rule uses the exact dir == 0 test of the torch projector
    zr = "thresh" if tables else "exact"
    ref = O.forward(_np(img), minv, b, src, dst, mode="torch", zero_rule=zr)
    got = _np(proj(img))
    assert got.shape == (src.shape[0],)
    assert rel_err(got, ref) < 1e-5
    y = np.random.default_rng(3).uniform(0, 1, src.shape[0]).astype(np.float32)
    bref = O.back(y, (n, n), minv, b, src, dst, mode="torch", zero_rule=zr)
    assert rel_err(_np(proj.back_project(torch.from_numpy(y).to(dev))), bref) < 1e-5
    # batched image -> [B, n_ray]
    two = torch.stack([img, 2 * img])
    out2 = _np(proj.forward_project(two))
    assert out2.shape == (2, src.shape[0]) and rel_err(out2[1], 2 * ref) < 1e-5


def test_own_table_runs_on_the_fly(dev):
    """A 2D module projecting with the table it computed runs the on-the-fly
    kernels with the table's rule: the forward equals the table consumer bit for
    bit, the back projection to 1e-5; once the table is written in place (or
    replaced) the module consumes it as given."""
    from ct_laboratory_amd.tomography.ct_projector_2d_hip import back_project_2d_hip, forward_project_2d_hip

    n = 80  # 1 x 2 tiles of 128 x 64
    proj = FanBeam2DProjector(n, n, 45, 120, 300.0, 450.0, det_spacing=1.0, pixel_size=1.0, device=dev)
    img = torch.from_numpy(ellipse_phantom_2d(n, n, 1.0, seed=4)).to(dev)
    assert proj._rule()[0] is None
    table_fwd = forward_project_2d_hip(img, proj.tvals, proj.M, proj.b, proj.src, proj.dst)
    assert torch.equal(proj.forward_project(img), table_fwd)
    assert torch.equal(proj(img), table_fwd)
    y = torch.from_numpy(np.random.default_rng(9).uniform(0, 1, proj.src.shape[0]).astype(np.float32)).to(dev)
    table_back = back_project_2d_hip(y, proj.tvals, proj.M, proj.b, proj.src, proj.dst, n, n)
    assert rel_err(_np(proj.back_project(y)), _np(table_back)) < 1e-5
    # an edited table is honoured: ray 0 loses its crossings
    k = int(torch.nonzero(table_fwd > 0)[0])
    with torch.no_grad():
        proj.tvals[k] = float("inf")
    assert proj._rule()[0] is proj.tvals
    edited = proj.forward_project(img)
    assert float(edited[k]) == 0.0
    assert torch.equal(edited, forward_project_2d_hip(img, proj.tvals, proj.M, proj.b, proj.src, proj.dst))


def test_autograd_2d_gradient_is_back_projection(dev):
    n = 48
    proj = FanBeam2DProjector(n, n, 40, 72, 300.0, 450.0, device=dev)
    x_true = torch.from_numpy(ellipse_phantom_2d(n, n, 1.0, seed=0)).to(dev)
    y = proj.forward_project(x_true)
    x = (0.5 * x_true).requires_grad_(True)
    loss = 0.5 * ((proj(x) - y) ** 2).sum()
    loss.backward()
    expect = proj.back_project(proj.forward_project(x.detach()) - y)
    assert rel_err(_np(x.grad), _np(expect)) < 1e-6

    def L(z):
        r = proj.forward_project(z).double() - y.double()
        return 0.5 * float((r * r).sum())

    _fd_check(L, x.grad, x.detach(), seed=5)


def test_ring_staticct_2d_matches_oracle(dev):
    n = 48
    M, b = standard_image_transform_2d(n, n, 1.0, 1.0)
    n_frame, n_source = 16, 8
    Mg = torch.eye(2).repeat(n_frame, 1, 1)
    bg = torch.zeros(n_frame, 2)
    active = torch.zeros(n_frame, n_source, dtype=torch.bool)
    active[torch.arange(n_frame), torch.arange(n_frame) % n_source] = True
    proj = UniformStaticCTProjector2D(n, n, M, b, n_source, 80.0, 32, 70.0, 8, 1.5, Mg, bg, active, device=dev)
    assert proj.src.shape[0] == n_frame * 32 * 8
    img = torch.from_numpy(ellipse_phantom_2d(n, n, 1.0, seed=1)).to(dev)
    ref = O.forward(_np(img), _minv(proj.M), _np(proj.b), _np(proj.src), _np(proj.dst), mode="torch",
                    zero_rule="thresh")
    assert rel_err(_np(proj(img)), ref) < 1e-5


def test_known_answers_2d(dev):
    """32x32 image, row 16 set: the ray along it integrates to 32
    (python_examples/test_ct_projector_2d_line_integral.py:31-43, 70); 16x16 image,
    row 8 set, rays giving 16, sqrt(2) and sqrt(15^2+7^2)/15
    (python_examples/test_ct_projector_2d_module.py:17-34)."""
    M, b = torch.eye(2), torch.zeros(2)
    img = torch.zeros(32, 32)
    img[16, :] = 1.0
    src, dst = torch.tensor([[16.0, -5.0]]), torch.tensor([[16.0, 37.0]])
    for tables in (True, False):
        p = CTProjector2DModule(32, 32, M.to(dev), b.to(dev), src.to(dev), dst.to(dev),
                                precomputed_intersections=tables)
        assert abs(float(p(img.to(dev))[0]) - 32.0) < 1e-5
    img = torch.zeros(16, 16)
    img[8, :] = 1.0
    src = torch.tensor([[8.0, -5.0], [0.0, 0.0], [15.0, 15.0]])
    dst = torch.tensor([[8.0, 25.0], [15.0, 15.0], [0.0, 8.0]])
    expect = [16.0, math.sqrt(2.0), math.sqrt(15.0 ** 2 + 7.0 ** 2) / 15.0]
    for tables in (True, False):
        p = CTProjector2DModule(16, 16, M.to(dev), b.to(dev), src.to(dev), dst.to(dev),
                                precomputed_intersections=tables)
        got = _np(p(img.to(dev)))
        assert np.allclose(got, expect, rtol=1e-5, atol=1e-6), got


# ---------------------------------------------------------------------------
# 3D
# ---------------------------------------------------------------------------


def _cone(dev, n=24, n_view=12, det=(20, 16), semantics="torch"):
    return ConeBeam3DProjector(n, n, n, n_view, det[1], det[0], 120.0, 180.0, det_spacing=2.0, voxel_size=1.0,
                               device=dev, semantics=semantics)


@pytest.mark.parametrize("semantics", ["torch", "otf"])
def test_conebeam_module_forward_back_adjoint(dev, semantics):
    proj = _cone(dev, semantics=semantics)
    assert proj._options().plan is not None, "cone-beam views should get a brick plan"
    vol = torch.from_numpy(ellipsoid_phantom_3d(24, 24, 24, 1.0, seed=1)).to(dev)
    minv, b, src, dst = _minv(proj.M), _np(proj.b), _np(proj.src), _np(proj.dst)
    ref = O.forward(_np(vol), minv, b, src, dst, mode=semantics)
    sino = proj(vol)
    assert rel_err(_np(sino), ref) < 1e-5
    y = np.random.default_rng(4).uniform(0, 1, src.shape[0]).astype(np.float32)
    bref = O.back(y, (24, 24, 24), minv, b, src, dst, mode=semantics)
    bk = _np(proj.back_project(torch.from_numpy(y).to(dev)))
    assert rel_err(bk, bref) < 1e-5
    x = np.random.default_rng(5).standard_normal((24, 24, 24)).astype(np.float32)
    yy = np.random.default_rng(6).standard_normal(src.shape[0]).astype(np.float32)
    Ax = _np(proj.forward_project(torch.from_numpy(x).to(dev))).astype(np.float64)
    Aty = _np(proj.back_project(torch.from_numpy(yy).to(dev))).astype(np.float64)
    lhs, rhs = float(Ax @ yy.astype(np.float64)), float((x.astype(np.float64) * Aty).sum())
    assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_autograd_3d_gradient_fd_and_transpose(dev):
    proj = _cone(dev)
    x_true = torch.from_numpy(ellipsoid_phantom_3d(24, 24, 24, 1.0, seed=1)).to(dev)
    y = proj.forward_project(x_true).detach()
    x = (0.7 * x_true + 0.01).requires_grad_(True)
    loss = 0.5 * ((proj(x) - y) ** 2).sum()
    loss.backward()
    expect = proj.back_project(proj.forward_project(x.detach()).detach() - y)
    assert rel_err(_np(x.grad), _np(expect)) < 1e-6

    def L(z):
        r = proj.forward_project(z).double() - y.double()
        return 0.5 * float((r * r).sum())

    _fd_check(L, x.grad, x.detach(), seed=7)
    # batched volumes go through the same Function
    xb = torch.stack([x.detach(), x_true]).requires_grad_(True)
    out = proj(xb)
    assert out.shape == (2, proj.src.shape[0])
    out.sum().backward()
    ones = torch.ones_like(out[0])
    assert rel_err(_np(xb.grad[1]), _np(proj.back_project(ones))) < 1e-6


@pytest.mark.parametrize("compress", [True, False])
def test_module_with_tables(dev, compress):
    n = 20
    src, dst = build_conebeam_3d_geometry(6, 14, 12, 100.0, 150.0, det_spacing=2.0)
    M, b = standard_image_transform_3d(n, n, n, 1.0, 1.0, 1.0)
    p = CTProjector3DModule(n, n, n, M, b, src, dst, device=dev, precomputed_intersections=True,
                            use_compression=compress)
    assert p.precomputed_intersections
    vol = torch.from_numpy(ellipsoid_phantom_3d(n, n, n, 1.0, seed=2)).to(dev)
    minv, bb, s, d = _minv(p.M), _np(p.b), _np(p.src), _np(p.dst)
    if compress:
        ref = O.forward_u16(_np(vol), _np(p.tvals_uint16), _np(p.tvals_start), _np(p.tvals_scale), minv, bb, s, d)
    else:
        ref = O.forward_tvals(_np(vol), _np(p.tvals_full), minv, bb, s, d)
    got = _np(p(vol))
    assert rel_err(got, ref) < 1e-5
    # a float table gives the line integrals of the exact traversal.  The
    # reference's uint16 format clamps deltas at 32000
    # (ct_projector_3d_torch.py:58-61): a cone-beam ray that crosses a plane
    # long before it reaches the grid decodes its later crossings shifted, so
    # the compressed table is held to the exact traversal only on rays whose
    # crossings all lie near the grid: chords of the grid box widened by a
    # third of a voxel (a ray ending inside the grid would also pick up the
    # decoded padding crossings past t = 1, as the reference's decode does).
    exact = O.forward(_np(vol), minv, bb, s, d, mode="torch")
    if not compress:
        assert rel_err(got, exact) < 1e-5
    else:
        rng = np.random.default_rng(9)
        lo, hi = bb - 0.5 - 0.3, bb + (n - 0.5) + 0.3
        p0 = rng.uniform(lo, hi, (500, 3))
        u = rng.standard_normal((500, 3))
        with np.errstate(divide="ignore"):
            t1, t2 = (lo - p0) / u, (hi - p0) / u
        tin, tout = np.minimum(t1, t2).max(axis=1), np.maximum(t1, t2).min(axis=1)
        si = (p0 + tin[:, None] * u).astype(np.float32)
        di = (p0 + tout[:, None] * u).astype(np.float32)
        q = CTProjector3DModule(n, n, n, M, b, torch.from_numpy(si), torch.from_numpy(di), device=dev,
                                precomputed_intersections=True, use_compression=True)
        ex = O.forward(_np(vol), minv, bb, si, di, mode="torch")
        assert rel_err(_np(q(vol)), ex) < 2e-3
    # gradient through the table path is its back projector
    x = vol.clone().requires_grad_(True)
    p(x).sum().backward()
    assert rel_err(_np(x.grad), _np(p.back_project(torch.ones(s.shape[0], device=dev)))) < 1e-6


def test_tvals_save_load_set(dev, tmp_path):
    n = 12
    src, dst = build_conebeam_3d_geometry(4, 10, 8, 80.0, 120.0, det_spacing=2.0)
    M, b = standard_image_transform_3d(n, n, n, 1.0, 1.0, 1.0)
    p = CTProjector3DModule(n, n, n, M, b, src, dst, device=dev, precomputed_intersections=True)
    f = tmp_path / "tv.pt"
    p.save_tvals(str(f))
    q = CTProjector3DModule(n, n, n, M, b, src, dst, device=dev)
    assert not q.precomputed_intersections
    q.load_tvals(str(f))
    assert torch.equal(q.tvals_uint16, p.tvals_uint16) and torch.equal(q.tvals_scale, p.tvals_scale)
    vol = torch.from_numpy(ellipsoid_phantom_3d(n, n, n, 1.0, seed=3)).to(dev)
    assert torch.equal(p(vol), q(vol))
    # decompressed tables match the float table they came from to the t quantum
    full = CTProjector3DModule(n, n, n, M, b, src, dst, device=dev, precomputed_intersections=True,
                               use_compression=False)
    tv = full.tvals_full
    dec = decompress_tvals_from_uint16(p.tvals_uint16, p.tvals_start, p.tvals_scale)
    # each crossing decodes within half a quantum (plus float rounding) up to
    # the first delta that hit the format's 32000 clamp
    deltas = p.tvals_uint16.to(torch.int64) & 0xFFFF
    clamped = torch.cumsum((deltas >= 32000).to(torch.int64), dim=1) > 0
    fin = torch.isfinite(tv) & ~clamped
    assert int(fin[:, 1:].sum()) > 0
    err = (dec - torch.where(fin, tv, dec)).abs() - 0.5 * p.tvals_scale[:, None]
    assert float(err.max()) <= 1e-6
    # set_tvals(None) returns to the on-the-fly traversal
    q.set_tvals(None)
    assert q.tvals is None and not q.precomputed_intersections
    q.set_tvals(tv)  # compressed on the way in
    assert q.tvals_uint16 is not None
    # stitched host tables: the reference's full width and GPU compression rule
    t16, ts, tsc = precompute_tvals_stitched(n, n, n, M, b, src, dst, chunk_size=100, verbose=False)
    assert t16.shape == (src.shape[0], 3 * (n + 1)) and t16.device.type == "cpu"


def test_staticct_3d_ring_matches_oracle(dev):
    n = 16
    M, b = standard_image_transform_3d(n, n, n, 1.0, 1.0, 1.0)
    Mg, bg, act = build_circular_sequence(6, 6)
    proj = UniformStaticCTProjector3D(n, n, n, 6, 40.0, 0.0, 24, 36.0, 6, 5, 1.5, 1.5, 0.0, Mg, bg, act, 8, M, b,
                                      device=dev)
    assert proj.src.shape[0] == 6 * 8 * 6 * 5
    vol = torch.from_numpy(ellipsoid_phantom_3d(n, n, n, 1.0, seed=4)).to(dev)
    minv, bb, s, d = _minv(proj.M), _np(proj.b), _np(proj.src), _np(proj.dst)
    assert rel_err(_np(proj(vol)), O.forward(_np(vol), minv, bb, s, d, mode="torch")) < 1e-5
    y = np.random.default_rng(8).uniform(0, 1, s.shape[0]).astype(np.float32)
    bref = O.back(y, (n, n, n), minv, bb, s, d, mode="torch")
    assert rel_err(_np(proj.back_project(torch.from_numpy(y).to(dev))), bref) < 1e-5


def test_known_answers_3d(dev):
    """32^3 volume, line (., 16, 16) set, ray along x: 32
    (python_examples/test_ct_projector_3d_line_integral.py:33-45, 74); 8^3 volume,
    slab x = 4 set: rays giving 8 and sqrt(3) (python_examples/test_ct_projector_3d_module.py:17-32)."""
    M, b = torch.eye(3), torch.zeros(3)
    vol = torch.zeros(32, 32, 32)
    vol[:, 16, 16] = 1.0
    src, dst = torch.tensor([[200.0, 16.0, 16.0]]), torch.tensor([[-200.0, 16.0, 16.0]])
    for sem in ("torch", "otf"):
        p = CTProjector3DModule(32, 32, 32, M, b, src, dst, device=dev, semantics=sem)
        assert abs(float(p(vol.to(dev))[0]) - 32.0) < 1e-4
    vol = torch.zeros(8, 8, 8)
    vol[4, :, :] = 1.0
    src = torch.tensor([[4.0, -5.0, 4.0], [0.0, 0.0, 0.0]])
    dst = torch.tensor([[4.0, 15.0, 4.0], [7.0, 7.0, 7.0]])
    p = CTProjector3DModule(8, 8, 8, M, b, src, dst, device=dev)
    got = _np(p(vol.to(dev)))
    assert abs(got[0] - 8.0) < 1e-5 and abs(got[1] - math.sqrt(3.0)) < 1e-5, got


def test_function_apply_reference_signature(dev):
    """CTProjector3DFunction.apply with the reference's positional arguments
    (volume, t_main, t_start, t_scale, M, b, src, dst, backend)."""
    n = 10
    src, dst = build_conebeam_3d_geometry(3, 6, 5, 60.0, 90.0, det_spacing=2.0)
    M, b = standard_image_transform_3d(n, n, n, 1.0, 1.0, 1.0)
    src, dst, M, b = src.to(dev), dst.to(dev), M.to(dev), b.to(dev)
    vol = torch.rand(n, n, n, device=dev, requires_grad=True)
    out = CTProjector3DFunction.apply(vol, None, None, None, M, b, src, dst, "cuda")
    ref = O.forward(_np(vol), _minv(M), _np(b), _np(src), _np(dst), mode="torch")
    assert rel_err(_np(out), ref) < 1e-5
    out.sum().backward()
    assert vol.grad is not None and vol.grad.shape == vol.shape
    with pytest.raises(ValueError):
        CTProjector3DFunction.apply(vol, None, None, None, M, b, src, dst, "torch")
    img = torch.rand(8, 8, device=dev)
    M2, b2 = torch.eye(2, device=dev), torch.zeros(2, device=dev)
    s2, d2 = torch.tensor([[4.0, -1.0]], device=dev), torch.tensor([[4.0, 9.0]], device=dev)
    assert CTProjector2DFunction.apply(img, None, M2, b2, s2, d2, "hip").shape == (1,)


def test_fanbeam_full_size_config(dev):
    """The 2D fan-beam configuration at full size: 256x256 image at 1 mm, 360 views,
    512 detector elements at 1 mm, SID 1000 mm, SDD 1500 mm (184,320 rays), seeded
    10-ellipse phantom with 0.02/mm background; forward and back projection against
    the oracle through the table path and the on-the-fly path, and the adjoint."""
    n = 256
    img = torch.from_numpy(ellipse_phantom_2d(n, n, 1.0, seed=0)).to(dev)
    y = np.random.default_rng(2).standard_normal(360 * 512).astype(np.float32)
    for tables in (True, False):
        proj = FanBeam2DProjector(n, n, 360, 512, 1000.0, 1500.0, det_spacing=1.0, pixel_size=1.0, device=dev,
                                  precomputed_intersections=tables)
        assert proj.src.shape[0] == 184320
        minv, b, src, dst = _minv(proj.M), _np(proj.b), _np(proj.src), _np(proj.dst)
        zr = "thresh" if tables else "exact"
        ref = O.forward(_np(img), minv, b, src, dst, mode="torch", zero_rule=zr)
        got = _np(proj.forward_project(img))
        assert rel_err(got, ref) < 1e-5
        bk = _np(proj.back_project(torch.from_numpy(y).to(dev))).astype(np.float64)
        bref = O.back(y, (n, n), minv, b, src, dst, mode="torch", zero_rule=zr)
        bound = O.back(np.abs(y), (n, n), minv, b, src, dst, mode="torch", zero_rule=zr)
        assert np.all(np.abs(bk - bref) <= 1e-5 * bound + 1e-6)
        lhs = float(got.astype(np.float64) @ y.astype(np.float64))
        rhs = float((_np(img).astype(np.float64) * bk).sum())
        assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_copied_table_runs_on_the_fly(dev):
    """A 2D module whose table is a copy (as after loading it from disk) is
    compared once with the table computed for its rays, grid and map; being
    equal, it is projected on the fly with the table's rule: the same results
    as the module's own table (the back projection on LDS tiles)."""
    from ct_laboratory_amd import _C
    from ct_laboratory_amd import _native as N

    n = 96
    proj = FanBeam2DProjector(n, n, 30, 140, 300.0, 450.0, det_spacing=1.0, pixel_size=1.0, device=dev)
    img = torch.from_numpy(ellipse_phantom_2d(n, n, 1.0, seed=6)).to(dev)
    y = torch.from_numpy(np.random.default_rng(10).uniform(0, 1, proj.src.shape[0]).astype(np.float32)).to(dev)
    own_fwd, own_back = proj.forward_project(img), proj.back_project(y)
    proj.tvals = proj.tvals.clone()
    assert proj._rule()[0] is proj.tvals
    assert torch.equal(proj.forward_project(img), own_fwd)
    assert rel_err(_np(proj.back_project(y)), _np(own_back)) < 1e-6
    assert _C._TABLES2D[id(proj.tvals)][3] == N.ZERO_THRESH
    x = img.clone().requires_grad_(True)
    proj(x).backward(y)
    assert rel_err(_np(x.grad), _np(own_back)) < 1e-6


def test_conebeam_rays_built_on_the_device_equal_the_host_ones(dev):
    """The cone-beam builder expands the per-view quantities (computed on the
    CPU) over the detector on the device: the same float32 multiplies and adds,
    so the rays equal the CPU-built ones bit for bit."""
    from ct_laboratory_amd.tomography.conebeam_projector_3d import build_conebeam_3d_geometry

    for args in ((37, 64, 48, 1000.0, 1500.0, 1.5, 0.3), (8, 384, 384, 600.0, 900.0, 2.0, 0.0)):
        s_cpu, d_cpu = build_conebeam_3d_geometry(*args[:5], det_spacing=args[5], angle_offset=args[6])
        s_dev, d_dev = build_conebeam_3d_geometry(*args[:5], det_spacing=args[5], angle_offset=args[6], device=dev)
        assert s_dev.is_cuda and d_dev.is_cuda
        assert torch.equal(s_dev.cpu(), s_cpu) and torch.equal(d_dev.cpu(), d_cpu)
