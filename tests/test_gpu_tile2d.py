"""GPU tests of the 2D LDS tile back projection behind ctlab_back_project_2d_otf
and the tiled multi-scan launch: ragged tile grids, one part per tile (plain
stores) and several (float atomics per pixel), accumulate, both segment rules,
rotated maps, signed data with the fixed-point bound, repeat determinism,
value scales near both ends of the float range, INF/NaN rays, zero and empty
inputs."""
import numpy as np
import pytest
import torch

import oracle as O
from ct_laboratory_amd import _C
from ct_laboratory_amd import _native as N
from helpers import random_rays, rel_err

pytestmark = pytest.mark.gpu

MODES = [("torch", N.MODE_TORCH), ("otf", N.MODE_OTF)]


def _t(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)


def _geometry(rng, shape, rotated):
    s = rng.uniform(0.6, 1.4, 2)
    M = np.diag(s)
    if rotated:
        th = rng.uniform(0, np.pi)
        M = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]]) @ M
    minv = np.linalg.inv(M).astype(np.float32)
    b = (-(M @ ((np.array(shape) - 1) / 2.0))).astype(np.float32)
    return minv, b, s.astype(np.float32)


def _back(dev, y, src, dst, minv, b, shape, mode, zr, accumulate=0, out=None):
    # every device buffer is held by a name until the launch is enqueued
    ty, ts, td, tm, tb = (_t(a, dev) for a in (y, src, dst, minv, b))
    out = torch.zeros((y.shape[0],) + shape, device=dev) if out is None else out
    N.call("ctlab_back_project_2d_otf", N.ptr(ty), y.shape[0], y.shape[1], N.ptr(ts), N.ptr(td), N.ptr(tm), N.ptr(tb),
           shape[0], shape[1], mode, zr, accumulate, N.ptr(out), N.stream_of(ty))
    torch.cuda.synchronize()
    return out


# n_ray 3000: one part per tile (plain stores); 20000: several parts (atomics)
@pytest.mark.parametrize("n_ray", [3000, 20000])
@pytest.mark.parametrize("mname,mode", MODES)
def test_tiles_vs_oracle(dev, mname, mode, n_ray):
    rng = np.random.default_rng(700 + n_ray + mode)
    shape = (150, 97)  # 2 x 2 tiles of 128 x 64, ragged in both directions
    minv, b, s = _geometry(rng, shape, rotated=(n_ray == 20000))
    src, dst = random_rays(rng, shape, s, n_ray, "far")
    zr = N.ZERO_EXACT if mname == "torch" else N.ZERO_THRESH
    ozr = "exact" if zr == N.ZERO_EXACT else "thresh"
    y = rng.standard_normal((2, n_ray)).astype(np.float32)
    got = _back(dev, y, src, dst, minv, b, shape, mode, zr).cpu().numpy()
    ref = O.back(y, shape, minv, b, src, dst, mode=mname, zero_rule=ozr)
    bound = O.back(np.abs(y), shape, minv, b, src, dst, mode=mname, zero_rule=ozr)
    floor = 1e-10 * float(np.abs(y).max())
    assert np.all(np.abs(got - ref) <= 1e-5 * bound + floor)
    # accumulate adds to what the output holds
    base = rng.uniform(-1, 1, (2,) + shape).astype(np.float32)
    acc = _back(dev, y, src, dst, minv, b, shape, mode, zr, accumulate=1, out=_t(base, dev)).cpu().numpy()
    assert np.all(np.abs(acc - (base + ref)) <= 1e-5 * (bound + np.abs(base)) + floor)
    if n_ray == 3000:  # one part per tile: fixed-point sums and plain stores, bit-identical repeats
        assert np.array_equal(got, _back(dev, y, src, dst, minv, b, shape, mode, zr).cpu().numpy())


@pytest.mark.parametrize("kind", ["inside", "distant", "lattice"])
@pytest.mark.parametrize("mname,mode", MODES)
def test_tiles_ray_kinds(dev, mname, mode, kind):
    """Rays that start or end inside the image, rays from ~4e4 pixels away
    (midpoint error bounds above one half: the exhaustive walk and the widest
    filter margins), and half-integer lattice rays on the identity grid (rays
    along tile and pixel boundaries, through corners, axis-parallel)."""
    rng = np.random.default_rng(900 + 7 * mode + len(kind))
    shape = (200, 150)  # 2 x 3 tiles
    if kind == "lattice":
        minv, b, s = np.eye(2, dtype=np.float32), (-(np.array(shape) - 1) / 2).astype(np.float32), np.ones(2)
    else:
        minv, b, s = _geometry(rng, shape, rotated=True)
    src, dst = random_rays(rng, shape, s, 6000, kind)
    zr = N.ZERO_EXACT if mname == "torch" else N.ZERO_THRESH
    ozr = "exact" if zr == N.ZERO_EXACT else "thresh"
    y = rng.standard_normal((1, src.shape[0])).astype(np.float32)
    got = _back(dev, y, src, dst, minv, b, shape, mode, zr).cpu().numpy()
    ref = O.back(y, shape, minv, b, src, dst, mode=mname, zero_rule=ozr)
    bound = O.back(np.abs(y), shape, minv, b, src, dst, mode=mname, zero_rule=ozr)
    assert float(bound.max()) > 0.0
    assert np.all(np.abs(got - ref) <= 1e-5 * bound + 1e-10 * float(np.abs(y).max()))


@pytest.mark.parametrize("mname,mode", MODES)
def test_tiles_scales_and_nonfinite(dev, mname, mode):
    rng = np.random.default_rng(77 + mode)
    shape = (70, 130)
    minv, b, s = _geometry(rng, shape, rotated=False)
    src, dst = random_rays(rng, shape, s, 2500, "far")
    zr = N.ZERO_EXACT if mname == "torch" else N.ZERO_THRESH
    ozr = "exact" if zr == N.ZERO_EXACT else "thresh"
    y = rng.standard_normal((1, src.shape[0])).astype(np.float32)
    ref = O.back(y, shape, minv, b, src, dst, mode=mname, zero_rule=ozr)
    bound = O.back(np.abs(y), shape, minv, b, src, dst, mode=mname, zero_rule=ozr)
    floor = 1e-10 * float(np.abs(y).max())
    for sc in (1e-30, 1e30):  # power-of-two-free scales near both ends of the float range
        ys = (y * np.float32(sc)).astype(np.float32)
        gs = _back(dev, ys, src, dst, minv, b, shape, mode, zr).cpu().numpy().astype(np.float64)
        assert np.all(np.abs(gs - ref * sc) <= (1e-5 * bound + floor) * sc)
    # an INF and a NaN ray: pixels they cross carry INF / NaN, the rest is unchanged
    through = np.nonzero(O.forward(np.ones((1,) + shape, np.float32), minv, b, src, dst, mode=mname,
                                   zero_rule=ozr)[0] > 0)[0]
    k_inf, k_nan = int(through[0]), int(through[len(through) // 2])
    hits = O.back(np.eye(src.shape[0], dtype=np.float32)[[k_inf, k_nan]], shape, minv, b, src, dst, mode=mname,
                  zero_rule=ozr) > 0
    assert hits[0].any() and hits[1].any()
    yn = y.copy()
    yn[0, k_inf], yn[0, k_nan] = np.inf, np.nan
    gn = _back(dev, yn, src, dst, minv, b, shape, mode, zr).cpu().numpy()[0]
    assert np.all(np.isnan(gn[hits[1]]))
    assert np.all(np.isposinf(gn[hits[0] & ~hits[1]]))
    clean = ~(hits[0] | hits[1])
    assert np.all(np.abs(gn[clean] - ref[0][clean]) <= 1e-5 * bound[0][clean] + floor)


def test_tiles_zero_and_empty(dev):
    shape = (80, 80)
    minv, b = np.eye(2, dtype=np.float32), np.full(2, -39.5, np.float32)
    src = np.array([[-100.0, 0.3], [0.2, -100.0]], np.float32)
    dst = -src
    # all-zero sinogram: zero image (also over a non-zero output buffer)
    out = _back(dev, np.zeros((1, 2), np.float32), src, dst, minv, b, shape, N.MODE_TORCH, N.ZERO_EXACT,
                out=torch.full((1,) + shape, 7.0, device=dev))
    assert float(out.abs().sum()) == 0.0
    # no rays at all
    e = np.zeros((0, 2), np.float32)
    out = _C.back_project_2d_on_the_fly_cuda(_t(np.zeros((1, 0)), dev), _t(e, dev), _t(e, dev), _t(minv, dev),
                                             _t(b, dev), *shape)
    assert out.shape == (1,) + shape and float(out.abs().sum()) == 0.0
    # one ray along pixel row 40 (through the pixel centres): length 1 in each of
    # its 80 pixels, nothing elsewhere (across both tiles of the 1 x 2 grid)
    got = _back(dev, np.ones((1, 1), np.float32), np.array([[0.5, -100.0]], np.float32),
                np.array([[0.5, 100.0]], np.float32), minv, b, shape, N.MODE_TORCH, N.ZERO_EXACT).cpu().numpy()[0]
    # (the torch rule measures lengths between float world points near |y| = 100,
    # each rounded by up to 2^-17 mm; the row's lengths telescope to its extent)
    np.testing.assert_allclose(got[40], 1.0, rtol=0, atol=2 * 2.0 ** -16)
    assert abs(float(got[40].astype(np.float64).sum()) - 80.0) < 1e-4
    assert float(np.abs(np.delete(got, 40, axis=0)).sum()) == 0.0


@pytest.mark.parametrize("mname,mode", MODES)
def test_tiles_nonfinite_geometry(dev, mname, mode):
    """Rays with INF or NaN end points contribute nothing (every world point of
    such a ray is INF or NaN and masked), the other rays as usual; a map with a
    NaN entry gives a zero image (every index is out of range)."""
    rng = np.random.default_rng(31 + mode)
    shape = (130, 70)
    minv, b, s = _geometry(rng, shape, rotated=True)
    src, dst = random_rays(rng, shape, s, 3000, "far")
    src[::7, 0] = np.inf
    dst[3::11, 1] = np.nan
    src[5::13, 1] = -np.inf
    zr = N.ZERO_EXACT if mname == "torch" else N.ZERO_THRESH
    ozr = "exact" if zr == N.ZERO_EXACT else "thresh"
    y = rng.uniform(0.5, 1.5, (1, src.shape[0])).astype(np.float32)
    got = _back(dev, y, src, dst, minv, b, shape, mode, zr).cpu().numpy()
    ref = O.back(y, shape, minv, b, src, dst, mode=mname, zero_rule=ozr)
    assert np.isfinite(got).all() and float(ref.max()) > 0.0
    assert rel_err(got, ref) < 1e-5
    bad = minv.copy()
    bad[0, 1] = np.nan
    out = _back(dev, y, src, dst, bad, b, shape, mode, zr, out=torch.full((1,) + shape, 3.0, device=dev))
    assert float(out.abs().sum()) == 0.0
