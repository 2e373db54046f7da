"""Pins the CPU oracle (test infrastructure) before it is trusted as the checker.

* Golden vectors of the reference's torch projector (tests/golden/*.npz, written
  by tests/golden/make_golden.py from ct_projector_{2d,3d}_torch.py): crossing
  tables and uint16 compression bit for bit, projections to float rounding.
* Closed forms from the reference's example scripts.
* Exact rational arithmetic on tiny grids: the segment lists of the FP64 oracle
  equal the exact ones, and those of the FP32 oracle too except on rays within
  rounding distance of a tie.
* Invariants: A.1 is the chord length through the grid box; adjoint identity.
"""
import glob
import math
import os
from fractions import Fraction

import numpy as np
import pytest
import torch

import oracle as O
from helpers import positive_lists, random_geometry, random_rays

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(pattern):
    files = sorted(glob.glob(os.path.join(GOLDEN, pattern)))
    assert files, f"missing golden vectors {pattern}"
    return files


def _bits_equal(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    return np.array_equal(a.view(np.int32), b.view(np.int32))


# ---------------------------------------------------------------------------
# Golden vectors
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("path", _golden("torch3d_*.npz"))
def test_golden_3d(path):
    z = np.load(path)
    shape = tuple(int(x) for x in z["shape"])
    minv, b, src, dst = z["minv"], z["b"], z["src"], z["dst"]
    # crossing tables: compute_intersections_3d_torch, bit for bit
    assert _bits_equal(O.intersections(shape, minv, b, src, dst, zero_rule="thresh"), z["tvals"])
    # projections: same segments, float sums in another order
    fwd = O.forward(z["vol"], minv, b, src, dst, mode="torch", zero_rule="thresh")
    assert np.abs(fwd - z["fwd"]).max() <= 2e-6 * np.abs(z["fwd"]).max()
    assert np.abs(O.forward_tvals(z["vol"], z["tvals"], minv, b, src, dst) - z["fwd"]).max() <= 2e-6 * np.abs(
        z["fwd"]).max()
    bk = O.back(z["sino_in"], shape, minv, b, src, dst, mode="torch", zero_rule="thresh")
    assert np.abs(bk - z["back"]).max() <= 2e-6 * np.abs(z["back"]).max()
    # compress_tvals_to_uint16: deltas, starts and scales bit for bit
    u16, st, sc = O.compress_torch(z["tvals"], src, dst, O.voxel_diag_len(z["M"]))
    assert np.array_equal(u16, z["u16"]) and _bits_equal(st, z["start"]) and _bits_equal(sc, z["scale"])
    # decompress_tvals_from_uint16 of the package, bit for bit
    from ct_laboratory_amd.tomography import decompress_tvals_from_uint16

    dec = decompress_tvals_from_uint16(torch.from_numpy(z["u16"]), torch.from_numpy(z["start"]),
                                       torch.from_numpy(z["scale"]))
    assert _bits_equal(dec.numpy(), z["decompressed"])
    # the torch path projects the decoded table like any float table
    fu = O.forward_tvals(z["vol"], z["decompressed"], minv, b, src, dst)
    assert np.abs(fu - z["fwd_u16"]).max() <= 2e-6 * np.abs(z["fwd_u16"]).max()


@pytest.mark.parametrize("path", _golden("torch2d_*.npz"))
def test_golden_2d(path):
    z = np.load(path)
    shape = tuple(int(x) for x in z["shape"])
    minv, b, src, dst = z["minv"], z["b"], z["src"], z["dst"]
    # 2D torch uses the exact dir == 0 test (2d_torch.py:42-44)
    assert _bits_equal(O.intersections(shape, minv, b, src, dst, zero_rule="exact"), z["tvals"])
    fwd = O.forward(z["img"], minv, b, src, dst, mode="torch", zero_rule="exact")
    assert np.abs(fwd - z["fwd"]).max() <= 2e-6 * np.abs(z["fwd"]).max()
    bk = O.back(z["sino_in"], shape, minv, b, src, dst, mode="torch", zero_rule="exact")
    assert np.abs(bk - z["back"]).max() <= 2e-6 * np.abs(z["back"]).max()


def test_voxel_diag_len_known_values():
    assert O.voxel_diag_len(np.eye(3, dtype=np.float32)) == np.float32(math.sqrt(3.0))
    M = np.diag([0.5, 2.0, 1.0]).astype(np.float32)
    assert abs(float(O.voxel_diag_len(M)) - math.sqrt(0.25 + 4.0 + 1.0)) < 1e-6


# ---------------------------------------------------------------------------
# Closed forms
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("mode", ["torch", "otf"])
def test_closed_forms(mode):
    """python_examples/test_ct_projector_{2d,3d}_line_integral.py and *_module.py."""
    img = np.zeros((32, 32), np.float32)
    img[16, :] = 1
    got = O.forward(img, np.eye(2, dtype=np.float32), np.zeros(2, np.float32), np.array([[16, -5]], np.float32),
                    np.array([[16, 37]], np.float32), mode=mode)
    assert abs(got[0] - 32.0) < 1e-5
    img = np.zeros((16, 16), np.float32)
    img[8, :] = 1
    src = np.array([[8, -5], [0, 0], [15, 15]], np.float32)
    dst = np.array([[8, 25], [15, 15], [0, 8]], np.float32)
    got = O.forward(img, np.eye(2, dtype=np.float32), np.zeros(2, np.float32), src, dst, mode=mode)
    assert np.allclose(got, [16.0, math.sqrt(2), math.sqrt(15 ** 2 + 7 ** 2) / 15], rtol=2e-6)
    vol = np.zeros((32, 32, 32), np.float32)
    vol[:, 16, 16] = 1
    got = O.forward(vol, np.eye(3, dtype=np.float32), np.zeros(3, np.float32), np.array([[200, 16, 16]], np.float32),
                    np.array([[-200, 16, 16]], np.float32), mode=mode)
    assert abs(got[0] - 32.0) < 1e-4
    vol = np.zeros((8, 8, 8), np.float32)
    vol[4] = 1
    got = O.forward(vol, np.eye(3, dtype=np.float32), np.zeros(3, np.float32),
                    np.array([[4, -5, 4], [0, 0, 0]], np.float32), np.array([[4, 15, 4], [7, 7, 7]], np.float32),
                    mode=mode)
    assert np.allclose(got, [8.0, math.sqrt(3)], rtol=2e-6)


# ---------------------------------------------------------------------------
# Exact arithmetic on tiny grids
# ---------------------------------------------------------------------------


def _rne(q: Fraction) -> int:
    f = math.floor(q)
    r = q - f
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and f % 2 == 1):
        return f + 1
    return f


def _exact_segments(shape, minv, b, src, dst):
    """Segments with positive length of one ray in exact rational arithmetic
    over the float32 inputs: (flat index list, lengths, degenerate flag)."""
    nd = len(shape)
    F = [[Fraction(float(minv[a, j])) for j in range(nd)] for a in range(nd)]
    sb = [Fraction(float(src[j])) - Fraction(float(b[j])) for j in range(nd)]
    db = [Fraction(float(dst[j])) - Fraction(float(b[j])) for j in range(nd)]
    s = [sum(F[a][j] * sb[j] for j in range(nd)) for a in range(nd)]
    d = [sum(F[a][j] * db[j] for j in range(nd)) for a in range(nd)]
    v = [d[a] - s[a] for a in range(nd)]
    ts = set()
    for a in range(nd):
        if v[a] == 0:
            continue
        for p in range(shape[a] + 1):
            t = (Fraction(2 * p - 1, 2) - s[a]) / v[a]
            if 0 <= t <= 1:
                ts.add(t)
    ts = sorted(ts)
    wl = math.sqrt(sum((float(dst[j]) - float(src[j])) ** 2 for j in range(nd)))
    idx, lens, degen = [], [], False
    for t0, t1 in zip(ts[:-1], ts[1:]):
        if t1 - t0 < Fraction(1, 10 ** 5) or t0 < Fraction(1, 10 ** 6) or t1 > 1 - Fraction(1, 10 ** 6):
            degen = True
        m = [s[a] + (t0 + t1) / 2 * v[a] for a in range(nd)]
        for a in range(nd):
            frac = m[a] - math.floor(m[a])
            if abs(frac - Fraction(1, 2)) < Fraction(1, 10 ** 3):
                degen = True
        ii = [_rne(m[a]) for a in range(nd)]
        if all(0 <= ii[a] < shape[a] for a in range(nd)):
            flat = ii[0] * shape[1] + ii[1] if nd == 2 else (ii[0] * shape[1] + ii[1]) * shape[2] + ii[2]
            idx.append(flat)
            lens.append(float(t1 - t0) * wl)
    return np.array(idx, np.int64), np.array(lens), degen


@pytest.mark.parametrize("nd", [2, 3])
def test_exact_rational_brute_force(nd):
    rng = np.random.default_rng(40 + nd)
    n_exact_fp32 = n_rays = 0
    for trial in range(6):
        shape, M, minv, b, s = random_geometry(rng, nd, max_n=7, rotated=(trial % 2 == 1))
        src, dst = random_rays(rng, shape, s, 40, "far" if trial < 3 else "inside")
        f64 = positive_lists(*O.trace_f64(shape, minv, b, src, dst, zero_rule="exact"))
        f32 = positive_lists(*O.trace(shape, minv, b, src, dst, mode="torch", zero_rule="exact"))
        for r in range(src.shape[0]):
            ei, el, degen = _exact_segments(shape, minv, b, src[r], dst[r])
            n_rays += 1
            # FP64 oracle: the exact segments unless a tie is within rounding
            if not degen:
                gi, gl = f64[r]
                assert np.array_equal(gi, ei), (r, gi, ei)
                assert np.allclose(gl, el, rtol=1e-9, atol=1e-12)
                gi32, gl32 = f32[r]
                if np.array_equal(gi32, ei):
                    n_exact_fp32 += 1
                    assert np.allclose(gl32, el, rtol=1e-4, atol=1e-5)
    # the FP32 (reference-rounding) lists equal the exact ones on almost every ray
    assert n_exact_fp32 >= 0.9 * n_rays


# ---------------------------------------------------------------------------
# Invariants
# ---------------------------------------------------------------------------


def _chord(shape, minv, b, src, dst):
    """|dst - src| * length of t in [0,1] with the index-space point inside the box."""
    s = (src.astype(np.float64) - b) @ minv.astype(np.float64).T
    d = (dst.astype(np.float64) - b) @ minv.astype(np.float64).T
    v = d - s
    lo = np.full(s.shape[0], 0.0)
    hi = np.full(s.shape[0], 1.0)
    for a in range(len(shape)):
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = (-0.5 - s[:, a]) / v[:, a]
            t2 = (shape[a] - 0.5 - s[:, a]) / v[:, a]
        inside = (s[:, a] >= -0.5) & (s[:, a] <= shape[a] - 0.5)
        par = v[:, a] == 0
        lo = np.where(par, np.where(inside, lo, 2.0), np.maximum(lo, np.minimum(t1, t2)))
        hi = np.where(par, hi, np.minimum(hi, np.maximum(t1, t2)))
    return np.maximum(hi - lo, 0.0) * np.linalg.norm(dst.astype(np.float64) - src.astype(np.float64), axis=1)


@pytest.mark.parametrize("nd", [2, 3])
def test_ones_give_chord_lengths_and_adjoint(nd):
    rng = np.random.default_rng(60 + nd)
    for trial in range(3):
        shape, M, minv, b, s = random_geometry(rng, nd, max_n=20, rotated=(trial == 2))
        src, dst = random_rays(rng, shape, s, 500, "far")
        ones = np.ones(shape, np.float32)
        ch = _chord(shape, minv, b, src, dst)
        # The torch rule has segments only between consecutive crossings, so a
        # ray ending inside the box loses its end pieces (3d_torch.py:161-274);
        # the otf rule inserts t = 0 and 1 (3d.cu:890-891).
        out = np.ones(src.shape[0], bool)
        for p in (src, dst):
            c = (p.astype(np.float64) - b) @ minv.astype(np.float64).T
            out &= ((c < -0.5) | (c > np.array(shape) - 0.5)).any(axis=1)
        assert out.sum() > 300
        f64 = O.forward_f64(ones, minv, b, src, dst, zero_rule="exact")
        assert np.allclose(f64[out], ch[out], rtol=1e-9, atol=1e-9)
        assert (f64 <= ch * (1 + 1e-9) + 1e-9).all()
        assert np.allclose(O.forward(ones, minv, b, src, dst, mode="torch")[out], ch[out], rtol=2e-5, atol=2e-5)
        assert np.allclose(O.forward(ones, minv, b, src, dst, mode="otf"), ch, rtol=2e-5, atol=2e-5)
        x = rng.standard_normal(shape)
        y = rng.standard_normal(src.shape[0])
        # signed data cancel in the dot products: the error is bounded relative
        # to sum |Ax|_i |y|_i, the scale of the terms
        ax = O.forward_f64(x, minv, b, src, dst, zero_rule="exact")
        aty = O.back_f64(y, shape, minv, b, src, dst, zero_rule="exact")
        assert abs(ax @ y - (x * aty).sum()) <= 1e-12 * (np.abs(ax) @ np.abs(y))
        for mode in ("torch", "otf"):
            ax = O.forward(x.astype(np.float32), minv, b, src, dst, mode=mode).astype(np.float64)
            aty = O.back(y.astype(np.float32), shape, minv, b, src, dst, mode=mode).astype(np.float64)
            assert abs(ax @ y - (x * aty).sum()) <= 1e-6 * (np.abs(ax) @ np.abs(y))


def test_u16_decode_error_bound():
    """CUDA-rule compression: every crossing decodes within half a quantum until
    a delta hits the 32000 clamp (3d.cu:160-212, 128-129)."""
    rng = np.random.default_rng(70)
    shape, M, minv, b, s = random_geometry(rng, 3, max_n=16)
    src, dst = random_rays(rng, shape, s, 400, "inside")
    tv = O.intersections(shape, minv, b, src, dst)
    K = int(np.isfinite(tv).sum(axis=1).max())
    tv = np.ascontiguousarray(tv[:, :K])
    u16, st, sc = O.compress_cuda(tv, src, dst, O.voxel_diag_len(M))
    k = np.cumsum(u16.astype(np.int64) & 0xFFFF, axis=1)
    dec = st[:, None].astype(np.float64) + k * sc[:, None].astype(np.float64)
    clamped = np.cumsum((u16.astype(np.int64) & 0xFFFF) >= 32000, axis=1) > 0
    ok = np.isfinite(tv) & ~clamped
    assert ok.sum() > 0.9 * np.isfinite(tv).sum()
    err = np.abs(dec - tv)[ok] - 0.5 * np.broadcast_to(sc[:, None], tv.shape)[ok]
    assert err.max() <= 1e-6
