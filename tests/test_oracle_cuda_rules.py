"""Pins of the oracle's restatements of the reference's CUDA rules, which no
golden vector covers (the CUDA code cannot be compiled here):

* compress_tvals_3d_kernel (src/ct_projector_3d.cu:160-212): roundf ties away
  from zero, INF entries -> delta 0, start = t[0] even when t[0] is INF; known
  answers built with power-of-two scales, where the torch rule
  (ct_projector_3d_torch.py:4-65) answers differently.
* The compressed consumers' decode (3d.cu:126-129): start + k*scale with k the
  integer running sum, rounded once (a fused multiply-add), checked against the
  correctly rounded exact value in rational arithmetic; the consumers'
  projections equal the float-table projections of the decoded table; and the
  skip of segments shorter than 1e-9 (3d.cu:137) on a known answer.
* The on-the-fly rule (3d.cu:834-1012): index lists against exact rational
  arithmetic with the end points 0 and 1 inserted.
* The 2D on-the-fly march (src/ct_projector_2d.cu:246-413): hand-derived known
  answers (floor-indexed entry voxel, the negative first step of a ray that
  starts past a voxel centre) and the chord length through the box.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from helpers import positive_lists, random_geometry, random_rays

INF = np.float32(np.inf)


def _ray4():
    # |dst - src| = 4 exactly, so scale = (voxel_diag / 10000) / 4
    return np.array([[0, 0, 0]], np.float32), np.array([[4, 0, 0]], np.float32)


# ---------------------------------------------------------------------------
# Compression
# ---------------------------------------------------------------------------


def test_compress_cuda_known_answers():
    src, dst = _ray4()
    vd = 10000.0 * 2.0 ** -10  # voxel_diag / 10000 = 2^-10 exactly; scale = 2^-12, 1/scale = 4096
    q = 2.0 ** -12
    rows = np.array([
        [0.25, 0.25 + 2.5 * q, 0.25 + 3.5 * q, 0.25 + 4.5 * q, INF, INF],  # ties
        [0.5, INF, 0.5 + 7 * q, 0.5 + 9 * q, INF, INF],                      # INF inside a row
        [INF] * 6,                                                            # no crossing at all
        [0.125, 0.125 + 40000 * q, 0.125 + 40001 * q, INF, INF, INF],        # clamp at 32000
    ], np.float32)
    u, st, sc = O.compress_cuda(rows, np.repeat(src, 4, 0), np.repeat(dst, 4, 0), vd)
    assert np.all(sc == np.float32(q))
    # roundf: 2.5 -> 3, 3.5 -> 4, 4.5 -> 5 (ties away from zero); deltas 3, 1, 1
    assert u[0].tolist() == [0, 3, 1, 1, 0, 0]
    # INF -> delta 0 without moving the running sum: 7 - 0, then 9 - 7
    assert u[1].tolist() == [0, 0, 7, 2, 0, 0]
    # an all-INF row keeps start = INF and all deltas 0
    assert np.isposinf(st[2]) and u[2].tolist() == [0] * 6
    # the running sum follows the clamped deltas: 32000, then 40001 - 32000
    assert u[3].tolist() == [0, 32000, 8001, 0, 0, 0]
    assert st[0] == np.float32(0.25) and st[1] == np.float32(0.5)

    # the torch rule on the same rows: round half to even, INF -> 10.0
    ut, stt, sct = O.compress_torch(rows, np.repeat(src, 4, 0), np.repeat(dst, 4, 0), np.float32(vd))
    ut = ut.view(np.uint16)
    assert ut[0, :4].tolist() == [0, 2, 2, 0]  # 2.5 -> 2, 3.5 -> 4, 4.5 -> 4
    assert ut[0, 4] == 32000  # INF -> 10.0: (10 - 0.25) * 4096 - 4, clamped
    assert stt[2] == np.float32(10.0) and ut[2].tolist() == [0] * 6


def _correctly_rounded_f32(x: Fraction) -> np.float32:
    """The float32 nearest to x (ties to even), by exact comparison with the neighbours."""
    c = np.float32(float(x))
    best = c
    for cand in (np.nextafter(c, np.float32(-np.inf)), np.nextafter(c, np.float32(np.inf))):
        d_best, d_cand = abs(Fraction(float(best)) - x), abs(Fraction(float(cand)) - x)
        if d_cand < d_best or (d_cand == d_best and (int(np.float32(cand).view(np.int32)) & 1) == 0):
            best = cand
    return np.float32(best)


def test_decode_cuda_is_one_rounding_of_the_integer_sum():
    rng = np.random.default_rng(81)
    n, K = 40, 24
    u = rng.integers(0, 32001, (n, K)).astype(np.uint16)
    u[:, 0] = 0
    u[:4, 1:] = 31999  # running sums beyond 2^24, where a float cumsum would round
    st = rng.uniform(0, 1, n).astype(np.float32)
    sc = (rng.uniform(0.5, 1.5, n) * 1e-7).astype(np.float32)
    dec = O.decode_u16_cuda(u, st, sc)
    differs_from_two_roundings = 0
    for r in range(n):
        k = 0
        for i in range(1, K):
            k += int(u[r, i])
            exact = Fraction(float(st[r])) + k * Fraction(float(sc[r]))
            assert dec[r, i] == _correctly_rounded_f32(exact), (r, i)
            two = np.float32(np.float32(np.float32(k) * sc[r]) + st[r])
            differs_from_two_roundings += int(two != dec[r, i])
    assert dec[:, 0].tolist() == st.tolist()
    # the test tells the fused form from a multiply-then-add decode
    assert differs_from_two_roundings > 0


def test_u16_projections_equal_table_projections_of_the_decoded_table():
    rng = np.random.default_rng(82)
    shape, M, minv, b, s = random_geometry(rng, 3, max_n=14, rotated=True)
    src, dst = random_rays(rng, shape, s, 300, "far")
    tv = O.intersections(shape, minv, b, src, dst)
    K = int(np.isfinite(tv).sum(axis=1).max())
    tv = np.ascontiguousarray(tv[:, :K])
    u16, st, sc = O.compress_cuda(tv, src, dst, O.voxel_diag_len(M))
    dec = O.decode_u16_cuda(u16, st, sc)
    vol = rng.uniform(0.5, 1.5, shape).astype(np.float32)
    fu = O.forward_u16(vol, u16, st, sc, minv, b, src, dst)
    ft = O.forward_tvals(vol, dec, minv, b, src, dst)
    assert np.array_equal(fu, ft)
    y = rng.uniform(0, 1, src.shape[0]).astype(np.float32)
    bu = O.back_u16(y, u16, st, sc, shape, minv, b, src, dst)
    bt = O.back_tvals(y, dec, shape, minv, b, src, dst)
    assert np.allclose(bu, bt, rtol=1e-12, atol=1e-12)


def test_u16_consumer_skips_segments_shorter_than_1e9():
    """3d.cu:137: a decoded pair whose world length is below 1e-9 contributes
    nothing, where the float-table consumer (3d.cu:295) adds value * length."""
    shape = (4, 4, 4)
    minv, b = np.eye(3, dtype=np.float32), np.zeros(3, np.float32)
    # near the origin the world points resolve steps of ~1e-11: |dst - src| = 1e-5, inside voxel (0, 0, 0)
    src = np.array([[0.0, 0.0, 0.0]], np.float32)
    dst = np.array([[1e-5, 0.0, 0.0]], np.float32)
    u16 = np.array([[0, 1, 1]], np.uint16)
    st, sc = np.array([0.25], np.float32), np.array([2.0 ** -20], np.float32)  # steps of ~1e-11 in length
    vol = np.zeros(shape, np.float32)
    vol[0, 0, 0] = 1000.0
    assert O.forward_u16(vol, u16, st, sc, minv, b, src, dst)[0] == 0.0
    dec = O.decode_u16_cuda(u16, st, sc)
    ft = O.forward_tvals(vol, dec, minv, b, src, dst)[0]
    assert 1e-9 < ft < 1e-7  # 1000 * 2 * ~1e-11: the float-table consumer counts the two short segments


# ---------------------------------------------------------------------------
# On-the-fly rule against exact arithmetic
# ---------------------------------------------------------------------------


def _rne(q: Fraction) -> int:
    f = math.floor(q)
    r = q - f
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and f % 2 == 1):
        return f + 1
    return f


def _exact_otf_segments(shape, minv, b, src, dst):
    """Exact segments of the on-the-fly rule: crossings in [0, 1] plus 0 and 1,
    consecutive pairs with positive length, voxel = RNE of the exact midpoint,
    length (t1 - t0)|dst - src|; (index list, lengths, degenerate flag)."""
    nd = len(shape)
    F = [[Fraction(float(minv[a, j])) for j in range(nd)] for a in range(nd)]
    sb = [Fraction(float(src[j])) - Fraction(float(b[j])) for j in range(nd)]
    db = [Fraction(float(dst[j])) - Fraction(float(b[j])) for j in range(nd)]
    s = [sum(F[a][j] * sb[j] for j in range(nd)) for a in range(nd)]
    d = [sum(F[a][j] * db[j] for j in range(nd)) for a in range(nd)]
    v = [d[a] - s[a] for a in range(nd)]
    ts = {Fraction(0), Fraction(1)}
    degen = False
    for a in range(nd):
        if abs(v[a]) <= Fraction(1, 10 ** 11):
            degen = degen or v[a] != 0
            continue
        for p in range(shape[a] + 1):
            t = (Fraction(2 * p - 1, 2) - s[a]) / v[a]
            if 0 <= t <= 1:
                ts.add(t)
            if abs(t) < Fraction(1, 10 ** 6) or abs(t - 1) < Fraction(1, 10 ** 6):
                degen = True
    ts = sorted(ts)
    wl = math.sqrt(sum((float(dst[j]) - float(src[j])) ** 2 for j in range(nd)))
    idx, lens = [], []
    for t0, t1 in zip(ts[:-1], ts[1:]):
        if t1 - t0 < Fraction(1, 10 ** 5):
            degen = True
        m = [s[a] + (t0 + t1) / 2 * v[a] for a in range(nd)]
        for a in range(nd):
            if abs(m[a] - math.floor(m[a]) - Fraction(1, 2)) < Fraction(1, 10 ** 3):
                degen = True
        ii = [_rne(m[a]) for a in range(nd)]
        if all(0 <= ii[a] < shape[a] for a in range(nd)):
            flat = ii[0] * shape[1] + ii[1] if nd == 2 else (ii[0] * shape[1] + ii[1]) * shape[2] + ii[2]
            idx.append(flat)
            lens.append(float(t1 - t0) * wl)
    return np.array(idx, np.int64), np.array(lens), degen


@pytest.mark.parametrize("nd", [2, 3])
def test_otf_exact_rational_brute_force(nd):
    rng = np.random.default_rng(90 + nd)
    n_rays = n_checked = 0
    for trial in range(6):
        shape, M, minv, b, s = random_geometry(rng, nd, max_n=7, rotated=(trial % 2 == 1))
        src, dst = random_rays(rng, shape, s, 40, "far" if trial < 3 else "inside")
        got = positive_lists(*O.trace(shape, minv, b, src, dst, mode="otf"))
        for r in range(src.shape[0]):
            ei, el, degen = _exact_otf_segments(shape, minv, b, src[r], dst[r])
            n_rays += 1
            if degen:
                continue
            n_checked += 1
            gi, gl = got[r]
            assert np.array_equal(gi, ei), (trial, r, gi, ei)
            assert np.allclose(gl, el, rtol=1e-4, atol=1e-5)
    # rays inside the box start and end at 0 and 1, so every one has end pieces to check
    assert n_checked >= 0.8 * n_rays


# ---------------------------------------------------------------------------
# 2D on-the-fly march
# ---------------------------------------------------------------------------

E2 = np.eye(2, dtype=np.float32)
Z2 = np.zeros(2, np.float32)


def _march(img, s, d):
    return O.forward(np.asarray(img, np.float32), E2, Z2, np.array([s], np.float32), np.array([d], np.float32),
                     mode="march")[0]


def test_march_known_answers():
    # identity map, b = 0: row/column coordinates are the world coordinates
    col4, col5 = np.zeros((8, 8)), np.zeros((8, 8))
    col4[:, 4] = 1
    col5[:, 5] = 1
    # constant column coordinate 3.7: entry column floor(3.7) = 3, its crossing
    # (3.5) lies behind the entry and is taken first, so the rows are walked in
    # column 4, as round-half-even would place them
    assert abs(_march(col4, (-2, 3.7), (10, 3.7)) - 8.0) < 1e-5
    # at 4.5 the same correction lands in column 5; the torch rule rounds 4.5 to 4
    assert abs(_march(col5, (-2, 4.5), (10, 4.5)) - 8.0) < 1e-5
    assert _march(col4, (-2, 4.5), (10, 4.5)) == 0.0
    t = O.forward(col4.astype(np.float32), E2, Z2, np.array([[-2, 4.5]], np.float32),
                  np.array([[10, 4.5]], np.float32), mode="torch")[0]
    assert abs(t - 8.0) < 1e-5
    # going down the rows: entry row ceil(7.5) - 1 = 7, column floor(2.3) = 2
    col2 = np.zeros((8, 8))
    col2[:, 2] = 1
    assert abs(_march(col2, (10, 2.3), (-2, 2.3)) - 8.0) < 1e-5
    # a ray starting inside the image at row coordinate 3.7: entry row 3, whose
    # crossing 3.5 lies 0.2 behind the start, a step of length -0.2 in pixel
    # (3, 2); pixel (4, 2) then gets 0.8 + 0.2 = 1
    one = np.zeros((8, 8))
    one[3, 2] = 1
    assert abs(_march(one, (3.7, 2.0), (9.7, 2.0)) - (-0.2)) < 1e-6
    one = np.zeros((8, 8))
    one[4, 2] = 1
    assert abs(_march(one, (3.7, 2.0), (9.7, 2.0)) - 1.0) < 1e-6
    # over all pixels the lengths add up to the chord from the start, 7.5 - 3.7
    assert abs(_march(np.ones((8, 8)), (3.7, 2.0), (9.7, 2.0)) - 3.8) < 1e-5
    # the step list visits every in-grid step, in order, zero and negative lengths included
    i, ln, c = O.trace((8, 8), E2, Z2, np.array([[3.7, 2.0]], np.float32), np.array([[9.7, 2.0]], np.float32),
                       mode="march")
    assert i[0, :c[0]].tolist() == [26, 34, 42, 50, 58]
    assert np.allclose(ln[0, :c[0]], [-0.2, 1, 1, 1, 1], atol=1e-6)
    # the reference's closed form: 32 x 32 image, row 16 = 1 (test_ct_projector_2d_line_integral.py:31-43, 70)
    img = np.zeros((32, 32))
    img[16, :] = 1
    assert abs(_march(img, (16.0, -5.0), (16.0, 37.0)) - 32.0) < 1e-4


def test_march_chord_lengths_and_adjoint():
    """A.1 is the chord length through the box (the steps' lengths telescope)
    over rays that enter from outside and whose entry voxel the floor rule gets
    right: the other axis's coordinate at entry has a fractional part below 1/2.
    Above 1/2 the entry voxel is one low on that axis: going up, the first
    crossing lies behind the entry point, and when its step falls outside the
    grid its length is added to the next visited step; going down, every later
    plane is one voxel low as well.  Such rays differ from the chord on this ray
    set.  Forward and back are transposes."""
    rng = np.random.default_rng(95)
    n_ok = n_low = 0
    for trial in range(3):
        shape, M, minv, b, s = random_geometry(rng, 2, max_n=20, rotated=(trial == 2))
        src, dst = random_rays(rng, shape, s, 1500, "far")
        n = np.array(shape, np.float64)
        cs = (src.astype(np.float64) - b) @ minv.astype(np.float64).T
        cd = (dst.astype(np.float64) - b) @ minv.astype(np.float64).T
        v = cd - cs
        with np.errstate(divide="ignore", invalid="ignore"):
            t1 = (-0.5 - cs) / v
            t2 = (n - 0.5 - cs) / v
        lo_t, hi_t = np.minimum(t1, t2), np.maximum(t1, t2)
        t_in, t_out = lo_t.max(axis=1), hi_t.min(axis=1)
        entry_axis = lo_t.argmax(axis=1)
        other = 1 - entry_axis
        rows = np.arange(src.shape[0])
        c_other = cs[rows, other] + t_in * v[rows, other]
        frac = c_other - np.floor(c_other)
        clean = (np.abs(v).min(axis=1) > 1e-3) & (t_in > 1e-3) & (t_out < 1 - 1e-3) & (t_out > t_in + 1e-3)
        clean &= (np.abs(frac - 0.5) > 1e-3) & (frac > 1e-3) & (frac < 1 - 1e-3)
        good = clean & (frac < 0.5)
        low = clean & (frac > 0.5)
        chord = (t_out - t_in) * np.linalg.norm(dst.astype(np.float64) - src.astype(np.float64), axis=1)
        a1 = O.forward(np.ones(shape, np.float32), minv, b, src, dst, mode="march")
        assert np.allclose(a1[good], chord[good], rtol=2e-5, atol=2e-4)
        n_ok += int(good.sum())
        n_low += int((low & ~np.isclose(a1, chord, rtol=1e-3, atol=1e-3)).sum())
        x = rng.standard_normal(shape).astype(np.float32)
        y = rng.standard_normal(src.shape[0]).astype(np.float32)
        ax = O.forward(x, minv, b, src, dst, mode="march")
        aty = O.back(y, shape, minv, b, src, dst, mode="march")
        assert abs(ax @ y - (x * aty).sum()) <= 1e-6 * (np.abs(ax) @ np.abs(y))
    assert n_ok > 200 and n_low > 0
