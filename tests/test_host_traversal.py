"""The product traversal (siddon_traverse.h), compiled for the host, against the
oracle's sort-based restatement: segment lists bit-exact (index and length)."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from helpers import lattice_rays, positive_lists, random_geometry, random_rays

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "_build", "libtrace_host.so")


@pytest.fixture(scope="module")
def host():
    assert os.path.exists(LIB), "run __graft_entry__.build() first"
    L = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    L.host_trace.restype = ctypes.c_int64
    L.host_trace.argtypes = [ctypes.c_int] * 3 + [P] * 5 + [ctypes.c_int64] * 2 + [P] * 3

    def run(shape, minv, b, src, dst, mode, zr):
        nd = len(shape)
        n = np.array(shape, np.int32)
        mi = np.ascontiguousarray(minv, np.float32).reshape(-1)
        bb = np.ascontiguousarray(b, np.float32)
        src = np.ascontiguousarray(src, np.float32)
        dst = np.ascontiguousarray(dst, np.float32)
        N = src.shape[0]
        cap = sum(shape) + nd + 2
        idx = np.zeros((N, cap), np.int64)
        ln = np.zeros((N, cap), np.float32)
        cnt = np.zeros(N, np.int64)
        p = lambda a: a.ctypes.data_as(P)  # noqa: E731
        L.host_trace(nd, {"torch": 0, "otf": 1, "march": 2}[mode], {"thresh": 0, "exact": 1}[zr], p(n), p(mi), p(bb), p(src),
                     p(dst), N, cap, p(idx), p(ln), p(cnt))
        return [(idx[r, :cnt[r]], ln[r, :cnt[r]]) for r in range(N)]

    return run


def _check(host, shape, minv, b, src, dst, mode):
    zr = "exact" if (len(shape) == 2 and mode == "torch") else "thresh"
    got = host(shape, minv, b, src, dst, mode, zr)
    ref = positive_lists(*O.trace(shape, minv, b, src, dst, mode=mode, zero_rule=zr))
    bad = [r for r in range(len(ref)) if len(got[r][0]) != len(ref[r][0]) or (got[r][0] != ref[r][0]).any()
           or (got[r][1].view(np.int32) != ref[r][1].view(np.int32)).any()]
    assert not bad, f"{len(bad)} rays differ (first {bad[:5]})"


@pytest.mark.parametrize("mode", ["torch", "otf"])
@pytest.mark.parametrize("nd", [2, 3])
def test_random_geometries(host, mode, nd):
    rng = np.random.default_rng(11 * nd + (mode == "otf"))
    for trial in range(16):
        shape, M, minv, b, s = random_geometry(rng, nd, rotated=(trial % 4 == 3))
        kind = ["far", "inside", "lattice", "distant"][trial % 4]
        src, dst = random_rays(rng, shape, s, 1500, kind)
        _check(host, shape, minv, b, src, dst, mode)


@pytest.mark.parametrize("mode", ["torch", "otf"])
@pytest.mark.parametrize("nd", [2, 3])
def test_plane_and_corner_rays(host, mode, nd):
    rng = np.random.default_rng(3 + nd)
    shape = (9, 9, 9) if nd == 3 else (13, 11)
    src, dst = lattice_rays(rng, nd, 20000)
    _check(host, shape, np.eye(nd), np.zeros(nd), src, dst, mode)


def _check_march(host, shape, minv, b, src, dst):
    """Every in-grid step of the 2D march (zero and negative lengths included), bit for bit."""
    got = host(shape, minv, b, src, dst, "march", "thresh")
    idx, ln, cnt = O.trace(shape, minv, b, src, dst, mode="march")
    bad = [r for r in range(src.shape[0]) if cnt[r] != len(got[r][0]) or (got[r][0] != idx[r, :cnt[r]]).any()
           or (got[r][1].view(np.int32) != ln[r, :cnt[r]].view(np.int32)).any()]
    assert not bad, f"{len(bad)} rays differ (first {bad[:5]})"
    return int(cnt.sum()), int((ln <= 0).sum() - (cnt.size * ln.shape[1] - cnt.sum()))


def test_march_host_equals_oracle(host):
    """The product's march_ray_2d (host compilation) against the oracle's
    restatement of src/ct_projector_2d.cu:246-413: random maps and rays, and
    half-integer lattice rays (entry coordinates at voxel edges and centres)."""
    rng = np.random.default_rng(71)
    steps = nonpos = 0
    for trial in range(16):
        shape, M, minv, b, s = random_geometry(rng, 2, rotated=(trial % 4 == 3))
        kind = ["far", "inside", "lattice", "distant"][trial % 4]
        src, dst = random_rays(rng, shape, s, 1500, kind)
        a, z = _check_march(host, shape, minv, b, src, dst)
        steps, nonpos = steps + a, nonpos + z
    src, dst = lattice_rays(rng, 2, 20000)
    a, z = _check_march(host, (13, 11), np.eye(2), np.zeros(2), src, dst)
    steps, nonpos = steps + a, nonpos + z
    assert steps > 100000 and nonpos > 0  # the comparison covers zero and negative steps


@pytest.mark.parametrize("mode", ["torch", "otf"])
def test_conebeam_rays(host, mode):
    from ct_laboratory_amd.tomography.conebeam_projector_3d import build_conebeam_3d_geometry

    src, dst = build_conebeam_3d_geometry(10, 40, 24, 1000.0, 1500.0, det_spacing=1.5 * 256 / 40)
    b = np.full(3, -(256 - 1) / 2, np.float32)
    _check(host, (256, 256, 256), np.eye(3), b, src.numpy(), dst.numpy(), mode)


@pytest.fixture(scope="module")
def host_window():
    L = ctypes.CDLL(LIB)
    P = ctypes.c_void_p
    L.host_trace_window.restype = ctypes.c_int64
    L.host_trace_window.argtypes = [ctypes.c_int] * 3 + [P] * 7 + [ctypes.c_int64] * 2 + [P] * 3

    def run(shape, minv, b, lo, ext, src, dst, mode, zr):
        nd = len(shape)
        arrs = [np.array(shape, np.int32), np.ascontiguousarray(minv, np.float32).reshape(-1),
                np.ascontiguousarray(b, np.float32), np.array(lo, np.int32), np.array(ext, np.int32),
                np.ascontiguousarray(src, np.float32), np.ascontiguousarray(dst, np.float32)]
        N = arrs[5].shape[0]
        cap = sum(shape) + nd + 2
        idx = np.zeros((N, cap), np.int64)
        ln = np.zeros((N, cap), np.float32)
        cnt = np.zeros(N, np.int64)
        p = lambda a: a.ctypes.data_as(P)  # noqa: E731
        L.host_trace_window(nd, {"torch": 0, "otf": 1}[mode], {"thresh": 0, "exact": 1}[zr], *[p(a) for a in arrs],
                            N, cap, p(idx), p(ln), p(cnt))
        return [(idx[r, :cnt[r]], ln[r, :cnt[r]]) for r in range(N)]

    return run


def _partition(rng, n, parts):
    cuts = np.sort(rng.choice(np.arange(1, n), size=min(parts - 1, n - 1), replace=False)) if n > 1 else []
    edges = [0] + list(cuts) + [n]
    return [(edges[i], edges[i + 1] - edges[i]) for i in range(len(edges) - 1)]


@pytest.mark.parametrize("mode", ["torch", "otf"])
@pytest.mark.parametrize("nd", [2, 3])
def test_windows_partition_the_traversal(host, host_window, mode, nd):
    """Splitting the grid into windows visits every segment exactly once: the
    per-ray union over the windows equals the full traversal."""
    import itertools

    rng = np.random.default_rng(77 + nd)
    for trial in range(8):
        shape, M, minv, b, s = random_geometry(rng, nd, rotated=(trial == 5))
        kind = ["far", "inside", "lattice", "distant"][trial % 4]
        src, dst = random_rays(rng, shape, s, 600, kind)
        zr = "exact" if (nd == 2 and mode == "torch") else "thresh"
        full = host(shape, minv, b, src, dst, mode, zr)
        parts = [_partition(rng, shape[a], 3) for a in range(nd)]
        got = [([], []) for _ in range(len(src))]
        for combo in itertools.product(*parts):
            lo = [c[0] for c in combo]
            ext = [c[1] for c in combo]
            res = host_window(shape, minv, b, lo, ext, src, dst, mode, zr)
            for r in range(len(src)):
                got[r][0].extend(res[r][0].tolist())
                got[r][1].extend(res[r][1].tolist())
        for r in range(len(src)):
            a = sorted(zip(full[r][0].tolist(), full[r][1].tolist()))
            bb = sorted(zip(got[r][0], got[r][1]))
            assert a == bb, (trial, r)


@pytest.mark.parametrize("nd", [2, 3])
def test_power_of_two_voxel_sizes(host, host_window, nd):
    """Voxel sizes +-2^k make the midpoint scale exact, and the torch-rule
    traversal then forms each rounded midpoint with one fma
    (MidIndex::from_ends_biased).  Index lists and lengths equal the oracle's
    bit for bit, on random and half-integer-lattice rays (midpoints on voxel
    faces, where the rounding ties), over the whole grid and over windows."""
    import itertools

    rng = np.random.default_rng(130 + nd)
    zr = "exact" if nd == 2 else "thresh"
    for trial in range(8):
        shape = tuple(int(x) for x in rng.integers(2, 20, nd))
        s = (2.0 ** rng.integers(-3, 3, nd) * rng.choice([-1.0, 1.0], nd)).astype(np.float32)
        minv = np.diag(1.0 / s).astype(np.float32)
        b = (-(np.array(shape) - 1) / 2 * s + rng.integers(-4, 5, nd) * 0.25).astype(np.float32)
        if trial % 2 == 0:
            src, dst = lattice_rays(rng, nd, 4000, lo=-40, hi=40)
            src, dst = src * np.abs(s), dst * np.abs(s)
        else:
            src, dst = random_rays(rng, shape, np.abs(s), 1500, "far" if trial % 4 == 1 else "inside")
        _check(host, shape, minv, b, src, dst, "torch")
        full = host(shape, minv, b, src, dst, "torch", zr)
        parts = [_partition(rng, shape[a], 2) for a in range(nd)]
        got = [[] for _ in range(len(src))]
        for combo in itertools.product(*parts):
            res = host_window(shape, minv, b, [c[0] for c in combo], [c[1] for c in combo], src, dst, "torch", zr)
            for r in range(len(src)):
                got[r].extend(zip(res[r][0].tolist(), res[r][1].tolist()))
        for r in range(len(src)):
            assert sorted(zip(full[r][0].tolist(), full[r][1].tolist())) == sorted(got[r]), (trial, r)


def test_half_and_twice_exact_on_every_float():
    """The traversal's integer forms of 0.5*x and 2*x (used for its midpoint
    constants) equal the FPU products bit for bit on all 2^32 patterns."""
    assert os.path.exists(LIB), "run __graft_entry__.build() first"
    L = ctypes.CDLL(LIB)
    L.host_scale_mismatches.restype = ctypes.c_int64
    L.host_scale_mismatches.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    assert L.host_scale_mismatches(0, 1 << 32) == 0
