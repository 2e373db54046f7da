"""Shared test helpers: random ray sets and comparisons."""
import numpy as np


def random_geometry(rng, nd, max_n=24, rotated=False):
    shape = tuple(int(x) for x in rng.integers(1, max_n, nd))
    s = rng.uniform(0.3, 2.0, nd).astype(np.float32)
    M = np.diag(s).astype(np.float64)
    if rotated:
        th = rng.uniform(0, np.pi)
        R = np.eye(nd)
        R[0, 0] = R[1, 1] = np.cos(th)
        R[0, 1], R[1, 0] = -np.sin(th), np.sin(th)
        M = R @ M
    M = M.astype(np.float32)
    b = (-(np.array(shape) - 1) / 2 * s + rng.uniform(-2, 2, nd)).astype(np.float32)
    minv = np.linalg.inv(M.astype(np.float64)).astype(np.float32)
    return shape, M, minv, b, s


def random_rays(rng, shape, s, n, kind):
    nd = len(shape)
    ext = np.array(shape) * s
    if kind == "far":
        src = rng.uniform(-3, 3, (n, nd)) * ext
        dst = rng.uniform(-3, 3, (n, nd)) * ext
    elif kind == "distant":
        # endpoints ~4e4 voxels away on opposite sides, lines through
        # the grid: midpoint error bounds above one half voxel
        c = rng.uniform(-0.6, 0.6, (n, nd)) * ext
        u = rng.normal(size=(n, nd))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        D = 40000.0 * float(np.max(s))  # ~4e4 voxels: E near 1.2
        src = c + u * D
        dst = c - u * D + rng.uniform(-0.5, 0.5, (n, nd)) * ext
    elif kind == "inside":
        src = rng.uniform(-0.6, 0.6, (n, nd)) * ext
        dst = rng.uniform(-0.6, 0.6, (n, nd)) * ext
    else:  # half-unit lattice endpoints, some axis-parallel
        src = np.round(rng.uniform(-1, 1, (n, nd)) * ext * 2) / 2
        dst = np.round(rng.uniform(-1, 1, (n, nd)) * ext * 2) / 2
        ax = rng.integers(0, nd, n)
        dst[np.arange(n), ax] = src[np.arange(n), ax]
    return src.astype(np.float32), dst.astype(np.float32)


def lattice_rays(rng, nd, n, lo=-8, hi=24):
    """Identity-grid rays whose endpoints sit on half-integers: on-plane,
    through-corner and axis-parallel cases."""
    src = rng.integers(lo, hi, (n, nd)).astype(np.float32) / 2
    dst = rng.integers(lo, hi, (n, nd)).astype(np.float32) / 2
    m = rng.random(n) < 0.4
    ax = rng.integers(0, nd, n)
    dst[m, ax[m]] = src[m, ax[m]]
    return src, dst


def rel_err(got, ref, floor_frac=1e-3):
    """max |got-ref| / max(|ref|, floor_frac * max|ref|)."""
    got = np.asarray(got, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if ref.size == 0:
        return 0.0
    scale = np.maximum(np.abs(ref), floor_frac * max(float(np.abs(ref).max()), 1e-30))
    return float((np.abs(got - ref) / scale).max())


def positive_lists(idx, ln, cnt):
    """Per-ray arrays of (idx, len) restricted to len > 0 from padded tables."""
    out = []
    for r in range(idx.shape[0]):
        k = int(cnt[r])
        sel = ln[r, :k] > 0
        out.append((np.asarray(idx[r, :k])[sel].astype(np.int64), np.asarray(ln[r, :k])[sel]))
    return out
