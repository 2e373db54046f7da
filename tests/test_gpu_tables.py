"""GPU parity of the crossing-table paths: compute_intersections_{2d,3d},
float-table and uint16-table projectors, compress_tvals_3d."""
import numpy as np
import pytest
import torch

import oracle as O
from ct_laboratory_amd import _C
from ct_laboratory_amd import _native as N
from ct_laboratory_amd.tomography.conebeam_projector_3d import build_conebeam_3d_geometry
from helpers import lattice_rays, random_geometry, random_rays, rel_err

pytestmark = pytest.mark.gpu


def _t(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)


def _same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and (a.view(np.int32) == b.view(np.int32)).all()


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("zero_rule", [N.ZERO_THRESH, N.ZERO_EXACT])
def test_intersections_bit_exact(dev, nd, zero_rule):
    rng = np.random.default_rng(20 + nd)
    zr = "exact" if zero_rule == N.ZERO_EXACT else "thresh"
    for trial in range(4):
        shape, M, minv, b, s = random_geometry(rng, nd, rotated=(trial == 3))
        src, dst = random_rays(rng, shape, s, 2000, ["far", "inside", "lattice", "far"][trial])
        fn = _C.compute_intersections_3d if nd == 3 else _C.compute_intersections_2d
        got = fn(*shape, _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev), zero_rule=zero_rule).cpu().numpy()
        ref = O.intersections(shape, minv, b, src, dst, zero_rule=zr)
        # -0.0 and +0.0 compare equal and may sort either way; compare as values there.
        assert got.shape == ref.shape
        assert _same_bits(got + 0.0, ref + 0.0)


@pytest.mark.parametrize("nd", [2, 3])
def test_table_consumer_equals_on_the_fly(dev, nd):
    """Full table + table projector == on-the-fly projector (torch semantics),
    bit for bit: same segments, same accumulation order."""
    rng = np.random.default_rng(30 + nd)
    shape, M, minv, b, s = random_geometry(rng, nd, max_n=30)
    src, dst = random_rays(rng, shape, s, 4000, "far")
    zr = N.ZERO_THRESH
    ts, td, tm, tb = _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev)
    vol = _t(rng.uniform(0, 1, (2,) + shape), dev)
    sino = _t(rng.uniform(0, 1, (2, src.shape[0])), dev)
    if nd == 3:
        tab = _C.compute_intersections_3d(*shape, ts, td, tm, tb, zero_rule=zr)
        f_tab = _C.forward_project_3d_cuda(vol, tab, ts, td, tm, tb)
        f_otf = _C.forward_project_3d_on_the_fly_cuda(vol, ts, td, tm, tb, mode=N.MODE_TORCH)
        b_tab = _C.back_project_3d_cuda(sino, tab, ts, td, tm, tb, *shape)
        b_otf = _C.back_project_3d_on_the_fly_cuda(sino, ts, td, tm, tb, *shape, mode=N.MODE_TORCH)
    else:
        own = _C.compute_intersections_2d(*shape, ts, td, tm, tb, zero_rule=zr)
        tab = own.clone()
        _C.AUTO_LAYOUT = False  # the table kernels
        try:
            f_tab = _C.forward_project_2d_cuda(vol, tab, ts, td, tm, tb)
            b_tab = _C.back_project_2d_cuda(sino, tab, ts, td, tm, tb, *shape)
        finally:
            _C.AUTO_LAYOUT = True
        f_otf = _C.forward_project_2d_on_the_fly_cuda(vol, ts, td, tm, tb, mode=N.MODE_TORCH, zero_rule=zr)
        b_otf = _C.back_project_2d_on_the_fly_cuda(sino, ts, td, tm, tb, *shape, mode=N.MODE_TORCH, zero_rule=zr)
        # the table as returned runs on the fly with its own rule
        assert torch.equal(_C.forward_project_2d_cuda(vol, own, ts, td, tm, tb), f_otf)
        assert torch.equal(_C.back_project_2d_cuda(sino, own, ts, td, tm, tb, *shape), b_otf)
        # a copy is compared once with the table computed here, found equal and run on the fly;
        # compared in chunks of rays, a difference in the last chunk is found
        saved = _C._TABLE2D_COMPARE_BYTES
        _C._TABLE2D_COMPARE_BYTES = 4 * tab.shape[1] * 300
        try:
            late = tab.clone()
            with torch.no_grad():
                late[-1, 0] = float("nan")
            assert _C._table2d_rule(late, ts, td, tm, tb, *shape) is None
            assert _C._table2d_rule(tab, ts, td, tm, tb, *shape) == zr
        finally:
            _C._TABLE2D_COMPARE_BYTES = saved
        assert torch.equal(_C.forward_project_2d_cuda(vol, tab, ts, td, tm, tb), f_otf)
        assert torch.equal(_C.back_project_2d_cuda(sino, tab, ts, td, tm, tb, *shape), b_otf)
        r = int(torch.nonzero(torch.isfinite(own[:, 1]))[0])
        with torch.no_grad():
            own[r, 0] = float("inf")  # written: differs from the computed table, consumed as given
        assert _C._table2d_rule(own, ts, td, tm, tb, *shape) is None
        _C.AUTO_LAYOUT = False
        try:
            f_edit = _C.forward_project_2d_cuda(vol, own, ts, td, tm, tb)
        finally:
            _C.AUTO_LAYOUT = True
        assert torch.equal(_C.forward_project_2d_cuda(vol, own, ts, td, tm, tb), f_edit)
        # each version is compared once; after two differing versions the table is not compared again
        with torch.no_grad():
            own[r, 1] = float("inf")
        assert _C._table2d_rule(own, ts, td, tm, tb, *shape) is None
        assert _C._TABLE2D_MISSES[id(own)][1] == 2
        with torch.no_grad():
            own.copy_(_C.compute_intersections_2d(*shape, ts, td, tm, tb, zero_rule=zr))
        assert _C._table2d_rule(own, ts, td, tm, tb, *shape) is None
    assert torch.equal(f_tab, f_otf)
    assert rel_err(b_tab.cpu().numpy(), b_otf.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("nd", [2, 3])
def test_table_projectors_vs_oracle(dev, nd):
    rng = np.random.default_rng(40 + nd)
    shape, M, minv, b, s = random_geometry(rng, nd, max_n=30)
    src, dst = lattice_rays(rng, nd, 3000) if nd == 2 else random_rays(rng, shape, s, 3000, "inside")
    if nd == 2:
        shape, minv, b = (13, 11), np.eye(2, dtype=np.float32), np.zeros(2, np.float32)
    zr = "exact" if nd == 2 else "thresh"
    tab = O.intersections(shape, minv, b, src, dst, zero_rule=zr)
    vol = rng.uniform(0, 1, (2,) + shape).astype(np.float32)
    sino = rng.uniform(0, 1, (2, src.shape[0])).astype(np.float32)
    ts, td, tm, tb, tt = _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev), _t(tab, dev)
    if nd == 3:
        f = _C.forward_project_3d_cuda(_t(vol, dev), tt, ts, td, tm, tb)
        bk = _C.back_project_3d_cuda(_t(sino, dev), tt, ts, td, tm, tb, *shape)
    else:
        _C.AUTO_LAYOUT = False  # the table kernels
        try:
            f = _C.forward_project_2d_cuda(_t(vol, dev), tt, ts, td, tm, tb)
            bk = _C.back_project_2d_cuda(_t(sino, dev), tt, ts, td, tm, tb, *shape)
        finally:
            _C.AUTO_LAYOUT = True
        # the oracle's table equals the one computed here: the same call runs on the fly, on tiles
        assert _C._table2d_rule(tt, ts, td, tm, tb, *shape) == N.ZERO_THRESH
        f2 = _C.forward_project_2d_cuda(_t(vol, dev), tt, ts, td, tm, tb)
        bk2 = _C.back_project_2d_cuda(_t(sino, dev), tt, ts, td, tm, tb, *shape)
        assert torch.equal(f2, f)
        assert rel_err(bk2.cpu().numpy(), O.back_tvals(sino, tab, shape, minv, b, src, dst)) < 1e-5
    assert rel_err(f.cpu().numpy(), O.forward_tvals(vol, tab, minv, b, src, dst)) < 1e-5
    assert rel_err(bk.cpu().numpy(), O.back_tvals(sino, tab, shape, minv, b, src, dst)) < 1e-5


@pytest.mark.parametrize("semantics,name", [(N.COMPRESS_CUDA, "cuda"), (N.COMPRESS_TORCH, "torch")])
def test_compress_bit_exact(dev, semantics, name):
    n = 40
    src, dst = build_conebeam_3d_geometry(9, 33, 21, 1000.0, 1500.0, det_spacing=3.5)
    b = np.full(3, -(n - 1) / 2 * 1.25, np.float32)
    M = np.eye(3, dtype=np.float32) * np.float32(1.25)
    minv = np.linalg.inv(M.astype(np.float64)).astype(np.float32)
    tab = O.intersections((n, n, n), minv, b, src.numpy(), dst.numpy())
    vdl = O.voxel_diag_len(M)
    u, st, sc = _C.compress_tvals_3d_cuda(_t(tab, dev), src.to(dev), dst.to(dev), float(vdl), semantics=semantics)
    ref = (O.compress_cuda if name == "cuda" else O.compress_torch)(tab, src.numpy(), dst.numpy(), vdl)
    assert (u.cpu().numpy().view(np.uint16) == ref[0].view(np.uint16)).all()
    assert _same_bits(st.cpu().numpy(), ref[1]) and _same_bits(sc.cpu().numpy(), ref[2])


def test_compressed_projectors_vs_oracle(dev):
    n = 40
    src, dst = build_conebeam_3d_geometry(9, 33, 21, 1000.0, 1500.0, det_spacing=3.5)
    b = np.full(3, -(n - 1) / 2 * 1.25, np.float32)
    M = np.eye(3, dtype=np.float32) * np.float32(1.25)
    minv = np.linalg.inv(M.astype(np.float64)).astype(np.float32)
    s_np, d_np = src.numpy(), dst.numpy()
    tab = O.intersections((n, n, n), minv, b, s_np, d_np)
    u, st, sc = O.compress_cuda(tab, s_np, d_np, O.voxel_diag_len(M))
    rng = np.random.default_rng(3)
    vol = rng.uniform(0, 1, (2, n, n, n)).astype(np.float32)
    sino = rng.uniform(0, 1, (2, s_np.shape[0])).astype(np.float32)
    tu = torch.from_numpy(u.view(np.int16)).to(dev)
    ts, td, tm, tb = src.to(dev), dst.to(dev), _t(minv, dev), _t(b, dev)
    f = _C.forward_project_3d_compressed_cuda(_t(vol, dev), tu, _t(st, dev), _t(sc, dev), ts, td, tm, tb)
    bk = _C.back_project_3d_compressed_cuda(_t(sino, dev), tu, _t(st, dev), _t(sc, dev), ts, td, tm, tb, n, n, n)
    assert rel_err(f.cpu().numpy(), O.forward_u16(vol, u, st, sc, minv, b, s_np, d_np)) < 1e-5
    assert rel_err(bk.cpu().numpy(), O.back_u16(sino, u, st, sc, (n, n, n), minv, b, s_np, d_np)) < 1e-5
    # Against the uncompressed table: decode error is at most half a quantum per
    # crossing, on rows where no delta hit the 32000 clamp.  (Rays crossing
    # planes far outside the grid, e.g. z-planes along a whole low-slope
    # cone-beam ray, clamp and then lag: that is the reference's format.)
    f_full = O.forward_tvals(vol, tab, minv, b, s_np, d_np)
    ok = (u.max(axis=1) < 32000) & (f_full[0] > 0)
    assert ok.sum() > 100
    assert rel_err(f.cpu().numpy()[:, ok], f_full[:, ok]) < 2e-3


def test_count_crossings(dev):
    rng = np.random.default_rng(9)
    shape, M, minv, b, s = random_geometry(rng, 3)
    src, dst = random_rays(rng, shape, s, 2000, "far")
    cnt = _C.count_crossings(shape, _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev)).cpu().numpy()
    ref = np.isfinite(O.intersections(shape, minv, b, src, dst)).sum(axis=1)
    assert (cnt == ref).all()


@pytest.mark.parametrize("nd", [2, 3])
def test_table_inf_across_stage_boundaries(dev, nd):
    """The table kernels read rows through LDS 16 entries at a time and carry
    each row's walk across the chunks: INF entries placed at and around chunk
    boundaries (a pair with an INF end is skipped, the walk resumes after it)
    give the oracle's line integrals and back projections."""
    rng = np.random.default_rng(70 + nd)
    shape, M, minv, b, s = random_geometry(rng, nd, max_n=30)
    src, dst = random_rays(rng, shape, s, 1500, "far")
    tab = O.intersections(shape, minv, b, src, dst, zero_rule="thresh")
    K = tab.shape[1]
    for r in range(0, src.shape[0], 3):
        for k in rng.choice([15, 16, 17, 31, 32, 33], size=2, replace=False):
            if k < K:
                tab[r, k] = np.inf
    vol = rng.uniform(0, 1, (1,) + shape).astype(np.float32)
    sino = rng.uniform(-1, 1, (1, src.shape[0])).astype(np.float32)
    ts, td, tm, tb, tt = _t(src, dev), _t(dst, dev), _t(minv, dev), _t(b, dev), _t(tab, dev)
    fwd = _C.forward_project_3d_cuda if nd == 3 else _C.forward_project_2d_cuda
    back = _C.back_project_3d_cuda if nd == 3 else _C.back_project_2d_cuda
    _C.AUTO_LAYOUT = False  # the table kernels, whatever the table
    try:
        f = fwd(_t(vol, dev), tt, ts, td, tm, tb)
        bk = back(_t(sino, dev), tt, ts, td, tm, tb, *shape)
    finally:
        _C.AUTO_LAYOUT = True
    assert rel_err(f.cpu().numpy(), O.forward_tvals(vol, tab, minv, b, src, dst)) < 1e-5
    # signed data: voxel sums may cancel, so the error is bounded by the largest value
    ref = O.back_tvals(sino, tab, shape, minv, b, src, dst)
    assert np.abs(bk.cpu().numpy() - ref).max() <= 1e-5 * np.abs(ref).max()
