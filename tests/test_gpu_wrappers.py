"""GPU tests of the ray-subset and multi-rotation wrappers and of the batched
Gram product: each against the same operators applied piece by piece (the
reference's loops), with the kernels' own bars (bit-identical forwards,
back projections within 1e-5, adjoints in FP64 within 1e-5)."""
import numpy as np
import pytest
import torch

from ct_laboratory_amd.gram import gram_matmul
from ct_laboratory_amd.phantoms import ellipsoid_phantom_3d
from ct_laboratory_amd.tomography import (CTProjector3DModule, ConeBeam3DProjector, MultiRotationProjector,
                                          OrderedSubsetProjector, RandomSubsetProjector,
                                          build_conebeam_3d_geometry, precompute_tvals_stitched,
                                          standard_image_transform_3d)
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _table_module(dev, n=16, compress=True):
    src, dst = build_conebeam_3d_geometry(8, 12, 10, 100.0, 150.0, det_spacing=2.0)
    M, b = standard_image_transform_3d(n, n, n, 1.0, 1.0, 1.0)
    return CTProjector3DModule(n, n, n, M, b, src, dst, device=dev, precomputed_intersections=True,
                               use_compression=compress)


def test_ordered_subsets_slice_core_tables(dev):
    core = _table_module(dev)
    osp = OrderedSubsetProjector(core, max_subset_size=300)
    assert osp.n_subsets == int(np.ceil(core.src.shape[0] / 300)) and osp.n_subsets > 1
    vol = torch.from_numpy(ellipsoid_phantom_3d(16, 16, 16, 1.0, seed=3)).to(dev)
    assert torch.equal(osp(vol), core(vol))  # same rows, same kernel: identical line integrals
    y = torch.from_numpy(np.random.default_rng(1).uniform(0, 1, core.src.shape[0]).astype(np.float32)).to(dev)
    assert rel_err(_np(osp.back_project(y)), _np(core.back_project(y))) < 1e-5
    vb = torch.stack([vol, 2 * vol])
    assert torch.equal(osp(vb), core(vb))


def test_ordered_subsets_precompute_and_cache(dev, tmp_path):
    n = 16
    src, dst = build_conebeam_3d_geometry(6, 12, 10, 100.0, 150.0, det_spacing=2.0)
    M, b = standard_image_transform_3d(n, n, n, 1.0, 1.0, 1.0)
    core = CTProjector3DModule(n, n, n, M, b, src, dst, device=dev)  # on-the-fly core: subsets build tables
    osp = OrderedSubsetProjector(core, max_subset_size=250, cache_dir=str(tmp_path))
    vol = torch.from_numpy(ellipsoid_phantom_3d(n, n, n, 1.0, seed=4)).to(dev)
    first = osp(vol)
    files = sorted(p.name for p in tmp_path.iterdir())
    assert len(files) == osp.n_subsets and files[0].startswith("tvals_subset0000_rays0-")
    # a second projector loads the cached tables
    osp2 = OrderedSubsetProjector(core, max_subset_size=250, cache_dir=str(tmp_path))
    assert torch.equal(osp2(vol), first)
    # each subset's table holds the same rows as one table of all the rays
    # (the reference's uint16 format from precompute_tvals_stitched)
    whole = CTProjector3DModule(n, n, n, M, b, src, dst, device=dev,
                                tvals=precompute_tvals_stitched(n, n, n, M, b, src, dst, verbose=False))
    assert torch.equal(first, whole(vol))
    y = torch.from_numpy(np.random.default_rng(2).uniform(0, 1, src.shape[0]).astype(np.float32)).to(dev)
    assert rel_err(_np(osp.back_project(y)), _np(whole.back_project(y))) < 1e-5


def test_random_subset(dev):
    core = _table_module(dev, compress=False)
    rsp = RandomSubsetProjector(core, subset_size=200, seed=5)
    vol = torch.from_numpy(ellipsoid_phantom_3d(16, 16, 16, 1.0, seed=5)).to(dev)
    idx = rsp.indices.to(dev)
    full = CTProjector3DModule(16, 16, 16, core.M, core.b, core.src, core.dst, device=dev)
    assert rel_err(_np(rsp(vol)), _np(full(vol)[idx])) < 1e-6
    rsp.randomize_projector(seed=6)
    assert rsp.indices.numel() == 200


@pytest.mark.parametrize("mode", ["repeat", "reflect", "zeros"])
def test_multirotation_matches_per_rotation_loop(dev, mode):
    nz_win, n_rot, shift = 8, 4, 3.0
    base = ConeBeam3DProjector(16, 16, nz_win, 10, 12, 8, 120.0, 180.0, det_spacing=2.0, voxel_size=1.0,
                               device=dev)
    mr = MultiRotationProjector(base, n_rot, shift, nz_win, padding_mode=mode,
                                reflection_padding_slices=1 if mode == "reflect" else None)
    Z = 14
    vol = torch.from_numpy(np.random.default_rng(2).uniform(0, 1, (16, 16, Z)).astype(np.float32)).to(dev)
    sino = mr(vol)
    assert sino.shape == (base.src.shape[0], n_rot)
    idx, keep, _ = mr._window_index(Z, dev)
    for r in range(n_rot):
        w = vol[..., idx[r]] * keep[r].to(vol.dtype)
        assert torch.equal(sino[:, r], base(w.contiguous()))
    # back projection: the adjoint of the forward in FP64
    y = torch.from_numpy(np.random.default_rng(3).standard_normal(sino.shape).astype(np.float32)).to(dev)
    x = torch.from_numpy(np.random.default_rng(4).standard_normal(vol.shape).astype(np.float32)).to(dev)
    lhs = float((mr(x).double() * y.double()).sum())
    rhs = float((x.double() * mr.back_project(y, Z).double()).sum())
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    # several volumes at once
    outs = mr([vol, 2 * vol])
    assert torch.equal(outs[0], sino) and rel_err(_np(outs[1]), _np(2 * sino)) < 1e-6


def test_multirotation_over_ordered_subsets(dev):
    nz_win = 8
    src, dst = build_conebeam_3d_geometry(6, 12, 8, 120.0, 180.0, det_spacing=2.0)
    M, b = standard_image_transform_3d(16, 16, nz_win, 1.0, 1.0, 1.0)
    core = CTProjector3DModule(16, 16, nz_win, M, b, src, dst, device=dev, precomputed_intersections=True)
    direct = MultiRotationProjector(core, 3, 4.0, nz_win)
    subs = MultiRotationProjector(OrderedSubsetProjector(core, max_subset_size=200), 3, 4.0, nz_win)
    vol = torch.from_numpy(np.random.default_rng(8).uniform(0, 1, (16, 16, 12)).astype(np.float32)).to(dev)
    assert torch.equal(subs(vol), direct(vol))
    y = torch.from_numpy(np.random.default_rng(9).uniform(0, 1, (src.shape[0], 3)).astype(np.float32)).to(dev)
    assert rel_err(_np(subs.back_project(y, 12)), _np(direct.back_project(y, 12))) < 1e-5


def test_gram_matmul_batches_columns(dev):
    proj = ConeBeam3DProjector(12, 12, 12, 8, 10, 10, 100.0, 150.0, det_spacing=2.0, voxel_size=1.0, device=dev)
    N = 12 ** 3
    Q = torch.from_numpy(np.random.default_rng(10).standard_normal((N, 5)).astype(np.float32)).to(dev)
    G = gram_matmul(proj, Q)
    for j in range(5):
        col = proj.back_project(proj.forward_project(Q[:, j].contiguous().reshape(12, 12, 12))).reshape(-1)
        assert rel_err(_np(G[:, j]), _np(col)) < 1e-5
    # symmetric: <G q0, q1> = <q0, G q1>
    a = float((G[:, 0].double() * Q[:, 1].double()).sum())
    c = float((Q[:, 0].double() * G[:, 1].double()).sum())
    assert abs(a - c) / abs(a) < 1e-5
