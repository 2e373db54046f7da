"""The reference's torch-projector names (ct_projector_{2d,3d}_torch.py), run by
the HIP kernels, against the oracle: crossing tables and the torch
compression bit for bit, table, compressed-table and table-free projections
within 1e-5, and the autograd modules' gradients equal to the back
projection."""
import numpy as np
import pytest
import torch

import oracle as O
from ct_laboratory_amd import _C
from ct_laboratory_amd.tomography import ct_projector_2d_torch as T2
from ct_laboratory_amd.tomography import ct_projector_3d_torch as T3
from helpers import random_geometry, random_rays, rel_err

pytestmark = pytest.mark.gpu


def _t(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)


@pytest.mark.parametrize("chunk_bytes", [1 << 30, 4096])
def test_3d_torch_names(dev, monkeypatch, chunk_bytes):
    monkeypatch.setattr(T3, "_CHUNK_BYTES", chunk_bytes)  # 4096: many decompression chunks
    rng = np.random.default_rng(120)
    shape, M, minv, b, s = random_geometry(rng, 3, max_n=14, rotated=True)
    src, dst = random_rays(rng, shape, s, 500, "far")
    tm, tb, ts, td = _t(M, dev), _t(b, dev), _t(src, dev), _t(dst, dev)
    minv = _C.cached_inverse(tm).cpu().numpy()  # the inverse these functions compute from M, as the reference does
    tv = T3.compute_intersections_3d_torch(*shape, tm, tb, ts, td)
    ref_tv = O.intersections(shape, minv, b, src, dst, zero_rule="thresh")
    assert np.array_equal(tv.cpu().numpy().view(np.int32), ref_tv.view(np.int32))
    vol = rng.uniform(0, 1, shape).astype(np.float32)
    y = rng.uniform(0, 1, src.shape[0]).astype(np.float32)
    f = T3.forward_project_3d_torch(_t(vol, dev), tv, tm, tb, ts, td)
    assert f.shape == (src.shape[0],)
    assert rel_err(f.cpu().numpy(), O.forward_tvals(vol, ref_tv, minv, b, src, dst)) < 1e-5
    bk = T3.back_project_3d_torch(_t(y, dev), tv, tm, tb, ts, td, *shape)
    assert bk.shape == shape
    assert rel_err(bk.cpu().numpy(), O.back_tvals(y, ref_tv, shape, minv, b, src, dst)) < 1e-5
    # the torch compression rule, and projections through its float decode
    u16, st, sc = T3.compress_tvals_to_uint16(tv, tm, ts, td)
    ou, ost, osc = O.compress_torch(ref_tv, src, dst, O.voxel_diag_len(M))
    assert np.array_equal(u16.cpu().numpy(), ou) and np.array_equal(st.cpu().numpy(), ost)
    assert np.array_equal(sc.cpu().numpy(), osc)
    dec = T3.decompress_tvals_from_uint16(u16, st, sc).cpu().numpy()
    fc = T3.forward_project_3d_torch(_t(vol, dev), (u16, st, sc), tm, tb, ts, td)
    assert rel_err(fc.cpu().numpy(), O.forward_tvals(vol, dec, minv, b, src, dst)) < 1e-5
    bc = T3.back_project_3d_torch(_t(y, dev), (u16, st, sc), tm, tb, ts, td, *shape)
    assert rel_err(bc.cpu().numpy(), O.back_tvals(y, dec, shape, minv, b, src, dst)) < 1e-5
    # no table: the on-the-fly traversal with the same rule
    fo = T3.forward_project_3d_torch(_t(vol, dev), None, tm, tb, ts, td)
    assert rel_err(fo.cpu().numpy(), O.forward(vol, minv, b, src, dst, mode="torch")) < 1e-5
    # the module's gradient is the back projection
    proj = T3.CTProjector3D(tm, tb, ts, td)
    x = _t(vol, dev).requires_grad_(True)
    (proj(x) * _t(y, dev)).sum().backward()
    assert rel_err(x.grad.cpu().numpy(), O.back(y, shape, minv, b, src, dst, mode="torch")) < 1e-5


def test_2d_torch_names(dev):
    rng = np.random.default_rng(121)
    shape, M, minv, b, s = random_geometry(rng, 2, max_n=30, rotated=True)
    src, dst = random_rays(rng, shape, s, 800, "far")
    tm, tb, ts, td = _t(M, dev), _t(b, dev), _t(src, dev), _t(dst, dev)
    minv = _C.cached_inverse(tm).cpu().numpy()
    tv = T2.compute_intersections_2d_torch(*shape, tm, tb, ts, td)
    ref_tv = O.intersections(shape, minv, b, src, dst, zero_rule="exact")
    assert np.array_equal(tv.cpu().numpy().view(np.int32), ref_tv.view(np.int32))
    img = rng.uniform(0, 1, shape).astype(np.float32)
    y = rng.uniform(0, 1, src.shape[0]).astype(np.float32)
    f = T2.forward_project_2d_torch(_t(img, dev), tv, tm, tb, ts, td)
    assert f.shape == (src.shape[0],)
    assert rel_err(f.cpu().numpy(), O.forward_tvals(img, ref_tv, minv, b, src, dst)) < 1e-5
    bk = T2.back_project_2d_torch(_t(y, dev), tv, tm, tb, ts, td, *shape)
    assert bk.shape == shape
    assert rel_err(bk.cpu().numpy(), O.back_tvals(y, ref_tv, shape, minv, b, src, dst)) < 1e-5
    proj = T2.CTProjector2D(tm, tb, ts, td)
    x = _t(img, dev).requires_grad_(True)
    out = proj(x)
    assert rel_err(out.detach().cpu().numpy(), O.forward(img, minv, b, src, dst, mode="torch", zero_rule="exact")) < 1e-5
    (out * _t(y, dev)).sum().backward()
    assert rel_err(x.grad.cpu().numpy(),
                   O.back(y, shape, minv, b, src, dst, mode="torch", zero_rule="exact")) < 1e-5
