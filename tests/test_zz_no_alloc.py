"""The brick back projection's entry point that allocates nothing
(ctlab_back_project_3d_brick_noalloc, per-view maxima in a static device
array), called through the C ABI, against the entry point that takes a
workspace: bit-identical results on consecutive calls, the second on a smaller
input, then an accumulating call; sizes beyond the static array are rejected
before anything is launched."""
import numpy as np
import pytest
import torch

from ct_laboratory_amd import _C
from ct_laboratory_amd import _native as N
from ct_laboratory_amd.tomography.conebeam_projector_3d import build_conebeam_3d_geometry

pytestmark = pytest.mark.gpu


def _case(dev, shape, n_view, det, seed, batch):
    rows, cols = det
    src, dst = build_conebeam_3d_geometry(n_view, cols, rows, 600.0, 900.0, det_spacing=2.0, angle_offset=0.1 * seed)
    ts, td = src.to(dev).contiguous(), dst.to(dev).contiguous()
    b = torch.tensor([-(n - 1) / 2 for n in shape], dtype=torch.float32, device=dev)
    m = torch.eye(3, device=dev)
    plan = _C.brick_plan(_C.cone_frames(ts, td, det, m, b, shape), det, shape)
    g = torch.Generator().manual_seed(seed)
    y = torch.randn((batch, ts.shape[0]), generator=g).to(dev)
    return ts, td, b, m, plan, y


def _noalloc(y, ts, td, det, plan, m, b, shape, out, mode, accumulate=0):
    n_view = ts.shape[0] // (det[0] * det[1])
    N.call("ctlab_back_project_3d_brick_noalloc", N.ptr(y), int(y.shape[0]), N.ptr(ts), N.ptr(td), n_view, det[0],
           det[1], N.ptr(plan), N.ptr(m), N.ptr(b), *shape, mode, accumulate, N.ptr(out), N.stream_of(y))
    return out


@pytest.mark.parametrize("mode", [N.MODE_TORCH, N.MODE_OTF])
def test_noalloc_brick_matches_workspace_entry(dev, mode):
    big = ((50, 42, 36), 12, (30, 44), 1, 2)
    small = ((33, 29, 20), 5, (18, 24), 2, 1)
    for shape, n_view, det, seed, batch in (big, small):
        print(f"[no_alloc] mode {mode} grid {shape} views {n_view} det {det} batch {batch}", flush=True)
        ts, td, b, m, plan, y = _case(dev, shape, n_view, det, seed, batch)
        ref = _C.back_project_3d_brick(y, ts, td, det, plan, m, b, *shape, mode=mode)
        out = torch.full((batch,) + shape, 7.0, device=dev)  # overwritten, not added to
        got = _noalloc(y, ts, td, det, plan, m, b, shape, out, mode)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
    # accumulate into an existing output, as the workspace entry point does
    shape, n_view, det, seed, batch = small
    ts, td, b, m, plan, y = _case(dev, shape, n_view, det, seed, batch)
    base = torch.randn((batch,) + shape, generator=torch.Generator().manual_seed(5)).to(dev)
    ref = _C.back_project_3d_brick(y, ts, td, det, plan, m, b, *shape, mode=mode, out=base.clone(), accumulate=True)
    got = _noalloc(y, ts, td, det, plan, m, b, shape, base.clone(), mode, accumulate=1)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)


def test_noalloc_rejects_sizes_beyond_the_static_array(dev):
    L = N.load()
    # batch * n_view = 2 * 40000 > 65536: refused on the host, nothing launched
    st = L.ctlab_back_project_3d_brick_noalloc(None, 2, None, None, 40000, 4, 4, None, None, None, 8, 8, 8,
                                               N.MODE_TORCH, 0, None, None)
    assert st != 0 and b"static workspace" in L.ctlab_last_error()
