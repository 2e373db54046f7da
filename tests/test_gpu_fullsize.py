"""GPU parity at the benchmark configurations' real sizes, against the oracle.

* The 256^3 cone-beam workload (1 mm voxels, 720 views, 384 x 384 detector at
  1.5 mm, SID 1000 mm, SDD 1500 mm; 106,168,320 rays): segment lists of a
  50,000-ray sample bit-exact; the detector-tiled forward of all 720 views, and
  the brick back projection with the full 720-view plan, within 1e-5 of the
  oracle on 8 evenly spaced views (the oracle cannot run all 106 M rays in a
  test).  The back projection's sinogram is nonzero on those 8 views only, so
  every other view's rays are skipped by value while the plan, the error
  bounds, the view-half work units and the window clipping are those of the
  full workload.
* The served stream's smallest requests at their real sizes (a 256^2 fan-beam
  request, 360 views x 512 bins, and a 128^3 cone-beam request, 180 views x
  256^2 panel, 20 iterations each): alone equal to in the mix, and the first
  iteration's back projection of the residual against the oracle's.

Bar: voxel index lists bit-exact (lengths within 3 ulp); projections within
1e-5 relative (floor 1e-3 of the max); back projections of signed data within
1e-5 of the back projection of their magnitudes, voxel by voxel.
"""
import numpy as np
import pytest
import torch

import oracle as O
from ct_laboratory_amd import _C
from ct_laboratory_amd import _native as N
from ct_laboratory_amd.phantoms import ellipsoid_phantom_3d
from ct_laboratory_amd.tomography.conebeam_projector_3d import build_conebeam_3d_geometry
from helpers import positive_lists, rel_err

pytestmark = pytest.mark.gpu

N_VOX, N_VIEW, DET, PITCH, SID, SDD = 256, 720, 384, 1.5, 1000.0, 1500.0
SAMPLE_VIEWS = np.linspace(0, N_VIEW, 8, endpoint=False).astype(int)


@pytest.fixture(scope="module")
def cone256(dev):
    print(f"[fullsize] building the {N_VOX}^3 / {N_VIEW}-view / {DET}^2 cone-beam rays", flush=True)
    src, dst = build_conebeam_3d_geometry(N_VIEW, DET, DET, SID, SDD, det_spacing=PITCH)
    b = np.full(3, -(N_VOX - 1) / 2.0, np.float32)
    minv = np.eye(3, dtype=np.float32)
    vol = ellipsoid_phantom_3d(N_VOX, N_VOX, N_VOX, voxel_size=1.0, seed=1)
    per_view = DET * DET
    sel = np.concatenate([np.arange(v * per_view, (v + 1) * per_view) for v in SAMPLE_VIEWS])
    g = dict(src=src, dst=dst, b=b, minv=minv, vol=vol, sel=sel,
             ts=src.to(dev), td=dst.to(dev), tb=torch.from_numpy(b).to(dev), tm=torch.from_numpy(minv).to(dev),
             tv=torch.from_numpy(vol).to(dev))
    yield g
    del g
    torch.cuda.empty_cache()


def test_c2_segment_lists_sample(dev, cone256):
    g = cone256
    rng = np.random.default_rng(31)
    pick = np.sort(rng.choice(g["src"].shape[0], 50_000, replace=False))
    s, d = g["src"].numpy()[pick], g["dst"].numpy()[pick]
    shape = (N_VOX,) * 3
    print("[fullsize] segment lists of 50,000 sampled rays", flush=True)
    idx, ln, cnt = _C.trace_segments(shape, torch.from_numpy(s).to(dev), torch.from_numpy(d).to(dev), g["tm"], g["tb"],
                                     mode=N.MODE_TORCH)
    gl = positive_lists(idx.cpu().numpy(), ln.cpu().numpy(), cnt.cpu().numpy())
    oi, ol, oc = O.trace(shape, g["minv"], g["b"], s, d, mode="torch")
    ol_ = positive_lists(oi, ol, oc)
    bad = [r for r in range(len(gl)) if len(gl[r][0]) != len(ol_[r][0]) or (gl[r][0] != ol_[r][0]).any()
           or (np.abs(gl[r][1].view(np.int32) - ol_[r][1].view(np.int32)) > 3).any()]
    assert not bad, f"{len(bad)} of 50000 rays differ, first {bad[:5]}"
    assert sum(len(x[0]) for x in gl) > 50_000 * 100  # most sampled rays cross the volume


@pytest.mark.parametrize("mname,mode", [("torch", N.MODE_TORCH), ("otf", N.MODE_OTF)])
def test_c2_tiled_forward_and_brick_back(dev, cone256, mname, mode):
    g = cone256
    det = (DET, DET)
    shape = (N_VOX,) * 3
    print(f"[fullsize] {mname}: tiled forward of all {N_VIEW} views", flush=True)
    sino = _C.forward_project_3d_on_the_fly_cuda(g["tv"], g["ts"], g["td"], g["tm"], g["tb"], mode=mode, det_shape=det)
    got_f = sino[0].cpu().numpy()[g["sel"]]
    s, d = g["src"].numpy()[g["sel"]], g["dst"].numpy()[g["sel"]]
    ref_f = O.forward(g["vol"], g["minv"], g["b"], s, d, mode=mname)
    assert rel_err(got_f, ref_f) < 1e-5

    print(f"[fullsize] {mname}: brick back projection, full {N_VIEW}-view plan", flush=True)
    frames = _C.cone_frames(g["ts"], g["td"], det, g["tm"], g["tb"], shape)
    assert int((frames[:, 15] == 1).sum()) == N_VIEW
    plan = _C.brick_plan(frames, det, shape)
    y = np.zeros(g["src"].shape[0], np.float32)
    ys = np.random.default_rng(32).standard_normal(len(g["sel"])).astype(np.float32)
    y[g["sel"]] = ys
    ty = torch.from_numpy(y).to(dev)[None]
    got_b = _C.back_project_3d_brick(ty, g["ts"], g["td"], det, plan, g["tm"], g["tb"], *shape, mode=mode)
    ref_b = O.back(ys[None], shape, g["minv"], g["b"], s, d, mode=mname)
    # signed data cancel in a voxel's sum: the rounding bound scales with the back
    # projection of |y| (the sum of the magnitudes of the voxel's terms)
    bound = O.back(np.abs(ys)[None], shape, g["minv"], g["b"], s, d, mode=mname)
    err = np.abs(got_b.cpu().numpy().astype(np.float64) - ref_b)
    assert np.all(err <= 1e-5 * bound + 1e-30), float((err / np.maximum(bound, 1e-30)).max())
    del sino, got_b, ty


def _serve_pair(dev):
    from ct_laboratory_amd.serving import MAPRequest, build_problem

    reqs = [MAPRequest(0, "fan2d", 256, 360, 512, iters=20, seed=40, arrival=0.0),
            MAPRequest(1, "cone3d", 128, 180, 256, iters=20, seed=41, arrival=0.0)]
    for r in reqs:
        build_problem(r, dev)
    return reqs


def test_c5_smallest_requests_alone_equal_mix(dev):
    from ct_laboratory_amd.serving import MAPRequest, ReconServer

    reqs = _serve_pair(dev)
    print("[fullsize] served pair: 256^2 fan-beam and 128^3 cone-beam, 20 iterations each", flush=True)
    mixed, _ = ReconServer(dev, slots=2).run(reqs)
    for r in reqs:
        alone, _ = ReconServer(dev, slots=1).run([MAPRequest(**{**r.__dict__, "arrival": 0.0})])
        a, m = alone[r.rid][0].cpu().numpy(), mixed[r.rid][0].cpu().numpy()
        if r.kind == "cone3d":
            assert np.array_equal(a, m)
        else:
            assert rel_err(m, a) < 1e-5
        xt = r.x_true.cpu().numpy()
        assert np.linalg.norm(a - xt) < np.linalg.norm(xt)


def test_c5_first_gradient_against_oracle(dev):
    """The fused iteration's back projection of the first residual, (A 0 - y) / var,
    against the oracle's back projection of the same residual, for both requests;
    and the first update x1 = -d2 * that back projection."""
    from ct_laboratory_amd.map_estimator import QuadraticMAP
    from ct_laboratory_amd.serving import _projector

    for r in _serve_pair(dev):
        print(f"[fullsize] first gradient of the {r.kind} request against the oracle", flush=True)
        proj = _projector(r, dev)
        m = QuadraticMAP(proj, r.y, r.var, r.beta, torch.zeros_like(r.x_true))
        m.step()
        torch.cuda.synchronize()
        resid = (-r.y.reshape(1, -1) / r.var).cpu().numpy().astype(np.float32)
        # the kernel forms (Ax - y) * (1/var) in float32: compare against the oracle fed that residual
        r_dev = m.r.cpu().numpy()
        assert rel_err(r_dev, resid) < 1e-6
        minv = torch.inverse(proj.M.float()).cpu().numpy()
        b = proj.b.cpu().numpy()
        shape = tuple(r.x_true.shape)
        zr = "exact" if r.kind == "fan2d" else None
        src, dst = proj.src.cpu().numpy(), proj.dst.cpu().numpy()
        ref = O.back(r_dev, shape, minv, b, src, dst, mode="torch", zero_rule=zr)
        bound = O.back(np.abs(r_dev), shape, minv, b, src, dst, mode="torch", zero_rule=zr)
        err = np.abs(m.bp.cpu().numpy().astype(np.float64) - ref)
        assert np.all(err <= 1e-5 * bound + 1e-30), float((err / np.maximum(bound, 1e-30)).max())
        x1 = m.estimate.cpu().numpy().reshape(shape)
        want = -(m.d2.cpu().numpy().astype(np.float64) * m.bp.cpu().numpy().reshape(shape))
        assert rel_err(x1, want) < 1e-6
