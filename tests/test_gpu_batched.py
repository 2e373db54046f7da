"""GPU tests of the persistent multi-scan 2D launch: per-scan forward projections
bit-identical to the single-scan kernel, back projections (LDS tiles and the
atomic launch) within 1e-5 of the single-scan kernel and of the oracle, queue
exhaustion with uneven scans, and the tuner."""
import numpy as np
import pytest
import torch

import oracle as O
from ct_laboratory_amd import _C
from ct_laboratory_amd import _native as N
from ct_laboratory_amd.batched import BatchedProjector2D, Scan2D
from ct_laboratory_amd.phantoms import ellipse_phantom_2d
from ct_laboratory_amd.tomography import (build_fanbeam_2d_geometry, build_static_2d_geometry,
                                          build_uniform_static_2d_geometry, standard_image_transform_2d)
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _scans(dev):
    scans, images = [], []
    specs = [(32, 20, 48, "fan"), (48, 36, 0, "ring"), (17, 9, 31, "fan"), (64, 50, 96, "fan"), (40, 24, 0, "ring"),
             (8, 3, 5, "fan")]
    for k, (n, n_view, bins, cls) in enumerate(specs):
        px = 64.0 / n
        M, b = standard_image_transform_2d(n, n, px, px)
        if cls == "fan":
            src, dst = build_fanbeam_2d_geometry(n_view, bins, 100.0, 150.0, det_spacing=0.7)
        else:
            n_src = 5
            Mg = torch.eye(2).repeat(n_view, 1, 1)
            bg = torch.zeros(n_view, 2)
            act = torch.zeros(n_view, n_src, dtype=torch.bool)
            act[torch.arange(n_view), torch.arange(n_view) % n_src] = True
            sp, mc, mo, mask = build_uniform_static_2d_geometry(n_src, 60.0, 16, 55.0, 4, 5.0, Mg, bg, act)
            src, dst = build_static_2d_geometry(sp, mc, mo, 4, 5.0, mask, Mg, bg, act)
        sem = "otf" if k == 3 else "torch"
        scans.append(Scan2D(n, n, M, b, src.to(dev), dst.to(dev), semantics=sem, geometry_class=cls))
        images.append(torch.from_numpy(ellipse_phantom_2d(n, n, px, seed=k)).to(dev))
    return scans, images


def _single(sc, img, dev):
    minv, b = sc.minv.to(dev), sc.b.to(dev)
    return _C.forward_project_2d_on_the_fly_cuda(img, sc.src, sc.dst, minv, b, mode=sc.mode, zero_rule=sc.zero_rule)[0]


@pytest.mark.parametrize("rpi,gpc,rpp,tiled,tr", [(256, 1, 64, True, 128), (512, 4, 4096, True, 64),
                                                 (4096, 16, 1 << 20, True, 128), (512, 4, 0, False, 128)])
def test_batch_matches_single_scan_kernels(dev, rpi, gpc, rpp, tiled, tr):
    scans, images = _scans(dev)
    bp = BatchedProjector2D(scans, dev, rays_per_item=rpi, groups_per_cu=gpc, rays_per_part=max(rpp, 1),
                            tiled_back=tiled, tile_rows=tr)
    sinos = bp.forward(images)
    ys = []
    for sc, img, s in zip(scans, images, sinos):
        ref = _single(sc, img, dev)
        assert torch.equal(s, ref), "forward must be bit-identical to the single-scan kernel"
        ys.append(torch.from_numpy(np.random.default_rng(sc.n_ray).uniform(0, 1, sc.n_ray).astype(np.float32)).to(dev))
    backs = bp.back(ys)
    for sc, y, bk in zip(scans, ys, backs):
        ref = _C.back_project_2d_on_the_fly_cuda(y, sc.src, sc.dst, sc.minv.to(dev), sc.b.to(dev), sc.n_row,
                                                 sc.n_col, mode=sc.mode, zero_rule=sc.zero_rule)[0]
        assert rel_err(bk.cpu().numpy(), ref.cpu().numpy()) < 1e-5
        ref_o = O.back(y.cpu().numpy()[None], (sc.n_row, sc.n_col), sc.minv.numpy(), sc.b.numpy(),
                       sc.src.cpu().numpy(), sc.dst.cpu().numpy(), mode="otf" if sc.mode == N.MODE_OTF else "torch",
                       zero_rule="exact" if sc.zero_rule == N.ZERO_EXACT else "thresh")[0]
        assert rel_err(bk.cpu().numpy(), ref_o) < 1e-5
    # repeated launches reuse the resident table and clear the queue counter
    again = bp.forward(images)
    assert all(torch.equal(a, _single(sc, im, dev)) for a, sc, im in zip(again, scans, images))


def test_autotune_and_cache(dev, tmp_path):
    scans, images = _scans(dev)
    f = tmp_path / "tuned.json"
    bp = BatchedProjector2D(scans, dev, tuner_file=str(f))
    best, timing = bp.autotune(images, rays_per_item=(256, 1024), groups_per_cu=(1, 4), rays_per_part=(512, 8192),
                               reps=1)
    assert best[:2] in timing["forward"] and len(timing["forward"]) == 4 and f.exists()
    assert (best[3], best[2]) in timing["back"] and len(timing["back"]) == 4
    bp2 = BatchedProjector2D(scans, dev, tuner_file=str(f))
    assert (bp2.rays_per_item, bp2.groups_per_cu, bp2.rays_per_part, bp2.tile_rows) == best
    sinos = bp2.forward(images)
    assert all(torch.equal(s, _single(sc, im, dev)) for s, sc, im in zip(sinos, scans, images))


def test_tiled_batch_accumulate_and_spare_units(dev):
    """The tiled multi-scan back projection called through the C ABI: accumulate
    adds to the images, and a table whose total counts spare units past the
    last scan gives the same images (the spare units are skipped)."""
    scans, images = _scans(dev)
    bp = BatchedProjector2D(scans, dev, rays_per_part=512)
    ys = [torch.from_numpy(np.random.default_rng(k).uniform(0, 1, sc.n_ray).astype(np.float32)).to(dev)
          for k, sc in enumerate(scans)]
    ref = [b.clone() for b in bp.back(ys)]
    table, units = bp._build(bp.images, ys, 512, tiled=True, tile_rows=128)
    base = [torch.full_like(im, 0.25) for im in bp.images]
    for k in range(len(scans)):
        bp.images[k].copy_(base[k])
    stream = N.stream_of(bp.counter)
    N.call("ctlab_back_project_2d_batch_tiled", N.ptr(table), len(scans), units + 37, 512, 128, 2 * bp.n_cu,
           N.ptr(bp.counter), 1, stream)
    torch.cuda.synchronize()
    for k in range(len(scans)):
        got = bp.images[k].cpu().numpy()
        assert rel_err(got, base[k].cpu().numpy() + ref[k].cpu().numpy()) < 1e-5
