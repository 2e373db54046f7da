"""CPU tests of the 2D module's own-table bookkeeping (no kernels run): the
module projects on the fly only while its table is the one it computed and
nothing has written to it, and that state survives deepcopy and pickling."""
import copy
import io
import torch

from ct_laboratory_amd import _native as N
from ct_laboratory_amd.tomography.ct_projector_2d_module import CTProjector2DModule


def _module_with_own_table():
    m = CTProjector2DModule(8, 8, torch.eye(2), torch.zeros(2), torch.zeros(3, 2), torch.ones(3, 2),
                            precomputed_intersections=False)
    t = torch.zeros(3, 18)  # stands in for the table the constructor computes on the GPU
    m.register_buffer("tvals", t)
    m._own_table = m._record(t)
    return m


def test_own_table_rule_and_edits():
    m = _module_with_own_table()
    assert m._rule() == (None, "torch", N.ZERO_THRESH)
    with torch.no_grad():
        m.tvals[0, 0] = 0.5  # written in place: consumed as given from now on
    assert m._rule()[0] is m.tvals
    m2 = _module_with_own_table()
    m2.tvals = torch.zeros(3, 18)  # replaced: consumed as given
    assert m2._rule()[0] is m2.tvals
    # the rays or the map written in place or replaced: the table is applied to them as given
    m3 = _module_with_own_table()
    with torch.no_grad():
        m3.src[0, 0] = 2.0
    assert m3._rule()[0] is m3.tvals
    m4 = _module_with_own_table()
    m4.M = torch.eye(2) * 2.0
    assert m4._rule()[0] is m4.tvals
    m5 = _module_with_own_table()
    with torch.no_grad():
        m5.b += 1.0
    assert m5._rule()[0] is m5.tvals


def test_own_table_survives_copies():
    m = _module_with_own_table()
    c = copy.deepcopy(m)
    assert c._rule()[0] is None and c.tvals is not m.tvals
    buf = io.BytesIO()
    torch.save(m, buf)
    buf.seek(0)
    r = torch.load(buf, weights_only=False)  # a file this test wrote
    assert r._rule()[0] is None
    with torch.no_grad():
        m.tvals.add_(1.0)
    assert copy.deepcopy(m)._rule()[0] is not None


def test_no_table_uses_module_semantics():
    m = CTProjector2DModule(8, 8, torch.eye(2), torch.zeros(2), torch.zeros(3, 2), torch.ones(3, 2),
                            precomputed_intersections=False, semantics="otf")
    assert m._rule() == (None, "otf", None)


def test_detect_layout_of_ray_stacks():
    """The _C shim recognises [view][rows][cols] stacks of flat-panel views (one
    source per view) and nothing else."""
    import numpy as np

    from ct_laboratory_amd import _C
    from ct_laboratory_amd.tomography import build_conebeam_3d_geometry

    src, dst = build_conebeam_3d_geometry(5, 24, 10, 1000.0, 1500.0, det_spacing=2.0)  # 5 views, 10 x 24
    assert _C.detect_layout(src.contiguous(), dst.contiguous()) == (10, 24)
    one_row = build_conebeam_3d_geometry(3, 80, 1, 1000.0, 1500.0, det_spacing=2.0)
    assert _C.detect_layout(*(t.contiguous() for t in one_row)) == (1, 80)
    rng = np.random.default_rng(0)
    r = torch.from_numpy(rng.uniform(-50, 50, (240, 3)).astype(np.float32))
    assert _C.detect_layout(r, -r) is None  # a different source per ray
    s2 = src.clone()
    s2[7] += 1.0  # one view whose source moves within the view
    assert _C.detect_layout(s2, dst.contiguous()) is None
    assert _C.detect_layout(src[:40].contiguous(), dst[:40].contiguous()) is None  # too few rays per view
