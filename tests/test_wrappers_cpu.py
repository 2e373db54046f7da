"""Host logic of the multi-rotation and subset wrappers (no GPU): window
bounds and padding index maps, pinned by hand-worked small cases, and the
subset cache names of ct_laboratory/tomography/ct_projector_3d_subsets.py:84-94."""
import numpy as np
import torch

from ct_laboratory_amd.tomography.ct_projector_3d_multirot import MultiRotationProjector


def _mr(mode, limit=None):
    # 3 rotations, 2 slices apart, 4-slice windows: on a 6-slice volume the
    # centre is 2.5, the first window centre 0.5, so the windows start at
    # round(-1.0) = -1, round(1.0) = 1 and round(3.0) = 3.
    return MultiRotationProjector(None, 3, 2.0, 4, padding_mode=mode, reflection_padding_slices=limit)


def test_rotation_bounds_worked_example():
    m = _mr("repeat")
    assert [m._get_rotation_z_bounds(6, i) for i in range(3)] == [(-1, 3), (1, 5), (3, 7)]
    # scan_center_slice moves the centre to scan_center_slice - 0.5
    m2 = MultiRotationProjector(None, 1, 0.0, 3, scan_center_slice=4)
    assert m2._get_rotation_z_bounds(10, 0) == (2, 5)  # round(3.5 - 1.0) = 2 (ties to even)


def test_padding_index_maps():
    idx, keep, det = _mr("repeat")._window_index(6, "cpu")
    assert idx.tolist() == [[0, 0, 1, 2], [1, 2, 3, 4], [3, 4, 5, 5]]
    assert keep.all() and det.tolist() == [[True, False, False, False], [False] * 4, [False, False, False, True]]
    idx, keep, det = _mr("reflect")._window_index(6, "cpu")
    assert idx.tolist() == [[1, 0, 1, 2], [1, 2, 3, 4], [3, 4, 5, 4]]  # -1 -> 1, 6 -> 4
    assert keep.all() and not det.any()
    idx, keep, _ = _mr("reflect", limit=0)._window_index(6, "cpu")
    assert keep.tolist() == [[False, True, True, True], [True] * 4, [True, True, True, False]]
    idx, keep, _ = _mr("zeros")._window_index(6, "cpu")
    assert keep.tolist() == [[False, True, True, True], [True] * 4, [True, True, True, False]]
    # reflections repeat until in range: a 2-slice volume seen by a 6-slice window
    m = MultiRotationProjector(None, 1, 0.0, 6, padding_mode="reflect")
    idx, _, _ = m._window_index(2, "cpu")
    # centre 0.5, start round(0.5 - 2.5) = -2: -2 -> 2 -> 0, -1 -> 1, 0, 1, 2 -> 0, 3 -> -1 -> 1
    assert idx.tolist() == [[0, 1, 0, 1, 0, 1]]


def test_windows_and_detach():
    vol = torch.arange(2 * 1 * 6, dtype=torch.float32).reshape(2, 1, 6).requires_grad_(True)
    m = _mr("repeat")
    w = m._windows(vol)
    assert w.shape == (3, 2, 1, 4)
    assert torch.equal(w[0, 0, 0], torch.tensor([0.0, 0.0, 1.0, 2.0]))
    w.sum().backward()
    # the padded copies are detached: slice 0 gets one unit from rotation 0's
    # in-range slot only, slice 5 one from rotation 2
    g = vol.grad[0, 0].tolist()
    assert g == [1.0, 2.0, 2.0, 2.0, 2.0, 1.0]
    vz = torch.ones(1, 1, 6)
    wz = _mr("zeros")._windows(vz)
    assert wz[0, 0, 0].tolist() == [0.0, 1.0, 1.0, 1.0] and wz[2, 0, 0].tolist() == [1.0, 1.0, 1.0, 0.0]


def test_subset_cache_name():
    from ct_laboratory_amd.tomography.ct_projector_3d_subsets import OrderedSubsetProjector
    o = OrderedSubsetProjector.__new__(OrderedSubsetProjector)
    o.cache_dir, o.n_x, o.n_y, o.n_z = "/tmp/c", 4, 5, 6
    o.subset_boundaries = np.array([0, 10, 20])
    assert o._get_subset_cache_path(1) == "/tmp/c/tvals_subset0001_rays10-20_nx4_ny5_nz6.pt"
