"""CPU-side checks of the C ABI: the library loads and exports every entry point
declared in include/ctlab_hip.h, and the ctypes binding declares the same set."""
import ctypes
import os
import re

from ct_laboratory_amd import _native
from ct_laboratory_amd.build import lib_path

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ctlab_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ctlab_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(lib_path()), "build the library first (python -m ct_laboratory_amd.build)"
    lib = ctypes.CDLL(lib_path())
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(_native.exported_symbols())


def test_version_and_error_string():
    L = _native.load()
    assert L.ctlab_version() >= 1
    assert isinstance(L.ctlab_last_error(), bytes)


def test_argument_errors_are_reported_without_gpu():
    """Argument checks run on the host before any HIP call."""
    L = _native.load()
    st = L.ctlab_forward_project_3d_otf(None, 1, 0, 4, 4, None, None, 1, None, None, 0, 0, 0, None, None)
    assert st == 1 and b"grid dimension" in L.ctlab_last_error()
    st = L.ctlab_back_project_3d_otf(None, 1, 1, None, None, None, None, 4, 4, 4, 7, 0, 0, 0, None, None)
    assert st == 1 and b"mode" in L.ctlab_last_error()
    st = L.ctlab_forward_project_2d_otf(None, 0, 4, 4, None, None, 1, None, None, 0, 0, None, None)
    assert st == 1 and b"batch" in L.ctlab_last_error()


def test_scan2d_record_layout():
    """The numpy record matches the C struct ctlab_scan2d field by field."""
    from ct_laboratory_amd.batched import SCAN_DTYPE

    expect = {"image": 0, "sino": 8, "src": 16, "dst": 24, "minv": 32, "b": 48, "n_row": 56, "n_col": 60,
              "n_ray": 64, "mode": 72, "zero_rule": 76, "item_begin": 80}
    assert {k: SCAN_DTYPE.fields[k][1] for k in expect} == expect and SCAN_DTYPE.itemsize == 88


def test_brick_plan_size():
    """Plan = per brick and view 4 ints, per brick a total, per view E and a 16-float frame."""
    from ct_laboratory_amd import _native as N

    L = N.load()
    for (nx, ny, nz, nv) in ((256, 256, 256, 720), (70, 45, 37, 11), (1, 1, 1, 1)):
        nb = -(-nx // 32) * -(-ny // 32) * -(-nz // 8)
        assert L.ctlab_brick_plan_ints_3d(nx, ny, nz, nv) == nb * (4 * nv + 1) + 17 * nv
    assert L.ctlab_brick_plan_ints_3d(0, 4, 4, 1) == -1
