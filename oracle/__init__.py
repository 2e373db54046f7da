"""CPU oracle for the Siddon projector semantics.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product package
(ct_laboratory_amd) never imports this module.

The C restatement lives in siddon_oracle.c (built by oracle/Makefile into
oracle/_build/libsiddon_oracle.so).  Every function here takes numpy arrays and
returns numpy arrays.  Modes:

  "torch"  ct_laboratory/tomography/ct_projector_{2d,3d}_torch.py semantics
  "otf"    src/ct_projector_3d.cu ray_kernel_plane_sort semantics (3d.cu:834-1012)
  "march"  src/ct_projector_2d.cu 2D on-the-fly march (2d.cu:246-413), 2D only

Parity status of each function is listed in ``PINNED`` below and in the tests
that pin it (tests/test_oracle.py).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libsiddon_oracle.so")
_lib = None

MODES = {"torch": 0, "otf": 1, "march": 2}
ZERO_RULES = {"thresh": 0, "exact": 1}

# Which check pins each oracle function (tests/test_oracle.py).
PINNED = {
    "intersections (2D, 3D)": "golden vectors of compute_intersections_{2d,3d}_torch, bit for bit "
    "(identity, anisotropic and rotated grids)",
    "trace/forward/back torch": "golden vectors of forward/back_project_{2d,3d}_torch; closed forms of the "
    "reference examples; exact rational arithmetic on tiny grids; A.1 = chord length; adjoint identity",
    "trace_f64/forward_f64/back_f64": "exact rational arithmetic (index lists equal, lengths to 1e-9); "
    "chord length; adjoint identity to 1e-12",
    "trace/forward/back otf": "closed forms; A.1 = chord length with the end points inserted; adjoint; "
    "index lists against exact rational arithmetic with 0 and 1 inserted. The exact nvcc contraction of "
    "the reference kernel cannot be compiled here: bit patterns of otf crossings are parity unpinned",
    "compress torch, voxel_diag_len": "golden vectors of compress_tvals_to_uint16, bit for bit",
    "forward/back tvals": "golden projections of the reference's own tables",
    "compress cuda, forward/back u16": "known answers for roundf ties, INF entries and rows, the clamp "
    "(tests/test_oracle_cuda_rules.py); decode = one rounding of start + k*scale against exact rationals; "
    "u16 projections = table projections of the decoded table; the 1e-9 skip by a known answer",
    "trace/forward/back march (2D)": "hand-derived known answers of the floor-indexed march, chord length "
    "where the entry voxel is right, adjoint; the product traversal compiled for the host equals it bit for bit",
}


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_INT = ctypes.c_int


def _declare(L):
    L.orc_trace.restype = _I64
    L.orc_trace.argtypes = [_INT, _P, _P, _P, _INT, _INT, _P, _P, _I64, _I64, _P, _P, _P]
    L.orc_trace_f64.restype = _I64
    L.orc_trace_f64.argtypes = [_INT, _P, _P, _P, _INT, _P, _P, _I64, _I64, _P, _P, _P]
    for name in ("orc_forward", "orc_back"):
        f = getattr(L, name)
        f.restype = _INT
        f.argtypes = [_INT, _P, _P, _P, _INT, _INT, _P, _I64, _P, _P, _I64, _P]
    for name in ("orc_forward_f64", "orc_back_f64"):
        f = getattr(L, name)
        f.restype = _INT
        f.argtypes = [_INT, _P, _P, _P, _INT, _P, _P, _P, _I64, _P]
    L.orc_intersections.restype = _INT
    L.orc_intersections.argtypes = [_INT, _P, _P, _P, _INT, _P, _P, _I64, _P]
    for name in ("orc_forward_tvals", "orc_back_tvals"):
        f = getattr(L, name)
        f.restype = _INT
        f.argtypes = [_INT, _P, _P, _P, _P, _I64, _P, _I64, _P, _P, _I64, _P]
    L.orc_compress_cuda.restype = _INT
    L.orc_compress_cuda.argtypes = [_P, _I64, _I64, _P, _P, ctypes.c_double, _P, _P, _P]
    L.orc_compress_torch.restype = _INT
    L.orc_compress_torch.argtypes = [_P, _I64, _I64, _P, _P, ctypes.c_float, _P, _P, _P]
    L.orc_voxel_diag_len.restype = ctypes.c_float
    L.orc_voxel_diag_len.argtypes = [_P]
    for name in ("orc_forward_u16", "orc_back_u16"):
        f = getattr(L, name)
        f.restype = _INT
        f.argtypes = [_P, _P, _P, _P, _I64, _P, _P, _P, _I64, _P, _P, _I64, _P]
    L.orc_decode_u16_cuda.restype = _INT
    L.orc_decode_u16_cuda.argtypes = [_P, _P, _P, _I64, _I64, _P]
    L.orc_num_threads.restype = _INT
    L.orc_set_num_threads.argtypes = [_INT]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _geom(shape, minv, b, dtype):
    n = np.ascontiguousarray(np.asarray(shape, dtype=np.int64))
    return n, np.ascontiguousarray(np.asarray(minv, dtype=dtype).reshape(-1)), np.ascontiguousarray(
        np.asarray(b, dtype=dtype).reshape(-1)
    )


def _default_zero_rule(ndim, zero_rule):
    if zero_rule is None:
        return "exact" if ndim == 2 else "thresh"
    return zero_rule


def set_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


def trace(shape, minv, b, src, dst, mode="torch", zero_rule=None, cap=None):
    """Per-ray (flat index, length) lists of every in-range segment (zero-length
    included).  Returns (idx [n_ray, cap] int64, len [n_ray, cap] f32, counts)."""
    ndim = len(shape)
    zr = _default_zero_rule(ndim, zero_rule)
    n, mi, bb = _geom(shape, minv, b, np.float32)
    src, dst = _f32(src).reshape(-1, ndim), _f32(dst).reshape(-1, ndim)
    n_ray = src.shape[0]
    if cap is None:
        cap = int(sum(shape)) + ndim + 2
    idx = np.zeros((n_ray, cap), dtype=np.int64)
    ln = np.zeros((n_ray, cap), dtype=np.float32)
    cnt = np.zeros(n_ray, dtype=np.int64)
    lib().orc_trace(ndim, _ptr(n), _ptr(mi), _ptr(bb), ZERO_RULES[zr], MODES[mode], _ptr(src), _ptr(dst), n_ray,
                    cap, _ptr(idx), _ptr(ln), _ptr(cnt))
    return idx, ln, cnt


def trace_f64(shape, minv, b, src, dst, zero_rule=None, cap=None):
    ndim = len(shape)
    zr = _default_zero_rule(ndim, zero_rule)
    n, mi, bb = _geom(shape, minv, b, np.float64)
    src, dst = _f64(src).reshape(-1, ndim), _f64(dst).reshape(-1, ndim)
    n_ray = src.shape[0]
    if cap is None:
        cap = int(sum(shape)) + ndim + 2
    idx = np.zeros((n_ray, cap), dtype=np.int64)
    ln = np.zeros((n_ray, cap), dtype=np.float64)
    cnt = np.zeros(n_ray, dtype=np.int64)
    lib().orc_trace_f64(ndim, _ptr(n), _ptr(mi), _ptr(bb), ZERO_RULES[zr], _ptr(src), _ptr(dst), n_ray, cap,
                        _ptr(idx), _ptr(ln), _ptr(cnt))
    return idx, ln, cnt


def index_lists(shape, minv, b, src, dst, mode="torch", zero_rule=None):
    """List per ray of flat voxel indices of the segments with length > 0."""
    idx, ln, cnt = trace(shape, minv, b, src, dst, mode=mode, zero_rule=zero_rule)
    out = []
    for r in range(idx.shape[0]):
        k = int(cnt[r])
        sel = ln[r, :k] > 0
        out.append(idx[r, :k][sel])
    return out


def forward(vol, minv, b, src, dst, mode="torch", zero_rule=None):
    """Line integrals. vol [*shape] or [B,*shape] (f32); returns f64 [n_ray] or [B,n_ray]."""
    ndim = np.asarray(src).shape[-1]
    vol = _f32(vol)
    batched = vol.ndim == ndim + 1
    shape = vol.shape[1:] if batched else vol.shape
    B = vol.shape[0] if batched else 1
    zr = _default_zero_rule(ndim, zero_rule)
    n, mi, bb = _geom(shape, minv, b, np.float32)
    src, dst = _f32(src).reshape(-1, ndim), _f32(dst).reshape(-1, ndim)
    out = np.zeros((B, src.shape[0]), dtype=np.float64)
    lib().orc_forward(ndim, _ptr(n), _ptr(mi), _ptr(bb), ZERO_RULES[zr], MODES[mode], _ptr(vol), B, _ptr(src),
                      _ptr(dst), src.shape[0], _ptr(out))
    return out if batched else out[0]


def back(sino, shape, minv, b, src, dst, mode="torch", zero_rule=None):
    """Back projection. sino [n_ray] or [B,n_ray] (f32); returns f64 volume."""
    ndim = len(shape)
    sino = _f32(sino)
    batched = sino.ndim == 2
    B = sino.shape[0] if batched else 1
    zr = _default_zero_rule(ndim, zero_rule)
    n, mi, bb = _geom(shape, minv, b, np.float32)
    src, dst = _f32(src).reshape(-1, ndim), _f32(dst).reshape(-1, ndim)
    out = np.zeros((B,) + tuple(shape), dtype=np.float64)
    lib().orc_back(ndim, _ptr(n), _ptr(mi), _ptr(bb), ZERO_RULES[zr], MODES[mode], _ptr(sino), B, _ptr(src),
                   _ptr(dst), src.shape[0], _ptr(out))
    return out if batched else out[0]


def forward_f64(vol, minv, b, src, dst, zero_rule=None):
    ndim = np.asarray(src).shape[-1]
    vol = _f64(vol)
    zr = _default_zero_rule(ndim, zero_rule)
    n, mi, bb = _geom(vol.shape, minv, b, np.float64)
    src, dst = _f64(src).reshape(-1, ndim), _f64(dst).reshape(-1, ndim)
    out = np.zeros(src.shape[0], dtype=np.float64)
    lib().orc_forward_f64(ndim, _ptr(n), _ptr(mi), _ptr(bb), ZERO_RULES[zr], _ptr(vol), _ptr(src), _ptr(dst),
                          src.shape[0], _ptr(out))
    return out


def back_f64(sino, shape, minv, b, src, dst, zero_rule=None):
    ndim = len(shape)
    sino = _f64(sino)
    zr = _default_zero_rule(ndim, zero_rule)
    n, mi, bb = _geom(shape, minv, b, np.float64)
    src, dst = _f64(src).reshape(-1, ndim), _f64(dst).reshape(-1, ndim)
    out = np.zeros(tuple(shape), dtype=np.float64)
    lib().orc_back_f64(ndim, _ptr(n), _ptr(mi), _ptr(bb), ZERO_RULES[zr], _ptr(sino), _ptr(src), _ptr(dst),
                       src.shape[0], _ptr(out))
    return out


def intersections(shape, minv, b, src, dst, zero_rule=None):
    ndim = len(shape)
    zr = _default_zero_rule(ndim, zero_rule)
    n, mi, bb = _geom(shape, minv, b, np.float32)
    src, dst = _f32(src).reshape(-1, ndim), _f32(dst).reshape(-1, ndim)
    out = np.zeros((src.shape[0], int(sum(shape)) + ndim), dtype=np.float32)
    lib().orc_intersections(ndim, _ptr(n), _ptr(mi), _ptr(bb), ZERO_RULES[zr], _ptr(src), _ptr(dst),
                            src.shape[0], _ptr(out))
    return out


def forward_tvals(vol, tvals, minv, b, src, dst):
    ndim = np.asarray(src).shape[-1]
    vol = _f32(vol)
    batched = vol.ndim == ndim + 1
    shape = vol.shape[1:] if batched else vol.shape
    B = vol.shape[0] if batched else 1
    n, mi, bb = _geom(shape, minv, b, np.float32)
    tv = _f32(tvals)
    src, dst = _f32(src).reshape(-1, ndim), _f32(dst).reshape(-1, ndim)
    out = np.zeros((B, src.shape[0]), dtype=np.float64)
    lib().orc_forward_tvals(ndim, _ptr(n), _ptr(mi), _ptr(bb), _ptr(vol), B, _ptr(tv), tv.shape[1], _ptr(src),
                            _ptr(dst), src.shape[0], _ptr(out))
    return out if batched else out[0]


def back_tvals(sino, tvals, shape, minv, b, src, dst):
    ndim = len(shape)
    sino = _f32(sino)
    batched = sino.ndim == 2
    B = sino.shape[0] if batched else 1
    n, mi, bb = _geom(shape, minv, b, np.float32)
    tv = _f32(tvals)
    src, dst = _f32(src).reshape(-1, ndim), _f32(dst).reshape(-1, ndim)
    out = np.zeros((B,) + tuple(shape), dtype=np.float64)
    lib().orc_back_tvals(ndim, _ptr(n), _ptr(mi), _ptr(bb), _ptr(sino), B, _ptr(tv), tv.shape[1], _ptr(src),
                         _ptr(dst), src.shape[0], _ptr(out))
    return out if batched else out[0]


def compress_cuda(tvals, src, dst, voxel_diag_len: float):
    tv = _f32(tvals)
    src, dst = _f32(src).reshape(-1, 3), _f32(dst).reshape(-1, 3)
    N, K = tv.shape
    u = np.zeros((N, K), dtype=np.uint16)
    st = np.zeros(N, dtype=np.float32)
    sc = np.zeros(N, dtype=np.float32)
    lib().orc_compress_cuda(_ptr(tv), N, K, _ptr(src), _ptr(dst), float(voxel_diag_len), _ptr(u), _ptr(st),
                            _ptr(sc))
    return u, st, sc


def decode_u16_cuda(u16, start, scale):
    """The crossings the CUDA compressed consumers decode (3d.cu:126-129):
    start + k * scale in one fused multiply-add, k the integer running sum."""
    u = np.ascontiguousarray(np.asarray(u16).view(np.uint16) if np.asarray(u16).dtype == np.int16 else
                             np.asarray(u16, dtype=np.uint16))
    st, sc = _f32(start), _f32(scale)
    N, K = u.shape
    out = np.zeros((N, K), dtype=np.float32)
    lib().orc_decode_u16_cuda(_ptr(u), _ptr(st), _ptr(sc), N, K, _ptr(out))
    return out


def compress_torch(tvals, src, dst, voxel_diag_len: float):
    tv = _f32(tvals)
    src, dst = _f32(src).reshape(-1, 3), _f32(dst).reshape(-1, 3)
    N, K = tv.shape
    u = np.zeros((N, K), dtype=np.int16)
    st = np.zeros(N, dtype=np.float32)
    sc = np.zeros(N, dtype=np.float32)
    lib().orc_compress_torch(_ptr(tv), N, K, _ptr(src), _ptr(dst), ctypes.c_float(voxel_diag_len), _ptr(u),
                             _ptr(st), _ptr(sc))
    return u, st, sc


def voxel_diag_len(M) -> np.float32:
    """max ||M d|| over the four body diagonals, float32, in torch's CPU
    evaluation order (3d_torch.py:27-35; orc_voxel_diag_len)."""
    M = np.ascontiguousarray(np.asarray(M, dtype=np.float32).reshape(3, 3))
    return np.float32(lib().orc_voxel_diag_len(_ptr(M)))


def forward_u16(vol, u16, start, scale, minv, b, src, dst):
    vol = _f32(vol)
    batched = vol.ndim == 4
    shape = vol.shape[1:] if batched else vol.shape
    B = vol.shape[0] if batched else 1
    n, mi, bb = _geom(shape, minv, b, np.float32)
    u = np.ascontiguousarray(np.asarray(u16).view(np.uint16))
    st, sc = _f32(start), _f32(scale)
    src, dst = _f32(src).reshape(-1, 3), _f32(dst).reshape(-1, 3)
    out = np.zeros((B, src.shape[0]), dtype=np.float64)
    lib().orc_forward_u16(_ptr(n), _ptr(mi), _ptr(bb), _ptr(vol), B, _ptr(u), _ptr(st), _ptr(sc), u.shape[1],
                          _ptr(src), _ptr(dst), src.shape[0], _ptr(out))
    return out if batched else out[0]


def back_u16(sino, u16, start, scale, shape, minv, b, src, dst):
    sino = _f32(sino)
    batched = sino.ndim == 2
    B = sino.shape[0] if batched else 1
    n, mi, bb = _geom(shape, minv, b, np.float32)
    u = np.ascontiguousarray(np.asarray(u16).view(np.uint16))
    st, sc = _f32(start), _f32(scale)
    src, dst = _f32(src).reshape(-1, 3), _f32(dst).reshape(-1, 3)
    out = np.zeros((B,) + tuple(shape), dtype=np.float64)
    lib().orc_back_u16(_ptr(n), _ptr(mi), _ptr(bb), _ptr(sino), B, _ptr(u), _ptr(st), _ptr(sc), u.shape[1],
                       _ptr(src), _ptr(dst), src.shape[0], _ptr(out))
    return out if batched else out[0]
