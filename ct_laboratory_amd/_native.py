"""ctypes binding of libctlab_hip.so (the C ABI declared in include/ctlab_hip.h).

The library is built in-tree by ``python -m ct_laboratory_amd.build`` into
``ct_laboratory_amd/_lib/libctlab_hip.so``.  There is no fallback: if the
library is missing or fails to load, every projector call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch


def lib_path() -> str:
    """Path of the built library (ct_laboratory_amd/_lib/libctlab_hip.so, written by build.py)."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libctlab_hip.so")

_lock = threading.Lock()
_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_INT = ctypes.c_int

MODE_TORCH = 0
MODE_OTF = 1
MODE_MARCH = 2
ZERO_THRESH = 0
ZERO_EXACT = 1

MODES = {"torch": MODE_TORCH, "otf": MODE_OTF, "march": MODE_MARCH}
COMPRESS_CUDA = 0
COMPRESS_TORCH = 1

# name -> argtypes (restype is int status unless listed in _RESTYPES)
_F32 = ctypes.c_float

_SIGS = {
    "ctlab_map_residual": [_P, _P, _I64, _F32, _P, _P, _P],
    "ctlab_map_prior_update": [_P, _P, _P, _F32, _F32, _I64, _I64, _I64, _P, _P, _P],
    "ctlab_map_sqs_weights": [_P, _F32, _F32, _I64, _I64, _I64, _P, _P],
    "ctlab_zero_f64": [_P, _I64, _P],
    "ctlab_forward_project_2d_batch": [_P, _I64, _I64, _I64, _I64, _P, _P],
    "ctlab_back_project_2d_batch": [_P, _I64, _I64, _I64, _I64, _P, _INT, _P],
    "ctlab_back_project_2d_batch_tiled": [_P, _I64, _I64, _I64, _I64, _I64, _P, _INT, _P],
    "ctlab_back_project_2d_tile_units": [_I64, _I64, _I64, _I64, _I64],
    "ctlab_forward_project_3d_otf": [_P, _I64, _I64, _I64, _I64, _P, _P, _I64, _P, _P, _INT, _I64, _I64, _P, _P],
    "ctlab_back_project_3d_otf": [_P, _I64, _I64, _P, _P, _P, _P, _I64, _I64, _I64, _INT, _I64, _I64, _INT, _P,
                                  _P],
    "ctlab_forward_project_2d_otf": [_P, _I64, _I64, _I64, _P, _P, _I64, _P, _P, _INT, _INT, _P, _P],
    "ctlab_back_project_2d_otf": [_P, _I64, _I64, _P, _P, _P, _P, _I64, _I64, _INT, _INT, _INT, _P, _P],
    "ctlab_forward_project_3d_otf_f64": [_P, _I64, _I64, _I64, _I64, _P, _P, _I64, _P, _P, _INT, _P, _P],
    "ctlab_back_project_3d_otf_f64": [_P, _I64, _I64, _P, _P, _P, _P, _I64, _I64, _I64, _INT, _INT, _P, _P],
    "ctlab_forward_project_2d_otf_f64": [_P, _I64, _I64, _I64, _P, _P, _I64, _P, _P, _INT, _INT, _P, _P],
    "ctlab_back_project_2d_otf_f64": [_P, _I64, _I64, _P, _P, _P, _P, _I64, _I64, _INT, _INT, _INT, _P, _P],
    "ctlab_trace_3d": [_I64, _I64, _I64, _P, _P, _I64, _P, _P, _INT, _INT, _I64, _P, _P, _P, _P],
    "ctlab_trace_2d": [_I64, _I64, _P, _P, _I64, _P, _P, _INT, _INT, _I64, _P, _P, _P, _P],
    "ctlab_compute_intersections_2d": [_I64, _I64, _P, _P, _I64, _P, _P, _INT, _I64, _P, _P],
    "ctlab_compute_intersections_3d": [_I64, _I64, _I64, _P, _P, _I64, _P, _P, _INT, _I64, _P, _P],
    "ctlab_count_crossings_2d": [_I64, _I64, _P, _P, _I64, _P, _P, _INT, _P, _P],
    "ctlab_count_crossings_3d": [_I64, _I64, _I64, _P, _P, _I64, _P, _P, _INT, _P, _P],
    "ctlab_forward_project_2d_tvals": [_P, _I64, _I64, _I64, _P, _I64, _P, _P, _I64, _P, _P, _P, _P],
    "ctlab_back_project_2d_tvals": [_P, _I64, _P, _I64, _P, _P, _I64, _P, _P, _I64, _I64, _INT, _P, _P],
    "ctlab_forward_project_3d_tvals": [_P, _I64, _I64, _I64, _I64, _P, _I64, _P, _P, _I64, _P, _P, _P, _P],
    "ctlab_back_project_3d_tvals": [_P, _I64, _P, _I64, _P, _P, _I64, _P, _P, _I64, _I64, _I64, _INT, _P, _P],
    "ctlab_compress_tvals_3d": [_P, _I64, _I64, _P, _P, ctypes.c_float, _INT, _P, _P, _P, _P],
    "ctlab_forward_project_3d_u16": [_P, _I64, _I64, _I64, _I64, _P, _P, _P, _I64, _P, _P, _I64, _P, _P, _P, _P],
    "ctlab_back_project_3d_u16": [_P, _I64, _P, _P, _P, _I64, _P, _P, _I64, _P, _P, _I64, _I64, _I64, _INT, _P,
                                  _P],
    "ctlab_cone_frames_3d": [_P, _P, _I64, _I64, _I64, _P, _P, _I64, _I64, _I64, _P, _P],
    "ctlab_brick_plan_ints_3d": [_I64, _I64, _I64, _I64],
    "ctlab_brick_plan_3d": [_P, _I64, _I64, _I64, _I64, _I64, _I64, _P, _P],
    "ctlab_back_project_3d_brick": [_P, _I64, _P, _P, _I64, _I64, _I64, _P, _P, _P, _I64, _I64, _I64, _INT, _INT,
                                    _P, _P, _P],
    "ctlab_back_project_3d_brick_noalloc": [_P, _I64, _P, _P, _I64, _I64, _I64, _P, _P, _P, _I64, _I64, _I64, _INT,
                                            _INT, _P, _P],
    "ctlab_brick_count_3d": [_I64, _I64, _I64],
    "ctlab_brick_bins_count_3d": [_P, _P, _I64, _P, _P, _I64, _I64, _I64, _INT, _P, _P, _P],
    "ctlab_brick_bins_fill_3d": [_P, _P, _I64, _P, _P, _I64, _I64, _I64, _INT, _P, _P, _I64, _P],
    "ctlab_brick_bins_slots_3d": [_I64, _I64, _I64, _I64],
    "ctlab_brick_bins_file_3d": [_P, _P, _I64, _P, _P, _I64, _I64, _I64, _P, _P, _P, _P, _P, _P],
    "ctlab_back_project_3d_brick_bins": [_P, _I64, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _INT,
                                         _INT, _P, _P, _P],
    "ctlab_brick_filed_slots_3d": [_I64, _I64, _I64, _I64],
    "ctlab_brick_filed_count_3d": [_P, _P, _I64, _I64, _I64, _P, _P, _P, _I64, _I64, _I64, _P, _P],
    "ctlab_brick_filed_fill_3d": [_P, _P, _I64, _I64, _I64, _P, _P, _P, _I64, _I64, _I64, _P, _P, _I64, _P],
    "ctlab_back_project_3d_brick_filed": [_P, _I64, _P, _P, _I64, _I64, _I64, _P, _P, _P, _P, _P, _P, _I64, _I64,
                                          _I64, _INT, _INT, _P, _P, _P],
    "ctlab_table_brick_count": [_I64, _I64, _I64],
    "ctlab_table_runs_count_3d": [_I64, _I64, _I64, _P, _P, _P, _P, _I64, _P, _P, _I64, _P, _P, _P, _P],
    "ctlab_table_runs_fill_3d": [_I64, _I64, _I64, _P, _P, _P, _P, _I64, _P, _P, _I64, _P, _P, _P, _P, _P, _I64,
                                 _P],
    "ctlab_back_project_3d_table_brick": [_P, _I64, _P, _P, _P, _P, _I64, _P, _P, _I64, _P, _P, _I64, _I64, _I64,
                                          _P, _I64, _P, _I64, _INT, _P, _P],
    "ctlab_forward_project_3d_table_ranged": [_P, _I64, _I64, _I64, _I64, _P, _P, _P, _P, _I64, _P, _P, _P, _P,
                                              _P, _I64, _P, _P, _P, _P],
    "ctlab_version": [],
    "ctlab_last_error": [],
}
_RESTYPES = {"ctlab_last_error": ctypes.c_char_p, "ctlab_brick_plan_ints_3d": ctypes.c_int64,
             "ctlab_brick_count_3d": ctypes.c_int64, "ctlab_table_brick_count": ctypes.c_int64,
             "ctlab_brick_filed_slots_3d": ctypes.c_int64, "ctlab_brick_bins_slots_3d": ctypes.c_int64,
             "ctlab_back_project_2d_tile_units": ctypes.c_int64}


def exported_symbols() -> list[str]:
    return list(_SIGS)


def load() -> ctypes.CDLL:
    """Loads the library (once).  Raises RuntimeError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(
                f"libctlab_hip.so not found at {path}; build it with `python -m ct_laboratory_amd.build`"
            )
        L = ctypes.CDLL(path)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def call(name: str, *args) -> None:
    """Calls an entry point; raises RuntimeError with the library's message on failure."""
    L = load()
    st = getattr(L, name)(*args)
    if st != 0:
        msg = L.ctlab_last_error()
        raise RuntimeError(f"{name} failed (status {st}): {msg.decode() if msg else ''}")


def ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_of(t: torch.Tensor):
    """The current torch stream on t's device, as the library's stream handle."""
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)
