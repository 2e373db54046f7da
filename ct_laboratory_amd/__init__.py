"""Siddon ray-driven projectors for MI355X (gfx950): hand-written HIP kernels
behind a C ABI (libctlab_hip.so), the reference extension's ``_C`` functions
over it, and the reference's projector layer in ``tomography``.

As in the reference package (ct_laboratory/__init__.py:26-33), the projector
layer is also importable flat (``ct_laboratory_amd.ConeBeam3DProjector``), and
the flat module names the reference keeps for backward compatibility
(ct_laboratory/ct_projector_3d_module.py:1-4 and its siblings) name the
``tomography`` modules: ``import ct_laboratory_amd.ct_projector_3d_module`` is
``ct_laboratory_amd.tomography.ct_projector_3d_module``.
"""
import importlib as _importlib
import sys as _sys

from . import tomography  # noqa: F401
from .tomography import *  # noqa: F401,F403

FLAT_MODULES = (
    "conebeam_projector_3d", "ct_projector_2d_autograd", "ct_projector_2d_cuda", "ct_projector_2d_module",
    "ct_projector_2d_torch", "ct_projector_3d_autograd", "ct_projector_3d_cuda", "ct_projector_3d_module",
    "ct_projector_3d_multirot", "ct_projector_3d_subsets", "ct_projector_3d_torch", "fanbeam_projector_2d",
    "standard_image_transform", "staticct_projector_2d", "staticct_projector_3d",
)

for _name in FLAT_MODULES:
    _module = _importlib.import_module(f"{__name__}.tomography.{_name}")
    _sys.modules[f"{__name__}.{_name}"] = _module
    globals()[_name] = _module
del _name, _module
