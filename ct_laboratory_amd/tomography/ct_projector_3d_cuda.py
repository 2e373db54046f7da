"""The reference's module name for the 3D dispatch functions
(ct_laboratory/tomography/ct_projector_3d_cuda.py): the same functions as
ct_projector_3d_hip, which defines them under both names."""
from .ct_projector_3d_hip import *  # noqa: F401,F403
from .ct_projector_3d_hip import __all__  # noqa: F401
