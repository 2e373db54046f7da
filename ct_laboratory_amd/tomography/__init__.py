"""Projector layer over the HIP kernels, with the reference's module names
(ct_laboratory/tomography/__init__.py:8-22): projector functions, autograd
Functions, nn.Modules, the fan-beam, cone-beam and StaticCT geometries, and
the ray-subset and multi-rotation wrappers.  Besides the ``*_hip`` modules,
the reference's module names ``ct_projector_{2d,3d}_cuda`` and
``ct_projector_{2d,3d}_torch`` import here too (the latter with the torch
projector's function names and rule, computed by the HIP kernels)."""
from .standard_image_transform import *  # noqa: F401,F403
# the torch-rule names first: the autograd Functions below take the shared names,
# as in the reference's package (tomography/__init__.py:8-15)
from .ct_projector_2d_torch import *  # noqa: F401,F403
from .ct_projector_3d_torch import *  # noqa: F401,F403
from .ct_projector_2d_hip import *  # noqa: F401,F403
from .ct_projector_3d_hip import *  # noqa: F401,F403
from .ct_projector_3d_hip import ConeBeamPlan  # noqa: F401
from .ct_projector_2d_autograd import *  # noqa: F401,F403
from .ct_projector_3d_autograd import *  # noqa: F401,F403
from .ct_projector_2d_module import *  # noqa: F401,F403
from .ct_projector_3d_module import *  # noqa: F401,F403
from .fanbeam_projector_2d import *  # noqa: F401,F403
from .conebeam_projector_3d import *  # noqa: F401,F403
from .staticct_projector_2d import *  # noqa: F401,F403
from .staticct_projector_3d import *  # noqa: F401,F403
from .ct_projector_3d_subsets import *  # noqa: F401,F403
from .ct_projector_3d_multirot import *  # noqa: F401,F403
