"""Projector layer over the HIP kernels, with the reference's module names
(ct_laboratory/tomography/__init__.py:8-22): projector functions, autograd
Functions, nn.Modules, the fan-beam, cone-beam and StaticCT geometries, and
the ray-subset and multi-rotation wrappers."""
from .standard_image_transform import *  # noqa: F401,F403
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
