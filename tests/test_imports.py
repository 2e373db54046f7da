"""The reference's import surface: every public name of its tomography modules
(ct_laboratory/tomography/*.py), under the same module names, in the package's
tomography namespace, flat at the top level (ct_laboratory/__init__.py:26-33)
and through the flat backward-compatibility module names
(ct_laboratory/ct_projector_3d_module.py:1-4 and siblings), including the
imports of the reference's timing benchmark
(benchmark/projector_timing_benchmark.py:28-33)."""
import importlib

import pytest

REFERENCE_NAMES = {
    "conebeam_projector_3d": ["build_conebeam_3d_geometry", "ConeBeam3DProjector"],
    "ct_projector_2d_autograd": ["CTProjector2DFunction"],
    "ct_projector_2d_cuda": ["compute_intersections_2d_cuda", "forward_project_2d_cuda", "back_project_2d_cuda"],
    "ct_projector_2d_module": ["CTProjector2DModule"],
    "ct_projector_2d_torch": ["compute_intersections_2d_torch", "forward_project_2d_torch", "back_project_2d_torch",
                              "CTProjector2DFunction", "CTProjector2D"],
    "ct_projector_3d_autograd": ["CTProjector3DFunction", "CTBackProjector3DFunction"],
    "ct_projector_3d_cuda": ["compute_intersections_3d_cuda", "compress_tvals_3d_cuda", "forward_project_3d_cuda",
                             "forward_project_3d_compressed_cuda", "back_project_3d_cuda",
                             "back_project_3d_compressed_cuda", "forward_project_3d_on_the_fly_cuda",
                             "back_project_3d_on_the_fly_cuda"],
    "ct_projector_3d_module": ["CTProjector3DModule", "precompute_tvals_stitched"],
    "ct_projector_3d_multirot": ["MultiRotationProjector"],
    "ct_projector_3d_subsets": ["SubsetProjector", "OrderedSubsetProjector", "RandomSubsetProjector"],
    "ct_projector_3d_torch": ["compress_tvals_to_uint16", "decompress_tvals_from_uint16",
                              "compute_intersections_3d_torch", "forward_project_3d_torch", "back_project_3d_torch",
                              "CTProjector3DFunction", "CTProjector3D"],
    "fanbeam_projector_2d": ["build_fanbeam_2d_geometry", "FanBeam2DProjector"],
    "standard_image_transform": ["standard_image_transform_3d", "standard_image_transform_2d"],
    "staticct_projector_2d": ["build_static_2d_geometry", "StaticCTProjector2D", "build_uniform_static_2d_geometry",
                              "UniformStaticCTProjector2D"],
    "staticct_projector_3d": ["build_circular_sequence", "build_helical_sequence", "build_static_3d_geometry",
                              "StaticCTProjector3D", "build_uniform_static_3d_geometry",
                              "UniformStaticCTProjector3D"],
}


@pytest.mark.parametrize("module", sorted(REFERENCE_NAMES))
def test_module_names(module):
    sub = importlib.import_module(f"ct_laboratory_amd.tomography.{module}")
    flat = importlib.import_module(f"ct_laboratory_amd.{module}")
    assert flat is sub
    for name in REFERENCE_NAMES[module]:
        assert hasattr(sub, name), f"{module}.{name}"


def test_package_namespaces():
    import ct_laboratory_amd
    import ct_laboratory_amd.tomography as tomo

    for module, names in REFERENCE_NAMES.items():
        for name in names:
            assert hasattr(tomo, name) and hasattr(ct_laboratory_amd, name), name
    # as in the reference package, the autograd Functions take the names the torch modules share
    assert tomo.CTProjector3DFunction.__module__.endswith("ct_projector_3d_autograd")
    assert tomo.CTProjector2DFunction.__module__.endswith("ct_projector_2d_autograd")


def test_timing_benchmark_imports():
    from ct_laboratory_amd.ct_projector_3d_module import CTProjector3DModule, precompute_tvals_stitched  # noqa: F401
    from ct_laboratory_amd.ct_projector_3d_torch import (back_project_3d_torch,  # noqa: F401
                                                         compute_intersections_3d_torch, forward_project_3d_torch)
