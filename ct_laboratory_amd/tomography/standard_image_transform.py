"""Voxel-index -> world affine maps (ct_laboratory/tomography/standard_image_transform.py:3-50).

``world = M @ index + b`` with M = diag(voxel size) and b chosen so that the grid
centre maps to the origin.  Returned as float32 CPU tensors, like the reference.
"""
import torch

__all__ = ["standard_image_transform_3d", "standard_image_transform_2d"]


def standard_image_transform_3d(n_x, n_y, n_z, s_x, s_y, s_z, device=None):
    """M = diag(s_x, s_y, s_z), b = -((n-1)/2) * s (standard_image_transform.py:3-28)."""
    M = torch.diag(torch.tensor([s_x, s_y, s_z], dtype=torch.float32))
    b = torch.tensor([-(n_x - 1) / 2.0 * s_x, -(n_y - 1) / 2.0 * s_y, -(n_z - 1) / 2.0 * s_z],
                     dtype=torch.float32)
    if device is not None:
        M, b = M.to(device), b.to(device)
    return M, b


def standard_image_transform_2d(n_x, n_y, s_x, s_y):
    """M = diag(s_x, s_y), b = -((n-1)/2) * s (standard_image_transform.py:31-50)."""
    M = torch.diag(torch.tensor([s_x, s_y], dtype=torch.float32))
    b = torch.tensor([-(n_x - 1) / 2.0 * s_x, -(n_y - 1) / 2.0 * s_y], dtype=torch.float32)
    return M, b
