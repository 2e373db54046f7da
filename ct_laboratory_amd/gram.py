"""Gram products A^T A for many vectors at once.

The reference applies the Gram operator of a projector column by column,
one forward and one back projection per column
(ct_laboratory/sparse_eigen_preconditioner/sparse_eigen_decomposition.py:109-128).
Here the k columns of Q (N x k, N = n_x*n_y*n_z) are one batch: one batched
forward launch ([k, X, Y, Z] -> [k, n_ray]) and one batched back launch, so k
launches each way become one.
"""
from __future__ import annotations

import torch

__all__ = ["gram_matmul", "apply_gram"]


def _shape(projector, volume_shape):
    if volume_shape is not None:
        return tuple(int(x) for x in volume_shape)
    return (int(projector.n_x), int(projector.n_y), int(projector.n_z))


def gram_matmul(projector, Q, volume_shape=None):
    """A^T A Q for Q of shape (N, k): column j of the result is
    back_project(forward_project(Q[:, j])).  ``projector`` exposes
    forward_project (or forward) and back_project over batched volumes."""
    shape = _shape(projector, volume_shape)
    N = 1
    for d in shape:
        N *= d
    if Q.dim() != 2 or Q.shape[0] != N:
        raise ValueError(f"Q must be (N, k) with N = {N}")
    k = Q.shape[1]
    vols = Q.t().contiguous().reshape(k, *shape)
    fwd = getattr(projector, "forward_project", None) or projector.forward
    with torch.no_grad():
        sino = fwd(vols)
        back = projector.back_project(sino.reshape(k, -1).contiguous())
    return back.reshape(k, N).t().contiguous()


def apply_gram(projector, x, volume_shape=None):
    """A^T A x for one flattened volume x (N,)."""
    return gram_matmul(projector, x.reshape(-1, 1), volume_shape).reshape(-1)
