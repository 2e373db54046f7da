"""Backend names accepted by the projector layer."""

_HIP_NAMES = ("hip", "cuda")


def check_backend(backend: str) -> str:
    """The reference's ``backend='cuda'`` and ``'hip'`` select the HIP kernels.
    ``'torch'`` (the reference's CPU tensor implementation) has no counterpart
    here: there is no CPU projector in this package, so it raises."""
    if backend in _HIP_NAMES:
        return "hip"
    if backend == "torch":
        raise ValueError("backend='torch' (the CPU tensor projector) is not provided: this package runs its "
                         "projectors only as HIP kernels; use backend='hip' (or 'cuda')")
    raise ValueError(f"unknown backend {backend!r}; expected 'hip' or 'cuda'")
