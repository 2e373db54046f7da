"""GPU tests of the FP64-valued projectors (double values and sums over the
float32 traversal's segments): the same operator as the float32 kernels, an
exact transpose pair in double (adjoint identity to FP64 rounding,
torch.autograd.gradcheck through the autograd Functions), and the MAP gradient
of the 64^3 / 90-view / 96x96 sub-problem against central differences to 1e-8."""
import numpy as np
import pytest
import torch

from ct_laboratory_amd import _C
from ct_laboratory_amd import _native as N
from ct_laboratory_amd.map_estimator import QuadraticSmoothnessRandomVariable
from ct_laboratory_amd.phantoms import ellipsoid_phantom_3d
from ct_laboratory_amd.tomography import (CTProjector2DFunction, CTProjector3DFunction, ConeBeam3DProjector,
                                          build_conebeam_3d_geometry, standard_image_transform_2d)
from ct_laboratory_amd.tomography.ct_projector_3d_autograd import OnTheFlyOptions
from helpers import random_rays

pytestmark = pytest.mark.gpu


def _t(a, dev, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), device=dev, dtype=dtype)


@pytest.mark.parametrize("mode", [N.MODE_TORCH, N.MODE_OTF])
def test_f64_same_operator_and_exact_transpose(dev, mode):
    shape = (30, 26, 22)
    src, dst = build_conebeam_3d_geometry(6, 20, 16, 300.0, 450.0, det_spacing=2.0)
    ts, td = src.to(dev).contiguous(), dst.to(dev).contiguous()
    tm, tb = _t(np.eye(3), dev), _t([-(n - 1) / 2 for n in shape], dev)
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 1, shape)
    y = rng.standard_normal(src.shape[0])
    ax32 = _C.forward_project_3d_on_the_fly_cuda(_t(x, dev), ts, td, tm, tb, mode=mode)[0].double()
    ax = _C.forward_project_3d_on_the_fly_f64(_t(x, dev, torch.float64), ts, td, tm, tb, mode=mode)[0]
    # the same matrix: float32 and float64 values differ only by float32 rounding of the sums
    assert float((ax - ax32).abs().max()) <= 1e-5 * float(ax.abs().max())
    aty = _C.back_project_3d_on_the_fly_f64(_t(y, dev, torch.float64), ts, td, tm, tb, *shape, mode=mode)[0]
    lhs = float((ax * _t(y, dev, torch.float64)).sum())
    rhs = float((_t(x, dev, torch.float64) * aty).sum())
    assert abs(lhs - rhs) <= 1e-12 * float((ax.abs() * _t(np.abs(y), dev, torch.float64)).sum())


def test_gradcheck_3d_and_2d(dev):
    """torch.autograd.gradcheck: the backward (hand-written adjoint) equals the
    numerical Jacobian of the forward, in double."""
    src, dst = build_conebeam_3d_geometry(3, 5, 4, 40.0, 60.0, det_spacing=3.0)
    n = 6
    M = torch.eye(3, device=dev)
    b = torch.full((3,), -(n - 1) / 2, device=dev)
    x = torch.rand((n, n, n), device=dev, dtype=torch.float64, requires_grad=True)
    ts, td = src.to(dev).contiguous(), dst.to(dev).contiguous()
    for sem in ("torch", "otf"):
        opts = OnTheFlyOptions(sem)
        assert torch.autograd.gradcheck(
            lambda v: CTProjector3DFunction.apply(v, None, None, None, M, b, ts, td, "cuda", opts), (x,),
            nondet_tol=1e-12)  # the backward adds with double atomics
    rng = np.random.default_rng(8)
    m = 7
    M2, b2 = standard_image_transform_2d(m, m, 1.0, 1.0)
    s2, d2 = random_rays(rng, (m, m), np.ones(2, np.float32), 40, "far")
    img = torch.rand((m, m), device=dev, dtype=torch.float64, requires_grad=True)
    for sem in ("torch", "otf"):
        assert torch.autograd.gradcheck(
            lambda v: CTProjector2DFunction.apply(v, None, M2.to(dev), b2.to(dev), _t(s2, dev), _t(d2, dev), "cuda",
                                                  sem), (img,), nondet_tol=1e-12)


def test_map_gradient_fd_subproblem_f64(dev):
    """The 64^3 volume at 1 mm, 90 views, 96x96 detector at 1 mm, in FP64: the MAP
    gradient through the autograd Function (hand-written adjoint) against central
    differences of the loss, to 1e-8 relative (the float32 pair reaches 2e-4)."""
    n = 64
    proj = ConeBeam3DProjector(n, n, n, 90, 96, 96, 300.0, 450.0, det_spacing=1.0, device=dev)
    x_true = torch.from_numpy(ellipsoid_phantom_3d(n, n, n, 1.0, seed=1)).to(dev).double()
    y = proj.forward_project(x_true).detach()
    assert y.dtype == torch.float64
    var, beta = 0.5, 1e-2
    x = (0.5 * x_true + 0.01).requires_grad_(True)
    prior = QuadraticSmoothnessRandomVariable(beta)
    loss = 0.5 * ((proj(x) - y) ** 2).sum() / var - prior.log_prob(x)
    loss.backward()
    g = x.grad.detach()

    def L(z):
        r = proj.forward_project(z) - y
        e = sum((torch.diff(z, dim=d) ** 2).sum() for d in range(3))
        return float(0.5 * (r * r).sum() / var + 0.5 * beta * e)

    gen = torch.Generator().manual_seed(11)
    for _ in range(3):
        v = torch.randn(x.shape, generator=gen, dtype=torch.float64).to(dev)
        h = 1e-3 * float(x.detach().norm()) / float(v.norm())
        fd = (L(x.detach() + h * v) - L(x.detach() - h * v)) / (2 * h)
        an = float((g * v).sum())
        assert abs(fd - an) <= 1e-8 * abs(an), (fd, an)
