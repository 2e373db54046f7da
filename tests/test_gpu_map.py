"""GPU tests of the MAP iteration: the fused, graph-captured iteration against the
reference-style autograd loop (same gradient through the hand-written adjoint),
graph replays against eager iterations, monotone descent with the separable
quadratic surrogate, and a finite-difference check of the MAP gradient on the
64^3 / 90-view / 96x96 sub-problem."""
import math

import numpy as np
import pytest
import torch

from ct_laboratory_amd.map_estimator import (DiagonalPreconditioner, LinearGaussianRandomVariable,
                                             MaximumAPosterioriEstimator, ProjectorOps, QuadraticMAP,
                                             QuadraticSmoothnessRandomVariable, sqs_preconditioner)
from ct_laboratory_amd.phantoms import ellipse_phantom_2d, ellipsoid_phantom_3d
from ct_laboratory_amd.tomography import ConeBeam3DProjector, FanBeam2DProjector
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _problem(dev, kind):
    if kind == "cone":
        n = 20
        proj = ConeBeam3DProjector(n, n, n, 10, 16, 14, 100.0, 150.0, det_spacing=2.0, device=dev)
        x_true = torch.from_numpy(ellipsoid_phantom_3d(n, n, n, 1.0, seed=1)).to(dev)
    else:
        n = 40
        proj = FanBeam2DProjector(n, n, 30, 64, 200.0, 300.0, device=dev)
        x_true = torch.from_numpy(ellipse_phantom_2d(n, n, 1.0, seed=0)).to(dev)
    y = proj.forward_project(x_true)
    g = torch.Generator(device="cpu").manual_seed(3)
    sigma = 0.01 * float(y.max())
    y = (y + sigma * torch.randn(y.shape, generator=g).to(dev)).contiguous()
    return proj, x_true, y, sigma ** 2


@pytest.mark.parametrize("kind", ["cone", "fan"])
def test_fused_matches_autograd_loop(dev, kind):
    proj, x_true, y, var = _problem(dev, kind)
    beta_w, pw = 0.5, 2.0
    ops = ProjectorOps(proj)
    d2 = sqs_preconditioner(ops, var, pw * beta_w)
    pre = DiagonalPreconditioner(d2.sqrt())
    x0 = torch.zeros_like(x_true)
    lik = LinearGaussianRandomVariable(proj, var)
    prior = QuadraticSmoothnessRandomVariable(beta_w)
    fused = MaximumAPosterioriEstimator(lik, prior, y, x0, preconditioner=pre, lr=1.0, prior_weight=pw)
    slow = MaximumAPosterioriEstimator(lik, prior, y, x0, preconditioner=pre, lr=1.0, prior_weight=pw, fused=False)
    assert fused.fast is not None and slow.fast is None
    tol = 1e-5 if kind == "cone" else 1e-4  # the fan-beam table back projector uses float atomics
    for _ in range(4):
        lf, llf, lpf = fused.step()
        ls, lls, lps = slow.step()
        assert abs(lf - ls) <= 1e-5 * abs(ls)
        assert abs(llf - lls) <= 1e-5 * abs(lls) and abs(lpf - lps) <= 1e-4 * abs(lps) + 1e-6
    assert rel_err(_np(fused.x), _np(slow.x)) < tol


@pytest.mark.parametrize("kind", ["cone", "fan"])
def test_graph_replay_equals_eager(dev, kind):
    proj, x_true, y, var = _problem(dev, kind)
    x0 = torch.zeros_like(x_true)
    eager = QuadraticMAP(proj, y, var, 0.3, x0)
    graph = QuadraticMAP(proj, y, var, 0.3, x0, d2=eager.d2)
    graph.capture()
    for rep in range(3):  # replays with eager work and synchronisations between them
        eager.step()
        eager.step()
        graph.run(2)
        torch.cuda.synchronize()
        a, b = _np(eager.estimate), _np(graph.estimate)
        if kind == "cone":  # forward, brick back and stencil are deterministic
            assert np.array_equal(a, b), rep
        else:
            assert rel_err(b, a) < 1e-5, rep
        ea, eb = eager.energies(), graph.energies()
        assert abs(ea[0] - eb[0]) <= 1e-6 * abs(ea[0]) and abs(ea[1] - eb[1]) <= 1e-6 * abs(ea[1]) + 1e-9


def test_sqs_descent_and_reconstruction(dev):
    proj, x_true, y, var = _problem(dev, "cone")
    m = QuadraticMAP(proj, y, var, 0.05, torch.zeros_like(x_true))
    prev = math.inf
    errs = []
    for k in range(12):
        m.step()
        d, p = m.energies()
        assert d + p <= prev * (1 + 1e-6), k  # the surrogate step never increases the objective
        prev = d + p
        errs.append(float((m.estimate - x_true).norm() / x_true.norm()))
    assert errs[-1] < 0.9 * errs[0], errs


def test_map_gradient_fd_subproblem(dev):
    """64^3 volume at 1 mm, 90 views, 96x96 detector at 1 mm: the gradient of the MAP
    loss through the autograd Function (hand-written adjoint backward) against central
    differences of the loss accumulated in FP64."""
    n = 64
    proj = ConeBeam3DProjector(n, n, n, 90, 96, 96, 300.0, 450.0, det_spacing=1.0, device=dev)
    x_true = torch.from_numpy(ellipsoid_phantom_3d(n, n, n, 1.0, seed=1)).to(dev)
    y = proj.forward_project(x_true).detach()
    var, beta = 0.5, 1e-2
    x = (0.5 * x_true + 0.01).requires_grad_(True)
    prior = QuadraticSmoothnessRandomVariable(beta)
    loss = 0.5 * ((proj(x) - y) ** 2).sum() / var - prior.log_prob(x)
    loss.backward()
    g = x.grad.detach()

    def L(z):
        r = proj.forward_project(z).double() - y.double()
        e = sum((torch.diff(z.double(), dim=d) ** 2).sum() for d in range(3))
        return float(0.5 * (r * r).sum() / var + 0.5 * beta * e)

    gen = torch.Generator().manual_seed(11)
    for _ in range(3):
        v = torch.randn(x.shape, generator=gen).to(dev)
        h = 1e-2 * float(x.detach().norm()) / float(v.norm())
        fd = (L(x.detach() + h * v) - L(x.detach() - h * v)) / (2 * h)
        an = float((g.double() * v.double()).sum())
        assert abs(fd - an) <= 2e-4 * abs(an), (fd, an)
