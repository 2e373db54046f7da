"""MAP reconstruction through the hand-written adjoint.

Mirrors the reference's estimator interface (bayesian_estimation/map_estimator.py:36-143)
and the two random variables it is used with for CT
(random_variable/gaussian.py:54-63: y | x ~ N(A x, var);
random_variable/image_energies.py:18-29: log p(x) = -(w/2) sum ||diff(x)||^2).

``MaximumAPosterioriEstimator`` keeps the reference's semantics:

    loss = -(log p(y | x) + prior_weight * log p(x))
    x <- x - lr_k * P(P(grad))      (SGD, optional linear warm-up)

For the pair the reference uses in CT (a Gaussian likelihood around one of this
package's projectors and the quadratic smoothness prior), the estimator runs a
fused iteration instead of autograd (``QuadraticMAP``):

1. forward projection into a resident buffer;
2. ``ctlab_map_residual``: residual (A x - y)/var and the data energy;
3. back projection of the residual (the hand-written adjoint kernel);
4. ``ctlab_map_prior_update``: prior gradient beta L x, preconditioned step and
   the prior energy.

The gradient is the same as autograd's through the projector Functions. No
iteration synchronises with the host, nothing is allocated per iteration, and
two iterations (ping-pong between two buffers) are captured once in a HIP graph
and replayed.  Any other likelihood or prior falls back to the reference's
autograd loop, whose backward is the same adjoint kernel.
"""
from __future__ import annotations

import math

import torch

from . import _C
from . import _native as N
from .tomography.ct_projector_2d_module import CTProjector2DModule
from .tomography.ct_projector_3d_hip import _collapse, semantics_mode
from .tomography.ct_projector_3d_module import CTProjector3DModule

__all__ = ["LinearGaussianRandomVariable", "QuadraticSmoothnessRandomVariable", "Preconditioner",
           "IdentityPreconditioner", "DiagonalPreconditioner", "MaximumAPosterioriEstimator", "QuadraticMAP",
           "ProjectorOps", "sqs_preconditioner"]


# ---------------------------------------------------------------------------
# Random variables and preconditioners (the reference's interface)
# ---------------------------------------------------------------------------


class LinearGaussianRandomVariable:
    """y | x ~ N(op(x), diag(var)) (random_variable/gaussian.py:54-63)."""

    def __init__(self, op, var=1.0):
        self.op = op
        self.var = var

    def log_prob(self, y, x):
        mu = self.op(x)
        var = torch.as_tensor(self.var, dtype=mu.dtype, device=mu.device)
        return torch.distributions.Normal(mu, var.sqrt(), validate_args=False).log_prob(y).sum()


class QuadraticSmoothnessRandomVariable:
    """log p(x) = -(w/2) sum_d ||diff_d(x)||^2 (random_variable/image_energies.py:18-29)."""

    normalized = False

    def __init__(self, weight=1.0):
        self.weight = float(weight)

    def log_prob(self, x):
        energy = sum(torch.diff(x, dim=d).pow(2).sum() for d in range(x.dim()))
        return -0.5 * self.weight * energy


class Preconditioner(torch.nn.Module):
    """Symmetric change of variables x = P(z) (optimization/preconditioner.py:19-39)."""

    def forward(self, x):  # pragma: no cover - abstract
        raise NotImplementedError

    def inverse(self, x):  # pragma: no cover - abstract
        raise NotImplementedError


class IdentityPreconditioner(Preconditioner):
    def forward(self, x):
        return x

    def inverse(self, x):
        return x


class DiagonalPreconditioner(Preconditioner):
    """P = diag(scale), scale > 0 (optimization/preconditioner.py:52-66)."""

    def __init__(self, scale):
        super().__init__()
        scale = torch.as_tensor(scale)
        if torch.any(scale <= 0):
            raise ValueError("DiagonalPreconditioner scale must be positive")
        self.register_buffer("scale", scale)

    def forward(self, x):
        return x * self.scale

    def inverse(self, x):
        return x / self.scale


# ---------------------------------------------------------------------------
# Allocation-free projector operations
# ---------------------------------------------------------------------------


class ProjectorOps:
    """Forward and back projection of one of this package's projector modules
    into caller buffers, with M^-1, tables, layout and brick plan resolved once."""

    def __init__(self, proj):
        if isinstance(proj, CTProjector3DModule):
            self.nd = 3
            self.shape = (proj.n_x, proj.n_y, proj.n_z)
        elif isinstance(proj, CTProjector2DModule):
            self.nd = 2
            self.shape = (proj.n_row, proj.n_col)
        else:
            raise TypeError(f"expected a CTProjector2DModule or CTProjector3DModule, got {type(proj).__name__}")
        self.proj = proj
        M, b = _collapse(proj.M, proj.b)
        self.minv = torch.inverse(M.to(torch.float32)).to(torch.float32).contiguous()
        self.b = b.to(torch.float32).contiguous()
        self.src, self.dst = proj.src, proj.dst
        self.n_ray = int(proj.src.shape[0])
        self.device = proj.src.device
        self.mode = semantics_mode(getattr(proj, "semantics", "torch"))
        if self.nd == 3:
            self.table = proj._table()
            opts = proj._options() if self.table is None else None
            self.det_shape = None if opts is None else opts.det_shape
            self.plan = None if opts is None else opts.plan
            self.bins = None
            if self.table is None and self.plan is None:
                # rays given without a layout: recognise it now, outside any graph capture,
                # else list the rays of each brick
                layout = _C._layout_of(self.src, self.dst)
                plan = _C._plan_of(self.src, self.dst, layout, self.minv, self.b, tuple(self.shape)) if layout else None
                if plan is not None:
                    self.det_shape, self.plan = layout, plan
                elif self.det_shape is None:
                    self.bins = _C._bins_of(self.src, self.dst, self.minv, self.b, tuple(self.shape), self.mode)
            self.ws = None
            if self.plan is not None:
                n_view = self.n_ray // (self.det_shape[0] * self.det_shape[1])
                self.ws = torch.empty(n_view, device=self.device)
            elif self.bins is not None:
                self.ws = torch.empty(1, device=self.device)
        else:
            # the module's own untouched table runs on the fly with the table's rule
            self.table, sem, zr = proj._rule()
            self.mode = semantics_mode(sem)
            self.zero_rule = (N.ZERO_EXACT if sem == "torch" else N.ZERO_THRESH) if zr is None else zr

    def forward(self, x, out):
        if self.nd == 3:
            if isinstance(self.table, tuple):
                return _C.forward_project_3d_compressed_cuda(x, *self.table, self.src, self.dst, self.minv, self.b,
                                                             out=out)
            if self.table is not None:
                return _C.forward_project_3d_cuda(x, self.table, self.src, self.dst, self.minv, self.b, out=out)
            return _C.forward_project_3d_on_the_fly_cuda(x, self.src, self.dst, self.minv, self.b, mode=self.mode,
                                                        det_shape=self.det_shape, out=out)
        if self.table is not None:
            return _C.forward_project_2d_cuda(x, self.table, self.src, self.dst, self.minv, self.b, out=out)
        return _C.forward_project_2d_on_the_fly_cuda(x, self.src, self.dst, self.minv, self.b, mode=self.mode,
                                                    zero_rule=self.zero_rule, out=out)

    def back(self, y, out):
        if self.nd == 3:
            n_x, n_y, n_z = self.shape
            if isinstance(self.table, tuple):
                return _C.back_project_3d_compressed_cuda(y, *self.table, self.src, self.dst, self.minv, self.b, n_x,
                                                          n_y, n_z, out=out)
            if self.table is not None:
                return _C.back_project_3d_cuda(y, self.table, self.src, self.dst, self.minv, self.b, n_x, n_y, n_z,
                                               out=out)
            if self.plan is not None:
                return _C.back_project_3d_brick(y, self.src, self.dst, self.det_shape, self.plan, self.minv, self.b,
                                                n_x, n_y, n_z, mode=self.mode, out=out, workspace=self.ws)
            if self.bins is not None:
                return _C.back_project_3d_brick_bins(y, self.src, self.dst, self.bins, self.minv, self.b, n_x, n_y, n_z,
                                                     mode=self.mode, out=out, workspace=self.ws)
            return _C.back_project_3d_on_the_fly_cuda(y, self.src, self.dst, self.minv, self.b, n_x, n_y, n_z,
                                                      mode=self.mode, det_shape=self.det_shape, out=out)
        n_row, n_col = self.shape
        if self.table is not None:
            return _C.back_project_2d_cuda(y, self.table, self.src, self.dst, self.minv, self.b, n_row, n_col,
                                           out=out)
        return _C.back_project_2d_on_the_fly_cuda(y, self.src, self.dst, self.minv, self.b, n_row, n_col,
                                                  mode=self.mode, zero_rule=self.zero_rule, out=out)

    def grid3(self):
        """The image grid as (n0, n1, n2) for the stencil kernels (2D: n0 = 1)."""
        return self.shape if self.nd == 3 else (1,) + tuple(self.shape)


def sqs_preconditioner(ops: ProjectorOps, var: float, beta: float) -> torch.Tensor:
    """Diagonal of the separable quadratic surrogate of the MAP objective:
    d2 = 1 / (A^T A 1 / var + 2 beta deg), so that x <- x - d2 * grad never
    increases the objective (A has nonnegative entries; 2 beta deg bounds the
    row sums of beta L).  Returned as d2 (= P^2 of a diagonal P)."""
    dev = ops.device
    ones = torch.ones((1,) + tuple(ops.shape), device=dev)
    a1 = torch.empty((1, ops.n_ray), device=dev)
    ops.forward(ones, a1)
    ata1 = torch.empty((1,) + tuple(ops.shape), device=dev)
    ops.back(a1, ata1)
    d2 = torch.empty(tuple(ops.shape), device=dev)
    n0, n1, n2 = ops.grid3()
    with torch.cuda.device(dev):
        N.call("ctlab_map_sqs_weights", N.ptr(ata1), 1.0 / float(var), float(beta), n0, n1, n2, N.ptr(d2),
               N.stream_of(ata1))
    return d2


# ---------------------------------------------------------------------------
# Fused iteration
# ---------------------------------------------------------------------------


_WARMED = set()


class QuadraticMAP:
    """x_{k+1} = x_k - lr * d2 * (A^T (A x_k - y) / var + beta * L x_k), resident on
    the device.  ``d2`` is the squared diagonal preconditioner (None: identity;
    "sqs": ``sqs_preconditioner``).  ``energies()`` returns the data and prior
    energies of the iterate the last iteration started from."""

    def __init__(self, projector, measurements, var, beta, x_init, lr=1.0, d2="sqs"):
        self.ops = projector if isinstance(projector, ProjectorOps) else ProjectorOps(projector)
        ops = self.ops
        dev = ops.device
        shape = tuple(ops.shape)
        if tuple(x_init.shape) != shape:
            raise ValueError(f"x_init must have shape {shape}")
        self.var, self.beta, self.lr = float(var), float(beta), float(lr)
        self.y = measurements.reshape(1, ops.n_ray).to(dev, torch.float32).contiguous()
        self.x = [x_init.detach().to(dev, torch.float32).reshape((1,) + shape).contiguous().clone(),
                  torch.empty((1,) + shape, device=dev)]
        self.cur = 0
        self.ax = torch.empty((1, ops.n_ray), device=dev)
        self.r = torch.empty((1, ops.n_ray), device=dev)
        self.bp = torch.empty((1,) + shape, device=dev)
        self.energy = torch.zeros(2, dtype=torch.float64, device=dev)
        if isinstance(d2, str):
            if d2 != "sqs":
                raise ValueError("d2 must be a tensor, None or 'sqs'")
            d2 = sqs_preconditioner(ops, self.var, self.beta)
        self.d2 = None if d2 is None else d2.to(dev, torch.float32).reshape(shape).contiguous()
        self.graph = None
        self.iterations = 0

    # one iteration from buffer s to buffer 1 - s, on the current stream
    def _enqueue(self, s: int, lr: float):
        ops = self.ops
        n0, n1, n2 = ops.grid3()
        stream = N.stream_of(self.ax)
        N.call("ctlab_zero_f64", N.ptr(self.energy), 2, stream)
        ops.forward(self.x[s], self.ax)
        N.call("ctlab_map_residual", N.ptr(self.ax), N.ptr(self.y), ops.n_ray, 1.0 / self.var, N.ptr(self.r),
               ctypes_offset(self.energy, 0), stream)
        ops.back(self.r, self.bp)
        N.call("ctlab_map_prior_update", N.ptr(self.x[s]), N.ptr(self.bp), N.ptr(self.d2), self.beta, lr, n0, n1,
               n2, N.ptr(self.x[1 - s]), ctypes_offset(self.energy, 1), stream)

    def step(self, lr: float | None = None):
        """One eager iteration."""
        with torch.cuda.device(self.ops.device):
            self._enqueue(self.cur, self.lr if lr is None else float(lr))
        self.cur = 1 - self.cur
        self.iterations += 1

    def _kind(self):
        ops = self.ops
        return (ops.nd, type(ops.table).__name__, getattr(ops, "plan", None) is not None, ops.mode, str(ops.device))

    def capture(self):
        """Captures two iterations (buffer 0 -> 1 -> 0) in a HIP graph.  Capture runs
        on a side stream with capture_begin/capture_end, which (unlike the
        torch.cuda.graph context) does not synchronise the device, so other
        streams keep running.  The first capture of a kind of iteration in the
        process runs it eagerly once before, so that every kernel has been
        launched outside a capture."""
        if self.cur != 0:
            self.step()  # the graph starts from buffer 0
        dev = self.ops.device
        cur = torch.cuda.current_stream(dev)
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(cur)
        g = torch.cuda.CUDAGraph()
        warm = self._kind() not in _WARMED
        keep = None
        if warm:
            keep = self.x[0].clone()
            keep.record_stream(side)
        with torch.cuda.device(dev), torch.cuda.stream(side):
            if warm:
                self._enqueue(0, self.lr)
                self._enqueue(1, self.lr)
                self.x[0].copy_(keep)
                _WARMED.add(self._kind())
            g.capture_begin()
            try:
                self._enqueue(0, self.lr)
                self._enqueue(1, self.lr)
            finally:
                g.capture_end()
        cur.wait_stream(side)
        self.graph = g
        return g

    def run(self, num_iters: int, use_graph: bool = True):
        """Runs ``num_iters`` iterations without host synchronisation."""
        if use_graph and num_iters >= 2:
            if self.graph is None:
                self.capture()
            if self.cur != 0:
                self.step()
                num_iters -= 1
            for _ in range(num_iters // 2):
                self.graph.replay()
                self.iterations += 2
            if num_iters % 2:
                self.step()
        else:
            for _ in range(num_iters):
                self.step()
        return self.estimate

    @property
    def estimate(self) -> torch.Tensor:
        return self.x[self.cur][0]

    def energies(self):
        """(data energy, prior energy) of the iterate the last iteration started from (syncs)."""
        e = self.energy.cpu()
        return float(e[0]), float(e[1])


def ctypes_offset(t: torch.Tensor, k: int):
    import ctypes

    return ctypes.c_void_p(t.data_ptr() + k * t.element_size())


# ---------------------------------------------------------------------------
# Reference-compatible estimator
# ---------------------------------------------------------------------------


def _projector_of(op):
    target = getattr(op, "__self__", op)
    return target if isinstance(target, (CTProjector2DModule, CTProjector3DModule)) else None


class MaximumAPosterioriEstimator:
    """x* = argmax log p(y | x) + prior_weight log p(x) by preconditioned SGD
    (bayesian_estimation/map_estimator.py:36-143).  ``fused=True`` uses
    ``QuadraticMAP`` when the likelihood is a ``LinearGaussianRandomVariable``
    around one of this package's projectors and the prior is a
    ``QuadraticSmoothnessRandomVariable`` (identity or diagonal preconditioner)."""

    def __init__(self, likelihood, prior, measurements, x_init, preconditioner=None, lr=1.0, warmup_iters=None,
                 prior_weight=1.0, fused=True):
        self.likelihood, self.prior = likelihood, prior
        self.measurements = measurements
        self.preconditioner = preconditioner if preconditioner is not None else IdentityPreconditioner()
        self.prior_weight = float(prior_weight)
        self.lr = float(lr)
        self.warmup_iters = warmup_iters
        self._shape = x_init.shape
        self._it = 0
        self.fast = None
        proj = _projector_of(likelihood.op) if isinstance(likelihood, LinearGaussianRandomVariable) else None
        var = likelihood.var if proj is not None else None
        if (fused and proj is not None and isinstance(prior, QuadraticSmoothnessRandomVariable)
                and isinstance(self.preconditioner, (IdentityPreconditioner, DiagonalPreconditioner))
                and not torch.is_tensor(var) and x_init.dim() in (2, 3)):
            d2 = None
            if isinstance(self.preconditioner, DiagonalPreconditioner):
                sc = self.preconditioner.scale.to(x_init.device, torch.float32)
                d2 = torch.broadcast_to(sc * sc, x_init.shape).contiguous()
            self.fast = QuadraticMAP(proj, measurements, float(var), self.prior_weight * prior.weight, x_init,
                                     lr=self.lr, d2=d2)
            self._n_meas = measurements.numel()
        else:
            self.x_var = x_init.detach().clone().requires_grad_(True)

    def _lr_k(self):
        if self.warmup_iters is None:
            return self.lr
        return self.lr * min(1.0, float(self._it) / self.warmup_iters)

    @property
    def x(self):
        if self.fast is not None:
            return self.fast.estimate.reshape(self._shape)
        return self.x_var.detach().reshape(self._shape)

    def negative_log_posterior(self, x):
        log_likelihood = self.likelihood.log_prob(self.measurements, x)
        log_prior = self.prior.log_prob(x)
        return -(log_likelihood + self.prior_weight * log_prior), log_likelihood, log_prior

    def step(self):
        """One step; returns (loss, log_lik, log_prior) at the iterate before the step."""
        lr = self._lr_k()
        self._it += 1
        if self.fast is not None:
            self.fast.step(lr)
            data, prior_e = self.fast.energies()
            const = self._n_meas * (0.5 * math.log(2 * math.pi) + 0.5 * math.log(self.fast.var))
            log_lik = -data - const
            log_prior = -prior_e / self.prior_weight if self.prior_weight != 0 else float("nan")
            return data + const + prior_e, log_lik, log_prior
        if self.x_var.grad is not None:
            self.x_var.grad = None
        loss, log_likelihood, log_prior = self.negative_log_posterior(self.x_var.reshape(self._shape))
        loss.backward()
        with torch.no_grad():
            g = self.x_var.grad.reshape(self._shape)
            direction = self.preconditioner(self.preconditioner(g))
            self.x_var -= lr * direction.reshape(self.x_var.shape)
        return loss.item(), log_likelihood.item(), log_prior.item()

    def estimate(self, num_iters=100, use_graph=True):
        """Runs ``num_iters`` steps and returns the estimate.  The fused path
        runs warm-up steps eagerly and the rest as graph replays, with no host
        synchronisation."""
        if self.fast is None:
            for _ in range(num_iters):
                self.step()
            return self.x
        while num_iters > 0 and self.warmup_iters is not None and self._it < self.warmup_iters:
            self.fast.step(self._lr_k())
            self._it += 1
            num_iters -= 1
        if num_iters > 0:
            self.fast.lr = self.lr
            self.fast.run(num_iters, use_graph=use_graph)
            self._it += num_iters
        return self.x
