"""GPU tests of the continuous-batching MAP server: each result is the same alone
and in the mix (cone-beam requests bit for bit: their kernels are deterministic;
fan-beam requests within 1e-5, the parts of a back-projection tile combine with float atomics),
and the same as the plain per-request autograd loop within 1e-4 (same iterations;
autograd forms the Gaussian and smoothness gradients with other float
operations than the fused kernels)."""
import numpy as np
import pytest
import torch

from ct_laboratory_amd.serving import MAPRequest, ReconServer, build_problem, latency_stats, plain_loop
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _mix(dev):
    reqs = [MAPRequest(0, "fan2d", 64, 60, 128, iters=6, seed=0, arrival=0.0),
            MAPRequest(1, "cone3d", 24, 16, 32, iters=6, seed=1, arrival=0.0),
            MAPRequest(2, "fan2d", 48, 40, 96, iters=6, seed=2, arrival=0.001),
            MAPRequest(3, "cone3d", 32, 12, 32, iters=6, seed=3, arrival=0.002),
            MAPRequest(4, "fan2d", 80, 72, 160, iters=6, seed=4, arrival=0.002)]
    for r in reqs:
        build_problem(r, dev)
    return reqs


def test_results_alone_equal_mix_and_plain(dev):
    reqs = _mix(dev)
    server = ReconServer(dev, slots=3)
    mixed, span = server.run(reqs)
    assert sorted(mixed) == [r.rid for r in reqs]
    stats = latency_stats(mixed)
    assert 0 < stats["p50_s"] <= stats["p99_s"] <= span + 1e-3
    plain, _ = plain_loop(reqs, dev)
    for r in reqs:
        alone, _ = ReconServer(dev, slots=1).run([MAPRequest(**{**r.__dict__, "arrival": 0.0})])
        a, m, p = (x.cpu().numpy() for x in (alone[r.rid][0], mixed[r.rid][0], plain[r.rid][0]))
        if r.kind == "cone3d":
            assert np.array_equal(a, m), r.rid
        else:
            assert rel_err(m, a) < 1e-5, r.rid
        assert rel_err(p, a) < 1e-4, r.rid
        # the iterations reconstruct: closer to the phantom than the zero start
        xt = r.x_true.cpu().numpy()
        assert np.linalg.norm(a - xt) < np.linalg.norm(xt), r.rid


def test_budget_limits_concurrency(dev):
    reqs = _mix(dev)
    from ct_laboratory_amd.serving import resident_bytes

    budget = max(resident_bytes(r) for r in reqs) * 1.01  # one request at a time
    out, _ = ReconServer(dev, hbm_budget=budget, slots=4).run(reqs)
    ref, _ = ReconServer(dev, slots=4).run(reqs)
    for r in reqs:
        a, b = out[r.rid][0].cpu().numpy(), ref[r.rid][0].cpu().numpy()
        assert rel_err(a, b) < 1e-5


def test_plain_loop_on_reference_kernel_structure(dev):
    """The baseline loop on the reference's kernel structure (fan-beam through a
    table consumed as given, cone-beam without a detector layout, the shim's
    schedules off) reconstructs the same images as the plain loop within 1e-4
    (the segments are the same; atomics sum them in other orders), and leaves
    the shim's schedule switch as it found it."""
    from ct_laboratory_amd import _C
    from ct_laboratory_amd.serving import _projector, _reference_projector

    reqs = _mix(dev)
    for r in reqs:
        p, q = _projector(r, dev), _reference_projector(r, dev)
        if r.kind == "fan2d":
            assert p.tvals is None and q.tvals is not None and not q._own_untouched()
        else:
            assert p.det_shape is not None and q.det_shape is None
    saved = _C.AUTO_LAYOUT
    plain, _ = plain_loop(reqs, dev)
    ref, _ = plain_loop(reqs, dev, reference_kernels=True)
    assert _C.AUTO_LAYOUT == saved
    for r in reqs:
        a, b = plain[r.rid][0].cpu().numpy(), ref[r.rid][0].cpu().numpy()
        assert rel_err(b, a) < 1e-4, r.rid
