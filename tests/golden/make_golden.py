"""Writes the golden vectors of the reference's torch projector.

Run in a container that holds the reference checkout at /root/reference:

    python tests/golden/make_golden.py

It loads ct_laboratory/tomography/ct_projector_{2d,3d}_torch.py by file path
(they import only torch), evaluates them on small seeded cases and stores
inputs and outputs in tests/golden/*.npz.  Only these data files are used by
the tests; this script is not run by them.
"""
import importlib.util
import os

import numpy as np
import torch

REF = "/root/reference/ct_laboratory/tomography"
OUT = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    spec = importlib.util.spec_from_file_location(name, os.path.join(REF, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _rays(rng, shape, M, b, n_each):
    """World-space rays: far endpoints, endpoints inside, half-integer lattice
    endpoints in index space (on-plane and through-corner cases, some
    axis-parallel), and a few zero-length rays."""
    nd = len(shape)
    n = np.array(shape, np.float64)
    out = []
    far = rng.uniform(-2.0, 3.0, (2, n_each, nd)) * n
    ins = rng.uniform(0.0, 1.0, (2, n_each, nd)) * (n - 1)
    lat = np.round(rng.uniform(-1.0, 2.0, (2, n_each, nd)) * n * 2) / 2 - 0.5 * rng.integers(0, 2, (2, n_each, nd))
    ax = rng.integers(0, nd, n_each)
    lat[1, np.arange(n_each), ax] = lat[0, np.arange(n_each), ax]
    for idx_pair in (far, ins, lat):
        out.append(idx_pair)
    idx = np.concatenate(out, axis=1)  # [2, N, nd] index space
    zl = idx[:, :4].copy()
    zl[1] = zl[0]
    idx = np.concatenate([idx, zl], axis=1)
    world = idx @ M.T.astype(np.float64) + b.astype(np.float64)
    return world[0].astype(np.float32), world[1].astype(np.float32)


def _affine(rng, shape, kind):
    nd = len(shape)
    if kind == "identity":
        M = np.eye(nd)
        b = -(np.array(shape) - 1) / 2.0
    elif kind == "aniso":
        s = np.array([0.7, 1.3, 1.1][:nd])
        M = np.diag(s)
        b = -(np.array(shape) - 1) / 2.0 * s + np.array([0.25, -0.4, 0.1][:nd])
    else:  # rotated about the last axis, scaled
        th = 0.3
        R = np.eye(nd)
        R[0, 0] = R[1, 1] = np.cos(th)
        R[0, 1], R[1, 0] = -np.sin(th), np.sin(th)
        M = R @ np.diag([0.9, 1.2, 1.0][:nd])
        b = np.array([-2.0, 1.5, -0.75][:nd])
    return M.astype(np.float32), b.astype(np.float32)


def make_3d(t3):
    cases = [((7, 5, 6), "identity", 11), ((6, 8, 5), "aniso", 12), ((5, 6, 7), "rotated", 13)]
    for shape, kind, seed in cases:
        rng = np.random.default_rng(seed)
        M, b = _affine(rng, shape, kind)
        src, dst = _rays(rng, shape, M, b, 40)
        Mt, bt, st, dt = (torch.from_numpy(x) for x in (M, b, src, dst))
        vol = rng.uniform(0.0, 1.0, shape).astype(np.float32)
        sino_in = rng.uniform(0.0, 1.0, src.shape[0]).astype(np.float32)
        tv = t3.compute_intersections_3d_torch(*shape, Mt, bt, st, dt)
        fwd = t3.forward_project_3d_torch(torch.from_numpy(vol), tv, Mt, bt, st, dt)
        bck = t3.back_project_3d_torch(torch.from_numpy(sino_in), tv, Mt, bt, st, dt, *shape)
        u16, start, scale = t3.compress_tvals_to_uint16(tv, Mt, st, dt)
        dec = t3.decompress_tvals_from_uint16(u16, start, scale)
        fwd_u16 = t3.forward_project_3d_torch(torch.from_numpy(vol), (u16, start, scale), Mt, bt, st, dt)
        minv = torch.inverse(Mt)
        np.savez_compressed(os.path.join(OUT, f"torch3d_{kind}.npz"), shape=np.array(shape), M=M, b=b,
                            minv=minv.numpy(), src=src, dst=dst, vol=vol, sino_in=sino_in, tvals=tv.numpy(),
                            fwd=fwd.numpy(), back=bck.numpy(), u16=u16.numpy(), start=start.numpy(),
                            scale=scale.numpy(), decompressed=dec.numpy(), fwd_u16=fwd_u16.numpy())


def make_2d(t2):
    cases = [((9, 7), "identity", 21), ((8, 10), "rotated", 22)]
    for shape, kind, seed in cases:
        rng = np.random.default_rng(seed)
        M, b = _affine(rng, shape, kind)
        src, dst = _rays(rng, shape, M, b, 50)
        Mt, bt, st, dt = (torch.from_numpy(x) for x in (M, b, src, dst))
        img = rng.uniform(0.0, 1.0, shape).astype(np.float32)
        sino_in = rng.uniform(0.0, 1.0, src.shape[0]).astype(np.float32)
        tv = t2.compute_intersections_2d_torch(*shape, Mt, bt, st, dt)
        fwd = t2.forward_project_2d_torch(torch.from_numpy(img), tv, Mt, bt, st, dt)
        bck = t2.back_project_2d_torch(torch.from_numpy(sino_in), tv, Mt, bt, st, dt, *shape)
        minv = torch.inverse(Mt)
        np.savez_compressed(os.path.join(OUT, f"torch2d_{kind}.npz"), shape=np.array(shape), M=M, b=b,
                            minv=minv.numpy(), src=src, dst=dst, img=img, sino_in=sino_in, tvals=tv.numpy(),
                            fwd=fwd.numpy(), back=bck.numpy())


if __name__ == "__main__":
    torch.set_num_threads(1)
    make_3d(_load("ct_projector_3d_torch"))
    make_2d(_load("ct_projector_2d_torch"))
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))
