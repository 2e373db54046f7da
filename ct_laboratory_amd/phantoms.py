"""Seeded synthetic phantoms for the benchmark configurations (numpy, CPU).

* ``ellipse_phantom_2d``: 0.02/mm background plus 10 random ellipses, centres
  U(-100,100) mm, semi-axes U(5,80) mm, rotation U(0,pi), attenuation
  U(0.005,0.04)/mm, evaluated at pixel centres, added up.  Row index <-> x.
* ``ellipsoid_phantom_3d``: 20 random ellipsoids, centres U(-100,100) mm,
  semi-axes U(5,100) mm, uniform random rotation, attenuation U(0.005,0.04)/mm,
  no background.
"""
import numpy as np

__all__ = ["ellipse_phantom_2d", "ellipsoid_phantom_3d"]


def _centres(n, s):
    return (np.arange(n, dtype=np.float64) - (n - 1) / 2.0) * s


def ellipse_phantom_2d(n_row, n_col, pixel_size=1.0, n_ellipse=10, seed=0, background=0.02):
    rng = np.random.default_rng(seed)
    x = _centres(n_row, pixel_size)[:, None]
    y = _centres(n_col, pixel_size)[None, :]
    img = np.full((n_row, n_col), background, dtype=np.float64)
    for _ in range(n_ellipse):
        cx, cy = rng.uniform(-100, 100, 2)
        a, b = rng.uniform(5, 80, 2)
        th = rng.uniform(0, np.pi)
        mu = rng.uniform(0.005, 0.04)
        c, s = np.cos(th), np.sin(th)
        xr = (x - cx) * c + (y - cy) * s
        yr = -(x - cx) * s + (y - cy) * c
        img += mu * ((xr / a) ** 2 + (yr / b) ** 2 <= 1.0)
    return img.astype(np.float32)


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def ellipsoid_phantom_3d(n_x, n_y, n_z, voxel_size=1.0, n_ellipsoid=20, seed=1, background=0.0, device=None):
    """With ``device`` (a torch device) the same float32 operations run there
    and a tensor on that device is returned (the ellipsoids' parameters still
    come from the seeded numpy generator)."""
    if device is not None:
        return _ellipsoid_phantom_3d_torch(n_x, n_y, n_z, voxel_size, n_ellipsoid, seed, background, device)
    rng = np.random.default_rng(seed)
    x = _centres(n_x, voxel_size).astype(np.float32)[:, None, None]
    y = _centres(n_y, voxel_size).astype(np.float32)[None, :, None]
    z = _centres(n_z, voxel_size).astype(np.float32)[None, None, :]
    vol = np.full((n_x, n_y, n_z), background, dtype=np.float32)
    for _ in range(n_ellipsoid):
        c = rng.uniform(-100, 100, 3).astype(np.float32)
        ax = rng.uniform(5, 100, 3).astype(np.float32)
        R = _random_rotation(rng).astype(np.float32)
        mu = np.float32(rng.uniform(0.005, 0.04))
        dx, dy, dz = x - c[0], y - c[1], z - c[2]
        q = np.zeros((n_x, n_y, n_z), dtype=np.float32)
        for k in range(3):
            comp = R[0, k] * dx + R[1, k] * dy + R[2, k] * dz
            q += (comp / ax[k]) ** 2
        vol += mu * (q <= 1.0)
    return vol


def _ellipsoid_phantom_3d_torch(n_x, n_y, n_z, voxel_size, n_ellipsoid, seed, background, device):
    import torch

    rng = np.random.default_rng(seed)

    def centres(n, shape):
        return torch.from_numpy(_centres(n, voxel_size).astype(np.float32)).to(device).view(shape)

    x, y, z = centres(n_x, (-1, 1, 1)), centres(n_y, (1, -1, 1)), centres(n_z, (1, 1, -1))
    vol = torch.full((n_x, n_y, n_z), float(background), dtype=torch.float32, device=device)
    for _ in range(n_ellipsoid):
        c = rng.uniform(-100, 100, 3).astype(np.float32)
        ax = rng.uniform(5, 100, 3).astype(np.float32)
        R = _random_rotation(rng).astype(np.float32)
        mu = np.float32(rng.uniform(0.005, 0.04))
        dx, dy, dz = x - float(c[0]), y - float(c[1]), z - float(c[2])
        q = torch.zeros((n_x, n_y, n_z), dtype=torch.float32, device=device)
        for k in range(3):
            comp = float(R[0, k]) * dx + float(R[1, k]) * dy + float(R[2, k]) * dz
            q += (comp / float(ax[k])) ** 2
        vol += float(mu) * (q <= 1.0)
    return vol
