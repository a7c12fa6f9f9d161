"""Seeded random-ellipsoid phantoms (synthetic inputs for the benchmark configurations).

An ellipsoid is (centre[3] mm, semi_axes[3] mm, R[3,3] rotation, mu mm^-1); the phantom is the
sum of mu * indicator(|diag(1/a) R^T (x - c)| <= 1) sampled at voxel centres, where the voxel
with index g along an axis of N voxels sits at (g - N//2) * voxel_mm (the reference's centring).
Arrays are [z, y, x]; centres and axes are given in (x, y, z).
"""
import math

import numpy as np


def random_rotation(rng):
    """Uniform random rotation from a uniform unit quaternion."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def random_ellipsoids(seed, n, axes_mm, mu_range, centre_radius_mm, z_axes_mm=None):
    """n ellipsoids: semi-axes U[axes_mm], attenuation U[mu_range] (mm^-1), centres uniform in a
    ball of radius centre_radius_mm, uniform random orientation. z_axes_mm, when given, draws the
    third semi-axis from its own range and keeps the rotation about z only (elongated along z)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        a = rng.uniform(axes_mm[0], axes_mm[1], 3)
        if z_axes_mm is not None:
            a[2] = rng.uniform(z_axes_mm[0], z_axes_mm[1])
            phi = rng.uniform(0, 2 * math.pi)
            R = np.array([[math.cos(phi), -math.sin(phi), 0.0],
                          [math.sin(phi), math.cos(phi), 0.0], [0.0, 0.0, 1.0]])
        else:
            R = random_rotation(rng)
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        c = d * centre_radius_mm * rng.uniform() ** (1.0 / 3.0)
        mu = rng.uniform(mu_range[0], mu_range[1])
        out.append((c, a, R, mu))
    return out


def voxelize(ellipsoids, shape, voxel_mm, device=None, dtype=None, texture_sigma=0.0,
             texture_seed=0):
    """Sample the phantom on a [nz, ny, nx] grid. Uses torch on `device` (a ROCm device for the
    large configurations); returns a float32 torch tensor. texture_sigma adds seeded N(0, sigma^2)
    per-voxel texture (mm^-1)."""
    import torch
    device = torch.device("cpu") if device is None else torch.device(device)
    dtype = torch.float32 if dtype is None else dtype
    nz, ny, nx = shape
    vol = torch.zeros(shape, dtype=torch.float32, device=device)
    cz, cy, cx = nz // 2, ny // 2, nx // 2
    for c, a, R, mu in ellipsoids:
        c = np.asarray(c, np.float64)
        a = np.asarray(a, np.float64)
        # axis-aligned bounding box of the rotated ellipsoid (mm)
        half = np.sqrt((np.asarray(R) ** 2) @ (a ** 2))
        lo = np.floor((c - half) / voxel_mm).astype(int) + np.array([cx, cy, cz])
        hi = np.ceil((c + half) / voxel_mm).astype(int) + np.array([cx, cy, cz])
        lo = np.maximum(lo, 0)
        hi = np.minimum(hi, np.array([nx - 1, ny - 1, nz - 1]))
        if np.any(hi < lo):
            continue
        xs = (torch.arange(lo[0], hi[0] + 1, device=device, dtype=torch.float64) - cx) * voxel_mm
        ys = (torch.arange(lo[1], hi[1] + 1, device=device, dtype=torch.float64) - cy) * voxel_mm
        zs = (torch.arange(lo[2], hi[2] + 1, device=device, dtype=torch.float64) - cz) * voxel_mm
        Z, Y, X = torch.meshgrid(zs - c[2], ys - c[1], xs - c[0], indexing="ij")
        M = torch.as_tensor(np.asarray(R).T / a[:, None], device=device)  # diag(1/a) R^T
        q = (M[0, 0] * X + M[0, 1] * Y + M[0, 2] * Z) ** 2
        q += (M[1, 0] * X + M[1, 1] * Y + M[1, 2] * Z) ** 2
        q += (M[2, 0] * X + M[2, 1] * Y + M[2, 2] * Z) ** 2
        sub = vol[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        sub += (q <= 1.0).to(torch.float32) * float(mu)
    if texture_sigma > 0:
        g = torch.Generator(device=device)
        g.manual_seed(int(texture_seed))
        vol += texture_sigma * torch.randn(shape, generator=g, device=device, dtype=torch.float32)
    return vol.to(dtype)


# ---- the benchmark configurations' phantoms -------------------------------------------------
def config_phantom(config, device=None, seed=None):
    """Phantom of configuration 0, 1, 2 or 3 (sizes and distributions from the workload
    definitions). The centre radius of config 1's ellipsoids is not specified: 60 mm is used."""
    if config == 0:
        e = random_ellipsoids(0 if seed is None else seed, 10, (5, 40), (0.005, 0.02), 40.0)
        return voxelize(e, (128, 128, 128), 1.0, device), e
    if config == 1:
        e = random_ellipsoids(1 if seed is None else seed, 40, (3, 80), (0.002, 0.03), 60.0)
        return voxelize(e, (512, 512, 512), 0.5, device, texture_sigma=0.001,
                        texture_seed=1 if seed is None else seed), e
    if config == 2:  # as configuration 1, elongated along z
        e = random_ellipsoids(2 if seed is None else seed, 40, (3, 80), (0.002, 0.03), 60.0,
                              z_axes_mm=(20, 120))
        return voxelize(e, (256, 512, 512), 0.5, device, texture_sigma=0.001,
                        texture_seed=2 if seed is None else seed), e
    if config == 3:
        e = random_ellipsoids(3 if seed is None else seed, 10, (5, 40), (0.005, 0.02), 40.0)
        return voxelize(e, (256, 256, 256), 1.0, device), e
    raise ValueError(f"no phantom for configuration {config}")
