"""Seeded synthetic inputs for the greedy-fluid workloads (numpy PCG64, float32).

Template: zero background plus Gaussian blobs (centres uniform in the inner 80 %
of the grid, sigma U(4, 10) voxels, amplitude U(0.5, 1)).  Displacement: a sum of
random sinusoidal modes per component (integer wavenumbers 1-3 per axis, random
phase) rescaled to a given maximum length.  Arrays are indexed [z, y, x]; vector
fields [3, z, y, x] (SoA, components x, y, z).
"""
import numpy as np


def blobs(shape, n_blobs=8, seed=0, sigma=(4.0, 10.0), amp=(0.5, 1.0)):
    nz, ny, nx = shape
    rng = np.random.Generator(np.random.PCG64(seed))
    img = np.zeros(shape, dtype=np.float64)
    z = np.arange(nz, dtype=np.float64)[:, None, None]
    y = np.arange(ny, dtype=np.float64)[None, :, None]
    x = np.arange(nx, dtype=np.float64)[None, None, :]
    for _ in range(n_blobs):
        cx = rng.uniform(0.1 * nx, 0.9 * nx)
        cy = rng.uniform(0.1 * ny, 0.9 * ny)
        cz = rng.uniform(0.1 * nz, 0.9 * nz)
        s = rng.uniform(*sigma)
        a = rng.uniform(*amp)
        gx = np.exp(-((x - cx) ** 2) / (2 * s * s))
        gy = np.exp(-((y - cy) ** 2) / (2 * s * s))
        gz = np.exp(-((z - cz) ** 2) / (2 * s * s))
        img += a * (gz * gy * gx)
    return img.astype(np.float32)


def sinusoidal_displacement(shape, max_disp=3.0, n_modes=4, seed=0):
    """u[3, z, y, x] with max |u| = max_disp (voxels)."""
    nz, ny, nx = shape
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    z = np.arange(nz, dtype=np.float64)[:, None, None] / nz
    y = np.arange(ny, dtype=np.float64)[None, :, None] / ny
    x = np.arange(nx, dtype=np.float64)[None, None, :] / nx
    u = np.zeros((3,) + tuple(shape), dtype=np.float64)
    for c in range(3):
        for _ in range(n_modes):
            k = rng.integers(1, 4, size=3)
            phi = rng.uniform(0, 2 * np.pi)
            a = rng.uniform(0.5, 1.0)
            u[c] += a * np.sin(2 * np.pi * (k[0] * x + k[1] * y + k[2] * z) + phi)
    norm = np.sqrt((u ** 2).sum(axis=0)).max()
    u *= max_disp / norm
    return u.astype(np.float32)


def rician(img, sigma, seed=0):
    """Rician noise: |img + n1 + i n2| with n1, n2 ~ N(0, sigma^2)."""
    rng = np.random.default_rng(seed)
    n1 = rng.normal(0.0, sigma, img.shape)
    n2 = rng.normal(0.0, sigma, img.shape)
    return np.sqrt((img.astype(np.float64) + n1) ** 2 + n2 ** 2).astype(np.float32)


def identity(shape):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([x, y, z]).astype(np.float32)


def smooth_noise_field(shape, sigma=2.0, seed=1):
    """3 x N(0,1) noise, Gaussian-smoothed with sigma (separable, FP64)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty((3,) + tuple(shape), dtype=np.float32)
    r = int(np.ceil(3 * sigma))
    k = np.exp(-np.arange(-r, r + 1) ** 2 / (2 * sigma * sigma))
    k /= k.sum()
    for c in range(3):
        a = rng.standard_normal(shape)
        for ax in range(3):
            a = _conv_axis(a, k, ax, r)
        out[c] = a
    return out


def _conv_axis(a, k, ax, r):
    a = np.moveaxis(a, ax, -1)
    n = a.shape[-1]
    p = np.take(a, np.arange(-r, n + r) % n, axis=-1)  # periodic padding, any n
    res = np.zeros_like(a)
    for i, kv in enumerate(k):
        res += kv * p[..., i:i + n]
    return np.moveaxis(res, -1, ax)


def pair(shape, seed=0, max_disp=3.0):
    """(template, hfield x + u, u) for the pairwise workload; target = template(x + u)."""
    t = blobs(shape, seed=seed)
    u = sinusoidal_displacement(shape, max_disp=max_disp, seed=seed)
    h = identity(shape) + u
    return t, h, u
