"""Synthetic inputs of the benchmark workloads: seeded ellipse phantoms, relative Gaussian noise, integer images.

The ellipse rasteriser follows the Shepp-Logan pattern of exp_shepp.py:44-50 (an ``indexing="ij"`` meshgrid of
linspace(-1, 1, N), the rotated indicator (xr/a)^2 + (yr/b)^2 <= 1, intensities added).  The noise normalisation
is the one of exp_compiled_tables.py:52-53: N(0,1) over the whole (4, 2N-1, N) array, structural zeros included,
scaled by level * ||b0|| / ||nz||.
"""

import numpy as np


def ellipses(seed, N, count=10):
    """Sum of ``count`` random ellipses on an N x N grid of [-1, 1]^2.

    Draw order (one numpy default_rng(seed)): u, v (centre sqrt(u) (cos 2 pi v, sin 2 pi v), uniform in the unit
    disc), semi-axes a, b ~ U[0.05, 0.4], angle ~ U[0, pi), intensity ~ U[-0.5, 1]; each a vector of ``count``."""
    rng = np.random.default_rng(seed)
    u = rng.random(count)
    v = rng.random(count)
    ax = rng.uniform(0.05, 0.4, count)
    bx = rng.uniform(0.05, 0.4, count)
    phi = rng.uniform(0.0, np.pi, count)
    amp = rng.uniform(-0.5, 1.0, count)
    x0 = np.sqrt(u) * np.cos(2 * np.pi * v)
    y0 = np.sqrt(u) * np.sin(2 * np.pi * v)
    y, x = np.meshgrid(np.linspace(-1, 1, N), np.linspace(-1, 1, N), indexing="ij")
    img = np.zeros((N, N))
    for i in range(count):
        c, s = np.cos(phi[i]), np.sin(phi[i])
        xr = (x - x0[i]) * c + (y - y0[i]) * s
        yr = -(x - x0[i]) * s + (y - y0[i]) * c
        img[(xr / ax[i]) ** 2 + (yr / bx[i]) ** 2 <= 1.0] += amp[i]
    return img


def add_noise(b0, level, seed):
    """b0 + level * ||b0|| / ||nz|| * nz with nz ~ N(0, 1) of b0's shape (sigma = level * RMS(b0))."""
    if level == 0:
        return np.array(b0, dtype=float)
    nz = np.random.default_rng(seed).standard_normal(np.shape(b0))
    return b0 + level * np.linalg.norm(b0) / np.linalg.norm(nz) * nz


def int_image(seed, N):
    """Integer image uniform in [0, 255] (int32)."""
    return np.random.default_rng(seed).integers(0, 256, (N, N), dtype=np.int32)
