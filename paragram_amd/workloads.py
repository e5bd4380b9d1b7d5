"""Seeded synthetic point clouds for the five measured configurations.

Every generator uses ``numpy.random.default_rng(seed)`` (PCG64), draws in
float64 and casts the result once with ``.astype(np.float32)``.  The draw order
below is part of the definition: changing it changes the cloud.

Configurations
--------------
* ``uniform(n, seed)``       C1 (10k, seed 0) and C2 (1M, seed 1): U[0,1)^3.
* ``clustered_power(...)``   C3 (2M, seed 2): 90% from 64 isotropic Gaussians
  (sigma 0.02, centres U[0,1]^3, clamped to the cube) + 10% uniform, permuted;
  weights U[0, 3e-5].
* ``lidar_batch(seed)``      C4 (256 clouds, seed 3): LiDAR-scan-like rings.
* ``learned_power(...)``     C5 (10M, seed 4): 50% uniform + 50% from 256
  Gaussians (sigma 0.01, clamped), permuted; weights U[0, 5e-6].
"""
from __future__ import annotations

import math

import numpy as np


def uniform(n: int, seed: int) -> np.ndarray:
    """``n`` points i.i.d. uniform in [0,1)^3 as float32 ``(n, 3)``."""
    rng = np.random.default_rng(seed)
    return rng.random((n, 3)).astype(np.float32)


def _gaussian_mixture(rng, n: int, n_centres: int, sigma: float) -> np.ndarray:
    centres = rng.random((n_centres, 3))
    idx = rng.integers(0, n_centres, n)
    pts = centres[idx] + sigma * rng.standard_normal((n, 3))
    return np.clip(pts, 0.0, 1.0)


def clustered_power(n: int = 2_000_000, seed: int = 2, w_max: float = 3e-5):
    """C3: 90% clustered (64 Gaussians, sigma 0.02, clamped) + 10% uniform.

    Returns ``(points float32 (n,3), weights float32 (n,))``.
    """
    rng = np.random.default_rng(seed)
    n_cl = (n * 9) // 10
    clustered = _gaussian_mixture(rng, n_cl, 64, 0.02)
    unif = rng.random((n - n_cl, 3))
    pts = np.concatenate([clustered, unif], axis=0)
    pts = pts[rng.permutation(n)]
    w = rng.uniform(0.0, w_max, n)
    return pts.astype(np.float32), w.astype(np.float32)


def learned_power(n: int = 10_000_000, seed: int = 4, w_max: float = 5e-6):
    """C5: 50% uniform + 50% from 256 Gaussians (sigma 0.01, clamped), permuted.

    Returns ``(points float32 (n,3), weights float32 (n,))``.
    """
    rng = np.random.default_rng(seed)
    n_u = n // 2
    unif = rng.random((n_u, 3))
    clustered = _gaussian_mixture(rng, n - n_u, 256, 0.01)
    pts = np.concatenate([unif, clustered], axis=0)
    pts = pts[rng.permutation(n)]
    w = rng.uniform(0.0, w_max, n)
    return pts.astype(np.float32), w.astype(np.float32)


def lidar_sizes(seed: int = 3, n_clouds: int = 256) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(math.log(2e4), math.log(2e5), n_clouds))).astype(np.int64)


def lidar_batch(seed: int = 3, n_clouds: int = 256, max_clouds: int | None = None):
    """C4: LiDAR-scan-like clouds.

    Sizes are log-uniform in [2e4, 2e5].  Per point: ring k in [0,64), elevation
    e = -25deg + 28deg*k/63, azimuth U[0, 2pi); range = ground hit at z=-1.8 m
    capped at 80 m when e < 0, else U[5, 80] m; plus N(0, 0.02) range noise and
    N(0, 1e-4) per-coordinate jitter.  Returns a list of float32 ``(n_i, 3)``.
    """
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(math.log(2e4), math.log(2e5), n_clouds))).astype(np.int64)
    clouds = []
    for i, n in enumerate(sizes):
        if max_clouds is not None and i >= max_clouds:
            break
        n = int(n)
        k = rng.integers(0, 64, n)
        a = rng.uniform(0.0, 2.0 * math.pi, n)
        r_far = rng.uniform(5.0, 80.0, n)
        noise = rng.normal(0.0, 0.02, n)
        jitter = rng.normal(0.0, 1e-4, (n, 3))
        e = np.deg2rad(-25.0 + 28.0 * k / 63.0)
        with np.errstate(divide="ignore"):
            r_ground = np.minimum(1.8 / np.sin(-e), 80.0)
        r = np.where(e < 0.0, r_ground, r_far) + noise
        xyz = np.stack([r * np.cos(e) * np.cos(a), r * np.cos(e) * np.sin(a), r * np.sin(e)], axis=1) + jitter
        clouds.append(xyz.astype(np.float32))
    return clouds


def oracle_sample(n: int, seed: int, count: int = 100_000) -> np.ndarray:
    """Seeded sample of cell ids checked against the oracle above 1M points."""
    rng = np.random.default_rng(seed + 10_000)
    if count >= n:
        return np.arange(n, dtype=np.int64)
    return np.sort(rng.choice(n, count, replace=False)).astype(np.int64)
