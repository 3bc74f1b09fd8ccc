"""Seeded synthetic workloads for the LBVH, traversal and render paths.

Every generator draws from ``numpy.random.Generator(PCG64(seed))`` so the same
seed gives the same float32 arrays on every machine with this numpy.

Triangle soups are ``float32[N, 9]`` rows ``v0.xyz, v1.xyz, v2.xyz``; rays are
``float32[R, 6]`` rows ``origin.xyz, direction.xyz``.
"""
from __future__ import annotations

import numpy as np


def _rng(seed: int, stream: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64([seed, stream]))


def _unit_sphere(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform directions on S^2 (z uniform in [-1,1], phi uniform)."""
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def gmm_triangles(n: int, seed: int = 1, components: int = 8, edge: float = 1e-3) -> np.ndarray:
    """Random triangle soup: centroids from a Gaussian mixture in the unit cube,
    vertex offsets of log-normal length (mu = ln edge, sigma = 0.7; edge = 1e-3
    by default) and uniform orientation. Returns float32[n, 9]."""
    rng = _rng(seed, 0)
    w = rng.uniform(0.5, 1.5, components)
    w /= w.sum()
    means = rng.uniform(0.15, 0.85, (components, 3))
    sig = rng.uniform(0.03, 0.12, components)
    comp = rng.choice(components, size=n, p=w)
    c = means[comp] + rng.standard_normal((n, 3)) * sig[comp, None]
    np.clip(c, 0.0, 1.0, out=c)
    out = np.empty((n, 9), dtype=np.float32)
    for j in range(3):
        length = rng.lognormal(np.log(edge), 0.7, n)
        u = _unit_sphere(rng, n)
        out[:, 3 * j:3 * j + 3] = (c + length[:, None] * u).astype(np.float32)
    return out


def random_rays(n: int, seed: int = 1, stream: int = 1, lo: float = -0.5,
                hi: float = 1.5) -> np.ndarray:
    """Origins uniform in [lo,hi]^3, directions uniform on S^2. float32[n, 6]."""
    rng = _rng(seed, stream)
    o = rng.uniform(lo, hi, (n, 3))
    d = _unit_sphere(rng, n)
    return np.concatenate([o, d], axis=1).astype(np.float32)


def shadow_limits(n: int, seed: int = 1, stream: int = 3) -> np.ndarray:
    """Shadow-ray distance limits uniform in (0, 2*sqrt(3)). float32[n]."""
    rng = _rng(seed, stream)
    return rng.uniform(0.0, 2.0 * np.sqrt(3.0), n).astype(np.float32)
