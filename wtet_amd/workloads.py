"""Synthetic point sets of the benchmark configurations (f32, seeded).

uniform      : i.i.d. U[0,1)^3, numpy PCG64 ``default_rng(seed).random(..., float32)``
lattice      : k^3 grid i/(k-1), ix-major, iz fastest (src/tests.rs:1103-1117)
sphere       : i.i.d. Gaussian triples (f64) normalized to the unit sphere, cast to f32
shell        : points in a thin shell of radius 0.4 around (0.5, 0.5, 0.5)
clusters     : mixture of isotropic Gaussians, samples outside [0,1]^3 redrawn
anisotropic  : mixture of k anisotropic Gaussians (random rotations,
               covariance eigenvalues U[1e-4, 1e-2]) plus a uniform share,
               samples outside [0,1]^3 redrawn
mixed_batch  : many clouds, log-normal sizes (median 2000, sigma 0.63, clipped
               to [64, 50000]), each a ``clusters`` cloud with K ~ U{1..8}
               Gaussians, sigma_k ~ U[0.02, 0.2]; returns (points, offsets)
"""
from __future__ import annotations

import hashlib

import numpy as np


def uniform(n: int, seed: int) -> np.ndarray:
    return np.random.default_rng(seed).random((n, 3), dtype=np.float32)


def lattice(k: int) -> np.ndarray:
    i = np.arange(k, dtype=np.float32) / np.float32(max(k, 2) - 1)
    x, y, z = np.meshgrid(i, i, i, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32)


def sphere(n: int, seed: int) -> np.ndarray:
    g = np.random.default_rng(seed).standard_normal((n, 3))
    return (g / np.linalg.norm(g, axis=1, keepdims=True)).astype(np.float32)


def shell(n: int, seed: int, radius: float = 0.4, thickness: float = 1e-4) -> np.ndarray:
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((n, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    r = radius + (rng.random(n) - 0.5) * thickness
    return (0.5 + g * r[:, None]).astype(np.float32)


def clusters(n: int, seed: int, k: int | None = None, sigma=(0.02, 0.2)) -> np.ndarray:
    rng = np.random.default_rng(seed)
    k = k or int(rng.integers(1, 9))
    centers = rng.random((k, 3))
    sig = rng.uniform(sigma[0], sigma[1], k)
    out = np.empty((0, 3))
    while len(out) < n:
        c = rng.integers(0, k, n)
        s = centers[c] + rng.standard_normal((n, 3)) * sig[c, None]
        s = s[np.all((s >= 0) & (s <= 1), axis=1)]
        out = np.vstack([out, s])
    return out[:n].astype(np.float32)


def anisotropic(n: int, seed: int, k: int = 32, uniform_share: float = 0.1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    nu = int(round(n * uniform_share))
    centers = rng.random((k, 3))
    rots = []
    for _ in range(k):
        q, r = np.linalg.qr(rng.standard_normal((3, 3)))
        rots.append(q * np.sign(np.diag(r)))
    scales = np.sqrt(rng.uniform(1e-4, 1e-2, (k, 3)))
    out = np.empty((0, 3))
    m = n - nu
    while len(out) < m:
        c = rng.integers(0, k, m)
        z = rng.standard_normal((m, 3)) * scales[c]
        s = centers[c] + np.einsum("nij,nj->ni", np.stack(rots)[c], z)
        s = s[np.all((s >= 0) & (s <= 1), axis=1)]
        out = np.vstack([out, s])
    pts = np.vstack([out[:m], rng.random((nu, 3))])
    return pts[rng.permutation(n)].astype(np.float32)


def mixed_batch(num_clouds: int, seed: int, median: float = 2000.0, sigma: float = 0.63,
                lo: int = 64, hi: int = 50000):
    """(points (N, 3) f32, offsets (C+1,) u64): cloud c is points[off[c]:off[c+1]],
    drawn by ``clusters(size_c, seed * 1_000_003 + c)``."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.rint(np.exp(rng.normal(np.log(median), sigma, num_clouds))), lo, hi).astype(np.int64)
    clouds = [clusters(int(sizes[c]), seed * 1_000_003 + c) for c in range(num_clouds)]
    off = np.zeros(num_clouds + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    return np.ascontiguousarray(np.concatenate(clouds), np.float32), off


def sha256(points: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(points, np.float32).tobytes()).hexdigest()
