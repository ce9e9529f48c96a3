"""Synthetic workloads (numpy, seeded) for the benchmark configurations.

Surface generator G(n_u, n_v, p, q): control points x_i = i/(n_u-1), y_j = j/(n_v-1),
z = 0.2 sin(2 pi x) cos(2 pi y) + N(0, 0.05); weights U[0.5, 2]; clamped knots on [0, 1] with
uniform interior knots k/(n-p) (or jittered by up to +-0.4 of the spacing for the uneven batch).
Knots stay on [0, 1], the domain on which the reference's derivative scaling
(src/NURBS/Evaluate.cu:120) is exact.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SurfaceDef:
    p: int
    q: int
    ku: np.ndarray
    kv: np.ndarray
    cp: np.ndarray  # (nu, nv, 3), u-major
    w: np.ndarray   # (nu, nv)

    @property
    def nu(self):
        return self.w.shape[0]

    @property
    def nv(self):
        return self.w.shape[1]


@dataclass
class CurveDef:
    p: int
    k: np.ndarray
    cp: np.ndarray  # (n, 3)
    w: np.ndarray   # (n,)


def clamped_knots(n: int, p: int, rng=None, jitter: float = 0.0) -> np.ndarray:
    """Clamped knot vector on [0, 1] with n - p - 1 interior knots k/(n-p), optionally jittered
    uniformly by up to +-jitter*spacing (kept strictly increasing)."""
    m = n - p  # number of spans
    interior = np.arange(1, m, dtype=np.float64) / m
    if jitter and m > 1:
        interior = interior + rng.uniform(-jitter, jitter, size=interior.shape) / m
        interior = np.clip(interior, 1e-6, 1 - 1e-6)
        interior.sort()
    return np.concatenate([np.zeros(p + 1), interior, np.ones(p + 1)])


def surface_G(nu: int, nv: int, p: int, q: int, seed: int, jitter: float = 0.0) -> SurfaceDef:
    rng = np.random.default_rng(seed)
    x = np.arange(nu, dtype=np.float64) / (nu - 1)
    y = np.arange(nv, dtype=np.float64) / (nv - 1)
    X, Y = np.meshgrid(x, y, indexing="ij")
    Z = 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + rng.normal(0.0, 0.05, size=X.shape)
    cp = np.stack([X, Y, Z], axis=-1)
    w = rng.uniform(0.5, 2.0, size=(nu, nv))
    ku = clamped_knots(nu, p, rng, jitter)
    kv = clamped_knots(nv, q, rng, jitter)
    return SurfaceDef(p, q, ku, kv, np.ascontiguousarray(cp), np.ascontiguousarray(w))


def uniform_uv(n: int, seed: int):
    rng = np.random.default_rng(seed)
    return rng.random(n), rng.random(n)


def offset_queries(S: np.ndarray, Su: np.ndarray, Sv: np.ndarray, sigma: float, seed: int):
    """P = S + delta * n_hat, delta ~ N(0, sigma); S, Su, Sv of shape (3, n)."""
    rng = np.random.default_rng(seed)
    nrm = np.cross(Su.T, Sv.T).T
    nrm = nrm / np.linalg.norm(nrm, axis=0, keepdims=True)
    delta = rng.normal(0.0, sigma, size=S.shape[1]) if sigma > 0 else np.zeros(S.shape[1])
    return S + delta * nrm, delta


def uneven_batch(count: int = 1024, seed: int = 1003, total_scale: float = 1.0):
    """Uneven-batch generator: degrees U{2..5}, nets U{4..48} (n >= p+1), jittered clamped knots,
    points per surface from a power law ~ N^-1.5 truncated to [256, 65536] (scaled by
    total_scale)."""
    rng = np.random.default_rng(seed)
    surfs, npts = [], []
    lo, hi, a = 256.0, 65536.0, 1.5
    for i in range(count):
        p, q = int(rng.integers(2, 6)), int(rng.integers(2, 6))
        nu = int(rng.integers(4, 49))
        while nu < p + 1:
            nu = int(rng.integers(4, 49))
        nv = int(rng.integers(4, 49))
        while nv < q + 1:
            nv = int(rng.integers(4, 49))
        surfs.append(surface_G(nu, nv, p, q, seed=seed * 100003 + i, jitter=0.4))
        r = rng.random()
        # inverse CDF of p(N) ~ N^-a on [lo, hi]
        n = (lo ** (1 - a) + r * (hi ** (1 - a) - lo ** (1 - a))) ** (1 / (1 - a))
        npts.append(max(1, int(round(n * total_scale))))
    return surfs, np.array(npts, dtype=np.int64)


def helix_curve(n: int, p: int, seed: int) -> CurveDef:
    """Cubic rational curve along a helix with the z-formula of G (for the serving objects)."""
    rng = np.random.default_rng(seed)
    s = np.arange(n, dtype=np.float64) / (n - 1)
    ang = 4 * np.pi * s
    x = 0.5 + 0.4 * np.cos(ang)
    y = 0.5 + 0.4 * np.sin(ang)
    z = 0.2 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y) + rng.normal(0.0, 0.05, size=n) + s
    w = rng.uniform(0.5, 2.0, size=n)
    return CurveDef(p, clamped_knots(n, p), np.stack([x, y, z], axis=-1), w)
