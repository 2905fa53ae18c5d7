"""Seeded synthetic inputs shared by the parity tests (particles on meshes, candidate lists)."""
import numpy as np

from oracle import pyoracle


def face_areas(X, T):
    return 0.5 * np.linalg.norm(np.cross(X[T[:, 1]] - X[T[:, 0]], X[T[:, 2]] - X[T[:, 0]]), axis=1)


def area_uniform(X, T, n, rng):
    """Particles uniform by area: face from the area CDF, corner weights by the square-root rule (Osada et al. 2002)."""
    A = face_areas(X, T)
    cdf = np.cumsum(A) / A.sum()
    f = np.minimum(np.searchsorted(cdf, rng.uniform(size=n)), len(T) - 1).astype(np.int32)
    r1, r2 = rng.uniform(size=n), rng.uniform(size=n)
    s = np.sqrt(r1)
    b = np.stack([1 - s, s * (1 - r2), s * r2], 1)
    return f, b


def positions(X, T, f, b):
    return (X[T[f]] * b[:, :, None]).sum(1) / b.sum(1)[:, None]


def candidates(X, T, f, b, rc):
    """Reference cell-list candidates (box seeded at the origin) and per-source max candidate distance."""
    P = positions(X, T, f, b)
    lo = np.minimum(X.min(0), 0.0)
    hi = np.maximum(X.max(0), 0.0)
    return pyoracle.cell_candidates(P, lo, hi, rc, rc)


def tangent_frame_velocity(X, T, f, rng, scale=1.0):
    A = X[T[f, 0]]
    n = np.cross(X[T[f, 1]] - A, X[T[f, 2]] - A)
    n /= np.linalg.norm(n, axis=1)[:, None]
    v = rng.normal(size=(len(f), 3))
    v -= (v * n).sum(1)[:, None] * n
    return v * scale
