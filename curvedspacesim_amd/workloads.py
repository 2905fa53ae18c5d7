"""Synthetic workloads of the benchmark configurations: seeded meshes, particles and velocities.

config 1: unit icosphere L4 (2,562 V / 5,120 F), N = 64 placed as triangulatedMeshSpace::randomPosition does
          (uniform face index, u ~ U(0,1), v ~ U(0,1-u); src/models/triangulatedMeshSpace.cpp:108-116), all-to-all
          exact geodesics on the whole mesh, harmonic U = (k/2)(1 - d/sigma)^2 with sigma = 0.3, k = 1 (the reference's
          harmonicRepulsion stiffness is k/sigma^2 in its (sigma - d) convention), velocity Verlet dt = 0.005, T = 0.1.
config 2: torus R = 3, r = 1, 200 x 100 quads as 40,000 triangles, vertex coordinates jittered by U(+-0.02 l),
          N = 2,000 uniform by area, harmonic sigma = 0.25, patches of radius sigma, velocity Verlet dt = 0.005.
config 3: bumpy sphere r = 1 + 0.15 sum_k exp(-g_k^2 / (2 * 0.3^2)) over 20 seeded bumps on icosphere L6 (81,920 F),
          N = 20,000 uniform by area, truncated Gaussian epsilon = 1, s = 0.05, cutoff 3s = 0.15, patches of radius
          0.15, Nose-Hoover chain (M = 2, tau = 1) at T = 0.5, dt = 0.005.
Velocities follow randomVectorAtPosition + setMaxwellBoltzmannVelocities (triangulatedMeshSpace.cpp:118-138,
simpleModel.cpp:223-230): g1 * o + g2 * (n x o) with g1, g2 ~ N(0,1) and o the reference's in-plane orthogonaliser,
scaled by sqrt(T). Draws use numpy's PCG64 with the given seed.
"""
from dataclasses import dataclass, field

import numpy as np

from . import _native as nat


@dataclass
class Workload:
    name: str
    X: np.ndarray
    T: np.ndarray
    face: np.ndarray
    bary: np.ndarray
    vel: np.ndarray
    potential: tuple
    range: float
    cell_list: bool
    submeshing: bool
    maximum_distance: float
    rmax: float
    integrator: str
    dt: float
    temperature: float
    tau: float = 1.0
    chain: int = 2
    meta: dict = field(default_factory=dict)


def face_normals(X, T):
    n = np.cross(X[T[:, 1]] - X[T[:, 0]], X[T[:, 2]] - X[T[:, 0]])
    return n / np.linalg.norm(n, axis=1)[:, None]


def area_uniform(X, T, n, rng):
    A = 0.5 * np.linalg.norm(np.cross(X[T[:, 1]] - X[T[:, 0]], X[T[:, 2]] - X[T[:, 0]]), axis=1)
    cdf = np.cumsum(A) / A.sum()
    f = np.minimum(np.searchsorted(cdf, rng.uniform(size=n)), len(T) - 1).astype(np.int32)
    s = np.sqrt(rng.uniform(size=n))
    r2 = rng.uniform(size=n)
    return f, np.stack([1 - s, s * (1 - r2), s * r2], 1)


def reference_placement(nF, n, rng):
    u = rng.uniform(size=n)
    v = rng.uniform(size=n) * (1 - u)
    f = rng.integers(0, nF, size=n).astype(np.int32)
    return f, np.stack([u, v, 1 - u - v], 1)


def tangent_velocities(X, T, f, temperature, rng):
    n = face_normals(X, T)[f]
    orth = np.stack([n[:, 1], -n[:, 0], np.zeros(len(f))], 1)
    polar = (n[:, 0] == 0) & (n[:, 1] == 0)
    orth[polar] = np.stack([n[polar, 1] - n[polar, 2], n[polar, 2] - n[polar, 0], n[polar, 0] - n[polar, 1]], 1)
    orth /= np.linalg.norm(orth, axis=1)[:, None]
    t2 = np.cross(n, orth)
    g = rng.normal(size=(len(f), 2))
    return (g[:, :1] * orth + g[:, 1:] * t2) * np.sqrt(temperature)


def config(k, seed=13377, n=None):
    rng = np.random.default_rng(seed)
    if k == 1:
        X, T = nat.icosphere(4)
        N = n or 64
        f, b = reference_placement(len(T), N, rng)
        sigma = 0.3
        return Workload("icosphere L4 (5,120 faces), N=64, all-to-all, harmonic sigma=0.3, VV-NVE", X, T, f, b,
                        tangent_velocities(X, T, f, 0.1, rng), (nat.HARMONIC, 1.0 / sigma ** 2, sigma), sigma, False,
                        False, sigma, sigma, "nve", 0.005, 0.1)
    if k == 2:
        X, T = nat.jittered_torus(3.0, 1.0, 200, 100, 0.02, seed)
        N = n or 2000
        f, b = area_uniform(X, T, N, rng)
        sigma = 0.25
        return Workload("jittered torus (40,000 faces), N=2,000, harmonic sigma=0.25, patches, VV-NVE", X, T, f, b,
                        tangent_velocities(X, T, f, 0.1, rng), (nat.HARMONIC, 1.0 / sigma ** 2, sigma), sigma, True,
                        True, sigma, sigma, "nve", 0.005, 0.1)
    if k == 3:
        X, T = nat.bumpy_sphere(6, 20, 0.15, 0.3, seed)
        N = n or 20000
        f, b = area_uniform(X, T, N, rng)
        s, rc = 0.05, 0.15
        return Workload("bumpy sphere L6 (81,920 faces), N=20,000, truncated Gaussian s=0.05 rc=0.15, NH-NVT T=0.5", X, T,
                        f, b, tangent_velocities(X, T, f, 0.5, rng), (nat.TRUNC_GAUSSIAN, 1.0, s, rc), rc, True, True,
                        rc, rc, "nvt", 0.005, 0.5, 1.0, 2)
    raise ValueError(f"unknown configuration {k}")


def make_system(ctx, w, max_patch_faces=0):
    ctx.upload_mesh(w.X, w.T)
    ctx.set_submeshing(w.submeshing, w.maximum_distance)
    sysm = nat.System(ctx, w.face, w.bary, w.vel)
    sysm.configure(w.potential, w.range, w.cell_list, w.rmax, max_patch_faces)
    return sysm


def step(sysm, w, steps):
    if w.integrator == "nvt":
        sysm.nvt(w.dt, w.temperature, w.tau, w.chain, steps)
    else:
        sysm.nve(w.dt, steps)
