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
    if k == 5:
        return replicas(seed, count=32)
    raise ValueError(f"unknown configuration {k}")


REPLICA_MESHES = ("icosphere L5", "jittered torus", "bumpy sphere L6")


def replica_mesh(kind, seed):
    if kind == 0:
        return nat.icosphere(5)
    if kind == 1:
        return nat.jittered_torus(3.0, 1.0, 200, 100, 0.02, seed)
    return nat.bumpy_sphere(6, 20, 0.15, 0.3, seed)


def replicas(seed=13377, count=32, n_min=256, n_max=8192, spacing=12.0, kinds=None):
    """Many uneven independent systems stepped together: `count` replicas, each on a seeded mesh from {icosphere L5,
    jittered torus (40,000 faces), bumpy sphere L6}, with N drawn log-uniformly from [n_min, n_max], area-uniform
    placement, harmonic repulsion with sigma ~ U(0.05, 0.3) (patch radius and range sigma), velocities at T = 0.1, VV-NVE.
    The replicas' meshes are uploaded as one union, replica r translated by r * spacing along x so that no two replicas'
    particles are Euclidean neighbours; each replica is a particle group (csg_system_set_groups). meta["replicas"]
    keeps every replica's own mesh and index ranges."""
    rng = np.random.default_rng(seed)
    Xs, Ts, fs, bs, vs, groups, pots, ranges, reps = [], [], [], [], [], [], [], [], []
    v0 = f0 = p0 = 0
    for r in range(count):
        kind = int(rng.integers(0, 3)) if kinds is None else int(kinds[r % len(kinds)])
        X, T = replica_mesh(kind, seed + 1 + r)
        n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        sigma = float(rng.uniform(0.05, 0.3))
        f, b = area_uniform(X, T, n, rng)
        v = tangent_velocities(X, T, f, 0.1, rng)
        Xs.append(X + np.array([r * spacing, 0.0, 0.0]))
        Ts.append(T + v0)
        fs.append(f + f0)
        bs.append(b)
        vs.append(v)
        groups.append(np.full(n, r, np.int32))
        pots.append((nat.HARMONIC, 1.0 / sigma ** 2, sigma))
        ranges.append(sigma)
        reps.append(dict(mesh=REPLICA_MESHES[kind], X=X, T=T, n=n, sigma=sigma, vert_offset=v0, face_offset=f0,
                         particle_offset=p0))
        v0 += len(X)
        f0 += len(T)
        p0 += n
    smax = max(ranges)
    return Workload(f"{count} replicas on icosphere L5 / jittered torus / bumpy sphere L6, N={p0} in total "
                    f"(log-uniform {n_min}-{n_max} each), harmonic sigma ~ U(0.05, 0.3), VV-NVE",
                    np.concatenate(Xs), np.concatenate(Ts).astype(np.int32), np.concatenate(fs).astype(np.int32),
                    np.concatenate(bs), np.concatenate(vs), (nat.HARMONIC, 1.0 / smax ** 2, smax), smax, True, True, smax,
                    smax, "nve", 0.005, 0.1,
                    meta=dict(groups=np.concatenate(groups), pots=pots, ranges=ranges, maxd=list(ranges), replicas=reps))


def make_system(ctx, w, max_patch_faces=0):
    ctx.upload_mesh(w.X, w.T)
    ctx.set_submeshing(w.submeshing, w.maximum_distance)
    sysm = nat.System(ctx, w.face, w.bary, w.vel)
    sysm.configure(w.potential, w.range, w.cell_list, w.rmax, max_patch_faces)
    if w.meta.get("groups") is not None:
        sysm.set_groups(w.meta["groups"], w.meta["pots"], w.meta["ranges"], w.meta["maxd"])
    return sysm


def step(sysm, w, steps):
    if w.integrator == "nvt":
        sysm.nvt(w.dt, w.temperature, w.tau, w.chain, steps)
    else:
        sysm.nve(w.dt, steps)
