"""Parity of the HIP walk on open meshes (absorbing and tangential border conditions, through the C-ABI) with the
oracle's restatement of openMeshSpace::transportParticleAndVectors and its border handlers
(src/models/openMeshSpace.cpp, absorbingOpenMeshSpace.cpp, tangentialOpenMeshSpace.cpp), and of a particle system that
steps on an open mesh.

Tolerances: walk positions and transported vectors 1e-10 (both FP64, the same formulas in a different operation
order); system steps as the closed-mesh system tests (positions 1e-8, velocities 1e-6 of their scale)."""
import numpy as np
import pytest

import curvedspacesim_amd as csa
import small_meshes as sm
from curvedspacesim_amd import workloads as pw
import workloads as wl
from oracle.pyoracle import OracleMesh, OracleSystem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = csa.Context(0)
    yield c
    c.close()


def _mesh(name):
    if name == "flat":
        return sm.grid(6, 6)
    if name == "cut_sphere":
        # the bumpy sphere without its polar cap (faces whose centroid has z >= 0.5): a curved open mesh with saddle
        # vertices and a jagged border; faces that would leave a border vertex with a single face are dropped too
        X, T = csa.bumpy_sphere(4, 20, 0.15, 0.3, 13377)
        T = T[X[T].mean(1)[:, 2] < 0.5]
        while True:
            lone = (np.bincount(T.ravel(), minlength=len(X))[T] == 1).any(1)
            if not lone.any():
                break
            T = T[~lone]
        used = np.unique(T)
        remap = -np.ones(len(X), int)
        remap[used] = np.arange(len(used))
        return X[used], remap[T]
    return sm.grid(24, 24, 0.25, 0.25, z=lambda x, y: 0.3 * np.sin(1.3 * x) * np.cos(0.9 * y))


@pytest.mark.parametrize("mode", ["absorbing", "tangential"])
@pytest.mark.parametrize("mesh", ["flat", "heightfield", "cut_sphere"])
@pytest.mark.parametrize("length", [0.3, 2.0, 8.0])
def test_walk_matches_oracle(ctx, rng, mode, mesh, length):
    X, T = _mesh(mesh)
    om = OracleMesh(X, T)
    om.set_boundary(mode)
    ctx.upload_mesh(X, T)
    ctx.set_boundary(mode)
    n = 2000
    f, b = wl.area_uniform(X, T, n, rng)
    v = wl.tangent_frame_velocity(X, T, f, rng)
    v /= np.linalg.norm(v, axis=1)[:, None]
    disp = v * length * rng.uniform(0.2, 1.0, n)[:, None]
    vecs = np.stack([v, rng.normal(size=(n, 3))], 1)
    print(f"case open walk {mode} {mesh} length={length} n={n}", flush=True)
    try:
        gf, gb, gd, gv, gs = ctx.transport_batch(f, b, disp, vecs)
    finally:
        ctx.set_boundary("closed")
    of, ob, od, ov, os_ = om.transport(f, b, disp, vecs)
    assert np.array_equal(gs, os_)
    ok = os_ >= 0
    assert ok.mean() > 0.99
    assert (gf == of)[ok].mean() > 0.999
    np.testing.assert_allclose(wl.positions(X, T, gf[ok], gb[ok]), wl.positions(X, T, of[ok], ob[ok]), atol=1e-10)
    np.testing.assert_allclose(gv[ok], ov[ok], atol=1e-10)
    np.testing.assert_allclose(gd[ok], od[ok], atol=1e-10)


@pytest.mark.parametrize("mode", ["absorbing", "tangential"])
def test_border_vertex_constructed(ctx, mode):
    """The hand-worked border-vertex walk of the oracle tests (exact dyadic inputs): from face 36 of the unit grid at
    barycentrics (1/4, 1/2, 1/4) with displacement (-1.5, -0.5) through the border vertex (0, 3). Absorbing: stop on
    the vertex in face 25; tangential: end at (0.5, 2.0) in face 24 (the next target is placed with the previous face's
    corner copy, as the reference does). A vector with x <= 0 loses its x part."""
    X, T = sm.grid(6, 6)
    ctx.upload_mesh(X, T)
    ctx.set_boundary(mode)
    vecs = np.array([[[-0.7, 0.2, 0.0], [0.4, 0.5, 0.0]]])
    try:
        gf, gb, gd, gv, gs = ctx.transport_batch([36], [[0.25, 0.5, 0.25]], [[-1.5, -0.5, 0.0]], vecs)
    finally:
        ctx.set_boundary("closed")
    assert gs[0] >= 1
    P = wl.positions(X, T, gf, gb)[0]
    if mode == "absorbing":
        assert gf[0] == 25
        np.testing.assert_allclose(P, [0.0, 3.0, 0.0], atol=1e-10)
    else:
        assert gf[0] == 24
        np.testing.assert_allclose(P, [0.5, 2.0, 0.0], atol=1e-10)
    np.testing.assert_allclose(gv[0, 0], [0.0, 0.2, 0.0], atol=1e-15)
    np.testing.assert_allclose(gv[0, 1], [0.4, 0.5, 0.0], atol=1e-15)


def test_closed_mode_still_reports_the_border(ctx):
    X, T = sm.grid(3, 3)
    ctx.upload_mesh(X, T)
    ctx.set_boundary("closed")
    gf, gb, gd, gv, gs = ctx.transport_batch([4], [[0.3, 0.3, 0.4]], [[10.0, 0.0, 0.0]])
    assert gs[0] < 0 and gf[0] == 4
    with pytest.raises(ValueError):
        ctx.set_boundary("periodic")


@pytest.mark.parametrize("mode", ["absorbing", "tangential"])
@pytest.mark.parametrize("mesh", ["heightfield", "cut_sphere"])
def test_system_on_open_mesh_matches_oracle(mode, mesh):
    """Harmonic discs on an open mesh, velocity Verlet: forces (geodesics near the border included) and three steps in
    which particles meet the border, against the oracle's system with the same border condition."""
    X, T = _mesh(mesh)
    r = np.random.default_rng(11)
    n = 300
    f, b = wl.area_uniform(X, T, n, r)
    v = pw.tangent_velocities(X, T, f, 4.0, r)
    pot = (csa.HARMONIC, 1.0 / 0.3 ** 2, 0.3)
    dt = 0.05
    c = csa.Context(0)
    try:
        c.upload_mesh(X, T)
        c.set_submeshing(True, 0.3)
        c.set_boundary(mode)
        sysm = csa.System(c, f, b, v)
        sysm.configure(pot, 0.3, True, 0.3, 0)
        om = OracleMesh(X, T)
        om.set_boundary(mode)
        osys = OracleSystem(om, f, b, v, pot, 0.3, True, True, 0.3, threads=8)
        print(f"case open-mesh system {mode} {mesh} N={n}", flush=True)
        for _ in range(3):
            sysm.nve(dt, 1)
            osys.nve_step(dt)
        gf, gb, gv, gF = sysm.get()
        np.testing.assert_allclose(om.positions(gf, gb), om.positions(osys.face, osys.bary), atol=1e-8)
        np.testing.assert_allclose(gv, osys.vel, atol=1e-6 * np.abs(osys.vel).max())
        st = sysm.status()
        assert st["walk_failures"] == 0 and st["unsolved"] == 0
        if mesh == "heightfield":
            # some particles did meet the border: they sit on it now
            P = om.positions(gf, gb)
            on_border = (np.minimum(P[:, 0], P[:, 1]) < 1e-9) | (np.maximum(P[:, 0], P[:, 1]) > 6 - 1e-9)
            assert on_border.sum() > 0
    finally:
        c.close()
