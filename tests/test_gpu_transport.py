"""Parity of the HIP displacement walk (through the C-ABI) with the FP64 oracle restatement of
transportParticleAndVectors. Positions and transported vectors must agree to 1e-10 (both are FP64; the only
differences are operation order inside the same formulas)."""
import numpy as np
import pytest

import curvedspacesim_amd as csa
import small_meshes as sm
import walk_cases
import workloads as wl
from oracle.pyoracle import OracleMesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = csa.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("mesh", ["torus", "bumpy", "cube"])
@pytest.mark.parametrize("length", [0.02, 0.3, 2.0])
def test_walk_matches_oracle(ctx, rng, mesh, length):
    if mesh == "torus":
        X, T = csa.jittered_torus(3.0, 1.0, 100, 50, 0.02, 11)
    elif mesh == "bumpy":
        X, T = csa.bumpy_sphere(5, 20, 0.15, 0.3, 13377)
    else:
        X, T = sm.cube()
    om = OracleMesh(X, T)
    ctx.upload_mesh(X, T)
    n = 2000
    f, b = wl.area_uniform(X, T, n, rng)
    v = wl.tangent_frame_velocity(X, T, f, rng)
    v /= np.linalg.norm(v, axis=1)[:, None]
    disp = v * length
    vecs = np.stack([v, rng.normal(size=(n, 3))], 1)
    print(f"case walk {mesh} length={length} n={n}", flush=True)
    gf, gb, gd, gv, gs = ctx.transport_batch(f, b, disp, vecs)
    of, ob, od, ov, os_ = om.transport(f, b, disp, vecs)
    assert np.array_equal(gs, os_)
    ok = os_ >= 0
    # a walk that ends on an edge or vertex may be reported in either face that holds that point: every face
    # mismatch is counted and must lie on a corner or edge of both faces (a barycentric coordinate at the clamp level)
    diff = ok & (gf != of)
    on_edge = (gb.min(1) < 1e-9) & (ob.min(1) < 1e-9)
    print(f"face mismatches {int(diff.sum())} of {int(ok.sum())}, all on edges: {bool(np.all(on_edge[diff]))}")
    assert np.all(on_edge[diff])
    P1 = wl.positions(X, T, gf[ok], gb[ok])
    P2 = wl.positions(X, T, of[ok], ob[ok])
    np.testing.assert_allclose(P1, P2, atol=1e-10)
    np.testing.assert_allclose(gv[ok], ov[ok], atol=1e-10)


def test_vertex_crossing_constructed(ctx):
    """A walk aimed exactly at a vertex of a planar grid: both edges register and the vertex rule applies."""
    X, T = sm.grid(6, 6)
    om = OracleMesh(X, T)
    ctx.upload_mesh(X, T)
    # from the centre of face 14 straight through the grid vertex at (3, 3)
    f0 = 2 * (2 * 6 + 2)
    b0 = np.array([1 / 3, 1 / 3, 1 / 3])
    S = wl.positions(X, T, np.array([f0]), b0[None])[0]
    disp = (np.array([3.0, 3.0, 0.0]) - S) * 1.8
    v = disp / np.linalg.norm(disp)
    gf, gb, gd, gv, gs = ctx.transport_batch([f0], [b0], [disp], [[v]])
    of, ob, od, ov, os_ = om.transport([f0], [b0], [disp], [[v]])
    assert gs[0] == os_[0]
    assert gf[0] == of[0]
    np.testing.assert_allclose(gb, ob, atol=1e-12)
    np.testing.assert_allclose(gv, ov, atol=1e-12)


@pytest.mark.parametrize("case", sorted(walk_cases.CASES))
def test_vertex_crossing_worked_by_hand(ctx, case):
    """The interior vertex crossings worked through the reference's steps by hand (tests/walk_cases.py)."""
    c = walk_cases.CASES[case]()
    ctx.upload_mesh(c["X"], c["T"])
    print(f"case vertex crossing {case}", flush=True)
    f2, b2, d2, v2, st = ctx.transport_batch([c["face"]], [c["bary"]], [c["disp"]], [[c["vec"]]])
    assert st[0] >= 1
    assert f2[0] == c["end_face"]
    np.testing.assert_allclose(b2[0], c["end_bary"], atol=1e-9)
    np.testing.assert_allclose(wl.positions(c["X"], c["T"], f2, b2)[0], c["end_point"], atol=1e-9)
    np.testing.assert_allclose(v2[0, 0], c["end_vec"], atol=1e-9)


def test_border_error_is_reported(ctx):
    X, T = sm.grid(3, 3)
    ctx.upload_mesh(X, T)
    gf, gb, gd, gv, gs = ctx.transport_batch([4], [[0.3, 0.3, 0.4]], [[10.0, 0.0, 0.0]])
    assert gs[0] < 0 and gf[0] == 4
