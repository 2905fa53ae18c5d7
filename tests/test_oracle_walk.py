"""Pins the oracle's displacement walk (restating triangulatedMeshSpace::transportParticleAndVectors) by exact
geometry: flat meshes, cube unfoldings, speed preservation and reversibility."""
import numpy as np
import pytest

import small_meshes as sm
import walk_cases
from oracle.pyoracle import OracleMesh


def test_flat_grid_translation(rng):
    X, T = sm.grid(10, 10, 0.3, 0.3)
    m = OracleMesh(X, T)
    n = 200
    f, b = sm.random_points(rng, len(T), n)
    P = m.positions(f, b)
    ang = rng.uniform(0, 2 * np.pi, n)
    L = rng.uniform(0, 0.6, n)
    disp = np.stack([L * np.cos(ang), L * np.sin(ang), np.zeros(n)], 1)
    goal = P + disp
    inside = (goal[:, 0] > 0.01) & (goal[:, 0] < 2.99) & (goal[:, 1] > 0.01) & (goal[:, 1] < 2.99)
    vel = np.stack([np.cos(ang), np.sin(ang), np.zeros(n)], 1)
    f2, b2, d2, v2, st = m.transport(f[inside], b[inside], disp[inside], vel[inside][:, None, :])
    assert np.all(st >= 0)
    P2 = m.positions(f2, b2)
    np.testing.assert_allclose(P2, goal[inside], atol=1e-12)
    np.testing.assert_allclose(v2[:, 0, :], vel[inside], atol=1e-13)
    assert np.all(b2 >= 0)


def test_out_of_plane_component_is_removed():
    X, T = sm.grid(4, 4)
    m = OracleMesh(X, T)
    f, b = np.array([10], np.int32), np.array([[0.3, 0.3, 0.4]])
    P = m.positions(f, b)[0]
    f2, b2, *_ = m.transport(f, b, [[0.2, 0.1, 5.0]])
    np.testing.assert_allclose(m.positions(f2, b2)[0], P + [0.2, 0.1, 0.0], atol=1e-12)


def _locate_on(X, T, p):
    for f, t in enumerate(T):
        A, B, Cc = X[t]
        n = np.cross(B - A, Cc - A)
        if abs((np.array(p) - A) @ n) > 1e-12:
            continue
        M = np.stack([B - A, Cc - A], 1)
        uv, *_ = np.linalg.lstsq(M, np.array(p) - A, rcond=None)
        w = np.array([1 - uv.sum(), uv[0], uv[1]])
        if (w >= -1e-12).all():
            return f, w
    raise ValueError(p)


def test_cube_edge_unfolding():
    """Walking off the z=0 face across the edge x=1 continues up the x=1 face; the velocity turns by 90 degrees."""
    X, T = sm.cube(1.0)
    m = OracleMesh(X, T)
    f, b = _locate_on(X, T, (0.7, 0.4, 0.0))
    disp = np.array([0.5, 0.1, 0.0])  # crosses x = 1 after 0.3 in x; 0.2 of x-travel remains
    v = disp / np.linalg.norm(disp)
    f2, b2, d2, v2, st = m.transport([f], [b], [disp], [[v]])
    assert st[0] == 1
    P2 = m.positions(f2, b2)[0]
    np.testing.assert_allclose(P2, [1.0, 0.4 + 0.1, 0.2], atol=1e-12)
    np.testing.assert_allclose(v2[0, 0], np.array([0.0, 0.1, 0.5]) / np.linalg.norm(disp), atol=1e-12)


def test_speed_preserved_and_tangent(rng):
    X, T = sm.torus(40, 20)
    m = OracleMesh(X, T)
    n = 300
    f, b = sm.random_points(rng, len(T), n)
    A = X[T[f, 0]]
    nrm = np.cross(X[T[f, 1]] - A, X[T[f, 2]] - A)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    v = rng.normal(size=(n, 3))
    v -= (v * nrm).sum(1)[:, None] * nrm
    disp = v * 0.5
    f2, b2, d2, v2, st = m.transport(f, b, disp, v[:, None, :])
    assert np.all(st >= 0)
    np.testing.assert_allclose(np.linalg.norm(v2[:, 0], axis=1), np.linalg.norm(v, axis=1), rtol=1e-12)
    A2 = X[T[f2, 0]]
    n2 = np.cross(X[T[f2, 1]] - A2, X[T[f2, 2]] - A2)
    n2 /= np.linalg.norm(n2, axis=1)[:, None]
    # across the coplanar triangle pairs of this grid the reference rotates by acos(1 - 2^-53) ~ 1.5e-8 rad about a
    # rounding-noise axis (triangulatedMeshSpace.cpp:627-630), which leaves residues of that size
    edge_only = st > 0
    rel = np.abs((v2[:, 0] * n2).sum(1)) / np.linalg.norm(v2[:, 0], axis=1)
    assert np.all(rel[edge_only] < 1e-6)


def test_reverse_walk_returns(rng):
    X, T = sm.torus(30, 15)
    m = OracleMesh(X, T)
    n = 100
    f, b = sm.random_points(rng, len(T), n)
    A = X[T[f, 0]]
    nrm = np.cross(X[T[f, 1]] - A, X[T[f, 2]] - A)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    u = rng.normal(size=(n, 3))
    u -= (u * nrm).sum(1)[:, None] * nrm
    u /= np.linalg.norm(u, axis=1)[:, None]
    L = 0.4
    f2, b2, d2, v2, st = m.transport(f, b, u * L, u[:, None, :])
    f3, b3, d3, v3, st3 = m.transport(f2, b2, -v2[:, 0] * L, None)
    P0, P3 = m.positions(f, b), m.positions(f3, b3)
    np.testing.assert_allclose(P3, P0, atol=1e-9)


@pytest.mark.parametrize("case", sorted(walk_cases.CASES))
def test_vertex_crossing_worked_by_hand(case):
    """Interior vertex crossings worked through the reference's steps by hand (tests/walk_cases.py): the over-travel,
    the stale corner copy, the vertex taken from the first edge hit and the heading and rotation rules."""
    c = walk_cases.CASES[case]()
    m = OracleMesh(c["X"], c["T"])
    f2, b2, d2, v2, st = m.transport([c["face"]], [c["bary"]], [c["disp"]], [[c["vec"]]])
    assert st[0] >= 1
    assert f2[0] == c["end_face"]
    np.testing.assert_allclose(b2[0], c["end_bary"], atol=1e-9)
    np.testing.assert_allclose(m.positions(f2, b2)[0], c["end_point"], atol=1e-9)
    np.testing.assert_allclose(v2[0, 0], c["end_vec"], atol=1e-9)
