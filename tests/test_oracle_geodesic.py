"""Pins the oracle's exact geodesics (the checker of the HIP solver) without trusting its own formulas:
brute-force unfolding + Dijkstra on tiny meshes, closed forms, and invariants of the metric."""
import numpy as np
import pytest

import bruteforce_geodesic as bf
import small_meshes as sm
from oracle.pyoracle import OracleMesh

CASES = {
    "tetrahedron": sm.tetrahedron,
    "cube": sm.cube,
    "octahedron": sm.octahedron,
    "icosahedron": sm.icosahedron,
    "planar_grid": lambda: sm.grid(4, 3),
    "l_shape": lambda: sm.l_shape(2),
    "torus_with_saddles": lambda: sm.torus(6, 4),
    "bumpy_open_grid": lambda: sm.grid(5, 4, z=lambda x, y: 0.3 * np.sin(2 * x) * np.cos(1.5 * y)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_matches_bruteforce(name, rng):
    X, T = CASES[name]()
    m = OracleMesh(X, T)
    for _ in range(2):
        sf = int(rng.integers(0, len(T)))
        sb = rng.dirichlet([1.0, 1.0, 1.0])
        tf, tb = sm.random_points(rng, len(T), 8)
        d, st, et, fl = m.geodesic(sf, sb, tf, tb)
        db = bf.geodesic_distances(X, T, sf, sb, tf, tb)
        np.testing.assert_allclose(d, db, rtol=1e-12, atol=1e-13)


def test_patch_matches_bruteforce_on_patch(rng):
    """Distances restricted to a face subset equal brute force on that subset as an open surface."""
    X, T = sm.torus(6, 4)
    m = OracleMesh(X, T)
    sf = 22
    sb = np.array([0.2, 0.5, 0.3])
    patch = m.patch(sf, sb, [sf], 10.0)  # only the source face: every target there
    assert list(patch) == [sf]
    # a strip of faces around the tube plus the source face neighbourhood
    keep = np.array(sorted(set(range(16, 34))), np.int32)
    tf = rng.choice(keep, 6).astype(np.int32)
    tb = rng.dirichlet([1.0, 1.0, 1.0], 6)
    d, _, _, fl = m.geodesic(sf, sb, tf, tb, patch=keep)
    db = bf.geodesic_distances(X, T, sf, sb, tf, tb, patch=keep)
    np.testing.assert_allclose(d, db, rtol=1e-12, atol=1e-13)


def test_planar_is_euclidean(rng):
    X, T = sm.grid(6, 5, 0.7, 0.9)
    m = OracleMesh(X, T)
    sf, sb = 13, np.array([0.1, 0.6, 0.3])
    tf, tb = sm.random_points(rng, len(T), 40)
    d, st, et, fl = m.geodesic(sf, sb, tf, tb)
    S = m.positions([sf], [sb])[0]
    P = m.positions(tf, tb)
    e = np.linalg.norm(P - S, axis=1)
    np.testing.assert_allclose(d, e, rtol=1e-13, atol=1e-14)
    # tangents: chord direction at both ends
    u = (P - S) / e[:, None]
    np.testing.assert_allclose(st, u, atol=1e-12)
    np.testing.assert_allclose(et, u, atol=1e-12)


def test_l_shape_bends_at_reflex_corner():
    X, T = sm.l_shape(2)  # reflex corner at (2,2)
    m = OracleMesh(X, T)
    # source near (3.5, 0.5), target near (0.5, 3.5): straight segment leaves the region
    def locate(p):
        for f, t in enumerate(T):
            A, B, Cc = X[t]
            M = np.array([[B[0] - A[0], Cc[0] - A[0]], [B[1] - A[1], Cc[1] - A[1]]])
            uv = np.linalg.solve(M, np.array(p) - A[:2])
            w = np.array([1 - uv.sum(), uv[0], uv[1]])
            if (w >= -1e-12).all():
                return f, w
        raise ValueError(p)

    sf, sb = locate((3.6, 0.4))
    tf, tb = locate((0.3, 3.7))
    d, st, et, fl = m.geodesic(sf, sb, [tf], [tb])
    corner = np.array([2.0, 2.0])
    expect = np.linalg.norm(np.array([3.6, 0.4]) - corner) + np.linalg.norm(np.array([0.3, 3.7]) - corner)
    assert abs(d[0] - expect) < 1e-12
    # the start tangent points at the corner, the end tangent away from it
    s_dir = (corner - [3.6, 0.4]) / np.linalg.norm(corner - [3.6, 0.4])
    e_dir = ([0.3, 3.7] - corner) / np.linalg.norm([0.3, 3.7] - corner)
    np.testing.assert_allclose(st[0, :2], s_dir, atol=1e-12)
    np.testing.assert_allclose(et[0, :2], e_dir, atol=1e-12)


def test_cube_adjacent_faces_closed_form():
    """Two points near a shared cube edge: the geodesic is the straight line of the two-face unfolding."""
    X, T = sm.cube(1.0)
    m = OracleMesh(X, T)
    # faces of the z=0 side and x=0 side share the edge x=0,z=0
    def locate(p):
        for f, t in enumerate(T):
            A, B, Cc = X[t]
            n = np.cross(B - A, Cc - A)
            if abs((np.array(p) - A) @ n) > 1e-12:
                continue
            Mx = np.stack([B - A, Cc - A], 1)
            uv, *_ = np.linalg.lstsq(Mx, np.array(p) - A, rcond=None)
            w = np.array([1 - uv.sum(), uv[0], uv[1]])
            if (w >= -1e-12).all():
                return f, w
        raise ValueError(p)

    sf, sb = locate((0.2, 0.45, 0.0))   # on z = 0, 0.2 from the edge x = 0
    tf, tb = locate((0.0, 0.55, 0.3))   # on x = 0, 0.3 from the same edge
    d, st, et, fl = m.geodesic(sf, sb, [tf], [tb])
    expect = np.hypot(0.2 + 0.3, 0.55 - 0.45)
    assert abs(d[0] - expect) < 1e-13


def test_metric_invariants(rng):
    X, T = sm.torus(10, 6)
    m = OracleMesh(X, T)
    f, b = sm.random_points(rng, len(T), 6)
    D = np.zeros((6, 6))
    for i in range(6):
        D[i] = m.geodesic(f[i], b[i], f, b)[0]
    np.testing.assert_allclose(D, D.T, rtol=1e-12, atol=1e-13)   # symmetry
    assert np.all(np.diag(D) < 1e-12)
    for i in range(6):                                          # triangle inequality
        assert np.all(D[i][:, None] <= D[i][None, :] + D + 1e-12)
    P = m.positions(f, b)
    E = np.linalg.norm(P[:, None] - P[None], axis=2)
    assert np.all(D >= E - 1e-12)                                # never shorter than the chord


def test_start_tangent_is_minus_gradient(rng):
    """Moving the source by h along the start tangent shortens the distance by h (first order)."""
    X, T = sm.torus(12, 8)
    m = OracleMesh(X, T)
    sf, sb = 40, np.array([0.3, 0.35, 0.35])
    tf, tb = sm.random_points(rng, len(T), 10)
    d, st, et, fl = m.geodesic(sf, sb, tf, tb)
    A, B, Cc = X[T[sf]]
    Minv = np.linalg.pinv(np.stack([B - A, Cc - A], 1))
    h = 1e-6
    for i in np.where(fl == 0)[0]:
        S = m.positions([sf], [sb])[0]
        uv = Minv @ (S + h * st[i] - A)
        nb = np.array([1 - uv.sum(), uv[0], uv[1]])
        if (nb < 0).any():
            continue
        d2 = m.geodesic(sf, nb, tf[i:i + 1], tb[i:i + 1])[0][0]
        assert abs((d[i] - d2) / h - 1.0) < 1e-4
        assert abs(np.linalg.norm(st[i]) - 1) < 1e-12
        assert abs(np.linalg.norm(et[i]) - 1) < 1e-12


def test_patch_rule_early_exits_and_targets():
    X, T = sm.grid(8, 8)
    m = OracleMesh(X, T)
    adj, _ = m.adjacency()
    sf, sb = 70, np.array([1 / 3, 1 / 3, 1 / 3])
    # every target in the source face -> source face only
    assert list(m.patch(sf, sb, [sf, sf], 5.0)) == [sf]
    # targets only in edge neighbours -> source face + its neighbours, whatever the radius
    nbs = [int(g) for g in adj[sf] if g >= 0]
    assert sorted(m.patch(sf, sb, nbs[:1], 1e-9)) == sorted([sf] + nbs)
    # far target with a tiny radius: neighbours + the target face itself (not connected)
    p = m.patch(sf, sb, [0], 1e-9)
    assert sorted(p) == sorted([sf] + nbs + [0])
    # big radius: everything reachable
    assert len(m.patch(sf, sb, [0], 100.0)) == len(T)
    # radius rule: a face is added iff one of its corners is within r of the source point (or it is a target
    # face, the source face or one of its neighbours) and it is connected to them through such faces
    S = m.positions([sf], [sb])[0]
    r = 2.3
    p = set(m.patch(sf, sb, [0], r).tolist())
    near = {f for f in range(len(T)) if (np.linalg.norm(X[T[f]] - S, axis=1) <= r).any()}
    assert p - {0} <= near | {sf} | set(nbs)
    assert near <= p  # the grid is connected through near faces here


def test_unreachable_sentinel():
    X, T = sm.grid(8, 8)
    m = OracleMesh(X, T)
    sf, sb = 70, np.array([1 / 3, 1 / 3, 1 / 3])
    d, st, et, fl = m.distance(sf, sb, [0], [[0.2, 0.3, 0.5]], submeshing=True, maximum_distance=0.5, threshold=0.1)
    assert d[0] == 1.0 and np.all(st[0] == [0, 0, 1]) and np.all(et[0] == [0, 0, 1])


def test_full_mesh_equals_patch_for_short_pairs(rng):
    """Reference harness branch 0 (meshTesting.cpp:188-203): submesh and full-mesh distances agree for d < r."""
    X, T = sm.torus(24, 12)
    m = OracleMesh(X, T)
    r = 0.9
    for _ in range(3):
        sf = int(rng.integers(0, len(T)))
        sb = rng.dirichlet([1, 1, 1])
        tf, tb = sm.random_points(rng, len(T), 200)
        P = m.positions(tf, tb)
        S = m.positions([sf], [sb])[0]
        near = np.linalg.norm(P - S, axis=1) < r
        tf, tb = tf[near], tb[near]
        dfull = m.distance(sf, sb, tf, tb, submeshing=False, maximum_distance=r)[0]
        dsub = m.distance(sf, sb, tf, tb, submeshing=True, maximum_distance=r, threshold=r)[0]
        ok = dfull < r - 0.35  # well inside the ball: patch faces cover every short path
        np.testing.assert_allclose(dsub[ok], dfull[ok], rtol=1e-12)
        assert np.all(dsub >= dfull - 1e-12)


def test_truncation_keeps_short_pairs(rng):
    X, T = sm.torus(24, 12)
    m = OracleMesh(X, T)
    sf = 100
    sb = np.array([0.2, 0.2, 0.6])
    tf, tb = sm.random_points(rng, len(T), 100)
    d, *_ = m.geodesic(sf, sb, tf, tb)
    rmax = np.median(d)
    d2, _, _, fl2 = m.geodesic(sf, sb, tf, tb, rmax=rmax)
    inside = d <= rmax
    np.testing.assert_allclose(d2[inside], d[inside], rtol=1e-13)
    assert np.all((fl2[~inside] & 5) != 0)
