"""Pins the oracle's open-mesh walk (openMeshSpace::transportParticleAndVectors with the absorbing and tangential
border handlers, src/models/openMeshSpace.cpp, absorbingOpenMeshSpace.cpp, tangentialOpenMeshSpace.cpp) by closed forms
on a flat square grid [0,6]^2:

* absorbing border edge: the particle stops where the straight path first meets the border; a transported vector that
  points out of the mesh loses its component along the border's outward normal, others are unchanged;
* tangential border edge: from the border point h the particle moves by the projection onto the border line of the
  displacement that remained when the path entered its last face, h + ((1 - t_in) d.e) e with t_in the path parameter
  of that entry (the reference measures the handler's displacement from there, not from h); vectors as above;
* tangential, displacement exactly orthogonal to the border: the walk stops and keeps the barycentrics computed for the
  unclamped target, so the particle ends at start + d, outside its face, as the reference leaves it;
* absorbing border vertex: the particle stops on the vertex;
* on a curved open mesh: transported vectors never grow, and absorbed particles end inside a face.
* border vertex, both modes: a hand-worked exact case, including the tangential continuation with the previous face's
  corner copy (the reference passes it to the vertex handler by value).
The choice among candidate headings at a border vertex follows a rule of this restatement (the reference's pairing of
neighbours and faces depends on CGAL's stored halfedges): parity unpinned for that choice where candidates differ."""
import numpy as np
import pytest

import small_meshes as sm
from oracle.pyoracle import OracleMesh

SIDE = 6.0


def _flat(mode):
    X, T = sm.grid(6, 6)
    m = OracleMesh(X, T)
    m.set_boundary(mode)
    return X, T, m


def _exit(P, d):
    """First parameter t in (0, 1] at which P + t d leaves the square, the side's outward normal and unit direction."""
    best = (np.inf, None, None)
    for axis in (0, 1):
        if d[axis] > 0:
            t = (SIDE - P[axis]) / d[axis]
            nrm = np.eye(3)[axis]
        elif d[axis] < 0:
            t = -P[axis] / d[axis]
            nrm = -np.eye(3)[axis]
        else:
            continue
        if t < best[0]:
            best = (t, nrm, np.eye(3)[1 - axis])
    return best


def _last_entry(P, d, t_exit):
    """Largest path parameter below t_exit at which P + t d crosses a grid line (x = k, y = k or the diagonals
    y - x = k of the unit cells), i.e. where the path entered the face in which it meets the border; 0 if none."""
    ts = [0.0]
    for a, b in ((d[0], P[0]), (d[1], P[1]), (d[1] - d[0], P[1] - P[0])):
        if a == 0:
            continue
        for k in range(-7, 8):
            t = (k - b) / a
            if 1e-12 < t < t_exit - 1e-9:
                ts.append(t)
    return max(ts)


def _project_out(v, nrm):
    return v - nrm * (v @ nrm) if v @ nrm > 0 else v


def _cases(rng, n, lo, hi):
    X, T = sm.grid(6, 6)
    f, b = sm.random_points(rng, len(T), n)
    P = (X[T[f]] * b[:, :, None]).sum(1)
    ang = rng.uniform(0, 2 * np.pi, n)
    L = rng.uniform(lo, hi, n)
    d = np.stack([L * np.cos(ang), L * np.sin(ang), np.zeros(n)], 1)
    vecs = np.concatenate([rng.normal(size=(n, 2, 2)), np.zeros((n, 2, 1))], 2)
    return f, b, P, d, vecs


def test_absorbing_edges_closed_form(rng):
    X, T, m = _flat("absorbing")
    f, b, P, d, vecs = _cases(rng, 400, 0.5, 9.0)
    f2, b2, d2, v2, st = m.transport(f, b, d, vecs)
    assert np.all(st >= 0)
    P2 = m.positions(f2, b2)
    exited = 0
    for i in range(len(f)):
        t, nrm, _ = _exit(P[i], d[i])
        if t >= 1:
            np.testing.assert_allclose(P2[i], P[i] + d[i], atol=1e-12)
            np.testing.assert_allclose(v2[i], vecs[i], atol=1e-14)
            continue
        h = P[i] + t * d[i]
        along = h[1 - int(np.argmax(np.abs(nrm)))]
        if min(abs(along - np.round(along)), along, SIDE - along) < 1e-6:
            continue  # within rounding of a border vertex
        exited += 1
        np.testing.assert_allclose(P2[i], h, atol=1e-9)
        for j in range(2):
            np.testing.assert_allclose(v2[i, j], _project_out(vecs[i, j], nrm), atol=1e-13)
        assert np.all(b2[i] >= 0) and abs(b2[i].sum() - 1) < 1e-12
    assert exited > 150


def test_tangential_edges_closed_form(rng):
    X, T, m = _flat("tangential")
    f, b, P, d, vecs = _cases(rng, 600, 0.5, 5.0)
    f2, b2, d2, v2, st = m.transport(f, b, d, vecs)
    assert np.all(st >= 0)
    P2 = m.positions(f2, b2)
    slid = 0
    for i in range(len(f)):
        t, nrm, e = _exit(P[i], d[i])
        if t >= 1:
            continue
        h = P[i] + t * d[i]
        goal = h + e * ((1 - _last_entry(P[i], d[i], t)) * (d[i] @ e))
        a = int(np.argmax(np.abs(e)))
        if not (0.01 < goal[a] < SIDE - 0.01) or min(abs(h[a] - np.round(h[a])), h[a], SIDE - h[a]) < 1e-6:
            continue  # slides off the side, or meets the border within rounding of a vertex
        slid += 1
        np.testing.assert_allclose(P2[i], goal, atol=1e-9)
        for j in range(2):
            np.testing.assert_allclose(v2[i, j], _project_out(vecs[i, j], nrm), atol=1e-13)
    assert slid > 100


def test_tangential_orthogonal_hit_stops_outside():
    """Displacement exactly along +x onto the side x = 6: both edge overlaps are zero, the walk stops and keeps the
    barycentrics of the unclamped target (tangentialOpenMeshSpace.cpp:29-34 and openMeshSpace.cpp:235-236)."""
    X, T, m = _flat("tangential")
    f = 2 * (2 * 6 + 5)  # lower triangle of the cell [5,6]x[2,3]
    b = np.array([0.5, 0.3, 0.2])
    P = m.positions([f], [b])[0]
    d = np.array([1.5, 0.0, 0.0])
    f2, b2, d2, v2, st = m.transport([f], [b], [d])
    assert st[0] == 1 and f2[0] == f
    np.testing.assert_allclose(m.positions(f2, b2)[0], P + d, atol=1e-12)
    assert b2[0].min() < 0
    np.testing.assert_allclose(d2[0], 0.0, atol=1e-15)


@pytest.mark.parametrize("mode", ["absorbing", "tangential"])
def test_border_vertex_constructed(mode):
    """A walk that meets the border vertex A = (0, 3) exactly, worked by hand through the reference's steps (all
    numbers are dyadic, so the floating-point run is exact). Face 36 has corners A, (1, 3), (1, 4); the particle starts
    at barycentrics (1/4, 1/2, 1/4) = (0.75, 3.25) with displacement (-1.5, -0.5), target barycentrics (7/4, -1/2, -1/4).
    Both edges at corner 0 register, and the reference hands corner 0 (A, a border vertex) to the vertex handler.

    Heading: the edges from A in its other faces (25: (0,2), (1,3), A and 37: A, (1,4), (0,4)); the largest overlap with
    the displacement is towards (0, 2), so the next face is 25 and the heading (0, -1, 0). Vectors: the edge normal
    cross((0,0,1), heading) = (1, 0, 0) points into the mesh (towards (1, 3)), so a vector with x <= 0 loses its x part.

    Absorbing: the particle stops on A, in face 25.
    Tangential: the displacement becomes (0, -0.5, 0) and the target (0, 2.5); the next barycentrics are taken against the
    corners of face 36 (the corner copy the handler received by value), (1, 1/2, -1/2), and placed in face 25, which moves
    the target to (0.5, 2.0). The path then crosses the diagonal into face 24 and ends there, at (0.5, 2.0)."""
    X, T, m = _flat(mode)
    f0 = 36
    b0 = np.array([0.25, 0.5, 0.25])
    np.testing.assert_array_equal(X[T[f0]][:, :2], [[0, 3], [1, 3], [1, 4]])
    d = np.array([-1.5, -0.5, 0.0])
    vecs = np.array([[[-0.7, 0.2, 0.0], [0.4, 0.5, 0.0]]])
    f2, b2, d2, v2, st = m.transport([f0], [b0], [d], vecs)
    assert st[0] >= 1
    if mode == "absorbing":
        assert f2[0] == 25
        np.testing.assert_allclose(m.positions(f2, b2)[0], [0.0, 3.0, 0.0], atol=1e-10)
    else:
        assert f2[0] == 24
        np.testing.assert_allclose(m.positions(f2, b2)[0], [0.5, 2.0, 0.0], atol=1e-10)
    np.testing.assert_allclose(v2[0, 0], [0.0, 0.2, 0.0], atol=1e-15)
    np.testing.assert_allclose(v2[0, 1], [0.4, 0.5, 0.0], atol=1e-15)


@pytest.mark.parametrize("mode", ["absorbing", "tangential"])
def test_curved_open_mesh_invariants(rng, mode):
    X, T = sm.grid(12, 12, 0.25, 0.25, z=lambda x, y: 0.3 * np.sin(1.3 * x) * np.cos(0.9 * y))
    m = OracleMesh(X, T)
    m.set_boundary(mode)
    n = 400
    f, b = sm.random_points(rng, len(T), n)
    A = X[T[f, 0]]
    nrm = np.cross(X[T[f, 1]] - A, X[T[f, 2]] - A)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    v = rng.normal(size=(n, 3))
    v -= (v * nrm).sum(1)[:, None] * nrm
    d = v * rng.uniform(0.2, 3.0, n)[:, None]
    f2, b2, d2, v2, st = m.transport(f, b, d, v[:, None, :])
    assert np.all(st >= 0)
    assert np.all(np.linalg.norm(v2[:, 0], axis=1) <= np.linalg.norm(v, axis=1) * (1 + 1e-12))
    if mode == "absorbing":
        assert np.all(b2 >= -1e-12)


def test_closed_mode_reports_the_border():
    X, T, m = _flat("closed")
    f2, b2, d2, v2, st = m.transport([4], [[0.3, 0.3, 0.4]], [[10.0, 0.0, 0.0]])
    assert st[0] < 0
