"""Tiny hand-built meshes for the exactness tests (consistently oriented, counter-clockwise seen from outside)."""
import numpy as np


def tetrahedron():
    X = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    T = np.array([[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]])
    return X, T


def cube(side=1.0):
    X = np.array([[x, y, z] for x in (0, side) for y in (0, side) for z in (0, side)], float)
    # vertex id = 4x + 2y + z (each 0/1)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    T = []
    for a, b, c, d in quads:
        T += [[a, b, c], [a, c, d]]
    T = np.array(T)
    return X, _orient_outward(X, T)


def octahedron():
    X = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)
    T = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]])
    return X, _orient_outward(X, T)


def icosahedron():
    t = (1 + 5 ** 0.5) / 2
    X = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    T = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    return X, _orient_outward(X, T)


def grid(nx, ny, hx=1.0, hy=1.0, z=None):
    """Planar (or height-field when z is a function) open grid of 2*nx*ny triangles."""
    X = []
    for j in range(ny + 1):
        for i in range(nx + 1):
            x, y = i * hx, j * hy
            X.append([x, y, 0.0 if z is None else z(x, y)])
    X = np.array(X)
    T = []
    for j in range(ny):
        for i in range(nx):
            a = j * (nx + 1) + i
            b, c, d = a + 1, a + nx + 2, a + nx + 1
            T += [[a, b, c], [a, c, d]]
    return X, np.array(T)


def l_shape(n=2):
    """Planar L-shaped region made of unit squares: [0,2n]x[0,n] plus [0,n]x[n,2n]; reflex corner at (n,n)."""
    X, T = grid(2 * n, 2 * n)
    keep = []
    for f, t in enumerate(T):
        c = X[t].mean(0)
        if not (c[0] > n and c[1] > n):
            keep.append(f)
    T = T[keep]
    used = np.unique(T)
    remap = -np.ones(len(X), int)
    remap[used] = np.arange(len(used))
    return X[used], remap[T]


def torus(nu, nv, R=2.0, r=0.8):
    X = []
    for i in range(nu):
        for j in range(nv):
            u, v = 2 * np.pi * i / nu, 2 * np.pi * j / nv
            X.append([(R + r * np.cos(v)) * np.cos(u), (R + r * np.cos(v)) * np.sin(u), r * np.sin(v)])
    X = np.array(X)
    T = []
    for i in range(nu):
        for j in range(nv):
            a = i * nv + j
            b = ((i + 1) % nu) * nv + j
            c = ((i + 1) % nu) * nv + (j + 1) % nv
            d = i * nv + (j + 1) % nv
            T += [[a, b, c], [a, c, d]]
    return X, np.array(T)


def _orient_outward(X, T):
    T = np.array(T).copy()
    c = X.mean(0)
    for f in range(len(T)):
        p = X[T[f]]
        n = np.cross(p[1] - p[0], p[2] - p[0])
        if n @ (p.mean(0) - c) < 0:
            T[f] = T[f][[0, 2, 1]]
    return T


def random_points(rng, nF, n):
    f = rng.integers(0, nF, n).astype(np.int32)
    b = rng.dirichlet([1.0, 1.0, 1.0], n)
    return f, b


def cylinder(n_around, n_along, radius, h):
    """Open prism-cylinder whose faces and corner orders are those of grid(n_around, n_along, chord, h) with the last
    column of vertices identified with the first: the flat grid is its exact unrolling (chord = 2 R sin(pi/n)), so a
    (face, barycentrics) pair names corresponding points on both. Outward normals."""
    X = []
    for j in range(n_along + 1):
        for i in range(n_around):
            t = 2 * np.pi * i / n_around
            X.append([radius * np.cos(t), radius * np.sin(t), j * h])
    X = np.array(X)
    T = []
    for j in range(n_along):
        for i in range(n_around):
            a = j * n_around + i
            b = j * n_around + (i + 1) % n_around
            c = (j + 1) * n_around + (i + 1) % n_around
            d = (j + 1) * n_around + i
            T += [[a, b, c], [a, c, d]]
    return X, np.array(T)


def cylinder_chord(n_around, radius):
    return 2 * radius * np.sin(np.pi / n_around)
