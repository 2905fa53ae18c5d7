"""Brute-force exact geodesics on tiny triangle meshes — an independent check of the oracle.

A shortest path on a polyhedral surface is a chain of straight segments in the unfolding of the faces it crosses,
bending only at vertices (Mitchell, Mount & Papadimitriou 1987). This module finds, from a point or a vertex, every
straight path through a strip of distinct faces by depth-first unfolding (each strip keeps the cone of directions that
pass through all of its edges so far), and then joins such straight pieces through *every* vertex with Dijkstra.
Allowing bends at every vertex (not only at saddle or border vertices) only adds valid paths, so the result is the
exact geodesic distance without classifying vertices. Nothing here shares code with the oracle.

Cost grows quickly with the number of faces; use meshes of at most ~100 faces.
"""
import heapq

import numpy as np


def _cross2(a, b):
    return a[0] * b[1] - a[1] * b[0]


class BruteMesh:
    def __init__(self, xyz, tri):
        self.X = np.asarray(xyz, np.float64).reshape(-1, 3)
        self.T = np.asarray(tri, np.int64).reshape(-1, 3)
        edges = {}
        for f, t in enumerate(self.T):
            for k in range(3):
                a, b = int(t[k]), int(t[(k + 1) % 3])
                edges.setdefault((min(a, b), max(a, b)), []).append(f)
        self.nb = [[] for _ in range(len(self.T))]  # (neighbour face, a, b)
        for (a, b), fs in edges.items():
            if len(fs) == 2:
                f, g = fs
                self.nb[f].append((g, a, b))
                self.nb[g].append((f, a, b))
        self.vfaces = [[] for _ in range(len(self.X))]
        for f, t in enumerate(self.T):
            for v in t:
                self.vfaces[int(v)].append(f)

    def point(self, f, bary):
        b = np.asarray(bary, np.float64)
        return (self.X[self.T[f]] * b[:, None]).sum(0) / b.sum()

    def _layout(self, f):
        """2D coordinates of the corners of face f in its own plane."""
        p = self.X[self.T[f]]
        e0 = p[1] - p[0]
        ex = e0 / np.linalg.norm(e0)
        n = np.cross(e0, p[2] - p[0])
        ey = np.cross(n / np.linalg.norm(n), ex)
        return {int(self.T[f][k]): np.array([(p[k] - p[0]) @ ex, (p[k] - p[0]) @ ey]) for k in range(3)}, ex, ey, p[0]

    def _unfold(self, g, a, b, A2, B2, other2):
        """2D position of the corner of g opposite edge (a,b), placed across line A2B2 from point other2."""
        c = [int(v) for v in self.T[g] if v != a and v != b][0]
        A, B, Cc = self.X[a], self.X[b], self.X[c]
        ab = B - A
        lab = np.linalg.norm(ab)
        x = (Cc - A) @ ab / lab
        y = np.linalg.norm(np.cross(Cc - A, ab)) / lab
        u = (B2 - A2) / np.linalg.norm(B2 - A2)
        w = np.array([-u[1], u[0]])
        if (other2 - A2) @ w > 0:
            w = -w
        return c, A2 + x * u + y * w

    def straight_from(self, start_faces, want_points, want_vertices=True):
        """Straight-strip distances from a source.

        start_faces: list of (face, S2 position in that face's layout, layout dict) — the strip starts there with the
        full set of directions into that face (a point source inside the face, or a vertex source at a corner, whose
        directions into the face already are a cone bounded by the two corner edges).
        want_points: dict face -> list of (key, 3D point) targets lying in that face.
        Returns dict key -> distance for points and ('v', vertex) -> distance for vertices.
        """
        best = {}

        def offer(key, d):
            if d < best.get(key, np.inf):
                best[key] = d

        for f0, S2, lay, skip in start_faces:
            stack = [(f0, lay, None, None, frozenset([f0]), skip)]
            while stack:
                f, lay, R, L, visited, entry = stack.pop()
                # targets and corners of this face inside the cone
                for key, P3 in want_points.get(f, []):
                    Q2 = self._to2d(f, lay, P3)
                    if self._in_cone(S2, R, L, Q2):
                        offer(key, float(np.linalg.norm(Q2 - S2)))
                if want_vertices:
                    for v, V2 in lay.items():
                        if self._in_cone(S2, R, L, V2):
                            offer(("v", v), float(np.linalg.norm(V2 - S2)))
                for g, a, b in self.nb[f]:
                    if g in visited:
                        continue
                    if entry is not None and (entry == {a, b} or (isinstance(entry, int) and entry in (a, b))):
                        continue
                    A2, B2 = lay[a], lay[b]
                    clip = self._clip(S2, R, L, A2, B2)
                    if clip is None:
                        continue
                    P, Q = clip
                    r, l = P - S2, Q - S2
                    if _cross2(r, l) < 0:
                        r, l = l, r
                    other = [lay[v] for v in lay if v != a and v != b][0]
                    c, C2 = self._unfold(g, a, b, A2, B2, other)
                    nl = {a: A2, b: B2, c: C2}
                    stack.append((g, nl, r, l, visited | {g}, {a, b}))
        return best

    def _to2d(self, f, lay, P3):
        t = self.T[f]
        # barycentric coordinates of P3 in face f, then the same weights on the 2D corners
        A, B, Cc = self.X[t[0]], self.X[t[1]], self.X[t[2]]
        M = np.stack([B - A, Cc - A], 1)
        uv, *_ = np.linalg.lstsq(M, P3 - A, rcond=None)
        w = np.array([1 - uv[0] - uv[1], uv[0], uv[1]])
        return sum(w[k] * lay[int(t[k])] for k in range(3))

    @staticmethod
    def _in_cone(S2, R, L, Q2, rtol=1e-10):
        if R is None:
            return True
        q = Q2 - S2
        nq = np.linalg.norm(q)
        if nq == 0:
            return True
        nR, nL = np.linalg.norm(R), np.linalg.norm(L)
        if q @ (R / nR + L / nL) <= 0:  # behind the apex: cones here are narrower than a half-turn
            return False
        return (_cross2(R, q) >= -rtol * nR * nq) and (_cross2(q, L) >= -rtol * nL * nq)

    @staticmethod
    def _clip(S2, R, L, A2, B2):
        """Part of segment A2B2 inside the cone (S2; R..L), or None."""
        t0, t1 = 0.0, 1.0
        d = B2 - A2
        if R is not None:
            for (u, sgn) in ((R, 1.0), (L, -1.0)):
                # constraint sgn * cross(u, X - S2) >= 0 for X = A2 + t d
                c0 = sgn * _cross2(u, A2 - S2)
                c1 = sgn * _cross2(u, d)
                tol = 1e-12 * np.linalg.norm(u) * (np.linalg.norm(A2 - S2) + np.linalg.norm(d))
                if abs(c1) < 1e-300:
                    if c0 < -tol:
                        return None
                    continue
                t = -c0 / c1
                if c1 > 0:
                    t0 = max(t0, t)
                else:
                    t1 = min(t1, t)
            # a cone narrower than this only grazes a vertex; paths through vertices are joined by Dijkstra
            if t1 - t0 <= 1e-12:
                return None
        return A2 + t0 * d, A2 + t1 * d

    def point_start(self, f, P3):
        lay, ex, ey, o = self._layout(f)
        S2 = self._to2d(f, lay, P3)
        return [(f, S2, lay, None)]

    def vertex_start(self, v):
        out = []
        for f in self.vfaces[v]:
            lay, ex, ey, o = self._layout(f)
            out.append((f, lay[v], lay, v))  # from a corner, only the opposite edge leads on
        return out


def geodesic_distances(xyz, tri, src_face, src_bary, tgt_face, tgt_bary, patch=None):
    """Exact geodesic distances from one source point to targets; patch (face list) restricts the surface."""
    xyz = np.asarray(xyz, np.float64).reshape(-1, 3)
    tri = np.asarray(tri, np.int64).reshape(-1, 3)
    if patch is not None:
        keep = sorted(set(int(f) for f in patch))
        remap = {f: i for i, f in enumerate(keep)}
        tri = tri[keep]
        src_face = remap[int(src_face)]
        tgt_face = [remap.get(int(f), -1) for f in tgt_face]
    M = BruteMesh(xyz, tri)
    S3 = M.point(src_face, src_bary)
    targets = {}
    pts = []
    for i, (f, b) in enumerate(zip(tgt_face, np.asarray(tgt_bary, np.float64).reshape(-1, 3))):
        if f < 0:
            pts.append(None)
            continue
        P3 = M.point(f, b)
        pts.append(P3)
        targets.setdefault(f, []).append((("t", i), P3))
    used = sorted(set(int(v) for v in tri.ravel()))
    # straight pieces from the source and from every vertex
    from_src = M.straight_from(M.point_start(src_face, S3), targets)
    from_v = {v: M.straight_from(M.vertex_start(v), targets) for v in used}
    dist = {v: from_src.get(("v", v), np.inf) for v in used}
    heap = [(d, v) for v, d in dist.items() if d < np.inf]
    heapq.heapify(heap)
    done = set()
    while heap:
        d, v = heapq.heappop(heap)
        if v in done or d > dist[v]:
            continue
        done.add(v)
        for key, w in from_v[v].items():
            if key[0] == "v":
                u = key[1]
                if d + w < dist[u]:
                    dist[u] = d + w
                    heapq.heappush(heap, (d + w, u))
    out = np.full(len(pts), np.inf)
    for i in range(len(pts)):
        if pts[i] is None:
            continue
        best = from_src.get(("t", i), np.inf)
        for v in used:
            if dist[v] < np.inf:
                best = min(best, dist[v] + from_v[v].get(("t", i), np.inf))
        out[i] = best
    return out
