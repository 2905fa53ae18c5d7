"""Interior vertex crossings worked by hand through the reference's steps (triangulatedMeshSpace::
transportParticleAndVectors, src/models/triangulatedMeshSpace.cpp:448-564, updateForVertexIntersection :566-611,
throughVertex :247-407, findTriangleEdgeIntersectionInformation src/utility/meshUtilities.cpp:283-331). Each case
returns the mesh, the walk's input and the expected end state, derived below without running any walk code.

Flat grid (grid(6, 6), unit squares; square (i, j) holds face 2(6j + i) with corners (i,j), (i+1,j), (i+1,j+1) and
face 2(6j + i) + 1 with corners (i,j), (i+1,j+1), (i,j+1)). Start at (1.75, 1.25) in face 14 (barycentrics
(0.25, 0.5, 0.25)) with displacement (0.375, 1.125), which passes exactly through the grid vertex (2, 2), corner 3 of
face 14, at two thirds of its length:
  1. target (2.125, 2.375) has barycentrics (-0.125, -0.25, 1.375) in face 14; the V2V3 test (edge (2,1)-(2,2)) and
     the V3V1 test (edge (2,2)-(1,1)) both hit, at the end and the start of their edges: a vertex crossing;
  2. the vertex handed on is involvedVertex[0], the first vertex of the first edge hit: (2, 1), not (2, 2);
  3. throughVertex around (2, 1) on a flat mesh keeps the heading (total angle 2 pi, half of it travelled from the
     reversed incoming direction), so the heading is (1, 3)/sqrt(10); the angle from the reversed incoming direction
     to the first ring edge (towards (2, 2)) is 161.6 degrees, one more ring step (45 degrees, towards (3, 2)) passes
     180, so the new face is the one with corners {(2,1), (2,2), (3,2)}: face 17, corners (2,1), (3,2), (2,2);
  4. the remaining length is the whole pre-crossing |displacement| (sqrt(1.40625)), so the target becomes
     (2, 2) + (0.375, 1.125) = (2.375, 3.125);
  5. the caller's corner copy is still face 14's ((1,1), (2,1), (2,2): passed by value), so the next target
     barycentrics are taken against it: (-0.375, -0.75, 2.125); read in face 17 they give the point
     -0.375 (2,1) - 0.75 (3,2) + 2.125 (2,2) = (1.25, 2.375);
  6. the walk continues straight from (2, 2) to (1.25, 2.375) (direction (-2, 1)), crossing edges near (2, 2) and then
     the diagonal of square (1, 2) at (4/3, 7/3), and stops in face 27 (corners (1,2), (2,3), (1,3)) with barycentrics
     (0.625, 0.25, 0.125). The faces are coplanar, so the velocity is not rotated.
Without the quirks the particle would end at (2.125, 2.375).

Cone apex: a fan of six faces [C, R_i, R_i+1] around C = (0, 0, 1.5) with R_i = (cos(i pi/3), sin(i pi/3), 0).
Start at C + (m0 - C)/2, m0 the midpoint of R0 R1 (barycentrics (0.5, 0.25, 0.25) in face 0), displacement
-0.75 (m0 - C): the walk reaches the apex, corner 1 of face 0, at two thirds of it.
  1. the V1V2 and V3V1 tests hit; the first edge hit is (C, R0), so involvedVertex[0] is the apex itself;
  2. the ring around C starts at R0 and runs R0, R5, R4, R3, ... (clockwise from outside); each face angle is alpha,
     the reversed incoming direction lies alpha/2 from the first ring edge, so the walk passes half the total angle
     (3 alpha) after three ring steps, between R4 and R3; the heading is rotated back by alpha/2 from the edge towards
     R3: it is the median of face 3 = [C, R3, R4], (m3 - C)/l with l = |m0 - C|;
  3. the target becomes C + 0.75 (m3 - C) (the remaining length is the full 0.75 l);
  4. the barycentrics against face 0's stale corners: face 3 is face 0 turned by pi about the z axis, so the target is
     that rotation of q = C + 0.75 (m0' - C) where m0' lies on the vertical plane through m0; projecting onto face 0's
     plane keeps the median component, 0.75 (h^2 - a^2)/(h^2 + a^2) (m0 - C) with a^2 = 3/4 the squared apothem and
     h = 1.5: beta = 0.375, so the barycentrics are (1 - beta, beta/2, beta/2) = (0.625, 0.1875, 0.1875);
  5. read in face 3 they give C + beta (m3 - C), inside face 3: the walk stops there, face 3, (0.625, 0.1875, 0.1875);
  6. the face normals satisfy n0.n3 = (3/4 - h^2)/(3/4 + h^2) = -1/2 < 0, so transported vectors turn by acos(-1/2) =
     120 degrees about n0 x n3 (a horizontal axis perpendicular to m0); that rotation takes the incoming direction
     (C - m0)/l onto the outgoing median (m3 - C)/l, which is where a vector along the path ends.
"""
import numpy as np

import small_meshes as sm


def flat_grid_case():
    X, T = sm.grid(6, 6)
    return dict(X=X, T=T, face=14, bary=[0.25, 0.5, 0.25], disp=[0.375, 1.125, 0.0], vec=[0.6, 0.8, 0.0],
                end_face=27, end_bary=[0.625, 0.25, 0.125], end_point=[1.25, 2.375, 0.0], end_vec=[0.6, 0.8, 0.0])


def cone_case(h=1.5):
    k = 6
    th = np.arange(k) * 2 * np.pi / k
    X = np.vstack([[0.0, 0.0, h], np.stack([np.cos(th), np.sin(th), np.zeros(k)], 1)])
    T = np.array([[0, 1 + i, 1 + (i + 1) % k] for i in range(k)])
    C, m0, m3 = X[0], 0.5 * (X[1] + X[2]), 0.5 * (X[4] + X[5])
    ell = np.linalg.norm(m0 - C)
    return dict(X=X, T=T, face=0, bary=[0.5, 0.25, 0.25], disp=list(-0.75 * (m0 - C)), vec=list((C - m0) / ell),
                end_face=3, end_bary=[0.625, 0.1875, 0.1875], end_point=list(C + 0.375 * (m3 - C)),
                end_vec=list((m3 - C) / ell))


CASES = {"flat_grid": flat_grid_case, "cone_apex": cone_case}
