// ORACLE — test infrastructure only (see oracle_internal.hpp).
//
// FP64 restatement of the reference's straight walk across faces with parallel transport,
// triangulatedMeshSpace::transportParticleAndVectors (src/models/triangulatedMeshSpace.cpp:448-564) and its helpers:
//   clamps                        src/utility/meshUtilities.cpp:138-182
//   edge-crossing predicates      src/utility/meshUtilities.cpp:229-331
//   edge crossing                 src/models/triangulatedMeshSpace.cpp:614-658
//   vertex crossing               src/models/triangulatedMeshSpace.cpp:566-611, throughVertex :247-407
//   Rodrigues rotation, angle     src/utility/functionUtilities.cpp:8-45
// Behaviour kept on purpose, because the reference has it:
//   * after a vertex crossing the caller's copy of the face corner positions is not refreshed (the helper receives it
//     by value, :568), so the next target barycentrics are taken against the previous face's corners;
//   * the vertex crossing keeps the pre-crossing remaining length (:578-581) and rotates transported vectors only when
//     the two face normals point into opposite half-spaces (:589);
//   * the vertex handed to the vertex crossing is the first vertex of the first edge found (:526-528).
// One deliberate difference: a rotation about a zero-length axis (coplanar neighbouring faces) is skipped, where the
// reference produces NaN coordinates and throws (:632-633 and src/utility/meshUtilities.cpp:184-195).
//
// Open meshes (Mesh::boundary != 0) follow openMeshSpace::transportParticleAndVectors (src/models/openMeshSpace.cpp:
// 114-238), which calls a border handler where the closed walk would fail:
//   border edge, absorbing     absorbingOpenMeshSpace::updateAtBoundaryEdge (src/models/absorbingOpenMeshSpace.cpp:2-25)
//   border edge, tangential    tangentialOpenMeshSpace::updateAtBoundaryEdge (src/models/tangentialOpenMeshSpace.cpp:3-39)
//   border vertex, both        getBoundaryVertexHeading and projectVectorsForBoundaryVertex
//                              (src/models/openMeshSpace.cpp:3-96), then the absorbing (:27-50) or tangential (:41-64)
//                              vertex handler
//   vector projection          projectVectorsIfOverBoundary (src/models/triangulatedMeshSpace.cpp:411-426),
//                              projectVectorOntoDirection (src/utility/meshUtilities.cpp:119-123)
// Kept as the reference has them: the displacement handed to a handler runs from the pre-crossing source (so the
// remaining length includes the distance already walked to the border); the tangential edge handler stops only when
// the displacement is orthogonal to the edge and then leaves the barycentrics computed at the top of that iteration;
// the vertex handlers receive the corner copy and the face normal by value, so after a border-vertex continuation both
// still describe the previous face.
// Deliberate difference at a border vertex: the candidate headings are the edges from the vertex to the other corners
// of every incident face except the current one (faces in increasing index, corners in face order), and the first
// edge of largest overlap with the displacement wins. The reference pairs neighbours and faces through CGAL's stored
// vertex halfedges (src/models/openMeshSpace.cpp:24-32), an order that depends on how the mesh was built.
#include <algorithm>
#include <cmath>
#include <set>
#include <vector>

#include "oracle.h"
#include "oracle_internal.hpp"

namespace orc {

static V3 corner(const Mesh& m, int f, int k) { return m.P(f, k); }

static V3 facePoint(const Mesh& m, int f, const double* b) {
    double s = b[0] + b[1] + b[2];
    return (corner(m, f, 0) * b[0] + corner(m, f, 1) * b[1] + corner(m, f, 2) * b[2]) / s;
}

static V3 pointIn(const std::array<V3, 3>& c, const double* b) {
    double s = b[0] + b[1] + b[2];
    return (c[0] * b[0] + c[1] * b[1] + c[2] * b[2]) / s;
}

// barycentric coordinates of q with respect to triangle (a,b,c): projection form (Ericson, Real-Time Collision
// Detection, section 3.4)
static void baryOf(const V3& a, const V3& b, const V3& c, const V3& q, double* out) {
    V3 e0 = b - a, e1 = c - a, e2 = q - a;
    double d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1), d20 = dot(e2, e0), d21 = dot(e2, e1);
    double den = d00 * d11 - d01 * d01;
    double v = (d11 * d20 - d01 * d21) / den;
    double w = (d00 * d21 - d01 * d20) / den;
    out[0] = 1.0 - v - w;
    out[1] = v;
    out[2] = w;
}

static void clampBelowZero(double* b, double tol = 1e-11) {
    double s = 0;
    for (int i = 0; i < 3; ++i) {
        b[i] = std::max(b[i], tol);
        s += b[i];
    }
    for (int i = 0; i < 3; ++i) b[i] /= s;
}

static void clampNearZero(double* b, double tol = 1e-13) {
    double s = 0;
    for (int i = 0; i < 3; ++i) {
        if (b[i] > -tol && b[i] < tol) b[i] = tol;
        s += b[i];
    }
    for (int i = 0; i < 3; ++i) b[i] /= s;
}

static V3 rotateAbout(const V3& p, const V3& base, const V3& axisDir, double angle) {
    V3 a = axisDir / norm(axisDir);
    V3 q = p - base;
    double ad = dot(a, q);
    double s = std::sin(angle), c = std::cos(angle);
    V3 r = a * (ad * (1.0 - c)) + q * c + cross(a, q) * s;
    return r + base;
}

static double angleOf(V3 u, V3 v) {
    u = normalized(u);
    v = normalized(v);
    double d = dot(u, v);
    if (d >= 1) return 0;
    if (d <= -1) return M_PI;
    return std::acos(d);
}

// crossing of the barycentric segment s->t with the face edge on which coordinate c vanishes; the edge runs from
// corner (c+1)%3 to corner (c+2)%3 and mu is the position along it
static bool crossEdge(const double* s, const double* t, int c, double* hit) {
    double den = s[c] - t[c];
    if (den == 0) return false;
    double lam = s[c] / den;
    int b = (c + 2) % 3;
    double mu = s[b] + lam * (t[b] - s[b]);
    for (int i = 0; i < 3; ++i) hit[i] = s[i] + lam * (t[i] - s[i]);
    return lam >= 0 && lam <= 1 && mu >= 0 && mu <= 1;
}

// triangulatedMeshSpace::projectVectorsIfOverBoundary (src/models/triangulatedMeshSpace.cpp:411-426): a vector that
// points over the border line (orthogonal gives the line's normal direction, inward points into the mesh) loses its
// component along that normal
static void projectIfOverBoundary(std::vector<V3>& vecs, const V3& orthogonal, const V3& inward) {
    bool pointsOut = dot(orthogonal, inward) < 0;
    for (auto& v : vecs) {
        bool alongPerp = dot(v, orthogonal) > 0;
        if ((pointsOut && alongPerp) || (!pointsOut && !alongPerp)) {
            V3 dh = normalized(orthogonal);
            v = v - dh * dot(v, dh);
        }
    }
}

struct Walker {
    const Mesh& m;
    explicit Walker(const Mesh& mm) : m(mm) {}

    int across(int f, int a, int b) const {  // face on the other side of edge {a,b} of face f
        for (int k = 0; k < 3; ++k) {
            int u = m.fv(f, k), v = m.fv(f, k + 1);
            if ((u == a && v == b) || (u == b && v == a)) return m.adj[3 * f + k];
        }
        return -1;
    }

    // straightest continuation through vertex vI (reference throughVertex)
    int throughVertex(int vI, const V3& toIntersection, int sourceFace, V3& heading) const {
        V3 vi = m.X[vI];
        int start = -1;
        for (int k = 0; k < 3; ++k)
            if (m.fv(sourceFace, k) == vI) start = m.fv(sourceFace, k + 1);
        // one-ring of vI, clockwise seen from outside (the order of a CGAL Vertex_around_target_circulator),
        // beginning at `start`
        std::vector<int> ring;
        int f = sourceFace, cur = start;
        for (int guard = 0; guard < 1000; ++guard) {
            ring.push_back(cur);
            // in face f, the corner after vI is cur; step to the face across edge (vI, cur)
            int g = across(f, vI, cur);
            if (g < 0) throw OracleError(-5, "border reached while circling a vertex");
            int nxt = -1;
            for (int k = 0; k < 3; ++k)
                if (m.fv(g, k) == vI) nxt = m.fv(g, k + 1);
            f = g;
            if (nxt == start) break;
            cur = nxt;
        }
        int n = (int)ring.size();
        std::vector<V3> e(n);
        for (int i = 0; i < n; ++i) e[i] = normalized(m.X[ring[i]] - vi);
        double total = 0;
        for (int i = 0; i < n; ++i) total += angleOf(e[i], e[(i + 1) % n]);
        double half = total / 2.0;
        V3 back = normalized(-toIntersection);
        double travelled = angleOf(back, e[0]);
        int counter = 0;
        while (travelled < half) {
            travelled += angleOf(e[counter], e[(counter + 1) % n]);
            counter++;
            if (counter >= n) throw OracleError(-6, "no angular halfway found at a vertex");
        }
        if (counter == 0) throw OracleError(-6, "no angular halfway found at a vertex");
        double diff = half - travelled;
        V3 axis = normalized(cross(e[counter - 1], e[counter]));
        V3 pr = rotateAbout(vi + e[counter], vi, axis, diff);
        heading = normalized(pr - vi);
        std::set<int> want = {vI, ring[counter - 1], ring[counter]};
        int target = sourceFace;
        for (int q = m.vfOff[vI]; q < m.vfOff[vI + 1]; ++q) {
            int g = m.vfList[q] >> 2;
            if (g == sourceFace) continue;
            std::set<int> have = {m.fv(g, 0), m.fv(g, 1), m.fv(g, 2)};
            if (have == want) {
                target = g;
                break;
            }
        }
        if (target == sourceFace) throw OracleError(-7, "throughVertex found no face");
        return target;
    }

    // openMeshSpace::getBoundaryVertexHeading (src/models/openMeshSpace.cpp:3-70) with the candidate rule stated at
    // the top of this file: returns the next face, writes the unit heading and the neighbour vertex it points to
    int borderVertexHeading(int v, int sourceFace, const V3& disp, V3& heading, int& towards) const {
        V3 dn = normalized(disp);
        double best = -1;
        int next = sourceFace;
        towards = v;
        heading = V3(0, 0, 0);
        for (int q = m.vfOff[v]; q < m.vfOff[v + 1]; ++q) {
            int g = m.vfList[q] >> 2, k = m.vfList[q] & 3;
            if (g == sourceFace) continue;
            for (int j = 1; j <= 2; ++j) {
                int w = m.fv(g, k + j);
                V3 out = normalized(m.X[w] - m.X[v]);
                double ov = dot(out, dn);
                if (ov > best) {
                    heading = out;
                    best = ov;
                    next = g;
                    towards = w;
                }
            }
        }
        if (next == sourceFace) throw OracleError(-7, "no face found for a border vertex crossing");
        return next;
    }

    // returns number of face-boundary crossings, or throws
    int run(int& face, double* bary, V3& disp, std::vector<V3>& vecs, int maxSteps) {
        int curFace = face;
        double srcB[3] = {bary[0], bary[1], bary[2]};
        V3 src = facePoint(m, curFace, srcB);
        std::array<V3, 3> corners = {corner(m, curFace, 0), corner(m, curFace, 1), corner(m, curFace, 2)};
        V3 n = m.normal(curFace);
        double nd = dot(n, disp);
        if (std::fabs(nd) > 1e-14) disp = disp - n * nd;
        V3 target = src + disp;
        int lastA = -1, lastB = -1;  // last edge crossed, as an unordered vertex pair
        double tgtB[3];
        int crossings = 0;
        for (int iter = 0;; ++iter) {
            if (iter > maxSteps) throw OracleError(-8, "too many face crossings");
            baryOf(corners[0], corners[1], corners[2], target, tgtB);
            clampNearZero(tgtB);
            target = facePoint(m, curFace, tgtB);
            clampBelowZero(srcB);
            src = facePoint(m, curFace, srcB);
            disp = target - src;
            if (tgtB[0] != tgtB[0]) throw OracleError(-4, "NaN barycentric coordinate");
            if (tgtB[0] >= 0 && tgtB[1] >= 0 && tgtB[2] >= 0) break;
            int vl[3] = {m.fv(curFace, 0), m.fv(curFace, 1), m.fv(curFace, 2)};
            double hit[3] = {0, 0, 0}, h[3];
            std::vector<int> uninvolved, involved;
            const int order[3] = {2, 0, 1};  // V1V2, V2V3, V3V1
            for (int oi = 0; oi < 3; ++oi) {
                int c = order[oi];
                int a = vl[(c + 1) % 3], b = vl[(c + 2) % 3];
                bool same = (a == lastA && b == lastB) || (a == lastB && b == lastA);
                if (same) continue;
                if (crossEdge(srcB, tgtB, c, h)) {
                    for (int i = 0; i < 3; ++i) hit[i] = h[i];
                    uninvolved.push_back(c);
                    involved.push_back(a);
                    involved.push_back(b);
                }
            }
            clampBelowZero(hit);
            V3 edgePt = facePoint(m, curFace, hit);
            V3 toI = edgePt - src;
            for (int i = 0; i < 3; ++i) srcB[i] = hit[i];
            src = edgePt;
            ++crossings;
            if (uninvolved.size() == 2 && m.border[involved[0]]) {
                int vI = involved[0];
                if (m.boundary == 0) throw OracleError(-9, "border vertex reached on a closed surface");
                // openMeshSpace.cpp:186-192 and the vertex handlers (absorbingOpenMeshSpace.cpp:27-50,
                // tangentialOpenMeshSpace.cpp:41-64)
                V3 heading;
                int towards;
                curFace = borderVertexHeading(vI, curFace, disp, heading, towards);
                std::array<V3, 3> fresh = {corner(m, curFace, 0), corner(m, curFace, 1), corner(m, curFace, 2)};
                baryOf(fresh[0], fresh[1], fresh[2], m.X[vI], srcB);
                if (!vecs.empty()) {
                    // projectVectorsForBoundaryVertex (openMeshSpace.cpp:72-96), with the caller's (previous) normal
                    V3 orth = normalized(cross(n, heading));
                    V3 inside(0, 0, 0);
                    for (int k = 0; k < 3; ++k) {
                        int sv = m.fv(curFace, k);
                        if (sv == vI || sv == towards) continue;
                        inside = m.X[sv];
                    }
                    projectIfOverBoundary(vecs, orth, inside - m.X[vI]);
                }
                lastA = lastB = -1;
                if (m.boundary == 1) {
                    for (int i = 0; i < 3; ++i) tgtB[i] = srcB[i];
                    break;
                }
                V3 dh = normalized(heading);
                disp = dh * dot(disp, dh);
                target = facePoint(m, curFace, srcB) + disp;
                continue;
            }
            if (uninvolved.size() == 2) {
                int vI = involved[0];
                V3 heading;
                int tf = throughVertex(vI, toI, curFace, heading);
                double len = norm(disp);
                V3 tn = m.normal(tf);
                disp = heading * len;
                target = src + disp;
                if (!vecs.empty()) {
                    double nd2 = dot(n, tn);
                    double ang = std::acos(nd2);
                    V3 axr = cross(n, tn);
                    if (nd2 < 0 && norm(axr) > 0) {
                        V3 ax = normalized(axr);
                        for (auto& v : vecs) v = rotateAbout(src + v, src, ax, ang) - src;
                    }
                }
                curFace = tf;
                // the caller's corner copy is left as it was (reference passes it by value)
                std::array<V3, 3> fresh = {corner(m, curFace, 0), corner(m, curFace, 1), corner(m, curFace, 2)};
                baryOf(fresh[0], fresh[1], fresh[2], src, srcB);
                n = tn;
                lastA = lastB = -1;
                continue;
            }
            if (uninvolved.size() != 1) throw OracleError(-10, "target outside the face but neither 1 nor 2 edge crossings found");
            int a = involved[0], b = involved[1];
            int nf = across(curFace, a, b);
            if (nf < 0 && m.boundary == 0) throw OracleError(-11, "border edge reached on a closed surface");
            if (nf < 0) {
                // openMeshSpace.cpp:217-222 and the edge handlers
                V3 ev1 = m.X[a], ev2 = m.X[b];
                V3 at = facePoint(m, curFace, srcB);
                V3 inward = corners[uninvolved[0]] - at;
                lastA = lastB = -1;
                if (m.boundary == 1) {
                    // absorbingOpenMeshSpace.cpp:2-25: stop on the edge
                    if (!vecs.empty()) projectIfOverBoundary(vecs, normalized(cross(m.normal(curFace), ev2 - ev1)), inward);
                    for (int i = 0; i < 3; ++i) tgtB[i] = srcB[i];
                    break;
                }
                // tangentialOpenMeshSpace.cpp:3-39: slide along the edge
                V3 fwd = normalized(ev2 - ev1), bwd = normalized(ev1 - ev2);
                V3 dn = normalized(disp);
                double fd = dot(dn, fwd), bd = dot(dn, bwd);
                double len = norm(disp);
                projectIfOverBoundary(vecs, normalized(cross(n, ev2 - ev1)), inward);
                bool stop = fd <= 0 && bd <= 0;
                disp = fd > bd ? fwd * (fd * len) : bwd * (bd * len);
                target = at + disp;
                if (stop) break;
                continue;
            }
            src = facePoint(m, curFace, hit);
            V3 tn = m.normal(nf);
            double nd2 = dot(n, tn);
            double ang = std::acos(nd2);
            V3 ax = cross(n, tn);
            double axl = norm(ax);
            // coplanar neighbours whose normals' dot product rounds below 1 give a zero axis; the reference divides
            // by it and then throws on the NaN coordinates, here the rotation is skipped instead
            if (nd2 < 1 && axl > 0) {
                ax = ax / axl;
                target = rotateAbout(target, src, ax, ang);
                disp = target - src;
                for (auto& v : vecs) v = rotateAbout(src + v, src, ax, ang) - src;
            }
            lastA = a;
            lastB = b;
            curFace = nf;
            corners = {corner(m, curFace, 0), corner(m, curFace, 1), corner(m, curFace, 2)};
            baryOf(corners[0], corners[1], corners[2], src, srcB);
            n = tn;
        }
        face = curFace;
        for (int i = 0; i < 3; ++i) bary[i] = tgtB[i];
        return crossings;
    }
};

}  // namespace orc

using namespace orc;

extern "C" {

int orc_mesh_set_boundary(orc_mesh* om, int mode) {
    if (mode < 0 || mode > 2) return -1;
    om->m.boundary = mode;
    return 0;
}

// One particle: face/bary/disp are in-out (3 doubles each); vecs holds nVec transported vectors (3 doubles each).
// Returns the number of face-boundary crossings (>= 0) or a negative error code.
int orc_transport(const orc_mesh* om, int* face, double* bary, double* disp, int nVec, double* vecs) {
    try {
        Walker w(om->m);
        V3 d(disp[0], disp[1], disp[2]);
        std::vector<V3> vs(nVec);
        for (int i = 0; i < nVec; ++i) vs[i] = V3(vecs[3 * i], vecs[3 * i + 1], vecs[3 * i + 2]);
        int c = w.run(*face, bary, d, vs, 1000000);
        disp[0] = d.x;
        disp[1] = d.y;
        disp[2] = d.z;
        for (int i = 0; i < nVec; ++i) {
            vecs[3 * i] = vs[i].x;
            vecs[3 * i + 1] = vs[i].y;
            vecs[3 * i + 2] = vs[i].z;
        }
        return c;
    } catch (OracleError& e) {
        return e.code;
    }
}

// batched form: n particles, each with nVec vectors stored contiguously (vecs: n*nVec*3). status[i] receives the
// per-particle return value of orc_transport.
int orc_transport_batch(const orc_mesh* om, int n, int* face, double* bary, double* disp, int nVec, double* vecs, int* status) {
    int bad = 0;
    for (int i = 0; i < n; ++i) {
        status[i] = orc_transport(om, face + i, bary + 3 * i, disp + 3 * i, nVec, vecs ? vecs + 3 * nVec * i : nullptr);
        if (status[i] < 0) ++bad;
    }
    return bad;
}

// point with construct_point semantics
void orc_positions(const orc_mesh* om, int n, const int* face, const double* bary, double* out) {
    for (int i = 0; i < n; ++i) {
        V3 p = facePoint(om->m, face[i], bary + 3 * i);
        out[3 * i] = p.x;
        out[3 * i + 1] = p.y;
        out[3 * i + 2] = p.z;
    }
}

}  // extern "C"
