// Batched straight walk across faces with parallel transport on gfx950: one lane per particle, FP64 throughout.
//
// Device restatement of triangulatedMeshSpace::transportParticleAndVectors (src/models/triangulatedMeshSpace.cpp:448-564)
// with its clamps (src/utility/meshUtilities.cpp:138-182), edge predicates (:229-331), edge crossing (:614-658),
// vertex crossing (:566-611, throughVertex :247-407) and Rodrigues rotation (src/utility/functionUtilities.cpp:8-45).
// Kept because the reference behaves so: the face corner copy is not refreshed after a vertex crossing; a vertex
// crossing keeps the pre-crossing remaining length and rotates transported vectors only when the normals point into
// opposite half-spaces; the crossing vertex is the first vertex of the first edge found. One deliberate difference:
// a rotation about a zero-length axis (coplanar neighbours) is skipped where the reference produces NaN and throws.
//
// Open meshes: with an absorbing or tangential boundary the walk follows openMeshSpace::transportParticleAndVectors
// (src/models/openMeshSpace.cpp:114-238). At a border edge the absorbing handler (absorbingOpenMeshSpace.cpp:2-25) stops
// the particle on the edge and the tangential one (tangentialOpenMeshSpace.cpp:3-39) turns the remaining displacement
// along the edge; at a border vertex (openMeshSpace.cpp:3-96, absorbingOpenMeshSpace.cpp:27-50,
// tangentialOpenMeshSpace.cpp:41-64) the particle stops on the vertex or turns along the heading of largest overlap. As in
// the reference, the handlers see the displacement measured from the pre-crossing source, and after a border vertex the
// corner copy and the face normal still describe the previous face (the reference passes both by value). The candidate
// headings at a border vertex are the edges to the other corners of every incident face except the current one, faces in
// increasing index and corners in face order, first largest overlap winning (the reference pairs neighbours and faces
// through CGAL's stored vertex halfedges, an order that depends on how the mesh was built).
#include <math.h>

#include "csg_common.h"
#include "transport_kernel.h"

// No fused multiply-adds are formed from separate products and sums in this file: the walk takes decisions on exact
// zeros and signs (a displacement along a border edge has an exactly zero normal component, so the tangential handler
// stops at a convex corner; barycentric signs decide the crossed edge), and separately rounded operations make those
// decisions the same as in the reference's IEEE arithmetic.
#pragma clang fp contract(off)

namespace csg {

namespace {

struct D3 {
    double x, y, z;
};
__device__ inline D3 mk(double x, double y, double z) { return D3{x, y, z}; }
__device__ inline D3 add(D3 a, D3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline D3 sub(D3 a, D3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline D3 mul(D3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline D3 cross(D3 a, D3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ inline double len(D3 a) { return sqrt(dot(a, a)); }

__device__ inline D3 divd(D3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
// the reference's normalize (src/utility/functionUtilities.cpp:6): components divided by the length
__device__ inline D3 unit(D3 a) { return divd(a, len(a)); }
__device__ inline D3 normalized(D3 a) { return unit(a); }

__device__ inline D3 vert(const DeviceMesh& M, int v) { return mk(M.vx[3 * v], M.vx[3 * v + 1], M.vx[3 * v + 2]); }
__device__ inline D3 corner(const DeviceMesh& M, int f, int k) { return vert(M, M.fv[3 * f + k]); }
__device__ inline D3 normalOf(const DeviceMesh& M, int f) { return mk(M.fnrm[3 * f], M.fnrm[3 * f + 1], M.fnrm[3 * f + 2]); }

// point of a face from weights, normalised by their sum (construct_point semantics)
__device__ inline D3 pointOf(D3 c0, D3 c1, D3 c2, const double* b) {
    double s = b[0] + b[1] + b[2];
    return mk((c0.x * b[0] + c1.x * b[1] + c2.x * b[2]) / s, (c0.y * b[0] + c1.y * b[1] + c2.y * b[2]) / s,
              (c0.z * b[0] + c1.z * b[1] + c2.z * b[2]) / s);
}

// barycentric coordinates by projection (Ericson, Real-Time Collision Detection, section 3.4)
__device__ inline void baryOf(D3 a, D3 b, D3 c, D3 q, double* out) {
    D3 e0 = sub(b, a), e1 = sub(c, a), e2 = sub(q, a);
    double d00 = dot(e0, e0), d01 = dot(e0, e1), d11 = dot(e1, e1), d20 = dot(e2, e0), d21 = dot(e2, e1);
    double den = d00 * d11 - d01 * d01;
    double v = (d11 * d20 - d01 * d21) / den;
    double w = (d00 * d21 - d01 * d20) / den;
    out[0] = 1.0 - v - w;
    out[1] = v;
    out[2] = w;
}

__device__ inline void clampBelowZero(double* b) {
    double s = 0;
    for (int i = 0; i < 3; ++i) {
        b[i] = fmax(b[i], 1e-11);
        s += b[i];
    }
    for (int i = 0; i < 3; ++i) b[i] /= s;
}

__device__ inline void clampNearZero(double* b) {
    double s = 0;
    for (int i = 0; i < 3; ++i) {
        if (b[i] > -1e-13 && b[i] < 1e-13) b[i] = 1e-13;
        s += b[i];
    }
    for (int i = 0; i < 3; ++i) b[i] /= s;
}

__device__ inline D3 rotateAbout(D3 p, D3 base, D3 axisDir, double ang) {
    D3 a = divd(axisDir, len(axisDir));
    D3 q = sub(p, base);
    double ad = dot(a, q);
    double s = sin(ang), c = cos(ang);
    D3 r = add(add(mul(a, ad * (1.0 - c)), mul(q, c)), mul(cross(a, q), s));
    return add(r, base);
}

__device__ inline double angleOf(D3 u, D3 v) {
    u = unit(u);
    v = unit(v);
    double d = dot(u, v);
    if (d >= 1) return 0;
    if (d <= -1) return M_PI;
    return acos(d);
}

// crossing of the barycentric segment s->t with the edge where coordinate c vanishes (from corner c+1 to c+2)
__device__ inline bool crossEdge(const double* s, const double* t, int c, double* hit) {
    double den = s[c] - t[c];
    if (den == 0) return false;
    double lam = s[c] / den;
    int b = (c + 2) % 3;
    double mu = s[b] + lam * (t[b] - s[b]);
    for (int i = 0; i < 3; ++i) hit[i] = s[i] + lam * (t[i] - s[i]);
    return lam >= 0 && lam <= 1 && mu >= 0 && mu <= 1;
}

__device__ inline int across(const DeviceMesh& M, int f, int a, int b) {
    for (int k = 0; k < 3; ++k) {
        int u = M.fv[3 * f + k], v = M.fv[3 * f + (k + 1) % 3];
        if ((u == a && v == b) || (u == b && v == a)) return M.fadj[3 * f + k] < 0 ? -1 : (M.fadj[3 * f + k] >> 2);
    }
    return -1;
}

// projectVectorsIfOverBoundary (src/models/triangulatedMeshSpace.cpp:411-426): a vector pointing over the border line
// (orthogonal: the line's normal direction, inward: towards the mesh) loses its component along that normal
__device__ inline void projectIfOverBoundary(D3* vec, int nv, D3 orthogonal, D3 inward) {
    const bool pointsOut = dot(orthogonal, inward) < 0;
    for (int j = 0; j < nv; ++j) {
        const bool alongPerp = dot(vec[j], orthogonal) > 0;
        if ((pointsOut && alongPerp) || (!pointsOut && !alongPerp)) {
            D3 dh = normalized(orthogonal);
            vec[j] = sub(vec[j], mul(dh, dot(vec[j], dh)));
        }
    }
}

// getBoundaryVertexHeading (src/models/openMeshSpace.cpp:3-70) with the candidate rule in the header comment: returns
// the next face (or TRANSPORT_ERR_NO_FACE) and writes the unit heading and the vertex it points to
__device__ int borderVertexHeading(const DeviceMesh& M, int v, int sourceFace, D3 disp, D3& heading, int& towards) {
    const D3 dn = normalized(disp);
    const D3 pv = vert(M, v);
    double best = -1;
    int next = sourceFace;
    towards = v;
    heading = mk(0, 0, 0);
    for (int q = M.vfoff[v]; q < M.vfoff[v + 1]; ++q) {
        const int g = M.vflist[q] >> 2, k = M.vflist[q] & 3;
        if (g == sourceFace) continue;
        for (int j = 1; j <= 2; ++j) {
            const int w = M.fv[3 * g + (k + j) % 3];
            const D3 out = normalized(sub(vert(M, w), pv));
            const double ov = dot(out, dn);
            if (ov > best) {
                heading = out;
                best = ov;
                next = g;
                towards = w;
            }
        }
    }
    return next == sourceFace ? TRANSPORT_ERR_NO_FACE : next;
}

constexpr int kMaxRing = 32;

// straightest continuation through vertex vI: returns the next face (or a negative error), writes the heading
__device__ int throughVertex(const DeviceMesh& M, int vI, D3 toI, int sourceFace, D3& heading) {
    D3 vi = vert(M, vI);
    int start = -1;
    for (int k = 0; k < 3; ++k)
        if (M.fv[3 * sourceFace + k] == vI) start = M.fv[3 * sourceFace + (k + 1) % 3];
    if (start < 0) return TRANSPORT_ERR_NO_FACE;
    int ring[kMaxRing];
    int n = 0;
    int f = sourceFace, cur = start;
    while (true) {
        if (n >= kMaxRing) return TRANSPORT_ERR_RING;
        ring[n++] = cur;
        int g = across(M, f, vI, cur);
        if (g < 0) return TRANSPORT_ERR_BORDER;
        int nxt = -1;
        for (int k = 0; k < 3; ++k)
            if (M.fv[3 * g + k] == vI) nxt = M.fv[3 * g + (k + 1) % 3];
        f = g;
        if (nxt == start) break;
        cur = nxt;
    }
    double total = 0;
    for (int i = 0; i < n; ++i) total += angleOf(sub(vert(M, ring[i]), vi), sub(vert(M, ring[(i + 1) % n]), vi));
    double half = total / 2.0;
    double travelled = angleOf(mul(toI, -1.0), sub(vert(M, ring[0]), vi));
    int counter = 0;
    while (travelled < half) {
        travelled += angleOf(sub(vert(M, ring[counter]), vi), sub(vert(M, ring[(counter + 1) % n]), vi));
        counter++;
        if (counter >= n) return TRANSPORT_ERR_NO_HALFWAY;
    }
    if (counter == 0) return TRANSPORT_ERR_NO_HALFWAY;
    double diff = half - travelled;
    D3 e0 = unit(sub(vert(M, ring[counter - 1]), vi)), e1 = unit(sub(vert(M, ring[counter]), vi));
    D3 axis = unit(cross(e0, e1));
    D3 pr = rotateAbout(add(vi, e1), vi, axis, diff);
    heading = unit(sub(pr, vi));
    int a = ring[counter - 1], b = ring[counter];
    for (int q = M.vfoff[vI]; q < M.vfoff[vI + 1]; ++q) {
        int g = M.vflist[q] >> 2;
        if (g == sourceFace) continue;
        int c0 = M.fv[3 * g], c1 = M.fv[3 * g + 1], c2 = M.fv[3 * g + 2];
        bool hasA = c0 == a || c1 == a || c2 == a, hasB = c0 == b || c1 == b || c2 == b;
        bool hasV = c0 == vI || c1 == vI || c2 == vI;
        if (hasA && hasB && hasV) return g;
    }
    return TRANSPORT_ERR_NO_FACE;
}

__global__ __launch_bounds__(256) void transportKernel(TransportParams P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const DeviceMesh& M = P.mesh;
    int curFace = P.face[i];
    double srcB[3] = {P.bary[3 * i], P.bary[3 * i + 1], P.bary[3 * i + 2]};
    D3 cc0 = corner(M, curFace, 0), cc1 = corner(M, curFace, 1), cc2 = corner(M, curFace, 2);
    D3 src = pointOf(cc0, cc1, cc2, srcB);
    D3 n = normalOf(M, curFace);
    D3 disp = mk(P.disp[3 * i], P.disp[3 * i + 1], P.disp[3 * i + 2]);
    if (P.scale != 1.0) disp = mul(disp, P.scale);
    const int nv = P.nVec;
    D3 vec[kMaxTransportVectors];
    for (int j = 0; j < nv; ++j) {
        const double* v = P.vecs + 3 * (i * nv + j);
        vec[j] = mk(v[0], v[1], v[2]);
    }
    double nd = dot(n, disp);
    if (fabs(nd) > 1e-14) disp = sub(disp, mul(n, nd));
    D3 target = add(src, disp);
    int lastA = -1, lastB = -1;
    double tgtB[3];
    int crossings = 0, err = 0;
    for (int iter = 0;; ++iter) {
        if (iter > P.maxCrossings) {
            err = TRANSPORT_ERR_TOO_MANY;
            break;
        }
        // corner copy `cc*` is the reference's vertexPositions, stale after a vertex crossing
        baryOf(cc0, cc1, cc2, target, tgtB);
        clampNearZero(tgtB);
        D3 f0 = corner(M, curFace, 0), f1 = corner(M, curFace, 1), f2 = corner(M, curFace, 2);
        target = pointOf(f0, f1, f2, tgtB);
        clampBelowZero(srcB);
        src = pointOf(f0, f1, f2, srcB);
        disp = sub(target, src);
        if (tgtB[0] != tgtB[0]) {
            err = TRANSPORT_ERR_NAN;
            break;
        }
        if (tgtB[0] >= 0 && tgtB[1] >= 0 && tgtB[2] >= 0) break;
        int vl[3] = {M.fv[3 * curFace], M.fv[3 * curFace + 1], M.fv[3 * curFace + 2]};
        double hit[3] = {0, 0, 0}, h[3];
        int nHits = 0, firstA = -1, firstB = -1, firstC = 0;
        const int order[3] = {2, 0, 1};  // V1V2, V2V3, V3V1
        for (int oi = 0; oi < 3; ++oi) {
            int c = order[oi];
            int a = vl[(c + 1) % 3], b = vl[(c + 2) % 3];
            if ((a == lastA && b == lastB) || (a == lastB && b == lastA)) continue;
            if (crossEdge(srcB, tgtB, c, h)) {
                for (int q = 0; q < 3; ++q) hit[q] = h[q];
                if (nHits == 0) {
                    firstA = a;
                    firstB = b;
                    firstC = c;
                }
                ++nHits;
            }
        }
        clampBelowZero(hit);
        D3 edgePt = pointOf(f0, f1, f2, hit);
        D3 toI = sub(edgePt, src);
        for (int q = 0; q < 3; ++q) srcB[q] = hit[q];
        src = edgePt;
        ++crossings;
        if (nHits == 2 && (M.vflag[firstA] & 2)) {
            // border vertex (openMeshSpace.cpp:186-192)
            if (P.boundary == CSG_BOUNDARY_CLOSED) {
                err = TRANSPORT_ERR_BORDER;
                break;
            }
            const int vI = firstA;
            D3 heading;
            int towards;
            const int nf = borderVertexHeading(M, vI, curFace, disp, heading, towards);
            if (nf < 0) {
                err = nf;
                break;
            }
            curFace = nf;
            const D3 pv = vert(M, vI);
            const D3 g0 = corner(M, curFace, 0), g1 = corner(M, curFace, 1), g2 = corner(M, curFace, 2);
            baryOf(g0, g1, g2, pv, srcB);
            if (nv > 0) {
                // projectVectorsForBoundaryVertex (openMeshSpace.cpp:72-96), with the previous face's normal
                D3 inside = mk(0, 0, 0);
                for (int k = 0; k < 3; ++k) {
                    const int sv = M.fv[3 * curFace + k];
                    if (sv != vI && sv != towards) inside = vert(M, sv);
                }
                projectIfOverBoundary(vec, nv, normalized(cross(n, heading)), sub(inside, pv));
            }
            lastA = lastB = -1;
            if (P.boundary == CSG_BOUNDARY_ABSORBING) {
                for (int q = 0; q < 3; ++q) tgtB[q] = srcB[q];
                break;
            }
            const D3 dh = normalized(heading);
            disp = mul(dh, dot(disp, dh));
            target = add(pointOf(g0, g1, g2, srcB), disp);
            continue;
        }
        if (nHits == 2) {
            int vI = firstA;
            D3 heading;
            int tf = throughVertex(M, vI, toI, curFace, heading);
            if (tf < 0) {
                err = tf;
                break;
            }
            double l = len(disp);
            D3 tn = normalOf(M, tf);
            disp = mul(heading, l);
            target = add(src, disp);
            if (nv > 0) {
                double nd2 = dot(n, tn);
                double ang = acos(nd2);
                D3 axr = cross(n, tn);
                if (nd2 < 0 && len(axr) > 0) {
                    D3 ax = unit(axr);
                    for (int j = 0; j < nv; ++j) vec[j] = sub(rotateAbout(add(src, vec[j]), src, ax, ang), src);
                }
            }
            curFace = tf;
            baryOf(corner(M, curFace, 0), corner(M, curFace, 1), corner(M, curFace, 2), src, srcB);
            n = tn;
            lastA = lastB = -1;
            continue;
        }
        if (nHits != 1) {
            err = TRANSPORT_ERR_NO_CROSSING;
            break;
        }
        int nf = across(M, curFace, firstA, firstB);
        if (nf < 0) {
            // border edge (openMeshSpace.cpp:217-222)
            if (P.boundary == CSG_BOUNDARY_CLOSED) {
                err = TRANSPORT_ERR_BORDER;
                break;
            }
            const D3 ev1 = vert(M, firstA), ev2 = vert(M, firstB);
            const D3 at = pointOf(f0, f1, f2, srcB);
            const D3 inward = sub(firstC == 0 ? cc0 : (firstC == 1 ? cc1 : cc2), at);
            lastA = lastB = -1;
            if (P.boundary == CSG_BOUNDARY_ABSORBING) {
                if (nv > 0) projectIfOverBoundary(vec, nv, normalized(cross(normalOf(M, curFace), sub(ev2, ev1))), inward);
                for (int q = 0; q < 3; ++q) tgtB[q] = srcB[q];
                break;
            }
            const D3 fwd = normalized(sub(ev2, ev1)), bwd = normalized(sub(ev1, ev2));
            const D3 dn = normalized(disp);
            const double fd = dot(dn, fwd), bd = dot(dn, bwd);
            const double l = len(disp);
            projectIfOverBoundary(vec, nv, normalized(cross(n, sub(ev2, ev1))), inward);
            const bool stop = fd <= 0 && bd <= 0;
            disp = fd > bd ? mul(fwd, fd * l) : mul(bwd, bd * l);
            target = add(at, disp);
            if (stop) break;  // the barycentrics computed at the top of this iteration stand
            continue;
        }
        src = pointOf(f0, f1, f2, hit);
        D3 tn = normalOf(M, nf);
        double nd2 = dot(n, tn);
        double ang = acos(nd2);
        D3 ax = cross(n, tn);
        double axl = len(ax);
        if (nd2 < 1 && axl > 0) {
            ax = divd(ax, axl);
            target = rotateAbout(target, src, ax, ang);
            disp = sub(target, src);
            for (int j = 0; j < nv; ++j) vec[j] = sub(rotateAbout(add(src, vec[j]), src, ax, ang), src);
        }
        lastA = firstA;
        lastB = firstB;
        curFace = nf;
        cc0 = corner(M, curFace, 0);
        cc1 = corner(M, curFace, 1);
        cc2 = corner(M, curFace, 2);
        baryOf(cc0, cc1, cc2, src, srcB);
        n = tn;
    }
    if (err) {
        if (P.status) P.status[i] = err;
        return;  // the particle keeps its input state
    }
    P.face[i] = curFace;
    for (int q = 0; q < 3; ++q) P.bary[3 * i + q] = tgtB[q];
    if (P.dispOut) {
        P.dispOut[3 * i] = disp.x;
        P.dispOut[3 * i + 1] = disp.y;
        P.dispOut[3 * i + 2] = disp.z;
    }
    for (int j = 0; j < nv; ++j) {
        double* v = P.vecs + 3 * (i * nv + j);
        v[0] = vec[j].x;
        v[1] = vec[j].y;
        v[2] = vec[j].z;
    }
    if (P.status) P.status[i] = crossings;
}

}  // namespace

hipError_t launchTransport(const TransportParams& P, hipStream_t stream) {
    if (P.n <= 0) return hipSuccess;
    if (P.nVec < 0 || P.nVec > kMaxTransportVectors) return hipErrorInvalidValue;
    hipLaunchKernelGGL(transportKernel, dim3((P.n + 255) / 256), dim3(256), 0, stream, P);
    return hipGetLastError();
}

}  // namespace csg
