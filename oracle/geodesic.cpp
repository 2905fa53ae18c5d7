// ORACLE — test infrastructure only (see oracle_internal.hpp).
//
// Exact single-source polyhedral geodesics by window (interval) propagation, FP64, sequential, processed in order of
// each window's lower-bound distance (a priority queue), as in the improved Chen-Han method of Xin & Wang (2009,
// "Improving Chen and Han's algorithm on the discrete geodesic problem"), which is the method behind the CGAL
// Surface_mesh_shortest_path package that the reference calls (src/utility/meshUtilities.h:44-48; build at
// src/models/triangulatedMeshSpace.cpp:189-193 and :229-230; query at src/utility/meshUtilities.cpp:360-380).
//
// Windows live on a directed edge (face entered, local edge k) in a 2D frame: origin at face vertex k, x along the
// edge to vertex k+1, y into the entered face. A window holds an interval [b0,b1] of the edge, the unfolded
// pseudo-source (sx, sy<0) and sigma, the geodesic distance from the real source to that pseudo-source. Pseudo-sources
// are the source point and the vertices where paths may bend: saddle vertices (angle sum >= 2*pi) and the vertices
// on the border of the face set being solved. A window is dropped when the current vertex distance bounds prove that
// it cannot carry a shortest path anywhere on its interval (Xin & Wang's filter, stated here per window endpoint).
//
// Start tangents (at the source, pointing along the path) and end tangents (at the target, pointing away from the
// source) follow the reference's convention (src/utility/meshUtilities.cpp:373-379).
#include <algorithm>
#include <memory>
#include <cstring>
#include <limits>
#include <queue>
#include <unordered_set>
#include <stack>

#include "oracle.h"
#include "oracle_internal.hpp"

namespace orc {

static const double INF = std::numeric_limits<double>::infinity();

struct Frame {
    V3 o, ex, ey;
    double L, x2, y2;
};

static Frame edgeFrame(const Mesh& m, int f, int k) {
    Frame fr;
    V3 p0 = m.P(f, k), p1 = m.P(f, k + 1), p2 = m.P(f, k + 2);
    V3 e = p1 - p0;
    fr.L = norm(e);
    fr.o = p0;
    fr.ex = e / fr.L;
    V3 n = normalized(cross(p1 - p0, p2 - p0));
    fr.ey = cross(n, fr.ex);
    V3 d2 = p2 - p0;
    fr.x2 = dot(d2, fr.ex);
    fr.y2 = dot(d2, fr.ey);
    return fr;
}

// point with the reference's construct_point semantics: weights are normalised by their sum
static V3 baryPoint(const Mesh& m, int f, const double* b) {
    double s = b[0] + b[1] + b[2];
    return (m.P(f, 0) * b[0] + m.P(f, 1) * b[1] + m.P(f, 2) * b[2]) / s;
}

struct Win {
    int f, k;
    double b0, b1, sx, sy, sigma;
    int rooted;        // 1: start tangent = R(q - s); 0: fixed start tangent (tx,ty)
    double c, s;       // rotation from this edge frame to the source face frame
    double tx, ty;
    double key;
};

struct Cand {
    double d, tx, ty;
};

struct Solver {
    const Mesh& m;
    std::vector<char> inPatch;   // per face
    std::vector<char> pseudo;    // per vertex: may act as a pseudo-source
    std::vector<double> dv;      // vertex distance upper bounds
    std::vector<double> tvx, tvy;
    std::vector<Win> wins;
    struct Ev {
        double key;
        long long seq;
        int type, idx;
        bool operator<(const Ev& o) const { return key > o.key || (key == o.key && seq > o.seq); }
    };
    std::priority_queue<Ev> pq;
    long long seq = 0;
    double rMax;
    // source
    int sf;
    V3 S, t1, t2;
    // targets
    int nT;
    const int* tf;
    const double* tb;
    std::vector<std::vector<int>> targetsInFace;  // indexed by face (sparse via map below)
    std::vector<int> faceTargetHead;
    std::vector<int> targetNext;
    std::vector<std::vector<Cand>> cands;
    std::vector<double> bestD;
    std::vector<V3> bestEt;
    std::vector<double> bestTx, bestTy;
    long long nWindowsProcessed = 0, nWindowsCreated = 0, nVertexEvents = 0;
    static constexpr double FILTER_TOL = 1e-12;

    explicit Solver(const Mesh& mm) : m(mm) {}

    void setPatch(const int* faces, int n) {
        inPatch.assign(m.nF, n < 0 ? 1 : 0);
        if (n >= 0)
            for (int i = 0; i < n; ++i) inPatch[faces[i]] = 1;
        pseudo.assign(m.nV, 0);
        for (int v = 0; v < m.nV; ++v) {
            if (m.border[v] || m.angleSum[v] >= 2 * M_PI - 1e-12) pseudo[v] = 1;
        }
        // vertices of patch faces that touch a face outside the patch are on the patch border
        if (n >= 0) {
            for (int i = 0; i < n; ++i) {
                int f = faces[i];
                for (int k = 0; k < 3; ++k) {
                    int v = m.fv(f, k);
                    for (int q = m.vfOff[v]; q < m.vfOff[v + 1]; ++q)
                        if (!inPatch[m.vfList[q] >> 2]) {
                            pseudo[v] = 1;
                            break;
                        }
                }
            }
        }
    }

    void toSourceFrame(double wx, double wy, double& ox, double& oy) const {  // identity helper
        ox = wx;
        oy = wy;
    }

    void startTangent(const Win& w, double rx, double ry, double& tx, double& ty) const {
        if (w.rooted) {
            double l = std::sqrt(rx * rx + ry * ry);
            if (l == 0) {
                tx = ty = 0;
                return;
            }
            rx /= l;
            ry /= l;
            tx = w.c * rx - w.s * ry;
            ty = w.s * rx + w.c * ry;
        } else {
            tx = w.tx;
            ty = w.ty;
        }
    }

    void candVertex(int v, double d, double tx, double ty) {
        if (d < dv[v]) {
            dv[v] = d;
            tvx[v] = tx;
            tvy[v] = ty;
            if (pseudo[v] && d <= rMax) pq.push({d, seq++, 1, v});
        }
    }

    static double keyOf(const Win& w) {
        double dmin;
        if (w.sx < w.b0)
            dmin = std::hypot(w.b0 - w.sx, w.sy);
        else if (w.sx > w.b1)
            dmin = std::hypot(w.b1 - w.sx, w.sy);
        else
            dmin = -w.sy;
        return w.sigma + dmin;
    }

    bool filtered(const Win& w, double L) const {
        int A = m.fv(w.f, w.k), B = m.fv(w.f, w.k + 1);
        double toI1 = w.sigma + std::hypot(w.b1 - w.sx, w.sy);
        double toI0 = w.sigma + std::hypot(w.b0 - w.sx, w.sy);
        if (toI1 > dv[A] + w.b1 + FILTER_TOL * toI1) return true;
        if (toI0 > dv[B] + (L - w.b0) + FILTER_TOL * toI0) return true;
        return false;
    }

    // a new window on (f,k): endpoint updates, filter, push
    void addWindow(Win w, double L) {
        if (!(w.b1 > w.b0)) return;
        w.key = keyOf(w);
        if (w.key > rMax) return;
        const double tolE = 1e-12 * L;
        double tx, ty;
        if (w.b0 <= tolE) {
            startTangent(w, -w.sx, -w.sy, tx, ty);
            candVertex(m.fv(w.f, w.k), w.sigma + std::hypot(w.sx, w.sy), tx, ty);
        }
        if (w.b1 >= L - tolE) {
            startTangent(w, L - w.sx, -w.sy, tx, ty);
            candVertex(m.fv(w.f, w.k + 1), w.sigma + std::hypot(L - w.sx, w.sy), tx, ty);
        }
        if (filtered(w, L)) return;
        ++nWindowsCreated;
        wins.push_back(w);
        pq.push({w.key, seq++, 0, (int)wins.size() - 1});
    }

    // pseudo-source point Q (3D) placed in the frame of the window entering face h through local edge kh
    Win windowFromPoint(int h, int kh, const V3& Q, double sigma, const Frame& fr) const {
        Win w;
        w.f = h;
        w.k = kh;
        w.b0 = 0;
        w.b1 = fr.L;
        V3 r = Q - fr.o;
        w.sx = dot(r, fr.ex);
        w.sy = -norm(cross(r, fr.ex));
        if (w.sy > -1e-14 * fr.L) w.sy = -1e-14 * fr.L;
        w.sigma = sigma;
        w.rooted = 0;
        w.c = 1;
        w.s = 0;
        w.tx = w.ty = 0;
        return w;
    }

    void recordTarget(int t, double d, double tx, double ty, const V3& et) {
        cands[t].push_back({d, tx, ty});
        if (d < bestD[t]) {
            bestD[t] = d;
            bestTx[t] = tx;
            bestTy[t] = ty;
            bestEt[t] = et;
        }
    }

    void evalTargets(const Win& w, const Frame& fr) {
        for (int t = faceTargetHead.empty() ? -1 : faceTargetHead[w.f]; t >= 0; t = targetNext[t]) {
            const double* b = tb + 3 * t;
            double sb = b[0] + b[1] + b[2];
            double bk1 = b[(w.k + 1) % 3] / sb, bk2 = b[(w.k + 2) % 3] / sb;
            double qx = bk1 * fr.L + bk2 * fr.x2, qy = bk2 * fr.y2;
            double rx = qx - w.sx, ry = qy - w.sy;
            double ax = w.b0 - w.sx, ay = -w.sy, bx = w.b1 - w.sx, by = -w.sy;
            double rl = std::hypot(rx, ry);
            double c1 = bx * ry - by * rx;  // cross(b, r)
            double c2 = rx * ay - ry * ax;  // cross(r, a)
            double tol = 1e-12 * rl;
            if (c1 < -tol * std::hypot(bx, by) || c2 < -tol * std::hypot(ax, ay)) continue;
            double d = w.sigma + rl;
            double tx, ty;
            startTangent(w, rx, ry, tx, ty);
            V3 et = (fr.ex * rx + fr.ey * ry) / rl;
            recordTarget(t, d, tx, ty, et);
        }
    }

    void processWindow(const Win& w) {
        if (w.key > rMax) return;
        Frame fr = edgeFrame(m, w.f, w.k);
        if (filtered(w, fr.L)) return;
        ++nWindowsProcessed;
        evalTargets(w, fr);
        const double sx = w.sx, sy = w.sy, L = fr.L, x2 = fr.x2, y2 = fr.y2;
        // crossing of the ray s -> P2 with the entry edge
        double xs = sx + (x2 - sx) * (-sy) / (y2 - sy);
        const double tolE = 1e-12 * L;
        if (xs >= w.b0 - tolE && xs <= w.b1 + tolE) {
            double tx, ty;
            startTangent(w, x2 - sx, y2 - sy, tx, ty);
            candVertex(m.fv(w.f, w.k + 2), w.sigma + std::hypot(x2 - sx, y2 - sy), tx, ty);
        }
        // left child: edge P0 -> P2 (local edge k+2 of f, seen from the neighbour as P0 -> P2)
        if (xs > w.b0) {
            int ke = (w.k + 2) % 3;
            int g = m.adj[3 * w.f + ke];
            if (g >= 0 && inPatch[g]) {
                double lL = std::hypot(x2, y2);
                double ux = x2 / lL, uy = y2 / lL;
                double spx = sx * ux + sy * uy, spy = ux * sy - uy * sx;
                auto hit = [&](double bx) {
                    double dx = bx - sx, dy = -sy;
                    double dpx = dx * ux + dy * uy, dpy = ux * dy - uy * dx;
                    if (dpy <= 0) return dpx < 0 ? 0.0 : lL;
                    double t = spx + dpx * (-spy) / dpy;
                    return std::min(lL, std::max(0.0, t));
                };
                double t0 = (w.b0 <= 0) ? 0.0 : hit(w.b0);
                double t1 = (xs <= w.b1) ? lL : hit(w.b1);
                Win c;
                c.f = g;
                c.k = m.adjEdge[3 * w.f + ke];
                c.b0 = t0;
                c.b1 = t1;
                c.sx = spx;
                c.sy = std::min(spy, -1e-14 * lL);
                c.sigma = w.sigma;
                c.rooted = w.rooted;
                c.c = w.c * ux - w.s * uy;
                c.s = w.s * ux + w.c * uy;
                c.tx = w.tx;
                c.ty = w.ty;
                addWindow(c, lL);
            }
        }
        // right child: edge P2 -> P1 (local edge k+1 of f, seen from the neighbour as P2 -> P1)
        if (xs < w.b1) {
            int ke = (w.k + 1) % 3;
            int g = m.adj[3 * w.f + ke];
            if (g >= 0 && inPatch[g]) {
                double ex = L - x2, ey = -y2;
                double lR = std::hypot(ex, ey);
                double ux = ex / lR, uy = ey / lR;
                double rx0 = sx - x2, ry0 = sy - y2;
                double spx = rx0 * ux + ry0 * uy, spy = ux * ry0 - uy * rx0;
                auto hit = [&](double bx) {
                    double dx = bx - sx, dy = -sy;
                    double dpx = dx * ux + dy * uy, dpy = ux * dy - uy * dx;
                    if (dpy <= 0) return dpx < 0 ? 0.0 : lR;
                    double t = spx + dpx * (-spy) / dpy;
                    return std::min(lR, std::max(0.0, t));
                };
                double t0 = (xs >= w.b0) ? 0.0 : hit(w.b0);
                double t1 = (w.b1 >= L) ? lR : hit(w.b1);
                Win c;
                c.f = g;
                c.k = m.adjEdge[3 * w.f + ke];
                c.b0 = t0;
                c.b1 = t1;
                c.sx = spx;
                c.sy = std::min(spy, -1e-14 * lR);
                c.sigma = w.sigma;
                c.rooted = w.rooted;
                c.c = w.c * ux - w.s * uy;
                c.s = w.s * ux + w.c * uy;
                c.tx = w.tx;
                c.ty = w.ty;
                addWindow(c, lR);
            }
        }
    }

    void processVertex(int v, double d) {
        if (d != dv[v] || d > rMax) return;
        ++nVertexEvents;
        for (int q = m.vfOff[v]; q < m.vfOff[v + 1]; ++q) {
            int g = m.vfList[q] >> 2, j = m.vfList[q] & 3;
            if (!inPatch[g]) continue;
            int k = (j + 1) % 3;
            int h = m.adj[3 * g + k];
            if (h < 0 || !inPatch[h]) continue;
            int kh = m.adjEdge[3 * g + k];
            Frame fr = edgeFrame(m, h, kh);
            Win w = windowFromPoint(h, kh, m.X[v], d, fr);
            w.tx = tvx[v];
            w.ty = tvy[v];
            addWindow(w, fr.L);
        }
    }

    void solve(int srcFace, const double* srcBary, int nTargets, const int* tgtFace, const double* tgtBary, double rmax) {
        rMax = rmax;
        sf = srcFace;
        nT = nTargets;
        tf = tgtFace;
        tb = tgtBary;
        dv.assign(m.nV, INF);
        tvx.assign(m.nV, 0);
        tvy.assign(m.nV, 0);
        wins.clear();
        S = baryPoint(m, sf, srcBary);
        t1 = normalized(m.P(sf, 1) - m.P(sf, 0));
        t2 = cross(m.normal(sf), t1);
        faceTargetHead.assign(m.nF, -1);
        targetNext.assign(nT, -1);
        for (int t = nT - 1; t >= 0; --t) {
            targetNext[t] = faceTargetHead[tf[t]];
            faceTargetHead[tf[t]] = t;
        }
        cands.assign(nT, {});
        bestD.assign(nT, INF);
        bestEt.assign(nT, V3());
        bestTx.assign(nT, 0);
        bestTy.assign(nT, 0);
        // source face vertices are directly visible
        for (int k = 0; k < 3; ++k) {
            V3 r = m.P(sf, k) - S;
            double l = norm(r);
            candVertex(m.fv(sf, k), l, l > 0 ? dot(r, t1) / l : 0, l > 0 ? dot(r, t2) / l : 0);
        }
        // windows across the three edges of the source face
        for (int k = 0; k < 3; ++k) {
            int g = m.adj[3 * sf + k];
            if (g < 0 || !inPatch[g]) continue;
            int kg = m.adjEdge[3 * sf + k];
            Frame fr = edgeFrame(m, g, kg);
            Win w = windowFromPoint(g, kg, S, 0.0, fr);
            w.rooted = 1;
            w.c = dot(fr.ex, t1);
            w.s = dot(fr.ex, t2);
            addWindow(w, fr.L);
        }
        while (!pq.empty()) {
            Ev e = pq.top();
            pq.pop();
            if (e.type == 0) {
                Win w = wins[e.idx];  // by value: processing appends to `wins`
                processWindow(w);
            } else
                processVertex(e.idx, e.key);
        }
        // direct visibility inside the source face and paths through the target face's own vertices
        for (int t = 0; t < nT; ++t) {
            int f = tf[t];
            V3 Q = baryPoint(m, f, tb + 3 * t);
            if (f == sf) {
                V3 r = Q - S;
                double l = norm(r);
                if (l > 0)
                    recordTarget(t, l, dot(r, t1) / l, dot(r, t2) / l, r / l);
                else
                    recordTarget(t, 0.0, 0, 0, V3());
            }
            for (int k = 0; k < 3; ++k) {
                int v = m.fv(f, k);
                if (!(dv[v] < INF)) continue;
                V3 r = Q - m.X[v];
                double l = norm(r);
                recordTarget(t, dv[v] + l, tvx[v], tvy[v], l > 0 ? r / l : V3());
            }
        }
    }
};

}  // namespace orc

using namespace orc;

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

orc_mesh* orc_mesh_create(int nV, const double* X, int nF, const int* F) {
    try {
        std::unique_ptr<orc_mesh> om(new orc_mesh());  // freed if the mesh checks throw
        buildMesh(om->m, nV, X, nF, F);
        return om.release();
    } catch (std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_mesh_destroy(orc_mesh* m) { delete m; }

int orc_mesh_counts(const orc_mesh* om, int* nV, int* nF, int* nBorderVerts, int* nSaddle) {
    const Mesh& m = om->m;
    *nV = m.nV;
    *nF = m.nF;
    int b = 0, s = 0;
    for (int v = 0; v < m.nV; ++v) {
        b += m.border[v];
        s += (m.angleSum[v] > 2 * M_PI + 1e-12);
    }
    *nBorderVerts = b;
    *nSaddle = s;
    return 0;
}

int orc_mesh_adjacency(const orc_mesh* om, int* adj, int* adjEdge) {
    std::copy(om->m.adj.begin(), om->m.adj.end(), adj);
    std::copy(om->m.adjEdge.begin(), om->m.adjEdge.end(), adjEdge);
    return 0;
}

int orc_mesh_angle_sums(const orc_mesh* om, double* out) {
    std::copy(om->m.angleSum.begin(), om->m.angleSum.end(), out);
    return 0;
}

// Patch rule of submesher::constructSubmeshFromSourceAndTargets (src/utility/submesher.cpp:55-147):
// source face; early exit when every target lies in it; its edge neighbours; early exit when every target is covered;
// depth-first search over faces having at least one vertex within `radius` (squared distance not greater than radius^2)
// of the source point; finally every target face not yet visited.
int orc_patch(const orc_mesh* om, int srcFace, const double* srcBary, int nT, const int* tgtFace, double radius, int* outFaces,
              int cap) {
    const Mesh& m = om->m;
    V3 S = baryPoint(m, srcFace, srcBary);
    double r2 = radius * radius;
    std::unordered_set<int> visited, goal;
    visited.insert(srcFace);
    for (int t = 0; t < nT; ++t)
        if (tgtFace[t] != srcFace) goal.insert(tgtFace[t]);
    auto finish = [&]() -> int {
        std::vector<int> out(visited.begin(), visited.end());
        std::sort(out.begin(), out.end());
        if ((int)out.size() > cap) {
            g_err = "patch capacity exceeded";
            return -(int)out.size();
        }
        std::copy(out.begin(), out.end(), outFaces);
        return (int)out.size();
    };
    if (goal.empty()) return finish();
    std::stack<int> st;
    for (int k = 0; k < 3; ++k) {
        int g = m.adj[3 * srcFace + k];
        if (g < 0) continue;
        visited.insert(g);
        st.push(g);
        goal.erase(g);
    }
    if (goal.empty()) return finish();
    while (!st.empty()) {
        int f = st.top();
        st.pop();
        for (int k = 0; k < 3; ++k) {
            int g = m.adj[3 * f + k];
            if (g < 0 || visited.count(g)) continue;
            bool outside = true;
            for (int j = 0; j < 3; ++j) {
                V3 d = m.P(g, j) - S;
                if (!(dot(d, d) > r2)) outside = false;
            }
            if (outside) continue;
            visited.insert(g);
            goal.erase(g);
            st.push(g);
        }
    }
    for (int g : goal) visited.insert(g);
    return finish();
}

// Exact geodesics from one source to a list of targets over a face set (nPatch < 0: whole mesh).
// dist: +inf and flags bit0 when a target is not reached; flags bit1 when another path within a relative 1e-8 of the
// shortest leaves the source in a different direction (a tie: the start tangent is then not unique).
// Windows whose lower-bound distance exceeds rMax are not propagated; targets farther than rMax get flags bit2.
int orc_geodesic(const orc_mesh* om, const int* patchFaces, int nPatch, int srcFace, const double* srcBary, int nT, const int* tgtFace,
                 const double* tgtBary, double rMax, double* dist, double* startTan, double* endTan, int* flags, long long* stats) {
    try {
        const Mesh& m = om->m;
        Solver s(m);
        s.setPatch(patchFaces, nPatch);
        s.solve(srcFace, srcBary, nT, tgtFace, tgtBary, rMax);
        for (int t = 0; t < nT; ++t) {
            int fl = 0;
            double best = s.bestD[t];
            if (!(best < INF)) {
                fl |= 1;
                dist[t] = INF;
                for (int c = 0; c < 3; ++c) startTan[3 * t + c] = endTan[3 * t + c] = 0;
                flags[t] = fl;
                continue;
            }
            if (best > rMax) fl |= 4;
            double tol = 1e-8 * std::max(best, 1e-12);
            for (const Cand& c : s.cands[t]) {
                if (c.d <= best + tol) {
                    double cosang = c.tx * s.bestTx[t] + c.ty * s.bestTy[t];
                    if (cosang < 1 - 5e-9) fl |= 2;
                }
            }
            dist[t] = best;
            V3 st = s.t1 * s.bestTx[t] + s.t2 * s.bestTy[t];
            startTan[3 * t] = st.x;
            startTan[3 * t + 1] = st.y;
            startTan[3 * t + 2] = st.z;
            endTan[3 * t] = s.bestEt[t].x;
            endTan[3 * t + 1] = s.bestEt[t].y;
            endTan[3 * t + 2] = s.bestEt[t].z;
            flags[t] = fl;
        }
        if (stats) {
            stats[0] = s.nWindowsCreated;
            stats[1] = s.nWindowsProcessed;
            stats[2] = s.nVertexEvents;
        }
        return 0;
    } catch (OracleError& e) {
        g_err = e.what();
        return e.code;
    } catch (std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// triangulatedMeshSpace::distance semantics (src/models/triangulatedMeshSpace.cpp:155-238): with submeshing, the patch
// radius is min(maximumDistance, distanceThreshold), the solve covers the whole patch, and a target the patch does
// not connect to the source gets distance 2*maximumDistance and tangents (0,0,1). Without submeshing the whole mesh
// is used.
int orc_distance(const orc_mesh* om, int submeshing, double maximumDistance, int srcFace, const double* srcBary, int nT, const int* tgtFace,
                 const double* tgtBary, double distanceThreshold, double* dist, double* startTan, double* endTan, int* flags) {
    const Mesh& m = om->m;
    std::vector<int> patch;
    int np = -1;
    if (submeshing) {
        double r = maximumDistance;
        if (distanceThreshold < maximumDistance) r = distanceThreshold;
        patch.resize(m.nF);
        np = orc_patch(om, srcFace, srcBary, nT, tgtFace, r, patch.data(), m.nF);
        if (np < 0) return -3;
    }
    int rc = orc_geodesic(om, patch.data(), np, srcFace, srcBary, nT, tgtFace, tgtBary, INF, dist, startTan, endTan, flags, nullptr);
    if (rc) return rc;
    if (submeshing)
        for (int t = 0; t < nT; ++t)
            if (flags[t] & 1) {
                dist[t] = 2.0 * maximumDistance;
                for (int c = 0; c < 3; ++c) startTan[3 * t + c] = endTan[3 * t + c] = (c == 2 ? 1.0 : 0.0);
            }
    return 0;
}

}  // extern "C"
