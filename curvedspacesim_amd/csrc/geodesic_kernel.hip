// Batched exact polyhedral geodesics + pair forces on gfx950.
//
// One workgroup (four or eight 64-wide wavefronts) solves one source particle; two to four workgroups share a CU, so
// that one's waves run while another's wait at a barrier or on memory:
//   1. patch   : the face set of the reference's submesh rule (src/utility/submesher.cpp:55-147), prepared by the
//                patch kernel, is numbered into compact LDS slots through setup-time hash tables (global id -> slot);
//                every face slot caches its corners' vertex slots and its neighbours' face slots, so the propagation
//                touches HBM only for face geometry and coordinates.
//   2. windows : exact window (interval) propagation in FP64 (the method behind CGAL's Surface_mesh_shortest_path,
//                src/utility/meshUtilities.h:44-48), up to 256 work items per step from an LDS ring, one per lane;
//                items that do not fit the ring wait in a per-source HBM spill queue, oldest first, and fill the free
//                lanes of later batches (the result does not depend on the order in which windows are processed).
//                Vertex distance bounds are LDS 64-bit atomic minima on the IEEE bit pattern (distances are
//                non-negative), and a window is dropped when those bounds show it cannot carry a shortest path
//                anywhere on its interval (Xin & Wang's filter). Saddle vertices (angle sum >= 2 pi) and vertices on
//                the patch border are pseudo-sources: each opens a window on the far edge of each incident patch
//                face. The first LDS tables mark an improved vertex pending and fan it out once no window of a batch
//                can lower its distance much further (so that a vertex improved several times fans out about once);
//                the larger tables queue a vertex event per improvement.
//   3. targets : every window is tested against the targets of the face it enters; a target keeps the minimum
//                (atomic on the bit pattern) and, in a second pass, the tangents of the path that set it. The start
//                tangent is carried per window as a rotation from the window's edge frame to the source-face frame,
//                so no path is traced back (reference: computePathDistanceAndTangents,
//                src/utility/meshUtilities.cpp:360-380).
//   4. forces  : the pair force (src/forces/baseForce.cpp:12-28) summed per source with a wavefront reduction.
// A source whose patch, vertex set, target list or spill exceeds the capacities is flagged and solved again by the
// same code with its tables in a per-workgroup HBM scratch slice; that variant also serves the whole-mesh mode of
// triangulatedMeshSpace::distance (src/models/triangulatedMeshSpace.cpp:222-237).
#include <math.h>

#include <algorithm>

#include "csg_common.h"
#include "geodesic_kernel.h"

namespace csg {

namespace {

constexpr double kInf = __builtin_huge_val();
constexpr double kFilterTol = 1e-12;  // relative slack before a window is declared useless
constexpr double kEdgeTol = 1e-12;    // relative distance at which an interval end counts as the edge's endpoint
constexpr int kMaxSteps = 1 << 18;    // FIFO steps per source before it is abandoned as over capacity
constexpr int kMaxWaves = 16;         // wavefronts per workgroup
constexpr int kVertexFace = 1 << 30;  // tag of the work item that opens a pseudo-source window across one face
// Pseudo-source fan-out. Pending marks (the first LDS tables): a pending vertex fans out once its distance is at most
// the batch's smallest key plus this fraction of the truncation radius; the margin trades windows opened from
// distances that are superseded later (small margins wait longer and take more steps) against the steps themselves
// (measured on the bumpy sphere L6, N = 20,000, cutoff 0.15: 0.1 is the fastest, 4% below vertex events; 0.2 and more
// overflow the work queue). Vertex events (the larger LDS tables and the HBM tables, whose scan of the pending marks
// would cover up to 900 vertex slots a step; measured 5% faster there): a work item per improvement, which fans out
// when it comes up if the distance still stands.
constexpr double kReleaseSlack = 0.1;

__device__ inline unsigned long long dbits(double d) { return (unsigned long long)__double_as_longlong(d); }
__device__ inline double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ inline uint32_t hashOf(uint32_t key) { return (key * 2654435761u) >> 7; }

// Tables of one source. LDS variant: every pointer refers to __shared__ storage (the setup hashes and the vertex ids
// share the FIFO's storage, which is idle until the propagation starts); scratch variant: to a slice of a global buffer. Faces and
// vertices are numbered 0, 1, ... in insertion order (compact slots); hash capacities are powers of two.
template <typename IT>
struct Tables {
    static constexpr bool kSlots = true;  // the face hash also maps ids to compact slots
    int32_t* hkey;   // face hash: global face id or -1
    IT* hslot;       // face hash: compact slot of the key
    int32_t* vhkey;  // vertex hash: global vertex id or -1
    IT* vhslot;      // vertex hash: compact slot of the key
    int32_t* fglob;  // per face slot: global face id
    IT* fcv;         // 3 per face slot: corner vertex slots
    IT* fnb;         // 3 per face slot: neighbour face slot * 4 + its local edge index, or -1 outside the patch
    IT* fhead;       // per face slot: first target lying in that face, or -1
    int32_t* thead;  // setup only: the same lists' heads as 32-bit words, built with atomic exchanges
    int32_t* vglob;  // per vertex slot: global vertex id (setup only)
    unsigned long long* vdist;
    float2* vtan;    // start tangent of the best path to the vertex (source-face frame)
    uint8_t* vflg;   // bit 0: may act as a pseudo-source; bit 1 (with pending marks): pending fan-out, set and cleared
                     // with 32-bit atomics on the flag's word (the array is 4-byte aligned, padded to a multiple of 4)
    int32_t* vfo;    // per vertex slot: end of its range in vfl (ranges are consecutive, the first starts at 0)
    IT* vfl;         // patch faces around each vertex: face slot * 4 + the vertex's corner index
    // FIFO ring of work items: windows and vertex events
    double *qb0, *qb1, *qsx, *qsy, *qsg;
    int32_t* qfk;    // window: face slot * 8 + k * 2 + rooted (< 2^30); vertex event: -(vertex slot) - 1;
                     // vertex-face item: kVertexFace | (face slot * 4 + corner of the vertex)
    float2* qt;      // rooted window: rotation (cos, sin) to the source-face frame; otherwise its fixed start tangent
    // targets of the source
    IT* tnext;
    IT* tslot;       // face slot of the target
    double* tq;      // 3 per target: barycentric weights normalised by their sum
    unsigned long long* tbest;
    float2* tst;     // start tangent (source-face frame)
    float4* tet;     // end tangent (xyz); nullptr in variants that do not report it
    int hmask, vhmask;           // hash capacities - 1
    int fcap, vcap, qcap, tcap;  // compact faces, compact vertices, ring records, targets
};

struct Shared {
    int head, tail, nFaces, nVerts, overflow, nCur, nNext, goals, nSteps, nItems;
    int spillHead, spillTail;  // the source's HBM spill queue: records [spillHead, spillTail), ring positions mod spillCap
    int next;                  // the source the workgroup takes next (dynamic distribution)
    int srcSlot;   // face slot of the source face
    unsigned long long minKey;  // bit pattern of the smallest window key of the current batch (all keys are >= 0)
    int nPend;                  // pseudo-source vertices marked pending (flag bit 1) and not yet fanned out
    double S[3], t1[3], t2[3];
    int wsum[kMaxWaves];
};

// Exclusive prefix sum of a[0..n) by the whole workgroup: each lane scans a contiguous chunk, the chunk totals are
// scanned across the wavefront with lane shuffles and across wavefronts through LDS. Ends with a barrier.
template <class IT>
__device__ void blockExclusiveScan(IT* a, int n, Shared& sh) {
    const int nt = blockDim.x, lane = threadIdx.x;
    const int chunk = (n + nt - 1) / nt;
    const int b = min(n, lane * chunk), e = min(n, b + chunk);
    int sum = 0;
    for (int i = b; i < e; ++i) sum += a[i];
    const int wl = lane & 63, w = lane >> 6;
    int x = sum;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (wl >= o) x += y;
    }
    if (wl == 63) sh.wsum[w] = x;
    __syncthreads();
    int run = x - sum;
    for (int k = 0; k < w; ++k) run += sh.wsum[k];
    for (int i = b; i < e; ++i) {
        int c = a[i];
        a[i] = run;
        run += c;
    }
    __syncthreads();
}

// Face hash. Lookups after a barrier that follows the last insertion see every key and slot; during an insertion
// phase only the presence of a key is reliable (faceInsert's own probe).
template <class TT>
__device__ int faceLookup(const TT& T, int g) {
    uint32_t h = hashOf((uint32_t)g) & T.hmask;
    for (int p = 0; p <= T.hmask; ++p) {
        int k = T.hkey[h];
        if (k == g) return (int)T.hslot[h];
        if (k == -1) return -1;
        h = (h + 1) & T.hmask;
    }
    return -1;
}

// inserts g (if absent) and gives it the next compact slot; `fresh` reports whether this call inserted it. Past the
// face capacity the overflow bit is set.
template <class TT>
__device__ void faceInsert(const TT& T, Shared& sh, int g, bool& fresh) {
    fresh = false;
    uint32_t h = hashOf((uint32_t)g) & T.hmask;
    for (int p = 0; p <= T.hmask; ++p) {
        int k = T.hkey[h];
        if (k == g) return;
        if (k == -1) {
            int old = atomicCAS(&T.hkey[h], -1, g);
            if (old == -1) {
                int slot = atomicAdd(&sh.nFaces, 1);
                if (slot >= T.fcap) {
                    atomicOr(&sh.overflow, OVF_FACES);
                    return;
                }
                if constexpr (TT::kSlots) T.hslot[h] = slot;
                T.fglob[slot] = g;
                fresh = true;
                return;
            }
            if (old == g) return;
        }
        h = (h + 1) & T.hmask;
    }
    atomicOr(&sh.overflow, OVF_FACES);
}

// inserts a face known to be absent with a given slot (faces of a prepared list)
template <class TT>
__device__ void faceInsertAt(const TT& T, int g, int slot) {
    uint32_t h = hashOf((uint32_t)g) & T.hmask;
    for (int p = 0; p <= T.hmask; ++p) {
        if (atomicCAS(&T.hkey[h], -1, g) == -1) {
            T.hslot[h] = slot;
            T.fglob[slot] = g;
            return;
        }
        h = (h + 1) & T.hmask;
    }
}

// inserts vertex v (if absent) and gives it the next compact slot; returns its hash entry, whose slot (vhslot) is
// readable after the next barrier, or -1 past the vertex capacity (the overflow bit is then set)
template <class TT>
__device__ int vertInsert(const TT& T, Shared& sh, int v) {
    uint32_t h = hashOf((uint32_t)v) & T.vhmask;
    for (int p = 0; p <= T.vhmask; ++p) {
        int k = T.vhkey[h];
        if (k == v) return (int)h;
        if (k == -1) {
            int old = atomicCAS(&T.vhkey[h], -1, v);
            if (old == -1) {
                int slot = atomicAdd(&sh.nVerts, 1);
                if (slot >= T.vcap) {
                    atomicOr(&sh.overflow, OVF_VERTS);
                    return -1;
                }
                T.vhslot[h] = slot;
                T.vglob[slot] = v;
                return (int)h;
            }
            if (old == v) return (int)h;
        }
        h = (h + 1) & T.vhmask;
    }
    atomicOr(&sh.overflow, OVF_VERTS);
    return -1;
}

template <class TT>
__device__ int vertLookup(const TT& T, int v) {
    uint32_t h = hashOf((uint32_t)v) & T.vhmask;
    for (int p = 0; p <= T.vhmask; ++p) {
        int k = T.vhkey[h];
        if (k == v) return (int)T.vhslot[h];
        if (k == -1) return -1;
        h = (h + 1) & T.vhmask;
    }
    return -1;
}

__device__ inline void loadV(const DeviceMesh& M, int v, double* p) {
    p[0] = M.vx[3 * v];
    p[1] = M.vx[3 * v + 1];
    p[2] = M.vx[3 * v + 2];
}

// start tangent (source-face frame) of the ray direction (rx, ry) of a window. Computed in FP32, the precision it is
// stored in: the unit direction is within a few FP32 ulps of the FP64 one rounded.
__device__ inline float2 startTangent(bool rooted, float2 rot, double rx, double ry) {
    if (!rooted) return rot;
    const float fx = (float)rx, fy = (float)ry;
    const float q = fx * fx + fy * fy;
    if (q == 0.f) return make_float2(0.f, 0.f);
    const float inv = rsqrtf(q);
    const float ux = fx * inv, uy = fy * inv;
    return make_float2(rot.x * ux - rot.y * uy, rot.y * ux + rot.x * uy);
}

struct VCand {  // one pending vertex candidate of a lane, carried from pass 1 to pass 2
    int slot;
    double d;
    float2 t;
    bool improved;
};

__device__ inline VCand noCand() { return VCand{-1, 0.0, make_float2(0.f, 0.f), false}; }

__device__ inline void offerVertexTail(VCand& c, int slot, double d, float2 t, bool imp) {
    if (c.slot == slot) {
        c.improved = c.improved || imp;
    } else {
        c.slot = slot;
        c.improved = imp;
    }
    c.d = d;
    c.t = t;
}

// Offer to a vertex whose distance the lane has read (at the start of the step, lowered since by its own offers), made
// only when d is below that value: the atomic minimum's old value is not awaited, and the offer counts as an
// improvement. If another lane offers the identical distance in the same step, both lanes then take it for theirs
// and push the vertex's event twice, which costs a duplicate fan-out and changes no distance.
template <class TT>
__device__ inline void offerVertexKnown(const TT& T, VCand& c, int slot, double d, float2 t) {
    if (slot < 0) return;
    if (c.slot == slot && !(d < c.d)) return;
    atomicMin(&T.vdist[slot], dbits(d));
    offerVertexTail(c, slot, d, t, true);
}

// Offers distance d (start tangent t) to a vertex slot: atomic minimum, remembered in the lane's candidate c for the
// second pass.
template <class TT>
__device__ inline void offerVertex(const TT& T, VCand& c, int slot, double d, float2 t) {
    if (slot < 0) return;
    if (c.slot == slot && !(d < c.d)) return;
    unsigned long long old = atomicMin(&T.vdist[slot], dbits(d));
    bool imp = d < bitsd(old);
    offerVertexTail(c, slot, d, t, imp);
}


// FP64 quotient and square root for the propagation: the hardware reciprocal (square root) estimate, one Newton step and
// a final correction (Markstein, IBM J. Res. Develop. 34 (1990) 111-119). For operands in the normal range they agree
// with the correctly rounded results (no difference over 4M random operands of 10^-30..10^4 on gfx950); operands near
// the bottom of the range, where the estimate loses accuracy, take the correctly rounded form. The general forms also
// rescale every operand, which costs about half again as many instructions.
__device__ inline double quot(double a, double b) {
    if (fabs(b) < 1e-280) return a / b;
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}
__device__ inline double root(double v) {
    if (!(v > 1e-280)) return sqrt(v);  // also 0, a negative value or NaN
    double y = __builtin_amdgcn_rsq(v);
    y = fma(fma(-v * y, 0.5 * y, 0.5), y, y);
    const double s = v * y;
    return fma(fma(-s, s, v), 0.5 * y, s);
}

// distance from the window's source (sx, sy) to the point b of its edge
__device__ inline double endDist(double b, double sx, double sy) { return root((b - sx) * (b - sx) + sy * sy); }

// A lower bound of endDist from an FP32 square root: the FP32 rounding of the square and of its root stay far below
// kLowRel, so endDist >= endDistLow. The pruning tests (rMax and the edge-end filter) use it, so they can only keep
// windows that exact tests would drop, never drop one they would keep; such a window is useless, and propagating it
// changes no distance.
constexpr double kLowRel = 1e-6;
__device__ inline double endDistLow(double b, double sx, double sy) {
    const double q = (b - sx) * (b - sx) + sy * sy;
    return (double)sqrtf((float)q) * (1.0 - kLowRel);
}

// lower bound of the distances the window carries: its source distance plus the distance to the nearest interval point
// (r0, r1: endDist of b0 and b1, or lower bounds of them)
__device__ inline double keyOf(double b0, double b1, double sx, double sy, double sg, double r0, double r1) {
    const double dm = sx < b0 ? r0 : (sx > b1 ? r1 : -sy);
    return sg + dm;
}

// Xin-Wang filter against the edge's end vertices: the window is useless when one of its interval ends is reached
// more cheaply through the vertex at the other end of the edge
__device__ inline bool filteredBy(double dA, double dB, double L, double b0, double b1, double sg, double r0, double r1) {
    double toI1 = sg + r1;
    double toI0 = sg + r0;
    if (toI1 > dA + b1 + kFilterTol * toI1) return true;
    if (toI0 > dB + (L - b0) + kFilterTol * toI0) return true;
    return false;
}

template <class TT>
__device__ inline bool filtered(const TT& T, int vA, int vB, double L, double b0, double b1, double sg, double r0, double r1) {
    return filteredBy(bitsd(T.vdist[vA]), bitsd(T.vdist[vB]), L, b0, b1, sg, r0, r1);
}

template <class TT>
struct Ctx {
    const GeoParams& P;
    const TT& T;
    Shared& sh;
    int qBase;  // head at the start of the step: appends stay below qBase + qcap, so live records are never overwritten
    double rMax;
    int tb;     // first pair index of the source
    SpillRecord* spill;  // this source's HBM spill queue (nullptr: none)
};

// Appends one work item: to the LDS ring while it has room, else to the source's HBM spill queue. A ring position at
// or past qBase + qcap is handed back at once, so the positions below it are handed out exactly once.
template <class TT>
__device__ void pushItem(const Ctx<TT>& C, double b0, double b1, double sx, double sy, double sg, int32_t fk, float2 t) {
    const TT& T = C.T;
    int pos = atomicAdd(&C.sh.tail, 1);
    if (pos - C.qBase < T.qcap) {
        int q = pos % T.qcap;
        T.qb0[q] = b0;
        T.qb1[q] = b1;
        T.qsx[q] = sx;
        T.qsy[q] = sy;
        T.qsg[q] = sg;
        T.qfk[q] = fk;
        T.qt[q] = t;
        return;
    }
    atomicSub(&C.sh.tail, 1);
    if (C.spill) {
        // spillHead only moves between full barriers, never while items are appended
        int sp = atomicAdd(&C.sh.spillTail, 1);
        if (sp - C.sh.spillHead >= C.P.spillCap) {
            atomicSub(&C.sh.spillTail, 1);  // as for the ring: positions past the capacity are handed back
        } else {
            SpillRecord r;
            r.b0 = b0;
            r.b1 = b1;
            r.sx = sx;
            r.sy = sy;
            r.sg = sg;
            r.fk = fk;
            r.tx = t.x;
            r.ty = t.y;
            r.pad = 0;
            C.spill[sp % C.P.spillCap] = r;
            return;
        }
    }
    atomicOr(&C.sh.overflow, OVF_QUEUE);
}

// A new window on (face slot fs, local edge k): distance candidates for reached edge endpoints, filter, append.
template <class TT>
__device__ void addWindow(const Ctx<TT>& C, int fs, int k, double L, double b0, double b1, double sx, double sy, double sg, bool rooted,
                          float2 rot, VCand& cA, VCand& cB) {
    if (!(b1 > b0)) return;
    const double r0 = endDistLow(b0, sx, sy), r1 = endDistLow(b1, sx, sy);  // lower bounds: pruning tests only
    if (keyOf(b0, b1, sx, sy, sg, r0, r1) > C.rMax) return;
    const TT& T = C.T;
    int vA = T.fcv[3 * fs + k], vB = T.fcv[3 * fs + (k + 1) % 3];
    double tol = kEdgeTol * L;
    // an interval end on an edge end: the distance to that vertex
    if (b0 <= tol) offerVertex(T, cA, vA, sg + endDist(0.0, sx, sy), startTangent(rooted, rot, -sx, -sy));
    if (b1 >= L - tol) offerVertex(T, cB, vB, sg + endDist(L, sx, sy), startTangent(rooted, rot, L - sx, -sy));
    if (filtered(T, vA, vB, L, b0, b1, sg, r0, r1)) return;
    pushItem(C, b0, b1, sx, sy, sg, fs * 8 + k * 2 + (rooted ? 1 : 0), rot);
}

// A child window of a window being propagated: as addWindow, with its edge's end vertices (slots vA, vB) and the
// distances the lane last saw for them (dA, dB; lowered here by the lane's own offers) given by the caller. Any
// distances that a path reached are valid bounds for the filter.
template <class TT>
__device__ void addChild(const Ctx<TT>& C, int fs, int k, double L, double b0, double b1, double sx, double sy, double sg, bool rooted,
                         float2 rot, int vA, double& dA, int vB, double& dB, VCand& cA, VCand& cB) {
    if (!(b1 > b0)) return;
    const double r0 = endDistLow(b0, sx, sy), r1 = endDistLow(b1, sx, sy);  // lower bounds: pruning tests only
    if (keyOf(b0, b1, sx, sy, sg, r0, r1) > C.rMax) return;
    const TT& T = C.T;
    double tol = kEdgeTol * L;
    if (b0 <= tol) {
        const double d = sg + endDist(0.0, sx, sy);
        if (d < dA) {
            offerVertexKnown(T, cA, vA, d, startTangent(rooted, rot, -sx, -sy));
            dA = d;
        }
    }
    if (b1 >= L - tol) {
        const double d = sg + endDist(L, sx, sy);
        if (d < dB) {
            offerVertexKnown(T, cB, vB, d, startTangent(rooted, rot, L - sx, -sy));
            dB = d;
        }
    }
    if (filteredBy(dA, dB, L, b0, b1, sg, r0, r1)) return;
    pushItem(C, b0, b1, sx, sy, sg, fs * 8 + k * 2 + (rooted ? 1 : 0), rot);
}

template <class TT>
__device__ inline void pushVertexEvent(const Ctx<TT>& C, int vslot, double d) {
    pushItem(C, 0.0, 0.0, 0.0, 0.0, d, -vslot - 1, make_float2(0.f, 0.f));
}

// Marks a pseudo-source vertex whose distance improved as pending: its fan-out waits until no window of a batch can
// lower the distance further (propagateAll). The bit is set once (the 32-bit atomic's old value), so nPend counts
// each pending vertex once.
template <class TT>
__device__ inline void markPending(const TT& T, Shared& sh, int v) {
    const unsigned bit = 2u << (8 * (v & 3));
    const unsigned old = atomicOr(reinterpret_cast<unsigned*>(T.vflg) + (v >> 2), bit);
    if (!(old & bit)) atomicAdd(&sh.nPend, 1);
}

template <class TT>
__device__ inline void pushVertexFace(const Ctx<TT>& C, int gs, int j, double d, float2 tv) {
    pushItem(C, 0.0, 0.0, 0.0, 0.0, d, kVertexFace | (gs * 4 + j), tv);
}

// pseudo-source point Q (global coordinates) in the frame of the window entering face g through its edge k
__device__ inline void pointInFrame(const DeviceMesh& M, int g, int k, const double* Q, double& sx, double& sy, double& L) {
    double o[3], e[3];
    loadV(M, M.fv[3 * g + k], o);
    loadV(M, M.fv[3 * g + (k + 1) % 3], e);
    L = M.fgeo[9 * g + 3 * k];
    double ex[3] = {(e[0] - o[0]) / L, (e[1] - o[1]) / L, (e[2] - o[2]) / L};
    double r[3] = {Q[0] - o[0], Q[1] - o[1], Q[2] - o[2]};
    sx = r[0] * ex[0] + r[1] * ex[1] + r[2] * ex[2];
    double c0 = r[1] * ex[2] - r[2] * ex[1], c1 = r[2] * ex[0] - r[0] * ex[2], c2 = r[0] * ex[1] - r[1] * ex[0];
    sy = -sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    if (sy > -1e-14 * L) sy = -1e-14 * L;
}

// One window against the targets of the face it enters. Pass 1: atomic minimum. Pass 2: tangents of the winner.
template <bool kPass2, class TT>
__device__ void windowTargets(const Ctx<TT>& C, int fs, int g, int k, double L, double x2, double y2, double b0, double b1, double sx,
                              double sy, double sg, double r0, double r1, bool rooted, float2 rot) {
    const TT& T = C.T;
    const GeoParams& P = C.P;
    for (int t = T.fhead[fs]; t >= 0; t = T.tnext[t]) {
        const double* b = T.tq + 3 * t;
        double bk1 = b[(k + 1) % 3], bk2 = b[(k + 2) % 3];
        double qx = bk1 * L + bk2 * x2, qy = bk2 * y2;
        double rx = qx - sx, ry = qy - sy;
        double ax = b0 - sx, ay = -sy, bx = b1 - sx, by = -sy;
        double rl = root(rx * rx + ry * ry);
        double c1 = bx * ry - by * rx;  // cross(interval end ray, target ray)
        double c2 = rx * ay - ry * ax;  // cross(target ray, interval start ray)
        double tol = 1e-12 * rl;
        if (c1 < -tol * r1 || c2 < -tol * r0) continue;  // r1 = |(bx, by)|, r0 = |(ax, ay)|
        double d = sg + rl;
        if (!kPass2) {
            atomicMin(&T.tbest[t], dbits(d));
        } else if (T.tbest[t] == dbits(d)) {
            T.tst[t] = startTangent(rooted, rot, rx, ry);
            if (!P.pairEnd) continue;
            const DeviceMesh& M = P.mesh;
            double o[3], e[3];
            loadV(M, M.fv[3 * g + k], o);
            loadV(M, M.fv[3 * g + (k + 1) % 3], e);
            double ex[3] = {(e[0] - o[0]) / L, (e[1] - o[1]) / L, (e[2] - o[2]) / L};
            const double* n = M.fnrm + 3 * g;
            double ey[3] = {n[1] * ex[2] - n[2] * ex[1], n[2] * ex[0] - n[0] * ex[2], n[0] * ex[1] - n[1] * ex[0]};
            T.tet[t] = make_float4((float)((rx * ex[0] + ry * ey[0]) / rl), (float)((rx * ex[1] + ry * ey[1]) / rl),
                                   (float)((rx * ex[2] + ry * ey[2]) / rl), 0.f);
        }
    }
}

// Window propagation across the face it enters (pass 1): candidate for the opposite corner, up to two children.
// The corners of the face a window enters, in the order k, k+1, k+2 of its entry edge k: vertex slots, the distances
// the lane read for them, and the neighbours across edges k+2 and k+1 (the edges the children leave by). A consistently
// oriented mesh runs the shared edge the other way in the neighbour, so the left child's edge goes from corner k to
// corner k+2 and the right child's from corner k+2 to corner k+1.
struct Corners {
    int vA, vB, vC, nbL, nbR;
    double dA, dB, dC;
};

// Window propagation across the face it enters (pass 1): candidate for the opposite corner, up to two children.
template <class TT>
__device__ void propagate(const Ctx<TT>& C, int fk, int g, double L, double x2, double y2, double b0, double b1, double sx, double sy,
                          double sg, double r0, double r1, const double* cf, float2 rot, Corners& K, VCand& cP0, VCand& cP1,
                          VCand& cP2) {
    const TT& T = C.T;
    int fs = fk >> 3, k = (fk >> 1) & 3;
    bool rooted = fk & 1;
    windowTargets<false, TT>(C, fs, g, k, L, x2, y2, b0, b1, sx, sy, sg, r0, r1, rooted, rot);
    double xs = sx + quot((x2 - sx) * (-sy), y2 - sy);  // where the ray through the opposite corner crosses the entry edge
    double tolE = kEdgeTol * L;
    if (xs >= b0 - tolE && xs <= b1 + tolE) {
        const double d = sg + root((x2 - sx) * (x2 - sx) + (y2 - sy) * (y2 - sy));
        if (d < K.dC) {
            offerVertexKnown(T, cP2, K.vC, d, startTangent(rooted, rot, x2 - sx, y2 - sy));
            K.dC = d;
        }
    }
    if (xs > b0) {  // left child: edge corner k -> corner k+2, entered from the neighbour across local edge k+2
        int nb = K.nbL;
        if (nb >= 0) {
            const double lL = cf[0], ux = cf[1], uy = cf[2];
            double spx = sx * ux + sy * uy, spy = ux * sy - uy * sx;
            auto hit = [&](double bx) {
                double dx = bx - sx, dy = -sy;
                double dpx = dx * ux + dy * uy, dpy = ux * dy - uy * dx;
                if (dpy <= 0) return dpx < 0 ? 0.0 : lL;
                return fmin(lL, fmax(0.0, spx + quot(dpx * (-spy), dpy)));
            };
            double t0 = (b0 <= 0) ? 0.0 : hit(b0);
            double t1 = (xs <= b1) ? lL : hit(b1);
            float2 r2 = rooted ? make_float2((float)(rot.x * ux - rot.y * uy), (float)(rot.y * ux + rot.x * uy)) : rot;
            addChild(C, nb >> 2, nb & 3, lL, t0, t1, spx, fmin(spy, -1e-14 * lL), sg, rooted, r2, K.vA, K.dA, K.vC, K.dC, cP0, cP2);
        }
    }
    if (xs < b1) {  // right child: edge corner k+2 -> corner k+1, entered from the neighbour across local edge k+1
        int nb = K.nbR;
        if (nb >= 0) {
            const double lR = cf[3], ux = cf[4], uy = cf[5];
            double rx0 = sx - x2, ry0 = sy - y2;
            double spx = rx0 * ux + ry0 * uy, spy = ux * ry0 - uy * rx0;
            auto hit = [&](double bx) {
                double dx = bx - sx, dy = -sy;
                double dpx = dx * ux + dy * uy, dpy = ux * dy - uy * dx;
                if (dpy <= 0) return dpx < 0 ? 0.0 : lR;
                return fmin(lR, fmax(0.0, spx + quot(dpx * (-spy), dpy)));
            };
            double t0 = (xs >= b0) ? 0.0 : hit(b0);
            double t1 = (b1 >= L) ? lR : hit(b1);
            float2 r2 = rooted ? make_float2((float)(rot.x * ux - rot.y * uy), (float)(rot.y * ux + rot.x * uy)) : rot;
            addChild(C, nb >> 2, nb & 3, lR, t0, t1, spx, fmin(spy, -1e-14 * lR), sg, rooted, r2, K.vC, K.dC, K.vB, K.dB, cP2, cP1);
        }
    }
}

// A pseudo-source vertex event fans out into one work item per incident patch face; each item opens the window on the
// face's far edge. The vertex's unfolded position in that window's frame comes from the face's precomputed frame of
// that edge (the corner opposite it), so the item needs a single geometry load.
template <class TT>
__device__ void fanOutVertex(const Ctx<TT>& C, int vslot, double d) {
    const TT& T = C.T;
    float2 tv = T.vtan[vslot];
    const int q0 = vslot > 0 ? T.vfo[vslot - 1] : 0, q1 = T.vfo[vslot];
    for (int q = q0; q < q1; ++q) {
        int e = T.vfl[q];
        int gs = e >> 2, j = e & 3;
        if (T.fnb[3 * gs + (j + 1) % 3] >= 0) pushVertexFace(C, gs, j, d, tv);
    }
}

// Pass 2 for a lane's three vertex candidates (distinct corners of one face): a candidate whose offer still stands
// sets the vertex's tangent and, for a pseudo-source vertex, marks it pending or pushes its event. The six LDS reads
// are issued together (slots clamped to 0 for empty candidates), so they cost one round trip.
template <bool kPending, class TT>
__device__ inline void finishVertexCands(const Ctx<TT>& C, const VCand& c0, const VCand& c1, const VCand& c2) {
    const TT& T = C.T;
    const int s0 = c0.slot < 0 ? 0 : c0.slot, s1 = c1.slot < 0 ? 0 : c1.slot, s2 = c2.slot < 0 ? 0 : c2.slot;
    const unsigned long long d0 = T.vdist[s0], d1 = T.vdist[s1], d2 = T.vdist[s2];
    const uint8_t f0 = T.vflg[s0], f1 = T.vflg[s1], f2 = T.vflg[s2];
    auto finish = [&](const VCand& c, unsigned long long d, uint8_t f) {
        if (c.slot < 0 || !c.improved || d != dbits(c.d)) return;
        T.vtan[c.slot] = c.t;
        if ((f & 1) && c.d <= C.rMax) {
            if constexpr (kPending)
                markPending(T, C.sh, c.slot);
            else
                pushVertexEvent(C, c.slot, c.d);
        }
    };
    finish(c0, d0, f0);
    finish(c1, d1, f1);
    finish(c2, d2, f2);
}

__device__ inline double wsum(double x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Workgroup barrier that orders LDS accesses only: outstanding global loads and stores stay in flight across it
// (a full barrier would wait for them). Global data written before it is not guaranteed visible after it.
__device__ inline void ldsBarrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The overflow bits as every wave of the workgroup sees them: read between two barriers, so that no wave can set a
// bit (in the phase that follows) before the slowest wave has read them, and all waves take the same branch.
__device__ inline bool uniformOverflow(const Shared& sh) {
    __syncthreads();
    const bool b = sh.overflow != 0;
    __syncthreads();
    return b;
}

// Builds the patch and the per-slot tables. Returns false (with sh.overflow set) when a capacity is exceeded.
// source point (construct_point semantics: weights normalised by their sum) and the source face's tangent frame
__device__ void initSourceFrame(const GeoParams& P, Shared& sh, int pi, int sf) {
    const DeviceMesh& M = P.mesh;
    const double* bb = P.pbary + 3 * pi;
    double s = bb[0] + bb[1] + bb[2];
    double p0[3], p1[3], p2[3];
    loadV(M, M.fv[3 * sf], p0);
    loadV(M, M.fv[3 * sf + 1], p1);
    loadV(M, M.fv[3 * sf + 2], p2);
    for (int a = 0; a < 3; ++a) sh.S[a] = (p0[a] * bb[0] + p1[a] * bb[1] + p2[a] * bb[2]) / s;
    double e[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double l = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    for (int a = 0; a < 3; ++a) sh.t1[a] = e[a] / l;
    const double* n = M.fnrm + 3 * sf;
    sh.t2[0] = n[1] * sh.t1[2] - n[2] * sh.t1[1];
    sh.t2[1] = n[2] * sh.t1[0] - n[0] * sh.t1[2];
    sh.t2[2] = n[0] * sh.t1[1] - n[1] * sh.t1[0];
}

// Whether point p lies within the radius of the source point S (squared distance not greater than r2). The sum is
// written with explicit FMAs so that every search that applies the rule decides a corner the same way.
__device__ inline bool withinRadius(const double* p, const double* S, double r2) {
    const double dx = p[0] - S[0], dy = p[1] - S[1], dz = p[2] - S[2];
    return !(fma(dz, dz, fma(dy, dy, dx * dx)) > r2);
}

// What the submesh rule needs of a face f: its adjacency words, its corners (c[0..9)) and, per local edge k, the far
// corner of the neighbour across it (c[9 + 3k ..]), all from its 160-byte record in ten 16-byte loads. The neighbour
// across edge k has f's corners k and k+1 and that far corner, so whether it has a corner within the radius is decided
// from f's record alone.
struct RuleFace {
    int adj[3];
    double c[18];
};

__device__ inline void loadRuleFace(const DeviceMesh& M, int f, RuleFace& F) {
    const int4* rec = reinterpret_cast<const int4*>(M.frule + 20 * (size_t)f);
    const int4 w = rec[0];
    F.adj[0] = w.x;
    F.adj[1] = w.y;
    F.adj[2] = w.z;
    const double2* d = reinterpret_cast<const double2*>(rec + 1);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const double2 v = d[i];
        F.c[2 * i] = v.x;
        F.c[2 * i + 1] = v.y;
    }
}

// bit k: the neighbour across edge k of the face has a corner within the radius
__device__ inline int ruleBits(const RuleFace& F, const double* S, double r2) {
    bool in[3], inOpp[3];
    for (int k = 0; k < 3; ++k) {
        in[k] = withinRadius(F.c + 3 * k, S, r2);
        inOpp[k] = withinRadius(F.c + 9 + 3 * k, S, r2);
    }
    int bits = 0;
    for (int k = 0; k < 3; ++k)
        if (in[k] || in[k == 2 ? 0 : k + 1] || inOpp[k]) bits |= 1 << k;
    return bits;
}

// Whether the source's targets lie off its face (bit 0) and beyond its edge neighbours (bit 1), or'ed into sh.goals;
// nbr: the neighbour faces of the source face (-2 across a border edge)
__device__ inline void patchGoals(const GeoParams& P, Shared& sh, const int* nbr, int sf, int tb, int nT) {
    int anyGoal = 0, beyond = 0;
    for (int t = threadIdx.x; t < nT; t += blockDim.x) {
        int f = P.pface[P.tidx[tb + t]];
        if (f != sf) anyGoal = 1;
        if (f != sf && f != nbr[0] && f != nbr[1] && f != nbr[2]) beyond = 1;
    }
    if (anyGoal) atomicOr(&sh.goals, 1);
    if (beyond) atomicOr(&sh.goals, 2);
}

// The reference's submesh rule (src/utility/submesher.cpp:55-147): the source face; when some target lies elsewhere, its
// edge neighbours; when some target lies beyond those, a breadth-first search over faces with at least one corner
// within the radius of the source point (squared distance not greater than r^2) - the set reached equals the
// reference's depth-first set, since it is the closure of the same rule; finally every target face.
template <class TT>
__device__ void patchSearch(const GeoParams& P, const TT& T, Shared& sh, int* cur, int* nxt, int frontCap, int src, int sf, int tb,
                            int nT) {
    const int lane = threadIdx.x;
    const DeviceMesh& M = P.mesh;
    const double r = P.patchRadius[src];
    const double r2 = r * r;
    int nbr[3];
    for (int k = 0; k < 3; ++k) nbr[k] = M.fadj[3 * sf + k] < 0 ? -2 : (M.fadj[3 * sf + k] >> 2);
    patchGoals(P, sh, nbr, sf, tb, nT);
    __syncthreads();
    if (lane == 0) {
        bool fresh;
        faceInsert(T, sh, sf, fresh);
        if (sh.goals & 1)
            for (int k = 0; k < 3; ++k)
                if (nbr[k] >= 0) {
                    faceInsert(T, sh, nbr[k], fresh);
                    if (fresh) cur[sh.nCur++] = nbr[k];
                }
        if (sh.overflow) sh.nCur = 0;
    }
    __syncthreads();
    if (sh.goals & 2) {
        // breadth-first: the set reached equals the reference's depth-first set (order-independent closure)
        while (true) {
            // nCur is written only between barriers, so every wave takes the same branch
            const int nc = sh.nCur;
            if (nc == 0) break;
            for (int i = lane; i < nc; i += blockDim.x) {
                const int f = cur[i];
                RuleFace F;
                loadRuleFace(M, f, F);
                const int bits = ruleBits(F, sh.S, r2);
                for (int k = 0; k < 3; ++k) {
                    int a = F.adj[k];
                    if (a < 0) continue;
                    int g = a >> 2;
                    if (!((bits >> k) & 1)) continue;  // every corner outside the radius
                    bool fresh;  // (faceInsert finds a face already entered without a separate lookup)
                    faceInsert(T, sh, g, fresh);
                    if (fresh) {
                        int pos = atomicAdd(&sh.nNext, 1);
                        if (pos < frontCap)
                            nxt[pos] = g;
                        else
                            atomicOr(&sh.overflow, OVF_FRONTIER);
                    }
                }
            }
            __syncthreads();
            // the new level's list becomes the frontier: the two lists swap roles (the same on every lane)
            const int nn = min(sh.nNext, frontCap);
            int* const done = cur;
            cur = nxt;
            nxt = done;
            __syncthreads();  // every lane has read nNext before it is reset
            if (lane == 0) {
                sh.nCur = sh.overflow ? 0 : nn;
                sh.nNext = 0;
            }
            __syncthreads();
        }
    }
    // target faces not reached above (possibly disconnected from the source)
    for (int t = lane; t < nT; t += blockDim.x) {
        bool fresh;
        faceInsert(T, sh, P.pface[P.tidx[tb + t]], fresh);
    }
}

// Builds the patch and the per-slot tables. Returns false (with sh.overflow set) when a capacity is exceeded. cur/nxt
// are breadth-first frontier lists for a source without a prepared face list (nullptr: such a source overflows).
// kIt > 0: a prepared list holds at most kIt faces per lane, whose corner ids and neighbour words are loaded once, all
// in flight together, and kept in registers from the face set (A) to the per-slot tables (C).
template <int kIt, class TT>
__device__ bool buildPatch(const GeoParams& P, const TT& T, Shared& sh, int* cur, int* nxt, int frontCap, int src, int sf, int tb,
                           int nT) {
    const int lane = threadIdx.x;
    const int nt = blockDim.x;
    const DeviceMesh& M = P.mesh;
    constexpr int kR = kIt > 0 ? kIt : 1;
    int cv[kR][3], ca[kR][3];
    bool cached = false;
    // targets first, independent of the patch: weights normalised by their sum, and the global face id held in the
    // best-distance word until phase C turns it into a face slot (so these loads overlap the face set's)
    if (nT <= T.tcap)
        for (int t = lane; t < nT; t += blockDim.x) {
            const int pj = P.tidx[tb + t];
            const double* b = P.pbary + 3 * pj;
            const double sb = b[0] + b[1] + b[2];
            for (int a = 0; a < 3; ++a) T.tq[3 * t + a] = b[a] / sb;
            T.tbest[t] = (unsigned long long)(uint32_t)P.pface[pj];
        }
    // A. the face set (with a prepared list, also B. their corner vertices)
    bool vertsDone = false;
    const int listed = (P.patchRadius && P.patchList) ? P.patchCount[src] : -1;
    if (P.patchRadius == nullptr) {
        for (int g = lane; g < M.nF; g += blockDim.x) {
            bool fresh;
            faceInsert(T, sh, g, fresh);
        }
    } else if (listed >= 0) {  // prepared by the patch kernel (distinct faces): slot i is the list's i-th face
        const int32_t* lst = P.patchList + (size_t)src * P.patchStride;
        if (listed > T.fcap) {
            if (lane == 0) atomicOr(&sh.overflow, OVF_FACES);
        } else {
            if constexpr (kIt > 0) {
                int gl[kR];
#pragma unroll
                for (int j = 0; j < kR; ++j) {
                    const int i = lane + j * nt;
                    gl[j] = i < listed ? lst[i] : -1;
                }
#pragma unroll
                for (int j = 0; j < kR; ++j)
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        cv[j][k] = gl[j] >= 0 ? M.fv[3 * gl[j] + k] : -1;
                        ca[j][k] = gl[j] >= 0 ? M.fadj[3 * gl[j] + k] : -1;
                    }
#pragma unroll
                for (int j = 0; j < kR; ++j)
                    if (gl[j] >= 0) {
                        faceInsertAt(T, gl[j], lane + j * nt);
                        for (int k = 0; k < 3; ++k) cv[j][k] = vertInsert(T, sh, cv[j][k]);  // ids become hash entries
                    }
                cached = true;
            } else {
                for (int i = lane; i < listed; i += nt) {
                    const int g = lst[i];
                    faceInsertAt(T, g, i);
                    for (int k = 0; k < 3; ++k) vertInsert(T, sh, M.fv[3 * g + k]);
                }
            }
            if (lane == 0) sh.nFaces = listed;
            vertsDone = true;
        }
    } else if (cur) {
        patchSearch(P, T, sh, cur, nxt, frontCap, src, sf, tb, nT);
    } else if (lane == 0) {
        atomicOr(&sh.overflow, OVF_FACES);
    }
    if (uniformOverflow(sh)) return false;
    const int nF = sh.nFaces;
    if (!vertsDone) {  // B. their corner vertices
        for (int s = lane; s < nF; s += blockDim.x) {
            const int g = T.fglob[s];
            for (int k = 0; k < 3; ++k) vertInsert(T, sh, M.fv[3 * g + k]);
        }
        if (uniformOverflow(sh)) return false;
    }
    const int nV = sh.nVerts;
    // C. per face slot: corner and neighbour slots, corner counts per vertex; per target: face slot and weights.
    // Corners are given as vertex ids, or (entries) as the hash entries their insertion returned.
    auto faceSlotTables = [&](int s, const int* fvg, const int* adj, bool entries) {
        for (int k = 0; k < 3; ++k) {
            const int v = entries ? (int)T.vhslot[fvg[k]] : vertLookup(T, fvg[k]);
            T.fcv[3 * s + k] = v;
            atomicAdd(&T.vfo[v], 1);
            const int a = adj[k];
            const int ns = a < 0 ? -1 : faceLookup(T, a >> 2);
            T.fnb[3 * s + k] = ns < 0 ? -1 : ns * 4 + (a & 3);
        }
        T.thead[s] = -1;
    };
    if (cached) {
#pragma unroll
        for (int j = 0; j < kR; ++j)
            if (lane + j * nt < nF) faceSlotTables(lane + j * nt, cv[j], ca[j], true);
    } else {
        for (int s = lane; s < nF; s += nt) {
            const int g = T.fglob[s];
            faceSlotTables(s, M.fv + 3 * g, M.fadj + 3 * g, false);
        }
    }
    for (int t = lane; t < nT; t += blockDim.x) {
        T.tslot[t] = faceLookup(T, (int)T.tbest[t]);
        T.tbest[t] = dbits(kInf);
    }
    if (lane == 0) {
        sh.srcSlot = faceLookup(T, sf);
        if (sh.srcSlot < 0) atomicOr(&sh.overflow, OVF_FACES);  // the patch rule always holds the source face
    }
    if (uniformOverflow(sh)) return false;
    // D. pseudo-source candidates (saddle and mesh-border vertices, and vertices with an incident face outside the
    // patch) and the vertices' initial state; then the vertex -> face ranges
    for (int s = lane; s < nV; s += blockDim.x) {
        const int v = T.vglob[s];
        T.vflg[s] = ((M.vflag[v] & 3) || (int)T.vfo[s] < M.vfoff[v + 1] - M.vfoff[v]) ? 1 : 0;
        T.vdist[s] = dbits(kInf);
        T.vtan[s] = make_float2(0.f, 0.f);
    }
    blockExclusiveScan(T.vfo, nV, sh);
    // E. the vertex -> face lists; each target pushed onto its face's list (the order within a list does not matter:
    // every window tests all targets of its face)
    for (int s = lane; s < nF; s += blockDim.x)
        for (int k = 0; k < 3; ++k) T.vfl[atomicAdd(&T.vfo[T.fcv[3 * s + k]], 1)] = s * 4 + k;
    for (int t = lane; t < nT; t += blockDim.x) {
        const int s = T.tslot[t];
        if (s >= 0 && (!P.keep || P.keep[tb + t])) T.tnext[t] = atomicExch(&T.thead[s], t);  // screened-out pairs stay unlinked
    }
    __syncthreads();
    for (int s = lane; s < nF; s += blockDim.x) T.fhead[s] = T.thead[s];
    __syncthreads();
    return true;
}

// Corner k of the source face, and the neighbour across its edge k with the end points and length of that edge in the
// neighbour's numbering: what the first windows need, loaded before the tables are built so that the loads overlap it.
struct SourceEdge {
    double c[3];
    double o[3], e[3], L;
    int h, kh;  // neighbour face and its local index of the shared edge; h < 0 on a mesh border
};

__device__ inline SourceEdge loadSourceEdge(const DeviceMesh& M, int sf, int k) {
    SourceEdge r{};
    loadV(M, M.fv[3 * sf + k], r.c);
    const int a = M.fadj[3 * sf + k];
    r.h = a < 0 ? -1 : a >> 2;
    r.kh = a & 3;
    if (r.h >= 0) {
        loadV(M, M.fv[3 * r.h + r.kh], r.o);
        loadV(M, M.fv[3 * r.h + (r.kh + 1) % 3], r.e);
        r.L = M.fgeo[9 * r.h + 3 * r.kh];
    }
    return r;
}

// lane that holds corner/edge k of the source face (SourceEdge): lanes 64-66, a wavefront other than the one that sets
// up the source frame
__device__ inline int sourceEdgeOf(int lane) { return lane - (blockDim.x >= 128 ? 64 : 0); }

template <bool kPending, class TT>
__device__ void propagateAll(const GeoParams& P, const TT& T, Shared& sh, int sf, int tb, SpillRecord* spill, double rMax, double slack,
                             const SourceEdge& se) {
    const int lane = threadIdx.x;
    const int ek = sourceEdgeOf(lane);
    const DeviceMesh& M = P.mesh;
    Ctx<TT> C{P, T, sh, 0, rMax, tb, spill};
    const int ss = sh.srcSlot;
    // corners of the source face are visible directly
    if (ek >= 0 && ek < 3) {
        const double* p = se.c;
        double r[3] = {p[0] - sh.S[0], p[1] - sh.S[1], p[2] - sh.S[2]};
        double l = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        int vs = T.fcv[3 * ss + ek];
        T.vdist[vs] = dbits(l);  // three distinct corners, one lane each
        T.vtan[vs] = l > 0 ? make_float2((float)((r[0] * sh.t1[0] + r[1] * sh.t1[1] + r[2] * sh.t1[2]) / l),
                                         (float)((r[0] * sh.t2[0] + r[1] * sh.t2[1] + r[2] * sh.t2[2]) / l))
                           : make_float2(0.f, 0.f);
    }
    __syncthreads();
    if (ek >= 0 && ek < 3) {
        int nb = T.fnb[3 * ss + ek];
        if (nb >= 0) {  // the neighbour se.h lies in the patch
            const int hs = nb >> 2, kh = nb & 3;
            const double L = se.L;
            const double *o = se.o, *e = se.e;
            // the source point in the frame of edge kh of the neighbour (pointInFrame)
            double ex[3] = {(e[0] - o[0]) / L, (e[1] - o[1]) / L, (e[2] - o[2]) / L};
            const double rr[3] = {sh.S[0] - o[0], sh.S[1] - o[1], sh.S[2] - o[2]};
            const double sx = rr[0] * ex[0] + rr[1] * ex[1] + rr[2] * ex[2];
            const double c0 = rr[1] * ex[2] - rr[2] * ex[1], c1 = rr[2] * ex[0] - rr[0] * ex[2], c2 = rr[0] * ex[1] - rr[1] * ex[0];
            double sy = -sqrt(c0 * c0 + c1 * c1 + c2 * c2);
            if (sy > -1e-14 * L) sy = -1e-14 * L;
            float2 rot = make_float2((float)(ex[0] * sh.t1[0] + ex[1] * sh.t1[1] + ex[2] * sh.t1[2]),
                                     (float)(ex[0] * sh.t2[0] + ex[1] * sh.t2[1] + ex[2] * sh.t2[2]));
            VCand a = noCand(), b = noCand();
            addWindow(C, hs, kh, L, 0.0, L, sx, sy, 0.0, true, rot, a, b);
        }
    }
    __syncthreads();
    // the windows above may lower a corner's distance by rounding (it is recomputed in their edge frame), so the
    // pseudo-source events of the corners carry the final values; an event with a superseded value would be skipped
    if (lane < 3) {
        int vs = T.fcv[3 * ss + lane];
        double d = bitsd(T.vdist[vs]);
        if ((T.vflg[vs] & 1) && d <= C.rMax) {
            if constexpr (kPending)
                markPending(T, sh, vs);
            else
                pushVertexEvent(C, vs, d);
        }
    }
    __syncthreads();
    int h = 0;  // every lane advances its own copy of the head
    for (int step = 0;; ++step) {
        const int t = sh.tail;
        const int s0 = sh.spillHead;
        const int sp = sh.spillTail - s0;  // spilled records waiting
        const bool ovf = sh.overflow != 0;
        const int np = sh.nPend;
        if (lane == 0) sh.minKey = ~0ull;  // last read in the previous step's first pass
        ldsBarrier();  // every wave has read the tail and the flags before any wave appends or sets one
        if (ovf || (h >= t && sp == 0 && np == 0)) break;
        if (step >= kMaxSteps) {  // every source ends: an unfinished one is reported like a capacity overflow
            if (lane == 0) atomicOr(&sh.overflow, OVF_STEPS);
            break;
        }
        // the batch: the ring's items, then the oldest spilled records on the lanes left (their stores were made
        // visible by the full barrier that ended the step in which they were spilled)
        const int nRing = min((int)blockDim.x, t - h);
        const int take = min((int)blockDim.x - nRing, sp);
        const int nb = nRing + take;
        if (lane == 0 && take > 0) sh.spillHead = s0 + take;  // appends read the head only after the next barrier
        const bool mine = lane < nb;
        int fk = -1, kind = 0, g = -1, fs = 0, k = 0;  // kind: 0 window, 1 vertex event, 2 vertex-face item
        double b0 = 0, b1 = 0, sx = 0, sy = 0, sg = 0, L = 0, x2 = 0, y2 = 0;
        double cf[6] = {0, 0, 0, 0, 0, 0};
        float2 rot = make_float2(0.f, 0.f);
        Corners K;
        if (mine) {
            if (lane < nRing) {
                const int q = (h + lane) % T.qcap;
                fk = T.qfk[q];
                sg = T.qsg[q];
                rot = T.qt[q];
                b0 = T.qb0[q];
                b1 = T.qb1[q];
                sx = T.qsx[q];
                sy = T.qsy[q];
            } else if (spill) {  // (take > 0 only with a spill queue)
                const SpillRecord r = spill[(s0 + lane - nRing) % P.spillCap];
                fk = r.fk;
                sg = r.sg;
                rot = make_float2(r.tx, r.ty);
                b0 = r.b0;
                b1 = r.b1;
                sx = r.sx;
                sy = r.sy;
            }
            if (fk >= kVertexFace) {
                kind = 2;
                fs = (fk & (kVertexFace - 1)) >> 2;
                k = (fk & 3) + 1;  // the far edge of the face, seen from the vertex at corner (fk & 3)
                if (k == 3) k = 0;
            } else if (kPending || fk >= 0) {  // (no vertex events with pending marks)
                kind = 0;
                fs = fk >> 3;
                k = (fk >> 1) & 3;
            } else {
                kind = 1;
            }
            if (kPending || kind != 1) {  // geometry of the edge the item works on, fetched before the barrier
                g = T.fglob[fs];
                const double* geo = M.fgeo + 9 * g + 3 * k;
                L = geo[0];
                x2 = geo[1];
                y2 = geo[2];
                if (kind == 0)
                    for (int i = 0; i < 6; ++i) cf[i] = M.fchild[18 * g + 6 * k + i];
            }
            if (kind == 0) {
                // the face's corners and their distances as the previous step left them, read while the geometry
                // loads are in flight (no distance changes before the barrier)
                const int k1 = k == 2 ? 0 : k + 1, k2 = k == 0 ? 2 : k - 1;
                K.vA = T.fcv[3 * fs + k];
                K.vB = T.fcv[3 * fs + k1];
                K.vC = T.fcv[3 * fs + k2];
                K.nbL = T.fnb[3 * fs + k2];
                K.nbR = T.fnb[3 * fs + k1];
                K.dA = bitsd(T.vdist[K.vA]);
                K.dB = bitsd(T.vdist[K.vB]);
                K.dC = bitsd(T.vdist[K.vC]);
            }
        }
        // lower bounds of the interval-end distances: a window near the pruning thresholds is propagated rather than
        // dropped (propagating a useless window changes no distance); the same values scale the target tests'
        // tolerance in both passes. With pending marks they are needed here, for the batch's smallest key (a lower
        // bound of every distance its items and their descendants set); otherwise they are computed in the first
        // pass (measured faster for each).
        double r0 = 0, r1 = 0;
        if constexpr (kPending) {
            r0 = endDistLow(b0, sx, sy);
            r1 = endDistLow(b1, sx, sy);
            if (mine) atomicMin(&sh.minKey, dbits(kind == 0 ? keyOf(b0, b1, sx, sy, sg, r0, r1) : sg));
        }
        // once every lane holds its record, the ring slots [h, h + nRing) are free for this step's appends
        ldsBarrier();
        C.qBase = h + nRing;
        VCand cP0 = noCand(), cP1 = noCand(), cP2 = noCand();
        bool live = false;
        if (mine) {
            if (kind == 0) {
                if constexpr (!kPending) {
                    r0 = endDistLow(b0, sx, sy);
                    r1 = endDistLow(b1, sx, sy);
                }
                live = keyOf(b0, b1, sx, sy, sg, r0, r1) <= C.rMax && !filteredBy(K.dA, K.dB, L, b0, b1, sg, r0, r1);
                if (live) propagate(C, fk, g, L, x2, y2, b0, b1, sx, sy, sg, r0, r1, cf, rot, K, cP0, cP1, cP2);
            } else if (!kPending && kind == 1) {
                int vs = -fk - 1;
                if (T.vdist[vs] == dbits(sg) && sg <= C.rMax) fanOutVertex(C, vs, sg);
            } else {
                int vs = T.fcv[3 * fs + (fk & 3)];
                live = T.vdist[vs] == dbits(sg);  // a later, shorter path to the vertex supersedes this item
                if (live) {
                    int nbk = T.fnb[3 * fs + k];
                    // the vertex unfolded into the neighbour's frame of the same edge (origin at the edge's far end)
                    addWindow(C, nbk >> 2, nbk & 3, L, 0.0, L, L - x2, -y2, sg, false, rot, cP0, cP1);
                }
            }
        }
        if (kPending && np > 0 && lane >= (int)blockDim.x - 64) {
            // pending pseudo-source vertices, on the last wavefront (batches fill the lanes from the first): a vertex
            // fans out once no window of this batch can lower its distance (every distance they set is at least the
            // batch's smallest key; with no window in the batch, every pending vertex goes). A later improvement
            // marks it again; the items of the superseded fan-out are then dropped when they come up.
            const unsigned long long mk = sh.minKey;
            for (int v = lane - ((int)blockDim.x - 64); v < sh.nVerts; v += 64) {
                if (!(T.vflg[v] & 2)) continue;
                const unsigned long long db = T.vdist[v];
                // (the bit patterns of non-negative doubles order as their values; mk = ~0: no item in the batch)
                if (db > mk && bitsd(db) > bitsd(mk) + slack) continue;
                atomicAnd(reinterpret_cast<unsigned*>(T.vflg) + (v >> 2), ~(2u << (8 * (v & 3))));
                atomicSub(&sh.nPend, 1);
                fanOutVertex(C, v, bitsd(db));
            }
        }
        ldsBarrier();
        if (live) {
            if (kind == 0 && T.fhead[fs] >= 0)
                windowTargets<true, TT>(C, fs, g, k, L, x2, y2, b0, b1, sx, sy, sg, r0, r1, fk & 1, rot);
            finishVertexCands<kPending>(C, cP0, cP1, cP2);
        }
        __syncthreads();  // a full barrier: this step's spill stores are visible to the next batches
        h += nRing;
        if (lane == 0) {
            sh.nSteps += 1;
            sh.nItems += nb;
        }
    }
}

// kIt: faces per lane of a prepared patch list (0: no register caching, for tables of any size)
template <int kIt, bool kPending, class TT>
__device__ void solveSource(const GeoParams& P, const TT& T, Shared& sh, int src, int* cur, int* nxt, int frontCap) {
    const int lane = threadIdx.x;
    const DeviceMesh& M = P.mesh;
    const int pi = P.srcIdx ? P.srcIdx[src] : src;
    const int sf = P.pface[pi];
    const int tb = P.toff[src], nT = P.toff[src + 1] - tb;
    // the source's potential, truncation radius and sentinel (its group's, with particle groups)
    const int grp = P.srcGroup ? P.srcGroup[pi] : -1;
    const Potential pot = grp >= 0 ? P.groupPot[grp] : P.pot;
    const double rMax = grp >= 0 ? P.groupRMax[grp] : P.rMax;
    const double sentinelDist = grp >= 0 ? P.groupSentinel[grp] : P.sentinelDist;
    for (int i = lane; i <= T.hmask; i += blockDim.x) T.hkey[i] = -1;
    for (int i = lane; i <= T.vhmask; i += blockDim.x) T.vhkey[i] = -1;
    for (int i = lane; i < T.vcap; i += blockDim.x) T.vfo[i] = 0;
    if (lane == 0) {
        sh.head = sh.tail = sh.nFaces = sh.nVerts = sh.nSteps = sh.nItems = sh.spillHead = sh.spillTail = sh.nPend = 0;
        sh.srcSlot = -1;
        sh.overflow = nT > T.tcap ? OVF_TARGETS : 0;
        sh.nCur = sh.nNext = sh.goals = 0;
        initSourceFrame(P, sh, pi, sf);
    }
    SourceEdge se{};
    {
        const int ek = sourceEdgeOf(lane);
        if (ek >= 0 && ek < 3) se = loadSourceEdge(M, sf, ek);
    }
    SpillRecord* spill = P.spill ? P.spill + (size_t)(P.spillPerBlock ? blockIdx.x : src) * P.spillCap : nullptr;
    __syncthreads();
    bool ok = !sh.overflow;
    __syncthreads();
    const long long c0 = wall_clock64();
    if (ok) ok = buildPatch<kIt>(P, T, sh, cur, nxt, frontCap, src, sf, tb, nT);
    const long long c1 = wall_clock64();
    if (ok) propagateAll<kPending>(P, T, sh, sf, tb, spill, rMax, kReleaseSlack * rMax, se);
    __syncthreads();
    const long long c2 = wall_clock64();
    // ---- finalize: paths through the target face's corners, direct visibility, outputs, force
    const bool ovf = sh.overflow != 0;
    double F0 = 0, F1 = 0, F2 = 0, E = 0;
    for (int t = lane; t < nT; t += blockDim.x) {
        const int pj = P.tidx[tb + t];
        const int out = tb + t;
        if (ovf) {
            if (P.pairFlags) P.pairFlags[out] = PAIR_OVERFLOW;
            continue;
        }
        if (P.keep && !P.keep[out]) continue;  // a pair the screen dropped: not solved, no force
        const int f = P.pface[pj];
        const double* b = P.pbary + 3 * pj;
        double sb = b[0] + b[1] + b[2];
        double cc[3][3];
        for (int k = 0; k < 3; ++k) loadV(M, M.fv[3 * f + k], cc[k]);
        double Q[3];
        for (int a = 0; a < 3; ++a) Q[a] = (cc[0][a] * b[0] + cc[1][a] * b[1] + cc[2][a] * b[2]) / sb;
        double best = bitsd(T.tbest[t]);
        float2 st = T.tst[t];
        float4 et4 = T.tet ? T.tet[t] : make_float4(0.f, 0.f, 0.f, 0.f);
        double et[3] = {et4.x, et4.y, et4.z};
        if (f == sf) {
            double r[3] = {Q[0] - sh.S[0], Q[1] - sh.S[1], Q[2] - sh.S[2]};
            double l = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            if (l < best) {
                best = l;
                if (l > 0) {
                    st = make_float2((float)((r[0] * sh.t1[0] + r[1] * sh.t1[1] + r[2] * sh.t1[2]) / l),
                                     (float)((r[0] * sh.t2[0] + r[1] * sh.t2[1] + r[2] * sh.t2[2]) / l));
                    for (int a = 0; a < 3; ++a) et[a] = r[a] / l;
                } else {
                    st = make_float2(0.f, 0.f);
                    et[0] = et[1] = et[2] = 0;
                }
            }
        }
        const int fs = T.tslot[t];
        for (int k = 0; k < 3 && fs >= 0; ++k) {
            int vs = T.fcv[3 * fs + k];
            double dv = bitsd(T.vdist[vs]);
            if (!(dv < kInf)) continue;
            double r[3] = {Q[0] - cc[k][0], Q[1] - cc[k][1], Q[2] - cc[k][2]};
            double l = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            if (dv + l < best) {
                best = dv + l;
                st = T.vtan[vs];
                for (int a = 0; a < 3; ++a) et[a] = l > 0 ? r[a] / l : 0.0;
            }
        }
        uint32_t flags = 0;
        double d = best;
        double s3[3] = {st.x * sh.t1[0] + st.y * sh.t2[0], st.x * sh.t1[1] + st.y * sh.t2[1], st.x * sh.t1[2] + st.y * sh.t2[2]};
        double nn = sqrt(s3[0] * s3[0] + s3[1] * s3[1] + s3[2] * s3[2]);
        if (nn > 0)
            for (int a = 0; a < 3; ++a) s3[a] /= nn;
        if (!(best < kInf)) {
            flags |= PAIR_UNREACHED;
            s3[0] = s3[1] = s3[2] = 0;
            et[0] = et[1] = et[2] = 0;
            if (P.useSentinel) {
                flags |= PAIR_SENTINEL;
                d = sentinelDist;
                s3[2] = 1;
                et[2] = 1;
            }
        } else if (best > rMax) {
            flags |= PAIR_BEYOND;
        }
        if (P.pairDist) P.pairDist[out] = d;
        if (P.pairStart) {
            P.pairStart[3 * out] = (float)s3[0];
            P.pairStart[3 * out + 1] = (float)s3[1];
            P.pairStart[3 * out + 2] = (float)s3[2];
        }
        if (P.pairEnd) {
            P.pairEnd[3 * out] = (float)et[0];
            P.pairEnd[3 * out + 1] = (float)et[1];
            P.pairEnd[3 * out + 2] = (float)et[2];
        }
        if (P.pairFlags) P.pairFlags[out] = flags;
        if (pot.kind != POT_NONE && (!(flags & PAIR_UNREACHED) || (flags & PAIR_SENTINEL))) {
            double e;
            double sc = potentialScale(pot, d, &e);
            F0 += sc * s3[0];
            F1 += sc * s3[1];
            F2 += sc * s3[2];
            E += e;
        }
    }
    F0 = wsum(F0);
    F1 = wsum(F1);
    F2 = wsum(F2);
    E = wsum(E);
    if (blockDim.x > 64) {  // combine the wavefronts' partial sums in fixed order, through the idle ring storage
        double* red = T.qb0;
        __syncthreads();
        if ((lane & 63) == 0) {
            red[4 * (lane >> 6)] = F0;
            red[4 * (lane >> 6) + 1] = F1;
            red[4 * (lane >> 6) + 2] = F2;
            red[4 * (lane >> 6) + 3] = E;
        }
        __syncthreads();
        if (lane == 0) {
            F0 = F1 = F2 = E = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
                F0 += red[4 * w];
                F1 += red[4 * w + 1];
                F2 += red[4 * w + 2];
                E += red[4 * w + 3];
            }
        }
    }
    if (lane == 0) {
        if (P.forces) {
            P.forces[3 * src] = F0;
            P.forces[3 * src + 1] = F1;
            P.forces[3 * src + 2] = F2;
        }
        if (P.energy) P.energy[src] = E;
        if (P.status) P.status[src] = sh.overflow;
        // counts, then phase durations in ticks of the constant 100 MHz clock; a source these tables could not hold
        // keeps the figures of the tables that last solved it (the larger tiers order their work by that duration)
        if (P.patchStats && !sh.overflow) {
            int* ps = P.patchStats + kPatchStats * src;
            ps[0] = sh.nFaces;
            ps[1] = sh.nVerts;
            ps[2] = sh.nSteps;
            ps[3] = sh.nItems;
            ps[4] = (int)(c1 - c0);
            ps[5] = (int)(c2 - c1);
            ps[6] = (int)(wall_clock64() - c2);
        }
    }
    __syncthreads();
}

// Sources a table capacity stopped: the larger tiers take them
constexpr int kTierReasons = OVF_FACES | OVF_VERTS | OVF_QUEUE | OVF_TARGETS;

// The tier-list bucket of source i, by its last completed solve's duration (patchStats, ticks of 100 MHz): bucket b
// holds durations of about 2^(b+10) ticks (10 us) and up, the last everything from 1.3 ms on.
__device__ inline int tierBucket(const int* patchStats, int i) {
    if (!patchStats) return 0;
    const int* ps = patchStats + (size_t)kPatchStats * i;
    const unsigned t = (unsigned)max(ps[4], 0) + (unsigned)max(ps[5], 0) + (unsigned)max(ps[6], 0);
    return min(max((31 - __clz((int)(t | 1u))) - 10, 0), kTierBuckets - 1);
}

// appends source i to a tier list in its bucket
__device__ void appendBucketed(int32_t* list, int32_t* count, int n, const int* patchStats, int i) {
    const int b = tierBucket(patchStats, i);
    list[(size_t)b * n + atomicAdd(count + 1 + b, 1)] = i;
}

}  // namespace

// Source of this workgroup. Workgroups are dispatched round-robin over the 8 XCDs (workgroup i to XCD i mod 8), each
// with its own L2. With an order (sources sorted by spatial cell), XCD x takes the contiguous chunk
// [x * chunk, (x + 1) * chunk) of that order, so the mesh data its sources touch is a compact region that stays in its
// L2. Returns -1 for the padding workgroups past the last source.
__device__ inline int sourceOfBlock(const GeoParams& P) {
    const int i = blockIdx.x;
    if (P.blockOrder == nullptr) return i < P.nSrc ? i : -1;
    const int p = (i % kXcds) * P.blockChunk + i / kXcds;
    return p < P.nSrc ? P.blockOrder[p] : -1;
}

int gridForSources(const GeoParams& P) { return P.blockOrder ? kXcds * P.blockChunk : P.nSrc; }

// Dynamic distribution: the next source for a workgroup on XCD x, or -1 when every chunk is taken. Chunk c holds
// positions [c * chunk, (c + 1) * chunk) of the order; each position is handed out once by its chunk's counter.
__device__ int grabSource(const GeoParams& P, int x) {
    const int n = P.nSrc, chunk = (n + kXcds - 1) / kXcds;
    for (int d = 0; d < kXcds; ++d) {
        const int c = (x + d) % kXcds;
        const int lo = c * chunk, len = min(chunk, n - lo);
        if (len <= 0) continue;
        if (__hip_atomic_load(&P.grab[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len) continue;
        const int j = atomicAdd(&P.grab[c], 1);
        if (j < len) return P.blockOrder ? P.blockOrder[lo + j] : lo + j;
    }
    return -1;
}


// LDS variant: tables in __shared__ memory, one source per workgroup of NT lanes, WGS workgroups per CU (each within
// 160 KB / WGS of LDS). FC faces, VC vertices, QC ring records, TC targets; HF and HV face and vertex hash slots, which
// share the ring's storage with the vertex ids during setup and with the force reduction afterwards; kEnd: the variant
// that also reports end tangents.
template <int FC, int VC, int QC, int TC, int HF, int HV, int NT, int WGS, bool kEnd, bool kList, int FR = 0>
__global__ __launch_bounds__(NT, WGS) void geodesicForceLds(GeoParams P, const int32_t* list, const int32_t* count, int32_t* next,
                                                             int32_t* nextCount, int* reasons) {
    constexpr int kRingBytes = QC * (5 * 8 + 4 + 8);
    static_assert(HF * 6 + HV * 6 + VC * 4 + FC * 4 <= kRingBytes, "setup hashes, vertex ids and list heads live in the ring's storage");
    static_assert(QC >= 4 * kMaxWaves, "the epilogue's reduction uses the ring's storage");
    static_assert(QC % 2 == 0 && (HF & (HF - 1)) == 0 && (HV & (HV - 1)) == 0, "alignment and power-of-two hashes");
    static_assert(4 * FC <= 32767 && VC <= 32767 && TC <= 32767, "slots must fit int16");
    static_assert(VC % 4 == 0, "the vertex flags are updated as 32-bit words");
    static_assert(NT % 64 == 0 && NT / 64 <= kMaxWaves, "whole wavefronts");
    __shared__ __attribute__((aligned(16))) unsigned char ring[kRingBytes];
    __shared__ int32_t fglob[FC];
    __shared__ int16_t fcv[3 * FC], fnb[3 * FC], fhead[FC], vfl[3 * FC];
    __shared__ unsigned long long vdist[VC];
    __shared__ float2 vtan[VC];
    __shared__ __attribute__((aligned(4))) uint8_t vflg[VC];
    __shared__ int32_t vfo[VC];
    __shared__ int16_t tnext[TC], tslot[TC];
    __shared__ double tq[3 * TC];
    __shared__ unsigned long long tbest[TC];
    __shared__ float2 tst[TC];
    __shared__ float4 tet[kEnd ? TC : 1];
    __shared__ int32_t front[FR > 0 ? 2 * FR : 1];  // breadth-first frontiers for sources without a prepared face list
    __shared__ Shared sh;
    Tables<int16_t> T;
    T.qb0 = (double*)ring;
    T.qb1 = T.qb0 + QC;
    T.qsx = T.qb1 + QC;
    T.qsy = T.qsx + QC;
    T.qsg = T.qsy + QC;
    T.qfk = (int32_t*)(T.qsg + QC);
    T.qt = (float2*)(T.qfk + QC);
    T.hkey = (int32_t*)ring;
    T.hslot = (int16_t*)(T.hkey + HF);
    T.vhkey = (int32_t*)(T.hslot + HF);
    T.vhslot = (int16_t*)(T.vhkey + HV);
    T.vglob = (int32_t*)(T.vhslot + HV);  // read during setup only
    T.thead = T.vglob + VC;               // written during setup only
    T.fglob = fglob;
    T.fcv = fcv;
    T.fnb = fnb;
    T.fhead = fhead;
    T.vdist = vdist;
    T.vtan = vtan;
    T.vflg = vflg;
    T.vfo = vfo;
    T.vfl = vfl;
    T.tnext = tnext;
    T.tslot = tslot;
    T.tq = tq;
    T.tbest = tbest;
    T.tst = tst;
    T.tet = kEnd ? tet : nullptr;
    T.hmask = HF - 1;
    T.vhmask = HV - 1;
    T.fcap = FC;
    T.vcap = VC;
    T.qcap = QC;
    T.tcap = TC;
    if constexpr (kList) {
        // persistent over a device list of sources (the larger tables' pass), handed out one at a time by the counter
        // count[0] (zero at launch): the longest first, bucket kTierBuckets - 1 down to bucket 0 (count[1 + b] entries
        // at list + b * nSrc), so that no workgroup ends with a long source while the others idle
        for (;;) {
            if (threadIdx.x == 0) {
                // the counter is read before it is taken, so that workgroups that find the list used up (or empty)
                // leave without an atomic on the one address all of them share
                int total = 0, s = -1;
                for (int b = 0; b < kTierBuckets; ++b) total += count[1 + b];
                if (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < total) {
                    int i = atomicAdd(const_cast<int32_t*>(count), 1);
                    for (int b = kTierBuckets - 1; b >= 0; --b) {
                        const int nb = count[1 + b];
                        if (i < nb) {
                            s = list[(size_t)b * P.nSrc + i];
                            break;
                        }
                        i -= nb;
                    }
                }
                sh.next = s;
            }
            __syncthreads();
            const int src = sh.next;
            __syncthreads();  // every lane has read sh.next before the source's setup writes the shared state
            if (src < 0) break;
            solveSource<(FC + NT - 1) / NT, false>(P, T, sh, src, FR > 0 ? front : nullptr, FR > 0 ? front + FR : nullptr, FR);
            // a source these tables cannot hold either goes on to the next tier's list (sh.overflow is read after the
            // source's last barrier, before the next source resets it)
            if (next && threadIdx.x == 0 && (sh.overflow & kTierReasons)) appendBucketed(next, nextCount, P.nSrc, P.patchStats, src);
        }
        return;
    }
    int src;
    if (P.grab) {  // the source is taken dynamically (blockIdx.x % kXcds is the XCD); surplus workgroups exit
        if (threadIdx.x == 0) sh.next = grabSource(P, blockIdx.x % kXcds);
        __syncthreads();
        src = sh.next;
        __syncthreads();
    } else {
        src = sourceOfBlock(P);
    }
    if (src < 0) return;
    solveSource<(FC + NT - 1) / NT, true>(P, T, sh, src, nullptr, nullptr, 0);
    // a source these tables could not hold goes on to the first tier's list (sh.overflow is read after the source's
    // last barrier)
    if (next && threadIdx.x == 0 && sh.overflow) {
        if (reasons)
            for (int b = 0; b < 6; ++b)
                if (sh.overflow & (1 << b)) atomicAdd(&reasons[b], 1);
        if (sh.overflow & kTierReasons) appendBucketed(next, nextCount, P.nSrc, P.patchStats, src);
    }
}

// Scratch variant: the same solver with its tables in a per-workgroup slice of an HBM buffer, persistent over a
// list of sources (sources the LDS variant could not hold, or whole-mesh solves). NT lanes per source: 64 for the
// patches the LDS tables could not hold (many workgroups, few sources each), kWideLanes for whole-mesh solves, whose
// fronts of thousands of windows fill them (few sources, one each).
template <int NT>
__global__ __launch_bounds__(NT) void geodesicForceScratch(GeoParams P, ScratchLayout L, const int32_t* list, int nList, bool redoOnly) {
    __shared__ Shared sh;
    char* base = (char*)L.base + (size_t)blockIdx.x * L.bytesPerBlock;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    Tables<int32_t> T;
    T.hkey = (int32_t*)take(4 * (size_t)L.hfcap);
    T.hslot = (int32_t*)take(4 * (size_t)L.hfcap);
    T.vhkey = (int32_t*)take(4 * (size_t)L.hvcap);
    T.vhslot = (int32_t*)take(4 * (size_t)L.hvcap);
    T.fglob = (int32_t*)take(4 * (size_t)L.fcap);
    T.fhead = (int32_t*)take(4 * (size_t)L.fcap);
    T.thead = (int32_t*)take(4 * (size_t)L.fcap);
    T.fcv = (int32_t*)take(12 * (size_t)L.fcap);
    T.fnb = (int32_t*)take(12 * (size_t)L.fcap);
    T.vfl = (int32_t*)take(12 * (size_t)L.fcap);
    T.vglob = (int32_t*)take(4 * (size_t)L.vcap);
    T.vflg = (uint8_t*)take((size_t)L.vcap);
    T.vfo = (int32_t*)take(4 * (size_t)L.vcap);
    T.vdist = (unsigned long long*)take(8 * (size_t)L.vcap);
    T.vtan = (float2*)take(8 * (size_t)L.vcap);
    T.qb0 = (double*)take(8 * (size_t)L.qcap);
    T.qb1 = (double*)take(8 * (size_t)L.qcap);
    T.qsx = (double*)take(8 * (size_t)L.qcap);
    T.qsy = (double*)take(8 * (size_t)L.qcap);
    T.qsg = (double*)take(8 * (size_t)L.qcap);
    T.qfk = (int32_t*)take(4 * (size_t)L.qcap);
    T.qt = (float2*)take(8 * (size_t)L.qcap);
    T.tnext = (int32_t*)take(4 * (size_t)L.tcap);
    T.tslot = (int32_t*)take(4 * (size_t)L.tcap);
    T.tq = (double*)take(24 * (size_t)L.tcap);
    T.tbest = (unsigned long long*)take(8 * (size_t)L.tcap);
    T.tst = (float2*)take(8 * (size_t)L.tcap);
    T.tet = (float4*)take(16 * (size_t)L.tcap);
    int* cur = (int*)take(4 * (size_t)L.fcap);
    int* nxt = (int*)take(4 * (size_t)L.fcap);
    T.hmask = L.hfcap - 1;
    T.vhmask = L.hvcap - 1;
    T.fcap = L.fcap;
    T.vcap = L.vcap;
    T.qcap = L.qcap;
    T.tcap = L.tcap;
    GeoParams Q = P;
    Q.spill = nullptr;  // the scratch ring is sized for a whole source
    if constexpr (NT == 64) {
        if (redoOnly && !list) {
            // the sources the earlier kernels left unsolved (nonzero status): each workgroup tests a chunk of 64
            // statuses at once and solves the flagged ones in turn (the ballot is the same on every lane)
            for (int c = blockIdx.x * 64; c < nList; c += gridDim.x * 64) {
                const int i = c + (int)threadIdx.x;
                unsigned long long m = __ballot(i < nList && P.status[i] != 0);
                while (m) {
                    const int j = __ffsll((long long)m) - 1;
                    m &= m - 1;
                    solveSource<0, false>(Q, T, sh, c + j, cur, nxt, L.fcap);
                }
            }
            return;
        }
    }
    for (int i = blockIdx.x; i < nList; i += gridDim.x) {
        const int src = list ? list[i] : i;
        if (redoOnly && P.status[src] == 0) continue;  // solved by the LDS variant
        solveSource<0, false>(Q, T, sh, src, cur, nxt, L.fcap);
    }
}

// Patch kernel: the reference's submesh face set of every source, one wavefront and ~13 KB of LDS per source (so
// a dozen sources per CU hide the search's memory latency), written as a compact list (count -1: more faces than the
// list holds, or a breadth-first level wider than its buffer).
struct FaceSet {  // membership only: the list's order numbers the faces, no id -> slot map is kept
    static constexpr bool kSlots = false;
    int32_t* hkey;
    int32_t* fglob;  // the output list itself (HBM): faces are written where they are found
    int hmask, fcap;
};

// One wavefront per source; 6.2 KB of LDS (hash and frontier lists), so that LDS admits 26 sources per CU.
__global__ __launch_bounds__(64) void patchKernel(GeoParams P, int32_t* list, int32_t* count, int stride) {
    __shared__ int32_t hkey[kPatchHashCap];
    __shared__ int32_t cur[kPatchFrontCap], nxt[kPatchFrontCap];
    __shared__ Shared sh;
    const int src = sourceOfBlock(P);
    if (src < 0) return;
    const int lane = threadIdx.x;
    const int pi = P.srcIdx ? P.srcIdx[src] : src;
    const int sf = P.pface[pi];
    const int tb = P.toff[src], nT = P.toff[src + 1] - tb;
    // slots below fcap <= stride: every write stays inside this source's stride of the list
    FaceSet T{hkey, list + (size_t)src * stride, kPatchHashCap - 1, min(stride, kPatchListCap)};
    for (int i = lane; i < kPatchHashCap; i += blockDim.x) hkey[i] = -1;
    if (lane == 0) {
        sh.nFaces = sh.overflow = sh.nCur = sh.nNext = sh.goals = 0;
        initSourceFrame(P, sh, pi, sf);
    }
    __syncthreads();
    patchSearch(P, T, sh, cur, nxt, kPatchFrontCap, src, sf, tb, nT);
    __syncthreads();
    if (lane == 0) count[src] = sh.overflow ? -1 : sh.nFaces;
}

// Frames of the two edges a window entering face g through edge k leaves by, per (face, edge): 6 doubles
__global__ __launch_bounds__(256) void childFramesKernel(DeviceMesh M, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;  // 3 g + k
    if (i >= 3 * M.nF) return;
    const double* geo = M.fgeo + 3 * i;
    const double L = geo[0], x2 = geo[1], y2 = geo[2];
    const double lL = sqrt(x2 * x2 + y2 * y2);
    const double ex = L - x2, ey = -y2;
    const double lR = sqrt(ex * ex + ey * ey);
    double* o = out + 6 * (size_t)i;
    o[0] = lL;
    o[1] = x2 / lL;
    o[2] = y2 / lL;
    o[3] = lR;
    o[4] = ex / lR;
    o[5] = ey / lR;
}

hipError_t launchChildFrames(const DeviceMesh& mesh, double* out, hipStream_t stream) {
    if (mesh.nF <= 0) return hipSuccess;
    hipLaunchKernelGGL(childFramesKernel, dim3((3 * mesh.nF + 255) / 256), dim3(256), 0, stream, mesh, out);
    return hipGetLastError();
}

hipError_t launchPatch(const GeoParams& P, int32_t* list, int32_t* count, int stride, hipStream_t stream) {
    if (P.nSrc <= 0 || P.patchRadius == nullptr) return hipSuccess;
    if (stride > kPatchListCap) return hipErrorInvalidValue;
    hipLaunchKernelGGL(patchKernel, dim3(gridForSources(P)), dim3(64), 0, stream, P, list, count, stride);
    return hipGetLastError();
}

ScratchLayout makeScratchLayout(long faces, long targets) {
    auto p2 = [](long n) {
        long p = 1;
        while (p < n) p <<= 1;
        return (int)p;
    };
    ScratchLayout L{};
    L.fcap = (int)std::max(faces, 16L);
    L.vcap = L.fcap + 64;  // a patch has fewer vertices than faces plus its border
    L.hfcap = p2(2 * (long)L.fcap);
    L.hvcap = p2(2 * (long)L.vcap);
    L.qcap = (int)std::max(16 * (long)L.fcap, 4096L);
    L.tcap = (int)std::max(targets, 1L);
    auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t F = L.fcap, V = L.vcap, Q = L.qcap, Tn = L.tcap, HF = L.hfcap, HV = L.hvcap;
    L.bytesPerBlock = 2 * r(4 * HF) + 2 * r(4 * HV)                        // hashes
                      + 3 * r(4 * F) + 3 * r(12 * F)                       // fglob, fhead, thead; fcv, fnb, vfl
                      + r(4 * V) + r(V) + r(4 * V) + 2 * r(8 * V)          // vglob, vflg, vfo; vdist, vtan
                      + 5 * r(8 * Q) + r(4 * Q) + r(8 * Q)                 // ring
                      + 2 * r(4 * Tn) + r(24 * Tn) + 2 * r(8 * Tn) + r(16 * Tn)  // targets
                      + 2 * r(4 * F);                                      // frontiers
    return L;
}

template <int FC, int VC, int QC, int TC, int HF, int HV, int NT, int WGS>
static void launchLdsVariant(const GeoParams& P, int32_t* next, int32_t* nextCount, int* reasons, hipStream_t stream) {
    const int32_t* none = nullptr;
    // dynamic distribution: twice the sources' workgroups, so that an XCD whose chunk costs less can take sources from
    // the others (the surplus exits at once)
    const int grid = P.grab ? 2 * kXcds * ((P.nSrc + kXcds - 1) / kXcds) : gridForSources(P);
    if (P.pairEnd)
        hipLaunchKernelGGL((geodesicForceLds<FC, VC, QC, TC, HF, HV, NT, WGS, true, false>), dim3(grid), dim3(NT), 0, stream, P, none, none,
                           next, nextCount, reasons);
    else
        hipLaunchKernelGGL((geodesicForceLds<FC, VC, QC, TC, HF, HV, NT, WGS, false, false>), dim3(grid), dim3(NT), 0, stream, P, none, none,
                           next, nextCount, reasons);
}

template <int FC, int VC, int QC, int TC, int HF, int HV, int NT, int WGS, int FR>
static void launchLdsList(const GeoParams& P, const int32_t* list, const int32_t* count, int32_t* next, int32_t* nextCount, int grid,
                          hipStream_t stream) {
    if (P.pairEnd)
        hipLaunchKernelGGL((geodesicForceLds<FC, VC, QC, TC, HF, HV, NT, WGS, true, true, FR>), dim3(grid), dim3(NT), 0, stream, P, list,
                           count, next, nextCount, nullptr);
    else
        hipLaunchKernelGGL((geodesicForceLds<FC, VC, QC, TC, HF, HV, NT, WGS, false, true, FR>), dim3(grid), dim3(NT), 0, stream, P, list,
                           count, next, nextCount, nullptr);
}

hipError_t launchGeodesicLdsLarge(const GeoParams& Q, int tier, const int32_t* list, int32_t* count, int32_t* next, int32_t* nextCount,
                                  SpillRecord* spill, hipStream_t stream) {
    if (Q.nSrc <= 0) return hipSuccess;
    GeoParams P = Q;
    // CUs of the current device (cached per device)
    static int cusOf[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (cusOf[dev] <= 0 && (hipDeviceGetAttribute(&cusOf[dev], hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                            cusOf[dev] <= 0))
        cusOf[dev] = 256;
    const int cus = cusOf[dev];
    // the spill queue per workgroup: kTierSpillBytes over the grid (8,192 records for 512 workgroups, 16,384 for 256)
    const int grid = std::min(tier == 1 ? 2 * cus : cus, 512);
    P.spill = spill;
    P.spillPerBlock = 1;
    P.spillCap = (int)(kTierSpillBytes / sizeof(SpillRecord) / grid);
    if (tier == 1)  // two 256-lane workgroups per CU, 78 KB of LDS each: up to 1,024 faces, 640 vertices, 256 targets
        launchLdsList<1024, 640, 576, 256, 2048, 1024, 256, 2, 0>(P, list, count, next, nextCount, grid, stream);
    else  // one 512-lane workgroup per CU, 117 KB of LDS: up to 1,536 faces, 896 vertices, 256 targets; it searches the
          // patch itself when the patch kernel could not list it (more faces than a list holds)
        launchLdsList<1536, 896, 896, 256, 4096, 2048, 512, 1, 512>(P, list, count, next, nextCount, grid, stream);
    return hipGetLastError();
}

hipError_t launchGeodesicLds(const GeoParams& P, int variant, int32_t* next, int32_t* nextCount, int* reasons, hipStream_t stream) {
    if (P.nSrc <= 0) return hipSuccess;
    if (variant == 1)
        launchLdsVariant<kLdsFaceCap, kLdsVertCap, kLdsQueueCap, kLdsTargetCap, kLdsFaceHash, kLdsVertHash, 512, 2>(P, next, nextCount,
                                                                                                                  reasons, stream);
    else if (variant == 2)  // smaller tables and ring, three workgroups per CU
        launchLdsVariant<576, 352, 384, 224, 1024, 512, 256, 3>(P, next, nextCount, reasons, stream);
    else if (variant == 3)  // four workgroups per CU: 40 KB each
        launchLdsVariant<576, 320, 280, 128, 1024, 512, 256, 4>(P, next, nextCount, reasons, stream);
    else
        launchLdsVariant<kLdsFaceCap, kLdsVertCap, kLdsQueueCap, kLdsTargetCap, kLdsFaceHash, kLdsVertHash, 256, 2>(P, next, nextCount,
                                                                                                                  reasons, stream);
    return hipGetLastError();
}

hipError_t launchGeodesicScratch(const GeoParams& P, const ScratchLayout& L, const int32_t* list, int nList, int blocks,
                                 hipStream_t stream) {
    if (nList <= 0) return hipSuccess;
    hipLaunchKernelGGL(geodesicForceScratch<64>, dim3(blocks), dim3(64), 0, stream, P, L, list, nList, false);
    return hipGetLastError();
}

hipError_t launchGeodesicScratchWide(const GeoParams& P, const ScratchLayout& L, const int32_t* list, int nList, int blocks,
                                     hipStream_t stream) {
    if (nList <= 0) return hipSuccess;
    hipLaunchKernelGGL(geodesicForceScratch<kWideLanes>, dim3(blocks), dim3(kWideLanes), 0, stream, P, L, list, nList, false);
    return hipGetLastError();
}

hipError_t launchGeodesicScratchRedo(const GeoParams& P, const ScratchLayout& L, int blocks, hipStream_t stream) {
    if (P.nSrc <= 0) return hipSuccess;
    hipLaunchKernelGGL(geodesicForceScratch<64>, dim3(blocks), dim3(64), 0, stream, P, L, (const int32_t*)nullptr, P.nSrc, true);
    return hipGetLastError();
}

}  // namespace csg
