// Shared device-side definitions for the curvedSpaceSim geodesic force path on gfx950 (MI355X).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace csg {

// Mesh resident in HBM for the lifetime of a context. Index conventions: face f has corners fv[3f+k]; its local
// edge k runs from corner k to corner k+1; fadj[3f+k] = neighbour face * 4 + its local index of the same edge
// (-1 on a border). fgeo holds, per face and local edge k, {L, x2, y2}: the edge length and the opposite corner in
// the edge frame (origin corner k, x along the edge, y into the face).
struct DeviceMesh {
    int nV, nF;
    const double* __restrict__ vx;     // 3 nV
    const int32_t* __restrict__ fv;    // 3 nF
    const int32_t* __restrict__ fadj;  // 3 nF
    const double* __restrict__ fgeo;   // 9 nF
    const double* __restrict__ fnrm;   // 3 nF unit normals
    const uint8_t* __restrict__ vflag; // nV: bit0 saddle (angle sum >= 2 pi), bit1 mesh border
    const int32_t* __restrict__ vfoff; // nV + 1
    const int32_t* __restrict__ vflist;// face * 4 + local corner index
    // 18 nF: per face and local edge k, the frames of the two edges a window entering through edge k leaves by:
    // {lL, uxL, uyL, lR, uxR, uyR}, the length and unit direction (in the edge-k frame) of corner k -> corner k+2 and
    // of corner k+2 -> corner k+1
    const double* __restrict__ fchild;
    // 20 nF, for the submesh search: per face a 160-byte record of its three adjacency words (and a pad word) in the
    // first 16 bytes, its corner coordinates (9 doubles) and, per local edge k, the coordinates of the neighbour's corner
    // opposite the shared edge (9 doubles, zero on a border edge). One record is ten 16-byte loads.
    const double* __restrict__ frule;
};

// pair potential, reference conventions (src/forces/harmonicRepulsion.cpp, src/forces/gaussianRepulsion.cpp)
enum PotentialKind : int { POT_HARMONIC = 0, POT_GAUSSIAN = 1, POT_TRUNC_GAUSSIAN = 2, POT_NONE = 3 };

struct Potential {
    int kind;
    double p0, p1, p2;
};

// per-pair flags written by the geodesic kernel
enum PairFlags : uint32_t {
    PAIR_UNREACHED = 1u,     // no path found inside the patch (disconnected) or beyond the truncation radius
    PAIR_BEYOND = 4u,        // shortest path found but longer than the truncation radius
    PAIR_OVERFLOW = 8u,      // the source exceeded a capacity of the kernel variant; solved again by the fallback
    PAIR_SENTINEL = 16u,     // unreachable pair replaced by the reference sentinel (2 * maximumDistance, (0,0,1))
};

__host__ __device__ inline double potentialScale(const Potential& p, double d, double* energy) {
    double e = 0.0, sc = 0.0;
    if (p.kind == POT_HARMONIC) {
        if (d <= p.p1) sc = -p.p0 * (p.p1 - d);
        if (d < p.p1) e = 0.5 * p.p0 * (p.p1 - d) * (p.p1 - d);
    } else if (p.kind == POT_GAUSSIAN) {
        const double sqrt2pi = 2.50662827463100050241576528481104525300698674061;
        double ex = exp(-d * d / (2.0 * p.p1 * p.p1));
        e = p.p0 * ex / (sqrt2pi * p.p1);
        sc = -(d * p.p0 * ex / (sqrt2pi * p.p1 * sqrt(p.p1)));
    } else if (p.kind == POT_TRUNC_GAUSSIAN) {
        if (d < p.p2) {
            double ex = exp(-d * d / (2.0 * p.p1 * p.p1));
            e = p.p0 * ex;
            sc = -p.p0 * (d / (p.p1 * p.p1)) * ex;
        }
    }
    if (energy) *energy = e;
    return sc;
}

// Distance beyond which the potential gives no force and no energy (infinite for the Gaussian, 0 for none).
__host__ __device__ inline double potentialRange(const Potential& p) {
    if (p.kind == POT_HARMONIC) return p.p1;
    if (p.kind == POT_TRUNC_GAUSSIAN) return p.p2;
    if (p.kind == POT_GAUSSIAN) return 1.0 / 0.0;
    return 0.0;
}

// Truncation radius that keeps every force and energy exact. Windows beyond rMax are dropped, so a pair farther than
// rMax ends up unreached and, with the sentinel on, is given the sentinel distance. That changes nothing when the
// potential vanishes both beyond rMax and at the sentinel distance; otherwise truncation is switched off.
inline double exactTruncation(const Potential& p, double rMax, bool useSentinel, double sentinelDist) {
    if (p.kind == POT_NONE) return rMax;
    const double range = potentialRange(p);
    if (rMax < range) rMax = range;
    if (useSentinel && sentinelDist < range) return 1.0 / 0.0;
    return rMax;
}

}  // namespace csg
