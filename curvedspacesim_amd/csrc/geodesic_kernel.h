// Launch interface of the batched geodesic + pair-force kernels (geodesic_kernel.hip).
#pragma once

#include "../../include/csg.h"
#include "csg_common.h"

namespace csg {

// Capacities of the LDS variant (one source per workgroup of kLdsThreads lanes, under 80 KB so that two workgroups
// share a CU): patch faces and vertices, records of the work-item ring, targets per source, and the slots of the
// setup-time face and vertex hashes. A source that exceeds the first, second or fourth, or its spill queue, is solved
// again by the scratch variant.
constexpr int kLdsFaceCap = 640;
constexpr int kLdsVertCap = 384;
constexpr int kLdsQueueCap = 768;
constexpr int kLdsTargetCap = 256;
constexpr int kLdsFaceHash = 1024;
constexpr int kLdsVertHash = 1024;
// work items beyond the ring wait in a per-source HBM stack of this many records
constexpr int kSpillCap = 1024;
constexpr int kXcds = 8;  // accelerator complex dies of an MI355X, each with its own L2
// patch kernel: hash slots per source, capacity of the written face list, and of one breadth-first level
constexpr int kPatchHashCap = 1024;
constexpr int kPatchListCap = 768;
constexpr int kPatchFrontCap = 256;

// reasons a source exceeded the capacities of a kernel variant (bits of the per-source status)
enum OverflowReason : int { OVF_FACES = 1, OVF_VERTS = 2, OVF_QUEUE = 4, OVF_TARGETS = 8, OVF_STEPS = 16, OVF_FRONTIER = 32 };

constexpr int kPatchStats = 7;

// one spilled work item (the fields of a ring record)
struct SpillRecord {
    double b0, b1, sx, sy, sg;
    int32_t fk;
    float tx, ty;
    int32_t pad;
};

struct GeoParams {
    DeviceMesh mesh;
    const int32_t* __restrict__ pface;  // particle faces
    const double* __restrict__ pbary;   // particle barycentric coordinates, 3 per particle
    int nSrc;
    const int32_t* __restrict__ srcIdx;  // particle index of each source (nullptr: source i is particle i)
    const int32_t* __restrict__ toff;    // nSrc + 1 offsets into tidx and into the pair outputs
    const int32_t* __restrict__ tidx;    // particle index of each target
    const uint8_t* __restrict__ keep;    // optional, per pair: 0 = not solved (the learned screen dropped it)
    const double* __restrict__ patchRadius;  // per source; nullptr: whole mesh (no submeshing)
    const int32_t* __restrict__ blockOrder;  // optional: source order (spatially sorted), split over the XCDs in chunks
    int blockChunk;                          // ceil(nSrc / kXcds) when blockOrder is set
    // optional: kXcds counters, zero before the launch. With them each workgroup of the LDS kernel takes its source
    // dynamically, from its own XCD's chunk of the order first, then from the other chunks, so that every XCD stays
    // busy until the last sources (the chunks' costs differ by up to 25%).
    int* grab;
    const int32_t* __restrict__ patchList;   // optional: face sets from the patch kernel, patchStride per source
    const int32_t* __restrict__ patchCount;  // per source: faces listed, or -1 (the solver then searches itself)
    int patchStride;
    SpillRecord* spill;                  // optional: spillCap records per source for work items beyond the LDS ring
    int spillCap;
    int spillPerBlock;                   // 1: spillCap records per workgroup instead (persistent list kernels)
    double rMax;                         // windows beyond this lower-bound distance are not propagated
    // outputs, all optional
    double* pairDist;
    float* pairStart;   // 3 per pair
    float* pairEnd;     // 3 per pair
    uint32_t* pairFlags;
    double* forces;     // 3 per source
    double* energy;     // per source
    int* status;        // per source: 0 solved, else OverflowReason bits
    int* patchStats;    // kPatchStats per source: patch faces, patch vertices, FIFO steps, work items, and the durations
                        // of patch construction, propagation and the epilogue in 10 ns ticks (optional)
    Potential pot;
    int useSentinel;    // unreachable pairs get (sentinelDist, (0,0,1)) as the reference does
    double sentinelDist;
    // optional particle groups: group of each particle; per group the potential, truncation radius and sentinel
    // distance that replace pot, rMax and sentinelDist for the group's sources
    const int32_t* srcGroup;
    const Potential* groupPot;
    const double* groupRMax;
    const double* groupSentinel;
};

struct ScratchLayout {
    void* base;
    size_t bytesPerBlock;
    int fcap, vcap, qcap, tcap;  // faces, vertices, ring records, targets per source
    int hfcap, hvcap;            // face and vertex hash slots (powers of two)
};

// capacities and bytes per workgroup of the scratch variant for sources of up to `faces` patch faces and `targets`
// targets (base is left null)
ScratchLayout makeScratchLayout(long faces, long targets);
// face sets of every source (requires P.patchRadius); stride <= kPatchListCap
hipError_t launchPatch(const GeoParams& P, int32_t* list, int32_t* count, int stride, hipStream_t stream);
// LDS variants: 0 = 256 lanes (4 wavefronts) per source and 1 = 512 lanes (8 wavefronts), both two workgroups per CU
// with the capacities above; 2 = 256 lanes, three workgroups per CU with 576 faces, 352 vertices, a 384-record ring
// and 224 targets; 3 = 256 lanes, four workgroups per CU (40 KB each) with 576 faces, 320 vertices, a 280-record ring
// and 128 targets
constexpr int kLdsVariants = CSG_LDS_VARIANTS;
// The sources flagged for a table capacity (P.status), solved again with larger LDS tables, persistent over a list:
// tier 1 up to 1,024 faces (two workgroups per CU), tier 2 up to 1,536 faces (one per CU, searching the patch itself
// when the patch kernel could not list it), the longest first by their last solve's duration (P.patchStats, when
// given). A list is kTierBuckets * nSrc ints with 1 + kTierBuckets counts, zero before it is built: the counter that
// hands the entries out, then each bucket's length. launchGeodesicLds appends the sources its tables cannot hold to
// the first tier's list next/nextCount (nullable), and counts each bit of their status into reasons[0..6) (nullable);
// each tier appends the sources it cannot hold either to its next/nextCount (nullable). spill: kTierSpillBytes, shared
// by the workgroups (P.spill is not used).
constexpr int kTierBuckets = 8;
constexpr size_t kTierSpillBytes = (size_t)512 * 8192 * sizeof(SpillRecord);
hipError_t launchGeodesicLds(const GeoParams& P, int variant, int32_t* next, int32_t* nextCount, int* reasons, hipStream_t stream);
hipError_t launchGeodesicLdsLarge(const GeoParams& P, int tier, const int32_t* list, int32_t* count, int32_t* next, int32_t* nextCount,
                                  SpillRecord* spill, hipStream_t stream);
// fills mesh.fchild (out, 18 per face) from mesh.fgeo
hipError_t launchChildFrames(const DeviceMesh& mesh, double* out, hipStream_t stream);
hipError_t launchGeodesicScratch(const GeoParams& P, const ScratchLayout& L, const int32_t* list, int nList, int blocks,
                                 hipStream_t stream);
// the same with kWideLanes lanes per source, for whole-mesh solves
constexpr int kWideLanes = 512;
hipError_t launchGeodesicScratchWide(const GeoParams& P, const ScratchLayout& L, const int32_t* list, int nList, int blocks,
                                     hipStream_t stream);
// scratch variant over every source whose P.status is non-zero (left unsolved by the LDS variant); no host round trip
hipError_t launchGeodesicScratchRedo(const GeoParams& P, const ScratchLayout& L, int blocks, hipStream_t stream);

}  // namespace csg
