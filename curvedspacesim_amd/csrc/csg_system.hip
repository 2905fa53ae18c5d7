// Device-resident particle system: the batched form of simpleModel (positions, velocities, forces in HBM), its
// neighbour search, the fused force computation, and the velocity-Verlet and Nose-Hoover-chain integrators.
//
//   cell list      src/utility/hyperRectangularCellList.cpp:9-159, src/utility/cellListNeighborStructure.cpp:22-84
//                  (grid of floor(extent/range) cells per axis over the box seeded at the origin; candidates from the
//                  27 surrounding cells, x outermost then y then z, inside a cell by particle index; |xi-xj|^2 < r^2)
//   all-to-all     src/utility/baseNeighborStructure.cpp:17-36 (every other particle by index; threshold 1e20)
//   forces         src/models/simpleModel.cpp:68-112 + src/forces/baseForce.cpp:12-28, fused (geodesic_kernel.hip)
//   move           src/models/simpleModel.cpp:44-66 (velocity transported) via transport_kernel.hip
//   VV-NVE         src/updaters/velocityVerletNVE.cpp:3-29
//   NH-NVT         src/updaters/noseHooverNVT.cpp:3-150 (chain of M thermostats, tau, Q0 = 2(N-1) T tau^2, Qi = T tau^2;
//                  kinetic energy initialised to Q0 as the reference does)
// Every step is a fixed sequence of kernel launches on the context's stream with no host synchronisation; error
// conditions (candidate capacity, unsolved sources, failed walks) are counted on the device and read by
// csg_system_status().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/csg.h"
#include "csg_common.h"
#include "csg_context.h"
#include "geodesic_kernel.h"
#include "scan_math.h"
#include "screen_kernel.h"
#include "transport_kernel.h"

namespace csg {

// device counters; CNT_LDS_REASON + b counts sources the LDS variant handed to the HBM variant for OverflowReason bit b
enum Counter : int {
    CNT_CAND_OVERFLOW = 0, CNT_UNSOLVED = 1, CNT_WALK_FAIL = 2, CNT_PAIRS = 3, CNT_LDS_REASON = 4, CNT_BAND = 10, CNT_KEPT = 11, CNT_N = 12
};

struct SystemState {
    int N = 0;
    DevBuf face, bary, vel, force, mass, energy;
    DevBuf xyz, sxyz, key, cellCount, cellStart, cellFill, order;
    DevBuf cnt, off, toff, cand, maxd, rad;
    // redo: the step's work counters (the XCD chunks' counters, then the two tiers' list counts) and the tiers' lists
    DevBuf disp, tstatus, gstatus, partial, counters, ke, chain, pstat, plist, pcount, spill, redo, tspill;
    DevBuf scratch;
    DevBuf scanTmp;  // tile sums and offsets of the tiled scan
    // learned screen: features and logits of the candidate pairs, kept-pair lists; recorded pair distances
    DevBuf sfeat, sparams, pdist, skeep;
    // independent particle groups: group of each particle; per group potential, squared range, patch-radius cap,
    // truncation radius and sentinel distance
    int nGroups = 0;
    DevBuf grp, gpot, gr2, gmaxd, grmax, gsent, grange;
    std::vector<Potential> hgpot;
    std::vector<double> hgrange, hgmaxd;
    bool screenOn = false, recordPairs = false;
    int reported[3] = {0, 0, 0};  // error counters (candidate overflow, unsolved, walk failures) at the last check
    int ldsVariant = 0;  // workgroup size of the LDS geodesic kernel, chosen by csg_system_autotune
    // one MD step captured as a HIP graph (replayed per step), and what it was captured for
    bool useGraph = false;
    hipGraphExec_t graph = nullptr;
    double graphKey[6] = {0, 0, 0, 0, 0, 0};
    float screenThreshold = 0.f;
    ScratchLayout layout{};
    int scratchBlocks = 0;
    // the last tier of the step: HBM tables that hold any patch of the configured radius (patchFaceBound), few
    // workgroups, for the sources that outgrow every other table
    DevBuf scratchBig;
    ScratchLayout layoutBig{};
    int bigBlocks = 0;
    double bigRadius = 0;
    // whole-mesh mode with a finite truncation radius: the patch radius that holds every geodesic within it (0: the
    // whole-mesh HBM tables are used)
    double wholePatchRadius = 0;
    // configuration
    Potential pot{POT_NONE, 0, 0, 0};
    double range = 1.0, rMax = HUGE_VAL;
    int cellList = 0;
    int n3[3] = {1, 1, 1};
    double cellSize[3] = {1, 1, 1};
    int nCells = 1;
    long candCap = 0;
    bool configured = false;
    // Nose-Hoover chain
    int M = 0;
    double T = 0, tau = 1;
    bool chainReady = false;
    // optional event timing of the launches, by kind
    bool profiling = false;
    struct Mark {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    // asynchronous snapshots: device positions, and a pinned host copy of faces, barycentrics, positions, velocities
    // and forces written by copies queued on the context stream, with the event that follows them
    DevBuf snapXyz;
    void* snapHost = nullptr;
    size_t snapBytes = 0;
    hipEvent_t snapEvent = nullptr;
    bool snapTaken = false;
};

enum MarkKind : int { MARK_NEIGHBORS = 0, MARK_GEODESIC_LDS = 1, MARK_GEODESIC_HBM = 2, MARK_TRANSPORT = 3, MARK_PATCH = 4, MARK_SCREEN = 5, MARK_N = 6 };

// records an event pair around a group of launches on the context stream when profiling is on
struct MarkScope {
    SystemState* S;
    hipStream_t st;
    int kind;
    hipEvent_t a = nullptr, b = nullptr;
    MarkScope(SystemState* s, hipStream_t stream, int k) : S(s), st(stream), kind(k) {
        if (!S->profiling) return;
        hipEventCreate(&a);
        hipEventCreate(&b);
        hipEventRecord(a, st);
    }
    ~MarkScope() {
        if (!S->profiling) return;
        hipEventRecord(b, st);
        S->marks.push_back({kind, a, b});
    }
};

void destroySystem(SystemState* s) {
    if (s->graph) hipGraphExecDestroy(s->graph);
    DevBuf* all[] = {&s->face, &s->bary, &s->vel, &s->force, &s->mass, &s->energy, &s->xyz, &s->sxyz, &s->key, &s->cellCount, &s->cellStart,
                     &s->cellFill, &s->order, &s->cnt, &s->off, &s->toff, &s->cand, &s->maxd, &s->rad, &s->disp, &s->tstatus,
                     &s->gstatus, &s->partial, &s->counters, &s->ke, &s->chain, &s->pstat, &s->plist, &s->pcount, &s->spill, &s->redo, &s->tspill, &s->scratch, &s->scratchBig,
                     &s->sfeat, &s->sparams, &s->pdist, &s->skeep,
                     &s->grp, &s->gpot, &s->gr2, &s->gmaxd, &s->grmax, &s->gsent, &s->grange};
    for (DevBuf* b : all) b->release();
    s->snapXyz.release();
    if (s->snapHost) hipHostFree(s->snapHost);
    if (s->snapEvent) hipEventDestroy(s->snapEvent);
    for (auto& m : s->marks) {
        hipEventDestroy(m.a);
        hipEventDestroy(m.b);
    }
    delete s;
}

namespace {

constexpr int kBlock = 256;

inline int blocksFor(long n) { return (int)((n + kBlock - 1) / kBlock); }

__global__ void kFillInt(int* p, int v, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void kFillDouble(double* p, double v, long n) {
    long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void kPositions(DeviceMesh M, int N, const int32_t* face, const double* bary, double* xyz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int* c = M.fv + 3 * face[i];
    const double* b = bary + 3 * i;
    double s = b[0] + b[1] + b[2];
    for (int a = 0; a < 3; ++a) xyz[3 * i + a] = (M.vx[3 * c[0] + a] * b[0] + M.vx[3 * c[1] + a] * b[1] + M.vx[3 * c[2] + a] * b[2]) / s;
}

struct Grid {
    double lo[3], size[3];
    int n[3];
};

__device__ inline int cellCoord(const Grid& g, const double* p, int a) {
    int v = (int)floor((p[a] - g.lo[a]) / g.size[a]);
    return max(0, min(g.n[a] - 1, v));
}

__global__ void kCellKeys(Grid g, int N, const double* xyz, int* key, int* cellCount) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double* p = xyz + 3 * i;
    int c = cellCoord(g, p, 0) + g.n[0] * (cellCoord(g, p, 1) + g.n[1] * cellCoord(g, p, 2));
    key[i] = c;
    atomicAdd(&cellCount[c], 1);
}

// stores of a 64-bit running offset: saturated into int32 offset arrays, as is into the int64 tile offsets
__device__ inline void storeOffset(int* p, long long v) { *p = saturateOffset(v); }
__device__ inline void storeOffset(long long* p, long long v) { *p = v; }

// exclusive prefix sum of n counts into out[0..n] (out[n] = total), one workgroup of 1024 lanes: each lane sums a
// contiguous run of ceil(n / 1024) entries, the lanes' sums are scanned across each wavefront with shuffles and across
// the 16 wavefronts through LDS (two barriers in all), and each lane writes its run. Sums are 64-bit.
template <class TIn, class TOut>
__global__ __launch_bounds__(1024) void kScanExclusive(const TIn* in, TOut* out, int n) {
    __shared__ long long wtot[16];
    const int per = (n + 1023) / 1024;
    const int b = min((int)threadIdx.x * per, n), e = min(b + per, n);
    long long s = 0;
    for (int i = b; i < e; ++i) s += in[i];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long long x = s;  // inclusive scan of the lanes' sums within the wavefront
    for (int o = 1; o < 64; o <<= 1) {
        const long long y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wtot[w] = x;
    __syncthreads();
    if (w == 0) {  // inclusive scan of the 16 wavefront totals
        long long t = lane < 16 ? wtot[lane] : 0;
        for (int o = 1; o < 16; o <<= 1) {
            const long long y = __shfl_up(t, o, 64);
            if (lane >= o) t += y;
        }
        if (lane < 16) wtot[lane] = t;
    }
    __syncthreads();
    long long run = x - s + (w > 0 ? wtot[w - 1] : 0);
    for (int i = b; i < e; ++i) {
        const long long v = in[i];
        storeOffset(out + i, run);
        run += v;
    }
    if (threadIdx.x == 0) storeOffset(out + n, wtot[15]);
}

// Tiled exclusive scan for long inputs (cell counts of large boxes, candidate counts of large systems): the tiles'
// sums, the single-workgroup scan above over those sums, then each tile's own scan from its offset.
constexpr int kScanTile = 4096;  // elements per 256-lane workgroup, a run of 16 per lane
static long scanTiles(long n) { return (n + kScanTile - 1) / kScanTile; }

__global__ __launch_bounds__(256) void kTileSums(const int* in, long long* sums, int n) {
    __shared__ long long wsum[4];
    const long base = (long)blockIdx.x * kScanTile;
    long long s = 0;
    for (int k = threadIdx.x; k < kScanTile; k += 256) {
        const long i = base + k;
        if (i < n) s += in[i];
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) sums[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// tileOff[t]: the sum of the tiles before tile t; tileOff[nTiles]: the total, stored at out[n]
__global__ __launch_bounds__(256) void kTileScan(const int* in, int* out, int n, const long long* tileOff, int nTiles) {
    __shared__ long long wtot[4];
    const long b = (long)blockIdx.x * kScanTile + 16 * (long)threadIdx.x;
    int v[16];
    long long s = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        v[k] = b + k < n ? in[b + k] : 0;
        s += v[k];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long long x = s;  // inclusive scan of the lanes' sums within the wavefront
    for (int o = 1; o < 64; o <<= 1) {
        const long long y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wtot[w] = x;
    __syncthreads();
    long long run = tileOff[blockIdx.x] + x - s;
    for (int q = 0; q < w; ++q) run += wtot[q];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        if (b + k < n) out[b + k] = saturateOffset(run);
        run += v[k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = saturateOffset(tileOff[nTiles]);
}

// exclusive scan of in[0, n) into out[0, n), with the total at out[n] (saturating offsets); tmp holds
// 2 * (scanTiles(n) + 1) 64-bit entries
static void scanExclusive(hipStream_t st, const int* in, int* out, int n, long long* tmp) {
    const int tiles = (int)scanTiles(n);
    if (tiles <= 1) {
        hipLaunchKernelGGL((kScanExclusive<int, int>), dim3(1), dim3(1024), 0, st, in, out, n);
        return;
    }
    long long* sums = tmp;
    long long* offs = tmp + tiles + 1;
    hipLaunchKernelGGL(kTileSums, dim3(tiles), dim3(256), 0, st, in, sums, n);
    hipLaunchKernelGGL((kScanExclusive<long long, long long>), dim3(1), dim3(1024), 0, st, (const long long*)sums, offs, tiles);
    hipLaunchKernelGGL(kTileScan, dim3(tiles), dim3(256), 0, st, in, out, n, (const long long*)offs, tiles);
}

__global__ void kScatter(int N, const int* key, const int* cellStart, int* cellFill, int* order) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int c = key[i];
    order[cellStart[c] + atomicAdd(&cellFill[c], 1)] = i;
}

// particles of a cell in increasing index (the reference fills cells in particle order)
__global__ void kSortCells(int nCells, const int* cellStart, int* order) {
    // one wavefront per cell: every element's rank is the number of smaller ones (indices are distinct), found by
    // broadcasting the cell's elements across the lanes; all reads finish before the writes (one wavefront per cell)
    const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (c >= nCells) return;
    const int b = cellStart[c], n = cellStart[c + 1] - b;
    if (n <= 1) return;
    if (n <= 64) {
        const int v = lane < n ? order[b + lane] : 0x7fffffff;
        int rank = 0;
        for (int j = 0; j < n; ++j) rank += __shfl(v, j, 64) < v ? 1 : 0;
        if (lane < n) order[b + rank] = v;
        return;
    }
    if (lane == 0)  // a dense cell: insertion sort by one lane
        for (int i = b + 1; i < b + n; ++i) {
            int v = order[i], j = i - 1;
            while (j >= b && order[j] > v) {
                order[j + 1] = order[j];
                --j;
            }
            order[j + 1] = v;
        }
}

// Candidates of every particle in the reference's order (stencil cells x, then y, then z; ascending index within a
// cell), one wavefront per particle: the lanes test consecutive members of each stencil cell, and a ballot prefix places
// the hits in that order. Wavefront t takes the t-th particle in cell order and reads positions from their cell-ordered
// copy sxyz, contiguous within a cell.
// With particle groups (grp non-null) only particles of the same group are candidates, within that group's range
// (gr2: squared) and with its patch-radius cap (gmaxd).
template <bool kFill>
__global__ void kCandidatesWave(Grid g, int N, const double* sxyz, const int* key, const int* cellStart, const int* order, double r2,
                                int* cnt, const int* off, long cap, int* cand, double* maxd, double* rad, double maxDist, int* counters,
                                const int* grp, const double* gr2, const double* gmaxd) {
    const int t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (t >= N) return;
    const int i = order[t];
    const int gi = grp ? grp[i] : 0;
    if (grp) {
        r2 = gr2[gi];
        maxDist = gmaxd[gi];
    }
    const double p[3] = {sxyz[3 * t], sxyz[3 * t + 1], sxyz[3 * t + 2]};
    const int c = key[i];
    const int cx = c % g.n[0], cy = (c / g.n[0]) % g.n[1], cz = c / (g.n[0] * g.n[1]);
    const unsigned long long below = (1ull << lane) - 1ull;
    long k = 0;
    double m2 = 0;
    const long base = kFill ? off[i] : 0;
    for (int x = max(0, cx - 1); x <= min(g.n[0] - 1, cx + 1); ++x)
        for (int y = max(0, cy - 1); y <= min(g.n[1] - 1, cy + 1); ++y)
            for (int z = max(0, cz - 1); z <= min(g.n[2] - 1, cz + 1); ++z) {
                const int cc = x + g.n[0] * (y + g.n[1] * z);
                const int ce = cellStart[cc + 1];
                for (int q0 = cellStart[cc]; q0 < ce; q0 += 64) {
                    const int q = q0 + lane;
                    bool hit = false;
                    double d2 = 0;
                    if (q < ce && q != t) {
                        const double dx = p[0] - sxyz[3 * q], dy = p[1] - sxyz[3 * q + 1], dz = p[2] - sxyz[3 * q + 2];
                        d2 = dx * dx + dy * dy + dz * dz;
                        hit = d2 < r2 && (grp == nullptr || grp[order[q]] == gi);
                    }
                    const unsigned long long mask = __ballot(hit);
                    if (hit) {
                        const long pos = k + __popcll(mask & below);
                        if (kFill && candidateSlotValid(base, pos, cap)) cand[base + pos] = order[q];
                        if (d2 > m2) m2 = d2;
                    }
                    k += __popcll(mask);
                }
            }
    for (int o = 32; o > 0; o >>= 1) m2 = fmax(m2, __shfl_xor(m2, o, 64));
    if (lane != 0) return;
    if (!kFill) {
        cnt[i] = (int)k;
    } else {
        if (candidatesOverflow(base, k, cap)) atomicAdd(&counters[CNT_CAND_OVERFLOW], 1);
        const double md = sqrt(m2);
        maxd[i] = md;
        rad[i] = md < maxDist ? md : maxDist;
    }
}

// positions in cell order
__global__ void kGatherPositions(int N, const int* order, const double* xyz, double* sxyz) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N) return;
    const int i = order[t];
    for (int a = 0; a < 3; ++a) sxyz[3 * t + a] = xyz[3 * i + a];
}

// all-to-all candidates: every other particle by increasing index
__global__ void kAllToAll(int N, int* cand, int* off, double* maxd, double* rad, double maxDist) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > N) return;
    off[i] = i * (N - 1);
    if (i == N) return;
    int k = 0;
    for (int j = 0; j < N; ++j)
        if (j != i) cand[(long)i * (N - 1) + k++] = j;
    maxd[i] = 1e20;
    rad[i] = 1e20 < maxDist ? 1e20 : maxDist;
}

// offsets clamped to the candidate capacity, so a source never reads past it (its step is then flagged invalid)
__global__ void kClampOffsets(int N, const int* off, long cap, int* toff, int* counters) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > N) return;
    toff[i] = (int)min((long)off[i], cap);
    if (i == N) counters[CNT_PAIRS] = off[N];
}

__global__ void kCountNonzero(int N, const int* v, int* counters, int which, bool negativeOnly) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    bool bad = negativeOnly ? v[i] < 0 : v[i] != 0;
    if (bad) atomicAdd(&counters[which], 1);
}

// velocityVerletNVE first half: displacement and half kick
__global__ void kVVFirst(int N, double dt, double* v, const double* f, double* disp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int a = 0; a < 3; ++a) {
        disp[3 * i + a] = dt * v[3 * i + a] + 0.5 * dt * dt * f[3 * i + a];
        v[3 * i + a] += 0.5 * dt * f[3 * i + a];
    }
}

__global__ void kVVKick(int N, double dt, double* v, const double* f) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int a = 0; a < 3; ++a) v[3 * i + a] += 0.5 * dt * f[3 * i + a];
}

__global__ void kScaledCopy(int N, double s, const double* v, double* disp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * N) return;
    disp[i] = s * v[i];
}

// noseHooverNVT::propagatePositionsVelocities second half: kick, next displacement, per-block kinetic energy
__global__ __launch_bounds__(256) void kNHKick(int N, double dt, double dt2, double* v, const double* f, const double* m, double* disp,
                                               double* partial) {
    __shared__ double red[256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0;
    if (i < N) {
        double mi = m[i];
        double vv = 0;
        for (int a = 0; a < 3; ++a) {
            double x = v[3 * i + a] + (dt / mi) * f[3 * i + a];
            v[3 * i + a] = x;
            disp[3 * i + a] = dt2 * x;
            vv += x * x;
        }
        ke = 0.5 * mi * vv;
    }
    red[threadIdx.x] = ke;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// fixed-order sum of the per-block partials (deterministic)
__global__ void kSumPartials(int n, const double* partial, double* ke) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0;
    for (int i = 0; i < n; ++i) s += partial[i];
    ke[0] = s;
}

// noseHooverNVT::propagateChain; chain holds (x, y, z, w) = (position, velocity, acceleration, mass) for M+1 entries,
// ke[0] the kinetic energy, ke[1] receives the velocity scale factor
__global__ void kNHChain(double* chain, double* ke, int M, double T, double dt) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double dt2 = 0.5 * dt, dt4 = 0.25 * dt, dt8 = 0.125 * dt;
    double4* b = (double4*)chain;
    double K = ke[0];
    for (int ii = M - 1; ii > 0; --ii) {
        b[ii].z = (b[ii - 1].w * b[ii - 1].y * b[ii - 1].y - T) / b[ii].w;
        double ef = exp(-dt8 * b[ii + 1].y);
        b[ii].y *= ef;
        b[ii].y += b[ii].z * dt4;
        b[ii].y *= ef;
    }
    b[0].z = (2.0 * K / b[0].w - 1.0);
    double ef = exp(-dt8 * b[1].y);
    b[0].y *= ef;
    b[0].y += b[0].z * dt4;
    b[0].y *= ef;
    for (int ii = 0; ii < M; ++ii) b[ii].x += dt2 * b[ii].y;
    double s = exp(-dt2 * b[0].y);
    K = s * s * K;
    b[0].z = (2.0 * K / b[0].w - 1.0);
    ef = exp(-dt8 * b[1].y);
    b[0].y *= ef;
    b[0].y += b[0].z * dt4;
    b[0].y *= ef;
    for (int ii = 1; ii < M; ++ii) {
        b[ii].z = (b[ii - 1].w * b[ii - 1].y * b[ii - 1].y - T) / b[ii].w;
        ef = exp(-dt8 * b[ii + 1].y);
        b[ii].y *= ef;
        b[ii].y += b[ii].z * dt4;
        b[ii].y *= ef;
    }
    ke[0] = K;
    ke[1] = s;
}

__global__ void kScaleVel(int N, double* v, const double* ke) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * N) return;
    v[i] *= ke[1];
}

}  // namespace

// ---- step pieces (all asynchronous on the context stream) ------------------------------------------------------

static csg_status sysNeighbors(csg_ctx* c, SystemState* S) {
    hipStream_t st = c->stream;
    MarkScope mark(S, st, MARK_NEIGHBORS);
    const int N = S->N;
    if (!S->cellList) {
        hipLaunchKernelGGL(kAllToAll, dim3(blocksFor(N + 1)), dim3(kBlock), 0, st, N, S->cand.as<int>(), S->off.as<int>(),
                           S->maxd.as<double>(), S->rad.as<double>(), c->maximumDistance);
        hipLaunchKernelGGL(kClampOffsets, dim3(blocksFor(N + 1)), dim3(kBlock), 0, st, N, S->off.as<int>(), S->candCap, S->toff.as<int>(),
                           S->counters.as<int>());
        HIPCHECK(c, hipGetLastError());
        return CSG_OK;
    }
    Grid g;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = c->lo[a];
        g.size[a] = S->cellSize[a];
        g.n[a] = S->n3[a];
    }
    hipLaunchKernelGGL(kPositions, dim3(blocksFor(N)), dim3(kBlock), 0, st, c->dm, N, S->face.as<int32_t>(), S->bary.as<double>(),
                       S->xyz.as<double>());
    hipLaunchKernelGGL(kFillInt, dim3(blocksFor(S->nCells)), dim3(kBlock), 0, st, S->cellCount.as<int>(), 0, (long)S->nCells);
    hipLaunchKernelGGL(kFillInt, dim3(blocksFor(S->nCells)), dim3(kBlock), 0, st, S->cellFill.as<int>(), 0, (long)S->nCells);
    hipLaunchKernelGGL(kCellKeys, dim3(blocksFor(N)), dim3(kBlock), 0, st, g, N, S->xyz.as<double>(), S->key.as<int>(),
                       S->cellCount.as<int>());
    scanExclusive(st, S->cellCount.as<int>(), S->cellStart.as<int>(), S->nCells, S->scanTmp.as<long long>());
    hipLaunchKernelGGL(kScatter, dim3(blocksFor(N)), dim3(kBlock), 0, st, N, S->key.as<int>(), S->cellStart.as<int>(), S->cellFill.as<int>(),
                       S->order.as<int>());
    hipLaunchKernelGGL(kSortCells, dim3(blocksFor(64L * S->nCells)), dim3(kBlock), 0, st, S->nCells, S->cellStart.as<int>(), S->order.as<int>());
    hipLaunchKernelGGL(kGatherPositions, dim3(blocksFor(N)), dim3(kBlock), 0, st, N, S->order.as<int>(), S->xyz.as<double>(),
                       S->sxyz.as<double>());
    double r2 = S->range * S->range;
    const int* grp = S->nGroups ? S->grp.as<int>() : nullptr;
    const double* gr2 = S->nGroups ? S->gr2.as<double>() : nullptr;
    const double* gmaxd = S->nGroups ? S->gmaxd.as<double>() : nullptr;
    hipLaunchKernelGGL((kCandidatesWave<false>), dim3(blocksFor(64L * N)), dim3(kBlock), 0, st, g, N, S->sxyz.as<double>(), S->key.as<int>(),
                       S->cellStart.as<int>(), S->order.as<int>(), r2, S->cnt.as<int>(), (const int*)nullptr, S->candCap, (int*)nullptr,
                       (double*)nullptr, (double*)nullptr, c->maximumDistance, S->counters.as<int>(), grp, gr2, gmaxd);
    scanExclusive(st, S->cnt.as<int>(), S->off.as<int>(), N, S->scanTmp.as<long long>());
    hipLaunchKernelGGL((kCandidatesWave<true>), dim3(blocksFor(64L * N)), dim3(kBlock), 0, st, g, N, S->sxyz.as<double>(), S->key.as<int>(),
                       S->cellStart.as<int>(), S->order.as<int>(), r2, S->cnt.as<int>(), S->off.as<int>(), S->candCap, S->cand.as<int>(),
                       S->maxd.as<double>(), S->rad.as<double>(), c->maximumDistance, S->counters.as<int>(), grp, gr2, gmaxd);
    hipLaunchKernelGGL(kClampOffsets, dim3(blocksFor(N + 1)), dim3(kBlock), 0, st, N, S->off.as<int>(), S->candCap, S->toff.as<int>(),
                       S->counters.as<int>());
    HIPCHECK(c, hipGetLastError());
    return CSG_OK;
}

// feature inputs of the candidate pairs of the last neighbour search (the pair count is read on the device)
static ScreenFeatureParams featureParams(csg_ctx* c, SystemState* S) {
    ScreenFeatureParams F{};
    F.mesh = c->dm;
    F.curv = c->curv.as<double>();
    F.face = S->face.as<int32_t>();
    F.bary = S->bary.as<double>();
    F.xyz = S->xyz.as<double>();
    F.nSrc = S->N;
    F.toff = S->toff.as<int32_t>();
    F.tidx = S->cand.as<int32_t>();
    F.patchCount = c->submeshing ? S->pcount.as<int32_t>() : nullptr;
    F.cutoff = S->range;
    if (S->nGroups) {
        F.srcGroup = S->grp.as<int32_t>();
        F.groupRange = S->grange.as<double>();
    }
    F.nPairsDev = S->toff.as<int32_t>() + S->N;
    F.out = S->sfeat.as<float>();
    return F;
}

// features of every candidate pair of the last neighbour search
static void screenFeatures(csg_ctx* c, SystemState* S) {
    (void)launchScreenFeatures(featureParams(c, S), (int)std::min<long>(S->candCap, 0x7fffffff), c->stream);
}

// screen the candidate pairs: a keep flag per pair, which the solver reads (patches are still built from every
// candidate), and the kept count
static csg_status sysScreen(csg_ctx* c, SystemState* S, GeoParams& G) {
    hipStream_t st = c->stream;
    // counters CNT_BAND (pairs scored) and CNT_KEPT, adjacent, start at zero
    hipLaunchKernelGGL(kFillInt, dim3(1), dim3(kBlock), 0, st, S->counters.as<int>() + CNT_BAND, 0, 2L);
    HIPCHECK(c, launchScreenSources(featureParams(c, S), S->sparams.as<float>(), S->screenThreshold, S->skeep.as<uint8_t>(),
                                    S->counters.as<int>() + CNT_BAND, st));
    G.keep = S->skeep.as<uint8_t>();
    return CSG_OK;
}

// Whole-mesh mode solved over patches: a geodesic no longer than the truncation radius r stays within Euclidean
// distance r of its source, so every face it crosses has a corner within r plus the longest edge, and the submesh
// rule with that radius holds every such path; distances up to r are those of the whole mesh (pairs beyond r carry
// no force or energy, and windows beyond r are not propagated in either form). No sentinel, as in whole-mesh mode.
static bool wholeAsPatches(const csg_ctx* c, const SystemState* S) { return !c->submeshing && S->nGroups == 0 && S->wholePatchRadius > 0; }

// parameters of the force computation of the current step (the kept lists when the screen is on)
static GeoParams forceParams(csg_ctx* c, SystemState* S) {
    const bool patches = c->submeshing || wholeAsPatches(c, S);
    GeoParams G{};
    G.mesh = c->dm;
    G.pface = S->face.as<int32_t>();
    G.pbary = S->bary.as<double>();
    G.nSrc = S->N;
    G.srcIdx = nullptr;
    G.toff = S->toff.as<int32_t>();
    G.tidx = S->cand.as<int32_t>();
    G.patchRadius = patches ? S->rad.as<double>() : nullptr;
    G.rMax = S->rMax;
    G.forces = S->force.as<double>();
    G.energy = S->energy.as<double>();
    G.status = S->gstatus.as<int>();
    G.pot = S->pot;
    G.useSentinel = c->submeshing;
    G.sentinelDist = 2.0 * c->maximumDistance;
    G.rMax = exactTruncation(S->pot, S->rMax, G.useSentinel != 0, G.sentinelDist);
    G.patchStats = S->pstat.as<int>();
    G.pairDist = S->recordPairs ? S->pdist.as<double>() : nullptr;
    if (S->nGroups) {  // per-group potential, truncation radius and sentinel
        G.srcGroup = S->grp.as<int32_t>();
        G.groupPot = S->gpot.as<Potential>();
        G.groupRMax = S->grmax.as<double>();
        G.groupSentinel = S->gsent.as<double>();
    }
    if (S->cellList) {  // sources in cell order, one contiguous chunk per XCD
        G.blockOrder = S->order.as<int32_t>();
        G.blockChunk = (S->N + kXcds - 1) / kXcds;
    }
    if (patches) {
        G.patchList = S->plist.as<int32_t>();
        G.patchCount = S->pcount.as<int32_t>();
        G.patchStride = kPatchListCap;
        G.spill = S->spill.as<SpillRecord>();
        G.spillCap = kSpillCap;
    }
    if (S->screenOn && c->submeshing) G.keep = S->skeep.as<uint8_t>();
    return G;
}

static csg_status sysForces(csg_ctx* c, SystemState* S) {
    csg_status s = sysNeighbors(c, S);
    if (s) return s;
    GeoParams G = forceParams(c, S);
    G.keep = nullptr;  // set by the screen once this step's flags exist
    if (c->submeshing || wholeAsPatches(c, S)) {
        if (!c->submeshing)  // every source's patch radius: the truncation radius plus the longest edge
            hipLaunchKernelGGL(kFillDouble, dim3(blocksFor(S->N)), dim3(kBlock), 0, c->stream, S->rad.as<double>(), S->wholePatchRadius,
                               (long)S->N);
        {
            MarkScope mark(S, c->stream, MARK_PATCH);
            HIPCHECK(c, launchPatch(G, S->plist.as<int32_t>(), S->pcount.as<int32_t>(), kPatchListCap, c->stream));
        }
        if (S->screenOn) {  // the solver solves only the pairs the screen keeps; patches keep every candidate's face
            MarkScope mark(S, c->stream, MARK_SCREEN);
            csg_status s2 = sysScreen(c, S, G);
            if (s2) return s2;
        }
        // the step's counters, cleared by one kernel: the XCD chunks' (dynamic distribution of the sources over the
        // XCDs' chunks), then each tier list's counts; then the lists
        int32_t* grab = S->redo.as<int32_t>();
        int32_t* count1 = grab + kXcds;
        int32_t* count2 = count1 + 1 + kTierBuckets;
        int32_t* list1 = count2 + 1 + kTierBuckets;
        int32_t* list2 = list1 + (size_t)kTierBuckets * S->N;
        {
            MarkScope mark(S, c->stream, MARK_GEODESIC_LDS);
            G.grab = grab;
            hipLaunchKernelGGL(kFillInt, dim3(1), dim3(kBlock), 0, c->stream, grab, 0, (long)(kXcds + 2 * (1 + kTierBuckets)));
            // the sources these tables cannot hold go on to the first tier's list, counted by reason
            HIPCHECK(c, launchGeodesicLds(G, S->ldsVariant, list1, count1, S->counters.as<int>() + CNT_LDS_REASON, c->stream));
            G.grab = nullptr;
        }
        MarkScope mark(S, c->stream, MARK_GEODESIC_HBM);
        // the sources the LDS variant could not hold: two tiers of larger LDS tables, each handing on what it cannot
        // hold either, then HBM tables for what remains
        HIPCHECK(c, launchGeodesicLdsLarge(G, 1, list1, count1, list2, count2, S->tspill.as<SpillRecord>(), c->stream));
        HIPCHECK(c, launchGeodesicLdsLarge(G, 2, list2, count2, nullptr, nullptr, S->tspill.as<SpillRecord>(), c->stream));
        HIPCHECK(c, launchGeodesicScratchRedo(G, S->layout, S->scratchBlocks, c->stream));
        if (S->bigBlocks) HIPCHECK(c, launchGeodesicScratchRedo(G, S->layoutBig, S->bigBlocks, c->stream));
    } else {
        MarkScope mark(S, c->stream, MARK_GEODESIC_HBM);
        HIPCHECK(c, launchGeodesicScratchWide(G, S->layout, nullptr, S->N, S->scratchBlocks, c->stream));
    }
    hipLaunchKernelGGL(kCountNonzero, dim3(blocksFor(S->N)), dim3(kBlock), 0, c->stream, S->N, S->gstatus.as<int>(), S->counters.as<int>(),
                       (int)CNT_UNSOLVED, false);
    HIPCHECK(c, hipGetLastError());
    return CSG_OK;
}

static csg_status sysMove(csg_ctx* c, SystemState* S) {
    TransportParams T{};
    T.mesh = c->dm;
    T.n = S->N;
    T.face = S->face.as<int32_t>();
    T.bary = S->bary.as<double>();
    T.disp = S->disp.as<double>();
    T.scale = 1.0;
    T.dispOut = nullptr;
    T.nVec = 1;
    T.vecs = S->vel.as<double>();
    T.status = S->tstatus.as<int32_t>();
    T.maxCrossings = 100000;
    T.boundary = c->boundary;
    {
        MarkScope mark(S, c->stream, MARK_TRANSPORT);
        HIPCHECK(c, launchTransport(T, c->stream));
    }
    hipLaunchKernelGGL(kCountNonzero, dim3(blocksFor(S->N)), dim3(kBlock), 0, c->stream, S->N, S->tstatus.as<int>(), S->counters.as<int>(),
                       (int)CNT_WALK_FAIL, true);
    HIPCHECK(c, hipGetLastError());
    return CSG_OK;
}

}  // namespace csg

using namespace csg;

static SystemState* sysOf(csg_ctx* c) { return c ? c->sys : nullptr; }

// velocityVerletNVE::performUpdate (velocityVerletNVE.cpp:3-29): displacement and half kick, move, forces, half kick
static csg_status stepNve(csg_ctx* c, SystemState* S, double dt) {
    const int N = S->N;
    hipLaunchKernelGGL(kVVFirst, dim3(blocksFor(N)), dim3(kBlock), 0, c->stream, N, dt, S->vel.as<double>(), S->force.as<double>(),
                       S->disp.as<double>());
    csg_status s = sysMove(c, S);
    if (s) return s;
    s = sysForces(c, S);
    if (s) return s;
    hipLaunchKernelGGL(kVVKick, dim3(blocksFor(N)), dim3(kBlock), 0, c->stream, N, dt, S->vel.as<double>(), S->force.as<double>());
    HIPCHECK(c, hipGetLastError());
    return CSG_OK;
}

// noseHooverNVT::performUpdate (noseHooverNVT.cpp:42-139): chain half step, half move, forces, kick, half move, chain
static csg_status stepNvt(csg_ctx* c, SystemState* S, double dt, double T, int M) {
    const int N = S->N;
    const double dt2 = 0.5 * dt;
    const int nb = blocksFor(N);
    hipLaunchKernelGGL(kNHChain, dim3(1), dim3(1), 0, c->stream, S->chain.as<double>(), S->ke.as<double>(), M, T, dt);
    hipLaunchKernelGGL(kScaleVel, dim3(blocksFor(3 * N)), dim3(kBlock), 0, c->stream, N, S->vel.as<double>(), S->ke.as<double>());
    hipLaunchKernelGGL(kScaledCopy, dim3(blocksFor(3 * N)), dim3(kBlock), 0, c->stream, N, dt2, S->vel.as<double>(), S->disp.as<double>());
    csg_status s = sysMove(c, S);
    if (s) return s;
    s = sysForces(c, S);
    if (s) return s;
    hipLaunchKernelGGL(kNHKick, dim3(nb), dim3(kBlock), 0, c->stream, N, dt, dt2, S->vel.as<double>(), S->force.as<double>(),
                       S->mass.as<double>(), S->disp.as<double>(), S->partial.as<double>());
    hipLaunchKernelGGL(kSumPartials, dim3(1), dim3(1), 0, c->stream, nb, S->partial.as<double>(), S->ke.as<double>());
    s = sysMove(c, S);
    if (s) return s;
    hipLaunchKernelGGL(kNHChain, dim3(1), dim3(1), 0, c->stream, S->chain.as<double>(), S->ke.as<double>(), M, T, dt);
    hipLaunchKernelGGL(kScaleVel, dim3(blocksFor(3 * N)), dim3(kBlock), 0, c->stream, N, S->vel.as<double>(), S->ke.as<double>());
    HIPCHECK(c, hipGetLastError());
    return CSG_OK;
}

static void dropGraph(SystemState* S) {
    if (S->graph) hipGraphExecDestroy(S->graph);
    S->graph = nullptr;
}

// nSteps of `step`: launched directly, or, in graph mode (and without event timing, whose per-launch events are not
// captured), as replays of one step captured into a HIP graph for these parameters. The step launches only kernels
// (no memset or copy nodes) with arguments fixed for given parameters, so every replay repeats the same work.
template <class Step>
static csg_status runSteps(csg_ctx* c, SystemState* S, const double key[6], int nSteps, Step step) {
    if (!S->useGraph || S->profiling || nSteps < 1) {
        for (int i = 0; i < nSteps; ++i) {
            csg_status s = step();
            if (s) return s;
        }
        return CSG_OK;
    }
    if (!S->graph || !std::equal(key, key + 6, S->graphKey)) {
        dropGraph(S);
        HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        csg_status s = step();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (s) {
            if (g) hipGraphDestroy(g);
            return s;
        }
        if (e != hipSuccess) return fail(c, CSG_ERR_DEVICE, std::string("step capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&S->graph, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) {
            S->graph = nullptr;
            return fail(c, CSG_ERR_DEVICE, std::string("step graph: ") + hipGetErrorString(e));
        }
        std::copy(key, key + 6, S->graphKey);
    }
    for (int i = 0; i < nSteps; ++i) HIPCHECK(c, hipGraphLaunch(S->graph, c->stream));
    return CSG_OK;
}

// Sizes the step's last tier for patches of radius `radius` (the largest of the maximum distances and ranges in use):
// tables for patchFaceBound faces and up to 32,768 targets, in as many workgroups (8 .. 256) as 2 GB hold.
static csg_status sizeLastTier(csg_ctx* c, SystemState* S, double radius) {
    if (!c->submeshing) {
        S->bigBlocks = 0;
        return CSG_OK;
    }
    const long faces = std::max<long>(patchFaceBound(c, radius), S->layout.fcap);
    if (faces <= S->layout.fcap) {  // the fixed tier holds every patch already
        S->bigBlocks = 0;
        S->bigRadius = radius;
        return CSG_OK;
    }
    ScratchLayout L = makeScratchLayout(faces, std::min<long>(std::max<long>(S->candCap, 1), 32768));
    const int blocks = (int)std::max<long>(8, std::min<long>(256, (2L << 30) / (long)L.bytesPerBlock));
    HIPCHECK(c, S->scratchBig.ensure(L.bytesPerBlock * (size_t)blocks));
    L.base = S->scratchBig.p;
    S->layoutBig = L;
    S->bigBlocks = blocks;
    S->bigRadius = radius;
    return CSG_OK;
}

// Sources that every tier of the last force computation left unsolved (nonzero status: a patch, target list or
// work-item volume beyond the step's fixed HBM tables), solved again with HBM tables that grow until each fits, as
// csg_pair_forces does. Synchronous; the unsolved counter drops by the sources recovered.
static csg_status resolveLeftovers(csg_ctx* c, SystemState* S) {
    int cnt[CNT_N];
    HIPCHECK(c, hipMemcpyAsync(cnt, S->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (cnt[CNT_UNSOLVED] <= S->reported[CNT_UNSOLVED]) return CSG_OK;
    const int N = S->N;
    GeoParams G = forceParams(c, S);
    std::vector<int> st(N), off(N + 1);
    HIPCHECK(c, hipMemcpyAsync(st.data(), G.status, sizeof(int) * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(off.data(), G.toff, sizeof(int) * (N + 1), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    std::vector<int> redo;
    long maxT = 1;
    for (int i = 0; i < N; ++i)
        if (st[i]) {
            redo.push_back(i);
            maxT = std::max<long>(maxT, off[i + 1] - off[i]);
        }
    const int before = (int)redo.size();
    long faces = c->submeshing ? std::max<long>(2L * S->layout.fcap, 4096) : c->nF;
    for (int attempt = 0; attempt < 6 && !redo.empty(); ++attempt, faces *= 4) {
        faces = std::min<long>(faces, c->nF);
        const int blocks = std::min<int>((int)redo.size(), 512);
        ScratchLayout L = makeScratchLayout(faces, maxT);
        HIPCHECK(c, c->scratch.ensure(L.bytesPerBlock * (size_t)blocks));
        L.base = c->scratch.p;
        HIPCHECK(c, c->list.ensure(sizeof(int) * redo.size()));
        HIPCHECK(c, hipMemcpyAsync(c->list.p, redo.data(), sizeof(int) * redo.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, launchGeodesicScratch(G, L, c->list.as<int32_t>(), (int)redo.size(), blocks, c->stream));
        HIPCHECK(c, hipMemcpyAsync(st.data(), G.status, sizeof(int) * N, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        std::vector<int> next;
        for (int i : redo)
            if (st[i]) next.push_back(i);
        redo.swap(next);
        if (faces >= c->nF) break;
    }
    const int left = cnt[CNT_UNSOLVED] - (before - (int)redo.size());
    HIPCHECK(c, hipMemcpyAsync(S->counters.as<int>() + CNT_UNSOLVED, &left, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return CSG_OK;
}

// Reports what the error counters gained since the last check (csg_system_init starts them at zero): failed walks
// as CSG_ERR_WALK (those particles keep their positions), candidate-capacity overflows and unsolved sources as
// CSG_ERR_CAPACITY (an overflowed source lost candidates; an unsolved source got zero force). Waits for the stream.
static csg_status checkCounters(csg_ctx* c, SystemState* S) {
    int cnt[CNT_N];
    HIPCHECK(c, hipMemcpyAsync(cnt, S->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    int d[3];
    for (int k = 0; k < 3; ++k) {
        d[k] = cnt[k] - S->reported[k];
        S->reported[k] = cnt[k];
    }
    if (d[CNT_WALK_FAIL] > 0)
        return fail(c, CSG_ERR_WALK, std::to_string(d[CNT_WALK_FAIL]) + " particle walks failed since the last check (those particles kept their positions)");
    if (d[CNT_CAND_OVERFLOW] > 0)
        return fail(c, CSG_ERR_CAPACITY, std::to_string(d[CNT_CAND_OVERFLOW]) + " sources had more candidates than the candidate capacity since the last check");
    if (d[CNT_UNSOLVED] > 0)
        return fail(c, CSG_ERR_CAPACITY, std::to_string(d[CNT_UNSOLVED]) +
                                             " sources exceeded the step's geodesic table capacity since the last check and got zero force "
                                             "(raise maxPatchFaces of csg_system_configure, or compute forces with csg_system_compute_forces, which "
                                             "grows its tables)");
    return CSG_OK;
}

extern "C" {

csg_status csg_system_init(csg_ctx* c, int32_t N, const int32_t* face, const double* bary, const double* vel, const double* mass) {
    if (!c) return CSG_ERR_ARGUMENT;
    if (!c->haveMesh) return fail(c, CSG_ERR_STATE, "no mesh uploaded");
    if (N <= 1 || !face || !bary) return fail(c, CSG_ERR_ARGUMENT, "need at least two particles with positions");
    for (int i = 0; i < N; ++i)
        if (face[i] < 0 || face[i] >= c->nF) return fail(c, CSG_ERR_ARGUMENT, "particle face out of range");
    HIPCHECK(c, hipSetDevice(c->device));
    if (!c->sys) c->sys = new SystemState();
    SystemState* S = c->sys;
    if (S->snapTaken) HIPCHECK(c, hipEventSynchronize(S->snapEvent));  // a snapshot of the previous system is dropped
    S->snapTaken = false;
    S->N = N;
    S->configured = false;
    S->chainReady = false;
    std::vector<double> zeros(3 * (size_t)N, 0.0), ones(N, 1.0);
    HIPCHECK(c, S->face.ensure(4 * (size_t)N));
    HIPCHECK(c, S->bary.ensure(24 * (size_t)N));
    HIPCHECK(c, S->vel.ensure(24 * (size_t)N));
    HIPCHECK(c, S->force.ensure(24 * (size_t)N));
    HIPCHECK(c, S->mass.ensure(8 * (size_t)N));
    HIPCHECK(c, hipMemcpyAsync(S->face.p, face, 4 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(S->bary.p, bary, 24 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(S->vel.p, vel ? vel : zeros.data(), 24 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(S->force.p, zeros.data(), 24 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(S->mass.p, mass ? mass : ones.data(), 8 * (size_t)N, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, S->energy.ensure(8 * (size_t)N));
    HIPCHECK(c, S->xyz.ensure(24 * (size_t)N));
    HIPCHECK(c, S->sxyz.ensure(24 * (size_t)N));
    HIPCHECK(c, S->key.ensure(4 * (size_t)N));
    HIPCHECK(c, S->order.ensure(4 * (size_t)N));
    HIPCHECK(c, S->cnt.ensure(4 * ((size_t)N + 1)));
    HIPCHECK(c, S->off.ensure(4 * ((size_t)N + 1)));
    HIPCHECK(c, S->toff.ensure(4 * ((size_t)N + 1)));
    HIPCHECK(c, S->maxd.ensure(8 * (size_t)N));
    HIPCHECK(c, S->rad.ensure(8 * (size_t)N));
    HIPCHECK(c, S->disp.ensure(24 * (size_t)N));
    HIPCHECK(c, S->tstatus.ensure(4 * (size_t)N));
    HIPCHECK(c, S->gstatus.ensure(4 * (size_t)N));
    HIPCHECK(c, S->pstat.ensure(4 * kPatchStats * (size_t)N));
    HIPCHECK(c, hipMemsetAsync(S->pstat.p, 0, 4 * kPatchStats * (size_t)N, c->stream));  // no solve has a duration yet
    HIPCHECK(c, S->plist.ensure(4 * (size_t)kPatchListCap * N));
    HIPCHECK(c, S->pcount.ensure(4 * (size_t)N));
    HIPCHECK(c, S->spill.ensure(sizeof(SpillRecord) * (size_t)kSpillCap * N));
    HIPCHECK(c, S->redo.ensure(sizeof(int) * (kXcds + 2 * (1 + kTierBuckets) + 2 * (size_t)kTierBuckets * N)));
    HIPCHECK(c, S->tspill.ensure(kTierSpillBytes));
    HIPCHECK(c, S->partial.ensure(8 * (size_t)blocksFor(N)));
    HIPCHECK(c, S->counters.ensure(4 * CNT_N));
    HIPCHECK(c, S->ke.ensure(16));
    HIPCHECK(c, hipMemsetAsync(S->counters.p, 0, 4 * CNT_N, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int& r : S->reported) r = 0;
    return CSG_OK;
}

csg_status csg_system_configure(csg_ctx* c, const csg_potential* pot, double range, int32_t cellList, double rMax, int32_t maxPatchFaces) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    if (!(range > 0)) return fail(c, CSG_ERR_ARGUMENT, "range must be positive");
    HIPCHECK(c, hipSetDevice(c->device));
    dropGraph(S);  // buffers may move
    S->pot = pot ? Potential{pot->kind, pot->p0, pot->p1, pot->p2} : Potential{POT_NONE, 0, 0, 0};
    S->range = range;
    S->nGroups = 0;  // particle groups are set again after a new configuration
    S->cellList = cellList ? 1 : 0;
    S->rMax = rMax;
    const int N = S->N;
    if (S->cellList) {
        long cells = 1;
        for (int a = 0; a < 3; ++a) {
            double ext = c->hi[a] - c->lo[a];
            S->n3[a] = std::max(1, (int)std::floor(ext / range));
            S->cellSize[a] = ext / S->n3[a];
            cells *= S->n3[a];
        }
        if (cells > (1L << 24)) return fail(c, CSG_ERR_CAPACITY, "more than 2^24 cells: interaction range too small for the mesh box");
        S->nCells = (int)cells;
        HIPCHECK(c, S->cellCount.ensure(4 * (size_t)cells));
        HIPCHECK(c, S->cellFill.ensure(4 * (size_t)cells));
        HIPCHECK(c, S->cellStart.ensure(4 * ((size_t)cells + 1)));
        // capacity: 8x the mean expected candidates plus 64 per particle
        double dens = N / std::max(c->area, 1e-300);
        long per = (long)(8.0 * dens * M_PI * range * range) + 64;
        S->candCap = std::min<long>((long)N * std::min<long>(per, N - 1), (1L << 31) - 1);
    } else {
        S->candCap = (long)N * (N - 1);
        if (S->candCap >= (1L << 31)) return fail(c, CSG_ERR_CAPACITY, "all-to-all pair count exceeds 2^31");
    }
    HIPCHECK(c, S->cand.ensure(4 * (size_t)std::max<long>(S->candCap, 1)));
    HIPCHECK(c, S->scanTmp.ensure(8 * 2 * ((size_t)scanTiles(std::max<long>(N, S->cellList ? S->nCells : 0)) + 1)));
    // HBM scratch for sources the LDS tables cannot hold, or for whole-mesh solves
    // sources that outgrow the LDS tables are rare (a few per force evaluation); many narrow workgroups take them
    long faces = c->submeshing ? std::max<int>(maxPatchFaces > 0 ? maxPatchFaces : 1536, 1024) : c->nF;
    long maxT = c->submeshing && S->cellList ? std::min<long>(S->candCap, 4096) : N;
    int blocks = c->submeshing ? 1024 : std::min(N, 256);
    ScratchLayout L = makeScratchLayout(faces, maxT);
    HIPCHECK(c, S->scratch.ensure(L.bytesPerBlock * (size_t)blocks));
    L.base = S->scratch.p;
    S->layout = L;
    S->scratchBlocks = blocks;
    csg_status st = sizeLastTier(c, S, std::max(c->maximumDistance, range));
    if (st) return st;
    S->wholePatchRadius = 0;
    if (!c->submeshing) {  // whole-mesh mode: patches when the truncation radius is finite and they fit the LDS tiers
        const double r = exactTruncation(S->pot, S->rMax, false, 0.0);
        if (std::isfinite(r) && r > 0 && patchFaceBound(c, r + c->maxEdge) <= 1536) S->wholePatchRadius = r + c->maxEdge;
    }
    S->configured = true;
    return CSG_OK;
}

csg_status csg_system_compute_forces(csg_ctx* c, double* forces, double* energy) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S || !S->configured) return fail(c, CSG_ERR_STATE, "csg_system_init and csg_system_configure first");
    HIPCHECK(c, hipSetDevice(c->device));
    csg_status s = sysForces(c, S);
    if (s) return s;
    s = resolveLeftovers(c, S);
    if (s) return s;
    s = checkCounters(c, S);
    if (s) return s;
    if (forces) HIPCHECK(c, hipMemcpyAsync(forces, S->force.p, 24 * (size_t)S->N, hipMemcpyDeviceToHost, c->stream));
    if (energy) {
        std::vector<double> e(S->N);
        HIPCHECK(c, hipMemcpyAsync(e.data(), S->energy.p, 8 * (size_t)S->N, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        double t = 0;
        for (double x : e) t += x;
        *energy = t;
    }
    if (forces) HIPCHECK(c, hipStreamSynchronize(c->stream));
    return CSG_OK;
}

csg_status csg_system_move(csg_ctx* c, const double* disp, int32_t transportVelocity) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    if (!disp) return fail(c, CSG_ERR_ARGUMENT, "null displacements");
    HIPCHECK(c, hipSetDevice(c->device));
    HIPCHECK(c, hipMemcpyAsync(S->disp.p, disp, 24 * (size_t)S->N, hipMemcpyHostToDevice, c->stream));
    TransportParams T{};
    T.mesh = c->dm;
    T.n = S->N;
    T.face = S->face.as<int32_t>();
    T.bary = S->bary.as<double>();
    T.disp = S->disp.as<double>();
    T.scale = 1.0;
    T.nVec = transportVelocity ? 1 : 0;
    T.vecs = S->vel.as<double>();
    T.status = S->tstatus.as<int32_t>();
    T.maxCrossings = 100000;
    T.boundary = c->boundary;
    HIPCHECK(c, launchTransport(T, c->stream));
    hipLaunchKernelGGL(kCountNonzero, dim3(blocksFor(S->N)), dim3(kBlock), 0, c->stream, S->N, S->tstatus.as<int>(), S->counters.as<int>(),
                       (int)CNT_WALK_FAIL, true);
    return checkCounters(c, S);
}

csg_status csg_system_nve(csg_ctx* c, double dt, int32_t nSteps) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S || !S->configured) return fail(c, CSG_ERR_STATE, "csg_system_init and csg_system_configure first");
    HIPCHECK(c, hipSetDevice(c->device));
    const double key[6] = {1, dt, 0, 0, 0, (double)S->ldsVariant};
    return runSteps(c, S, key, nSteps, [=]() { return stepNve(c, S, dt); });
}

csg_status csg_system_nvt(csg_ctx* c, double dt, double T, double tau, int32_t M, int32_t nSteps) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S || !S->configured) return fail(c, CSG_ERR_STATE, "csg_system_init and csg_system_configure first");
    if (M < 1) return fail(c, CSG_ERR_ARGUMENT, "chain length must be at least 1");
    HIPCHECK(c, hipSetDevice(c->device));
    const int N = S->N;
    if (!S->chainReady || S->M != M || S->T != T || S->tau != tau) {
        // noseHooverNVT::setBathVariables (noseHooverNVT.cpp:28-36): Q0 = 2 (Ndof - 1) T tau^2, Qi = T tau^2, KE = Q0
        std::vector<double> chain(4 * (size_t)(M + 1), 0.0);
        chain[3] = 2.0 * (N - 1) * T * tau * tau;
        for (int i = 1; i <= M; ++i) chain[4 * i + 3] = T * tau * tau;
        double ke[2] = {chain[3], 1.0};
        HIPCHECK(c, S->chain.ensure(8 * chain.size()));
        HIPCHECK(c, hipMemcpyAsync(S->chain.p, chain.data(), 8 * chain.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, hipMemcpyAsync(S->ke.p, ke, 16, hipMemcpyHostToDevice, c->stream));
        S->M = M;
        S->T = T;
        S->tau = tau;
        S->chainReady = true;
    }
    const double key[6] = {2, dt, T, tau, (double)M, (double)S->ldsVariant};
    return runSteps(c, S, key, nSteps, [=]() { return stepNvt(c, S, dt, T, M); });
}

csg_status csg_system_sync(csg_ctx* c) {
    if (!c) return CSG_ERR_ARGUMENT;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    SystemState* S = sysOf(c);
    return S ? checkCounters(c, S) : CSG_OK;
}

csg_status csg_system_get(csg_ctx* c, int32_t* face, double* bary, double* vel, double* forces) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    HIPCHECK(c, hipSetDevice(c->device));
    if (face) HIPCHECK(c, hipMemcpyAsync(face, S->face.p, 4 * (size_t)S->N, hipMemcpyDeviceToHost, c->stream));
    if (bary) HIPCHECK(c, hipMemcpyAsync(bary, S->bary.p, 24 * (size_t)S->N, hipMemcpyDeviceToHost, c->stream));
    if (vel) HIPCHECK(c, hipMemcpyAsync(vel, S->vel.p, 24 * (size_t)S->N, hipMemcpyDeviceToHost, c->stream));
    if (forces) HIPCHECK(c, hipMemcpyAsync(forces, S->force.p, 24 * (size_t)S->N, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return CSG_OK;
}

csg_status csg_system_snapshot(csg_ctx* c) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    HIPCHECK(c, hipSetDevice(c->device));
    const size_t N = S->N, bytes = 4 * N + 4 * 24 * N;  // faces, then barycentrics, positions, velocities, forces
    if (S->snapBytes < bytes) {
        if (S->snapEvent && S->snapTaken) HIPCHECK(c, hipEventSynchronize(S->snapEvent));  // earlier copies still land here
        if (S->snapHost) HIPCHECK(c, hipHostFree(S->snapHost));
        S->snapHost = nullptr;
        S->snapBytes = 0;
        HIPCHECK(c, hipHostMalloc(&S->snapHost, bytes, hipHostMallocDefault));
        S->snapBytes = bytes;
    }
    if (!S->snapEvent) HIPCHECK(c, hipEventCreateWithFlags(&S->snapEvent, hipEventDisableTiming));
    HIPCHECK(c, S->snapXyz.ensure(24 * N));
    char* h = (char*)S->snapHost;
    hipLaunchKernelGGL(kPositions, dim3(blocksFor(S->N)), dim3(kBlock), 0, c->stream, c->dm, S->N, S->face.as<int32_t>(), S->bary.as<double>(),
                       S->snapXyz.as<double>());
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipMemcpyAsync(h, S->face.p, 4 * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(h + 4 * N, S->bary.p, 24 * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(h + 28 * N, S->snapXyz.p, 24 * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(h + 52 * N, S->vel.p, 24 * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(h + 76 * N, S->force.p, 24 * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipEventRecord(S->snapEvent, c->stream));
    S->snapTaken = true;
    return CSG_OK;
}

csg_status csg_system_snapshot_read(csg_ctx* c, int32_t* face, double* bary, double* xyz, double* vel, double* forces) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S || !S->snapTaken) return fail(c, CSG_ERR_STATE, "csg_system_snapshot first");
    HIPCHECK(c, hipSetDevice(c->device));
    HIPCHECK(c, hipEventSynchronize(S->snapEvent));  // the snapshot's copies only, not the steps queued after them
    const size_t N = S->N;
    const char* h = (const char*)S->snapHost;
    if (face) std::memcpy(face, h, 4 * N);
    if (bary) std::memcpy(bary, h + 4 * N, 24 * N);
    if (xyz) std::memcpy(xyz, h + 28 * N, 24 * N);
    if (vel) std::memcpy(vel, h + 52 * N, 24 * N);
    if (forces) std::memcpy(forces, h + 76 * N, 24 * N);
    return CSG_OK;
}

csg_status csg_system_set_state(csg_ctx* c, const int32_t* face, const double* bary, const double* vel, const double* forces) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    if (face)
        for (int i = 0; i < S->N; ++i)
            if (face[i] < 0 || face[i] >= c->nF) return fail(c, CSG_ERR_ARGUMENT, "particle face out of range");
    HIPCHECK(c, hipSetDevice(c->device));
    if (face) HIPCHECK(c, hipMemcpyAsync(S->face.p, face, 4 * (size_t)S->N, hipMemcpyHostToDevice, c->stream));
    if (bary) HIPCHECK(c, hipMemcpyAsync(S->bary.p, bary, 24 * (size_t)S->N, hipMemcpyHostToDevice, c->stream));
    if (vel) HIPCHECK(c, hipMemcpyAsync(S->vel.p, vel, 24 * (size_t)S->N, hipMemcpyHostToDevice, c->stream));
    if (forces) HIPCHECK(c, hipMemcpyAsync(S->force.p, forces, 24 * (size_t)S->N, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return CSG_OK;
}

csg_status csg_system_thermostat(csg_ctx* c, int32_t* M, double* T, double* tau, double* chain, double* kineticEnergy) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    const int m = S->chainReady ? S->M : 0;
    if (M) *M = m;
    if (T) *T = S->T;
    if (tau) *tau = S->tau;
    double ke[2] = {0, 0};
    HIPCHECK(c, hipSetDevice(c->device));
    if (m > 0) {
        if (chain) HIPCHECK(c, hipMemcpyAsync(chain, S->chain.p, 32 * (size_t)(m + 1), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipMemcpyAsync(ke, S->ke.p, 16, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
    }
    if (kineticEnergy) *kineticEnergy = ke[0];
    return CSG_OK;
}

csg_status csg_system_set_thermostat(csg_ctx* c, int32_t M, double T, double tau, const double* chain, double kineticEnergy) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    if (M == 0) {  // no chain: the next NVT step initialises one, as after csg_system_init
        S->chainReady = false;
        return CSG_OK;
    }
    if (M < 0 || !chain) return fail(c, CSG_ERR_ARGUMENT, "need a chain of length at least 1");
    HIPCHECK(c, hipSetDevice(c->device));
    dropGraph(S);
    const double ke[2] = {kineticEnergy, 1.0};
    HIPCHECK(c, S->chain.ensure(32 * (size_t)(M + 1)));
    HIPCHECK(c, hipMemcpyAsync(S->chain.p, chain, 32 * (size_t)(M + 1), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(S->ke.p, ke, 16, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    S->M = M;
    S->T = T;
    S->tau = tau;
    S->chainReady = true;
    return CSG_OK;
}

csg_status csg_system_status(csg_ctx* c, int64_t* pairsLastForce, int32_t* candOverflow, int32_t* unsolved, int32_t* walkFailures,
                             double* kineticEnergy) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    HIPCHECK(c, hipSetDevice(c->device));
    int cnt[CNT_N];
    double ke[2] = {0, 0};
    HIPCHECK(c, hipMemcpyAsync(cnt, S->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    if (S->chainReady) HIPCHECK(c, hipMemcpyAsync(ke, S->ke.p, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (pairsLastForce) *pairsLastForce = cnt[CNT_PAIRS];
    if (candOverflow) *candOverflow = cnt[CNT_CAND_OVERFLOW];
    if (unsolved) *unsolved = cnt[CNT_UNSOLVED];
    if (walkFailures) *walkFailures = cnt[CNT_WALK_FAIL];
    if (kineticEnergy) *kineticEnergy = ke[0];
    return CSG_OK;
}

csg_status csg_system_profile(csg_ctx* c, int32_t enable) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    S->profiling = enable != 0;
    return CSG_OK;
}

csg_status csg_system_kernel_times(csg_ctx* c, double* ms, int32_t* launches) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < MARK_N; ++k) {
        if (ms) ms[k] = 0;
        if (launches) launches[k] = 0;
    }
    for (auto& m : S->marks) {
        float t = 0;
        HIPCHECK(c, hipEventElapsedTime(&t, m.a, m.b));
        if (ms) ms[m.kind] += t;
        if (launches) launches[m.kind] += 1;
        hipEventDestroy(m.a);
        hipEventDestroy(m.b);
    }
    S->marks.clear();
    return CSG_OK;
}

csg_status csg_system_patch_stats(csg_ctx* c, int64_t* sums, int32_t* maxima) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    std::vector<int> st(kPatchStats * (size_t)S->N);
    HIPCHECK(c, hipMemcpy(st.data(), S->pstat.p, 4 * kPatchStats * (size_t)S->N, hipMemcpyDeviceToHost));
    int64_t sm[kPatchStats] = {};
    int mx[kPatchStats] = {};
    for (int i = 0; i < S->N; ++i)
        for (int a = 0; a < kPatchStats; ++a) {
            sm[a] += st[kPatchStats * i + a];
            mx[a] = std::max(mx[a], st[kPatchStats * i + a]);
        }
    for (int a = 0; a < kPatchStats; ++a) {
        if (sums) sums[a] = sm[a];
        if (maxima) maxima[a] = mx[a];
    }
    return CSG_OK;
}

csg_status csg_system_fallback_counts(csg_ctx* c, int32_t* reasons) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    int cnt[CNT_N];
    HIPCHECK(c, hipMemcpy(cnt, S->counters.p, sizeof(cnt), hipMemcpyDeviceToHost));
    for (int b = 0; b < 6; ++b) reasons[b] = cnt[CNT_LDS_REASON + b];
    return CSG_OK;
}

csg_status csg_system_autotune(csg_ctx* c, int32_t* chosen, double* ms) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S || !S->configured) return fail(c, CSG_ERR_STATE, "csg_system_init and csg_system_configure first");
    HIPCHECK(c, hipSetDevice(c->device));
    const bool prof = S->profiling;
    S->profiling = false;
    hipEvent_t a, b;
    HIPCHECK(c, hipEventCreate(&a));
    HIPCHECK(c, hipEventCreate(&b));
    int best = 0;
    double bestMs = HUGE_VAL;
    for (int v = 0; v < kLdsVariants; ++v) {
        S->ldsVariant = v;
        csg_status s = sysForces(c, S);  // untimed: first use of the variant
        if (s) return s;
        HIPCHECK(c, hipEventRecord(a, c->stream));
        for (int r = 0; r < 2; ++r) {
            s = sysForces(c, S);
            if (s) return s;
        }
        HIPCHECK(c, hipEventRecord(b, c->stream));
        HIPCHECK(c, hipEventSynchronize(b));
        float t = 0;
        HIPCHECK(c, hipEventElapsedTime(&t, a, b));
        if (ms) ms[v] = t / 2;
        if (t < bestMs) {
            bestMs = t;
            best = v;
        }
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    S->ldsVariant = best;
    S->profiling = prof;
    if (chosen) *chosen = best;
    return CSG_OK;
}

csg_status csg_set_boundary(csg_ctx* c, int32_t mode) {
    if (!c) return CSG_ERR_ARGUMENT;
    if (mode != CSG_BOUNDARY_CLOSED && mode != CSG_BOUNDARY_ABSORBING && mode != CSG_BOUNDARY_TANGENTIAL)
        return fail(c, CSG_ERR_ARGUMENT, "unknown boundary mode");
    c->boundary = mode;
    if (SystemState* S = sysOf(c)) dropGraph(S);  // a captured step holds the walk's parameters
    return CSG_OK;
}

csg_status csg_system_use_graph(csg_ctx* c, int32_t enable) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    dropGraph(S);
    S->useGraph = enable != 0;
    return CSG_OK;
}

csg_status csg_system_set_lds_variant(csg_ctx* c, int32_t variant) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    if (variant < 0 || variant >= kLdsVariants) return fail(c, CSG_ERR_ARGUMENT, "unknown LDS kernel variant");
    dropGraph(S);
    S->ldsVariant = variant;
    return CSG_OK;
}

csg_status csg_system_record_pairs(csg_ctx* c, int32_t enable) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S || !S->configured) return fail(c, CSG_ERR_STATE, "csg_system_init and csg_system_configure first");
    dropGraph(S);
    HIPCHECK(c, hipSetDevice(c->device));
    if (enable) HIPCHECK(c, S->pdist.ensure(8 * (size_t)std::max<long>(S->candCap, 1)));
    S->recordPairs = enable != 0;
    return CSG_OK;
}

csg_status csg_system_pair_distances(csg_ctx* c, double* dist, int64_t cap) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S || !S->recordPairs) return fail(c, CSG_ERR_STATE, "csg_system_record_pairs(ctx, 1) before the force computation");
    if (S->screenOn) return fail(c, CSG_ERR_STATE, "pair distances are recorded for the unscreened lists only");
    int n = 0;
    HIPCHECK(c, hipMemcpy(&n, S->toff.as<int>() + S->N, 4, hipMemcpyDeviceToHost));
    if (n > cap) return fail(c, CSG_ERR_CAPACITY, "distance buffer too small");
    HIPCHECK(c, hipMemcpy(dist, S->pdist.p, 8 * (size_t)n, hipMemcpyDeviceToHost));
    return CSG_OK;
}

csg_status csg_system_screen_features(csg_ctx* c, float* features, int64_t cap) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S || !S->configured) return fail(c, CSG_ERR_STATE, "csg_system_init and csg_system_configure first");
    if (!S->cellList) return fail(c, CSG_ERR_STATE, "the screen works on cell-list candidates");
    HIPCHECK(c, hipSetDevice(c->device));
    HIPCHECK(c, S->sfeat.ensure(sizeof(float) * kScreenIn * (size_t)std::max<long>(S->candCap, 1)));
    int n = 0;
    HIPCHECK(c, hipMemcpy(&n, S->toff.as<int>() + S->N, 4, hipMemcpyDeviceToHost));
    if (n > cap) return fail(c, CSG_ERR_CAPACITY, "feature buffer too small");
    screenFeatures(c, S);
    HIPCHECK(c, hipGetLastError());
    HIPCHECK(c, hipMemcpyAsync(features, S->sfeat.p, sizeof(float) * kScreenIn * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return CSG_OK;
}

// per-group truncation radius and sentinel from the group's potential, range and maximum distance (as exactTruncation
// does for a single system whose truncation radius is its range)
static csg_status uploadGroupRadii(csg_ctx* c, SystemState* S) {
    std::vector<double> rmax(S->nGroups), sent(S->nGroups);
    for (int g = 0; g < S->nGroups; ++g) {
        sent[g] = 2.0 * S->hgmaxd[g];
        rmax[g] = exactTruncation(S->hgpot[g], S->hgrange[g], c->submeshing != 0, sent[g]);
    }
    HIPCHECK(c, hipMemcpy(S->grmax.p, rmax.data(), 8 * (size_t)S->nGroups, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(S->gsent.p, sent.data(), 8 * (size_t)S->nGroups, hipMemcpyHostToDevice));
    return CSG_OK;
}

csg_status csg_system_set_groups(csg_ctx* c, int32_t nGroups, const int32_t* groupOf, const csg_potential* pots, const double* ranges,
                                 const double* maximumDistances) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S || !S->configured) return fail(c, CSG_ERR_STATE, "csg_system_init and csg_system_configure first");
    dropGraph(S);
    HIPCHECK(c, hipSetDevice(c->device));
    if (nGroups <= 0) {
        S->nGroups = 0;
        return CSG_OK;
    }
    if (!groupOf || !pots || !ranges || !maximumDistances) return fail(c, CSG_ERR_ARGUMENT, "null group arrays");
    if (!S->cellList) return fail(c, CSG_ERR_STATE, "particle groups need cell-list candidates");
    for (int i = 0; i < S->N; ++i)
        if (groupOf[i] < 0 || groupOf[i] >= nGroups) return fail(c, CSG_ERR_ARGUMENT, "group index out of range");
    for (int g = 0; g < nGroups; ++g)
        if (!(ranges[g] > 0) || ranges[g] > S->range) return fail(c, CSG_ERR_ARGUMENT, "a group range is not in (0, configured range]");
    S->hgpot.resize(nGroups);
    S->hgrange.assign(ranges, ranges + nGroups);
    S->hgmaxd.assign(maximumDistances, maximumDistances + nGroups);
    std::vector<double> r2(nGroups);
    for (int g = 0; g < nGroups; ++g) {
        S->hgpot[g] = Potential{pots[g].kind, pots[g].p0, pots[g].p1, pots[g].p2};
        r2[g] = ranges[g] * ranges[g];
    }
    HIPCHECK(c, S->grp.ensure(4 * (size_t)S->N));
    HIPCHECK(c, S->gpot.ensure(sizeof(Potential) * (size_t)nGroups));
    HIPCHECK(c, S->gr2.ensure(8 * (size_t)nGroups));
    HIPCHECK(c, S->gmaxd.ensure(8 * (size_t)nGroups));
    HIPCHECK(c, S->grmax.ensure(8 * (size_t)nGroups));
    HIPCHECK(c, S->gsent.ensure(8 * (size_t)nGroups));
    HIPCHECK(c, S->grange.ensure(8 * (size_t)nGroups));
    HIPCHECK(c, hipMemcpy(S->grp.p, groupOf, 4 * (size_t)S->N, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(S->grange.p, ranges, 8 * (size_t)nGroups, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(S->gpot.p, S->hgpot.data(), sizeof(Potential) * (size_t)nGroups, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(S->gr2.p, r2.data(), 8 * (size_t)nGroups, hipMemcpyHostToDevice));
    HIPCHECK(c, hipMemcpy(S->gmaxd.p, maximumDistances, 8 * (size_t)nGroups, hipMemcpyHostToDevice));
    S->nGroups = nGroups;
    double radius = std::max(c->maximumDistance, S->range);
    for (int g = 0; g < nGroups; ++g) radius = std::max(radius, std::max(ranges[g], maximumDistances[g]));
    if (radius > S->bigRadius) {
        csg_status st = sizeLastTier(c, S, radius);
        if (st) return st;
    }
    return uploadGroupRadii(c, S);
}

csg_status csg_system_set_screen(csg_ctx* c, const float* params, float threshold) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S || !S->configured) return fail(c, CSG_ERR_STATE, "csg_system_init and csg_system_configure first");
    if (params && (!S->cellList || !c->submeshing)) return fail(c, CSG_ERR_STATE, "the screen needs cell-list candidates and submeshing");
    dropGraph(S);
    HIPCHECK(c, hipSetDevice(c->device));
    if (!params) {
        S->screenOn = false;
        return CSG_OK;
    }
    const size_t cap = (size_t)std::max<long>(S->candCap, 1);
    HIPCHECK(c, S->skeep.ensure(cap));
    HIPCHECK(c, S->sparams.ensure(sizeof(float) * kScreenParams));
    HIPCHECK(c, hipMemcpyAsync(S->sparams.p, params, sizeof(float) * kScreenParams, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    S->screenThreshold = threshold;
    S->screenOn = true;
    return CSG_OK;
}

csg_status csg_system_screen_stats(csg_ctx* c, int64_t* candidates, int64_t* kept) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    int cnt[CNT_N];
    HIPCHECK(c, hipMemcpy(cnt, S->counters.p, sizeof(cnt), hipMemcpyDeviceToHost));
    if (candidates) *candidates = cnt[CNT_PAIRS];
    if (kept) *kept = S->screenOn ? cnt[CNT_KEPT] : cnt[CNT_PAIRS];
    return CSG_OK;
}

csg_status csg_system_pairs(csg_ctx* c, int32_t* offsets, int32_t* idx, int64_t cap) {
    SystemState* S = sysOf(c);
    if (!c) return CSG_ERR_ARGUMENT;
    if (!S) return fail(c, CSG_ERR_STATE, "csg_system_init first");
    std::vector<int> off(S->N + 1);
    HIPCHECK(c, hipMemcpy(off.data(), S->toff.p, 4 * (size_t)(S->N + 1), hipMemcpyDeviceToHost));
    if (offsets) std::copy(off.begin(), off.end(), offsets);
    if (idx) {
        if (off[S->N] > cap) return fail(c, CSG_ERR_CAPACITY, "index buffer too small");
        HIPCHECK(c, hipMemcpy(idx, S->cand.p, 4 * (size_t)off[S->N], hipMemcpyDeviceToHost));
    }
    return CSG_OK;
}

}  // extern "C"
