// C-ABI of the geodesic force path (include/csg.h): contexts, mesh preprocessing, device buffers and launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/csg.h"
#include "csg_common.h"
#include "csg_context.h"
#include "geodesic_kernel.h"
#include "transport_kernel.h"

// ---- csg_pair_forces_noalloc: its scratch is a static device array --------------------------------------------------
namespace {

constexpr int kNoAllocMaxSources = 2048;          // sources per call
constexpr int kNoAllocSpill = 256;                // spill records per source (beyond: the HBM-table tier)
constexpr int kNoAllocFallbackBlocks = 16;        // workgroups of the HBM-table tier
constexpr long kNoAllocFallbackFaces = 1536;      // faces and targets per source the HBM tables hold
constexpr long kNoAllocFallbackTargets = 4096;
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;

}  // namespace

// The scratch of csg_pair_forces_noalloc: one copy per GPU, held until the process ends. Host code uses the device
// address that hipGetSymbolAddress gives for it (csg_ctx::noallocScratch), never this name.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

// byte offsets of the regions of the scratch array
struct NoAllocLayout {
    size_t status, pcount, counter, forces, energy, plist, spill, tables, total;
    csg::ScratchLayout tablesLayout;
};

NoAllocLayout noAllocLayout() {
    auto r = [](size_t b) { return (b + 255) & ~(size_t)255; };
    NoAllocLayout o{};
    size_t off = 0;
    const size_t n = kNoAllocMaxSources;
    o.status = off, off += r(4 * n);
    o.pcount = off, off += r(4 * n);
    o.counter = off, off += r(4);
    o.forces = off, off += r(24 * n);
    o.energy = off, off += r(8 * n);
    o.plist = off, off += r(4 * n * csg::kPatchListCap);
    o.spill = off, off += r(sizeof(csg::SpillRecord) * n * kNoAllocSpill);
    o.tablesLayout = csg::makeScratchLayout(kNoAllocFallbackFaces, kNoAllocFallbackTargets);
    o.tables = off, off += o.tablesLayout.bytesPerBlock * kNoAllocFallbackBlocks;
    o.total = off;
    return o;
}

__global__ void kNoAllocZero(int* p) { *p = 0; }

__global__ void kNoAllocCountUnsolved(int n, const int* status, int* counter) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && status[i] != 0) atomicAdd(counter, 1);
}

}  // namespace

namespace csg {

long patchFaceBound(const csg_ctx* c, double radius) {
    const long nF = c->nF;
    const double cell = radius + c->maxEdge;
    if (nF == 0 || !(cell > 0) || !std::isfinite(cell)) return nF;
    long n[3];
    double lo[3];
    long double cells = 1;
    for (int a = 0; a < 3; ++a) {
        lo[a] = c->lo[a];
        n[a] = (long)std::floor((c->hi[a] - c->lo[a]) / cell) + 1;
        cells *= n[a];
    }
    if (cells > 1e15L) return nF;  // a radius tiny against the mesh: no useful bound from bins
    std::unordered_map<long long, int> count;
    count.reserve(2 * (size_t)nF);
    auto key = [&](long x, long y, long z) { return (long long)x + (long long)n[0] * ((long long)y + (long long)n[1] * z); };
    std::vector<long long> keys(nF);
    for (long f = 0; f < nF; ++f) {
        long q[3];
        for (int a = 0; a < 3; ++a)
            q[a] = std::min(n[a] - 1, std::max(0L, (long)std::floor((c->faceCentroid[3 * f + a] - lo[a]) / cell)));
        keys[f] = key(q[0], q[1], q[2]);
        ++count[keys[f]];
    }
    long best = 0;
    for (const auto& kv : count) {
        const long long k = kv.first;
        const long x = (long)(k % n[0]), y = (long)((k / n[0]) % n[1]), z = (long)(k / ((long long)n[0] * n[1]));
        long sum = 0;
        for (long dx = -1; dx <= 1; ++dx)
            for (long dy = -1; dy <= 1; ++dy)
                for (long dz = -1; dz <= 1; ++dz) {
                    const long X = x + dx, Y = y + dy, Z = z + dz;
                    if (X < 0 || Y < 0 || Z < 0 || X >= n[0] || Y >= n[1] || Z >= n[2]) continue;
                    auto it = count.find(key(X, Y, Z));
                    if (it != count.end()) sum += it->second;
                }
        best = std::max(best, sum);
    }
    return std::min(nF, best);
}

}  // namespace csg

namespace {

using csg::fail;
using csg::DevBuf;

template <class T>
hipError_t upload(csg_ctx* c, DevBuf& b, const T* host, size_t n) {
    hipError_t e = b.ensure(n * sizeof(T));
    if (e != hipSuccess || n == 0) return e;
    return hipMemcpyAsync(b.p, host, n * sizeof(T), hipMemcpyHostToDevice, c->stream);
}

template <class T>
hipError_t download(csg_ctx* c, T* host, const DevBuf& b, size_t n) {
    if (!host || n == 0) return hipSuccess;
    return hipMemcpyAsync(host, b.p, n * sizeof(T), hipMemcpyDeviceToHost, c->stream);
}

// Face adjacency, per-edge frames, normals, vertex classes, as the reference's Surface_mesh would provide them.
// Per vertex (mean curvature H, Gaussian curvature K), features of the learned screen: K = (2 pi - angle sum) / A and
// H from the cotangent Laplacian, Delta x = (1 / 2A) sum_faces cot(theta_k) (x_j - x_i) over the edge (i, j) opposite
// corner k, as H = -n . Delta x / 2 with the area-weighted vertex normal n; A is a third of the incident face area.
// Border vertices get (0, 0).
std::vector<double> vertexCurvatures(int nV, const double* X, int nF, const int32_t* F, const std::vector<double>& angle,
                                     const std::vector<uint8_t>& flag) {
    std::vector<double> lap(3 * (size_t)nV, 0.0), vn(3 * (size_t)nV, 0.0), area(nV, 0.0), out(2 * (size_t)nV, 0.0);
    for (int f = 0; f < nF; ++f) {
        const double* p[3] = {X + 3 * F[3 * f], X + 3 * F[3 * f + 1], X + 3 * F[3 * f + 2]};
        double e1[3], e2[3], n[3];
        for (int a = 0; a < 3; ++a) {
            e1[a] = p[1][a] - p[0][a];
            e2[a] = p[2][a] - p[0][a];
        }
        n[0] = e1[1] * e2[2] - e1[2] * e2[1];
        n[1] = e1[2] * e2[0] - e1[0] * e2[2];
        n[2] = e1[0] * e2[1] - e1[1] * e2[0];
        const double A = 0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; ++k) {
            const int vi = F[3 * f + (k + 1) % 3], vj = F[3 * f + (k + 2) % 3];
            const double* pk = p[k];
            const double* pi = p[(k + 1) % 3];
            const double* pj = p[(k + 2) % 3];
            double u[3], w[3];
            for (int a = 0; a < 3; ++a) {
                u[a] = pi[a] - pk[a];
                w[a] = pj[a] - pk[a];
            }
            const double dot = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
            const double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
            const double cot = dot / std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            for (int a = 0; a < 3; ++a) {
                lap[3 * vi + a] += cot * (pj[a] - pi[a]);
                lap[3 * vj + a] += cot * (pi[a] - pj[a]);
            }
            const int v = F[3 * f + k];
            area[v] += A / 3.0;
            for (int a = 0; a < 3; ++a) vn[3 * v + a] += n[a];
        }
    }
    for (int v = 0; v < nV; ++v) {
        if ((flag[v] & 2) || !(area[v] > 0)) continue;
        const double nl = std::sqrt(vn[3 * v] * vn[3 * v] + vn[3 * v + 1] * vn[3 * v + 1] + vn[3 * v + 2] * vn[3 * v + 2]);
        double h = 0;
        for (int a = 0; a < 3; ++a) h += -(vn[3 * v + a] / nl) * lap[3 * v + a] / (2.0 * area[v]) / 2.0;
        out[2 * v] = h;
        out[2 * v + 1] = (2 * M_PI - angle[v]) / area[v];
    }
    return out;
}

csg_status preprocessMesh(csg_ctx* c, int nV, const double* X, int nF, const int32_t* F) {
    if (nV <= 0 || nF <= 0 || !X || !F) return fail(c, CSG_ERR_ARGUMENT, "empty mesh");
    for (long i = 0; i < 3L * nF; ++i)
        if (F[i] < 0 || F[i] >= nV) return fail(c, CSG_ERR_MESH, "face references a vertex out of range");
    std::vector<int32_t> adj(3 * (size_t)nF, -1);
    std::unordered_map<uint64_t, int> directed;
    directed.reserve(6 * (size_t)nF);
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) {
            uint64_t a = (uint32_t)F[3 * f + k], b = (uint32_t)F[3 * f + (k + 1) % 3];
            if (a == b) return fail(c, CSG_ERR_MESH, "degenerate face");
            if (!directed.emplace((a << 32) | b, 4 * f + k).second)
                return fail(c, CSG_ERR_MESH, "directed edge used twice: non-manifold or inconsistently oriented mesh");
        }
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) {
            uint64_t a = (uint32_t)F[3 * f + k], b = (uint32_t)F[3 * f + (k + 1) % 3];
            auto it = directed.find((b << 32) | a);
            if (it != directed.end()) adj[3 * f + k] = it->second;
        }
    std::vector<double> geo(9 * (size_t)nF), nrm(3 * (size_t)nF);
    std::vector<double> angle(nV, 0.0);
    c->faceCentroid.assign(3 * (size_t)nF, 0.0f);  // for the bound on patch sizes (csg::patchFaceBound)
    c->maxEdge = 0;
    std::vector<uint8_t> flag(nV, 0);
    auto P = [&](int f, int k) { return X + 3 * F[3 * f + ((k % 3) + 3) % 3]; };
    c->area = 0;
    for (int f = 0; f < nF; ++f) {
        for (int k = 0; k < 3; ++k) {
            const double *p0 = P(f, k), *p1 = P(f, k + 1), *p2 = P(f, k + 2);
            double e[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
            double d2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
            double L = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            double ex[3] = {e[0] / L, e[1] / L, e[2] / L};
            double n[3] = {e[1] * d2[2] - e[2] * d2[1], e[2] * d2[0] - e[0] * d2[2], e[0] * d2[1] - e[1] * d2[0]};
            double nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            if (!(nl > 0)) return fail(c, CSG_ERR_MESH, "zero-area face");
            for (int a = 0; a < 3; ++a) n[a] /= nl;
            double ey[3] = {n[1] * ex[2] - n[2] * ex[1], n[2] * ex[0] - n[0] * ex[2], n[0] * ex[1] - n[1] * ex[0]};
            geo[9 * f + 3 * k] = L;
            c->maxEdge = std::max(c->maxEdge, L);
            for (int a = 0; a < 3; ++a) c->faceCentroid[3 * (size_t)f + a] += (float)(p0[a] / 3.0);
            geo[9 * f + 3 * k + 1] = d2[0] * ex[0] + d2[1] * ex[1] + d2[2] * ex[2];
            geo[9 * f + 3 * k + 2] = d2[0] * ey[0] + d2[1] * ey[1] + d2[2] * ey[2];
            if (k == 0) {
                for (int a = 0; a < 3; ++a) nrm[3 * f + a] = n[a];
                c->area += 0.5 * nl;
            }
            double l2 = std::sqrt(d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2]);
            double cs = (e[0] * d2[0] + e[1] * d2[1] + e[2] * d2[2]) / (L * l2);
            angle[F[3 * f + k]] += std::acos(std::max(-1.0, std::min(1.0, cs)));
            if (adj[3 * f + k] < 0) {
                flag[F[3 * f + k]] |= 2;
                flag[F[3 * f + (k + 1) % 3]] |= 2;
            }
        }
    }
    for (int v = 0; v < nV; ++v)
        if (angle[v] >= 2 * M_PI - 1e-12) flag[v] |= 1;
    std::vector<int32_t> vfoff(nV + 1, 0), vflist(3 * (size_t)nF);
    for (long i = 0; i < 3L * nF; ++i) vfoff[F[i] + 1]++;
    for (int v = 0; v < nV; ++v) vfoff[v + 1] += vfoff[v];
    std::vector<int32_t> fill(vfoff.begin(), vfoff.end() - 1);
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) vflist[fill[F[3 * f + k]]++] = 4 * f + k;
    for (int v = 0; v < nV; ++v)
        if (vfoff[v + 1] - vfoff[v] > 32) return fail(c, CSG_ERR_CAPACITY, "a vertex has more than 32 incident faces");
    for (int a = 0; a < 3; ++a) c->lo[a] = c->hi[a] = 0.0;  // seeded at the origin (triangulatedMeshSpace.cpp:9-14)
    for (int v = 0; v < nV; ++v)
        for (int a = 0; a < 3; ++a) {
            c->lo[a] = std::min(c->lo[a], X[3 * v + a]);
            c->hi[a] = std::max(c->hi[a], X[3 * v + a]);
        }
    HIPCHECK(c, upload(c, c->vx, X, 3 * (size_t)nV));
    HIPCHECK(c, upload(c, c->fv, F, 3 * (size_t)nF));
    HIPCHECK(c, upload(c, c->fadj, adj.data(), adj.size()));
    HIPCHECK(c, upload(c, c->fgeo, geo.data(), geo.size()));
    HIPCHECK(c, upload(c, c->fnrm, nrm.data(), nrm.size()));
    HIPCHECK(c, upload(c, c->vflag, flag.data(), flag.size()));
    HIPCHECK(c, upload(c, c->vfoff, vfoff.data(), vfoff.size()));
    HIPCHECK(c, upload(c, c->vflist, vflist.data(), vflist.size()));
    std::vector<double> rule(20 * (size_t)nF, 0.0);
    for (int f = 0; f < nF; ++f) {
        double* r = rule.data() + 20 * (size_t)f;
        const int32_t words[4] = {adj[3 * f], adj[3 * f + 1], adj[3 * f + 2], 0};
        std::memcpy(r, words, sizeof(words));
        for (int k = 0; k < 3; ++k) {
            for (int a = 0; a < 3; ++a) r[2 + 3 * k + a] = X[3 * F[3 * f + k] + a];
            const int n = adj[3 * f + k];
            if (n < 0) continue;
            const int g = n >> 2, opp = F[3 * g + ((n & 3) + 2) % 3];  // the neighbour's corner off the shared edge
            for (int a = 0; a < 3; ++a) r[11 + 3 * k + a] = X[3 * opp + a];
        }
    }
    HIPCHECK(c, upload(c, c->frule, rule.data(), rule.size()));
    std::vector<double> curv = vertexCurvatures(nV, X, nF, F, angle, flag);
    HIPCHECK(c, upload(c, c->curv, curv.data(), curv.size()));
    HIPCHECK(c, c->fchild.ensure(18 * sizeof(double) * (size_t)nF));
    c->nV = nV;
    c->nF = nF;
    c->dm = csg::DeviceMesh{nV, nF, c->vx.as<double>(), c->fv.as<int32_t>(), c->fadj.as<int32_t>(), c->fgeo.as<double>(),
                            c->fnrm.as<double>(), c->vflag.as<uint8_t>(), c->vfoff.as<int32_t>(), c->vflist.as<int32_t>(),
                            c->fchild.as<double>(), c->frule.as<double>()};
    HIPCHECK(c, csg::launchChildFrames(c->dm, c->fchild.as<double>(), c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->haveMesh = true;
    return CSG_OK;
}

int pow2AtLeast(long n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// scratch tables sized for `faces` patch faces and `targets` targets per source
csg_status ensureScratch(csg_ctx* c, long faces, long targets, int blocks) {
    csg::ScratchLayout L = csg::makeScratchLayout(faces, targets);
    size_t need = L.bytesPerBlock * (size_t)blocks;
    HIPCHECK(c, c->scratch.ensure(need));
    L.base = c->scratch.p;
    c->layout = L;
    c->scratchBlocks = blocks;
    return CSG_OK;
}

// Solves nSrc sources over staged particle buffers; overflowed sources of the LDS variant are re-solved in scratch.
csg_status runGeodesics(csg_ctx* c, csg::GeoParams& G, int maxTargets, bool wholeMesh, bool wantStatus) {
    if (G.nSrc == 0) return CSG_OK;
    HIPCHECK(c, c->status.ensure(sizeof(int) * (size_t)G.nSrc));
    G.status = c->status.as<int>();
    if (!wholeMesh) {
        if (G.patchRadius) {
            HIPCHECK(c, c->plist.ensure(4 * (size_t)csg::kPatchListCap * G.nSrc));
            HIPCHECK(c, c->pcount.ensure(4 * (size_t)G.nSrc));
            G.patchList = c->plist.as<int32_t>();
            G.patchCount = c->pcount.as<int32_t>();
            G.patchStride = csg::kPatchListCap;
            HIPCHECK(c, c->spill.ensure(sizeof(csg::SpillRecord) * (size_t)csg::kSpillCap * G.nSrc));
            G.spill = c->spill.as<csg::SpillRecord>();
            G.spillCap = csg::kSpillCap;
            HIPCHECK(c, csg::launchPatch(G, c->plist.as<int32_t>(), c->pcount.as<int32_t>(), G.patchStride, c->stream));
        }
        HIPCHECK(c, csg::launchGeodesicLds(G, 0, nullptr, nullptr, nullptr, c->stream));
        std::vector<int> st(G.nSrc);
        HIPCHECK(c, hipMemcpyAsync(st.data(), G.status, sizeof(int) * G.nSrc, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        std::vector<int> redo;
        for (int s = 0; s < G.nSrc; ++s)
            if (st[s]) redo.push_back(s);
        if (redo.empty()) return CSG_OK;
        long faces = 4096;  // grows until every listed source fits
        for (int attempt = 0; attempt < 6 && !redo.empty(); ++attempt, faces *= 4) {
            faces = std::min<long>(faces, c->nF);
            int blocks = std::min<int>((int)redo.size(), 512);
            csg_status s = ensureScratch(c, faces, maxTargets, blocks);
            if (s) return s;
            HIPCHECK(c, upload(c, c->list, redo.data(), redo.size()));
            HIPCHECK(c, csg::launchGeodesicScratch(G, c->layout, c->list.as<int32_t>(), (int)redo.size(), blocks, c->stream));
            HIPCHECK(c, hipMemcpyAsync(st.data(), G.status, sizeof(int) * G.nSrc, hipMemcpyDeviceToHost, c->stream));
            HIPCHECK(c, hipStreamSynchronize(c->stream));
            std::vector<int> next;
            for (int s2 : redo)
                if (st[s2]) next.push_back(s2);
            redo.swap(next);
            if (faces >= c->nF) break;
        }
        if (!redo.empty()) return fail(c, CSG_ERR_CAPACITY, "a source exceeded every scratch capacity");
        return CSG_OK;
    }
    int blocks = std::min(G.nSrc, 256);
    csg_status s = ensureScratch(c, c->nF, maxTargets, blocks);
    if (s) return s;
    HIPCHECK(c, csg::launchGeodesicScratchWide(G, c->layout, nullptr, G.nSrc, blocks, c->stream));
    std::vector<int> st(G.nSrc);
    HIPCHECK(c, hipMemcpyAsync(st.data(), G.status, sizeof(int) * G.nSrc, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    (void)wantStatus;
    for (int s2 = 0; s2 < G.nSrc; ++s2)
        if (st[s2]) return fail(c, CSG_ERR_CAPACITY, "whole-mesh solve exceeded the scratch window capacity");
    return CSG_OK;
}

}  // namespace

extern "C" {

csg_status csg_ctx_create(csg_ctx** out, int device) {
    if (!out) return CSG_ERR_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return CSG_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return CSG_ERR_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return CSG_ERR_DEVICE;
    auto* c = new csg_ctx();
    c->device = device;
    c->arch = prop.gcnArchName;
    if (c->arch.rfind("gfx950", 0) != 0) {
        delete c;
        return CSG_ERR_DEVICE;  // the code objects of this library are built for gfx950 only
    }
    if (hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return CSG_ERR_DEVICE;
    }
    // the device address of csg_pair_forces_noalloc's scratch on this GPU (this device is current)
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
        c->noallocErr = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
    else if (noAllocLayout().total > sizeof(scratch))
        c->noallocErr = "the scratch layout exceeds the static array";
    else
        c->noallocScratch = p;
    *out = c;
    return CSG_OK;
}

void csg_ctx_destroy(csg_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    delete ctx;
}

const char* csg_last_error(const csg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

const char* csg_device_name(const csg_ctx* ctx) { return ctx ? ctx->arch.c_str() : ""; }

csg_status csg_mesh_upload(csg_ctx* ctx, int32_t nV, const double* xyz, int32_t nF, const int32_t* tri) {
    if (!ctx) return CSG_ERR_ARGUMENT;
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    ctx->haveMesh = false;
    return preprocessMesh(ctx, nV, xyz, nF, tri);
}

csg_status csg_mesh_info(const csg_ctx* ctx, int32_t* nV, int32_t* nF, double* area, double* lo3, double* hi3) {
    if (!ctx) return CSG_ERR_ARGUMENT;
    if (!ctx->haveMesh) return CSG_ERR_STATE;
    if (nV) *nV = ctx->nV;
    if (nF) *nF = ctx->nF;
    if (area) *area = ctx->area;
    for (int a = 0; a < 3; ++a) {
        if (lo3) lo3[a] = ctx->lo[a];
        if (hi3) hi3[a] = ctx->hi[a];
    }
    return CSG_OK;
}

csg_status csg_set_submeshing(csg_ctx* ctx, int32_t enabled, double maximumDistance) {
    if (!ctx) return CSG_ERR_ARGUMENT;
    ctx->submeshing = enabled ? 1 : 0;
    ctx->maximumDistance = maximumDistance;
    return CSG_OK;
}

csg_status csg_pair_forces(csg_ctx* c, int32_t nP, const int32_t* face, const double* bary, int32_t nSrc, const int32_t* srcIdx,
                           const int32_t* tgtOffsets, const int32_t* tgtIdx, const double* patchRadius, double rMax,
                           const csg_potential* pot, int32_t useSentinel, double sentinelDistance, double* dist, float* startTan,
                           float* endTan, uint32_t* flags, double* forces, double* energy) {
    if (!c) return CSG_ERR_ARGUMENT;
    if (!c->haveMesh) return fail(c, CSG_ERR_STATE, "no mesh uploaded");
    if (nP < 0 || nSrc < 0 || (nP > 0 && (!face || !bary)) || (nSrc > 0 && !tgtOffsets))
        return fail(c, CSG_ERR_ARGUMENT, "bad particle or source arrays");
    HIPCHECK(c, hipSetDevice(c->device));
    long nPairs = nSrc > 0 ? tgtOffsets[nSrc] : 0;
    if (nPairs > 0 && !tgtIdx) return fail(c, CSG_ERR_ARGUMENT, "null target list");
    int maxT = 0;
    for (int s = 0; s < nSrc; ++s) {
        int n = tgtOffsets[s + 1] - tgtOffsets[s];
        if (n < 0 || tgtOffsets[s] < 0) return fail(c, CSG_ERR_ARGUMENT, "target offsets must be non-decreasing from 0");
        maxT = std::max(maxT, n);
        int p = srcIdx ? srcIdx[s] : s;
        if (p < 0 || p >= nP) return fail(c, CSG_ERR_ARGUMENT, "source index out of range");
    }
    if (tgtOffsets && nSrc > 0 && tgtOffsets[0] != 0) return fail(c, CSG_ERR_ARGUMENT, "tgtOffsets[0] must be 0");
    if (maxT > (1 << 20)) return fail(c, CSG_ERR_CAPACITY, "more than 2^20 targets for one source");
    for (long q = 0; q < nPairs; ++q)
        if (tgtIdx[q] < 0 || tgtIdx[q] >= nP) return fail(c, CSG_ERR_ARGUMENT, "target index out of range");
    for (int i = 0; i < nP; ++i)
        if (face[i] < 0 || face[i] >= c->nF) return fail(c, CSG_ERR_ARGUMENT, "particle face out of range");
    HIPCHECK(c, upload(c, c->pface, face, (size_t)nP));
    HIPCHECK(c, upload(c, c->pbary, bary, 3 * (size_t)nP));
    if (srcIdx) HIPCHECK(c, upload(c, c->srcIdx, srcIdx, (size_t)nSrc));
    HIPCHECK(c, upload(c, c->toff, tgtOffsets, (size_t)nSrc + 1));
    HIPCHECK(c, upload(c, c->tidx, tgtIdx, (size_t)nPairs));
    if (patchRadius) HIPCHECK(c, upload(c, c->prad, patchRadius, (size_t)nSrc));
    HIPCHECK(c, c->pdist.ensure(8 * (size_t)nPairs));
    HIPCHECK(c, c->pst.ensure(12 * (size_t)nPairs));
    HIPCHECK(c, c->pet.ensure(12 * (size_t)nPairs));
    HIPCHECK(c, c->pflags.ensure(4 * (size_t)nPairs));
    HIPCHECK(c, c->pforces.ensure(24 * (size_t)nSrc));
    HIPCHECK(c, c->penergy.ensure(8 * (size_t)nSrc));
    csg::GeoParams G{};
    G.mesh = c->dm;
    G.pface = c->pface.as<int32_t>();
    G.pbary = c->pbary.as<double>();
    G.nSrc = nSrc;
    G.srcIdx = srcIdx ? c->srcIdx.as<int32_t>() : nullptr;
    G.toff = c->toff.as<int32_t>();
    G.tidx = c->tidx.as<int32_t>();
    G.patchRadius = patchRadius ? c->prad.as<double>() : nullptr;
    G.rMax = rMax;
    G.pairDist = c->pdist.as<double>();
    G.pairStart = c->pst.as<float>();
    G.pairEnd = c->pet.as<float>();
    G.pairFlags = c->pflags.as<uint32_t>();
    G.forces = c->pforces.as<double>();
    G.energy = c->penergy.as<double>();
    G.pot = pot ? csg::Potential{pot->kind, pot->p0, pot->p1, pot->p2} : csg::Potential{csg::POT_NONE, 0, 0, 0};
    G.useSentinel = useSentinel;
    G.sentinelDist = sentinelDistance;
    G.rMax = csg::exactTruncation(G.pot, rMax, useSentinel != 0, sentinelDistance);
    csg_status s = runGeodesics(c, G, maxT, patchRadius == nullptr, true);
    if (s) return s;
    HIPCHECK(c, download(c, dist, c->pdist, (size_t)nPairs));
    HIPCHECK(c, download(c, startTan, c->pst, 3 * (size_t)nPairs));
    HIPCHECK(c, download(c, endTan, c->pet, 3 * (size_t)nPairs));
    HIPCHECK(c, download(c, flags, c->pflags, (size_t)nPairs));
    HIPCHECK(c, download(c, forces, c->pforces, 3 * (size_t)nSrc));
    HIPCHECK(c, download(c, energy, c->penergy, (size_t)nSrc));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return CSG_OK;
}

csg_status csg_pair_forces_noalloc(csg_ctx* c, int32_t nP, const int32_t* face, const double* bary, int32_t nSrc, const int32_t* srcIdx,
                                   const int32_t* tgtOffsets, const int32_t* tgtIdx, const double* patchRadius, double rMax,
                                   const csg_potential* pot, int32_t useSentinel, double sentinelDistance, double* dist, float* startTan,
                                   float* endTan, uint32_t* flags, double* forces, double* energy) {
    if (!c) return CSG_ERR_ARGUMENT;
    if (!c->haveMesh) return fail(c, CSG_ERR_STATE, "no mesh uploaded");
    if (!c->noallocScratch) return fail(c, CSG_ERR_DEVICE, "the static scratch of csg_pair_forces_noalloc is unavailable: " + c->noallocErr);
    if (nP < 0 || nSrc < 0) return fail(c, CSG_ERR_ARGUMENT, "negative particle or source count");
    if (nSrc > kNoAllocMaxSources)
        return fail(c, CSG_ERR_CAPACITY, "csg_pair_forces_noalloc takes at most " + std::to_string(kNoAllocMaxSources) +
                                             " sources per call (the size of its static scratch); split the sources over several calls");
    if (nSrc == 0) return CSG_OK;
    if (!face || !bary || !tgtOffsets || !tgtIdx) return fail(c, CSG_ERR_ARGUMENT, "null device arrays");
    if (!patchRadius) return fail(c, CSG_ERR_ARGUMENT, "csg_pair_forces_noalloc needs patch radii (whole-mesh solves need larger tables)");
    HIPCHECK(c, hipSetDevice(c->device));
    const NoAllocLayout o = noAllocLayout();
    char* base = (char*)c->noallocScratch;
    csg::GeoParams G{};
    G.mesh = c->dm;
    G.pface = face;
    G.pbary = bary;
    G.nSrc = nSrc;
    G.srcIdx = srcIdx;
    G.toff = tgtOffsets;
    G.tidx = tgtIdx;
    G.patchRadius = patchRadius;
    G.pairDist = dist;
    G.pairStart = startTan;
    G.pairEnd = endTan;
    G.pairFlags = flags;
    G.forces = forces ? forces : (double*)(base + o.forces);
    G.energy = energy ? energy : (double*)(base + o.energy);
    G.status = (int*)(base + o.status);
    int32_t* plist = (int32_t*)(base + o.plist);
    int32_t* pcount = (int32_t*)(base + o.pcount);
    G.patchList = plist;
    G.patchCount = pcount;
    G.patchStride = csg::kPatchListCap;
    G.spill = (csg::SpillRecord*)(base + o.spill);
    G.spillCap = kNoAllocSpill;
    G.pot = pot ? csg::Potential{pot->kind, pot->p0, pot->p1, pot->p2} : csg::Potential{csg::POT_NONE, 0, 0, 0};
    G.useSentinel = useSentinel;
    G.sentinelDist = sentinelDistance;
    G.rMax = csg::exactTruncation(G.pot, rMax, useSentinel != 0, sentinelDistance);
    // the same kernels as csg_pair_forces: patches, the LDS tables, then HBM tables of a fixed size for what they left
    HIPCHECK(c, csg::launchPatch(G, plist, pcount, G.patchStride, c->stream));
    HIPCHECK(c, csg::launchGeodesicLds(G, 0, nullptr, nullptr, nullptr, c->stream));
    csg::ScratchLayout L = o.tablesLayout;
    L.base = base + o.tables;
    HIPCHECK(c, csg::launchGeodesicScratchRedo(G, L, kNoAllocFallbackBlocks, c->stream));
    int* counter = (int*)(base + o.counter);
    hipLaunchKernelGGL(kNoAllocZero, dim3(1), dim3(1), 0, c->stream, counter);
    hipLaunchKernelGGL(kNoAllocCountUnsolved, dim3((nSrc + 255) / 256), dim3(256), 0, c->stream, nSrc, (const int*)G.status, counter);
    HIPCHECK(c, hipGetLastError());
    int unsolved = 0;
    HIPCHECK(c, hipMemcpyAsync(&unsolved, counter, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (unsolved)
        return fail(c, CSG_ERR_CAPACITY, std::to_string(unsolved) + " sources exceed the tables of csg_pair_forces_noalloc (" +
                                             std::to_string(kNoAllocFallbackFaces) + " patch faces, " + std::to_string(kNoAllocFallbackTargets) +
                                             " targets per source); their outputs are not valid (csg_pair_forces grows its tables)");
    return CSG_OK;
}

csg_status csg_distance_batch(csg_ctx* c, int32_t nSrc, const int32_t* srcFace, const double* srcBary, const int32_t* tgtOffsets,
                              const int32_t* tgtFace, const double* tgtBary, const double* distanceThreshold, double* dist,
                              double* startTan, double* endTan, uint32_t* flags) {
    if (!c) return CSG_ERR_ARGUMENT;
    if (nSrc < 0 || (nSrc > 0 && (!srcFace || !srcBary || !tgtOffsets))) return fail(c, CSG_ERR_ARGUMENT, "bad source arrays");
    long nT = nSrc > 0 ? tgtOffsets[nSrc] : 0;
    if (nT > 0 && (!tgtFace || !tgtBary)) return fail(c, CSG_ERR_ARGUMENT, "bad target arrays");
    // particles: sources first, then every target
    std::vector<int32_t> face(nSrc + nT), tidx(nT);
    std::vector<double> bary(3 * (size_t)(nSrc + nT));
    std::copy(srcFace, srcFace + nSrc, face.begin());
    std::copy(tgtFace, tgtFace + nT, face.begin() + nSrc);
    std::copy(srcBary, srcBary + 3 * (size_t)nSrc, bary.begin());
    std::copy(tgtBary, tgtBary + 3 * nT, bary.begin() + 3 * (size_t)nSrc);
    for (long q = 0; q < nT; ++q) tidx[q] = (int32_t)(nSrc + q);
    std::vector<double> rad;
    if (c->submeshing) {
        rad.resize(nSrc);
        for (int s = 0; s < nSrc; ++s) {
            double thr = distanceThreshold ? distanceThreshold[s] : 1e20;
            rad[s] = thr < c->maximumDistance ? thr : c->maximumDistance;
        }
    }
    std::vector<float> st(3 * nT), et(3 * nT);
    csg_status s = csg_pair_forces(c, (int32_t)(nSrc + nT), face.data(), bary.data(), nSrc, nullptr, tgtOffsets, tidx.data(),
                                   c->submeshing ? rad.data() : nullptr, HUGE_VAL, nullptr, c->submeshing, 2.0 * c->maximumDistance, dist,
                                   startTan ? st.data() : nullptr, endTan ? et.data() : nullptr, flags, nullptr, nullptr);
    if (s) return s;
    for (long q = 0; q < 3 * nT; ++q) {
        if (startTan) startTan[q] = st[q];
        if (endTan) endTan[q] = et[q];
    }
    return CSG_OK;
}

csg_status csg_transport_batch(csg_ctx* c, int32_t n, int32_t* face, double* bary, double* disp, int32_t nVec, double* vecs, int32_t* status) {
    if (!c) return CSG_ERR_ARGUMENT;
    if (!c->haveMesh) return fail(c, CSG_ERR_STATE, "no mesh uploaded");
    if (n < 0 || (n > 0 && (!face || !bary || !disp)) || nVec < 0 || nVec > csg::kMaxTransportVectors || (nVec > 0 && !vecs))
        return fail(c, CSG_ERR_ARGUMENT, "bad transport arrays");
    if (n == 0) return CSG_OK;
    for (int i = 0; i < n; ++i)
        if (face[i] < 0 || face[i] >= c->nF) return fail(c, CSG_ERR_ARGUMENT, "particle face out of range");
    HIPCHECK(c, hipSetDevice(c->device));
    HIPCHECK(c, upload(c, c->pface, face, (size_t)n));
    HIPCHECK(c, upload(c, c->pbary, bary, 3 * (size_t)n));
    HIPCHECK(c, upload(c, c->tdisp, disp, 3 * (size_t)n));
    if (nVec) HIPCHECK(c, upload(c, c->tvecs, vecs, 3 * (size_t)n * nVec));
    HIPCHECK(c, c->tstat.ensure(4 * (size_t)n));
    HIPCHECK(c, c->pdist.ensure(24 * (size_t)n));
    csg::TransportParams T{};
    T.mesh = c->dm;
    T.n = n;
    T.face = c->pface.as<int32_t>();
    T.bary = c->pbary.as<double>();
    T.disp = c->tdisp.as<double>();
    T.scale = 1.0;
    T.dispOut = c->pdist.as<double>();
    T.nVec = nVec;
    T.vecs = c->tvecs.as<double>();
    T.status = c->tstat.as<int32_t>();
    T.maxCrossings = 100000;
    T.boundary = c->boundary;
    HIPCHECK(c, csg::launchTransport(T, c->stream));
    std::vector<int32_t> st(n);
    HIPCHECK(c, download(c, face, c->pface, (size_t)n));
    HIPCHECK(c, download(c, bary, c->pbary, 3 * (size_t)n));
    HIPCHECK(c, download(c, st.data(), c->tstat, (size_t)n));
    std::vector<double> dout(3 * (size_t)n);
    HIPCHECK(c, download(c, dout.data(), c->pdist, 3 * (size_t)n));
    if (nVec) HIPCHECK(c, download(c, vecs, c->tvecs, 3 * (size_t)n * nVec));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    bool bad = false;
    for (int i = 0; i < n; ++i) {
        if (status) status[i] = st[i];
        if (st[i] < 0)
            bad = true;
        else
            for (int a = 0; a < 3; ++a) disp[3 * i + a] = dout[3 * i + a];
    }
    if (bad) return fail(c, CSG_ERR_WALK, "a displacement walk failed; see the per-particle status");
    return CSG_OK;
}

csg_status csg_positions(csg_ctx* c, int32_t n, const int32_t* face, const double* bary, double* xyz) {
    if (!c) return CSG_ERR_ARGUMENT;
    if (!c->haveMesh) return fail(c, CSG_ERR_STATE, "no mesh uploaded");
    if (n < 0 || (n > 0 && (!face || !bary || !xyz))) return fail(c, CSG_ERR_ARGUMENT, "bad arrays");
    // host evaluation: this is a format conversion, not part of the hot path
    std::vector<double> X(3 * (size_t)c->nV);
    std::vector<int32_t> F(3 * (size_t)c->nF);
    HIPCHECK(c, hipMemcpy(X.data(), c->vx.p, X.size() * 8, hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemcpy(F.data(), c->fv.p, F.size() * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        if (face[i] < 0 || face[i] >= c->nF) return fail(c, CSG_ERR_ARGUMENT, "particle face out of range");
        const double* b = bary + 3 * i;
        double s = b[0] + b[1] + b[2];
        for (int a = 0; a < 3; ++a)
            xyz[3 * i + a] = (X[3 * F[3 * face[i]] + a] * b[0] + X[3 * F[3 * face[i] + 1] + a] * b[1] + X[3 * F[3 * face[i] + 2] + a] * b[2]) / s;
    }
    return CSG_OK;
}

}  // extern "C"
