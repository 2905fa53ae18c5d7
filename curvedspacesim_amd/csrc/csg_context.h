// Internal definition of the C-ABI context (shared by the translation units of libcsg.so).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/csg.h"
#include "csg_common.h"
#include "geodesic_kernel.h"

namespace csg {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
        size_t want = std::max<size_t>(n, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
};

struct SystemState;  // device-resident particle system (csg_system.hip)
void destroySystem(SystemState* s);

}  // namespace csg

struct csg_ctx {
    using DevBuf = csg::DevBuf;
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::string arch;
    // mesh
    int nV = 0, nF = 0;
    double area = 0, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    DevBuf vx, fv, fadj, fgeo, fnrm, vflag, vfoff, vflist, curv, fchild, frule;
    csg::DeviceMesh dm{};
    std::vector<float> faceCentroid;  // host copy, 3 per face
    double maxEdge = 0;               // longest edge of the mesh
    bool haveMesh = false;
    int submeshing = 0;
    double maximumDistance = 0;
    int boundary = CSG_BOUNDARY_CLOSED;  // walk behaviour at the mesh border
    // per-call staging
    DevBuf pface, pbary, srcIdx, toff, tidx, prad, pdist, pst, pet, pflags, pforces, penergy, status, list, plist, pcount, spill;
    DevBuf tdisp, tvecs, tstat;
    DevBuf sX, sY, sOut, sPart, sGrad, sParams, sM, sV;  // learned-screen staging
    DevBuf scratch;
    int scratchBlocks = 0;
    csg::ScratchLayout layout{};
    csg::SystemState* sys = nullptr;
    // device address of the static scratch array of csg_pair_forces_noalloc on this context's GPU (null when it could
    // not be resolved; noallocErr then says why)
    void* noallocScratch = nullptr;
    std::string noallocErr;
    ~csg_ctx() {
        if (sys) csg::destroySystem(sys);
        DevBuf* all[] = {&vx, &fv, &fadj, &fgeo, &fnrm, &vflag, &vfoff, &vflist, &curv, &fchild, &frule, &sX, &sY, &sOut, &sPart, &sGrad, &sParams, &sM, &sV, &pface, &pbary, &srcIdx, &toff, &tidx, &prad,
                         &pdist, &pst, &pet, &pflags, &pforces, &penergy, &status, &list, &plist, &pcount, &spill, &tdisp, &tvecs, &tstat, &scratch};
        for (DevBuf* b : all) b->release();
        if (stream) hipStreamDestroy(stream);
    }
};

namespace csg {

// An upper bound on the faces of any submesh patch of radius `radius` (a face joins a patch when a corner lies within
// the radius of the source, so its centroid lies within radius + the longest edge): with the centroids binned in cubic
// cells of that size, every patch lies in the 27 cells around its source's cell. Host work, linear in the faces.
long patchFaceBound(const csg_ctx* c, double radius);

inline csg_status fail(csg_ctx* c, csg_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}

#define HIPCHECK(ctx, expr)                                                                                      \
    do {                                                                                                         \
        hipError_t e_ = (expr);                                                                                  \
        if (e_ != hipSuccess) return csg::fail(ctx, CSG_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

}  // namespace csg
