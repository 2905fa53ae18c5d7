// C facade of the host layer, used by the Python wrappers (tests, bench). Declared in include/csg_host.h.
#include <cstring>
#include <exception>
#include <string>

#include "../../../include/csg_host.h"
#include "meshGenerators.h"

struct csgh_mesh {
    csg::HostMesh m;
};

static thread_local std::string g_hostErr;

extern "C" {

const char* csgh_last_error(void) { return g_hostErr.c_str(); }

static int wrapMesh(csg::HostMesh&& m, csgh_mesh** out) {
    *out = new csgh_mesh{std::move(m)};
    return 0;
}

int csgh_mesh_icosphere(int level, double radius, csgh_mesh** out) {
    if (level < 0 || level > 9) {
        g_hostErr = "icosphere level out of range [0,9]";
        return -1;
    }
    return wrapMesh(csg::icosphere(level, radius), out);
}

int csgh_mesh_torus(double R, double r, int nu, int nv, double jitter, uint32_t seed, csgh_mesh** out) {
    if (nu < 3 || nv < 3 || !(r > 0) || !(R > r)) {
        g_hostErr = "torus parameters out of range";
        return -1;
    }
    return wrapMesh(csg::jitteredTorus(R, r, nu, nv, jitter, seed), out);
}

int csgh_mesh_bumpy_sphere(int level, int nBumps, double amplitude, double width, uint32_t seed, csgh_mesh** out) {
    if (level < 0 || level > 9 || nBumps < 0 || !(width > 0)) {
        g_hostErr = "bumpy sphere parameters out of range";
        return -1;
    }
    return wrapMesh(csg::bumpySphere(level, nBumps, amplitude, width, seed), out);
}

int csgh_mesh_read_off(const char* path, csgh_mesh** out) {
    try {
        return wrapMesh(csg::readOFF(path), out);
    } catch (std::exception& e) {
        g_hostErr = e.what();
        return -1;
    }
}

int csgh_mesh_from_arrays(int nV, const double* xyz, int nF, const int32_t* tri, csgh_mesh** out) {
    csg::HostMesh m;
    m.xyz.assign(xyz, xyz + 3 * (size_t)nV);
    m.tri.assign(tri, tri + 3 * (size_t)nF);
    return wrapMesh(std::move(m), out);
}

int csgh_mesh_size(const csgh_mesh* m, int* nV, int* nF) {
    *nV = m->m.nV();
    *nF = m->m.nF();
    return 0;
}

int csgh_mesh_copy(const csgh_mesh* m, double* xyz, int32_t* tri) {
    std::memcpy(xyz, m->m.xyz.data(), m->m.xyz.size() * sizeof(double));
    std::memcpy(tri, m->m.tri.data(), m->m.tri.size() * sizeof(int32_t));
    return 0;
}

void csgh_mesh_free(csgh_mesh* m) { delete m; }

}  // extern "C"
