/* C facade of the C++ host layer (libcsg_host.so): the meshes of the workloads (generators and an OFF reader). The
 * reference-API classes of that library (simpleModel, force, updater, simulation and the gpu* adapters) are C++ and
 * are used directly from C++ (tests/cpp/refapi_scenarios.cpp); the reference has no C interface for meshes, and this
 * facade exists so that scripting callers (the Python tests and the benchmark) can build them.
 * All functions return 0 on success and a negative value on failure; csgh_last_error() then describes it.
 */
#ifndef CSG_HOST_H
#define CSG_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct csgh_mesh csgh_mesh;

const char* csgh_last_error(void);

/* Meshes. The reference loads OFF files (src/models/triangulatedMeshSpace.cpp:42-72); these also generate the
 * benchmark surfaces. */
int csgh_mesh_icosphere(int level, double radius, csgh_mesh** out);
int csgh_mesh_torus(double R, double r, int nu, int nv, double jitter, uint32_t seed, csgh_mesh** out);
int csgh_mesh_bumpy_sphere(int level, int nBumps, double amplitude, double width, uint32_t seed, csgh_mesh** out);
int csgh_mesh_read_off(const char* path, csgh_mesh** out);
int csgh_mesh_from_arrays(int nV, const double* xyz, int nF, const int32_t* tri, csgh_mesh** out);
int csgh_mesh_size(const csgh_mesh* m, int* nV, int* nF);
int csgh_mesh_copy(const csgh_mesh* m, double* xyz, int32_t* tri);
void csgh_mesh_free(csgh_mesh* m);

#ifdef __cplusplus
}
#endif
#endif
