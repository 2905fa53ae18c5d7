/* libcsg_h5.so — trajectory and checkpoint files in the HDF5 layout of the reference's simpleModelDatabase
 * (src/databases/simpleModelDatabase.cpp:8-158, src/databases/baseHDF5Database.cpp:86-233): root datasets time,
 * R3position, barycentricPosition, faceIndex and optionally velocity, force and type, two-dimensional with a row per
 * record. The C++ class simpleModelDatabase (curvedspacesim_amd/csrc/host/simpleModelDatabase.h) reads and writes
 * simpleModel states; this facade does the same from arrays, for scripting callers. It also keeps the Nose-Hoover
 * chain, which the reference does not save (datasets nhChain and nhParams), so that a resumed run continues the same
 * trajectory. Built only where the HDF5 C library is available. Functions return 0 on success and a negative value on
 * failure; csgh5_last_error() then describes it.
 */
#ifndef CSG_H5_H
#define CSG_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct csgh5_db csgh5_db;

const char* csgh5_last_error(void);
/* simpleModelDatabase(numberOfParticles, fn, mode, saveVelocities, saveTypes, saveForces) (simpleModelDatabase.cpp:8-35);
 * mode 0 readonly, 1 readwrite (appends to an existing file), 2 replace */
int csgh5_open(const char* path, int32_t N, int32_t mode, int32_t saveVelocities, int32_t saveTypes, int32_t saveForces, csgh5_db** out);
void csgh5_close(csgh5_db* db);
/* currentNumberOfRecords (simpleModelDatabase.cpp:52) */
int64_t csgh5_records(const csgh5_db* db);
/* writeState (simpleModelDatabase.cpp:63-120) from arrays: face N, bary 3N, r3 3N; vel, force 3N and types N are
 * needed when the file saves them (nullable otherwise) */
int csgh5_write(csgh5_db* db, double time, const int32_t* face, const double* bary, const double* r3, const double* vel,
                const double* force, const int32_t* types);
/* readState (simpleModelDatabase.cpp:122-158) into arrays: record rec (the last for rec < 0); every pointer nullable;
 * returns the record read, or a negative value */
int64_t csgh5_read(const csgh5_db* db, int64_t rec, double* time, int32_t* face, double* bary, double* r3, double* vel, double* force,
                   int32_t* types);
/* Nose-Hoover chain (no reference counterpart): M, chain 4 (M + 1) doubles (position, velocity, acceleration, mass
 * per entry), T, tau and the chain's kinetic energy, appended as a record; csgh5_read_thermostat returns M of record
 * rec (the last for rec < 0; 0 when the file holds none) and fills chain (nullable) and params[4] = M, T, tau, KE */
int csgh5_write_thermostat(csgh5_db* db, int32_t M, const double* chain, double T, double tau, double kineticEnergy);
int csgh5_read_thermostat(const csgh5_db* db, int64_t rec, double* chain, double* params);
/* columns of a root dataset of a file (faceIndex: the number of particles), read without knowing N */
int64_t csgh5_columns(const char* path, const char* dataset);

#ifdef __cplusplus
}
#endif
#endif
