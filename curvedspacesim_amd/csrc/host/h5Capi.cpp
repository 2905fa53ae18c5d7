// C facade of simpleModelDatabase (include/csg_h5.h).
#include <string>

#include "../../../include/csg_h5.h"
#include "simpleModelDatabase.h"

struct csgh5_db {
    simpleModelDatabase db;
    csgh5_db(const char* path, int N, fileMode::Enum m, bool v, bool t, bool f) : db(N, path, m, v, t, f) {}
};

static thread_local std::string g_err;

template <class F>
static auto guard(F f, decltype(f()) bad) -> decltype(f()) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_err = e.what();
        return bad;
    }
}

extern "C" {

const char* csgh5_last_error(void) { return g_err.c_str(); }

int csgh5_open(const char* path, int32_t N, int32_t mode, int32_t saveVelocities, int32_t saveTypes, int32_t saveForces, csgh5_db** out) {
    if (!path || !out || mode < 0 || mode > 2) {
        g_err = "bad arguments";
        return -1;
    }
    *out = nullptr;
    return guard(
        [&]() {
            *out = new csgh5_db(path, N, (fileMode::Enum)mode, saveVelocities != 0, saveTypes != 0, saveForces != 0);
            return 0;
        },
        -1);
}

void csgh5_close(csgh5_db* db) { delete db; }

int64_t csgh5_records(const csgh5_db* db) {
    if (!db) return -1;
    return guard([&]() { return (int64_t)const_cast<simpleModelDatabase&>(db->db).currentNumberOfRecords(); }, (int64_t)-1);
}

int csgh5_write(csgh5_db* db, double time, const int32_t* face, const double* bary, const double* r3, const double* vel,
                const double* force, const int32_t* types) {
    if (!db) return -1;
    return guard(
        [&]() {
            db->db.writeRecord(time, face, bary, r3, vel, force, types);
            return 0;
        },
        -1);
}

int64_t csgh5_read(const csgh5_db* db, int64_t rec, double* time, int32_t* face, double* bary, double* r3, double* vel, double* force,
                   int32_t* types) {
    if (!db) return -1;
    return guard([&]() { return (int64_t)db->db.readRecord((long)rec, time, face, bary, r3, vel, force, types); }, (int64_t)-1);
}

int csgh5_write_thermostat(csgh5_db* db, int32_t M, const double* chain, double T, double tau, double kineticEnergy) {
    if (!db) return -1;
    return guard(
        [&]() {
            db->db.writeThermostat(M, chain, T, tau, kineticEnergy);
            return 0;
        },
        -1);
}

int csgh5_read_thermostat(const csgh5_db* db, int64_t rec, double* chain, double* params) {
    if (!db) return -1;
    return guard([&]() { return db->db.readThermostat((long)rec, chain, params); }, -1);
}

int64_t csgh5_columns(const char* path, const char* dataset) {
    if (!path || !dataset) return -1;
    return guard(
        [&]() {
            hdf5RecordFile f(path, fileMode::readonly);
            if (!f.hasDataset(dataset)) throw std::runtime_error(std::string("the file does not hold ") + dataset);
            return (int64_t)f.datasetColumns(dataset);
        },
        (int64_t)-1);
}

}  // extern "C"
