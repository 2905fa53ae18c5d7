#include "simpleModelDatabase.h"

#include <sys/stat.h>

#include <algorithm>

namespace {

template <typename T>
hid_t nativeType();
template <>
hid_t nativeType<int>() {
    return H5T_NATIVE_INT;
}
template <>
hid_t nativeType<double>() {
    return H5T_NATIVE_DOUBLE;
}

// closes an HDF5 identifier when it leaves scope
struct Handle {
    hid_t id;
    herr_t (*close)(hid_t);
    Handle(hid_t i, herr_t (*c)(hid_t), const char* what) : id(i), close(c) {
        if (id < 0) throw std::runtime_error(std::string("HDF5: cannot ") + what);
    }
    ~Handle() {
        if (id >= 0) close(id);
    }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
};

bool fileExists(const std::string& f) {
    struct stat st;
    return stat(f.c_str(), &st) == 0;
}

void check(herr_t e, const char* what) {
    if (e < 0) throw std::runtime_error(std::string("HDF5: ") + what + " failed");
}

// the HDF5 library's own error stack printing is off: failures are reported by exceptions
struct QuietErrors {
    QuietErrors() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};

}  // namespace

hdf5RecordFile::hdf5RecordFile(const std::string& fn, fileMode::Enum m) : filename(fn), mode(m) {
    static QuietErrors quiet;
    if (mode == fileMode::readonly) {
        if (!fileExists(filename)) throw std::runtime_error("trying to read a file that does not exist: " + filename);
        file = H5Fopen(filename.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    } else if (mode == fileMode::readwrite && fileExists(filename)) {
        file = H5Fopen(filename.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
    } else {
        file = H5Fcreate(filename.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    }
    if (file < 0) throw std::runtime_error("HDF5: cannot open " + filename);
}

hdf5RecordFile::~hdf5RecordFile() {
    if (file >= 0) H5Fclose(file);
}

bool hdf5RecordFile::hasDataset(const std::string& name) const { return H5Lexists(file, name.c_str(), H5P_DEFAULT) > 0; }

template <typename T>
void hdf5RecordFile::registerExtendableDataset(const std::string& name, int columns) {
    if (mode == fileMode::readonly) throw std::runtime_error("don't write on a readonly file");
    if (columns < 1) throw std::runtime_error("a dataset needs at least one column: " + name);
    const hsize_t dims[2] = {0, (hsize_t)columns}, maxDims[2] = {H5S_UNLIMITED, (hsize_t)columns}, chunk[2] = {1, (hsize_t)columns};
    Handle space(H5Screate_simple(2, dims, maxDims), H5Sclose, "make a data space");
    Handle props(H5Pcreate(H5P_DATASET_CREATE), H5Pclose, "make a property list");
    check(H5Pset_layout(props.id, H5D_CHUNKED), "set the chunked layout");
    check(H5Pset_chunk(props.id, 2, chunk), "set the chunk size");
    Handle ds(H5Dcreate2(file, name.c_str(), nativeType<T>(), space.id, H5P_DEFAULT, props.id, H5P_DEFAULT), H5Dclose,
              ("create dataset " + name).c_str());
}

template <typename T>
void hdf5RecordFile::extendDataset(const std::string& name, const std::vector<T>& data) {
    if (mode == fileMode::readonly) throw std::runtime_error("don't write on a readonly file");
    Handle ds(H5Dopen2(file, name.c_str(), H5P_DEFAULT), H5Dclose, ("open dataset " + name).c_str());
    hsize_t dims[2];
    {
        Handle space(H5Dget_space(ds.id), H5Sclose, "get a data space");
        if (H5Sget_simple_extent_ndims(space.id) != 2) throw std::runtime_error("dataset is not two-dimensional: " + name);
        H5Sget_simple_extent_dims(space.id, dims, nullptr);
    }
    if (data.size() != dims[1]) throw std::runtime_error("trying to write a vector of the wrong size for the dataset " + name);
    const hsize_t grown[2] = {dims[0] + 1, dims[1]};
    check(H5Dset_extent(ds.id, grown), "extend a dataset");
    const hsize_t row[2] = {1, dims[1]}, offset[2] = {dims[0], 0};
    Handle memory(H5Screate_simple(2, row, nullptr), H5Sclose, "make a memory space");
    Handle space(H5Dget_space(ds.id), H5Sclose, "get a data space");
    check(H5Sselect_hyperslab(space.id, H5S_SELECT_SET, offset, nullptr, row, nullptr), "select a row");
    check(H5Dwrite(ds.id, nativeType<T>(), memory.id, space.id, H5P_DEFAULT, data.data()), "write a row");
    check(H5Fflush(file, H5F_SCOPE_GLOBAL), "flush");
}

template <typename T>
void hdf5RecordFile::readDataset(const std::string& name, std::vector<T>& data, int record) const {
    Handle ds(H5Dopen2(file, name.c_str(), H5P_DEFAULT), H5Dclose, ("open dataset " + name).c_str());
    Handle space(H5Dget_space(ds.id), H5Sclose, "get a data space");
    hsize_t dims[2];
    if (H5Sget_simple_extent_ndims(space.id) != 2) throw std::runtime_error("reading a dataset that isn't two dimensional: " + name);
    H5Sget_simple_extent_dims(space.id, dims, nullptr);
    data.resize(dims[1]);
    const hsize_t r = record < 0 ? dims[0] - 1 : (hsize_t)record;
    if (dims[0] == 0 || r >= dims[0]) throw std::runtime_error("trying to read past the end of the dataset " + name);
    const hsize_t row[2] = {1, dims[1]}, offset[2] = {r, 0};
    check(H5Sselect_hyperslab(space.id, H5S_SELECT_SET, offset, nullptr, row, nullptr), "select a row");
    Handle memory(H5Screate_simple(2, row, nullptr), H5Sclose, "make a memory space");
    check(H5Dread(ds.id, nativeType<T>(), memory.id, space.id, H5P_DEFAULT, data.data()), "read a row");
}

static void extents(hid_t file, const std::string& name, hsize_t dims[2]) {
    Handle ds(H5Dopen2(file, name.c_str(), H5P_DEFAULT), H5Dclose, ("open dataset " + name).c_str());
    Handle space(H5Dget_space(ds.id), H5Sclose, "get a data space");
    if (H5Sget_simple_extent_ndims(space.id) != 2) throw std::runtime_error("dataset is not two-dimensional: " + name);
    H5Sget_simple_extent_dims(space.id, dims, nullptr);
}

unsigned long hdf5RecordFile::getDatasetDimensions(const std::string& name) const {
    hsize_t dims[2];
    extents(file, name, dims);
    return (unsigned long)dims[0];
}

unsigned long hdf5RecordFile::datasetColumns(const std::string& name) const {
    hsize_t dims[2];
    extents(file, name, dims);
    return (unsigned long)dims[1];
}

template void hdf5RecordFile::registerExtendableDataset<int>(const std::string&, int);
template void hdf5RecordFile::registerExtendableDataset<double>(const std::string&, int);
template void hdf5RecordFile::extendDataset<int>(const std::string&, const std::vector<int>&);
template void hdf5RecordFile::extendDataset<double>(const std::string&, const std::vector<double>&);
template void hdf5RecordFile::readDataset<int>(const std::string&, std::vector<int>&, int) const;
template void hdf5RecordFile::readDataset<double>(const std::string&, std::vector<double>&, int) const;

// --- simpleModelDatabase ------------------------------------------------------------------------------------------

simpleModelDatabase::simpleModelDatabase(int numberOfParticles, const std::string& fn, fileMode::Enum m, bool saveVelocities,
                                         bool saveTypes, bool saveForces)
    : hdf5RecordFile(fn, m), N(numberOfParticles), velocity(saveVelocities), type(saveTypes), force(saveForces) {
    if (N < 1) throw std::runtime_error("a database needs at least one particle");
    if (mode == fileMode::replace || (mode == fileMode::readwrite && !hasDataset("time"))) registerDatasets();
    if (hasDataset("faceIndex") && datasetColumns("faceIndex") != (unsigned long)N)
        throw std::runtime_error("the file holds records of another number of particles");
}

void simpleModelDatabase::registerDatasets() {
    registerExtendableDataset<double>("time", 1);
    registerExtendableDataset<double>("R3position", 3 * N);
    registerExtendableDataset<double>("barycentricPosition", 3 * N);
    registerExtendableDataset<int>("faceIndex", N);
    if (velocity) registerExtendableDataset<double>("velocity", 3 * N);
    if (force) registerExtendableDataset<double>("force", 3 * N);
    if (type) registerExtendableDataset<int>("type", N);
}

void simpleModelDatabase::writeRecord(double time, const int* face, const double* bary, const double* r3, const double* vel,
                                      const double* frc, const int* types) {
    if (!face || !bary || !r3 || (velocity && !vel) || (force && !frc) || (type && !types))
        throw std::runtime_error("a record needs every saved quantity");
    extendDataset("time", std::vector<double>{time});
    extendDataset("R3position", std::vector<double>(r3, r3 + 3 * N));
    extendDataset("barycentricPosition", std::vector<double>(bary, bary + 3 * N));
    extendDataset("faceIndex", std::vector<int>(face, face + N));
    if (velocity) extendDataset("velocity", std::vector<double>(vel, vel + 3 * N));
    if (force) extendDataset("force", std::vector<double>(frc, frc + 3 * N));
    if (type) extendDataset("type", std::vector<int>(types, types + N));
}

long simpleModelDatabase::readRecord(long rec, double* time, int* face, double* bary, double* r3, double* vel, double* frc,
                                     int* types) const {
    const long n = (long)getDatasetDimensions("time");
    const long r = rec < 0 ? n - 1 : rec;
    if (n == 0 || r >= n) throw std::runtime_error("trying to read past the end of the dataset time");
    std::vector<double> d;
    std::vector<int> iv;
    auto get = [&](const char* name, double* out, int cols) {
        if (!out) return;
        if (!hasDataset(name)) throw std::runtime_error(std::string("the file does not hold ") + name);
        readDataset(name, d, (int)r);
        if ((int)d.size() != cols) throw std::runtime_error(std::string("wrong record size in ") + name);
        std::copy(d.begin(), d.end(), out);
    };
    auto geti = [&](const char* name, int* out) {
        if (!out) return;
        if (!hasDataset(name)) throw std::runtime_error(std::string("the file does not hold ") + name);
        readDataset(name, iv, (int)r);
        if ((int)iv.size() != N) throw std::runtime_error(std::string("wrong record size in ") + name);
        std::copy(iv.begin(), iv.end(), out);
    };
    get("time", time, 1);
    geti("faceIndex", face);
    get("barycentricPosition", bary, 3 * N);
    get("R3position", r3, 3 * N);
    get("velocity", vel, 3 * N);
    get("force", frc, 3 * N);
    geti("type", types);
    return r;
}

void simpleModelDatabase::writeState(std::shared_ptr<simpleModel> s, double time, int record) {
    if (record >= 0) throw std::runtime_error("overwriting specific records not implemented at the moment");
    if (s->N != N) throw std::runtime_error("the model holds another number of particles than the database");
    s->fillEuclideanLocations();
    std::vector<double> r3(3 * N), bary(3 * N), vel(3 * N), frc(3 * N);
    std::vector<int> face(N), types(N, 0);
    for (int i = 0; i < N; ++i) {
        r3[3 * i] = s->euclideanLocations[i].x;
        r3[3 * i + 1] = s->euclideanLocations[i].y;
        r3[3 * i + 2] = s->euclideanLocations[i].z;
        for (int a = 0; a < 3; ++a) {
            bary[3 * i + a] = s->positions[i].x[a];
            if (velocity) vel[3 * i + a] = s->velocities[i][a];
            if (force) frc[3 * i + a] = s->forces[i][a];
        }
        face[i] = s->positions[i].faceIndex;
        if (type && i < (int)s->types.size()) types[i] = s->types[i];
    }
    writeRecord(time, face.data(), bary.data(), r3.data(), vel.data(), frc.data(), types.data());
}

void simpleModelDatabase::readState(std::shared_ptr<simpleModel> s, int rec) {
    if (s->N != N) throw std::runtime_error("the model holds another number of particles than the database");
    std::vector<double> bary(3 * N), vel(3 * N), frc(3 * N);
    std::vector<int> face(N), types(N);
    double t;
    readRecord(rec, &t, face.data(), bary.data(), nullptr, velocity ? vel.data() : nullptr, force ? frc.data() : nullptr,
               type ? types.data() : nullptr);
    for (int i = 0; i < N; ++i) {
        s->positions[i] = meshPosition(point3(bary[3 * i], bary[3 * i + 1], bary[3 * i + 2]), face[i]);
        if (velocity) s->velocities[i] = vector3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        if (force) s->forces[i] = vector3(frc[3 * i], frc[3 * i + 1], frc[3 * i + 2]);
        if (type) {
            if ((int)s->types.size() < N) s->types.resize(N);
            s->types[i] = types[i];
        }
    }
}

void simpleModelDatabase::writeThermostat(int M, const double* chain, double T, double tau, double kineticEnergy) {
    if (M < 0 || (M > 0 && !chain)) throw std::runtime_error("bad thermostat chain");
    if (!hasDataset("nhParams")) {
        registerExtendableDataset<double>("nhParams", 4);
        registerExtendableDataset<double>("nhChain", 4 * (std::max(M, 1) + 1));
    }
    const unsigned long cols = datasetColumns("nhChain");
    if (M > 0 && cols != 4ul * (M + 1)) throw std::runtime_error("the file's thermostat records hold another chain length");
    std::vector<double> c(cols, 0.0);
    if (M > 0) std::copy(chain, chain + 4 * (M + 1), c.begin());
    extendDataset("nhChain", c);
    extendDataset("nhParams", std::vector<double>{(double)M, T, tau, kineticEnergy});
}

int simpleModelDatabase::readThermostat(long rec, double* chain, double* params) const {
    if (!hasDataset("nhParams")) {
        if (params) std::fill(params, params + 4, 0.0);
        return 0;
    }
    const long n = (long)getDatasetDimensions("nhParams");
    const long r = rec < 0 ? n - 1 : rec;
    if (n == 0 || r >= n) throw std::runtime_error("trying to read past the end of the dataset nhParams");
    std::vector<double> p, c;
    readDataset("nhParams", p, (int)r);
    readDataset("nhChain", c, (int)r);
    const int M = (int)p[0];
    if (params) std::copy(p.begin(), p.end(), params);
    if (chain && M > 0) std::copy(c.begin(), c.begin() + 4 * (M + 1), chain);
    return M;
}
