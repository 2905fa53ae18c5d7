// Trajectory and checkpoint files in the layout of the reference's simpleModelDatabase
// (src/databases/simpleModelDatabase.cpp:8-158 over src/databases/baseHDF5Database.cpp): an HDF5 file whose root holds
// one two-dimensional dataset per quantity, a row per record, extendable without limit in the number of rows and
// chunked one row at a time:
//   time                 double, 1 column
//   R3position           double, 3N  (x, y, z of each particle)
//   barycentricPosition  double, 3N  (the barycentric weights of each particle in its face)
//   faceIndex            int,    N
//   velocity             double, 3N  (if saved)
//   force                double, 3N  (if saved)
//   type                 int,    N   (if saved)
// Files of either side can be read by the other. This build also records what the reference does not save, so that a
// resumed Nose-Hoover run continues the same trajectory: two more datasets, written by writeThermostat,
//   nhChain              double, 4 (M + 1): position, velocity, acceleration and mass of each chain entry
//   nhParams             double, 4: M, T, tau and the chain's kinetic energy
// Built only where the HDF5 C library is available (libcsg_h5.so).
#pragma once

#include <hdf5.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "simpleModel.h"

// how a database file is opened (src/databases/baseDatabase.h:7-17)
struct fileMode {
    enum Enum { readonly, readwrite, replace };
};

// Extendable two-dimensional datasets at the root of an HDF5 file: the part of the reference's baseHDF5Database that
// simpleModelDatabase uses (src/databases/baseHDF5Database.cpp:86-233). Errors throw std::runtime_error.
class hdf5RecordFile {
public:
    hdf5RecordFile(const std::string& filename, fileMode::Enum mode);
    ~hdf5RecordFile();
    hdf5RecordFile(const hdf5RecordFile&) = delete;
    hdf5RecordFile& operator=(const hdf5RecordFile&) = delete;

    bool hasDataset(const std::string& name) const;
    // a dataset of rows of `columns` values with no row yet
    template <typename T>
    void registerExtendableDataset(const std::string& name, int columns);
    // appends one row (data.size() must equal the dataset's columns)
    template <typename T>
    void extendDataset(const std::string& name, const std::vector<T>& data);
    // row `record` (the last for record < 0) into data, resized to the dataset's columns
    template <typename T>
    void readDataset(const std::string& name, std::vector<T>& data, int record = -1) const;
    unsigned long getDatasetDimensions(const std::string& name) const;  // number of rows
    unsigned long datasetColumns(const std::string& name) const;

    std::string filename;
    fileMode::Enum mode;

protected:
    hid_t file = -1;
};

// simpleModelDatabase (src/databases/simpleModelDatabase.h:15-48): records of a simpleModel's state
class simpleModelDatabase : public hdf5RecordFile {
public:
    simpleModelDatabase(int numberOfParticles, const std::string& fn = "temp.h5", fileMode::Enum mode = fileMode::readonly,
                        bool saveVelocities = true, bool saveTypes = false, bool saveForces = false);
    // appends the model's state as a record (record >= 0, overwriting a record, is refused as the reference does)
    virtual void writeState(std::shared_ptr<simpleModel> s, double time = -1.0, int record = -1);
    // positions (barycentric and face), and the saved ones of velocities, forces and types, of record `rec`
    // (the last for rec < 0); R3position is not read, as in the reference
    virtual void readState(std::shared_ptr<simpleModel> s, int rec);
    unsigned long currentNumberOfRecords() { return getDatasetDimensions("time"); }

    // the same as arrays: face N, bary 3N, r3 3N, vel 3N, force 3N (nullable where the dataset is not saved), types N
    void writeRecord(double time, const int* face, const double* bary, const double* r3, const double* vel, const double* force,
                     const int* types);
    // face, bary, vel, force, types and time of a record (every pointer nullable); returns the record read
    long readRecord(long rec, double* time, int* face, double* bary, double* r3, double* vel, double* force, int* types) const;
    // the Nose-Hoover chain (M + 1 entries of position, velocity, acceleration, mass) and M, T, tau, kinetic energy:
    // appended as a record of nhChain / nhParams (registered on first use), and read back; M = 0 writes an empty chain
    void writeThermostat(int M, const double* chain, double T, double tau, double kineticEnergy);
    // returns M (0 when no thermostat record exists); chain receives 4 (M + 1) values (nullable), params M, T, tau, KE
    int readThermostat(long rec, double* chain, double* params) const;

    int N;
    bool velocity, type, force;

private:
    void registerDatasets();
};
