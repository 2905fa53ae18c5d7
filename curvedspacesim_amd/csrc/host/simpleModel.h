// simpleModel (src/models/simpleModel.h:25-111): particle state plus the findNeighbors / moveParticles seams, with the
// reference's per-particle loops over any baseSpace; and gpuModel, which overrides both seams with one batched device
// call each when its space is a gpuMeshSpace.
#pragma once

#include <memory>
#include <vector>

#include "baseSpace.h"
#include "neighborStructures.h"

class simpleModel {
public:
    explicit simpleModel(int n);
    simpleModel() = default;
    virtual ~simpleModel() = default;
    void initializeSimpleModel(int n);
    virtual void setSpace(std::shared_ptr<baseSpace> _space) { space = _space; }
    virtual void setNeighborStructure(std::shared_ptr<baseNeighborStructure> s) { neighborStructure = s; }
    // displacements are in/out, as the space's transport mutates them (src/models/simpleModel.cpp:44-66)
    virtual void moveParticles(std::vector<vector3>& displacements);
    virtual int getNumberOfParticles() { return N; }
    virtual void computeForces(bool zeroOutForces = true);
    // candidates from the neighbour structure, then the space's geodesic distances and start tangents
    // (src/models/simpleModel.cpp:68-112)
    virtual void findNeighbors(double maximumInteractionRange);
    virtual void setParticlePositions(std::vector<meshPosition>& newPositions);
    virtual void setRandomParticlePositions(noiseSource& noise);
    virtual void setMaxwellBoltzmannVelocities(noiseSource& noise, double T);
    virtual void fillEuclideanLocations();
    void setVerbose(bool v) { verbose = v; }

    std::shared_ptr<baseSpace> space;
    int N = 0;
    std::vector<meshPosition> positions;
    std::vector<vector3> velocities;
    std::vector<vector3> forces;
    std::vector<std::vector<int>> neighbors;
    std::vector<std::vector<vector3>> neighborVectors;
    std::vector<std::vector<double>> neighborDistances;
    std::vector<int> types;
    std::vector<double> masses;
    std::vector<double3> euclideanLocations;
    bool particleShiftsRequireVelocityTransport = false;
    bool particleShiftsRequireForceTransport = false;

protected:
    // the candidate lists of every particle and their largest Euclidean distances
    void gatherCandidates(double maximumInteractionRange, std::vector<std::vector<meshPosition>>& targets, std::vector<double>& largest);
    std::shared_ptr<baseNeighborStructure> neighborStructure;
    std::vector<meshPosition> euclideanMeshPosition;
    bool verbose = false;
};

class gpuMeshSpace;

class gpuModel : public simpleModel {
public:
    explicit gpuModel(int n) : simpleModel(n) {}
    void setSpace(std::shared_ptr<baseSpace> _space) override;
    // one batched geodesic launch for all sources
    void findNeighbors(double maximumInteractionRange) override;
    // one batched walk launch for all particles
    void moveParticles(std::vector<vector3>& displacements) override;
    gpuMeshSpace* meshSpace() { return gpuSpace; }
    // the neighbour structure's candidates of every particle as CSR (offsets N+1, indices) and their largest
    // Euclidean distances; fills `neighbors` as findNeighbors does
    void candidateLists(double maximumInteractionRange, std::vector<int32_t>& offsets, std::vector<int32_t>& indices,
                        std::vector<double>& largest);

private:
    gpuMeshSpace* gpuSpace = nullptr;
};

typedef std::shared_ptr<simpleModel> ConfigPtr;
typedef std::weak_ptr<simpleModel> WeakConfigPtr;
