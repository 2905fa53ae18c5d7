#include "simpleModel.h"

#include <cmath>
#include <cstdio>
#include <stdexcept>

#include "gpuMeshSpace.h"

simpleModel::simpleModel(int n) { initializeSimpleModel(n); }

void simpleModel::initializeSimpleModel(int n) {
    N = n;
    positions.assign(n, meshPosition());
    euclideanLocations.assign(n, double3());
    velocities.assign(n, vector3(0, 0, 0));
    forces.assign(n, vector3(0, 0, 0));
    masses.assign(n, 1.0);
    types.assign(n, 0);
    neighbors.assign(n, {});
    neighborVectors.assign(n, {});
    neighborDistances.assign(n, {});
    neighborStructure = std::make_shared<baseNeighborStructure>();
}

void simpleModel::fillEuclideanLocations() { space->meshPositionToEuclideanLocation(positions, euclideanLocations); }

void simpleModel::moveParticles(std::vector<vector3>& disp) {
    for (int i = 0; i < N; ++i) {
        std::vector<vector3> transports;
        if (particleShiftsRequireForceTransport) transports.push_back(forces[i]);
        if (particleShiftsRequireVelocityTransport) transports.push_back(velocities[i]);
        space->transportParticleAndVectors(positions[i], disp[i], transports);
        int k = 0;
        if (particleShiftsRequireForceTransport) forces[i] = transports[k++];
        if (particleShiftsRequireVelocityTransport) velocities[i] = transports[k++];
    }
}

void simpleModel::gatherCandidates(double range, std::vector<std::vector<meshPosition>>& targets, std::vector<double>& largest) {
    neighborStructure->setInteractionRange(range);
    const bool euclid = neighborStructure->requireEuclideanPositions && !space->positionsAreEuclidean;
    if (euclid) {
        space->meshPositionToEuclideanLocation(positions, euclideanMeshPosition);
        neighborStructure->initialize(euclideanMeshPosition);
    } else {
        neighborStructure->initialize(positions);
    }
    targets.assign(N, {});
    largest.assign(N, 0.0);
    for (int i = 0; i < N; ++i) {
        largest[i] = neighborStructure->constructCandidateNeighborList(euclid ? euclideanMeshPosition[i] : positions[i], i, neighbors[i],
                                                                       targets[i]);
        if (euclid)
            for (size_t j = 0; j < neighbors[i].size(); ++j) targets[i][j] = positions[neighbors[i][j]];
    }
}

void simpleModel::findNeighbors(double range) {
    std::vector<std::vector<meshPosition>> targets;
    std::vector<double> largest;
    gatherCandidates(range, targets, largest);
    double meanNN = 0;
    for (int i = 0; i < N; ++i) {
        std::vector<vector3> placeholder;
        space->distance(positions[i], targets[i], neighborDistances[i], neighborVectors[i], placeholder, largest[i]);
        meanNN += neighborDistances[i].size();
    }
    if (verbose) printf("mean number of neighbors = %f\n", meanNN / N);
}

void simpleModel::computeForces(bool zeroOutForces) {
    if (zeroOutForces)
        for (vector3& f : forces) f = vector3(0, 0, 0);
}

void simpleModel::setParticlePositions(std::vector<meshPosition>& p) {
    if ((int)p.size() != N) initializeSimpleModel((int)p.size());
    positions = p;
}

void simpleModel::setRandomParticlePositions(noiseSource& noise) {
    for (int i = 0; i < N; ++i) space->randomPosition(positions[i], noise);
}

void simpleModel::setMaxwellBoltzmannVelocities(noiseSource& noise, double T) {
    for (int i = 0; i < N; ++i) {
        space->randomVectorAtPosition(positions[i], velocities[i], noise);
        velocities[i] *= std::sqrt(T);
    }
}

void gpuModel::setSpace(std::shared_ptr<baseSpace> _space) {
    simpleModel::setSpace(_space);
    gpuSpace = dynamic_cast<gpuMeshSpace*>(_space.get());
}

void gpuModel::findNeighbors(double range) {
    if (!gpuSpace) {
        simpleModel::findNeighbors(range);
        return;
    }
    std::vector<std::vector<meshPosition>> targets;
    std::vector<double> largest;
    gatherCandidates(range, targets, largest);
    gpuSpace->distanceBatch(positions, targets, largest, neighborDistances, neighborVectors, nullptr);
    if (verbose) {
        double meanNN = 0;
        for (auto& d : neighborDistances) meanNN += d.size();
        printf("mean number of neighbors = %f\n", meanNN / N);
    }
}

void gpuModel::moveParticles(std::vector<vector3>& disp) {
    if (!gpuSpace) {
        simpleModel::moveParticles(disp);
        return;
    }
    std::vector<std::vector<vector3>*> vecs;
    if (particleShiftsRequireForceTransport) vecs.push_back(&forces);
    if (particleShiftsRequireVelocityTransport) vecs.push_back(&velocities);
    gpuSpace->transportBatch(positions, disp, vecs);
}

void gpuModel::candidateLists(double range, std::vector<int32_t>& offsets, std::vector<int32_t>& indices, std::vector<double>& largest) {
    std::vector<std::vector<meshPosition>> targets;
    gatherCandidates(range, targets, largest);
    offsets.assign(N + 1, 0);
    indices.clear();
    for (int i = 0; i < N; ++i) {
        indices.insert(indices.end(), neighbors[i].begin(), neighbors[i].end());
        offsets[i + 1] = (int32_t)indices.size();
    }
}
