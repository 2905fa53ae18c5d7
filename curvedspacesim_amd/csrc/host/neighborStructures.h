// Candidate-neighbour structures of the reference: all-to-all (src/utility/baseNeighborStructure.cpp:17-36) and the
// 27-cell Euclidean cell list (src/utility/cellListNeighborStructure.cpp:22-84 over
// src/utility/hyperRectangularCellList.cpp:9-159). Candidate order is the reference's: cells in x-major, then y,
// then z stencil order; within a cell, ascending particle index.
#pragma once

#include <vector>

#include "pointDataType.h"

class baseNeighborStructure {
public:
    virtual ~baseNeighborStructure() = default;
    virtual void setInteractionRange(double range) { maximumInteractionRange = range; }
    virtual void initialize(std::vector<meshPosition>& _particles);
    // fills the candidate indices and positions of particle particleIndex; returns the largest candidate distance
    // (1e20 for all-to-all, as the reference's VERYLARGEDOUBLE)
    virtual double constructCandidateNeighborList(meshPosition& p, int particleIndex, std::vector<int>& candidateNeighborIndices,
                                                  std::vector<meshPosition>& candidateParticles, int offset = 0);
    std::vector<int> indices;
    std::vector<meshPosition> particles;
    bool requireEuclideanPositions = false;
    double maximumInteractionRange = 0;
};

class cellListNeighborStructure : public baseNeighborStructure {
public:
    cellListNeighborStructure(double3 minPos, double3 maxPos, double gridSize);
    cellListNeighborStructure(std::vector<double>& minPos, std::vector<double>& maxPos, double gridSize);
    void setInteractionRange(double range) override;
    void initialize(std::vector<meshPosition>& _particles) override;
    double constructCandidateNeighborList(meshPosition& p, int particleIndex, std::vector<int>& candidateNeighborIndices,
                                          std::vector<meshPosition>& candidateParticles, int offset = 0) override;

private:
    void setGridSize(double size);
    int cellOf(const meshPosition& p) const;
    double lo[3], hi[3];
    double gridSize = -1;
    int n[3] = {1, 1, 1};
    double cell[3] = {1, 1, 1};
    double squaredInteractionRange = 0;
    std::vector<int> cellStart, cellMembers;  // CSR of particle indices per cell, ascending within a cell
};
