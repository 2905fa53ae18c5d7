#include "neighborStructures.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

void baseNeighborStructure::initialize(std::vector<meshPosition>& _particles) {
    particles = _particles;
    indices.resize(particles.size());
    for (size_t i = 0; i < particles.size(); ++i) indices[i] = (int)i;
}

double baseNeighborStructure::constructCandidateNeighborList(meshPosition&, int particleIndex, std::vector<int>& candidateNeighborIndices,
                                                             std::vector<meshPosition>& candidateParticles, int offset) {
    const int N = (int)particles.size();
    candidateNeighborIndices.clear();
    candidateParticles.clear();
    for (int i = 0; i < N; ++i)
        if (particleIndex + offset != i) {
            candidateParticles.push_back(particles[i]);
            candidateNeighborIndices.push_back(indices[i]);
        }
    return 1e20;
}

cellListNeighborStructure::cellListNeighborStructure(double3 minPos, double3 maxPos, double size) {
    lo[0] = minPos.x, lo[1] = minPos.y, lo[2] = minPos.z;
    hi[0] = maxPos.x, hi[1] = maxPos.y, hi[2] = maxPos.z;
    setGridSize(size);
}

cellListNeighborStructure::cellListNeighborStructure(std::vector<double>& minPos, std::vector<double>& maxPos, double size) {
    for (int a = 0; a < 3; ++a) lo[a] = minPos[a], hi[a] = maxPos[a];
    setGridSize(size);
}

// as many whole cells of at least `size` as fit the box in each dimension (at least one)
void cellListNeighborStructure::setGridSize(double size) {
    if (size == gridSize) return;
    gridSize = size;
    for (int a = 0; a < 3; ++a) {
        double ext = hi[a] - lo[a];
        n[a] = std::max(1, (int)std::floor(ext / size));
        cell[a] = ext / n[a];
    }
    if ((long)n[0] * n[1] * n[2] == 1)
        printf("You have chosen a grid size that results in a cell list with only one cell. This will be functional, but slower than not "
               "using a grid at all.\n");
}

void cellListNeighborStructure::setInteractionRange(double range) {
    requireEuclideanPositions = true;
    maximumInteractionRange = range;
    squaredInteractionRange = range * range;
    setGridSize(range);
}

int cellListNeighborStructure::cellOf(const meshPosition& p) const {
    int c[3];
    for (int a = 0; a < 3; ++a) c[a] = std::max(0, std::min(n[a] - 1, (int)std::floor((p.x[a] - lo[a]) / cell[a])));
    return c[0] + n[0] * (c[1] + n[1] * c[2]);
}

void cellListNeighborStructure::initialize(std::vector<meshPosition>& _particles) {
    baseNeighborStructure::initialize(_particles);
    const int nc = n[0] * n[1] * n[2];
    cellStart.assign(nc + 1, 0);
    std::vector<int> key(particles.size());
    for (size_t i = 0; i < particles.size(); ++i) {
        key[i] = cellOf(particles[i]);
        ++cellStart[key[i] + 1];
    }
    for (int c = 0; c < nc; ++c) cellStart[c + 1] += cellStart[c];
    cellMembers.assign(particles.size(), 0);
    std::vector<int> fill(cellStart.begin(), cellStart.end() - 1);
    for (size_t i = 0; i < particles.size(); ++i) cellMembers[fill[key[i]]++] = (int)i;
}

double cellListNeighborStructure::constructCandidateNeighborList(meshPosition& p, int particleIndex, std::vector<int>& candidateNeighborIndices,
                                                                 std::vector<meshPosition>& candidateParticles, int offset) {
    candidateNeighborIndices.clear();
    candidateParticles.clear();
    const int c0 = cellOf(p);
    const int cx = c0 % n[0], cy = (c0 / n[0]) % n[1], cz = c0 / (n[0] * n[1]);
    double maxD2 = 0;
    for (int x = std::max(0, cx - 1); x <= std::min(n[0] - 1, cx + 1); ++x)
        for (int y = std::max(0, cy - 1); y <= std::min(n[1] - 1, cy + 1); ++y)
            for (int z = std::max(0, cz - 1); z <= std::min(n[2] - 1, cz + 1); ++z) {
                const int c = x + n[0] * (y + n[1] * z);
                for (int q = cellStart[c]; q < cellStart[c + 1]; ++q) {
                    const int j = cellMembers[q];
                    if (particleIndex + offset == j) continue;
                    double d2 = squared_distance(p.x, particles[j].x);
                    if (d2 < squaredInteractionRange) {
                        candidateParticles.push_back(particles[j]);
                        candidateNeighborIndices.push_back(indices[j]);
                        if (d2 > maxD2) maxD2 = d2;
                    }
                }
            }
    return std::sqrt(maxD2);
}
