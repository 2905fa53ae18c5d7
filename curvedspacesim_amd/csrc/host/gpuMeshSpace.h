// gpuMeshSpace: a triangulated mesh space whose geodesics and walks run on the GPU through libcsg.so. As the reference's
// triangulatedMeshSpace it treats the mesh as closed (a walk that meets a border fails); the open-mesh subclasses below
// give the reference's absorbing and tangential border conditions.
// It keeps the public interface of triangulatedMeshSpace (src/models/triangulatedMeshSpace.h:21-140) for the calls on
// the geodesic force path, and adds batched forms that the batched model uses. Failures throw std::runtime_error
// (the reference prints an error and throws std::exception, inc/debuggingHelp.h:19-31).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../../include/csg.h"
#include "baseSpace.h"

class gpuMeshSpace : public baseSpace {
public:
    explicit gpuMeshSpace(int device = 0);
    ~gpuMeshSpace() override;
    gpuMeshSpace(const gpuMeshSpace&) = delete;
    gpuMeshSpace& operator=(const gpuMeshSpace&) = delete;

    // triangulatedMeshSpace::loadMeshFromFile (src/models/triangulatedMeshSpace.cpp:42-72): OFF triangle meshes
    virtual void loadMeshFromFile(std::string filename, bool verbose = false);
    // the same from arrays (3 coordinates per vertex, 3 counter-clockwise corners per face)
    void loadMesh(const std::vector<double>& xyz, const std::vector<int32_t>& tri);

    void displaceParticle(meshPosition& pos, vector3& displacementVector) override;
    void transportParticleAndVectors(meshPosition& pos, vector3& displacementVector, std::vector<vector3>& transportVectors) override;
    void distance(meshPosition& p1, std::vector<meshPosition>& p2, std::vector<double>& distances, std::vector<vector3>& startPathTangent,
                  std::vector<vector3>& endPathTangent, double distanceThreshold = 1e20) override;
    void meshPositionToEuclideanLocation(std::vector<meshPosition>& p1, std::vector<double3>& result) override;
    void meshPositionToEuclideanLocation(std::vector<meshPosition>& p1, std::vector<meshPosition>& result) override;
    double getArea() override;
    // uniform face index, u ~ U(0,1), v ~ U(0,1-u) (src/models/triangulatedMeshSpace.cpp:108-116)
    void randomPosition(meshPosition& p, noiseSource& noise) override;
    // g1 o + g2 (n x o), g ~ N(0,1) (src/models/triangulatedMeshSpace.cpp:118-138)
    void randomVectorAtPosition(meshPosition& p, vector3& v, noiseSource& noise) override;

    // triangulatedMeshSpace::useSubmeshingRoutines (src/models/triangulatedMeshSpace.h:60-66)
    void useSubmeshingRoutines(bool useSubMesh, double maxDist = 1.0, bool danger = false);
    void setNewSubmeshCutoff(double newCutoff) { useSubmeshingRoutines(true, newCutoff); }

    // batched forms: every source with its own target list, one device launch
    void distanceBatch(const std::vector<meshPosition>& sources, const std::vector<std::vector<meshPosition>>& targets,
                       const std::vector<double>& distanceThreshold, std::vector<std::vector<double>>& distances,
                       std::vector<std::vector<vector3>>& startPathTangent, std::vector<std::vector<vector3>>* endPathTangent);
    void transportBatch(std::vector<meshPosition>& pos, std::vector<vector3>& displacement, std::vector<std::vector<vector3>*>& transportVectors);

    double3 minVertexPosition, maxVertexPosition;
    int numberOfFaces() const { return nF; }
    int numberOfVertices() const { return nV; }
    bool submeshingActivated() const { return submeshing; }
    double maximumDistance() const { return maxDistance; }
    csg_ctx* context() { return ctx; }
    const std::vector<double>& vertexCoordinates() const { return xyz; }
    const std::vector<int32_t>& faceCorners() const { return tri; }

protected:
    // CSG_BOUNDARY_*: what every walk of this space does at the mesh border
    void setBoundary(int mode) { check(csg_set_boundary(ctx, mode), "csg_set_boundary"); }

private:
    void check(csg_status s, const char* what) const;
    vector3 faceNormal(int f) const;
    csg_ctx* ctx = nullptr;
    std::vector<double> xyz;
    std::vector<int32_t> tri;
    int nV = 0, nF = 0;
    double area = 0;
    bool submeshing = false;
    double maxDistance = 0;
};

// absorbingOpenMeshSpace (src/models/absorbingOpenMeshSpace.h:8-30): a particle that meets the border stops there, and
// transported vectors lose their component pointing over it
class gpuAbsorbingOpenMeshSpace : public gpuMeshSpace {
public:
    explicit gpuAbsorbingOpenMeshSpace(int device = 0) : gpuMeshSpace(device) { setBoundary(CSG_BOUNDARY_ABSORBING); }
};

// tangentialOpenMeshSpace (src/models/tangentialOpenMeshSpace.h:8-33): a particle that meets the border slides along it
// for the rest of the step, and transported vectors lose their component pointing over it
class gpuTangentialOpenMeshSpace : public gpuMeshSpace {
public:
    explicit gpuTangentialOpenMeshSpace(int device = 0) : gpuMeshSpace(device) { setBoundary(CSG_BOUNDARY_TANGENTIAL); }
};
