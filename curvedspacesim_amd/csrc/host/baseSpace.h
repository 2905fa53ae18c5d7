// The reference's space interface (src/models/baseSpace.h:24-53) with the same virtual functions, argument meaning and
// in/out conventions; CGAL value types are replaced by those of pointDataType.h.
#pragma once

#include <vector>

#include "noiseSource.h"
#include "pointDataType.h"

class baseSpace {
public:
    virtual ~baseSpace() = default;
    // pos and displacementVector are in/out
    virtual void displaceParticle(meshPosition& pos, vector3& displacementVector) = 0;
    // pos, displacementVector and every transport vector are in/out
    virtual void transportParticleAndVectors(meshPosition& pos, vector3& displacementVector, std::vector<vector3>& transportVectors) = 0;
    // distances and unit start/end path tangents from p1 to every p2; the outputs are resized by the callee
    virtual void distance(meshPosition& p1, std::vector<meshPosition>& p2, std::vector<double>& distances,
                          std::vector<vector3>& startPathTangent, std::vector<vector3>& endPathTangent, double distanceThreshold) = 0;
    virtual void meshPositionToEuclideanLocation(std::vector<meshPosition>& p1, std::vector<double3>& result) = 0;
    virtual void meshPositionToEuclideanLocation(std::vector<meshPosition>& p1, std::vector<meshPosition>& result) = 0;
    virtual double getArea() = 0;
    bool positionsAreEuclidean = true;
    virtual void randomPosition(meshPosition& p, noiseSource& noise) = 0;
    virtual void randomVectorAtPosition(meshPosition& p, vector3& v, noiseSource& noise) = 0;
};
