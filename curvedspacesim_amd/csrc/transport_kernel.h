// Launch interface of the batched displacement walk (transport_kernel.hip).
#pragma once

#include "../../include/csg.h"
#include "csg_common.h"

namespace csg {

constexpr int kMaxTransportVectors = 2;  // the reference transports at most force and velocity (simpleModel.cpp:44-66)

enum TransportError : int {
    TRANSPORT_ERR_NAN = -4,          // reference: checkBaryNan throws (src/utility/meshUtilities.cpp:184-195)
    TRANSPORT_ERR_NO_FACE = -7,      // throughVertex found no face (src/models/triangulatedMeshSpace.cpp:404-405), or no
                                     // other face at a border vertex (src/models/openMeshSpace.cpp:56-57)
    TRANSPORT_ERR_TOO_MANY = -8,     // more face crossings than the configured maximum
    TRANSPORT_ERR_BORDER = -9,       // border edge or vertex on a closed surface (:522-523, :547-548)
    TRANSPORT_ERR_NO_CROSSING = -10, // neither 1 nor 2 crossings (:534-541)
    TRANSPORT_ERR_NO_HALFWAY = -6,   // no angular halfway at a vertex (:339-343)
    TRANSPORT_ERR_RING = -12,        // vertex with more than 32 incident faces
};

struct TransportParams {
    DeviceMesh mesh;
    int n;
    int32_t* face;   // in/out
    double* bary;    // in/out, 3 per particle
    const double* disp;  // 3 per particle
    double scale;    // displacement = scale * disp
    double* dispOut; // optional: the last segment vector, as the reference leaves its in/out displacement
    int nVec;        // transported vectors per particle
    double* vecs;    // in/out, 3 * nVec per particle
    int32_t* status; // optional: crossings (>= 0) or a TransportError; a failed particle keeps its input state
    int maxCrossings;
    int boundary;    // CSG_BOUNDARY_*: what a walk does at a border edge or vertex of an open mesh
};

hipError_t launchTransport(const TransportParams& P, hipStream_t stream);

}  // namespace csg
