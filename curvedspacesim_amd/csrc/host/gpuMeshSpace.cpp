#include "gpuMeshSpace.h"

#include <stdexcept>

#include "meshGenerators.h"

gpuMeshSpace::gpuMeshSpace(int device) {
    positionsAreEuclidean = false;
    csg_status s = csg_ctx_create(&ctx, device);
    if (s != CSG_OK) throw std::runtime_error("gpuMeshSpace: no usable gfx950 device " + std::to_string(device));
}

gpuMeshSpace::~gpuMeshSpace() { csg_ctx_destroy(ctx); }

void gpuMeshSpace::check(csg_status s, const char* what) const {
    if (s != CSG_OK) throw std::runtime_error(std::string(what) + ": " + csg_last_error(ctx));
}

void gpuMeshSpace::loadMeshFromFile(std::string filename, bool verbose) {
    csg::HostMesh m = csg::readOFF(filename);
    loadMesh(m.xyz, m.tri);
    if (verbose) printf("loaded %s: %d vertices, %d faces, area %f\n", filename.c_str(), nV, nF, area);
}

void gpuMeshSpace::loadMesh(const std::vector<double>& x, const std::vector<int32_t>& t) {
    if (x.size() % 3 || t.size() % 3) throw std::runtime_error("loadMesh: arrays must hold 3 values per vertex and per face");
    check(csg_mesh_upload(ctx, (int32_t)(x.size() / 3), x.data(), (int32_t)(t.size() / 3), t.data()), "csg_mesh_upload");
    xyz = x;
    tri = t;
    double lo[3], hi[3];
    int32_t v, f;
    check(csg_mesh_info(ctx, &v, &f, &area, lo, hi), "csg_mesh_info");
    nV = v;
    nF = f;
    minVertexPosition = {lo[0], lo[1], lo[2]};
    maxVertexPosition = {hi[0], hi[1], hi[2]};
}

void gpuMeshSpace::useSubmeshingRoutines(bool useSubMesh, double maxDist, bool) {
    check(csg_set_submeshing(ctx, useSubMesh ? 1 : 0, maxDist), "csg_set_submeshing");
    submeshing = useSubMesh;
    maxDistance = maxDist;
}

double gpuMeshSpace::getArea() { return area; }

vector3 gpuMeshSpace::faceNormal(int f) const {
    const double* a = &xyz[3 * tri[3 * f]];
    const double* b = &xyz[3 * tri[3 * f + 1]];
    const double* c = &xyz[3 * tri[3 * f + 2]];
    vector3 n = cross_product(vector3(b[0] - a[0], b[1] - a[1], b[2] - a[2]), vector3(c[0] - a[0], c[1] - a[1], c[2] - a[2]));
    return n / std::sqrt(n.squared_length());
}

void gpuMeshSpace::randomPosition(meshPosition& p, noiseSource& noise) {
    double u = noise.getRealUniform();
    double v = noise.getRealUniform(0, 1 - u);
    p.x = point3(u, v, 1 - u - v);
    p.faceIndex = noise.getInt(0, nF - 1);
}

void gpuMeshSpace::randomVectorAtPosition(meshPosition& p, vector3& v, noiseSource& noise) {
    if (p.faceIndex < 0 || p.faceIndex >= nF) throw std::runtime_error("randomVectorAtPosition: face index out of range");
    vector3 n = faceNormal(p.faceIndex);
    vector3 o = (n[0] == 0 && n[1] == 0) ? vector3(n[1] - n[2], n[2] - n[0], n[0] - n[1]) : vector3(n[1], -n[0], 0);
    o /= std::sqrt(o.squared_length());
    vector3 t2 = cross_product(n, o);
    // The reference writes both draws in one expression, getRealNormal() * o + getRealNormal() * t2, whose operand order
    // C++ leaves unspecified; g++ evaluates the right operand first, so the t2 weight is drawn first.
    double g2 = noise.getRealNormal();
    double g1 = noise.getRealNormal();
    v = g1 * o + g2 * t2;
}

void gpuMeshSpace::meshPositionToEuclideanLocation(std::vector<meshPosition>& p1, std::vector<double3>& result) {
    const int n = (int)p1.size();
    std::vector<int32_t> face(n);
    std::vector<double> bary(3 * (size_t)n), out(3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        face[i] = p1[i].faceIndex;
        for (int a = 0; a < 3; ++a) bary[3 * i + a] = p1[i].x[a];
    }
    check(csg_positions(ctx, n, face.data(), bary.data(), out.data()), "csg_positions");
    result.resize(n);
    for (int i = 0; i < n; ++i) result[i] = {out[3 * i], out[3 * i + 1], out[3 * i + 2]};
}

void gpuMeshSpace::meshPositionToEuclideanLocation(std::vector<meshPosition>& p1, std::vector<meshPosition>& result) {
    std::vector<double3> r;
    meshPositionToEuclideanLocation(p1, r);
    result.resize(p1.size());
    for (size_t i = 0; i < p1.size(); ++i) {
        result[i].x = point3(r[i].x, r[i].y, r[i].z);
        result[i].faceIndex = p1[i].faceIndex;
    }
}

void gpuMeshSpace::distance(meshPosition& p1, std::vector<meshPosition>& p2, std::vector<double>& distances,
                            std::vector<vector3>& startPathTangent, std::vector<vector3>& endPathTangent, double distanceThreshold) {
    std::vector<std::vector<double>> d;
    std::vector<std::vector<vector3>> st, et;
    distanceBatch({p1}, {p2}, {distanceThreshold}, d, st, &et);
    distances = std::move(d[0]);
    startPathTangent = std::move(st[0]);
    endPathTangent = std::move(et[0]);
}

void gpuMeshSpace::distanceBatch(const std::vector<meshPosition>& sources, const std::vector<std::vector<meshPosition>>& targets,
                                 const std::vector<double>& distanceThreshold, std::vector<std::vector<double>>& distances,
                                 std::vector<std::vector<vector3>>& startPathTangent, std::vector<std::vector<vector3>>* endPathTangent) {
    const int nS = (int)sources.size();
    if ((int)targets.size() != nS || (int)distanceThreshold.size() != nS) throw std::runtime_error("distanceBatch: size mismatch");
    std::vector<int32_t> sf(nS), off(nS + 1, 0), tf;
    std::vector<double> sb(3 * (size_t)nS), tb;
    for (int s = 0; s < nS; ++s) {
        sf[s] = sources[s].faceIndex;
        for (int a = 0; a < 3; ++a) sb[3 * s + a] = sources[s].x[a];
        off[s + 1] = off[s] + (int32_t)targets[s].size();
        for (const meshPosition& t : targets[s]) {
            tf.push_back(t.faceIndex);
            for (int a = 0; a < 3; ++a) tb.push_back(t.x[a]);
        }
    }
    const size_t nT = tf.size();
    std::vector<double> d(nT), st(3 * nT), et(3 * nT);
    check(csg_distance_batch(ctx, nS, sf.data(), sb.data(), off.data(), tf.data(), tb.data(), distanceThreshold.data(), d.data(), st.data(),
                             endPathTangent ? et.data() : nullptr, nullptr),
          "csg_distance_batch");
    distances.assign(nS, {});
    startPathTangent.assign(nS, {});
    if (endPathTangent) endPathTangent->assign(nS, {});
    for (int s = 0; s < nS; ++s) {
        for (int q = off[s]; q < off[s + 1]; ++q) {
            distances[s].push_back(d[q]);
            startPathTangent[s].emplace_back(st[3 * q], st[3 * q + 1], st[3 * q + 2]);
            if (endPathTangent) (*endPathTangent)[s].emplace_back(et[3 * q], et[3 * q + 1], et[3 * q + 2]);
        }
    }
}

void gpuMeshSpace::transportBatch(std::vector<meshPosition>& pos, std::vector<vector3>& displacement,
                                  std::vector<std::vector<vector3>*>& transportVectors) {
    const int n = (int)pos.size();
    const int nVec = (int)transportVectors.size();
    if ((int)displacement.size() != n) throw std::runtime_error("transportBatch: size mismatch");
    if (nVec > 2) throw std::runtime_error("transportBatch: at most 2 transported vectors per particle");
    std::vector<int32_t> face(n), status(n);
    std::vector<double> bary(3 * (size_t)n), disp(3 * (size_t)n), vecs(3 * (size_t)n * nVec);
    for (int i = 0; i < n; ++i) {
        face[i] = pos[i].faceIndex;
        for (int a = 0; a < 3; ++a) {
            bary[3 * i + a] = pos[i].x[a];
            disp[3 * i + a] = displacement[i][a];
            for (int k = 0; k < nVec; ++k) vecs[3 * ((size_t)i * nVec + k) + a] = (*transportVectors[k])[i][a];
        }
    }
    check(csg_transport_batch(ctx, n, face.data(), bary.data(), disp.data(), nVec, vecs.data(), status.data()), "csg_transport_batch");
    for (int i = 0; i < n; ++i) {
        pos[i].faceIndex = face[i];
        pos[i].x = point3(bary[3 * i], bary[3 * i + 1], bary[3 * i + 2]);
        displacement[i] = vector3(disp[3 * i], disp[3 * i + 1], disp[3 * i + 2]);
        for (int k = 0; k < nVec; ++k) {
            const double* v = &vecs[3 * ((size_t)i * nVec + k)];
            (*transportVectors[k])[i] = vector3(v[0], v[1], v[2]);
        }
    }
}

void gpuMeshSpace::transportParticleAndVectors(meshPosition& pos, vector3& displacementVector, std::vector<vector3>& transportVectors) {
    const int nVec = (int)transportVectors.size();
    std::vector<meshPosition> p{pos};
    std::vector<vector3> d{displacementVector};
    std::vector<std::vector<vector3>> vs(nVec);
    std::vector<std::vector<vector3>*> ptrs;
    for (int k = 0; k < nVec; ++k) {
        vs[k] = {transportVectors[k]};
        ptrs.push_back(&vs[k]);
    }
    transportBatch(p, d, ptrs);
    pos = p[0];
    displacementVector = d[0];
    for (int k = 0; k < nVec; ++k) transportVectors[k] = vs[k][0];
}

void gpuMeshSpace::displaceParticle(meshPosition& pos, vector3& displacementVector) {
    std::vector<vector3> none;
    transportParticleAndVectors(pos, displacementVector, none);
}
