// Scenarios of the reference-API host layer, written the way the reference's own drivers use its classes
// (curvedSpaceSimulation.cpp, curvedSpaceNVTSim.cpp, meshTesting.cpp). Each prints one JSON object that the pytest
// parity tests compare with the oracle.
//
//   refapi_scenarios cells <points.txt> <range> <lox> <loy> <loz> <hix> <hiy> <hiz>      (no GPU)
//   refapi_scenarios potentials                                                         (no GPU)
//   refapi_scenarios noise <seed>                                                       (no GPU)
//   refapi_scenarios distance <mesh.off> <submesh> <maxDist> <particles.txt> <source> <threshold>
//   refapi_scenarios simulate <mesh.off> <particles.txt> <kind> <p0> <p1> <p2> <cellList> <submesh> <maxDist>
//                             <nve|nvt> <dt> <T> <steps> <mode: host|batched|fused> [closed|absorbing|tangential]
//
// points.txt: N, then N lines "x y z". particles.txt: N, then N lines "face b0 b1 b2 v0 v1 v2".
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "forces.h"
#include "gpuMeshSpace.h"
#include "neighborStructures.h"
#include "noiseSource.h"
#include "simpleModel.h"
#include "simulation.h"

namespace {

void printArray(const char* name, const std::vector<double>& v, bool last = false) {
    printf("\"%s\": [", name);
    for (size_t i = 0; i < v.size(); ++i) printf("%s%.17g", i ? ", " : "", v[i]);
    printf("]%s\n", last ? "" : ",");
}

void printInts(const char* name, const std::vector<int>& v, bool last = false) {
    printf("\"%s\": [", name);
    for (size_t i = 0; i < v.size(); ++i) printf("%s%d", i ? ", " : "", v[i]);
    printf("]%s\n", last ? "" : ",");
}

struct Particles {
    std::vector<meshPosition> pos;
    std::vector<vector3> vel;
};

Particles readParticles(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    int n;
    in >> n;
    Particles p;
    p.pos.resize(n);
    p.vel.resize(n);
    for (int i = 0; i < n; ++i) {
        double b[3], v[3];
        in >> p.pos[i].faceIndex >> b[0] >> b[1] >> b[2] >> v[0] >> v[1] >> v[2];
        p.pos[i].x = point3(b[0], b[1], b[2]);
        p.vel[i] = vector3(v[0], v[1], v[2]);
    }
    if (!in) throw std::runtime_error("bad particle file " + path);
    return p;
}

int cells(char** a) {
    std::ifstream in(a[2]);
    int n;
    in >> n;
    std::vector<meshPosition> pts(n);
    for (int i = 0; i < n; ++i) {
        double x, y, z;
        in >> x >> y >> z;
        pts[i].x = point3(x, y, z);
    }
    double range = atof(a[3]);
    double3 lo{atof(a[4]), atof(a[5]), atof(a[6])}, hi{atof(a[7]), atof(a[8]), atof(a[9])};
    cellListNeighborStructure cl(lo, hi, range);
    cl.setInteractionRange(range);
    cl.initialize(pts);
    std::vector<int> off{0}, idx;
    std::vector<double> largest;
    for (int i = 0; i < n; ++i) {
        std::vector<int> nb;
        std::vector<meshPosition> cp;
        largest.push_back(cl.constructCandidateNeighborList(pts[i], i, nb, cp));
        idx.insert(idx.end(), nb.begin(), nb.end());
        off.push_back((int)idx.size());
    }
    printf("{\n");
    printInts("offsets", off);
    printInts("indices", idx);
    printArray("largest", largest, true);
    printf("}\n");
    return 0;
}

int potentials() {
    std::vector<std::shared_ptr<force>> fs{std::make_shared<harmonicRepulsion>(11.1, 0.3), std::make_shared<gaussianRepulsion>(1.3, 0.2),
                                           std::make_shared<truncatedGaussianRepulsion>(1.0, 0.05, 0.15)};
    const char* names[] = {"harmonic", "gaussian", "truncated_gaussian"};
    printf("{\n");
    for (size_t k = 0; k < fs.size(); ++k) {
        std::vector<double> d, e, f;
        for (int i = 0; i <= 40; ++i) {
            double x = 0.01 * i;
            d.push_back(x);
            e.push_back(fs[k]->pairwiseEnergy(vector3(1, 0, 0), x));
            f.push_back(fs[k]->pairwiseForce(vector3(1, 0, 0), x)[0]);
        }
        printf("\"%s\": {\n", names[k]);
        printArray("d", d);
        printArray("energy", e);
        printArray("force", f, true);
        printf("}%s\n", k + 1 < fs.size() ? "," : "");
    }
    printf("}\n");
    return 0;
}

int noise(char** a) {
    noiseSource n1(true), n2(true);
    n1.setReproducibleSeed(atoi(a[2]));
    n2.setReproducibleSeed(atoi(a[2]));
    std::vector<double> u1, u2, bc;
    std::vector<int> ints;
    for (int i = 0; i < 64; ++i) {
        u1.push_back(n1.getRealUniform());
        u2.push_back(n2.getRealUniform());
        ints.push_back(n1.getInt(0, 9));
        n2.getInt(0, 9);
        double3 b = n1.getRandomBarycentricSet();
        n2.getRandomBarycentricSet();
        bc.push_back(b.x);
        bc.push_back(b.y);
        bc.push_back(b.z);
    }
    printf("{\n");
    printArray("uniform_a", u1);
    printArray("uniform_b", u2);
    printInts("ints", ints);
    printArray("barycentric", bc, true);
    printf("}\n");
    return 0;
}

int distance(char** a) {
    auto space = std::make_shared<gpuMeshSpace>(0);
    space->loadMeshFromFile(a[2]);
    if (atoi(a[3])) space->useSubmeshingRoutines(true, atof(a[4]));
    Particles p = readParticles(a[5]);
    int src = atoi(a[6]);
    double thr = atof(a[7]);
    std::vector<meshPosition> targets;
    for (size_t j = 0; j < p.pos.size(); ++j)
        if ((int)j != src) targets.push_back(p.pos[j]);
    std::vector<double> d;
    std::vector<vector3> st, et;
    space->distance(p.pos[src], targets, d, st, et, thr);
    std::vector<double> s3, e3;
    for (size_t j = 0; j < d.size(); ++j)
        for (int c = 0; c < 3; ++c) {
            s3.push_back(st[j][c]);
            e3.push_back(et[j][c]);
        }
    printf("{\n");
    printArray("distance", d);
    printArray("start", s3);
    printArray("end", e3, true);
    printf("}\n");
    return 0;
}

int simulate(int argc, char** a) {
    // optional last argument: the space class (closed triangulated mesh, or an open mesh with absorbing or tangential
    // border conditions)
    const std::string boundary = argc > 16 ? a[16] : "closed";
    std::shared_ptr<gpuMeshSpace> space;
    if (boundary == "absorbing")
        space = std::make_shared<gpuAbsorbingOpenMeshSpace>(0);
    else if (boundary == "tangential")
        space = std::make_shared<gpuTangentialOpenMeshSpace>(0);
    else
        space = std::make_shared<gpuMeshSpace>(0);
    space->loadMeshFromFile(a[2]);
    Particles p = readParticles(a[3]);
    const int kind = atoi(a[4]);
    const double p0 = atof(a[5]), p1 = atof(a[6]), p2 = atof(a[7]);
    const bool cellList = atoi(a[8]) != 0, submesh = atoi(a[9]) != 0;
    const double maxDist = atof(a[10]);
    const std::string integrator = a[11];
    const double dt = atof(a[12]), T = atof(a[13]);
    const int steps = atoi(a[14]);
    const std::string mode = a[15];
    if (submesh) space->useSubmeshingRoutines(true, maxDist);

    const int N = (int)p.pos.size();
    std::shared_ptr<simpleModel> model = mode == "host" ? std::make_shared<simpleModel>(N) : std::make_shared<gpuModel>(N);
    model->setSpace(space);
    model->setParticlePositions(p.pos);
    model->velocities = p.vel;
    std::shared_ptr<force> pot;
    if (kind == 0)
        pot = std::make_shared<harmonicRepulsion>(p0, p1);
    else if (kind == 1)
        pot = std::make_shared<gaussianRepulsion>(p0, p1);
    else
        pot = std::make_shared<truncatedGaussianRepulsion>(p0, p1, p2);
    if (cellList)
        model->setNeighborStructure(
            std::make_shared<cellListNeighborStructure>(space->minVertexPosition, space->maxVertexPosition, pot->maximumInteractionRange));
    std::shared_ptr<force> frc = mode == "fused" ? std::make_shared<gpuPairForce>(pot) : pot;

    auto sim = std::make_shared<simulation>();
    sim->setConfiguration(model);
    sim->addForce(frc, model);
    std::shared_ptr<updater> upd;
    if (integrator == "nve")
        upd = std::make_shared<velocityVerletNVE>(dt);
    else
        upd = std::make_shared<noseHooverNVT>(dt, T, 1.0, 2);
    sim->addUpdater(upd, model);
    sim->setIntegrationTimestep(dt);
    sim->computeForces();
    for (int s = 0; s < steps; ++s) sim->performTimestep();

    std::vector<int> face;
    std::vector<double> bary, vel, frcOut;
    for (int i = 0; i < N; ++i) {
        face.push_back(model->positions[i].faceIndex);
        for (int c = 0; c < 3; ++c) {
            bary.push_back(model->positions[i].x[c]);
            vel.push_back(model->velocities[i][c]);
            frcOut.push_back(model->forces[i][c]);
        }
    }
    printf("{\n");
    printInts("face", face);
    printArray("bary", bary);
    printArray("vel", vel);
    printArray("forces", frcOut, true);
    printf("}\n");
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: refapi_scenarios <cells|potentials|noise|distance|simulate> ...\n");
        return 2;
    }
    const std::string s = argv[1];
    const int need = s == "cells" ? 10 : s == "noise" ? 3 : s == "distance" ? 8 : s == "simulate" ? 16 : 2;
    if (argc < need) {
        fprintf(stderr, "%s: expected %d arguments\n", s.c_str(), need - 1);
        return 2;
    }
    try {
        if (s == "cells") return cells(argv);
        if (s == "potentials") return potentials();
        if (s == "noise") return noise(argv);
        if (s == "distance") return distance(argv);
        if (s == "simulate") return simulate(argc, argv);
    } catch (std::exception& e) {
        fprintf(stderr, "ERRORERROR %s\n", e.what());
        return 1;
    }
    fprintf(stderr, "unknown scenario %s\n", s.c_str());
    return 2;
}
