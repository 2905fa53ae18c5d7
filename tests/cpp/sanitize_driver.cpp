// Drives the oracle (oracle/*.cpp) and the CPU-only pieces of the host layer (mesh generators and OFF I/O, the noise
// source, the cell-list neighbour structure) through their main paths, for a build with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_sanitizers.py). Every check also compares results, so the run is a test of
// its own: the host cell list against the oracle's candidates, an OFF round trip, walks against their inputs.
// Prints "ok" and exits 0, or names the failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "meshGenerators.h"
#include "neighborStructures.h"
#include "noiseSource.h"
#include "oracle.h"

static int failures = 0;
#define CHECK(cond, ...)                 \
    do {                                 \
        if (!(cond)) {                   \
            printf("FAILED: " __VA_ARGS__); \
            printf("\n");                \
            ++failures;                  \
        }                                \
    } while (0)

struct Particles {
    std::vector<int> face;
    std::vector<double> bary, xyz;
};

static Particles place(const csg::HostMesh& m, orc_mesh* om, int n, noiseSource& noise) {
    Particles p;
    for (int i = 0; i < n; ++i) {
        p.face.push_back(noise.getInt(0, m.nF() - 1));
        double3 b = noise.getRandomBarycentricSet();
        p.bary.insert(p.bary.end(), {b.x, b.y, b.z});
    }
    p.xyz.resize(3 * n);
    orc_positions(om, n, p.face.data(), p.bary.data(), p.xyz.data());
    return p;
}

static void forces(const csg::HostMesh& m, orc_mesh* om, const Particles& p, double range, int kind, const double* params) {
    const int n = (int)p.face.size();
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int v = 0; v < m.nV(); ++v)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::fmin(lo[a], m.xyz[3 * v + a]);
            hi[a] = std::fmax(hi[a], m.xyz[3 * v + a]);
        }
    std::vector<int> off(n + 1), idx(64 * n);
    std::vector<double> md(n);
    long long tot = orc_cell_candidates(n, p.xyz.data(), lo, hi, range, range, off.data(), idx.data(), md.data(), (long long)idx.size());
    CHECK(tot >= 0, "candidate capacity");
    if (tot < 0) return;
    // the host layer's cell list gives the same lists in the same order
    std::vector<double> vlo(lo, lo + 3), vhi(hi, hi + 3);
    cellListNeighborStructure cl(vlo, vhi, range);
    cl.setInteractionRange(range);
    std::vector<meshPosition> parts(n);
    for (int i = 0; i < n; ++i) {
        parts[i].faceIndex = p.face[i];
        parts[i].x = point3(p.xyz[3 * i], p.xyz[3 * i + 1], p.xyz[3 * i + 2]);
    }
    cl.requireEuclideanPositions = true;
    cl.initialize(parts);
    for (int i = 0; i < n; ++i) {
        std::vector<int> ci;
        std::vector<meshPosition> cp;
        cl.constructCandidateNeighborList(parts[i], i, ci, cp);
        bool same = (int)ci.size() == off[i + 1] - off[i];
        for (size_t k = 0; same && k < ci.size(); ++k) same = ci[k] == idx[off[i] + k];
        CHECK(same, "host cell list differs from the oracle's for particle %d", i);
    }
    std::vector<double> F(3 * n), dist(tot), st(3 * tot), et(3 * tot);
    std::vector<int> fl(tot);
    double e = 0;
    int rc = orc_compute_forces(om, n, p.face.data(), p.bary.data(), 1, range, off.data(), idx.data(), md.data(), kind, params, F.data(),
                                dist.data(), st.data(), et.data(), fl.data(), &e, 2);
    CHECK(rc == 0, "orc_compute_forces: %s", orc_last_error());
    for (long long q = 0; q < tot; ++q) CHECK(dist[q] > 0 && std::isfinite(dist[q]), "pair %lld distance %g", q, dist[q]);
}

static void walks(orc_mesh* om, const Particles& p, noiseSource& noise) {
    const int n = (int)p.face.size();
    std::vector<int> face = p.face, status(n);
    std::vector<double> bary = p.bary, disp(3 * n), vecs(3 * n);
    for (int i = 0; i < 3 * n; ++i) {
        disp[i] = 0.3 * noise.getRealNormal();
        vecs[i] = disp[i];
    }
    int bad = orc_transport_batch(om, n, face.data(), bary.data(), disp.data(), 1, vecs.data(), status.data());
    CHECK(bad == 0, "%d walks failed on a closed mesh", bad);
    for (int i = 0; i < n; ++i) {
        double s = bary[3 * i] + bary[3 * i + 1] + bary[3 * i + 2];
        CHECK(std::fabs(s - 1) < 1e-9, "walk %d barycentrics sum to %g", i, s);
    }
}

int main(int argc, char** argv) {
    const std::string tmp = argc > 1 ? argv[1] : "/tmp";
    noiseSource noise(true);
    const double harm[3] = {4.0, 0.3, 0}, gauss[3] = {1.0, 0.2, 0}, trunc[3] = {1.0, 0.05, 0.15};
    // meshes of the workloads, through the OFF writer and reader
    std::vector<csg::HostMesh> meshes = {csg::icosphere(3), csg::bumpySphere(3, 20, 0.15, 0.3, 13377),
                                         csg::jitteredTorus(3.0, 1.0, 40, 20, 0.02, 7)};
    for (size_t k = 0; k < meshes.size(); ++k) {
        const csg::HostMesh& m = meshes[k];
        const std::string path = tmp + "/mesh" + std::to_string(k) + ".off";
        csg::writeOFF(m, path);
        csg::HostMesh r = csg::readOFF(path);
        CHECK(r.tri == m.tri && r.xyz.size() == m.xyz.size(), "OFF round trip of mesh %zu", k);
        orc_mesh* om = orc_mesh_create(m.nV(), m.xyz.data(), m.nF(), m.tri.data());
        CHECK(om != nullptr, "orc_mesh_create: %s", orc_last_error());
        if (!om) continue;
        Particles p = place(m, om, 120, noise);
        forces(m, om, p, k == 2 ? 0.6 : 0.3, k == 1 ? 2 : (k == 2 ? 1 : 0), k == 1 ? trunc : (k == 2 ? gauss : harm));
        walks(om, p, noise);
        // whole-mesh distances from one source to every particle
        std::vector<double> d(120), st(360), et(360);
        std::vector<int> fl(120);
        int rc = orc_distance(om, 0, 1.0, p.face[0], p.bary.data(), 120, p.face.data(), p.bary.data(), 1e20, d.data(), st.data(), et.data(),
                              fl.data());
        CHECK(rc == 0 && d[0] < 1e-9, "whole-mesh distance: %s", orc_last_error());
        orc_mesh_destroy(om);
    }
    // error paths: a non-triangular OFF file, and a flipped face
    {
        const std::string path = tmp + "/quad.off";
        FILE* f = fopen(path.c_str(), "w");
        fprintf(f, "OFF\n4 1 0\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n4 0 1 2 3\n");
        fclose(f);
        bool threw = false;
        try {
            csg::readOFF(path);
        } catch (const std::exception&) {
            threw = true;
        }
        CHECK(threw, "a quad OFF file was accepted");
        csg::HostMesh t = csg::icosphere(0);
        std::swap(t.tri[1], t.tri[2]);
        CHECK(orc_mesh_create(t.nV(), t.xyz.data(), t.nF(), t.tri.data()) == nullptr, "a flipped face was accepted");
    }
    // walks that fail: off a border on a closed-surface walk, and a NaN displacement
    {
        std::vector<double> X = {0, 0, 0, 1, 0, 0, 1, 1, 0, 0, 1, 0};
        std::vector<int> T = {0, 1, 2, 0, 2, 3};
        orc_mesh* om = orc_mesh_create(4, X.data(), 2, T.data());
        int face = 0;
        double bary[3] = {0.5, 0.25, 0.25}, disp[3] = {5, 0.1, 0};
        CHECK(orc_transport(om, &face, bary, disp, 0, nullptr) < 0, "walk off the border succeeded");
        double nanDisp[3] = {NAN, 0, 0};
        CHECK(orc_transport(om, &face, bary, nanDisp, 0, nullptr) == -4, "NaN walk");
        orc_mesh_destroy(om);
    }
    printf(failures ? "%d failures\n" : "ok\n", failures);
    return failures ? 1 : 0;
}
