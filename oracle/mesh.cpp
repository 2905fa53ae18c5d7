// ORACLE — test infrastructure only (see oracle_internal.hpp).
//
// Face adjacency for a consistently oriented triangle mesh. The reference keeps this information in a CGAL
// Surface_mesh halfedge structure (src/models/triangulatedMeshSpace.cpp:42-72); here it is two flat arrays.
#include <algorithm>
#include <unordered_map>

#include "oracle_internal.hpp"

namespace orc {

void buildMesh(Mesh& m, int nV, const double* X, int nF, const int* F) {
    m.nV = nV;
    m.nF = nF;
    m.X.resize(nV);
    for (int i = 0; i < nV; ++i) m.X[i] = V3(X[3 * i], X[3 * i + 1], X[3 * i + 2]);
    m.F.assign(F, F + 3 * nF);
    m.adj.assign(3 * nF, -1);
    m.adjEdge.assign(3 * nF, -1);
    for (int i = 0; i < 3 * nF; ++i)
        if (m.F[i] < 0 || m.F[i] >= nV) throw OracleError(-2, "face references a vertex out of range");
    // directed edge (a,b) -> face*4+k
    std::unordered_map<uint64_t, int> directed;
    directed.reserve(6 * nF);
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) {
            uint64_t a = (uint32_t)m.fv(f, k), b = (uint32_t)m.fv(f, k + 1);
            if (a == b) throw OracleError(-2, "degenerate face");
            uint64_t key = (a << 32) | b;
            if (directed.count(key)) throw OracleError(-2, "directed edge used twice: non-manifold or inconsistently oriented mesh");
            directed[key] = f * 4 + k;
        }
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) {
            uint64_t a = (uint32_t)m.fv(f, k), b = (uint32_t)m.fv(f, k + 1);
            auto it = directed.find((b << 32) | a);
            if (it != directed.end()) {
                m.adj[3 * f + k] = it->second >> 2;
                m.adjEdge[3 * f + k] = it->second & 3;
            }
        }
    m.vfOff.assign(nV + 1, 0);
    for (int i = 0; i < 3 * nF; ++i) m.vfOff[m.F[i] + 1]++;
    for (int v = 0; v < nV; ++v) m.vfOff[v + 1] += m.vfOff[v];
    m.vfList.assign(3 * nF, 0);
    std::vector<int> fill(m.vfOff.begin(), m.vfOff.end() - 1);
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) m.vfList[fill[m.fv(f, k)]++] = f * 4 + k;
    m.border.assign(nV, 0);
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k)
            if (m.adj[3 * f + k] < 0) {
                m.border[m.fv(f, k)] = 1;
                m.border[m.fv(f, k + 1)] = 1;
            }
    m.angleSum.assign(nV, 0.0);
    for (int f = 0; f < nF; ++f)
        for (int k = 0; k < 3; ++k) {
            V3 a = m.P(f, k + 1) - m.P(f, k), b = m.P(f, k + 2) - m.P(f, k);
            double c = dot(a, b) / (norm(a) * norm(b));
            c = std::max(-1.0, std::min(1.0, c));
            m.angleSum[m.fv(f, k)] += std::acos(c);
        }
}

}  // namespace orc
