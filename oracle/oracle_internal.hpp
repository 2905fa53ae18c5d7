// ORACLE — test infrastructure only.
//
// This directory is a plain FP64 CPU restatement of the reference's geodesic force path, used solely as the checker
// for the HIP implementation (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg). No product code may
// include, link or call anything here.
//
// Internal mesh representation shared by the oracle translation units.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator*(double s) const { return {x * s, y * s, z * s}; }
    V3 operator/(double s) const { return {x / s, y / s, z / s}; }
    V3 operator-() const { return {-x, -y, -z}; }
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(const V3& a, const V3& b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(const V3& a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(const V3& a) { return a / norm(a); }

struct Mesh {
    int nV = 0, nF = 0;
    std::vector<V3> X;             // vertex positions
    std::vector<int> F;            // 3 per face, vertex order defines barycentric order
    std::vector<int> adj;          // 3 per face: face across local edge k (v[k] -> v[k+1]); -1 on a border
    std::vector<int> adjEdge;      // 3 per face: local edge index of the same edge inside the neighbour
    std::vector<int> vfOff, vfList;  // vertex -> incident faces (CSR), entries are face*4 + local index
    std::vector<char> border;      // vertex touches a border edge of the whole mesh
    std::vector<double> angleSum;  // total face angle at each vertex
    // what a walk does at the mesh border: 0 closed surface (an error), 1 absorbing, 2 tangential
    // (absorbingOpenMeshSpace / tangentialOpenMeshSpace, src/models/*OpenMeshSpace.cpp)
    int boundary = 0;

    V3 P(int f, int k) const { return X[F[3 * f + ((k % 3) + 3) % 3]]; }
    int fv(int f, int k) const { return F[3 * f + ((k % 3) + 3) % 3]; }
    V3 normal(int f) const { return normalized(cross(P(f, 1) - P(f, 0), P(f, 2) - P(f, 0))); }
};

// Builds adjacency; throws std::runtime_error on a non-manifold edge or inconsistent orientation.
void buildMesh(Mesh& m, int nV, const double* X, int nF, const int* F);

struct OracleError : std::runtime_error {
    int code;
    OracleError(int c, const std::string& s) : std::runtime_error(s), code(c) {}
};

}  // namespace orc

struct orc_mesh {
    orc::Mesh m;
};
