// Seeded synthetic surface meshes for the benchmark workloads. The reference reads OFF files
// (triangulatedMeshSpace::loadMeshFromFile, src/models/triangulatedMeshSpace.cpp:42-72); these generators produce the
// same kind of data (vertex coordinates, consistently oriented triangles with outward normals) without files.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace csg {

struct HostMesh {
    std::vector<double> xyz;  // 3 per vertex
    std::vector<int32_t> tri; // 3 per face, counter-clockwise seen from outside
    int nV() const { return (int)xyz.size() / 3; }
    int nF() const { return (int)tri.size() / 3; }
};

// icosahedron subdivided `level` times (each triangle into four), vertices projected to the sphere of `radius`.
// level 4: 2,562 vertices / 5,120 faces; level 5: 10,242 / 20,480; level 6: 40,962 / 81,920.
HostMesh icosphere(int level, double radius = 1.0);

// torus with major radius R and minor radius r, as an nu x nv parametric quad grid split into 2*nu*nv triangles;
// every vertex coordinate is then moved by an independent U(-jitter*l, +jitter*l) draw, l the mean edge length,
// from std::mt19937(seed).
HostMesh jitteredTorus(double R, double r, int nu, int nv, double jitter, uint32_t seed);

// icosphere whose vertex at unit direction u is moved to radius 1 + amplitude * sum_k exp(-g_k^2 / (2 width^2)),
// g_k the angle between u and bump centre c_k; the nBumps centres are uniform on the sphere from std::mt19937(seed).
HostMesh bumpySphere(int level, int nBumps, double amplitude, double width, uint32_t seed);

// OFF reader/writer (triangles only; polygons with more corners are rejected)
HostMesh readOFF(const std::string& path);
void writeOFF(const HostMesh& m, const std::string& path);

}  // namespace csg
