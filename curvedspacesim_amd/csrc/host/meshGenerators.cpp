#include "meshGenerators.h"

#include <cmath>
#include <fstream>
#include <map>
#include <random>
#include <sstream>
#include <stdexcept>

namespace csg {

namespace {

void normalizeTo(double* p, double radius) {
    double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    for (int a = 0; a < 3; ++a) p[a] *= radius / l;
}

}  // namespace

HostMesh icosphere(int level, double radius) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    HostMesh m;
    double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                          {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : base) {
        normalizeTo(p, 1.0);
        m.xyz.insert(m.xyz.end(), p, p + 3);
    }
    int faces[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                        {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                        {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (auto& f : faces) m.tri.insert(m.tri.end(), f, f + 3);
    for (int l = 0; l < level; ++l) {
        std::map<std::pair<int, int>, int> mid;
        auto midpoint = [&](int a, int b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double p[3];
            for (int c = 0; c < 3; ++c) p[c] = 0.5 * (m.xyz[3 * a + c] + m.xyz[3 * b + c]);
            normalizeTo(p, 1.0);
            int id = m.nV();
            m.xyz.insert(m.xyz.end(), p, p + 3);
            mid[key] = id;
            return id;
        };
        std::vector<int32_t> nt;
        nt.reserve(4 * m.tri.size());
        for (int f = 0; f < m.nF(); ++f) {
            int a = m.tri[3 * f], b = m.tri[3 * f + 1], c = m.tri[3 * f + 2];
            int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            int add[4][3] = {{a, ab, ca}, {b, bc, ab}, {c, ca, bc}, {ab, bc, ca}};
            for (auto& q : add) nt.insert(nt.end(), q, q + 3);
        }
        m.tri.swap(nt);
    }
    for (size_t i = 0; i < m.xyz.size(); ++i) m.xyz[i] *= radius;
    return m;
}

HostMesh jitteredTorus(double R, double r, int nu, int nv, double jitter, uint32_t seed) {
    HostMesh m;
    m.xyz.resize(3 * (size_t)nu * nv);
    for (int i = 0; i < nu; ++i)
        for (int j = 0; j < nv; ++j) {
            double u = 2 * M_PI * i / nu, v = 2 * M_PI * j / nv;
            double* p = &m.xyz[3 * ((size_t)i * nv + j)];
            p[0] = (R + r * std::cos(v)) * std::cos(u);
            p[1] = (R + r * std::cos(v)) * std::sin(u);
            p[2] = r * std::sin(v);
        }
    auto id = [&](int i, int j) { return ((i + nu) % nu) * nv + ((j + nv) % nv); };
    for (int i = 0; i < nu; ++i)
        for (int j = 0; j < nv; ++j) {
            // quad (i,j),(i+1,j),(i+1,j+1),(i,j+1); d/du x d/dv points outward on this parametrisation
            int a = id(i, j), b = id(i + 1, j), c = id(i + 1, j + 1), d = id(i, j + 1);
            int t1[3] = {a, b, c}, t2[3] = {a, c, d};
            m.tri.insert(m.tri.end(), t1, t1 + 3);
            m.tri.insert(m.tri.end(), t2, t2 + 3);
        }
    // mean edge length over the three edges of every triangle (interior edges counted twice)
    double sum = 0;
    for (int f = 0; f < m.nF(); ++f)
        for (int k = 0; k < 3; ++k) {
            const double* p = &m.xyz[3 * m.tri[3 * f + k]];
            const double* q = &m.xyz[3 * m.tri[3 * f + (k + 1) % 3]];
            sum += std::sqrt((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]));
        }
    double ell = sum / (3.0 * m.nF());
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> U(-jitter * ell, jitter * ell);
    for (double& x : m.xyz) x += U(gen);
    return m;
}

HostMesh bumpySphere(int level, int nBumps, double amplitude, double width, uint32_t seed) {
    HostMesh m = icosphere(level, 1.0);
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<double> c(3 * nBumps);
    for (int k = 0; k < nBumps; ++k) {
        double z = 2.0 * U(gen) - 1.0;
        double phi = 2.0 * M_PI * U(gen);
        double s = std::sqrt(std::max(0.0, 1.0 - z * z));
        c[3 * k] = s * std::cos(phi);
        c[3 * k + 1] = s * std::sin(phi);
        c[3 * k + 2] = z;
    }
    for (int v = 0; v < m.nV(); ++v) {
        double* p = &m.xyz[3 * v];
        double rad = 1.0;
        for (int k = 0; k < nBumps; ++k) {
            double d = p[0] * c[3 * k] + p[1] * c[3 * k + 1] + p[2] * c[3 * k + 2];
            d = std::max(-1.0, std::min(1.0, d));
            double g = std::acos(d);
            rad += amplitude * std::exp(-g * g / (2 * width * width));
        }
        for (int a = 0; a < 3; ++a) p[a] *= rad;
    }
    return m;
}

HostMesh readOFF(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::string tok;
    in >> tok;
    if (tok != "OFF") throw std::runtime_error("not an OFF file: " + path);
    long nv, nf, ne;
    in >> nv >> nf >> ne;
    HostMesh m;
    m.xyz.resize(3 * nv);
    for (long i = 0; i < 3 * nv; ++i) in >> m.xyz[i];
    m.tri.reserve(3 * nf);
    for (long f = 0; f < nf; ++f) {
        int k;
        in >> k;
        if (k != 3) throw std::runtime_error("non-triangular face in " + path);
        int a, b, c;
        in >> a >> b >> c;
        m.tri.push_back(a);
        m.tri.push_back(b);
        m.tri.push_back(c);
        std::string rest;
        std::getline(in, rest);
    }
    if (!in) throw std::runtime_error("truncated OFF file: " + path);
    return m;
}

void writeOFF(const HostMesh& m, const std::string& path) {
    std::ofstream out(path);
    out.precision(17);
    out << "OFF\n" << m.nV() << " " << m.nF() << " 0\n";
    for (int v = 0; v < m.nV(); ++v) out << m.xyz[3 * v] << " " << m.xyz[3 * v + 1] << " " << m.xyz[3 * v + 2] << "\n";
    for (int f = 0; f < m.nF(); ++f) out << "3 " << m.tri[3 * f] << " " << m.tri[3 * f + 1] << " " << m.tri[3 * f + 2] << "\n";
}

}  // namespace csg
