// ORACLE — test infrastructure only (see oracle_internal.hpp).
//
// The callers around the geodesic solve, restated in FP64:
//   cell-list candidates    src/utility/hyperRectangularCellList.cpp:9-159, src/utility/cellListNeighborStructure.cpp:45-84
//   all-to-all candidates   src/utility/baseNeighborStructure.cpp:17-36
//   per-source driver       src/models/simpleModel.cpp:68-112
//   pair forces             src/forces/baseForce.cpp:12-44, harmonicRepulsion.cpp, gaussianRepulsion.cpp
//   integrators             src/updaters/velocityVerletNVE.cpp:3-29, src/updaters/noseHooverNVT.cpp:3-150
// The truncated Gaussian, U = eps*exp(-d^2/(2 s^2)) for d < rc, has no reference counterpart.
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"
#include "oracle_internal.hpp"

using namespace orc;

namespace {

struct CellGrid {
    double lo[3], hi[3], size[3];
    int n[3];
    int idx(int x, int y, int z) const { return x + n[0] * (y + n[1] * z); }
    int cellOf(const double* p) const {
        int c[3];
        for (int a = 0; a < 3; ++a) {
            int v = (int)std::floor((p[a] - lo[a]) / size[a]);
            c[a] = std::max(0, std::min(n[a] - 1, v));
        }
        return idx(c[0], c[1], c[2]);
    }
};

}  // namespace

extern "C" {

// Euclidean candidates of every particle in the reference's order: the up-to-27 cells around the particle's cell,
// x outermost, then y, then z; inside a cell by increasing particle index; strict |xi-xj|^2 < range^2. maxDist[i] is
// the largest accepted candidate distance (0 when none). Returns the total number of candidates, or -(needed) when
// `cap` is too small.
long long orc_cell_candidates(int N, const double* pos, const double* lo, const double* hi, double gridSize, double range,
                              int* offsets, int* idx, double* maxDist, long long cap) {
    CellGrid g;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = lo[a];
        g.hi[a] = hi[a];
        double ext = hi[a] - lo[a];
        g.n[a] = std::max(1, (int)std::floor(ext / gridSize));
        g.size[a] = ext / g.n[a];
    }
    int nc = g.n[0] * g.n[1] * g.n[2];
    std::vector<int> cellStart(nc + 1, 0), order(N);
    std::vector<int> cellOf(N);
    for (int i = 0; i < N; ++i) {
        cellOf[i] = g.cellOf(pos + 3 * i);
        cellStart[cellOf[i] + 1]++;
    }
    for (int c = 0; c < nc; ++c) cellStart[c + 1] += cellStart[c];
    std::vector<int> fill(cellStart.begin(), cellStart.end() - 1);
    for (int i = 0; i < N; ++i) order[fill[cellOf[i]]++] = i;
    double r2 = range * range;
    long long total = 0;
    for (int i = 0; i < N; ++i) {
        offsets[i] = (int)std::min<long long>(total, cap);
        int c = cellOf[i];
        int cx = c % g.n[0], cy = (c / g.n[0]) % g.n[1], cz = c / (g.n[0] * g.n[1]);
        double m2 = 0;
        for (int x = std::max(0, cx - 1); x <= std::min(g.n[0] - 1, cx + 1); ++x)
            for (int y = std::max(0, cy - 1); y <= std::min(g.n[1] - 1, cy + 1); ++y)
                for (int z = std::max(0, cz - 1); z <= std::min(g.n[2] - 1, cz + 1); ++z) {
                    int cc = g.idx(x, y, z);
                    for (int q = cellStart[cc]; q < cellStart[cc + 1]; ++q) {
                        int j = order[q];
                        if (j == i) continue;
                        double dx = pos[3 * i] - pos[3 * j], dy = pos[3 * i + 1] - pos[3 * j + 1], dz = pos[3 * i + 2] - pos[3 * j + 2];
                        double d2 = dx * dx + dy * dy + dz * dz;
                        if (d2 < r2) {
                            if (total < cap) idx[total] = j;
                            ++total;
                            if (d2 > m2) m2 = d2;
                        }
                    }
                }
        maxDist[i] = std::sqrt(m2);
    }
    offsets[N] = (int)std::min<long long>(total, cap);
    return total <= cap ? total : -total;
}

// pair force and energy of one potential; kind 0 = harmonic (p0 = k, p1 = sigma), 1 = reference Gaussian
// (p0 = alpha, p1 = sigma), 2 = truncated Gaussian (p0 = eps, p1 = s, p2 = rc)
double orc_pair_force(int kind, const double* p, double d, double* scaleOut) {
    double e = 0, sc = 0;
    if (kind == 0) {
        double k = p[0], sg = p[1];
        if (d <= sg) sc = -k * (sg - d);
        if (d < sg) e = 0.5 * k * (sg - d) * (sg - d);
    } else if (kind == 1) {
        double a = p[0], sg = p[1];
        const double sqrt2pi = 2.50662827463100050241576528481104525300698674061;
        double ex = std::exp(-d * d / (2.0 * sg * sg));
        e = a * ex / (sqrt2pi * sg);
        sc = -(d * a * ex / (sqrt2pi * sg * std::sqrt(sg)));
    } else {
        double eps = p[0], s = p[1], rc = p[2];
        if (d < rc) {
            double ex = std::exp(-d * d / (2.0 * s * s));
            e = eps * ex;
            sc = -eps * (d / (s * s)) * ex;
        }
    }
    *scaleOut = sc;
    return e;
}

// findNeighbors + force::computeForces for every source: candidates (CSR) are given; each source's geodesics are
// solved with triangulatedMeshSpace::distance semantics; forces[i] = sum_j pairForce(startTangent_ij, d_ij).
// Pair outputs are aligned with the candidate CSR. Sources are distributed over `threads` OpenMP threads.
int orc_compute_forces(const orc_mesh* om, int N, const int* face, const double* bary, int submeshing, double maximumDistance,
                       const int* offsets, const int* idx, const double* maxDist, int kind, const double* params, double* forces,
                       double* pairDist, double* pairTan, double* pairEnd, int* pairFlags, double* energy, int threads) {
    int err = 0;
    double etot = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : etot)
#endif
    for (int i = 0; i < N; ++i) {
        int b = offsets[i], e = offsets[i + 1], nT = e - b;
        std::vector<int> tf(nT);
        std::vector<double> tb(3 * nT), d(nT), st(3 * nT), et(3 * nT);
        std::vector<int> fl(nT);
        for (int q = 0; q < nT; ++q) {
            int j = idx[b + q];
            tf[q] = face[j];
            for (int c = 0; c < 3; ++c) tb[3 * q + c] = bary[3 * j + c];
        }
        double thr = maxDist ? maxDist[i] : 1e20;
        int rc = 0;
        if (nT > 0)
            rc = orc_distance(om, submeshing, maximumDistance, face[i], bary + 3 * i, nT, tf.data(), tb.data(), thr, d.data(), st.data(),
                              et.data(), fl.data());
        if (rc) {
#ifdef _OPENMP
#pragma omp critical
#endif
            err = rc;
            continue;
        }
        double F[3] = {0, 0, 0};
        for (int q = 0; q < nT; ++q) {
            double sc;
            etot += orc_pair_force(kind, params, d[q], &sc);
            for (int c = 0; c < 3; ++c) F[c] += sc * st[3 * q + c];
            if (pairDist) pairDist[b + q] = d[q];
            if (pairTan)
                for (int c = 0; c < 3; ++c) pairTan[3 * (b + q) + c] = st[3 * q + c];
            if (pairEnd)
                for (int c = 0; c < 3; ++c) pairEnd[3 * (b + q) + c] = et[3 * q + c];
            if (pairFlags) pairFlags[b + q] = fl[q];
        }
        for (int c = 0; c < 3; ++c) forces[3 * i + c] = F[c];
    }
    if (energy) *energy = etot;
    return err;
}

// Moves every particle by disp (velocity transported when vel != nullptr), as simpleModel::moveParticles
// (src/models/simpleModel.cpp:44-66) with only velocity transport requested.
int orc_move(const orc_mesh* om, int N, int* face, double* bary, const double* disp, double* vel, int threads) {
    int bad = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static) reduction(+ : bad)
#endif
    for (int i = 0; i < N; ++i) {
        double d[3] = {disp[3 * i], disp[3 * i + 1], disp[3 * i + 2]};
        int rc = orc_transport(om, face + i, bary + 3 * i, d, vel ? 1 : 0, vel ? vel + 3 * i : nullptr);
        if (rc < 0) ++bad;
    }
    return bad;
}

}  // extern "C"
