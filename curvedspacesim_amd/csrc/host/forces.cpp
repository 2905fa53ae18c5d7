#include "forces.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "gpuMeshSpace.h"

void force::computeForces(std::vector<vector3>& forces, bool zeroOutForce, int) {
    if ((int)forces.size() != model->N) forces.resize(model->N);
    model->findNeighbors(maximumInteractionRange);
    for (int i = 0; i < model->N; ++i) {
        if (zeroOutForce) forces[i] = vector3(0, 0, 0);
        const int n = (int)model->neighbors[i].size();
        for (int j = 0; j < n; ++j) forces[i] += pairwiseForce(model->neighborVectors[i][j], model->neighborDistances[i][j]);
    }
}

double force::computeEnergy(bool) {
    model->findNeighbors(maximumInteractionRange);
    energy = 0;
    for (int i = 0; i < model->N; ++i) {
        const int n = (int)model->neighbors[i].size();
        for (int j = 0; j < n; ++j) energy += pairwiseEnergy(model->neighborVectors[i][j], model->neighborDistances[i][j]);
    }
    return energy;
}

harmonicRepulsion::harmonicRepulsion(double k, double s) : stiffness(k), sigma(s) { maximumInteractionRange = s; }

void harmonicRepulsion::setForceParameters(std::vector<double>& params) {
    if (params.size() != 2) throw std::runtime_error("harmonicRepulsion takes (k, sigma); only monodisperse systems are supported");
    stiffness = params[0];
    sigma = params[1];
    maximumInteractionRange = sigma;
}

double harmonicRepulsion::pairwiseEnergy(vector3, double d) { return d < sigma ? 0.5 * stiffness * (sigma - d) * (sigma - d) : 0.0; }

vector3 harmonicRepulsion::pairwiseForce(vector3 t, double d) {
    if (d <= sigma) return (-stiffness * (sigma - d)) * t;
    return vector3(0, 0, 0);
}

gaussianRepulsion::gaussianRepulsion(double a, double s) : alpha(a), sigma(s) {}

void gaussianRepulsion::setForceParameters(std::vector<double>& params) {
    if (params.size() != 2) throw std::runtime_error("gaussianRepulsion takes (alpha, sigma)");
    alpha = params[0];
    sigma = params[1];
}

static const double kSqrt2Pi = 2.50662827463100050241576528481104525300698674061;

double gaussianRepulsion::pairwiseEnergy(vector3, double d) { return alpha * std::exp(-d * d / (2.0 * sigma * sigma)) / (kSqrt2Pi * sigma); }

vector3 gaussianRepulsion::pairwiseForce(vector3 t, double d) {
    double ex = std::exp(-d * d / (2.0 * sigma * sigma));
    return (-(d * alpha * ex / (kSqrt2Pi * sigma * std::sqrt(sigma)))) * t;
}

truncatedGaussianRepulsion::truncatedGaussianRepulsion(double e, double s_, double c) : epsilon(e), s(s_), cutoff(c) {
    maximumInteractionRange = c;
}

double truncatedGaussianRepulsion::pairwiseEnergy(vector3, double d) { return d < cutoff ? epsilon * std::exp(-d * d / (2.0 * s * s)) : 0.0; }

vector3 truncatedGaussianRepulsion::pairwiseForce(vector3 t, double d) {
    if (d >= cutoff) return vector3(0, 0, 0);
    return (-epsilon * (d / (s * s)) * std::exp(-d * d / (2.0 * s * s))) * t;
}

gpuPairForce::gpuPairForce(std::shared_ptr<force> potential) : pot(std::move(potential)) {
    if (pot->devicePotential().kind == 3) throw std::runtime_error("gpuPairForce: the potential has no device form");
    maximumInteractionRange = pot->maximumInteractionRange;
}

// candidates on the host in the reference order, then one csg_pair_forces call over all sources
void gpuPairForce::evaluate(std::vector<double>* F, double* E) {
    auto* gm = dynamic_cast<gpuModel*>(model.get());
    gpuMeshSpace* space = gm ? gm->meshSpace() : nullptr;
    if (!space) throw std::runtime_error("gpuPairForce needs a gpuModel living in a gpuMeshSpace");
    const int N = model->N;
    std::vector<int32_t> off, idx;
    std::vector<double> largest;
    gm->candidateLists(maximumInteractionRange, off, idx, largest);
    std::vector<int32_t> face(N);
    std::vector<double> bary(3 * (size_t)N), radius;
    for (int i = 0; i < N; ++i) {
        face[i] = model->positions[i].faceIndex;
        for (int a = 0; a < 3; ++a) bary[3 * i + a] = model->positions[i].x[a];
    }
    const bool sub = space->submeshingActivated();
    if (sub) {
        radius.resize(N);
        for (int i = 0; i < N; ++i) radius[i] = std::min(space->maximumDistance(), largest[i]);
    }
    csg_potential p = pot->devicePotential();
    // windows beyond the potential's range carry no force or energy; the Gaussian has none, so nothing is cut
    const double rMax = p.kind == 0 ? p.p1 : p.kind == 2 ? p.p2 : HUGE_VAL;
    std::vector<double> en(N);
    if (F) F->assign(3 * (size_t)N, 0.0);
    csg_status s = csg_pair_forces(space->context(), N, face.data(), bary.data(), N, nullptr, off.data(), idx.data(),
                                   sub ? radius.data() : nullptr, rMax, &p, sub ? 1 : 0, 2.0 * space->maximumDistance(), nullptr, nullptr,
                                   nullptr, nullptr, F ? F->data() : nullptr, E ? en.data() : nullptr);
    if (s != CSG_OK) throw std::runtime_error(std::string("csg_pair_forces: ") + csg_last_error(space->context()));
    if (E) {
        *E = 0;
        for (double e : en) *E += e;
    }
}

void gpuPairForce::computeForces(std::vector<vector3>& forces, bool zeroOutForce, int) {
    std::vector<double> F;
    evaluate(&F, nullptr);
    if ((int)forces.size() != model->N) forces.resize(model->N);
    for (int i = 0; i < model->N; ++i) {
        vector3 f(F[3 * i], F[3 * i + 1], F[3 * i + 2]);
        if (zeroOutForce)
            forces[i] = f;
        else
            forces[i] += f;
    }
}

double gpuPairForce::computeEnergy(bool) {
    double E = 0;
    evaluate(nullptr, &E);
    energy = E;
    return E;
}
