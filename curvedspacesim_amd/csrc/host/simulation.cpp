#include "simulation.h"

#include <algorithm>
#include <cmath>
#include <cstdio>

void updater::Update(int timestep) {
    iterations = timestep;
    if (maxIterations > 0 && maxIterations < iterations) return;
    if (Period <= 0 || (timestep + Phase) % Period == 0) performUpdate();
}

void updater::performUpdate() { sim->computeForces(); }

double updater::getForceNorm() {
    std::vector<double> norm(1, 0.0);
    for (int i = 0; i < Ndof; ++i) norm[0] += model->forces[i].squared_length();
    sim->manipulateUpdaterData(norm, [](double a, double b) { return a + b; });
    squaredTotalForceNorm = norm[0];
    return std::sqrt(squaredTotalForceNorm);
}

double updater::getMaxForce() {
    std::vector<double> m(1, 0.0);
    for (int i = 0; i < Ndof; ++i) m[0] = std::max(m[0], model->forces[i].squared_length());
    sim->manipulateUpdaterData(m, [](double a, double b) { return std::max(a, b); });
    maximumForceNorm = std::sqrt(m[0]);
    return maximumForceNorm;
}

void simulation::addUpdater(UpdaterPtr u, ConfigPtr c) {
    u->setModel(c);
    u->setSimulation(getPointer());
    updaters.push_back(u);
}

void simulation::addForce(ForcePtr f, ConfigPtr c) {
    f->setModel(c);
    forceComputers.push_back(f);
}

void simulation::setIntegrationTimestep(double dt) {
    integrationTimestep = dt;
    for (auto& w : updaters) w.lock()->setDeltaT(dt);
}

void simulation::setReproducible(bool reproducible) {
    for (auto& w : updaters) w.lock()->setReproducible(reproducible);
}

void simulation::computeForces() {
    auto conf = configuration.lock();
    for (size_t f = 0; f < forceComputers.size(); ++f) forceComputers[f].lock()->computeForces(conf->forces, f == 0);
}

double simulation::computePotentialEnergy(bool verbose) {
    double e = 0;
    for (auto& w : forceComputers) e += w.lock()->computeEnergy(verbose);
    return e;
}

void simulation::moveParticles(std::vector<vector3>& displacements) { configuration.lock()->moveParticles(displacements); }

void simulation::performTimestep() {
    integerTimestep += 1;
    Time += integrationTimestep;
    for (auto& w : updaters) w.lock()->Update(integerTimestep);
}

void velocityVerletNVE::performUpdate() {
    if ((int)displacements.size() != Ndof) displacements.resize(Ndof);
    velocityVerletFirstHalfStep();
    velocityVerletSecondHalfStep();
}

void velocityVerletNVE::velocityVerletFirstHalfStep() {
    for (int i = 0; i < Ndof; ++i) {
        displacements[i] = deltaT * model->velocities[i] + 0.5 * deltaT * deltaT * model->forces[i];
        model->velocities[i] += 0.5 * deltaT * model->forces[i];
    }
}

void velocityVerletNVE::velocityVerletSecondHalfStep() {
    sim->moveParticles(displacements);
    sim->computeForces();
    for (int i = 0; i < Ndof; ++i) model->velocities[i] += 0.5 * deltaT * model->forces[i];
}

noseHooverNVT::noseHooverNVT(double dt, double T, double tau_, int M) : temperature(T), tau(tau_), chainLength(M) {
    deltaT = dt;
    dt2 = 0.5 * dt;
    dt4 = 0.25 * dt;
    dt8 = 0.125 * dt;
    bathVariables.assign(M + 1, chainVariable());
}

void noseHooverNVT::setBathVariables() {
    bathVariables[0].w = 2.0 * (Ndof - 1) * temperature * tau * tau;
    for (size_t i = 1; i < bathVariables.size(); ++i) bathVariables[i].w = temperature * tau * tau;
    kineticEnergy = bathVariables[0].w;
    kineticEnergyScaleFactor = 1.0;
}

// sum of |v|^2 over 2 Ndof (noseHooverNVT.cpp:141-150)
double noseHooverNVT::getTemperatureFromKE() {
    double v2 = 0;
    for (int i = 0; i < Ndof; ++i) v2 = v2 + model->velocities[i] * model->velocities[i];
    return v2 / (2 * Ndof);
}

void noseHooverNVT::performUpdate() {
    if ((int)displacements.size() != Ndof) displacements.resize(Ndof);
    propagateChain();
    for (int i = 0; i < Ndof; ++i) model->velocities[i] = kineticEnergyScaleFactor * model->velocities[i];
    propagatePositionsVelocities();
    propagateChain();
    for (int i = 0; i < Ndof; ++i) model->velocities[i] = kineticEnergyScaleFactor * model->velocities[i];
}

// quarter-step updates of the chain velocities from the top down, half-step chain positions, the velocity scale,
// then quarter-step updates from the bottom up (noseHooverNVT.cpp:65-108)
void noseHooverNVT::propagateChain() {
    auto& b = bathVariables;
    for (int i = chainLength - 1; i > 0; --i) {
        b[i].z = (b[i - 1].w * b[i - 1].y * b[i - 1].y - temperature) / b[i].w;
        double ef = std::exp(-dt8 * b[i + 1].y);
        b[i].y = (b[i].y * ef + b[i].z * dt4) * ef;
    }
    b[0].z = 2.0 * kineticEnergy / b[0].w - 1.0;
    double ef = std::exp(-dt8 * b[1].y);
    b[0].y = (b[0].y * ef + b[0].z * dt4) * ef;
    for (int i = 0; i < chainLength; ++i) b[i].x += dt2 * b[i].y;
    kineticEnergyScaleFactor = std::exp(-dt2 * b[0].y);
    kineticEnergy = kineticEnergyScaleFactor * kineticEnergyScaleFactor * kineticEnergy;
    b[0].z = 2.0 * kineticEnergy / b[0].w - 1.0;
    ef = std::exp(-dt8 * b[1].y);
    b[0].y = (b[0].y * ef + b[0].z * dt4) * ef;
    for (int i = 1; i < chainLength; ++i) {
        b[i].z = (b[i - 1].w * b[i - 1].y * b[i - 1].y - temperature) / b[i].w;
        double e2 = std::exp(-dt8 * b[i + 1].y);
        b[i].y = (b[i].y * e2 + b[i].z * dt4) * e2;
    }
}

// half-step move, forces, full velocity kick with the kinetic energy, half-step move (noseHooverNVT.cpp:114-139)
void noseHooverNVT::propagatePositionsVelocities() {
    kineticEnergy = 0.0;
    for (int i = 0; i < Ndof; ++i) displacements[i] = dt2 * model->velocities[i];
    sim->moveParticles(displacements);
    sim->computeForces();
    for (int i = 0; i < Ndof; ++i) {
        model->velocities[i] = model->velocities[i] + (deltaT / model->masses[i]) * model->forces[i];
        displacements[i] = dt2 * model->velocities[i];
        kineticEnergy += 0.5 * model->masses[i] * (model->velocities[i] * model->velocities[i]);
    }
    sim->moveParticles(displacements);
}
