// Simulation, updater and integrators of the reference (src/simulation/simulation.h:27-60,
// src/simulation/basicSimulation.h:9-47, src/updaters/baseUpdater.h:15-104, src/updaters/velocityVerletNVE.cpp:3-29,
// src/updaters/noseHooverNVT.cpp:3-139) with the same call pattern: an updater asks the simulation to move particles
// and to compute forces, the simulation forwards to the model and to every force computer.
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "forces.h"
#include "simpleModel.h"

class basicSimulation {
public:
    virtual ~basicSimulation() = default;
    virtual void computeForces() = 0;
    virtual void moveParticles(std::vector<vector3>& displacements) = 0;
    virtual double computePotentialEnergy(bool = false) { return 0.0; }
    virtual void manipulateUpdaterData(std::vector<double>&, std::function<double(double, double)>) {}
    WeakConfigPtr configuration;
    int integerTimestep = 0;
    double Time = 0.;
    double integrationTimestep = 0.01;
};

class updater {
public:
    updater() = default;
    explicit updater(int p) : Period(p) {}
    virtual ~updater() = default;
    // performUpdate when the period and phase select this timestep (src/updaters/baseUpdater.cpp:5-13)
    virtual void Update(int timestep);
    virtual void performUpdate();
    void setSimulation(std::shared_ptr<basicSimulation> s) { sim = s; }
    virtual void setModel(std::shared_ptr<simpleModel> m) {
        model = m;
        initializeFromModel();
    }
    virtual void initializeFromModel() { Ndof = model->getNumberOfParticles(); }
    void setPeriod(int p) { Period = p; }
    void setPhase(int p) { Phase = p; }
    virtual void setReproducible(bool rep) { reproducible = rep; }
    int getNdof() { return Ndof; }
    void setNdof(int n) { Ndof = n; }
    virtual void setDeltaT(double dt) { deltaT = dt; }
    virtual double getMaxForce();
    virtual double getForceNorm();
    void setMaximumIterations(int m = -1) { maxIterations = m; }
    int getCurrentIterations() { return iterations; }

    std::shared_ptr<basicSimulation> sim;
    std::shared_ptr<simpleModel> model;
    int iterations = 0;
    double squaredTotalForceNorm = 0, maximumForceNorm = 0;

protected:
    int Period = -1, Phase = 0, Ndof = 0, maxIterations = -1;
    bool reproducible = true;
    double deltaT = 0.001;
};

typedef std::shared_ptr<updater> UpdaterPtr;
typedef std::weak_ptr<updater> WeakUpdaterPtr;
typedef std::shared_ptr<force> ForcePtr;
typedef std::weak_ptr<force> WeakForcePtr;

class simulation : public basicSimulation, public std::enable_shared_from_this<simulation> {
public:
    void setConfiguration(ConfigPtr c) { configuration = c; }
    // the first force computer zeroes the forces, the others add (src/simulation/simulation.cpp:61-70)
    void computeForces() override;
    void moveParticles(std::vector<vector3>& displacements) override;
    void performTimestep();
    std::shared_ptr<simulation> getPointer() { return shared_from_this(); }
    void addUpdater(UpdaterPtr u) { updaters.push_back(u); }
    void addUpdater(UpdaterPtr u, ConfigPtr c);
    void addForce(ForcePtr f) { forceComputers.push_back(f); }
    void addForce(ForcePtr f, ConfigPtr c);
    void clearForceComputers() { forceComputers.clear(); }
    void clearUpdaters() { updaters.clear(); }
    void setIntegrationTimestep(double dt);
    void setReproducible(bool reproducible);
    double computePotentialEnergy(bool verbose = false) override;

    std::vector<WeakUpdaterPtr> updaters;
    std::vector<WeakForcePtr> forceComputers;
};

typedef std::shared_ptr<simulation> SimulationPtr;

// velocity Verlet NVE: x += dt v + dt^2/2 F, v += dt/2 F, forces, v += dt/2 F (masses ignored, as in the reference)
class velocityVerletNVE : public updater {
public:
    velocityVerletNVE() { deltaT = 0.001; }
    explicit velocityVerletNVE(double dt) { deltaT = dt; }
    void performUpdate() override;
    void velocityVerletFirstHalfStep();
    void velocityVerletSecondHalfStep();
    void setModel(std::shared_ptr<simpleModel> m) override {
        model = m;
        model->particleShiftsRequireVelocityTransport = true;
        initializeFromModel();
    }

protected:
    std::vector<vector3> displacements;
};

// Nose-Hoover chain NVT of Frenkel & Smit algorithms 30-32 with chain length M; Q0 = 2 (Ndof - 1) T tau^2, Qi = T tau^2
class noseHooverNVT : public updater {
public:
    noseHooverNVT(double dt, double T, double tau = 1.0, int M = 2);
    void performUpdate() override;
    void setModel(std::shared_ptr<simpleModel> m) override {
        model = m;
        model->particleShiftsRequireVelocityTransport = true;
        initializeFromModel();
        setBathVariables();
    }
    double getTemperatureFromKE();
    double getKineticEnergy() const { return kineticEnergy; }
    // (position, velocity, acceleration, mass) of each chain variable
    struct chainVariable {
        double x = 0, y = 0, z = 0, w = 0;
    };
    std::vector<chainVariable> bathVariables;

protected:
    void propagateChain();
    void propagatePositionsVelocities();
    void setBathVariables();
    std::vector<vector3> displacements;
    double temperature, tau;
    int chainLength;
    double kineticEnergy = 0, kineticEnergyScaleFactor = 1;
    double dt2, dt4, dt8;
};
