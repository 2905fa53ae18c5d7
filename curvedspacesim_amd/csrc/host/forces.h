// Force computers of the reference (src/forces/baseForce.h:17-63) with its pair potentials, plus gpuPairForce, which
// fuses neighbour search, geodesics and the pair sum in one device call (no host neighbour lists).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../../include/csg.h"
#include "simpleModel.h"

class force {
public:
    virtual ~force() = default;
    virtual std::string reportSelfName() { return "base force"; }
    // findNeighbors, then the sum of pairwiseForce over each particle's neighbours (src/forces/baseForce.cpp:12-28)
    virtual void computeForces(std::vector<vector3>& forces, bool zeroOutForce = true, int type = 0);
    virtual void setForceParameters(std::vector<double>&) {}
    virtual void setModel(std::shared_ptr<simpleModel> _model) { model = _model; }
    // src/forces/baseForce.cpp:33-44
    virtual double computeEnergy(bool verbose = false);
    virtual double pairwiseEnergy(vector3 separation, double distance) = 0;
    virtual vector3 pairwiseForce(vector3 separation, double distance) = 0;
    // the same potential as a device descriptor (kind 3 = none: not fusable)
    virtual csg_potential devicePotential() const { return {3, 0, 0, 0}; }
    std::shared_ptr<simpleModel> model;
    double energy = 0;
    double maximumInteractionRange = 1;
};

// U = k/2 (sigma - d)^2 for d < sigma; F = -k (sigma - d) t for d <= sigma (src/forces/harmonicRepulsion.cpp:3-34)
class harmonicRepulsion : public force {
public:
    harmonicRepulsion(double k = 1.0, double sigma = 1.0);
    std::string reportSelfName() override { return "harmonic repulsion"; }
    void setForceParameters(std::vector<double>& params) override;
    double pairwiseEnergy(vector3 separation, double distance) override;
    vector3 pairwiseForce(vector3 separation, double distance) override;
    csg_potential devicePotential() const override { return {0, stiffness, sigma, 0}; }
    double stiffness, sigma;
};

// U = alpha exp(-d^2 / 2 sigma^2) / (sigma sqrt(2 pi)); F = -d alpha exp(-d^2 / 2 sigma^2) / (sigma^(3/2) sqrt(2 pi)) t,
// no cutoff (src/forces/gaussianRepulsion.cpp:3-12, src/forces/gaussianRepulsion.h:16-24)
class gaussianRepulsion : public force {
public:
    gaussianRepulsion(double alpha = 1.0, double sigma = 1.0);
    std::string reportSelfName() override { return "gaussian repulsion"; }
    void setForceParameters(std::vector<double>& params) override;
    double pairwiseEnergy(vector3 separation, double distance) override;
    vector3 pairwiseForce(vector3 separation, double distance) override;
    csg_potential devicePotential() const override { return {1, alpha, sigma, 0}; }
    double alpha, sigma;
};

// U = epsilon exp(-d^2 / 2 s^2) and F = -epsilon (d / s^2) exp(-d^2 / 2 s^2) t for d < cutoff; zero beyond. No
// reference counterpart; its range is the cutoff.
class truncatedGaussianRepulsion : public force {
public:
    truncatedGaussianRepulsion(double epsilon = 1.0, double s = 1.0, double cutoff = 3.0);
    std::string reportSelfName() override { return "truncated gaussian repulsion"; }
    double pairwiseEnergy(vector3 separation, double distance) override;
    vector3 pairwiseForce(vector3 separation, double distance) override;
    csg_potential devicePotential() const override { return {2, epsilon, s, cutoff}; }
    double epsilon, s, cutoff;
};

// A force computer whose computeForces is one fused device call: candidates from the model's neighbour structure on
// the host, then geodesics, pair forces and the per-source sum on the device. The model must live in a gpuMeshSpace;
// the pair potential is that of the wrapped force computer.
class gpuPairForce : public force {
public:
    explicit gpuPairForce(std::shared_ptr<force> potential);
    std::string reportSelfName() override { return "gpu pair force (" + pot->reportSelfName() + ")"; }
    void computeForces(std::vector<vector3>& forces, bool zeroOutForce = true, int type = 0) override;
    double computeEnergy(bool verbose = false) override;
    double pairwiseEnergy(vector3 separation, double distance) override { return pot->pairwiseEnergy(separation, distance); }
    vector3 pairwiseForce(vector3 separation, double distance) override { return pot->pairwiseForce(separation, distance); }
    csg_potential devicePotential() const override { return pot->devicePotential(); }

private:
    void evaluate(std::vector<double>* F, double* E);
    std::shared_ptr<force> pot;
};
