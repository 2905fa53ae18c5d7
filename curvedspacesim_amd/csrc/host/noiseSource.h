// Random numbers with the reference's draw semantics (src/utility/noiseSource.h:18-29, src/utility/noiseSource.cpp:5-61):
// a reproducible std::mt19937 seeded 13377 and a non-reproducible one seeded from std::random_device; every draw
// builds a fresh standard-library distribution, so with libstdc++ the sequences equal the reference's.
#pragma once

#include <random>

#include "pointDataType.h"

class noiseSource {
public:
    explicit noiseSource(bool rep = false);
    // integer in [minimum, maximum], both ends included
    int getInt(int minimum, int maximum);
    double getRealUniform(double minimum = 0., double maximum = 1.);
    double getRealNormal(double mean = 0., double std = 1.);
    // u ~ U(0,1), v ~ U(0, 1-u), w = 1 - u - v
    double3 getRandomBarycentricSet();
    void setReproducible(bool rep) { Reproducible = rep; }
    void setReproducibleSeed(int seed);

    bool Reproducible;
    int RNGSeed = 13377;
    std::mt19937 gen;
    std::mt19937 genrd;

private:
    std::mt19937& engine() { return Reproducible ? gen : genrd; }
};
