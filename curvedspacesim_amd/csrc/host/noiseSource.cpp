#include "noiseSource.h"

noiseSource::noiseSource(bool rep) : Reproducible(rep), gen(13377) {
    std::random_device rd;
    genrd = std::mt19937(rd());
}

int noiseSource::getInt(int minimum, int maximum) {
    std::uniform_int_distribution<int> d(minimum, maximum);
    return d(engine());
}

double noiseSource::getRealUniform(double minimum, double maximum) {
    std::uniform_real_distribution<double> d(minimum, maximum);
    return d(engine());
}

double noiseSource::getRealNormal(double mean, double std) {
    std::normal_distribution<> d(mean, std);
    return d(engine());
}

double3 noiseSource::getRandomBarycentricSet() {
    double3 a;
    a.x = getRealUniform(0, 1);
    a.y = getRealUniform(0, 1 - a.x);
    a.z = 1 - a.x - a.y;
    return a;
}

void noiseSource::setReproducibleSeed(int seed) {
    RNGSeed = seed;
    gen = std::mt19937(seed);
}
