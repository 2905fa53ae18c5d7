// Learned pair screen: a 16 -> 64 -> 64 -> 1 MLP (ReLU) that scores how likely a candidate pair lies within the
// cutoff, in BF16 on the matrix cores with FP32 accumulation (no reference counterpart). Pairs scoring below a
// calibrated threshold are not solved; every other pair is solved exactly.
#pragma once

#include "csg_common.h"

namespace csg {

constexpr int kScreenIn = 16;
constexpr int kScreenHidden = 64;
// parameter vector (FP32): W1[16][64], b1[64], W2[64][64], b2[64], W3[64], b3
constexpr int kScreenW1 = 0;
constexpr int kScreenB1 = kScreenW1 + kScreenIn * kScreenHidden;
constexpr int kScreenW2 = kScreenB1 + kScreenHidden;
constexpr int kScreenB2 = kScreenW2 + kScreenHidden * kScreenHidden;
constexpr int kScreenW3 = kScreenB2 + kScreenHidden;
constexpr int kScreenB3 = kScreenW3 + kScreenHidden;
constexpr int kScreenParams = kScreenB3 + 1;  // 5,313
constexpr int kScreenThreads = 256;           // 4 wavefronts, 16 samples per wavefront per tile
constexpr int kScreenPersistentBlocks = 1024; // inference workgroups: four per CU, each wavefront looping over tiles or sources
// Pairs whose first feature, 100 (1 - |d| / rc), exceeds this are always kept, whatever their score: such a pair lies
// beyond the cutoff only if its geodesic exceeds the chord by more than 1/0.985 - 1 = 1.5%, over twice the largest
// excess of any candidate pair of the benchmark surfaces (0.71%).
constexpr float kScreenGuard = 1.5f;

// logits of n samples (features row-major n x 16); with nDev, the count is read on the device (n bounds it)
hipError_t launchScreenForward(const float* params, const float* X, int n, const int* nDev, float* logits, hipStream_t stream);
// binary cross-entropy with logits, averaged over the n samples: per-workgroup partial gradients (blocks x
// kScreenParams, plus the partial loss at index kScreenParams of each row) and their fixed-order sum into grad
// (kScreenParams + 1 floats, the last the loss)
int screenTrainBlocks(int n);
hipError_t launchScreenGradient(const float* params, const float* X, const float* y, int n, float* partial, float* grad,
                                hipStream_t stream);
// one Adam step on params with moments m, v (step counts from 1)
hipError_t launchScreenAdam(float* params, const float* grad, float* m, float* v, int step, float lr, hipStream_t stream);

// per candidate pair: the 16 features (row-major) of source i and target j
struct ScreenFeatureParams {
    DeviceMesh mesh;
    const double* curv;     // 2 per vertex: mean and Gaussian curvature
    const int32_t* face;
    const double* bary;
    const double* xyz;      // particle positions, 3 each
    int nSrc;
    const int32_t* toff;
    const int32_t* tidx;
    const int32_t* patchCount;  // faces of each source's patch (nullable)
    double cutoff;
    const int32_t* srcGroup;    // particle groups (nullable): the cutoff of source i is groupRange[srcGroup[i]]
    const double* groupRange;
    const int* nPairsDev;  // optional: pair count on the device (the launch size bounds it)
    float* out;
};
hipError_t launchScreenFeatures(const ScreenFeatureParams& P, int nPairs, hipStream_t stream);
// the screening of a force computation: per candidate pair a keep flag (pairs closer than the guard distance are kept
// unscored; the others are scored and kept when their logit reaches the threshold). counts[0] (pairs scored, the guard
// band) and counts[1] (pairs kept) must be zero before the call.
hipError_t launchScreenSources(const ScreenFeatureParams& P, const float* params, float threshold, uint8_t* keep, int* counts,
                               hipStream_t stream);

}  // namespace csg
