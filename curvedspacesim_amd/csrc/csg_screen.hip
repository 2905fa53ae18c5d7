// C-ABI of the learned pair screen's standalone calls (inference, gradient, training); the per-step use inside the
// device-resident system is in csg_system.hip.
#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/csg.h"
#include "csg_context.h"
#include "screen_kernel.h"

using csg::fail;

namespace {

csg_status stage(csg_ctx* c, const float* params, int32_t n, const float* X, const float* y) {
    HIPCHECK(c, c->sParams.ensure(sizeof(float) * csg::kScreenParams));
    HIPCHECK(c, hipMemcpyAsync(c->sParams.p, params, sizeof(float) * csg::kScreenParams, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, c->sX.ensure(sizeof(float) * csg::kScreenIn * (size_t)n));
    HIPCHECK(c, hipMemcpyAsync(c->sX.p, X, sizeof(float) * csg::kScreenIn * (size_t)n, hipMemcpyHostToDevice, c->stream));
    if (y) {
        HIPCHECK(c, c->sY.ensure(sizeof(float) * (size_t)n));
        HIPCHECK(c, hipMemcpyAsync(c->sY.p, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    }
    return CSG_OK;
}

}  // namespace

extern "C" {

csg_status csg_screen_forward(csg_ctx* c, const float* params, int32_t n, const float* features, float* logits) {
    if (!c) return CSG_ERR_ARGUMENT;
    if (n < 0 || (n > 0 && (!params || !features || !logits))) return fail(c, CSG_ERR_ARGUMENT, "bad screen arrays");
    if (n == 0) return CSG_OK;
    HIPCHECK(c, hipSetDevice(c->device));
    csg_status s = stage(c, params, n, features, nullptr);
    if (s) return s;
    HIPCHECK(c, c->sOut.ensure(sizeof(float) * (size_t)n));
    HIPCHECK(c, csg::launchScreenForward(c->sParams.as<float>(), c->sX.as<float>(), n, nullptr, c->sOut.as<float>(), c->stream));
    HIPCHECK(c, hipMemcpyAsync(logits, c->sOut.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return CSG_OK;
}

csg_status csg_screen_gradient(csg_ctx* c, const float* params, int32_t n, const float* features, const float* labels, float* grad,
                               float* loss) {
    if (!c) return CSG_ERR_ARGUMENT;
    if (n <= 0 || !params || !features || !labels) return fail(c, CSG_ERR_ARGUMENT, "bad screen arrays");
    HIPCHECK(c, hipSetDevice(c->device));
    csg_status s = stage(c, params, n, features, labels);
    if (s) return s;
    const int blocks = csg::screenTrainBlocks(n);
    HIPCHECK(c, c->sPart.ensure(sizeof(float) * (csg::kScreenParams + 1) * (size_t)blocks));
    HIPCHECK(c, c->sGrad.ensure(sizeof(float) * (csg::kScreenParams + 1)));
    HIPCHECK(c, csg::launchScreenGradient(c->sParams.as<float>(), c->sX.as<float>(), c->sY.as<float>(), n, c->sPart.as<float>(),
                                          c->sGrad.as<float>(), c->stream));
    std::vector<float> g(csg::kScreenParams + 1);
    HIPCHECK(c, hipMemcpyAsync(g.data(), c->sGrad.p, sizeof(float) * g.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    if (grad) std::copy(g.begin(), g.end() - 1, grad);
    if (loss) *loss = g.back();
    return CSG_OK;
}

csg_status csg_screen_train(csg_ctx* c, float* params, int32_t n, const float* features, const float* labels, int32_t steps, int32_t batch,
                            float lr, uint32_t seed, float* losses) {
    if (!c) return CSG_ERR_ARGUMENT;
    if (n <= 0 || steps < 0 || batch <= 0 || !params || !features || !labels) return fail(c, CSG_ERR_ARGUMENT, "bad training arguments");
    HIPCHECK(c, hipSetDevice(c->device));
    // one seeded shuffle; minibatches are then consecutive slices, wrapping around
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937 gen(seed);
    std::shuffle(perm.begin(), perm.end(), gen);
    std::vector<float> X(csg::kScreenIn * (size_t)n), y(n);
    for (int i = 0; i < n; ++i) {
        std::copy(features + (size_t)perm[i] * csg::kScreenIn, features + (size_t)(perm[i] + 1) * csg::kScreenIn,
                  X.begin() + (size_t)i * csg::kScreenIn);
        y[i] = labels[perm[i]];
    }
    csg_status s = stage(c, params, n, X.data(), y.data());
    if (s) return s;
    batch = std::min(batch, n);
    const int blocks = csg::screenTrainBlocks(batch);
    HIPCHECK(c, c->sPart.ensure(sizeof(float) * (csg::kScreenParams + 1) * (size_t)blocks));
    HIPCHECK(c, c->sGrad.ensure(sizeof(float) * (csg::kScreenParams + 1)));
    HIPCHECK(c, c->sM.ensure(sizeof(float) * csg::kScreenParams));
    HIPCHECK(c, c->sV.ensure(sizeof(float) * csg::kScreenParams));
    HIPCHECK(c, c->sOut.ensure(sizeof(float) * (size_t)std::max(steps, 1)));
    HIPCHECK(c, hipMemsetAsync(c->sM.p, 0, sizeof(float) * csg::kScreenParams, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->sV.p, 0, sizeof(float) * csg::kScreenParams, c->stream));
    for (int step = 1; step <= steps; ++step) {
        const long start = ((long)(step - 1) * batch) % n;
        const int len = (int)std::min<long>(batch, n - start);
        HIPCHECK(c, csg::launchScreenGradient(c->sParams.as<float>(), c->sX.as<float>() + start * csg::kScreenIn, c->sY.as<float>() + start,
                                              len, c->sPart.as<float>(), c->sGrad.as<float>(), c->stream));
        HIPCHECK(c, csg::launchScreenAdam(c->sParams.as<float>(), c->sGrad.as<float>(), c->sM.as<float>(), c->sV.as<float>(), step, lr,
                                          c->stream));
        HIPCHECK(c, hipMemcpyAsync(c->sOut.as<float>() + (step - 1), c->sGrad.as<float>() + csg::kScreenParams, sizeof(float),
                                   hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHECK(c, hipMemcpyAsync(params, c->sParams.p, sizeof(float) * csg::kScreenParams, hipMemcpyDeviceToHost, c->stream));
    if (losses && steps > 0)
        HIPCHECK(c, hipMemcpyAsync(losses, c->sOut.p, sizeof(float) * (size_t)steps, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return CSG_OK;
}

}  // extern "C"
