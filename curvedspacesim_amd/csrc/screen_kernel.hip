// Learned pair screen on gfx950: MLP 16 -> 64 -> 64 -> 1 with ReLU, BF16 operands on the matrix cores
// (v_mfma_f32_16x16x16_bf16), FP32 accumulation, FP32 master weights. No reference counterpart.
//
// A workgroup of four wavefronts takes 64 samples per iteration, 16 per wavefront. For a 16 x 16 x 16 product a lane l
// (r = l % 16, g = l / 16) holds A[r][4g .. 4g+3] and B[4g .. 4g+3][r] and receives D[4g .. 4g+3][r]; a layer output
// in that D layout is moved through LDS to become the A operand of the next layer. The backward pass reuses the D
// layouts directly: H1 in D layout is the A operand of H1^T dH2, dH2 and dH1 in D layout are B operands.
#include <math.h>

#include "screen_kernel.h"

namespace csg {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int H = kScreenHidden;

// float -> bfloat16, round to nearest even (finite inputs)
__device__ inline short toBf16(float x) {
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// weights staged in LDS once per workgroup: W1 and W2 as bfloat16, biases and W3 as float
struct Staged {
    short w1[kScreenIn * H];
    short w2[H * H];
    float b1[H], b2[H], w3[H];
    float b3;
};

__device__ void stageWeights(const float* p, Staged& s) {
    for (int i = threadIdx.x; i < kScreenIn * H; i += blockDim.x) s.w1[i] = toBf16(p[kScreenW1 + i]);
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) s.w2[i] = toBf16(p[kScreenW2 + i]);
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        s.b1[i] = p[kScreenB1 + i];
        s.b2[i] = p[kScreenB2 + i];
        s.w3[i] = p[kScreenW3 + i];
    }
    if (threadIdx.x == 0) s.b3 = p[kScreenB3];
}

// Forward pass of one 16-sample tile of this wavefront. x: the tile's features (16 x 16 floats, rows past the end
// zero); tr: 16 x 64 floats of wavefront-private LDS. Outputs in D layout: h1[j][i] = H1[4g+i][16j+r] and
// h2[j][i] = H2[4g+i][16j+r] (after ReLU), z[i] = logit of sample 4g+i (every lane of the row group holds it).
__device__ void forwardTile(const Staged& s, const float* x, float* tr, f32x4 h1[4], f32x4 h2[4], float z[4]) {
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    bf16x4 a;
    for (int i = 0; i < 4; ++i) a[i] = toBf16(x[r * kScreenIn + 4 * g + i]);
    for (int j = 0; j < 4; ++j) {
        bf16x4 b;
        for (int i = 0; i < 4; ++i) b[i] = s.w1[(4 * g + i) * H + 16 * j + r];
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
        c = mfma(a, b, c);
        for (int i = 0; i < 4; ++i) c[i] = fmaxf(c[i] + s.b1[16 * j + r], 0.f);
        h1[j] = c;
        for (int i = 0; i < 4; ++i) tr[(4 * g + i) * H + 16 * j + r] = c[i];
    }
    __syncthreads();
    bf16x4 a2[4];
    for (int kb = 0; kb < 4; ++kb)
        for (int i = 0; i < 4; ++i) a2[kb][i] = toBf16(tr[r * H + 16 * kb + 4 * g + i]);
    float zp[4] = {0.f, 0.f, 0.f, 0.f};
    for (int j = 0; j < 4; ++j) {
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
        for (int kb = 0; kb < 4; ++kb) {
            bf16x4 b;
            for (int i = 0; i < 4; ++i) b[i] = s.w2[(16 * kb + 4 * g + i) * H + 16 * j + r];
            c = mfma(a2[kb], b, c);
        }
        for (int i = 0; i < 4; ++i) {
            c[i] = fmaxf(c[i] + s.b2[16 * j + r], 0.f);
            zp[i] += c[i] * s.w3[16 * j + r];
        }
        h2[j] = c;
    }
    for (int i = 0; i < 4; ++i) {
        float v = zp[i];
        for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
        z[i] = v + s.b3;
    }
    __syncthreads();  // tr is rewritten by the next use
}

// samples [t0, t0 + 64) of the workgroup's tile into LDS, zero past n
__device__ void loadTile(const float* X, int n, int t0, float* xs) {
    for (int i = threadIdx.x; i < 64 * kScreenIn; i += blockDim.x) {
        const int s = t0 + i / kScreenIn;
        xs[i] = s < n ? X[(size_t)t0 * kScreenIn + i] : 0.f;
    }
}

__global__ __launch_bounds__(kScreenThreads) void screenForwardKernel(const float* params, const float* X, int n, const int* nDev,
                                                                     float* logits) {
    __shared__ Staged s;
    __shared__ float xs[64 * kScreenIn];
    __shared__ float tr[4][16 * H];
    if (nDev) n = *nDev;
    stageWeights(params, s);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    for (int t0 = blockIdx.x * 64; t0 < n; t0 += gridDim.x * 64) {
        __syncthreads();
        loadTile(X, n, t0, xs);
        __syncthreads();
        f32x4 h1[4], h2[4];
        float z[4];
        forwardTile(s, xs + w * 16 * kScreenIn, tr[w], h1, h2, z);
        if (r == 0)
            for (int i = 0; i < 4; ++i) {
                const int smp = t0 + 16 * w + 4 * g + i;
                if (smp < n) logits[smp] = z[i];
            }
    }
}

// sum over the four row groups of a wavefront (lanes r, r+16, r+32, r+48)
__device__ inline float groupSum(float v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

// Forward + backward over the workgroup's samples; writes the workgroup's partial gradient (and partial loss) to
// partial[blockIdx.x * (kScreenParams + 1) ...]. Gradients are of the mean loss over all n samples.
__global__ __launch_bounds__(kScreenThreads) void screenGradientKernel(const float* params, const float* X, const float* y, int n,
                                                                      float* partial) {
    __shared__ Staged s;
    __shared__ float xs[64 * kScreenIn];
    __shared__ float tr[4][16 * H];
    __shared__ float acc[kScreenParams + 1];
    stageWeights(params, s);
    for (int i = threadIdx.x; i <= kScreenParams; i += blockDim.x) acc[i] = 0.f;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    const float invN = 1.f / (float)n;
    f32x4 gW2[4][4], gW1[4];
    for (int a = 0; a < 4; ++a) {
        gW1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int b = 0; b < 4; ++b) gW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    float gW3[4] = {0.f, 0.f, 0.f, 0.f}, gB2[4] = {0.f, 0.f, 0.f, 0.f}, gB1[4] = {0.f, 0.f, 0.f, 0.f};
    float gB3 = 0.f, loss = 0.f;
    for (int t0 = blockIdx.x * 64; t0 < n; t0 += gridDim.x * 64) {
        __syncthreads();
        loadTile(X, n, t0, xs);
        __syncthreads();
        const float* x = xs + w * 16 * kScreenIn;
        f32x4 h1[4], h2[4];
        float z[4];
        forwardTile(s, x, tr[w], h1, h2, z);
        // dL/dz of the mean binary cross-entropy with logits
        float dz[4];
        for (int i = 0; i < 4; ++i) {
            const int smp = t0 + 16 * w + 4 * g + i;
            if (smp < n) {
                const float yy = y[smp];
                dz[i] = (1.f / (1.f + expf(-z[i])) - yy) * invN;
                if (r == 0) loss += (fmaxf(z[i], 0.f) - z[i] * yy + log1pf(expf(-fabsf(z[i])))) * invN;
            } else {
                dz[i] = 0.f;
            }
            if (r == 0) gB3 += dz[i];
        }
        // layer 3 and the ReLU of layer 2
        f32x4 d2[4];
        for (int j = 0; j < 4; ++j) {
            const float w3 = s.w3[16 * j + r];
            for (int i = 0; i < 4; ++i) {
                gW3[j] += h2[j][i] * dz[i];
                d2[j][i] = h2[j][i] > 0.f ? dz[i] * w3 : 0.f;
                gB2[j] += d2[j][i];
            }
        }
        // gW2 += H1^T dH2 (A: H1 in D layout, B: dH2 in D layout)
        bf16x4 bd2[4];
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) bd2[j][i] = toBf16(d2[j][i]);
        for (int kb = 0; kb < 4; ++kb) {
            bf16x4 ah;
            for (int i = 0; i < 4; ++i) ah[i] = toBf16(h1[kb][i]);
            for (int jn = 0; jn < 4; ++jn) gW2[kb][jn] = mfma(ah, bd2[jn], gW2[kb][jn]);
        }
        // dH1 = dH2 W2^T (A: dH2 through LDS, B: W2 transposed), then the ReLU of layer 1
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) tr[w][(4 * g + i) * H + 16 * j + r] = d2[j][i];
        __syncthreads();
        bf16x4 ad[4];
        for (int nb = 0; nb < 4; ++nb)
            for (int i = 0; i < 4; ++i) ad[nb][i] = toBf16(tr[w][r * H + 16 * nb + 4 * g + i]);
        f32x4 d1[4];
        for (int j = 0; j < 4; ++j) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
            for (int nb = 0; nb < 4; ++nb) {
                bf16x4 b;
                for (int i = 0; i < 4; ++i) b[i] = s.w2[(16 * j + r) * H + 16 * nb + 4 * g + i];
                c = mfma(ad[nb], b, c);
            }
            for (int i = 0; i < 4; ++i) {
                c[i] = h1[j][i] > 0.f ? c[i] : 0.f;
                gB1[j] += c[i];
            }
            d1[j] = c;
        }
        // gW1 += X^T dH1 (A: X transposed from LDS, B: dH1 in D layout)
        bf16x4 ax;
        for (int i = 0; i < 4; ++i) ax[i] = toBf16(x[(4 * g + i) * kScreenIn + r]);
        for (int j = 0; j < 4; ++j) {
            bf16x4 b;
            for (int i = 0; i < 4; ++i) b[i] = toBf16(d1[j][i]);
            gW1[j] = mfma(ax, b, gW1[j]);
        }
        __syncthreads();
    }
    // combine the wavefronts in fixed order
    for (int j = 0; j < 4; ++j) {
        gW3[j] = groupSum(gW3[j]);
        gB2[j] = groupSum(gB2[j]);
        gB1[j] = groupSum(gB1[j]);
    }
    gB3 = groupSum(gB3);
    loss = groupSum(loss);
    for (int ww = 0; ww < 4; ++ww) {
        __syncthreads();
        if (w != ww) continue;
        for (int kb = 0; kb < 4; ++kb)
            for (int jn = 0; jn < 4; ++jn)
                for (int i = 0; i < 4; ++i) acc[kScreenW2 + (16 * kb + 4 * g + i) * H + 16 * jn + r] += gW2[kb][jn][i];
        for (int j = 0; j < 4; ++j)
            for (int i = 0; i < 4; ++i) acc[kScreenW1 + (4 * g + i) * H + 16 * j + r] += gW1[j][i];
        if (g == 0)
            for (int j = 0; j < 4; ++j) {
                acc[kScreenW3 + 16 * j + r] += gW3[j];
                acc[kScreenB2 + 16 * j + r] += gB2[j];
                acc[kScreenB1 + 16 * j + r] += gB1[j];
            }
        if (l == 0) {
            acc[kScreenB3] += gB3;
            acc[kScreenParams] += loss;
        }
    }
    __syncthreads();
    float* out = partial + (size_t)blockIdx.x * (kScreenParams + 1);
    for (int i = threadIdx.x; i <= kScreenParams; i += blockDim.x) out[i] = acc[i];
}

__global__ void screenReduceKernel(const float* partial, int blocks, float* grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kScreenParams) return;
    float s = 0.f;
    for (int b = 0; b < blocks; ++b) s += partial[(size_t)b * (kScreenParams + 1) + p];
    grad[p] = s;
}

__global__ void screenAdamKernel(float* params, const float* grad, float* m, float* v, int step, float lr) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kScreenParams) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gr = grad[p];
    const float mm = b1 * m[p] + (1.f - b1) * gr;
    const float vv = b2 * v[p] + (1.f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.f - powf(b1, (float)step));
    const float vh = vv / (1.f - powf(b2, (float)step));
    params[p] -= lr * mh / (sqrtf(vh) + eps);
}

// features of pair q: its source by binary search over the offsets
__global__ void screenFeaturesKernel(ScreenFeatureParams P, int nPairs) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nPairs || (P.nPairsDev && q >= *P.nPairsDev)) return;
    int lo = 0, hi = P.nSrc - 1;
    while (lo < hi) {  // last source whose first pair is at or before q
        const int mid = (lo + hi + 1) >> 1;
        if (P.toff[mid] <= q) lo = mid;
        else hi = mid - 1;
    }
    const int i = lo, j = P.tidx[q];
    const DeviceMesh& M = P.mesh;
    const int fi = P.face[i], fj = P.face[j];
    const double* ni = M.fnrm + 3 * fi;
    const double* nj = M.fnrm + 3 * fj;
    double curvI[2] = {0, 0}, curvJ[2] = {0, 0};
    for (int side = 0; side < 2; ++side) {
        const int p = side ? j : i, f = side ? fj : fi;
        const double* b = P.bary + 3 * p;
        const double sb = b[0] + b[1] + b[2];
        double* c = side ? curvJ : curvI;
        for (int k = 0; k < 3; ++k) {
            const int v = M.fv[3 * f + k];
            c[0] += b[k] / sb * P.curv[2 * v];
            c[1] += b[k] / sb * P.curv[2 * v + 1];
        }
    }
    double d[3] = {P.xyz[3 * j] - P.xyz[3 * i], P.xyz[3 * j + 1] - P.xyz[3 * i + 1], P.xyz[3 * j + 2] - P.xyz[3 * i + 2]};
    const double dE = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double rc = P.cutoff;
    const double riseI = d[0] * ni[0] + d[1] * ni[1] + d[2] * ni[2];  // rise of the chord above the source's plane
    const double riseJ = d[0] * nj[0] + d[1] * nj[1] + d[2] * nj[2];
    const double dd = dE > 0 ? dE : 1e-300;
    float* o = P.out + (size_t)q * kScreenIn;
    // the distance as its margin to the cutoff in percent, so that bfloat16 resolves it near the cutoff
    o[0] = (float)(100.0 * (1.0 - dE / rc));
    o[1] = (float)(ni[0] * nj[0] + ni[1] * nj[1] + ni[2] * nj[2]);
    for (int a = 0; a < 3; ++a) {
        o[2 + a] = (float)ni[a];
        o[5 + a] = (float)nj[a];
    }
    o[8] = (float)(curvI[0] * rc);
    o[9] = (float)(curvI[1] * rc * rc);
    o[10] = (float)(curvJ[0] * rc);
    o[11] = (float)(curvJ[1] * rc * rc);
    o[12] = (float)(P.patchCount ? P.patchCount[i] * 1e-3 : 0.0);
    // arc-over-chord excess of a circle through both points with the chord rising h above one tangent plane,
    // h^2 / (6 |d|), in percent of the cutoff, seen from either end; and the rise's slope
    o[13] = (float)(100.0 * riseI * riseI / (6.0 * dd * rc));
    o[14] = (float)(100.0 * riseJ * riseJ / (6.0 * dd * rc));
    o[15] = (float)(10.0 * riseI / dd);
}

}  // namespace

int screenTrainBlocks(int n) { return n <= 0 ? 1 : (n + 63) / 64 < 1024 ? (n + 63) / 64 : 1024; }

hipError_t launchScreenForward(const float* params, const float* X, int n, const int* nDev, float* logits, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(screenForwardKernel, dim3(screenTrainBlocks(n)), dim3(kScreenThreads), 0, stream, params, X, n, nDev, logits);
    return hipGetLastError();
}

hipError_t launchScreenGradient(const float* params, const float* X, const float* y, int n, float* partial, float* grad,
                                hipStream_t stream) {
    if (n <= 0) return hipErrorInvalidValue;
    const int blocks = screenTrainBlocks(n);
    hipLaunchKernelGGL(screenGradientKernel, dim3(blocks), dim3(kScreenThreads), 0, stream, params, X, y, n, partial);
    hipLaunchKernelGGL(screenReduceKernel, dim3((kScreenParams + 256) / 256), dim3(256), 0, stream, partial, blocks, grad);
    return hipGetLastError();
}

hipError_t launchScreenAdam(float* params, const float* grad, float* m, float* v, int step, float lr, hipStream_t stream) {
    hipLaunchKernelGGL(screenAdamKernel, dim3((kScreenParams + 255) / 256), dim3(256), 0, stream, params, grad, m, v, step, lr);
    return hipGetLastError();
}

hipError_t launchScreenFeatures(const ScreenFeatureParams& P, int nPairs, hipStream_t stream) {
    if (nPairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(screenFeaturesKernel, dim3((nPairs + 255) / 256), dim3(256), 0, stream, P, nPairs);
    return hipGetLastError();
}

}  // namespace csg
