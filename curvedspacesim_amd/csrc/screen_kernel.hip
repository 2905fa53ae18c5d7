// Learned pair screen on gfx950: MLP 16 -> 64 -> 64 -> 1 with ReLU, BF16 operands on the matrix cores, FP32
// accumulation, FP32 master weights. No reference counterpart.
//
// Inference (screenForwardKernel, screenSourcesKernel) runs the layers transposed, H^T = W^T X^T, on
// v_mfma_f32_16x16x32_bf16: the weights are the A operand (a 16 x 32 block: lane l, r = l % 16, g = l / 16, holds
// A[r][8g .. 8g+7]), the 16 samples of a tile are the columns of the B operand (lane l holds B[8g .. 8g+7][r]) and of
// the result (lane l receives D[4g .. 4g+3][r]). Layer 1's output tile j holds hidden units 16j .. 16j+15; layer 2
// reads its K dimension in the order in which those outputs already lie in the lanes (K position 8g+t of chunk kb is
// hidden unit 16 (2kb + t/4) + 4g + t%4), so the B operand of layer 2 is the lane's own layer-1 results: no exchange
// between lanes and no barrier. Each wavefront keeps its weight fragments in registers for all its tiles.
//
// Training (screenGradientKernel) uses v_mfma_f32_16x16x16_bf16 in the untransposed orientation: a workgroup of four
// wavefronts takes 64 samples per iteration, 16 per wavefront; a lane holds A[r][4g .. 4g+3] and B[4g .. 4g+3][r] and
// receives D[4g .. 4g+3][r]; a layer output in that D layout is moved through LDS to become the A operand of the next
// layer, and the backward pass reuses the D layouts directly (H1 in D layout is the A operand of H1^T dH2, dH2 and dH1
// in D layout are B operands).
#include <math.h>

#include <algorithm>

#include "screen_kernel.h"

namespace csg {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int H = kScreenHidden;

// float -> bfloat16, round to nearest even (finite inputs)
__device__ inline short toBf16(float x) {
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
__device__ inline f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ inline __bf16 bf(float x) { return __builtin_bit_cast(__bf16, toBf16(x)); }
__device__ inline __bf16 bfBits(short v) { return __builtin_bit_cast(__bf16, v); }

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

// A wavefront's register copy of the network for the transposed layers (see the top of this file): lane (r, g) holds
//   w1[j]      A of layer 1, tile j:           element t = W1[8g + t][16j + r]  (zero for g >= 2: features 16 .. 31)
//   w2[j][kb]  A of layer 2, tile j, chunk kb: element t = W2[16 (2kb + t/4) + 4g + t%4][16j + r]
//   b1, b2, w3 [tile][i]: the values of hidden unit 16 tile + 4g + i, the rows of the lane's results
struct Frags {
    bf16x8 w1[4];
    bf16x8 w2[4][2];
    f32x4 b1[4], b2[4], w3[4];
    float b3;
    __device__ bf16x8 W1(int j) const { return w1[j]; }
    __device__ bf16x8 W2(int j, int kb) const { return w2[j][kb]; }
    __device__ f32x4 B1(int j) const { return b1[j]; }
    __device__ f32x4 B2(int j) const { return b2[j]; }
    __device__ f32x4 W3(int j) const { return w3[j]; }
    __device__ float B3() const { return b3; }
};

// lane l's layer-1 A fragment of output tile j, layer-2 A fragment of output tile j and K chunk kb, and the biases and
// output weights of the rows 16j + 4g .. 16j + 4g + 3 it receives; p: the FP32 parameter vector (global or staged)
template <class Get>
__device__ void fragmentsOf(Get w, int l, int j, bf16x8& w1, bf16x8 w2[2], f32x4& b1, f32x4& b2, f32x4& w3) {
    const int r = l & 15, g = l >> 4;
    for (int t = 0; t < 8; ++t) w1[t] = bfBits(g < 2 ? toBf16(w(kScreenW1 + (8 * g + t) * H + 16 * j + r)) : (short)0);
    for (int kb = 0; kb < 2; ++kb)
        for (int t = 0; t < 8; ++t) w2[kb][t] = bfBits(toBf16(w(kScreenW2 + (16 * (2 * kb + t / 4) + 4 * g + t % 4) * H + 16 * j + r)));
    for (int i = 0; i < 4; ++i) {
        b1[i] = w(kScreenB1 + 16 * j + 4 * g + i);
        b2[i] = w(kScreenB2 + 16 * j + 4 * g + i);
        w3[i] = w(kScreenW3 + 16 * j + 4 * g + i);
    }
}

__device__ void loadFrags(const Staged& s, Frags& F) {
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    for (int j = 0; j < 4; ++j) {
        for (int t = 0; t < 8; ++t) F.w1[j][t] = bfBits(g < 2 ? s.w1[(8 * g + t) * H + 16 * j + r] : (short)0);
        for (int kb = 0; kb < 2; ++kb)
            for (int t = 0; t < 8; ++t) F.w2[j][kb][t] = bfBits(s.w2[(16 * (2 * kb + t / 4) + 4 * g + t % 4) * H + 16 * j + r]);
        for (int i = 0; i < 4; ++i) {
            F.b1[j][i] = s.b1[16 * j + 4 * g + i];
            F.b2[j][i] = s.b2[16 * j + 4 * g + i];
            F.w3[j][i] = s.w3[16 * j + 4 * g + i];
        }
    }
    F.b3 = s.b3;
}

// The fragments of every lane in LDS, lane-minor (each lane's 16-byte fragment is one conflict-free read), for
// kernels that score few tiles between long stretches of other work and so do not hold them in registers.
struct FragTable {
    bf16x8 w1[4][64];
    bf16x8 w2[4][2][64];
    f32x4 b1[4][64], b2[4][64], w3[4][64];
    float b3;
};

// wavefront j of the workgroup (four wavefronts) writes output tile j's fragments of all 64 lanes
__device__ void buildFragTable(const float* p, FragTable& T) {
    const int l = threadIdx.x & 63, j = threadIdx.x >> 6;
    bf16x8 w2[2];
    fragmentsOf([p](int k) { return p[k]; }, l, j, T.w1[j][l], w2, T.b1[j][l], T.b2[j][l], T.w3[j][l]);
    T.w2[j][0][l] = w2[0];
    T.w2[j][1][l] = w2[1];
    if (threadIdx.x == 0) T.b3 = p[kScreenB3];
}

struct LdsFrags {
    const FragTable& T;
    int l;
    __device__ bf16x8 W1(int j) const { return T.w1[j][l]; }
    __device__ bf16x8 W2(int j, int kb) const { return T.w2[j][kb][l]; }
    __device__ f32x4 B1(int j) const { return T.b1[j][l]; }
    __device__ f32x4 B2(int j) const { return T.b2[j][l]; }
    __device__ f32x4 W3(int j) const { return T.w3[j][l]; }
    __device__ float B3() const { return T.b3; }
};

// Logit of sample r of a 16-sample tile whose features lane (r, g) holds in x (features 8g .. 8g+7, zero for g >= 2).
// Every lane of column r returns it. Operands are rounded to bfloat16 as in the layer-by-layer form: features,
// weights and the layer-1 activations; products accumulate in FP32. F: the lane's fragments (Frags or LdsFrags).
template <class FragSource>
__device__ inline float mlpTile(const FragSource& F, bf16x8 x) {
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    f32x4 h1[4];
    for (int j = 0; j < 4; ++j) {
        h1[j] = mfma32(F.W1(j), x, zero);
        const f32x4 b1 = F.B1(j);
        for (int i = 0; i < 4; ++i) h1[j][i] = fmaxf(h1[j][i] + b1[i], 0.f);
    }
    f32x4 h2[4] = {zero, zero, zero, zero};
    for (int kb = 0; kb < 2; ++kb) {
        bf16x8 b;
        for (int t = 0; t < 4; ++t) {
            b[t] = bf(h1[2 * kb][t]);
            b[4 + t] = bf(h1[2 * kb + 1][t]);
        }
        for (int j = 0; j < 4; ++j) h2[j] = mfma32(F.W2(j, kb), b, h2[j]);
    }
    float z = 0.f;
    for (int j = 0; j < 4; ++j) {
        const f32x4 b2 = F.B2(j), w3 = F.W3(j);
        for (int i = 0; i < 4; ++i) z += fmaxf(h2[j][i] + b2[i], 0.f) * w3[i];
    }
    z += __shfl_xor(z, 16, 64);
    z += __shfl_xor(z, 32, 64);
    return z + F.B3();
}

// Logits of n feature rows. Persistent wavefronts: one barrier to stage the weights, then each wavefront loads its
// fragments once and takes 16-row tiles independently (tile t by wavefront t mod the wavefront count).
__global__ __launch_bounds__(kScreenThreads) void screenForwardKernel(const float* params, const float* X, int n, const int* nDev,
                                                                     float* logits) {
    __shared__ Staged s;
    if (nDev) n = min(n, *nDev);
    stageWeights(params, s);
    __syncthreads();
    Frags F;
    loadFrags(s, F);
    const int l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    const int waves = gridDim.x * (kScreenThreads / 64);
    const int tiles = (n + 15) / 16;
    for (int t = blockIdx.x * (kScreenThreads / 64) + (threadIdx.x >> 6); t < tiles; t += waves) {
        const int row = 16 * t + r;
        bf16x8 x;
        if (g < 2 && row < n) {
            const float4* p = reinterpret_cast<const float4*>(X + (size_t)row * kScreenIn + 8 * g);
            const float4 a = p[0], b = p[1];
            x = bf16x8{bf(a.x), bf(a.y), bf(a.z), bf(a.w), bf(b.x), bf(b.y), bf(b.z), bf(b.w)};
        } else {
            for (int k = 0; k < 8; ++k) x[k] = bfBits(0);
        }
        const float z = mlpTile(F, x);
        if (g == 0 && row < n) logits[row] = z;
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

// fixed-order sum of the workgroups' partial gradients: a workgroup per 64 parameters, four lanes per parameter summing
// every fourth partial row with several loads in flight, then the four in a fixed order
__global__ __launch_bounds__(256) void screenReduceKernel(const float* partial, int blocks, float* grad) {
    __shared__ float part[4][64];
    const int pl = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + pl;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    if (p <= kScreenParams) {
        int b = q;
        for (; b + 12 < blocks; b += 16)
            for (int u = 0; u < 4; ++u) acc[u] += partial[(size_t)(b + 4 * u) * (kScreenParams + 1) + p];
        for (; b < blocks; b += 4) acc[0] += partial[(size_t)b * (kScreenParams + 1) + p];
    }
    part[q][pl] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    __syncthreads();
    if (q == 0 && p <= kScreenParams) grad[p] = (part[0][pl] + part[1][pl]) + (part[2][pl] + part[3][pl]);
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

// the cutoff of source i: its group's range, or the system's
__device__ inline double sourceCutoff(const ScreenFeatureParams& P, int i) { return P.srcGroup ? P.groupRange[P.srcGroup[i]] : P.cutoff; }

// the 16 features of the pair (source i, target j) with cutoff rc
__device__ void pairFeatures(const ScreenFeatureParams& P, int i, int j, double rc, float* o) {
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
    const double riseI = d[0] * ni[0] + d[1] * ni[1] + d[2] * ni[2];  // rise of the chord above the source's plane
    const double riseJ = d[0] * nj[0] + d[1] * nj[1] + d[2] * nj[2];
    const double dd = dE > 0 ? dE : 1e-300;
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

// the first feature alone: 100 (1 - |x_j - x_i| / rc), as pairFeatures computes it
__device__ inline float cutoffMargin(const ScreenFeatureParams& P, int i, int j, double rc) {
    const double d0 = P.xyz[3 * j] - P.xyz[3 * i], d1 = P.xyz[3 * j + 1] - P.xyz[3 * i + 1], d2 = P.xyz[3 * j + 2] - P.xyz[3 * i + 2];
    return (float)(100.0 * (1.0 - sqrt(d0 * d0 + d1 * d1 + d2 * d2) / rc));
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
    float o[kScreenIn];
    pairFeatures(P, lo, P.tidx[q], sourceCutoff(P, lo), o);
    float4* out = reinterpret_cast<float4*>(P.out + (size_t)q * kScreenIn);
    for (int k = 0; k < 4; ++k) out[k] = float4{o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]};
}

constexpr int kBandRing = 128;  // band pairs a wavefront holds before scoring: fewer than 16 + one 64-candidate chunk

// Screening of every source's candidates, one wavefront per source in turn (persistent): candidates closer than the
// guard distance are kept unscored; the others (the guard band) queue in the wavefront's LDS ring and are scored 16 at
// a time by the MLP as soon as 16 are waiting (the last partial tile at the end). No global list and no per-pair
// atomics: the band and kept counts reach counts[0] and counts[1] once per workgroup.
// Unconstrained it takes 208 VGPRs (two wavefronts per SIMD); capping it at three or four spills in the scoring code
// and measured 37% slower (0.057 against 0.042 ms for 20,000 sources and 1.6 M candidates).
__global__ __launch_bounds__(kScreenThreads) void screenSourcesKernel(ScreenFeatureParams P, const float* params, float threshold, uint8_t* keep, int* counts) {
    __shared__ FragTable T;
    __shared__ int2 ring[kScreenThreads / 64][kBandRing];
    __shared__ int totals[2][kScreenThreads / 64];
    buildFragTable(params, T);
    __syncthreads();
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, r = l & 15, g = l >> 4;
    const LdsFrags F{T, l};
    int2* Q = ring[w];
    const unsigned long long below = (1ull << l) - 1ull;
    unsigned head = 0, tail = 0;  // the ring's wavefront-uniform read and write positions
    int band = 0, kept = 0;
    auto scoreTile = [&](int n) {  // the n <= 16 oldest ring entries
        const bool mine = r < n;
        int2 e = make_int2(0, 0);
        bf16x8 x;
        for (int k = 0; k < 8; ++k) x[k] = bfBits(0);
        if (mine) e = Q[(head + r) & (kBandRing - 1)];
        if (mine && g < 2) {  // two lanes per pair, each computing the features and keeping its half
            float o[kScreenIn];
            pairFeatures(P, e.y, P.tidx[e.x], sourceCutoff(P, e.y), o);
            for (int k = 0; k < 8; ++k) x[k] = bf(g ? o[8 + k] : o[k]);
        }
        const float z = mlpTile(F, x);
        const bool keepIt = mine && g == 0 && z >= threshold;
        if (mine && g == 0) keep[e.x] = keepIt ? 1 : 0;
        kept += __popcll(__ballot(keepIt));
        head += n;
    };
    const int waves = gridDim.x * (kScreenThreads / 64);
    for (int i = blockIdx.x * (kScreenThreads / 64) + w; i < P.nSrc; i += waves) {
        const double rc = sourceCutoff(P, i);
        const int b0 = P.toff[i], b1 = P.toff[i + 1];
        for (int c0 = b0; c0 < b1; c0 += 64) {
            const int q = c0 + l;
            const bool valid = q < b1;
            const bool inBand = valid && cutoffMargin(P, i, P.tidx[q], rc) <= kScreenGuard;
            if (valid && !inBand) keep[q] = 1;
            const unsigned long long m = __ballot(inBand);
            if (inBand) Q[(tail + __popcll(m & below)) & (kBandRing - 1)] = make_int2(q, i);
            tail += __popcll(m);
            band += __popcll(m);
            kept += __popcll(__ballot(valid && !inBand));
            // the ring's entries were written by other lanes of this wavefront
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            while (tail - head >= 16) scoreTile(16);
        }
    }
    if (tail != head) scoreTile((int)(tail - head));
    if (l == 0) {
        totals[0][w] = band;
        totals[1][w] = kept;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        int t = 0;
        for (int k = 0; k < kScreenThreads / 64; ++k) t += totals[threadIdx.x][k];
        if (t) atomicAdd(counts + threadIdx.x, t);
    }
}

}  // namespace

int screenTrainBlocks(int n) { return n <= 0 ? 1 : (n + 63) / 64 < 1024 ? (n + 63) / 64 : 1024; }

hipError_t launchScreenForward(const float* params, const float* X, int n, const int* nDev, float* logits, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    const long tiles = ((long)n + 15) / 16;
    const int blocks = (int)std::min<long>((tiles + 3) / 4, kScreenPersistentBlocks);
    hipLaunchKernelGGL(screenForwardKernel, dim3(blocks), dim3(kScreenThreads), 0, stream, params, X, n, nDev, logits);
    return hipGetLastError();
}

hipError_t launchScreenSources(const ScreenFeatureParams& P, const float* params, float threshold, uint8_t* keep, int* counts,
                               hipStream_t stream) {
    if (P.nSrc <= 0) return hipSuccess;
    // as many workgroups as are resident at once, each wavefront looping over sources
    static int resident = 0;
    if (!resident) {
        int perCU = 0, dev = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, screenSourcesKernel, kScreenThreads, 0) != hipSuccess || perCU < 1)
            perCU = 1;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus < 1)
            cus = 256;
        resident = perCU * cus;
    }
    const int blocks = std::min(resident, (P.nSrc + kScreenThreads / 64 - 1) / (kScreenThreads / 64));
    hipLaunchKernelGGL(screenSourcesKernel, dim3(blocks), dim3(kScreenThreads), 0, stream, P, params, threshold, keep, counts);
    return hipGetLastError();
}

hipError_t launchScreenGradient(const float* params, const float* X, const float* y, int n, float* partial, float* grad,
                                hipStream_t stream) {
    if (n <= 0) return hipErrorInvalidValue;
    const int blocks = screenTrainBlocks(n);
    hipLaunchKernelGGL(screenGradientKernel, dim3(blocks), dim3(kScreenThreads), 0, stream, params, X, y, n, partial);
    hipLaunchKernelGGL(screenReduceKernel, dim3((kScreenParams + 64) / 64), dim3(256), 0, stream, partial, blocks, grad);
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
