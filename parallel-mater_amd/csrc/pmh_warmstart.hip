// SPDX-License-Identifier: MIT
//
// Learned warm start (no counterpart in the reference, whose cold contacts
// start from lambda = 0 and the pair stiffness, geometry_constraints.slang:
// 1366-1373): a 16 -> 64 -> 64 -> 2 perceptron with ReLU, BF16 weights and
// activations on MFMA (v_mfma_f32_16x16x16_bf16) with FP32 accumulation,
// predicting the normal row's dual (lambda, penalty) of each cold contact
// point from 16 features computed at contact preparation.  Trained on the
// GPU by the kernels below from contact samples that plain AVBD logs.
//
// Every product is a 16x16 MFMA tile C += A(16 x K) B(K x 16) whose operands
// are read from LDS through strides, so transposed operands need no copies.
// Lane l supplies A[l % 16][k0 + 4(l / 16) + j] and B[k0 + 4(l / 16) + j][l % 16]
// (j = 0..3) and receives C[4(l / 16) + e][l % 16] (e = 0..3).

#include "pmh_internal.hpp"
#include "pmh_warmstart.hpp"

#include <cmath>
#include <vector>

namespace pmh {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t to_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return uint16_t((u >> 16) | 0x40u);   // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                           // round to nearest even
    return uint16_t(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
    const uint32_t u = uint32_t(h) << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ f32x4 mma_tile(const uint16_t *A, int sai, int sak, const uint16_t *B, int sbk, int sbj,
                                          int K, f32x4 acc) {
    const int l = threadIdx.x & 63, r = l & 15, q = (l >> 4) * 4;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            a[j] = short(A[r * sai + (k0 + q + j) * sak]);
            b[j] = short(B[(k0 + q + j) * sbk + r * sbj]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// Operands of one 64-sample pass, bf16 in LDS.
struct MlpShared {
    uint16_t w1[kMlpHidden * kMlpIn];          // [out][in]
    uint16_t w2[kMlpHidden * kMlpHidden];      // [out][in]
    uint16_t w3[16 * kMlpHidden];              // [out][in], rows 2..15 zero
    float b1[kMlpHidden], b2[kMlpHidden], b3[16];
    uint16_t x[64 * kMlpIn];                   // [sample][feature]
    uint16_t h1[64 * kMlpHidden], h2[64 * kMlpHidden];
    uint16_t dy[64 * 16];                      // [sample][out], cols 2..15 zero
    uint16_t dz2[64 * kMlpHidden], dz1[64 * kMlpHidden];
    float y[64 * kMlpOut], t[64 * kMlpOut];
};

__device__ void load_weights(MlpShared &S, const float *p) {
    const int l = threadIdx.x;
    for (int k = l; k < kMlpHidden * kMlpIn; k += 64) S.w1[k] = to_bf16(p[kMlpW1 + k]);
    for (int k = l; k < kMlpHidden * kMlpHidden; k += 64) S.w2[k] = to_bf16(p[kMlpW2 + k]);
    for (int k = l; k < 16 * kMlpHidden; k += 64) S.w3[k] = k < kMlpOut * kMlpHidden ? to_bf16(p[kMlpW3 + k]) : 0;
    for (int k = l; k < kMlpHidden; k += 64) {
        S.b1[k] = p[kMlpB1 + k];
        S.b2[k] = p[kMlpB2 + k];
    }
    if (l < 16) S.b3[l] = l < kMlpOut ? p[kMlpB3 + l] : 0.0f;
}

// Forward pass of the 64 samples in S.x: h1, h2 (bf16, after ReLU) and y.
__device__ void forward(MlpShared &S) {
    const int l = threadIdx.x, r = l & 15, q = (l >> 4) * 4;
    for (int mi = 0; mi < 4; ++mi)
        for (int nj = 0; nj < 4; ++nj) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mma_tile(S.x + mi * 16 * kMlpIn, kMlpIn, 1, S.w1 + nj * 16 * kMlpIn, 1, kMlpIn, kMlpIn, acc);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int s = mi * 16 + q + e, o = nj * 16 + r;
                S.h1[s * kMlpHidden + o] = to_bf16(fmaxf(acc[e] + S.b1[o], 0.0f));
            }
        }
    __syncthreads();
    for (int mi = 0; mi < 4; ++mi)
        for (int nj = 0; nj < 4; ++nj) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mma_tile(S.h1 + mi * 16 * kMlpHidden, kMlpHidden, 1, S.w2 + nj * 16 * kMlpHidden, 1, kMlpHidden,
                           kMlpHidden, acc);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int s = mi * 16 + q + e, o = nj * 16 + r;
                S.h2[s * kMlpHidden + o] = to_bf16(fmaxf(acc[e] + S.b2[o], 0.0f));
            }
        }
    __syncthreads();
    for (int mi = 0; mi < 4; ++mi) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma_tile(S.h2 + mi * 16 * kMlpHidden, kMlpHidden, 1, S.w3, 1, kMlpHidden, kMlpHidden, acc);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int s = mi * 16 + q + e;
            if (r < kMlpOut) S.y[s * kMlpOut + r] = acc[e] + S.b3[r];
        }
    }
    __syncthreads();
}

// Mean-squared-error gradients of the 64 samples (weights of `count` of
// them, the rest padding) into g[kMlpParams]; returns the summed squared
// error.  Bias gradients are FP32 sums, in sample order, of the bf16 deltas.
__device__ float backward(MlpShared &S, uint32_t count, float *g) {
    const int l = threadIdx.x, r = l & 15, q = (l >> 4) * 4;
    const float scale = 2.0f / float(count * kMlpOut);
    float sq = 0.0f;
    {   // dY, one sample per lane
        const int s = l;
        for (int o = 0; o < 16; ++o) {
            float d = 0.0f;
            if (o < kMlpOut && uint32_t(s) < count) {
                const float e = S.y[s * kMlpOut + o] - S.t[s * kMlpOut + o];
                sq += e * e;
                d = scale * e;
            }
            S.dy[s * 16 + o] = to_bf16(d);
        }
    }
    __syncthreads();
    // dW3 = dY^T H2, db3
    for (int nj = 0; nj < 4; ++nj) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma_tile(S.dy, 1, 16, S.h2 + nj * 16, kMlpHidden, 1, 64, acc);
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (q + e < kMlpOut) g[kMlpW3 + (q + e) * kMlpHidden + nj * 16 + r] = acc[e];
    }
    if (l < kMlpOut) {
        float b = 0.0f;
        for (int s = 0; s < 64; ++s) b += from_bf16(S.dy[s * 16 + l]);
        g[kMlpB3 + l] = b;
    }
    // dZ2 = (dY W3) .* [H2 > 0]
    for (int mi = 0; mi < 4; ++mi)
        for (int nj = 0; nj < 4; ++nj) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mma_tile(S.dy + mi * 16 * 16, 16, 1, S.w3 + nj * 16, kMlpHidden, 1, 16, acc);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int s = mi * 16 + q + e, o = nj * 16 + r;
                S.dz2[s * kMlpHidden + o] = from_bf16(S.h2[s * kMlpHidden + o]) > 0.0f ? to_bf16(acc[e]) : 0;
            }
        }
    __syncthreads();
    // dW2 = dZ2^T H1, db2
    for (int mi = 0; mi < 4; ++mi)
        for (int nj = 0; nj < 4; ++nj) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mma_tile(S.dz2 + mi * 16, 1, kMlpHidden, S.h1 + nj * 16, kMlpHidden, 1, 64, acc);
#pragma unroll
            for (int e = 0; e < 4; ++e) g[kMlpW2 + (mi * 16 + q + e) * kMlpHidden + nj * 16 + r] = acc[e];
        }
    {
        float b = 0.0f;
        for (int s = 0; s < 64; ++s) b += from_bf16(S.dz2[s * kMlpHidden + l]);
        g[kMlpB2 + l] = b;
    }
    // dZ1 = (dZ2 W2) .* [H1 > 0]
    for (int mi = 0; mi < 4; ++mi)
        for (int nj = 0; nj < 4; ++nj) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mma_tile(S.dz2 + mi * 16 * kMlpHidden, kMlpHidden, 1, S.w2 + nj * 16, kMlpHidden, 1, kMlpHidden,
                           acc);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int s = mi * 16 + q + e, o = nj * 16 + r;
                S.dz1[s * kMlpHidden + o] = from_bf16(S.h1[s * kMlpHidden + o]) > 0.0f ? to_bf16(acc[e]) : 0;
            }
        }
    __syncthreads();
    // dW1 = dZ1^T X, db1
    for (int mi = 0; mi < 4; ++mi) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma_tile(S.dz1 + mi * 16, 1, kMlpHidden, S.x, kMlpIn, 1, 64, acc);
#pragma unroll
        for (int e = 0; e < 4; ++e) g[kMlpW1 + (mi * 16 + q + e) * kMlpIn + r] = acc[e];
    }
    {
        float b = 0.0f;
        for (int s = 0; s < 64; ++s) b += from_bf16(S.dz1[s * kMlpHidden + l]);
        g[kMlpB1 + l] = b;
    }
    // summed squared error over the wave
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) sq += __shfl_xor(sq, d, 64);
    return sq;
}

// Samples [first, first + count) (count <= 64) into S.x / S.t.
__device__ void load_samples(MlpShared &S, const float *samples, uint32_t first, uint32_t count) {
    const int l = threadIdx.x;
    const bool live = uint32_t(l) < count;
    const float *row = samples + size_t(first + (live ? l : 0)) * kMlpSample;
    for (int f = 0; f < kMlpIn; ++f) S.x[l * kMlpIn + f] = live ? to_bf16(row[f]) : 0;
    for (int o = 0; o < kMlpOut; ++o) S.t[l * kMlpOut + o] = live ? row[kMlpIn + o] : 0.0f;
    __syncthreads();
}

// Gradients of one minibatch (no update).
__global__ void __launch_bounds__(64) k_mlp_gradients(const float *samples, uint32_t count, const float *params,
                                                      float *grads, float *loss) {
    __shared__ MlpShared S;
    load_weights(S, params);
    load_samples(S, samples, 0, count);
    forward(S);
    const float sq = backward(S, count, grads);
    if (threadIdx.x == 0) *loss = sq / float(count * kMlpOut);
}

// Adam (beta 0.9, 0.999) over minibatches of 64 in the order given by
// `order`, for `steps` steps; one wavefront.  loss[e] receives the mean
// loss of each epoch.
__global__ void __launch_bounds__(64) k_mlp_train(const float *samples, uint32_t n, const uint32_t *order,
                                                  uint32_t epochs, float lr, float *params, float *m, float *v,
                                                  float *grads, float *loss, float *staged) {
    __shared__ MlpShared S;
    const int l = threadIdx.x;
    const uint32_t batches = (n + 63) / 64;
    uint32_t t = 0;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        float total = 0.0f;
        for (uint32_t bi = 0; bi < batches; ++bi) {
            const uint32_t first = bi * 64, count = min(64u, n - first);
            // gather the batch through the permutation into a staging block
            for (uint32_t k = l; k < count * kMlpSample; k += 64) {
                const uint32_t s = k / kMlpSample, f = k % kMlpSample;
                staged[k] = samples[size_t(order[ep * n + first + s]) * kMlpSample + f];
            }
            __syncthreads();
            load_weights(S, params);
            load_samples(S, staged, 0, count);
            forward(S);
            total += backward(S, count, grads);
            __syncthreads();
            ++t;
            const float c1 = 1.0f - powf(0.9f, float(t)), c2 = 1.0f - powf(0.999f, float(t));
            for (int k = l; k < kMlpBound; k += 64) {
                const float gk = grads[k];
                m[k] = 0.9f * m[k] + 0.1f * gk;
                v[k] = 0.999f * v[k] + 0.001f * gk * gk;
                params[k] -= lr * (m[k] / c1) / (sqrtf(v[k] / c2) + 1.0e-8f);
            }
            __syncthreads();
        }
        if (l == 0) loss[ep] = total / float(n * kMlpOut);
    }
}

// Inference for listed cold points: y0 = log10(1 + lambda / (m_eff |g|)),
// y1 = log10(k / k0); lambda >= 0, k0 <= k <= the penalty cap.
__global__ void __launch_bounds__(64) k_mlp_infer(const uint32_t *count_ptr, const uint32_t *list, const float *feat,
                                                  const float *scale, const float *params, float4 *c_lam,
                                                  float4 *c_pen, uint32_t cap, uint32_t policy) {
    __shared__ MlpShared S;
    load_weights(S, params);
    const uint32_t total = min(*count_ptr, cap);
    const int l = threadIdx.x;
    for (uint32_t first = blockIdx.x * 64; first < total; first += gridDim.x * 64) {
        const uint32_t count = min(64u, total - first);
        const bool live = uint32_t(l) < count;
        for (int f = 0; f < kMlpIn; ++f) S.x[l * kMlpIn + f] = live ? to_bf16(feat[size_t(first + l) * kMlpIn + f]) : 0;
        __syncthreads();
        forward(S);
        if (live) {
            const uint32_t c = list[first + l];
            const float lam_scale = scale[2 * size_t(first + l)], k0 = scale[2 * size_t(first + l) + 1];
            // predictions stay within the range of the training targets
            const float y0 = fminf(fmaxf(S.y[l * kMlpOut], 0.0f), params[kMlpBound]);
            const float y1 = fminf(fmaxf(S.y[l * kMlpOut + 1], 0.0f), params[kMlpBound + 1]);
            const float lam = (exp10f(y0) - 1.0f) * lam_scale;
            const float k = fminf(k0 * exp10f(y1), kMaxPenalty);
            // The penalty always; lambda only where the contact starts in
            // violation (initial normal error > 0, feature 5), i.e. where it
            // is active from the first iteration anyway.
            if (policy & kWarmPenalty) {
                float4 b = c_pen[c];
                b.x = fmaxf(k, k0);
                c_pen[c] = b;
            }
            if ((policy & kWarmLambda) && feat[size_t(first + l) * kMlpIn + 5] > 0.0f) {
                float4 a = c_lam[c];
                a.x = lam;
                c_lam[c] = a;
            }
        }
        __syncthreads();
    }
}

PmhResult mlp_infer(hipStream_t s, const uint32_t *count, const uint32_t *list, const float *feat,
                    const float *scale, const float *params, float4 *c_lam, float4 *c_pen, uint32_t cap,
                    uint32_t policy) {
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((cap + 63) / 64, 1024));
    k_mlp_infer<<<blocks, 64, 0, s>>>(count, list, feat, scale, params, c_lam, c_pen, cap, policy);
    PMH_CHECK(hipGetLastError());
    return PMH_SUCCESS;
}

// Forward pass of `n` feature rows (16 floats each) into y (2 floats each).
__global__ void __launch_bounds__(64) k_mlp_forward(const float *x, uint32_t n, const float *params, float *y) {
    __shared__ MlpShared S;
    load_weights(S, params);
    const int l = threadIdx.x;
    for (uint32_t first = blockIdx.x * 64; first < n; first += gridDim.x * 64) {
        const uint32_t count = min(64u, n - first);
        const bool live = uint32_t(l) < count;
        for (int f = 0; f < kMlpIn; ++f) S.x[l * kMlpIn + f] = live ? to_bf16(x[size_t(first + l) * kMlpIn + f]) : 0;
        __syncthreads();
        forward(S);
        if (live)
            for (int o = 0; o < kMlpOut; ++o) y[size_t(first + l) * kMlpOut + o] = S.y[l * kMlpOut + o];
        __syncthreads();
    }
}

}  // namespace pmh

using namespace pmh;

extern "C" {

void pmhWarmstartInit(uint32_t seed, float *params) {
    // He-uniform weights, zero biases, from a 32-bit LCG.
    uint32_t st = seed * 2654435761u + 12345u;
    auto uni = [&]() { st = st * 1664525u + 1013904223u; return (st >> 8) * (1.0f / 16777216.0f); };
    for (int k = 0; k < kMlpParams; ++k) params[k] = 0.0f;
    const float a1 = std::sqrt(6.0f / kMlpIn), a2 = std::sqrt(6.0f / kMlpHidden), a3 = std::sqrt(3.0f / kMlpHidden);
    for (int k = 0; k < kMlpHidden * kMlpIn; ++k) params[kMlpW1 + k] = (2.0f * uni() - 1.0f) * a1;
    for (int k = 0; k < kMlpHidden * kMlpHidden; ++k) params[kMlpW2 + k] = (2.0f * uni() - 1.0f) * a2;
    for (int k = 0; k < kMlpOut * kMlpHidden; ++k) params[kMlpW3 + k] = (2.0f * uni() - 1.0f) * a3;
    params[kMlpBound] = params[kMlpBound + 1] = 0.0f;
}

PmhResult pmhWarmstartForward(const float *params, const float *features, uint32_t n, float *outputs) {
    if (!params || !features || !outputs || n == 0) return PMH_ERROR_INVALID_ARGUMENT;
    DBuf<float> dx, dp, dy;
    PmhResult r;
    if ((r = dx.reserve(size_t(n) * kMlpIn)) || (r = dp.reserve(kMlpParams)) || (r = dy.reserve(size_t(n) * kMlpOut)))
        return r;
    PMH_CHECK(hipMemcpy(dx.p, features, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice));
    PMH_CHECK(hipMemcpy(dp.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    k_mlp_forward<<<std::max<uint32_t>(1, std::min<uint32_t>((n + 63) / 64, 1024)), 64>>>(dx.p, n, dp.p, dy.p);
    PMH_CHECK(hipGetLastError());
    PMH_CHECK(hipMemcpy(outputs, dy.p, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost));
    dx.release(); dp.release(); dy.release();
    return PMH_SUCCESS;
}

PmhResult pmhWarmstartGradients(const float *samples, uint32_t count, const float *params, float *grads,
                                float *loss) {
    if (!samples || !params || !grads || !loss || count == 0 || count > 64) return PMH_ERROR_INVALID_ARGUMENT;
    DBuf<float> ds, dp, dg, dl;
    PmhResult r;
    if ((r = ds.reserve(size_t(count) * kMlpSample)) || (r = dp.reserve(kMlpParams)) || (r = dg.reserve(kMlpParams)) ||
        (r = dl.reserve(1)))
        return r;
    PMH_CHECK(hipMemcpy(ds.p, samples, sizeof(float) * count * kMlpSample, hipMemcpyHostToDevice));
    PMH_CHECK(hipMemcpy(dp.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    PMH_CHECK(hipMemset(dg.p, 0, sizeof(float) * kMlpParams));
    k_mlp_gradients<<<1, 64>>>(ds.p, count, dp.p, dg.p, dl.p);
    PMH_CHECK(hipGetLastError());
    PMH_CHECK(hipMemcpy(grads, dg.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    PMH_CHECK(hipMemcpy(loss, dl.p, sizeof(float), hipMemcpyDeviceToHost));
    ds.release(); dp.release(); dg.release(); dl.release();
    return PMH_SUCCESS;
}

PmhResult pmhWarmstartTrain(const float *samples, uint32_t n, uint32_t epochs, float learning_rate, uint32_t seed,
                            float *params, float *epoch_loss) {
    if (!samples || !params || !epoch_loss || n == 0 || epochs == 0 || !(learning_rate > 0.0f))
        return PMH_ERROR_INVALID_ARGUMENT;
    // Epoch permutations from the seed (Fisher-Yates with an LCG).
    std::vector<uint32_t> order(size_t(epochs) * n);
    uint32_t st = seed * 747796405u + 2891336453u;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        uint32_t *o = order.data() + size_t(ep) * n;
        for (uint32_t k = 0; k < n; ++k) o[k] = k;
        for (uint32_t k = n - 1; k > 0; --k) {
            st = st * 1664525u + 1013904223u;
            const uint32_t j = uint32_t((uint64_t(st) * (k + 1)) >> 32);
            const uint32_t t = o[k];
            o[k] = o[j];
            o[j] = t;
        }
    }
    DBuf<float> ds, dp, dm, dv, dg, dl, dst;
    DBuf<uint32_t> dorder;
    PmhResult r;
    if ((r = ds.reserve(size_t(n) * kMlpSample)) || (r = dp.reserve(kMlpParams)) || (r = dm.reserve(kMlpParams)) ||
        (r = dv.reserve(kMlpParams)) || (r = dg.reserve(kMlpParams)) || (r = dl.reserve(epochs)) ||
        (r = dst.reserve(64 * kMlpSample)) || (r = dorder.reserve(order.size())))
        return r;
    PMH_CHECK(hipMemcpy(ds.p, samples, sizeof(float) * size_t(n) * kMlpSample, hipMemcpyHostToDevice));
    PMH_CHECK(hipMemcpy(dp.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    PMH_CHECK(hipMemset(dm.p, 0, sizeof(float) * kMlpParams));
    PMH_CHECK(hipMemset(dv.p, 0, sizeof(float) * kMlpParams));
    PMH_CHECK(hipMemset(dg.p, 0, sizeof(float) * kMlpParams));
    PMH_CHECK(hipMemcpy(dorder.p, order.data(), sizeof(uint32_t) * order.size(), hipMemcpyHostToDevice));
    k_mlp_train<<<1, 64>>>(ds.p, n, dorder.p, epochs, learning_rate, dp.p, dm.p, dv.p, dg.p, dl.p, dst.p);
    PMH_CHECK(hipGetLastError());
    PMH_CHECK(hipMemcpy(params, dp.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
    PMH_CHECK(hipMemcpy(epoch_loss, dl.p, sizeof(float) * epochs, hipMemcpyDeviceToHost));
    for (int o = 0; o < kMlpOut; ++o) {
        float hi = 0.0f;
        for (uint32_t k = 0; k < n; ++k) hi = std::max(hi, samples[size_t(k) * kMlpSample + kMlpIn + o]);
        params[kMlpBound + o] = hi;
    }
    ds.release(); dp.release(); dm.release(); dv.release(); dg.release(); dl.release(); dst.release(); dorder.release();
    return PMH_SUCCESS;
}

}  // extern "C"
