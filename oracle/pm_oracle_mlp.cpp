// SPDX-License-Identifier: MIT
//
// Test infrastructure: the CPU restatement of the learned warm start's
// network (16 -> 64 -> 64 -> 2, ReLU) that checks the HIP forward and
// gradient kernels.  Same rounding points as the device path: inputs,
// weights, activations and back-propagated deltas are rounded to bf16
// (nearest even); every product is accumulated in FP32, here in index
// order (the MFMA unit sums each 16-term block in its own order, hence the
// tolerance of the comparison).  No counterpart in the reference; pinned by
// finite differences of the FP32 network in the tests.
#include "pm_oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kIn = 16, kHid = 64, kOut = 2, kSample = kIn + kOut;
constexpr int kW1 = 0, kB1 = kW1 + kHid * kIn, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid, kW3 = kB2 + kHid,
              kB3 = kW3 + kOut * kHid, kParams = kB3 + kOut;

float bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return f;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

struct Pass {
    std::vector<float> x, h1, h2, y;   // bf16-rounded x, h1, h2; FP32 y
};

void forward(const float *p, const float *x, int n, Pass &P, bool round) {
    auto R = [&](float v) { return round ? bf(v) : v; };
    P.x.assign(size_t(n) * kIn, 0.0f);
    P.h1.assign(size_t(n) * kHid, 0.0f);
    P.h2.assign(size_t(n) * kHid, 0.0f);
    P.y.assign(size_t(n) * kOut, 0.0f);
    for (int s = 0; s < n; ++s) {
        for (int f = 0; f < kIn; ++f) P.x[s * kIn + f] = R(x[s * kIn + f]);
        for (int o = 0; o < kHid; ++o) {
            float z = 0.0f;
            for (int f = 0; f < kIn; ++f) z += P.x[s * kIn + f] * R(p[kW1 + o * kIn + f]);
            P.h1[s * kHid + o] = R(std::fmax(z + p[kB1 + o], 0.0f));
        }
        for (int o = 0; o < kHid; ++o) {
            float z = 0.0f;
            for (int f = 0; f < kHid; ++f) z += P.h1[s * kHid + f] * R(p[kW2 + o * kHid + f]);
            P.h2[s * kHid + o] = R(std::fmax(z + p[kB2 + o], 0.0f));
        }
        for (int o = 0; o < kOut; ++o) {
            float z = 0.0f;
            for (int f = 0; f < kHid; ++f) z += P.h2[s * kHid + f] * R(p[kW3 + o * kHid + f]);
            P.y[s * kOut + o] = z + p[kB3 + o];
        }
    }
}

}  // namespace

extern "C" {

void pmo_mlp_forward(const float *params, const float *x, uint32_t n, float *y, int round_bf16) {
    Pass P;
    forward(params, x, int(n), P, round_bf16 != 0);
    std::memcpy(y, P.y.data(), sizeof(float) * n * kOut);
}

// Mean-squared-error loss and gradients of `count` samples (features then
// targets), the same formulas and rounding points as the device kernel.
float pmo_mlp_gradients(const float *params, const float *samples, uint32_t count, float *grads, int round_bf16) {
    const int n = int(count);
    auto R = [&](float v) { return round_bf16 ? bf(v) : v; };
    std::vector<float> x(size_t(n) * kIn);
    for (int s = 0; s < n; ++s)
        for (int f = 0; f < kIn; ++f) x[s * kIn + f] = samples[s * kSample + f];
    Pass P;
    forward(params, x.data(), n, P, round_bf16 != 0);
    std::memset(grads, 0, sizeof(float) * kParams);
    const float scale = 2.0f / float(n * kOut);
    float sq = 0.0f;
    std::vector<float> dy(size_t(n) * kOut), dz2(size_t(n) * kHid), dz1(size_t(n) * kHid);
    for (int s = 0; s < n; ++s)
        for (int o = 0; o < kOut; ++o) {
            const float e = P.y[s * kOut + o] - samples[s * kSample + kIn + o];
            sq += e * e;
            dy[s * kOut + o] = R(scale * e);
        }
    for (int o = 0; o < kOut; ++o)
        for (int j = 0; j < kHid; ++j) {
            float g = 0.0f;
            for (int s = 0; s < n; ++s) g += dy[s * kOut + o] * P.h2[s * kHid + j];
            grads[kW3 + o * kHid + j] = g;
        }
    for (int o = 0; o < kOut; ++o) {
        float b = 0.0f;
        for (int s = 0; s < n; ++s) b += dy[s * kOut + o];
        grads[kB3 + o] = b;
    }
    for (int s = 0; s < n; ++s)
        for (int j = 0; j < kHid; ++j) {
            float d = 0.0f;
            for (int o = 0; o < kOut; ++o) d += dy[s * kOut + o] * R(params[kW3 + o * kHid + j]);
            dz2[s * kHid + j] = P.h2[s * kHid + j] > 0.0f ? R(d) : 0.0f;
        }
    for (int o = 0; o < kHid; ++o)
        for (int j = 0; j < kHid; ++j) {
            float g = 0.0f;
            for (int s = 0; s < n; ++s) g += dz2[s * kHid + o] * P.h1[s * kHid + j];
            grads[kW2 + o * kHid + j] = g;
        }
    for (int o = 0; o < kHid; ++o) {
        float b = 0.0f;
        for (int s = 0; s < n; ++s) b += dz2[s * kHid + o];
        grads[kB2 + o] = b;
    }
    for (int s = 0; s < n; ++s)
        for (int j = 0; j < kHid; ++j) {
            float d = 0.0f;
            for (int o = 0; o < kHid; ++o) d += dz2[s * kHid + o] * R(params[kW2 + o * kHid + j]);
            dz1[s * kHid + j] = P.h1[s * kHid + j] > 0.0f ? R(d) : 0.0f;
        }
    for (int o = 0; o < kHid; ++o)
        for (int f = 0; f < kIn; ++f) {
            float g = 0.0f;
            for (int s = 0; s < n; ++s) g += dz1[s * kHid + o] * P.x[s * kIn + f];
            grads[kW1 + o * kIn + f] = g;
        }
    for (int o = 0; o < kHid; ++o) {
        float b = 0.0f;
        for (int s = 0; s < n; ++s) b += dz1[s * kHid + o];
        grads[kB1 + o] = b;
    }
    return sq / float(n * kOut);
}

}  // extern "C"
