// SPDX-License-Identifier: MIT
// The learned warm start's network: 16 inputs, two hidden layers of 64
// (ReLU), 2 outputs; parameters packed as W1[64][16], b1[64], W2[64][64],
// b2[64], W3[2][64], b3[2] (FP32 master copy; BF16 on the MFMA path), then
// the largest target of each output seen in training, which bounds the
// predictions (not trained).
#pragma once

#include "pmh_internal.hpp"

namespace pmh {

constexpr int kMlpIn = 16, kMlpHidden = 64, kMlpOut = 2;
constexpr int kMlpSample = kMlpIn + kMlpOut;                  // features then targets
constexpr int kMlpW1 = 0;
constexpr int kMlpB1 = kMlpW1 + kMlpHidden * kMlpIn;
constexpr int kMlpW2 = kMlpB1 + kMlpHidden;
constexpr int kMlpB2 = kMlpW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpW3 = kMlpB2 + kMlpHidden;
constexpr int kMlpB3 = kMlpW3 + kMlpOut * kMlpHidden;
constexpr int kMlpBound = kMlpB3 + kMlpOut;                  // largest training target per output
constexpr int kMlpParams = kMlpBound + kMlpOut;               // 5380

// Which predictions the warm start applies (PMH_WARMSTART_* of the C API).
constexpr uint32_t kWarmPenalty = 1u, kWarmLambda = 2u;

// Predicted (lambda, k) of the listed cold points into their normal rows.
PmhResult mlp_infer(hipStream_t s, const uint32_t *count, const uint32_t *list, const float *feat,
                    const float *scale, const float *params, float4 *c_lam, float4 *c_pen, uint32_t cap,
                    uint32_t policy);

}  // namespace pmh
