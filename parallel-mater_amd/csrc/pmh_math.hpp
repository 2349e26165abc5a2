// SPDX-License-Identifier: MIT
// Vector, quaternion and AVBD block arithmetic for host and device.
//
// The reference evaluates every expression below unfused and in this
// operand order (src/slang/math.slang:6-248, constraint_math.slang:16-276);
// the library is compiled with -ffp-contract=off so gfx950 produces the
// same IEEE-754 single results as the reference's CPU driver.  sqrt and
// division are correctly rounded on both sides; rsqrt is 1/sqrt as on the
// reference's CPU targets (src/simulation/runtime.cpp:173-175).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PMH_HD __host__ __device__ __forceinline__

namespace pmh {

struct V3 { float x, y, z; };
struct Q { float x, y, z, w; };

constexpr float kEpsilon = 1.0e-6f;
constexpr uint32_t kInvalid = 0xffffffffu;
constexpr uint32_t kStatic = 0u, kKinematic = 1u, kDynamic = 2u;

PMH_HD float rsq(float v) { return 1.0f / sqrtf(v); }
PMH_HD float smin(float a, float b) { return fminf(a, b); }
PMH_HD float smax(float a, float b) { return fmaxf(a, b); }
PMH_HD V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
PMH_HD V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
PMH_HD V3 mul(V3 v, float s) { return {v.x * s, v.y * s, v.z * s}; }
PMH_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PMH_HD V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PMH_HD V3 cmin(V3 a, V3 b) { return {smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)}; }
PMH_HD V3 cmax(V3 a, V3 b) { return {smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)}; }
PMH_HD float len2(V3 v) { return dot(v, v); }
PMH_HD float len(V3 v) { return sqrtf(len2(v)); }
PMH_HD float clampf(float v, float lo, float hi) { return smin(smax(v, lo), hi); }
PMH_HD V3 normalized_or(V3 v, V3 f) {
    float s = len2(v);
    return s <= kEpsilon * kEpsilon ? f : mul(v, rsq(s));
}
PMH_HD Q conj(Q q) { return {-q.x, -q.y, -q.z, q.w}; }
PMH_HD Q qmul(Q a, Q b) {
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
PMH_HD Q qnormalize(Q v) {
    float s = v.x * v.x + v.y * v.y + v.z * v.z + v.w * v.w;
    if (s <= kEpsilon * kEpsilon) return {0.0f, 0.0f, 0.0f, 0.0f};
    float inv = rsq(s);
    return {v.x * inv, v.y * inv, v.z * inv, v.w * inv};
}
PMH_HD Q relative_quaternion(Q from, Q to, bool local) {
    float c = from.x * to.x + from.y * to.y + from.z * to.z + from.w * to.w;
    if (c < 0.0f) to = {-to.x, -to.y, -to.z, -to.w};
    Q d = {to.x - from.x, to.y - from.y, to.z - from.z, to.w - from.w};
    Q r = local ? qmul(conj(from), d) : qmul(d, conj(from));
    r.w = fabsf(c);
    return qnormalize(r);
}
PMH_HD V3 rotate(Q q, V3 v) {
    V3 u = {q.x, q.y, q.z};
    V3 t = mul(cross(u, v), 2.0f);
    return add(v, add(mul(t, q.w), cross(u, t)));
}
PMH_HD V3 inv_rotate(Q q, V3 v) { return rotate(conj(q), v); }
PMH_HD V3 rotation_displacement(Q from, Q to, V3 v) {
    if (from.x * to.x + from.y * to.y + from.z * to.z + from.w * to.w < 0.0f)
        to = {-to.x, -to.y, -to.z, -to.w};
    V3 ov = {from.x, from.y, from.z}, nv = {to.x, to.y, to.z};
    V3 d = sub(nv, ov);
    V3 oc = cross(ov, v), dc = cross(d, v);
    return mul(add(add(mul(oc, to.w - from.w), mul(dc, to.w)), add(cross(d, oc), cross(nv, dc))), 2.0f);
}
PMH_HD V3 quaternion_rotation_vector(Q d) {
    float s = d.x * d.x + d.y * d.y + d.z * d.z;
    float scale = 2.0f + s * (1.0f / 3.0f + s * (3.0f / 20.0f + s * (5.0f / 56.0f + s * (35.0f / 576.0f))));
    if (s > 0.01f) {
        float vs = sqrtf(s);
        scale = 2.0f * atan2f(vs, clampf(d.w, 0.0f, 1.0f)) / vs;
    }
    return {d.x * scale, d.y * scale, d.z * scale};
}
PMH_HD V3 quaternion_delta_velocity(Q from, Q to, float dt) {
    if (dt <= 0.0f) return {0.0f, 0.0f, 0.0f};
    return mul(quaternion_rotation_vector(relative_quaternion(from, to, false)), 1.0f / dt);
}
PMH_HD V3 clamp_length(V3 v, float m) {
    float s = len2(v);
    if (s <= m * m || s <= kEpsilon * kEpsilon) return v;
    return mul(v, m * rsq(s));
}
PMH_HD V3 basis_axis(uint32_t a) {
    return a == 0 ? V3{1.0f, 0.0f, 0.0f} : (a == 1 ? V3{0.0f, 1.0f, 0.0f} : V3{0.0f, 0.0f, 1.0f});
}
PMH_HD float comp(V3 v, uint32_t a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
// Contact tangent basis, geometry_constraints.slang:424-430.
PMH_HD void contact_basis(V3 n, V3 &t0, V3 &t1) {
    t0 = normalized_or(cross(n, fabsf(n.x) < 0.57735f ? V3{1.0f, 0.0f, 0.0f} : V3{0.0f, 1.0f, 0.0f}),
                       {0.0f, 0.0f, 1.0f});
    t1 = cross(n, t0);
}

// ---- AVBD numerical core, constraint_math.slang:8-276 -----------------
constexpr float kAlpha = 0.99f, kGamma = 0.999f;
constexpr float kMinPenalty = 1.0f, kMaxPenalty = 1.0e10f;

PMH_HD float amin(float a, float b) { return a < b ? a : b; }
PMH_HD float amax(float a, float b) { return a > b ? a : b; }
PMH_HD float aclamp(float v, float lo, float hi) { return amin(amax(v, lo), hi); }

struct Force3 { float n, t0, t1; };

PMH_HD void warm_start(float &lambda, float &penalty, float stiffness) {
    lambda *= kAlpha * kGamma;
    penalty = amin(stiffness, aclamp(penalty * kGamma, kMinPenalty, kMaxPenalty));
}
PMH_HD float friction_scale(float t0, float t1, float bound) {
    if (bound <= 0.0f) return 0.0f;
    float s = t0 * t0 + t1 * t1;
    return s <= bound * bound ? 1.0f : amin(1.0f, bound * rsq(s));
}
PMH_HD Force3 project_contact(Force3 f, float mu) {
    f.n = amax(f.n, 0.0f);
    float bound = mu * f.n;
    float s = friction_scale(f.t0, f.t1, bound);
    f.t0 *= s;
    f.t1 *= s;
    return f;
}
PMH_HD Force3 stiffness_scales(Force3 t, float mu) {
    float n = amax(t.n, 0.0f);
    float s = friction_scale(t.t0, t.t1, mu * n);
    return {t.n > 0.0f ? 1.0f : 0.0f, s, s};
}

// Packed symmetric 6x6 block: h[r*(r+1)/2 + c], c <= r.
struct Block { float h[21]; float g[6]; };

PMH_HD void block_zero(Block &b) {
#pragma unroll
    for (int i = 0; i < 21; ++i) b.h[i] = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; ++i) b.g[i] = 0.0f;
}
PMH_HD float dot6(const float *a, const float *b) {
    float r = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; ++i) r += a[i] * b[i];
    return r;
}
PMH_HD void add_row(Block &b, const float *j, float k, float f) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        b.g[i] += j[i] * f;
#pragma unroll
        for (int c = 0; c <= i; ++c) b.h[i * (i + 1) / 2 + c] += k * j[i] * j[c];
    }
}

// avbdSolve, constraint_math.slang:165-276: LDL^T with a pivot floor and
// an equilibrated retry.  Fully unrolled: every index is a constant.
PMH_HD bool block_solve(const Block &in, float *delta) {
    const float largest = 3.402823466e38f;
    const float eps = 1.1920928955078125e-7f;
    const float pivot_floor = 8.0f * eps;
    const float roundoff = 64.0f * eps;
#pragma unroll
    for (int i = 0; i < 6; ++i) delta[i] = 0.0f;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float d = in.h[i * (i + 1) / 2 + i];
        ok = ok && (d > 0.0f) && (d < largest) && (fabsf(in.g[i]) < largest);
#pragma unroll
        for (int j = 0; j < i; ++j) ok = ok && (fabsf(in.h[i * (i + 1) / 2 + j]) < largest);
    }
    if (!ok) return false;
#pragma unroll
    for (int attempt = 0; attempt < 2; ++attempt) {
        Block b = in;
        float scale[6], diag[6], fwd[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            scale[i] = attempt != 0 ? rsq(b.h[i * (i + 1) / 2 + i]) : 1.0f;
            diag[i] = 0.0f;
            fwd[i] = 0.0f;
        }
        if (attempt != 0) {
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                b.g[i] *= scale[i];
                if (!(fabsf(b.g[i]) < largest)) return false;
#pragma unroll
                for (int j = 0; j < i; ++j) {
                    b.h[i * (i + 1) / 2 + j] *= scale[i] * scale[j];
                    if (!(fabsf(b.h[i * (i + 1) / 2 + j]) <= 1.0f + roundoff)) return false;
                }
                b.h[i * (i + 1) / 2 + i] = 1.0f;
            }
        }
        bool retry = false;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            if (!retry) {
#pragma unroll
                for (int j = 0; j < i; ++j) {
                    float sum = b.h[i * (i + 1) / 2 + j];
#pragma unroll
                    for (int k = 0; k < j; ++k)
                        sum -= b.h[i * (i + 1) / 2 + k] * diag[k] * b.h[j * (j + 1) / 2 + k];
                    b.h[i * (i + 1) / 2 + j] = sum / diag[j];
                }
                float d = b.h[i * (i + 1) / 2 + i];
#pragma unroll
                for (int k = 0; k < i; ++k)
                    d -= b.h[i * (i + 1) / 2 + k] * b.h[i * (i + 1) / 2 + k] * diag[k];
                if (attempt == 0) {
                    if (!(d > pivot_floor * in.h[i * (i + 1) / 2 + i]) || !(d < largest)) retry = true;
                    else diag[i] = d;
                } else {
                    if (!(d >= -roundoff) || !(d < largest)) return false;
                    diag[i] = amax(d, pivot_floor);
                }
                if (!retry) {
                    fwd[i] = -b.g[i];
#pragma unroll
                    for (int k = 0; k < i; ++k) fwd[i] -= b.h[i * (i + 1) / 2 + k] * fwd[k];
                }
            }
        }
        if (retry) continue;
#pragma unroll
        for (int i = 5; i >= 0; --i) {
            delta[i] = fwd[i] / diag[i];
#pragma unroll
            for (int k = i + 1; k < 6; ++k) delta[i] -= b.h[k * (k + 1) / 2 + i] * delta[k];
        }
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            delta[i] *= scale[i];
            if (!(fabsf(delta[i]) < largest)) return false;
        }
        return true;
    }
    return false;
}

}  // namespace pmh
