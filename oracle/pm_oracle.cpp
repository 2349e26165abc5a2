// SPDX-License-Identifier: MIT
//
// CPU oracle for the ParallelMater rigid-body step.  TEST INFRASTRUCTURE:
// only the tests, the smoke check and the benchmark's CPU-baseline leg use
// it, as the checker.  It restates, serially and with O(n) storage, the
// reference's Slang path as its CPU driver executes it (one lane per
// dispatch, src/device/cpu_dispatch.hpp:30-60).  Every function cites the
// reference lines it follows.  Compile with -ffp-contract=off: the
// reference's CPU driver evaluates every float expression unfused.
//
// Deliberate storage changes (results unchanged):
//  * manifolds are indexed by active pair, not by the n(n-1)/2 table;
//  * the contact cache is keyed by pair code in a map, not an n^2 slot
//    array (geometry.slang:2941-3070 semantics kept: hit iff the pair had a
//    non-empty manifold in the immediately preceding step);
//  * free_mean reduces each island's members in ascending index order
//    instead of every root scanning all n bodies
//    (geometry_constraints_solve.slang:617-675).
//
// Parity status per part (how each is pinned) is listed in the header
// comment of each section below.

#include "pm_oracle.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <unordered_map>
#include <vector>

static_assert(sizeof(PmoState) == 52, "RigidBodyState layout");
static_assert(sizeof(PmoParams) == 64, "BodyParameters layout");
static_assert(sizeof(PmoHingeFrame) == 72, "HingeContactFrame layout");
static_assert(sizeof(PmoContact) == 284, "Contact layout");
static_assert(sizeof(PmoManifold) == 2304, "ContactManifold layout");
static_assert(sizeof(PmoAvbdBody) == 236, "AvbdBody layout");
static_assert(sizeof(PmoSchedule) == 264, "ContactSchedule layout");
static_assert(sizeof(PmoRow) == 88, "RigidAvbdRow layout");
static_assert(sizeof(PmoJointOptions) == 228, "RigidConstraintOptions layout");
static_assert(sizeof(PmoJoint) == 1852, "RigidConstraintResource layout");

namespace {

using V3 = PmoVec3;
using Q = PmoQuat;

constexpr float kEpsilon = 1.0e-6f;            // types.slang:19
constexpr uint32_t kInvalid = 0xffffffffu;     // types.slang:20
constexpr uint32_t kStatic = 0u, kKinematic = 1u, kDynamic = 2u;

// ---------------------------------------------------------------------
// math.slang:6-248.  rsqrt is 1/sqrt as on the reference's CPU targets
// (host reciprocal_sqrt, src/simulation/runtime.cpp:173-175).
// ---------------------------------------------------------------------
inline float rsq(float v) { return 1.0f / std::sqrt(v); }
inline float fmin_(float a, float b) { return std::fmin(a, b); }
inline float fmax_(float a, float b) { return std::fmax(a, b); }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 v, float s) { return {v.x * s, v.y * s, v.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
            a.x * b.y - a.y * b.x};
}
inline V3 cmin(V3 a, V3 b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
inline V3 cmax(V3 a, V3 b) { return {fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
inline float len2(V3 v) { return dot(v, v); }
inline float len(V3 v) { return std::sqrt(len2(v)); }
inline float clampf(float v, float lo, float hi) { return fmin_(fmax_(v, lo), hi); }
inline V3 normalized_or(V3 v, V3 fallback) {                    // math.slang:59-64
    float s = len2(v);
    return s <= kEpsilon * kEpsilon ? fallback : mul(v, rsq(s));
}
inline Q conj(Q q) { return {-q.x, -q.y, -q.z, q.w}; }
inline Q qmul(Q a, Q b) {                                       // math.slang:71-78
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w,
            a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
inline Q qnormalize(Q v) {                                       // math.slang:80-88
    float s = v.x * v.x + v.y * v.y + v.z * v.z + v.w * v.w;
    if (s <= kEpsilon * kEpsilon) return {0, 0, 0, 0};
    float inv = rsq(s);
    return {v.x * inv, v.y * inv, v.z * inv, v.w * inv};
}
inline Q relative_quaternion(Q from, Q to, bool local) {         // math.slang:90-105
    float c = from.x * to.x + from.y * to.y + from.z * to.z + from.w * to.w;
    if (c < 0.0f) to = {-to.x, -to.y, -to.z, -to.w};
    Q d = {to.x - from.x, to.y - from.y, to.z - from.z, to.w - from.w};
    Q r = local ? qmul(conj(from), d) : qmul(d, conj(from));
    r.w = std::fabs(c);
    return qnormalize(r);
}
inline V3 rotate(Q q, V3 v) {                                    // math.slang:107-113
    V3 u = {q.x, q.y, q.z};
    V3 t = mul(cross(u, v), 2.0f);
    return add(v, add(mul(t, q.w), cross(u, t)));
}
inline V3 inv_rotate(Q q, V3 v) { return rotate(conj(q), v); }
inline V3 rotation_displacement(Q from, Q to, V3 v) {            // math.slang:122-135
    if (from.x * to.x + from.y * to.y + from.z * to.z + from.w * to.w < 0.0f)
        to = {-to.x, -to.y, -to.z, -to.w};
    V3 ov = {from.x, from.y, from.z}, nv = {to.x, to.y, to.z};
    V3 d = sub(nv, ov);
    V3 oc = cross(ov, v), dc = cross(d, v);
    return mul(add(add(mul(oc, to.w - from.w), mul(dc, to.w)),
                   add(cross(d, oc), cross(nv, dc))), 2.0f);
}
inline V3 transform_point(const PmoState &s, V3 p) { return add(s.position, rotate(s.orientation, p)); }
inline V3 inverse_inertia_world(const PmoParams &p, const PmoState &s, V3 w) {  // math.slang:208-217
    V3 l = inv_rotate(s.orientation, w);
    V3 t = {l.x * p.inverse_inertia_local.x, l.y * p.inverse_inertia_local.y,
            l.z * p.inverse_inertia_local.z};
    return rotate(s.orientation, t);
}
inline V3 quaternion_rotation_vector(Q d) {                      // math.slang:221-233
    float s = d.x * d.x + d.y * d.y + d.z * d.z;
    float scale = 2.0f + s * (1.0f / 3.0f + s * (3.0f / 20.0f + s * (5.0f / 56.0f +
                  s * (35.0f / 576.0f))));
    if (s > 0.01f) {
        float vs = std::sqrt(s);
        scale = 2.0f * std::atan2(vs, clampf(d.w, 0.0f, 1.0f)) / vs;
    }
    return {d.x * scale, d.y * scale, d.z * scale};
}
inline V3 quaternion_delta_velocity(Q from, Q to, float dt) {    // math.slang:235-240
    if (dt <= 0.0f) return {0, 0, 0};
    return mul(quaternion_rotation_vector(relative_quaternion(from, to, false)), 1.0f / dt);
}
inline V3 clamp_length(V3 v, float m) {                          // math.slang:242-248
    float s = len2(v);
    if (s <= m * m || s <= kEpsilon * kEpsilon) return v;
    return mul(v, m * rsq(s));
}
inline bool finite3(V3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

// ---------------------------------------------------------------------
// Numerical core, constraint_math.slang:8-345.
// Pinned by the reference's 85 rigid AVBD fixtures
// (tests/solver/geometry_constraint_cases.hpp, block cases 0-28 and scalar
// cases 0-55) through tests/test_oracle_fixtures.py.
// ---------------------------------------------------------------------
constexpr float kAlpha = 0.99f, kGamma = 0.999f;
constexpr float kMinPenalty = 1.0f, kMaxPenalty = 1.0e10f;

struct Dual { float lambda, penalty; };
struct Force3 { float normal, t0, t1; };
struct Vec6 { float v[6]; };
struct Block { float h[21]; float g[6]; };
struct Row {
    Vec6 a, b; Dual dual; float error, velocity, lower, upper, stiffness,
    damping, beta, reference_force;
};

inline float amin(float a, float b) { return a < b ? a : b; }   // :16
inline float amax(float a, float b) { return a > b ? a : b; }   // :17
inline float aabs(float v) { return v < 0.0f ? -v : v; }         // :18
inline float aclamp(float v, float lo, float hi) { return amin(amax(v, lo), hi); }
inline uint32_t bidx(uint32_t r, uint32_t c) {                   // :134-139
    uint32_t hi = r > c ? r : c, lo = r > c ? c : r;
    return hi * (hi + 1u) / 2u + lo;
}
inline Dual warm_start(Dual d, float stiffness) {                // :30-37
    d.lambda *= kAlpha * kGamma;
    d.penalty = amin(stiffness, aclamp(d.penalty * kGamma, kMinPenalty, kMaxPenalty));
    return d;
}
inline float dual_force(Dual d, float e, float lo, float hi) { return aclamp(d.lambda + d.penalty * e, lo, hi); }
inline Dual update_dual(Dual d, float e, float beta, float lo, float hi, float stiffness) {  // :44-56
    float u = d.lambda + d.penalty * e;
    d.lambda = stiffness < kMaxPenalty ? 0.0f : aclamp(u, lo, hi);
    if (u > lo && u < hi)
        d.penalty = amin(stiffness, amin(kMaxPenalty, d.penalty + beta * aabs(e)));
    return d;
}
inline float friction_scale(float t0, float t1, float bound) {   // :65-73
    if (bound <= 0.0f) return 0.0f;
    float s = t0 * t0 + t1 * t1;
    return s <= bound * bound ? 1.0f : amin(1.0f, bound * rsq(s));
}
inline Force3 project_contact(Force3 f, float mu) {              // :75-83
    f.normal = amax(f.normal, 0.0f);
    float bound = mu * f.normal;
    float s = friction_scale(f.t0, f.t1, bound);
    f.t0 *= s; f.t1 *= s;
    return f;
}
inline Force3 stiffness_scales(Force3 trial, float mu) {         // :85-96
    float n = amax(trial.normal, 0.0f);
    float bound = mu * n;
    float s = friction_scale(trial.t0, trial.t1, bound);
    return {trial.normal > 0.0f ? 1.0f : 0.0f, s, s};
}
inline float dot6(const Vec6 &a, const Vec6 &b) {                // :116-123
    float r = 0.0f;
    for (int i = 0; i < 6; ++i) r += a.v[i] * b.v[i];
    return r;
}
inline Block zero_block() { Block b; std::memset(&b, 0, sizeof b); return b; }
inline Vec6 zero6() { Vec6 v; std::memset(&v, 0, sizeof v); return v; }
inline void add_row(Block &b, const Vec6 &j, float k, float f) { // :151-163
    for (uint32_t i = 0; i < 6; ++i) {
        b.g[i] += j.v[i] * f;
        for (uint32_t c = 0; c <= i; ++c) b.h[bidx(i, c)] += k * j.v[i] * j.v[c];
    }
}
bool block_solve(const Block &input, Vec6 &delta) {              // :165-276
    const float largest = 3.402823466e38f;
    const float eps = 1.1920928955078125e-7f;
    const float pivot_floor = 8.0f * eps;
    const float roundoff = 64.0f * eps;
    delta = zero6();
    for (uint32_t i = 0; i < 6; ++i) {
        if (!(input.h[bidx(i, i)] > 0.0f) || !(input.h[bidx(i, i)] < largest) ||
            !(aabs(input.g[i]) < largest)) return false;
        for (uint32_t j = 0; j < i; ++j)
            if (!(aabs(input.h[bidx(i, j)]) < largest)) return false;
    }
    for (uint32_t attempt = 0; attempt < 2; ++attempt) {
        Block b = input;
        float scale[6], diag[6], fwd[6];
        for (uint32_t i = 0; i < 6; ++i) {
            scale[i] = attempt != 0 ? rsq(b.h[bidx(i, i)]) : 1.0f;
            diag[i] = 0.0f; fwd[i] = 0.0f;
        }
        if (attempt != 0) {
            for (uint32_t i = 0; i < 6; ++i) {
                b.g[i] *= scale[i];
                if (!(aabs(b.g[i]) < largest)) return false;
                for (uint32_t j = 0; j < i; ++j) {
                    b.h[bidx(i, j)] *= scale[i] * scale[j];
                    if (!(aabs(b.h[bidx(i, j)]) <= 1.0f + roundoff)) return false;
                }
                b.h[bidx(i, i)] = 1.0f;
            }
        }
        bool retry = false;
        for (uint32_t i = 0; i < 6; ++i) {
            for (uint32_t j = 0; j < i; ++j) {
                float sum = b.h[bidx(i, j)];
                for (uint32_t k = 0; k < j; ++k)
                    sum -= b.h[bidx(i, k)] * diag[k] * b.h[bidx(j, k)];
                b.h[bidx(i, j)] = sum / diag[j];
            }
            float d = b.h[bidx(i, i)];
            for (uint32_t k = 0; k < i; ++k)
                d -= b.h[bidx(i, k)] * b.h[bidx(i, k)] * diag[k];
            if (attempt == 0) {
                if (!(d > pivot_floor * input.h[bidx(i, i)]) || !(d < largest)) {
                    retry = true;
                    break;
                }
                diag[i] = d;
            } else {
                if (!(d >= -roundoff) || !(d < largest)) return false;
                diag[i] = amax(d, pivot_floor);
            }
            fwd[i] = -b.g[i];
            for (uint32_t k = 0; k < i; ++k) fwd[i] -= b.h[bidx(i, k)] * fwd[k];
        }
        if (retry) continue;
        for (int i = 5; i >= 0; --i) {
            delta.v[i] = fwd[i] / diag[i];
            for (uint32_t k = uint32_t(i) + 1; k < 6; ++k)
                delta.v[i] -= b.h[bidx(k, uint32_t(i))] * delta.v[k];
        }
        for (uint32_t i = 0; i < 6; ++i) {
            delta.v[i] *= scale[i];
            if (!(aabs(delta.v[i]) < largest)) return false;
        }
        return true;
    }
    return false;
}
inline float row_error(const Row &r, const Vec6 &a, const Vec6 &b) {   // :312-315
    return r.error + dot6(r.a, a) + dot6(r.b, b);
}
inline float row_trial(const Row &r, const Vec6 &a, const Vec6 &b, float dt) {  // :317-322
    float d = dot6(r.a, a) + dot6(r.b, b);
    return r.dual.lambda + r.dual.penalty * (r.error + d) + r.damping * (r.velocity + d / dt);
}
inline float row_force(const Row &r, const Vec6 &a, const Vec6 &b, float dt) {
    return aclamp(row_trial(r, a, b, dt), r.lower, r.upper);
}
inline void row_accumulate(Block &blk, const Row &r, bool first, const Vec6 &a,
                           const Vec6 &b, float dt) {                     // :329-338
    float trial = row_trial(r, a, b, dt);
    bool two_sided = r.lower < 0.0f && r.upper > 0.0f;
    float k = two_sided || (trial > r.lower && trial < r.upper)
        ? r.dual.penalty + r.damping / dt : 0.0f;
    add_row(blk, first ? r.a : r.b, k, aclamp(trial, r.lower, r.upper) - r.reference_force);
}
inline void row_advance(Row &r, const Vec6 &a, const Vec6 &b) {          // :340-345
    r.dual = update_dual(r.dual, row_error(r, a, b), r.beta, r.lower, r.upper, r.stiffness);
}
Row row_from(const float *f) {
    Row r;
    for (int i = 0; i < 6; ++i) { r.a.v[i] = f[i]; r.b.v[i] = f[6 + i]; }
    r.dual = {f[12], f[13]};
    r.error = f[14]; r.velocity = f[15]; r.lower = f[16]; r.upper = f[17];
    r.stiffness = f[18]; r.damping = f[19]; r.beta = f[20]; r.reference_force = f[21];
    return r;
}
void row_to(const Row &r, float *f) {
    for (int i = 0; i < 6; ++i) { f[i] = r.a.v[i]; f[6 + i] = r.b.v[i]; }
    f[12] = r.dual.lambda; f[13] = r.dual.penalty; f[14] = r.error; f[15] = r.velocity;
    f[16] = r.lower; f[17] = r.upper; f[18] = r.stiffness; f[19] = r.damping;
    f[20] = r.beta; f[21] = r.reference_force;
}

// ---------------------------------------------------------------------
// Shapes.  Mesh upload preparation restates runtime_deformables.cu.inc:
// 104-281 (median-split BVH with <=4 triangles per leaf and the index
// buffer rewritten in BVH order), closed_convex_planes
// (runtime_domains.cu.inc:1061-1108) and unique_convex_planes
// (src/convex_plane_dedup.hpp).  Pinned by closed forms in the tests
// (unit cube planes and offsets, box bounding sphere, inertia formula).
// ---------------------------------------------------------------------
struct Plane { V3 normal; float offset; };
enum ShapeKind : uint32_t { kMesh = 0, kSphere = 1, kCapsule = 2 };

// Median-split BVH node (geometry_storage.hpp:174-181): leaves hold
// triangles [first, first + count) of the reordered index buffer.
struct BvhNode {
    V3 minimum, maximum;
    uint32_t left = 0, right = 0, first = 0, count = 0;
};

struct Shape {
    uint32_t kind = kMesh;
    std::vector<V3> vertices;
    std::vector<uint32_t> indices;
    std::vector<BvhNode> nodes;     // preorder
    std::vector<uint32_t> leaves;   // indices of leaf nodes, ascending
    std::vector<V3> shell;          // certified outward normal per triangle (may be empty)
    std::vector<Plane> planes;      // per triangle, then the unique tail
    uint32_t unique_planes = 0;
    bool has_planes = false;
    V3 minimum{}, maximum{}, center{};
    float radius = 0.0f;            // bounding radius
    V3 unit_inertia{};
    float sphere_radius = 0.0f, half_length = 0.0f;
    uint32_t index_count() const { return uint32_t(indices.size()); }
    uint32_t vertex_count() const { return uint32_t(vertices.size()); }
};

std::vector<Plane> closed_convex_planes(const std::vector<V3> &v,
                                        const std::vector<uint32_t> &idx) {
    std::map<std::array<float, 3>, uint32_t> welded;
    std::vector<uint32_t> remap(v.size());
    V3 center{0, 0, 0};
    for (uint32_t i = 0; i < v.size(); ++i) {
        auto e = welded.emplace(std::array<float, 3>{v[i].x, v[i].y, v[i].z},
                                uint32_t(welded.size()));
        remap[i] = e.first->second;
        if (e.second) center = add(center, v[i]);
    }
    if (welded.size() < 4u) return {};
    center = mul(center, 1.0f / float(welded.size()));
    std::unordered_map<uint64_t, uint32_t> edges;
    for (size_t i = 0; i < idx.size(); i += 3)
        for (size_t e = 0; e < 3; ++e) {
            uint32_t a = remap[idx[i + e]], b = remap[idx[i + (e + 1) % 3]];
            ++edges[(uint64_t(std::min(a, b)) << 32) | std::max(a, b)];
        }
    for (auto &e : edges) if (e.second != 2u) return {};
    float scale = 0.0f;
    for (auto &p : v) scale = std::max(scale, len(sub(p, center)));
    float tol = std::max(1.0e-7f, scale * 1.0e-5f);
    std::vector<Plane> planes;
    for (size_t i = 0; i < idx.size(); i += 3) {
        V3 a = v[idx[i]];
        V3 n = normalized_or(cross(sub(v[idx[i + 1]], a), sub(v[idx[i + 2]], a)), {0, 0, 0});
        float side = dot(n, sub(a, center));
        if (std::fabs(side) <= tol) return {};
        if (side < 0.0f) n = mul(n, -1.0f);
        float off = dot(n, a);
        for (auto &p : v) if (dot(n, p) - off > tol) return {};
        planes.push_back({n, off});
    }
    return planes;
}

// closed_shell_normals, runtime_domains.cu.inc:1111-1245: per connected
// component (vertices welded by position), convex components take their
// face planes' normals, closed orientable ones the signed-volume
// orientation, open volumetric ones a support-plane side or a quarter-length
// two-sided marker on interior triangles.  Empty when nothing is certified.
std::vector<V3> closed_shell_normals(const std::vector<V3> &v, const std::vector<uint32_t> &idx) {
    const uint32_t nv = uint32_t(v.size());
    std::map<std::array<float, 3>, uint32_t> weld;
    std::vector<uint32_t> rep(nv), up(nv);
    for (uint32_t i = 0; i < nv; ++i) {
        rep[i] = weld.emplace(std::array<float, 3>{v[i].x, v[i].y, v[i].z}, i).first->second;
        up[i] = i;
    }
    auto find = [&](uint32_t i) {
        while (up[i] != i) { up[i] = up[up[i]]; i = up[i]; }
        return i;
    };
    for (size_t t = 0; t < idx.size(); t += 3)
        for (size_t c = 1; c < 3; ++c) {
            uint32_t r0 = find(rep[idx[t]]), rc = find(rep[idx[t + c]]);
            up[rc] = r0;
        }
    const uint32_t tri = uint32_t(idx.size() / 3);
    std::map<uint32_t, std::vector<uint32_t>> components;
    for (uint32_t t = 0; t < tri; ++t) components[find(rep[idx[3 * t]])].push_back(t);
    std::vector<V3> out(tri, V3{0, 0, 0});
    bool any = false;
    for (auto &kv : components) {
        const std::vector<uint32_t> &ts = kv.second;
        std::map<uint32_t, uint32_t> local;
        std::vector<V3> pts;
        std::vector<uint32_t> faces;
        for (uint32_t t : ts)
            for (uint32_t c = 0; c < 3; ++c) {
                auto e = local.emplace(rep[idx[3 * t + c]], uint32_t(pts.size()));
                if (e.second) pts.push_back(v[rep[idx[3 * t + c]]]);
                faces.push_back(e.first->second);
            }
        auto planes = closed_convex_planes(pts, faces);
        if (!planes.empty()) {
            any = true;
            for (size_t i = 0; i < ts.size(); ++i) out[ts[i]] = planes[i].normal;
            continue;
        }
        std::map<std::pair<uint32_t, uint32_t>, std::pair<uint32_t, int>> edges;
        double volume = 0.0;
        for (size_t f = 0; f < faces.size(); f += 3) {
            for (size_t c = 0; c < 3; ++c) {
                uint32_t e0 = faces[f + c], e1 = faces[f + (c + 1) % 3];
                auto &use = edges[{std::min(e0, e1), std::max(e0, e1)}];
                use.first += 1;
                use.second += e0 < e1 ? 1 : -1;
            }
            double d[3][3];
            for (size_t c = 0; c < 3; ++c) {
                V3 q = pts[faces[f + c]];
                d[c][0] = double(q.x) - pts[0].x;
                d[c][1] = double(q.y) - pts[0].y;
                d[c][2] = double(q.z) - pts[0].z;
            }
            volume += d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) +
                      d[0][1] * (d[1][2] * d[2][0] - d[1][0] * d[2][2]) +
                      d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
        }
        bool closed = volume != 0.0 && std::isfinite(volume);
        for (auto &e : edges) closed = closed && e.second.first == 2u && e.second.second == 0;
        if (closed) {
            any = true;
            for (size_t i = 0; i < ts.size(); ++i) {
                V3 a = pts[faces[3 * i]], b = pts[faces[3 * i + 1]], c = pts[faces[3 * i + 2]];
                out[ts[i]] = mul(normalized_or(cross(sub(b, a), sub(c, a)), {0, 0, 0}), volume > 0.0 ? 1.0f : -1.0f);
            }
            continue;
        }
        float scale = 0.0f;
        for (V3 q : pts) scale = std::max(scale, len(sub(q, pts[0])));
        const float tol = std::max(1.0e-7f, scale * 1.0e-5f);
        for (size_t i = 0; i < ts.size(); ++i) {
            V3 a = pts[faces[3 * i]];
            V3 n = normalized_or(cross(sub(pts[faces[3 * i + 1]], a), sub(pts[faces[3 * i + 2]], a)), {0, 0, 0});
            float lo = 0.0f, hi = 0.0f;
            for (V3 q : pts) {
                float d = dot(n, sub(q, a));
                lo = std::min(lo, d);
                hi = std::max(hi, d);
            }
            if (hi <= tol && lo < -tol) {
                out[ts[i]] = n;
            } else if (lo >= -tol && hi > tol) {
                out[ts[i]] = mul(n, -1.0f);
            } else {
                bool interior = true;
                for (size_t c = 0; c < 3; ++c) {
                    uint32_t e0 = faces[3 * i + c], e1 = faces[3 * i + (c + 1) % 3];
                    auto use = edges.at({std::min(e0, e1), std::max(e0, e1)});
                    interior = interior && use.first == 2u && use.second == 0;
                }
                if (!interior) continue;
                out[ts[i]] = mul(n, 0.25f);
            }
            any = true;
        }
    }
    return any ? out : std::vector<V3>{};
}

bool prepare_mesh(Shape &s, const float *verts, uint32_t nv, const uint32_t *idx, uint32_t ni) {
    if (nv < 3 || ni < 3 || ni % 3 != 0) return false;
    s.kind = kMesh;
    s.vertices.resize(nv);
    for (uint32_t i = 0; i < nv; ++i) s.vertices[i] = {verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]};
    std::vector<uint32_t> in(idx, idx + ni);
    for (auto &p : s.vertices) if (!finite3(p)) return false;
    V3 mn = s.vertices[0], mx = s.vertices[0];
    for (uint32_t i = 1; i < nv; ++i) { mn = cmin(mn, s.vertices[i]); mx = cmax(mx, s.vertices[i]); }
    for (uint32_t i = 0; i < ni; i += 3) {
        if (in[i] >= nv || in[i + 1] >= nv || in[i + 2] >= nv) return false;
        V3 area = cross(sub(s.vertices[in[i + 1]], s.vertices[in[i]]),
                        sub(s.vertices[in[i + 2]], s.vertices[in[i]]));
        if (len2(area) <= kEpsilon * kEpsilon) return false;
    }
    // Median-split BVH order (runtime_deformables.cu.inc:140-200).
    const uint32_t tri = ni / 3;
    std::vector<uint32_t> order(tri);
    for (uint32_t i = 0; i < tri; ++i) order[i] = i;
    auto coord = [](V3 v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); };
    auto centroid = [&](uint32_t t) {
        uint32_t o = t * 3;
        return mul(add(add(s.vertices[in[o]], s.vertices[in[o + 1]]), s.vertices[in[o + 2]]), 1.0f / 3.0f);
    };
    struct Builder {
        std::vector<uint32_t> &order;
        std::vector<BvhNode> &nodes;
        decltype(coord) &coordf;
        decltype(centroid) &centroidf;
        const std::vector<V3> &verts;
        const std::vector<uint32_t> &in;
        uint32_t build(uint32_t begin, uint32_t end) {
            BvhNode node;
            node.minimum = {FLT_MAX, FLT_MAX, FLT_MAX};
            node.maximum = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            for (uint32_t it = begin; it < end; ++it) {
                uint32_t o = order[it] * 3;
                for (uint32_t c = 0; c < 3; ++c) {
                    node.minimum = cmin(node.minimum, verts[in[o + c]]);
                    node.maximum = cmax(node.maximum, verts[in[o + c]]);
                }
            }
            const uint32_t self = uint32_t(nodes.size());
            nodes.push_back(node);
            if (end - begin <= 4u) {
                nodes[self].first = begin;
                nodes[self].count = end - begin;
                return self;
            }
            V3 ext = sub(node.maximum, node.minimum);
            int axis = ext.x >= ext.y && ext.x >= ext.z ? 0 : (ext.y >= ext.z ? 1 : 2);
            std::stable_sort(order.begin() + begin, order.begin() + end, [&](uint32_t a, uint32_t b) {
                float fa = coordf(centroidf(a), axis), fb = coordf(centroidf(b), axis);
                return fa < fb || (fa == fb && a < b);
            });
            uint32_t mid = begin + (end - begin) / 2u;
            uint32_t l = build(begin, mid);
            uint32_t r = build(mid, end);
            nodes[self].left = l;
            nodes[self].right = r;
            return self;
        }
    } builder{order, s.nodes, coord, centroid, s.vertices, in};
    s.nodes.clear();
    builder.build(0, tri);
    s.leaves.clear();
    for (uint32_t k = 0; k < s.nodes.size(); ++k)
        if (s.nodes[k].count != 0u) s.leaves.push_back(k);
    s.indices.resize(ni);
    for (uint32_t t = 0; t < tri; ++t)
        for (uint32_t c = 0; c < 3; ++c) s.indices[t * 3 + c] = in[order[t] * 3 + c];
    auto planes = closed_convex_planes(s.vertices, s.indices);
    s.shell = closed_shell_normals(s.vertices, s.indices);
    s.has_planes = !planes.empty();
    if (s.has_planes) {
        std::map<std::array<float, 4>, bool> seen;
        std::vector<Plane> unique;
        for (auto &p : planes)
            if (seen.emplace(std::array<float, 4>{p.normal.x, p.normal.y, p.normal.z, p.offset}, true).second)
                unique.push_back(p);
        s.unique_planes = uint32_t(unique.size());
        s.planes = planes;
        s.planes.insert(s.planes.end(), unique.begin(), unique.end());
    }
    s.minimum = mn; s.maximum = mx;
    s.center = mul(add(mn, mx), 0.5f);
    float r2 = 0.0f;
    for (auto &p : s.vertices) r2 = std::fmax(r2, len2(sub(p, s.center)));
    s.radius = std::sqrt(r2);
    V3 h = mul(sub(mx, mn), 0.5f);
    s.unit_inertia = {std::fmax((h.y * h.y + h.z * h.z) / 3.0f, kEpsilon),
                      std::fmax((h.x * h.x + h.z * h.z) / 3.0f, kEpsilon),
                      std::fmax((h.x * h.x + h.y * h.y) / 3.0f, kEpsilon)};
    return true;
}

// ---------------------------------------------------------------------
// Contact bookkeeping shared by every narrowphase.
// ---------------------------------------------------------------------
void clear_manifold(PmoManifold &m) { std::memset(&m, 0, sizeof m); }   // geometry.slang:192-202

// rigid_add_contact, geometry.slang:1129-1220, for unguided contacts (no
// hinge joints exist on this path) and no swept-witness canonicalisation
// (the convex face path passes false).
// rigid_guided_pair / rigid_guided_contact, geometry.slang:373-382: an
// axially guided member (a slider or piston held by a static body) against
// a static body.
inline bool guided_pair(const PmoHingeFrame &a, const PmoHingeFrame &b) {
    return (a.axial != 0u && b.static_body != 0u) || (b.axial != 0u && a.static_body != 0u);
}
inline bool guided_contact(const PmoContact &c) { return guided_pair(c.body_hinge, c.collider_hinge); }

// rigid_guide_direction, geometry.slang:1118-1127: the contact's response
// along the guide's free directions (axial translation; rotation about the
// axis for a piston), each scaled by the square root of its inverse mass.
inline V3 guide_direction(const PmoContact &c) {
    const PmoHingeFrame &f = c.body_hinge.axial != 0u ? c.body_hinge : c.collider_hinge;
    return {dot(c.normal, f.axis) * std::sqrt(f.inverse_mass),
            f.axial_rotation != 0u ? dot(f.axis, cross(sub(c.point, f.anchor), c.normal)) * std::sqrt(f.inverse_moment)
                                   : 0.0f,
            0.0f};
}

// rigid_add_contact, geometry.slang:1129-1220.  A guided contact is merged
// by its guide direction: the first impact keeps its rows (later impacts
// are dropped, earlier ones replace the swept rows), and a contact whose
// direction matches a kept one replaces it when it is deeper per unit of
// direction.
void add_contact(PmoManifold &m, const PmoContact &c, float spacing) {
    const bool guided = guided_contact(c);
    const V3 cand_dir = guided ? guide_direction(c) : V3{0, 0, 0};
    const float cand_len = guided ? len(cand_dir) : 1.0f;
    if (guided && cand_len * cand_len <= kEpsilon) return;
    if (guided && c.impact_fraction > 0.0f) {
        float first_impact = 1.0f;
        for (uint32_t r = 0; r < m.count; ++r)
            if (m.contacts[r].impact_fraction > 0.0f) first_impact = fmin_(first_impact, m.contacts[r].impact_fraction);
        if (c.impact_fraction > first_impact + 1.0e-4f) return;
        if (c.impact_fraction < first_impact - 1.0e-4f) {
            uint32_t kept = 0;
            for (uint32_t r = 0; r < m.count; ++r)
                if (m.contacts[r].impact_fraction == 0.0f) m.contacts[kept++] = m.contacts[r];
            m.count = kept;
        }
    }
    float sp2 = spacing * spacing;
    for (uint32_t i = 0; i < m.count; ++i) {
        if (guided) {
            const V3 dir = guide_direction(m.contacts[i]);
            const float size = len(dir);
            if (dot(cand_dir, dir) > 0.9999f * cand_len * size) {
                if (c.penetration / cand_len > m.contacts[i].penetration / size) m.contacts[i] = c;
                return;
            }
            continue;
        }
        if (len2(sub(m.contacts[i].point, c.point)) < sp2) {
            float dd = c.penetration - m.contacts[i].penetration;
            if (dd > 0.0f) m.contacts[i] = c;
            return;
        }
    }
    if (m.count < 8u) { m.contacts[m.count++] = c; return; }
    uint32_t sh = 0;
    for (uint32_t i = 1; i < m.count; ++i)
        if (m.contacts[i].penetration < m.contacts[sh].penetration) sh = i;
    if (c.penetration > m.contacts[sh].penetration) m.contacts[sh] = c;
}

// rigid_reduce_face_patch, geometry.slang:1228-1270.
void reduce_face_patch(PmoManifold &m, float tol) {
    float t2 = tol * tol;
    bool removed = true;
    while (removed) {
        removed = false;
        for (uint32_t cand = 0; cand < m.count && !removed; ++cand) {
            V3 p = m.contacts[cand].point;
            for (uint32_t f = 0; f < m.count && !removed; ++f) {
                if (f == cand) continue;
                for (uint32_t s = f + 1; s < m.count; ++s) {
                    if (s == cand) continue;
                    V3 e = sub(m.contacts[s].point, m.contacts[f].point);
                    float e2 = len2(e);
                    if (e2 <= t2) continue;
                    float t = dot(sub(p, m.contacts[f].point), e) / e2;
                    if (t <= 1.0e-4f || t >= 0.9999f) continue;
                    V3 cl = add(m.contacts[f].point, mul(e, t));
                    if (len2(sub(p, cl)) > t2) continue;
                    for (uint32_t k = cand + 1; k < m.count; ++k) m.contacts[k - 1] = m.contacts[k];
                    --m.count;
                    removed = true;
                    break;
                }
            }
        }
    }
}

inline float rotational_motion_bound(const PmoState &prev, const PmoState &cur, const Shape &s) {  // math.slang:192-206
    Q d = relative_quaternion(prev.orientation, cur.orientation, false);
    float sh = len({d.x, d.y, d.z});
    V3 ma = {fmax_(std::fabs(s.minimum.x), std::fabs(s.maximum.x)),
             fmax_(std::fabs(s.minimum.y), std::fabs(s.maximum.y)),
             fmax_(std::fabs(s.minimum.z), std::fabs(s.maximum.z))};
    return 2.0f * len(ma) * sh;
}

bool convex_face_shape(const Shape &s) { return s.kind == kMesh && s.has_planes && s.index_count() <= 96u; }

// rigid_convex_face_manifold, geometry.slang:1606-1745 (hinge frames are
// absent on this path: both are zero).
bool convex_face_manifold(const PmoState &bs, const Shape &bm, const PmoState &cs,
                          const Shape &cm, float margin, float reach_corr, PmoManifold &out) {
    clear_manifold(out);
    float rest = fmin_(margin, 1.0e-3f);
    float reach = margin + reach_corr;
    if (!convex_face_shape(bm) || !convex_face_shape(cm)) return false;
    float best = -3.402823466e+38f;
    uint32_t ref_face = 0;
    bool ref_is_body = false;
    for (uint32_t side = 0; side < 2; ++side) {
        const Shape &rm = side == 0 ? cm : bm;
        const PmoState &rs = side == 0 ? cs : bs;
        const Shape &im = side == 0 ? bm : cm;
        const PmoState &is = side == 0 ? bs : cs;
        for (uint32_t f = 0; f < rm.index_count() / 3; ++f) {
            Plane pl = rm.planes[f];
            V3 n = rotate(rs.orientation, pl.normal);
            V3 inn = inv_rotate(is.orientation, n);
            float support = 3.402823466e+38f;
            for (uint32_t v = 0; v < im.vertex_count(); ++v) support = fmin_(support, dot(inn, im.vertices[v]));
            float sep = support + dot(n, sub(is.position, rs.position)) - pl.offset;
            if (sep > best) { best = sep; ref_face = f; ref_is_body = side != 0; }
        }
    }
    if (best > reach) return true;
    const Shape &rm = ref_is_body ? bm : cm;
    const PmoState &rs = ref_is_body ? bs : cs;
    const Shape &im = ref_is_body ? cm : bm;
    const PmoState &is = ref_is_body ? cs : bs;
    Plane reference = rm.planes[ref_face];
    V3 outward = rotate(rs.orientation, reference.normal);
    V3 iaxis = inv_rotate(is.orientation, outward);
    float alignment = 1.0f;
    uint32_t inc_face = 0;
    for (uint32_t f = 0; f < im.index_count() / 3; ++f) {
        float val = dot(iaxis, im.planes[f].normal);
        if (val < alignment) { alignment = val; inc_face = f; }
    }
    V3 normal = mul(outward, ref_is_body ? -1.0f : 1.0f);
    Plane incident = im.planes[inc_face];
    for (uint32_t t = 0; t < im.index_count() / 3; ++t) {
        Plane face = im.planes[t];
        if (dot(face.normal, incident.normal) < 0.99999f ||
            std::fabs(face.offset - incident.offset) > 1.0e-5f) continue;
        V3 poly[36], clipped[36];
        uint32_t count = 3;
        for (uint32_t c = 0; c < 3; ++c)
            poly[c] = inv_rotate(rs.orientation, sub(transform_point(is, im.vertices[im.indices[t * 3 + c]]), rs.position));
        for (uint32_t pi = 0; pi < rm.index_count() / 3 && count != 0; ++pi) {
            Plane pl = rm.planes[pi];
            float off = pl.offset + (dot(pl.normal, reference.normal) > 0.99999f ? reach : 0.0f);
            uint32_t cc = 0;
            V3 prev = poly[count - 1];
            float pd = dot(pl.normal, prev) - off;
            for (uint32_t v = 0; v < count; ++v) {
                V3 cur = poly[v];
                float d = dot(pl.normal, cur) - off;
                if ((d <= 0.0f) != (pd <= 0.0f))
                    clipped[cc++] = add(prev, mul(sub(cur, prev), pd / (pd - d)));
                if (d <= 0.0f) clipped[cc++] = cur;
                prev = cur; pd = d;
            }
            count = cc;
            for (uint32_t v = 0; v < count; ++v) poly[v] = clipped[v];
        }
        for (uint32_t v = 0; v < count; ++v) {
            float d = dot(reference.normal, poly[v]) - reference.offset;
            if (d > reach) continue;
            PmoContact c;
            std::memset(&c, 0, sizeof c);
            c.normal = normal;
            c.point = transform_point(rs, sub(poly[v], mul(reference.normal, d * 0.5f)));
            c.penetration = rest - d + 1.0e-5f;
            c.impact_fraction = 1.0f;
            add_contact(out, c, fmax_(margin * 2.0f, 1.0e-4f));
        }
    }
    if (out.count == 0u) return false;
    reduce_face_patch(out, fmax_(1.0e-6f, margin * 1.0e-3f));
    out.face_patch = 1u;
    return true;
}

// rigid_swept_convex_face_manifold, geometry.slang:1751-1853.
bool swept_convex_face_manifold(const PmoState &pb, const PmoState &b, const Shape &bm,
                                const PmoState &pc, const PmoState &c, const Shape &cm,
                                float margin, PmoManifold &out) {
    clear_manifold(out);
    Q bd = relative_quaternion(pb.orientation, b.orientation, false);
    Q cd = relative_quaternion(pc.orientation, c.orientation, false);
    if (bd.x != 0.0f || bd.y != 0.0f || bd.z != 0.0f || cd.x != 0.0f || cd.y != 0.0f || cd.z != 0.0f)
        return false;
    V3 bmo = sub(b.position, pb.position), cmo = sub(c.position, pc.position);
    float rest = fmin_(margin, 1.0e-3f);
    float entry = 0.0f, exit_time = 1.0f;
    V3 entry_normal{0, 0, 0};
    for (uint32_t side = 0; side < 2; ++side) {
        const Shape &rm = side == 0 ? cm : bm;
        const Shape &im = side == 0 ? bm : cm;
        const PmoState &rs = side == 0 ? pc : pb;
        const PmoState &is = side == 0 ? pb : pc;
        V3 rel = side == 0 ? sub(bmo, cmo) : sub(cmo, bmo);
        for (uint32_t f = 0; f < rm.index_count() / 3; ++f) {
            Plane pl = rm.planes[f];
            V3 n = rotate(rs.orientation, pl.normal);
            V3 inn = inv_rotate(is.orientation, n);
            float support = 3.402823466e+38f;
            for (uint32_t v = 0; v < im.vertex_count(); ++v) support = fmin_(support, dot(inn, im.vertices[v]));
            float gap = support + dot(n, sub(is.position, rs.position)) - pl.offset;
            float speed = dot(n, rel);
            if (speed == 0.0f) {
                if (gap > rest) return true;
            } else {
                float crossing = (rest - gap) / speed;
                if (speed < 0.0f && crossing > entry) {
                    entry = crossing;
                    entry_normal = mul(n, side == 0 ? 1.0f : -1.0f);
                }
                if (speed > 0.0f) exit_time = fmin_(exit_time, crossing);
            }
        }
    }
    if (entry > exit_time || entry > 1.0f || exit_time < 0.0f) return true;
    if (entry == 0.0f) {
        bool clipped = convex_face_manifold(b, bm, c, cm, margin, 0.0f, out);
        if (clipped && out.count != 0u) return true;
        if (!convex_face_manifold(pb, bm, pc, cm, margin, 0.0f, out) || out.count == 0u) return false;
        for (uint32_t p = 0; p < out.count; ++p) {
            float closing = -dot(sub(bmo, cmo), out.contacts[p].normal);
            if (closing <= 0.0f) return false;
            out.contacts[p].penetration += closing;
            out.contacts[p].impact_fraction = 0.0f;
        }
        return true;
    }
    PmoState ib = pb, ic = pc;
    ib.position = add(pb.position, mul(bmo, entry));
    ic.position = add(pc.position, mul(cmo, entry));
    if (!convex_face_manifold(ib, bm, ic, cm, rest + 1.0e-5f, 0.0f, out) || out.count == 0u) return false;
    for (uint32_t p = 0; p < out.count; ++p) {
        PmoContact &ct = out.contacts[p];
        if (dot(ct.normal, entry_normal) < 0.99999f) return false;
        float remaining = -dot(sub(bmo, cmo), ct.normal) * (1.0f - entry);
        ct.penetration = fmax_(0.0f, remaining + ct.penetration) + rest + 1.0e-5f;
        ct.impact_fraction = entry;
    }
    return true;
}

// rigid_requires_swept_pair_contact, geometry.slang:1586-1604.
bool requires_swept(const PmoState &pa, const PmoState &a, const Shape &ma,
                    const PmoState &pb, const PmoState &b, const Shape &mb, float threshold) {
    bool planes_a = ma.kind != kMesh || ma.has_planes;
    bool planes_b = mb.kind != kMesh || mb.has_planes;
    if (!planes_a || !planes_b) threshold = fmin_(threshold, 1.0e-3f);
    V3 rel = sub(sub(a.position, pa.position), sub(b.position, pb.position));
    return len(rel) + rotational_motion_bound(pa, a, ma) + rotational_motion_bound(pb, b, mb) > threshold;
}

// rigid_solver_correction_reach, geometry.slang:513-522.
float correction_reach(const PmoParams &p, const Shape &s) {
    if (p.motion != kDynamic) return 0.0f;
    V3 e = sub(s.maximum, s.minimum);
    return fmin_(0.5f, 0.25f * fmax_(0.0f, fmin_(e.x, fmin_(e.y, e.z))));
}

// rigid_contact_normal_speed, geometry.slang:1272-1280.
float contact_normal_speed(const PmoState &b, const PmoState &c, V3 p, V3 n) {
    V3 bv = add(b.linear_velocity, cross(b.angular_velocity, sub(p, b.position)));
    V3 cv = add(c.linear_velocity, cross(c.angular_velocity, sub(p, c.position)));
    return dot(sub(bv, cv), n);
}

// ---------------------------------------------------------------------
// Triangle-pair narrowphase (mesh pairs the convex face path does not
// take): closest features of triangle pairs, certified shell normals,
// bisection recovery and conservative advancement.  Restates
// geometry.slang:949-1117, 1272-1584, 1855-2285.  Pinned by brute force
// on small inputs (closest points against dense sampling) and closed
// forms (edge/vertex/face contacts of boxes on a floor, crossing edges,
// shell normals of closed, open and concave meshes) in the tests; the
// order in which candidates reach add_contact has no reference-held
// vector (parity unpinned beyond those).
// ---------------------------------------------------------------------
void closest_segments(V3 p0, V3 p1, V3 q0, V3 q1, V3 &on_p, V3 &on_q) {   // :949-998
    V3 d0 = sub(p1, p0), d1 = sub(q1, q0), r = sub(p0, q0);
    float a = dot(d0, d0), e = dot(d1, d1);
    float s = 0.0f, t = 0.0f;
    if (a <= 1.0e-12f && e <= 1.0e-12f) { on_p = p0; on_q = q0; return; }
    if (a <= 1.0e-12f) {
        t = clampf(dot(d1, r) / e, 0.0f, 1.0f);
    } else {
        float c = dot(d0, r);
        if (e <= 1.0e-12f) {
            s = clampf(-c / a, 0.0f, 1.0f);
        } else {
            float b = dot(d0, d1), f = dot(d1, r);
            float den = a * e - b * b;
            if (std::fabs(den) > kEpsilon * a * e) s = clampf((b * f - c * e) / den, 0.0f, 1.0f);
            t = (b * s + f) / e;
            if (t < 0.0f) { t = 0.0f; s = clampf(-c / a, 0.0f, 1.0f); }
            else if (t > 1.0f) { t = 1.0f; s = clampf((b - c) / a, 0.0f, 1.0f); }
        }
    }
    on_p = add(p0, mul(d0, s));
    on_q = add(q0, mul(d1, t));
}

V3 closest_on_triangle(V3 p, V3 a, V3 b, V3 c) {                    // :1000-1029
    V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    V3 bp = sub(p, b);
    float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return add(a, mul(ab, d1 / (d1 - d3)));
    V3 cp = sub(p, c);
    float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return add(a, mul(ac, d2 / (d2 - d6)));
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && d4 - d3 >= 0.0f && d5 - d6 >= 0.0f)
        return add(b, mul(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    V3 n = cross(ab, ac);
    return sub(p, mul(n, dot(ap, n) / len2(n)));
}

bool point_in_triangle(V3 q, V3 a, V3 b, V3 c, V3 n) {                 // :1031-1038
    float tol = -1.0e-5f * len2(n);
    return dot(cross(sub(b, a), sub(q, a)), n) >= tol &&
           dot(cross(sub(c, b), sub(q, b)), n) >= tol &&
           dot(cross(sub(a, c), sub(q, c)), n) >= tol;
}

bool segment_hits_triangle(V3 p, V3 q, V3 a, V3 b, V3 c, V3 &hit) {   // :1040-1053
    hit = {0, 0, 0};
    V3 n = cross(sub(b, a), sub(c, a));
    V3 dir = sub(q, p);
    float den = dot(n, dir);
    if (len2(n) <= kEpsilon * kEpsilon || std::fabs(den) <= kEpsilon) return false;
    float amount = dot(n, sub(a, p)) / den;
    if (amount < 0.0f || amount > 1.0f) return false;
    hit = add(p, mul(dir, amount));
    return point_in_triangle(hit, a, b, c, n);
}

inline void keep_closer(V3 a, V3 b, float &best, V3 &best_a, V3 &best_b) {   // :1055-1065
    float d = len2(sub(a, b));
    if (d < best) { best = d; best_a = a; best_b = b; }
}

void closest_triangle_pair(const V3 A[3], const V3 B[3], V3 &pa, V3 &pb) {   // :1067-1105
    pa = {0, 0, 0};
    pb = {0, 0, 0};
    V3 hit;
    for (uint32_t e = 0; e < 3; ++e)
        if (segment_hits_triangle(A[e], A[(e + 1) % 3], B[0], B[1], B[2], hit)) { pa = hit; pb = hit; return; }
    for (uint32_t e = 0; e < 3; ++e)
        if (segment_hits_triangle(B[e], B[(e + 1) % 3], A[0], A[1], A[2], hit)) { pa = hit; pb = hit; return; }
    float best = 3.402823466e+38f;
    for (uint32_t i = 0; i < 3; ++i) keep_closer(A[i], closest_on_triangle(A[i], B[0], B[1], B[2]), best, pa, pb);
    for (uint32_t i = 0; i < 3; ++i) keep_closer(closest_on_triangle(B[i], A[0], A[1], A[2]), B[i], best, pa, pb);
    for (uint32_t i = 0; i < 3; ++i)
        for (uint32_t j = 0; j < 3; ++j) {
            V3 u, w;
            closest_segments(A[i], A[(i + 1) % 3], B[j], B[(j + 1) % 3], u, w);
            keep_closer(u, w, best, pa, pb);
        }
}

inline bool boxes_touch(V3 amin, V3 amax, V3 bmin, V3 bmax) {       // math.slang:43-49
    return amin.x <= bmax.x && amax.x >= bmin.x && amin.y <= bmax.y && amax.y >= bmin.y &&
           amin.z <= bmax.z && amax.z >= bmin.z;
}

bool triangle_bounds_overlap(const V3 A[3], const V3 B[3], float margin) {   // :1107-1116
    V3 e = {margin, margin, margin};
    V3 amin = sub(cmin(A[0], cmin(A[1], A[2])), e), amax = add(cmax(A[0], cmax(A[1], A[2])), e);
    V3 bmin = cmin(B[0], cmin(B[1], B[2])), bmax = cmax(B[0], cmax(B[1], B[2]));
    return boxes_touch(amin, amax, bmin, bmax);
}

bool reaches_rest_offset(const PmoState &b, const PmoState &c, V3 point, V3 n, float distance,
                         float rest, float dt) {                       // :1282-1290
    if (distance <= rest + 1.0e-5f) return true;
    float speed = contact_normal_speed(b, c, point, n);
    if (speed > 1.0e-5f) return false;
    return -speed * dt + rest + 1.0e-5f >= distance;
}

V3 advect_triangle_point(V3 q, const V3 from[3], const V3 to[3]) {   // :1292-1304
    V3 ab = sub(from[1], from[0]), ac = sub(from[2], from[0]);
    V3 n = cross(ab, ac);
    float sq = len2(n);
    if (sq <= 1.0e-20f) return to[0];
    V3 rel = sub(q, from[0]);
    float v = dot(cross(rel, ac), n) / sq;
    float w = dot(cross(ab, rel), n) / sq;
    return add(to[0], add(mul(sub(to[1], to[0]), v), mul(sub(to[2], to[0]), w)));
}

// rigid_shell_outward, :1308-1322: the certified normal's side applied to
// the (possibly swept) world triangle; unit length for certified exterior
// faces, a quarter for two-sided interior features, zero otherwise.
V3 shell_outward(const Shape &m, uint32_t t, const V3 W[3]) {
    if (m.shell.empty()) return {0, 0, 0};
    V3 o = m.shell[t];
    float sq = len2(o);
    if (sq < 0.01f) return {0, 0, 0};
    V3 la = m.vertices[m.indices[3 * t]], lb = m.vertices[m.indices[3 * t + 1]], lc = m.vertices[m.indices[3 * t + 2]];
    float sign = dot(cross(sub(lb, la), sub(lc, la)), o) >= 0.0f ? 1.0f : -1.0f;
    return mul(normalized_or(cross(sub(W[1], W[0]), sub(W[2], W[0])), {0, 0, 0}), sign * (sq > 0.5f ? 1.0f : 0.25f));
}

bool shell_facing(V3 n, V3 oa, V3 ob) {                              // :1324-1337
    float sa = len2(oa), sb = len2(ob);
    return (sa < 0.01f || (sa > 0.5f ? dot(n, oa) < -1.0e-3f : std::fabs(dot(n, oa)) > 2.5e-4f)) &&
           (sb < 0.01f || (sb > 0.5f ? dot(n, ob) > 1.0e-3f : std::fabs(dot(n, ob)) > 2.5e-4f));
}

bool point_inside_convex(const Shape &m, const PmoState &st, V3 world) {   // :1339-1357
    if (!m.has_planes) return false;
    V3 l = inv_rotate(st.orientation, sub(world, st.position));
    uint32_t tri = m.index_count() / 3u;
    uint32_t count = m.unique_planes != 0u ? m.unique_planes : tri;
    uint32_t first = m.unique_planes != 0u ? tri : 0u;
    for (uint32_t f = 0; f < count; ++f) {
        const Plane &pl = m.planes[first + f];
        if (dot(pl.normal, l) > pl.offset + 1.0e-5f) return false;
    }
    return true;
}

V3 triangle_feature_normal(const V3 A[3], const V3 B[3], V3 pa, V3 pb, V3 fallback) {   // :1855-1908
    V3 delta = sub(pa, pb);
    V3 normal = normalized_or(delta, fallback);
    V3 faces[2] = {normalized_or(cross(sub(A[1], A[0]), sub(A[2], A[0])), normal),
                   normalized_or(cross(sub(B[1], B[0]), sub(B[2], B[0])), normal)};
    const V3 *tris[2] = {A, B};
    V3 pts[2] = {pa, pb};
    float align_best = -1.0f;
    V3 interior_normal = {0, 0, 0};
    for (uint32_t side = 0; side < 2; ++side) {
        const V3 *T = tris[side];
        V3 raw = cross(sub(T[1], T[0]), sub(T[2], T[0]));
        float tol = 1.0e-5f * len2(raw);
        bool interior = tol > 0.0f;
        for (uint32_t e = 0; e < 3; ++e)
            interior = interior && dot(cross(sub(T[(e + 1) % 3], T[e]), sub(pts[side], T[e])), raw) > tol;
        float alignment = std::fabs(dot(faces[side], normal));
        if (interior && alignment > align_best) {
            align_best = alignment;
            float proj = dot(delta, faces[side]);
            float dir = std::fabs(proj) > kEpsilon ? proj : dot(fallback, faces[side]);
            interior_normal = dir >= 0.0f ? faces[side] : mul(faces[side], -1.0f);
        }
    }
    if (align_best >= 0.0f) return interior_normal;
    float best = 4.0f * kEpsilon * kEpsilon;
    if (len2(delta) > kEpsilon * kEpsilon) {
        for (uint32_t side = 0; side < 2; ++side) {
            float proj = dot(delta, faces[side]);
            float err = len2(sub(delta, mul(faces[side], proj)));
            if (err < best) {
                best = err;
                normal = proj >= 0.0f ? faces[side] : mul(faces[side], -1.0f);
            }
        }
    }
    return normal;
}

// The per-pair inputs of the triangle path: shapes, previous and predicted
// states, the hinge contact frames of both bodies at the predicted states
// (rigid_hinge_contact_frame, geometry.slang:404-491) and the body's
// reference point at the previous and predicted states (its hinge anchor
// when it has a non-axial frame, else its position).
struct TrianglePair {
    const Shape *am, *bm;
    PmoState ap, a, bp, b;          // previous and predicted states
    PmoHingeFrame ha, hb;
    float margin, dt;
    V3 ref_prev, ref;
};

inline void world_triangle(const Shape &m, const PmoState &st, uint32_t t, V3 W[3]) {
    for (uint32_t c = 0; c < 3; ++c) W[c] = transform_point(st, m.vertices[m.indices[3 * t + c]]);
}

// rigid_collide_triangle_ranges, geometry.slang:1359-1584.
void collide_triangle_ranges(const TrianglePair &P, uint32_t a_first, uint32_t a_count, uint32_t b_first,
                             uint32_t b_count, PmoManifold &m) {
    const Shape &am = *P.am, &bm = *P.bm;
    if (guided_pair(P.ha, P.hb)) return;
    const float margin = P.margin;
    const float rest = fmin_(margin, 1.0e-3f);
    const V3 body_ref = P.ref;
    const bool inside_a = P.ha.static_body != 0u && point_inside_convex(am, P.ap, P.bp.position);
    const bool inside_b = P.hb.static_body != 0u && point_inside_convex(bm, P.bp, P.ap.position);
    for (uint32_t ai = 0; ai < a_count; ++ai) {
        const uint32_t at = a_first + ai;
        if (3u * at + 2u >= am.index_count()) break;
        V3 A[3];
        world_triangle(am, P.a, at, A);
        for (uint32_t bi = 0; bi < b_count; ++bi) {
            const uint32_t bt = b_first + bi;
            if (3u * bt + 2u >= bm.index_count()) break;
            V3 B[3];
            world_triangle(bm, P.b, bt, B);
            if (!triangle_bounds_overlap(A, B, margin)) continue;
            V3 oa = shell_outward(am, at, A), ob = shell_outward(bm, bt, B);
            V3 coa = P.ha.static_body != 0u && inside_a ? V3{0, 0, 0} : oa;
            V3 cob = P.hb.static_body != 0u && inside_b ? V3{0, 0, 0} : ob;
            if (dot(oa, ob) > 0.9999f) continue;
            V3 pa, pb;
            closest_triangle_pair(A, B, pa, pb);
            V3 delta = sub(pa, pb);
            float sq = len2(delta);
            if (sq > margin * margin) continue;
            V3 cn = normalized_or(cross(sub(B[1], B[0]), sub(B[2], B[0])), {0.0f, 1.0f, 0.0f});
            V3 fallback = dot(cn, sub(body_ref, pb)) >= 0.0f ? cn : mul(cn, -1.0f);
            if (len2(cob) > 0.5f) fallback = cob;
            float distance = std::sqrt(fmax_(sq, 0.0f));
            V3 point = mul(add(pa, pb), 0.5f);
            V3 proj = sub(pa, mul(cn, dot(sub(pa, B[0]), cn)));
            bool on_face = len2(sub(proj, closest_on_triangle(proj, B[0], B[1], B[2]))) <= 1.0e-10f;
            V3 normal = (distance <= 1.0e-5f || on_face) ? fallback : normalized_or(delta, fallback);
            if (on_face && len2(oa) < 0.01f && len2(ob) > 0.5f && bm.has_planes) {
                V3 face_a = normalized_or(cross(sub(A[1], A[0]), sub(A[2], A[0])), {0, 0, 0});
                if (std::fabs(dot(face_a, normal)) < 1.0e-3f) {
                    float depth = fmax_(0.0f, -fmin_(dot(sub(A[0], B[0]), normal),
                                                     fmin_(dot(sub(A[1], B[0]), normal), dot(sub(A[2], B[0]), normal))));
                    V3 axis = inv_rotate(P.b.orientation, face_a);
                    float offset = dot(sub(P.b.position, A[0]), face_a);
                    float lo = 3.402823466e+38f, hi = -3.402823466e+38f;
                    for (uint32_t v = 0; v < bm.vertex_count(); ++v) {
                        float pr = dot(bm.vertices[v], axis) + offset;
                        lo = fmin_(lo, pr);
                        hi = fmax_(hi, pr);
                    }
                    float alternate = fmax_(0.0f, fmin_(-lo, hi));
                    if (alternate + 1.0e-5f < depth) continue;
                }
            }
            float separation = on_face && len2(cob) > 0.5f ? dot(delta, normal) : distance;
            if (!reaches_rest_offset(P.a, P.b, point, normal, separation, rest, P.dt)) continue;
            PmoContact c;
            std::memset(&c, 0, sizeof c);
            c.normal = normal;
            c.point = point;
            c.body_hinge = P.ha;
            c.collider_hinge = P.hb;
            c.penetration = rest - separation + 1.0e-5f;
            bool recovered = false;
            if (!am.has_planes && !bm.has_planes && distance <= 1.0e-5f) {
                V3 oldA[3], oldB[3];
                world_triangle(am, P.ap, at, oldA);
                world_triangle(bm, P.bp, bt, oldB);
                V3 opa, opb;
                closest_triangle_pair(oldA, oldB, opa, opb);
                V3 od = sub(opa, opb);
                if (len2(od) > 1.0e-10f) {
                    V3 IA[3], IB[3];
                    float lo = 0.0f, hi = 1.0f;
                    float gap = fmin_(rest, 0.5f * std::sqrt(len2(od)));
                    for (uint32_t k = 0; k < 10u; ++k) {
                        float time = 0.5f * (lo + hi);
                        for (uint32_t v = 0; v < 3; ++v) {
                            IA[v] = add(oldA[v], mul(sub(A[v], oldA[v]), time));
                            IB[v] = add(oldB[v], mul(sub(B[v], oldB[v]), time));
                        }
                        V3 u, w;
                        closest_triangle_pair(IA, IB, u, w);
                        if (len2(sub(u, w)) > gap * gap) lo = time;
                        else hi = time;
                    }
                    for (uint32_t v = 0; v < 3; ++v) {
                        IA[v] = add(oldA[v], mul(sub(A[v], oldA[v]), lo));
                        IB[v] = add(oldB[v], mul(sub(B[v], oldB[v]), lo));
                    }
                    closest_triangle_pair(IA, IB, opa, opb);
                    od = sub(opa, opb);
                    c.normal = triangle_feature_normal(IA, IB, opa, opb, normalized_or(od, fallback));
                    V3 ca = advect_triangle_point(opa, IA, A), cb = advect_triangle_point(opb, IB, B);
                    c.point = mul(add(ca, cb), 0.5f);
                    c.penetration = rest - dot(sub(ca, cb), c.normal) + 1.0e-5f;
                    recovered = true;
                }
            }
            if (distance <= 1.0e-5f) {
                if (!recovered && (P.ha.fixed_member != 0u || P.hb.fixed_member != 0u || on_face)) {
                    float depth = fmax_(0.0f, -fmin_(dot(sub(A[0], pb), normal),
                                                     fmin_(dot(sub(A[1], pb), normal), dot(sub(A[2], pb), normal))));
                    c.penetration = fmin_(margin, depth + rest) + 1.0e-5f;
                } else if (!recovered) {
                    c.penetration = margin + 1.0e-5f;
                }
            }
            if (!shell_facing(c.normal, coa, cob)) continue;
            c.impact_fraction = 1.0f;
            c.initial_normal_speed = contact_normal_speed(P.a, P.b, c.point, c.normal);
            add_contact(m, c, fmax_(2.0f * margin, 1.0e-4f));
        }
    }
}

// rigid_triangle_pair_face_contact, geometry.slang:1910-1919: the
// separation lies along a face normal of either triangle.
inline bool triangle_pair_face_contact(const V3 A[3], const V3 B[3], V3 separation) {
    V3 n = normalized_or(separation, {0, 0, 0});
    V3 fa = normalized_or(cross(sub(A[1], A[0]), sub(A[2], A[0])), {0, 0, 0});
    V3 fb = normalized_or(cross(sub(B[1], B[0]), sub(B[2], B[0])), {0, 0, 0});
    return std::fabs(dot(n, fa)) > 0.9999f || std::fabs(dot(n, fb)) > 0.9999f;
}

// rigid_collide_triangle_ranges_swept, geometry.slang:1921-2154.  A guided
// pair rests at 0.1 mm, takes edge contacts that both slide along and turn
// about the guide axis at zero offset, and advances by the faces' closing
// speed along the feature normal.
void collide_triangle_ranges_swept(const TrianglePair &P, uint32_t a_first, uint32_t a_count, uint32_t b_first,
                                   uint32_t b_count, PmoManifold &m) {
    const Shape &am = *P.am, &bm = *P.bm;
    const float margin = P.margin;
    const bool guided = guided_pair(P.ha, P.hb);
    const float rest = guided ? fmin_(margin, 1.0e-4f) : fmin_(margin, 1.0e-3f);
    const bool inside_a = P.ha.static_body != 0u && point_inside_convex(am, P.ap, P.bp.position);
    const bool inside_b = P.hb.static_body != 0u && point_inside_convex(bm, P.bp, P.ap.position);
    for (uint32_t ai = 0; ai < a_count; ++ai) {
        const uint32_t at = a_first + ai;
        if (3u * at + 2u >= am.index_count()) break;
        V3 pA[3], cA[3];
        world_triangle(am, P.ap, at, pA);
        world_triangle(am, P.a, at, cA);
        V3 amin = cmin(pA[0], cA[0]), amax = cmax(pA[0], cA[0]);
        for (uint32_t v = 1; v < 3; ++v) {
            amin = cmin(amin, cmin(pA[v], cA[v]));
            amax = cmax(amax, cmax(pA[v], cA[v]));
        }
        for (uint32_t bi = 0; bi < b_count; ++bi) {
            const uint32_t bt = b_first + bi;
            if (3u * bt + 2u >= bm.index_count()) break;
            V3 pB[3], cB[3];
            world_triangle(bm, P.bp, bt, pB);
            world_triangle(bm, P.b, bt, cB);
            V3 bmin = cmin(pB[0], cB[0]), bmax = cmax(pB[0], cB[0]);
            for (uint32_t v = 1; v < 3; ++v) {
                bmin = cmin(bmin, cmin(pB[v], cB[v]));
                bmax = cmax(bmax, cmax(pB[v], cB[v]));
            }
            V3 e = {margin, margin, margin};
            if (!boxes_touch(sub(amin, e), add(amax, e), bmin, bmax)) continue;
            V3 dA[3], dB[3];
            float speed_abs = 0.0f;
            for (uint32_t v = 0; v < 3; ++v) {
                dA[v] = sub(cA[v], pA[v]);
                dB[v] = sub(cB[v], pB[v]);
                speed_abs = fmax_(speed_abs, len(dA[v]));
            }
            float speed_b = 0.0f;
            for (uint32_t v = 0; v < 3; ++v) speed_b = fmax_(speed_b, len(dB[v]));
            speed_abs += speed_b;
            V3 common = dB[0];
            float rel_a = 0.0f, rel_b = 0.0f;
            for (uint32_t v = 0; v < 3; ++v) {
                rel_a = fmax_(rel_a, len(sub(dA[v], common)));
                rel_b = fmax_(rel_b, len(sub(dB[v], common)));
            }
            float bound = fmin_(speed_abs, rel_a + rel_b);
            if (bound <= kEpsilon) continue;
            float time = 0.0f;
            bool starts_near = false;
            V3 sep_normal = {0, 0, 0};
            bool have_sep = false;
            for (uint32_t it = 0; it < 32u; ++it) {
                V3 A[3], B[3];
                for (uint32_t v = 0; v < 3; ++v) {
                    A[v] = add(pA[v], mul(dA[v], time));
                    B[v] = add(pB[v], mul(dB[v], time));
                }
                V3 pa, pb;
                closest_triangle_pair(A, B, pa, pb);
                V3 delta = sub(pa, pb);
                float distance = std::sqrt(fmax_(0.0f, len2(delta)));
                float offset = rest;
                if (guided && !triangle_pair_face_contact(A, B, delta)) {
                    const PmoHingeFrame &guide = P.ha.axial != 0u ? P.ha : P.hb;
                    V3 direction = normalized_or(delta, {0, 0, 0});
                    float axial = dot(direction, guide.axis);
                    float angular = dot(guide.axis, cross(sub(pa, guide.anchor), direction));
                    if (std::fabs(axial) > 1.0e-3f && std::fabs(angular) > 1.0e-3f) offset = 0.0f;
                }
                if (it == 0u) starts_near = distance <= offset + 5.0e-5f;
                if (distance <= offset + 1.0e-5f) {
                    V3 cn = normalized_or(cross(sub(B[1], B[0]), sub(B[2], B[0])), {0.0f, 1.0f, 0.0f});
                    V3 center = add(P.ap.position, mul(sub(P.a.position, P.ap.position), time));
                    V3 fallback = dot(cn, sub(center, pb)) >= 0.0f ? cn : mul(cn, -1.0f);
                    if (P.ha.present != 0u) {
                        V3 reference = add(P.ref_prev, mul(sub(P.ref, P.ref_prev), time));
                        fallback = dot(cn, sub(reference, pb)) >= 0.0f ? cn : mul(cn, -1.0f);
                    }
                    V3 oa = shell_outward(am, at, A), ob = shell_outward(bm, bt, B);
                    V3 coa = P.ha.static_body != 0u && inside_a ? V3{0, 0, 0} : oa;
                    V3 cob = P.hb.static_body != 0u && inside_b ? V3{0, 0, 0} : ob;
                    if (dot(oa, ob) > 0.9999f) break;
                    if (len2(cob) > 0.5f) fallback = cob;
                    else if (len2(coa) > 0.5f) fallback = mul(coa, -1.0f);
                    V3 normal = triangle_feature_normal(A, B, pa, pb, have_sep ? sep_normal : fallback);
                    if (!shell_facing(normal, coa, cob)) break;
                    V3 point = mul(add(pa, pb), 0.5f);
                    float vn = contact_normal_speed(P.a, P.b, point, normal);
                    if (vn > 1.0e-5f) break;
                    float remaining = -vn * P.dt * (1.0f - time) - distance;
                    PmoContact c;
                    std::memset(&c, 0, sizeof c);
                    c.normal = normal;
                    c.point = point;
                    c.penetration = fmax_(0.0f, remaining + offset) + 1.0e-5f;
                    c.body_hinge = P.ha;
                    c.collider_hinge = P.hb;
                    c.impact_fraction = starts_near ? 0.0f : time;
                    c.initial_normal_speed = vn;
                    add_contact(m, c, fmax_(margin * 2.0f, 1.0e-4f));
                    break;
                }
                sep_normal = triangle_feature_normal(A, B, pa, pb, normalized_or(delta, {0.0f, 1.0f, 0.0f}));
                have_sep = true;
                float advance = (distance - offset) / (bound + kEpsilon) * 0.9f;
                if (guided) {
                    V3 plane = triangle_feature_normal(A, B, pa, pb, sep_normal);
                    float min_a = 3.402823466e+38f, max_b = -3.402823466e+38f;
                    float min_speed_a = 3.402823466e+38f, max_speed_b = -3.402823466e+38f;
                    for (uint32_t v = 0; v < 3; ++v) {
                        min_a = fmin_(min_a, dot(sub(A[v], pb), plane));
                        max_b = fmax_(max_b, dot(sub(B[v], pb), plane));
                        min_speed_a = fmin_(min_speed_a, dot(dA[v], plane));
                        max_speed_b = fmax_(max_speed_b, dot(dB[v], plane));
                    }
                    float gap = min_a - max_b - offset;
                    float closing = max_speed_b - min_speed_a;
                    if (gap > 1.0e-5f && closing <= 0.0f) break;
                    if (gap > 0.0f && closing > kEpsilon) advance = fmax_(advance, 0.9f * gap / closing);
                }
                time += fmax_(advance, 1.0e-5f);
                if (time > 1.0f) break;
            }
        }
    }
}

// BoundsTransform and transformed_(motion_)bounds, math.slang:142-190.
struct BoundsFrame { V3 position, ax, ay, az; };
inline BoundsFrame bounds_frame(const PmoState &s) {
    return {s.position, rotate(s.orientation, {1, 0, 0}), rotate(s.orientation, {0, 1, 0}),
            rotate(s.orientation, {0, 0, 1})};
}
inline void frame_bounds(V3 lmin, V3 lmax, const BoundsFrame &f, float margin, V3 &mn, V3 &mx) {
    V3 c = mul(add(lmin, lmax), 0.5f), h = mul(sub(lmax, lmin), 0.5f);
    V3 wc = add(f.position, add(mul(f.ax, c.x), add(mul(f.ay, c.y), mul(f.az, c.z))));
    V3 wh = {std::fabs(f.ax.x) * h.x + std::fabs(f.ay.x) * h.y + std::fabs(f.az.x) * h.z,
             std::fabs(f.ax.y) * h.x + std::fabs(f.ay.y) * h.y + std::fabs(f.az.y) * h.z,
             std::fabs(f.ax.z) * h.x + std::fabs(f.ay.z) * h.y + std::fabs(f.az.z) * h.z};
    V3 e = {margin, margin, margin};
    mn = sub(sub(wc, wh), e);
    mx = add(add(wc, wh), e);
}
inline void motion_bounds(V3 lmin, V3 lmax, const BoundsFrame &prev, const BoundsFrame &cur, bool swept,
                          float margin, V3 &mn, V3 &mx) {
    frame_bounds(lmin, lmax, cur, margin, mn, mx);
    if (!swept) return;
    V3 pmn, pmx;
    frame_bounds(lmin, lmax, prev, margin, pmn, pmx);
    mn = cmin(mn, pmn);
    mx = cmax(mx, pmx);
}

// One leaf range pair: the discrete pass, then the swept pass when the
// pair moves past the margin.  Shared by the leaf-pair evaluation and the
// BVH traversal.
inline void collide_leaf_ranges(const TrianglePair &P, bool swept, const BvhNode &na, const BvhNode &nb,
                                PmoManifold &m) {
    collide_triangle_ranges(P, na.first, na.count, nb.first, nb.count, m);
    if (swept) collide_triangle_ranges_swept(P, na.first, na.count, nb.first, nb.count, m);
}

// rigid_collide_meshes, geometry.slang:2156-2285: simultaneous BVH descent
// with a 64-entry stack, all triangles when the stack would overflow.
void collide_meshes(const TrianglePair &P, bool swept, PmoManifold &m) {
    clear_manifold(m);
    const Shape &am = *P.am, &bm = *P.bm;
    BoundsFrame fa = bounds_frame(P.a), fb = bounds_frame(P.b);
    BoundsFrame pfa = swept ? bounds_frame(P.ap) : BoundsFrame{};
    BoundsFrame pfb = swept ? bounds_frame(P.bp) : BoundsFrame{};
    V3 amn, amx, bmn, bmx;
    motion_bounds(am.minimum, am.maximum, pfa, fa, swept, P.margin, amn, amx);
    motion_bounds(bm.minimum, bm.maximum, pfb, fb, swept, 0.0f, bmn, bmx);
    if (!boxes_touch(amn, amx, bmn, bmx)) return;
    constexpr uint32_t kStack = 64;
    uint32_t sa[kStack] = {}, sb[kStack] = {};
    uint32_t top = 1;
    bool overflow = am.nodes.empty() || bm.nodes.empty();
    while (top > 0u && !overflow) {
        --top;
        const uint32_t ia = sa[top], ib = sb[top];
        const BvhNode &na = am.nodes[ia], &nb = bm.nodes[ib];
        motion_bounds(na.minimum, na.maximum, pfa, fa, swept, P.margin, amn, amx);
        motion_bounds(nb.minimum, nb.maximum, pfb, fb, swept, 0.0f, bmn, bmx);
        if (!boxes_touch(amn, amx, bmn, bmx)) continue;
        const bool la = na.count != 0u, lb = nb.count != 0u;
        if (la && lb) {
            collide_leaf_ranges(P, swept, na, nb, m);
            continue;
        }
        const uint32_t need = la || lb ? 2u : 4u;
        if (top + need > kStack) { overflow = true; break; }
        if (la) {
            sa[top] = ia; sb[top++] = nb.right;
            sa[top] = ia; sb[top++] = nb.left;
        } else if (lb) {
            sa[top] = na.right; sb[top++] = ib;
            sa[top] = na.left; sb[top++] = ib;
        } else {
            sa[top] = na.right; sb[top++] = nb.right;
            sa[top] = na.right; sb[top++] = nb.left;
            sa[top] = na.left; sb[top++] = nb.right;
            sa[top] = na.left; sb[top++] = nb.left;
        }
    }
    if (overflow) {
        clear_manifold(m);
        collide_triangle_ranges(P, 0, am.index_count() / 3u, 0, bm.index_count() / 3u, m);
        if (swept) collide_triangle_ranges_swept(P, 0, am.index_count() / 3u, 0, bm.index_count() / 3u, m);
    }
}

// pm_generate_rigid_leaf_pairs (geometry.slang:2303-2492): overlapping
// (body leaf, collider leaf) pairs of motion bounds.  128 workers each take
// every 128th entry of the row-major leaf-pair grid and emit their hits
// contiguously in worker order; that order is the order in which
// candidates later meet add_contact.  Returns false when the pair has more
// than `limit` candidates.
bool leaf_candidates(const TrianglePair &P, bool swept, uint32_t limit,
                     std::vector<std::pair<uint32_t, uint32_t>> &out) {
    out.clear();
    const Shape &am = *P.am, &bm = *P.bm;
    BoundsFrame fa = bounds_frame(P.a), fb = bounds_frame(P.b);
    BoundsFrame pfa = swept ? bounds_frame(P.ap) : BoundsFrame{};
    BoundsFrame pfb = swept ? bounds_frame(P.bp) : BoundsFrame{};
    const uint32_t la = uint32_t(am.leaves.size()), lb = uint32_t(bm.leaves.size());
    const uint64_t cells = uint64_t(la) * lb;
    constexpr uint32_t kWorkers = 128;
    for (uint32_t w = 0; w < kWorkers; ++w)
        for (uint64_t f = w; f < cells; f += kWorkers) {
            const BvhNode &na = am.nodes[am.leaves[f / lb]], &nb = bm.nodes[bm.leaves[f % lb]];
            V3 amn, amx, bmn, bmx;
            motion_bounds(na.minimum, na.maximum, pfa, fa, swept, P.margin, amn, amx);
            motion_bounds(nb.minimum, nb.maximum, pfb, fb, swept, 0.0f, bmn, bmx);
            if (boxes_touch(amn, amx, bmn, bmx)) out.push_back({am.leaves[f / lb], bm.leaves[f % lb]});
        }
    return out.size() <= limit;
}

// ---------------------------------------------------------------------
// Analytic spheres and capsules: no reference counterpart.  Contacts
// follow the face path's conventions: the normal points from the collider
// towards the body, the point is the midpoint of the two surface points,
// penetration = rest offset - separation + 1e-5.  Pinned by closed forms
// (sphere on plane / sphere on sphere / box corner) in the tests.
// ---------------------------------------------------------------------
bool box_half_extents(const Shape &s, V3 &half) {
    if (s.kind != kMesh || !s.has_planes || s.vertex_count() != 8u) return false;
    V3 c = s.center;
    if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f) return false;
    half = mul(sub(s.maximum, s.minimum), 0.5f);
    for (auto &p : s.vertices)
        if (std::fabs(p.x) != half.x || std::fabs(p.y) != half.y || std::fabs(p.z) != half.z) return false;
    return true;
}

// Closest point of segment [p0,p1] to point q.
inline V3 closest_on_segment(V3 p0, V3 p1, V3 q) {
    V3 d = sub(p1, p0);
    float dd = len2(d);
    float t = dd > 0.0f ? clampf(dot(sub(q, p0), d) / dd, 0.0f, 1.0f) : 0.0f;
    return add(p0, mul(d, t));
}

// A sphere of radius r centred at world point c against a shape; returns
// the normal (from the shape towards c), the surface point on the shape
// and the signed distance from the shape surface to c.  Boxes, spheres
// and capsules are supported.
bool point_shape_distance(V3 c, const PmoState &ss, const Shape &sh, V3 &normal,
                          V3 &surface, float &distance) {
    if (sh.kind == kSphere) {
        V3 d = sub(c, ss.position);
        normal = normalized_or(d, {0.0f, 1.0f, 0.0f});
        distance = len(d) - sh.sphere_radius;
        surface = add(ss.position, mul(normal, sh.sphere_radius));
        return true;
    }
    if (sh.kind == kCapsule) {
        V3 axis = rotate(ss.orientation, {0.0f, sh.half_length, 0.0f});
        V3 p = closest_on_segment(sub(ss.position, axis), add(ss.position, axis), c);
        V3 d = sub(c, p);
        normal = normalized_or(d, rotate(ss.orientation, {1.0f, 0.0f, 0.0f}));
        distance = len(d) - sh.sphere_radius;
        surface = add(p, mul(normal, sh.sphere_radius));
        return true;
    }
    V3 half;
    if (!box_half_extents(sh, half)) return false;
    V3 l = inv_rotate(ss.orientation, sub(c, ss.position));
    V3 q = {clampf(l.x, -half.x, half.x), clampf(l.y, -half.y, half.y), clampf(l.z, -half.z, half.z)};
    bool inside = q.x == l.x && q.y == l.y && q.z == l.z;
    V3 ln;
    float dist;
    if (!inside) {
        V3 d = sub(l, q);
        float dl = len(d);
        ln = mul(d, 1.0f / dl);
        dist = dl;
    } else {
        // Deepest-face exit: the axis with the least remaining depth.
        float dx = half.x - std::fabs(l.x), dy = half.y - std::fabs(l.y), dz = half.z - std::fabs(l.z);
        if (dx <= dy && dx <= dz) { ln = {l.x < 0.0f ? -1.0f : 1.0f, 0, 0}; q.x = l.x < 0.0f ? -half.x : half.x; dist = -dx; }
        else if (dy <= dz) { ln = {0, l.y < 0.0f ? -1.0f : 1.0f, 0}; q.y = l.y < 0.0f ? -half.y : half.y; dist = -dy; }
        else { ln = {0, 0, l.z < 0.0f ? -1.0f : 1.0f}; q.z = l.z < 0.0f ? -half.z : half.z; dist = -dz; }
    }
    normal = rotate(ss.orientation, ln);
    surface = transform_point(ss, q);
    distance = dist;
    return true;
}

void push_sphere_contact(PmoManifold &m, V3 normal, V3 surface_collider, float gap,
                         float margin, float reach) {
    if (gap > reach) return;
    float rest = fmin_(margin, 1.0e-3f);
    PmoContact c;
    std::memset(&c, 0, sizeof c);
    c.normal = normal;
    c.point = add(surface_collider, mul(normal, gap * 0.5f));
    c.penetration = rest - gap + 1.0e-5f;
    c.impact_fraction = 1.0f;
    add_contact(m, c, fmax_(margin * 2.0f, 1.0e-4f));
}

// Round shapes (sphere, capsule) against anything supported.  Returns
// false when the pair has no analytic path.
bool analytic_manifold(const PmoState &as, const Shape &am, const PmoState &bs, const Shape &bm,
                       float margin, float reach, PmoManifold &out) {
    clear_manifold(out);
    bool a_round = am.kind == kSphere || am.kind == kCapsule;
    bool b_round = bm.kind == kSphere || bm.kind == kCapsule;
    if (!a_round && !b_round) return false;
    // Sample points on the round shape's core: the centre for a sphere,
    // the segment points closest to the other shape for a capsule.
    V3 n, s;
    float d;
    if (am.kind == kSphere) {
        if (!point_shape_distance(as.position, bs, bm, n, s, d)) return false;
        push_sphere_contact(out, n, s, d - am.sphere_radius, margin, reach);
        return true;
    }
    if (bm.kind == kSphere) {
        if (!point_shape_distance(bs.position, as, am, n, s, d)) return false;
        // Normal must point from the collider (b) towards the body (a).
        V3 nb = mul(n, -1.0f);
        V3 sb = sub(bs.position, mul(n, bm.sphere_radius));
        (void)s;
        push_sphere_contact(out, nb, sb, d - bm.sphere_radius, margin, reach);
        return true;
    }
    // Capsule body against a box or capsule collider: both segment end
    // points plus, for capsule-capsule, the mutual closest point.
    const Shape &cap = am.kind == kCapsule ? am : bm;
    const PmoState &cs = am.kind == kCapsule ? as : bs;
    const Shape &other = am.kind == kCapsule ? bm : am;
    const PmoState &os = am.kind == kCapsule ? bs : as;
    float sign = am.kind == kCapsule ? 1.0f : -1.0f;
    V3 axis = rotate(cs.orientation, {0.0f, cap.half_length, 0.0f});
    V3 ends[3] = {sub(cs.position, axis), add(cs.position, axis), cs.position};
    uint32_t samples = 2;
    if (other.kind == kCapsule) {
        V3 oaxis = rotate(os.orientation, {0.0f, other.half_length, 0.0f});
        ends[2] = closest_on_segment(ends[0], ends[1], closest_on_segment(sub(os.position, oaxis), add(os.position, oaxis), cs.position));
        samples = 3;
    }
    for (uint32_t k = 0; k < samples; ++k) {
        if (!point_shape_distance(ends[k], os, other, n, s, d)) return false;
        float gap = d - cap.sphere_radius;
        if (sign > 0.0f) push_sphere_contact(out, n, s, gap, margin, reach);
        else push_sphere_contact(out, mul(n, -1.0f), sub(ends[k], mul(n, cap.sphere_radius)), gap, margin, reach);
    }
    return true;
}

// ---------------------------------------------------------------------
// AVBD prepare and solve.  geometry_constraints.slang:374-445,
// 1213-1591 and geometry_constraints_solve.slang:211-865.  Pinned by the
// kernel-level physical gates of tests/solver/rigid_realtime_cuda_tests.cu
// re-run on this oracle (tests/test_oracle_physics.py) and by the
// invariants listed there; contact preparation, the crossing safeguard and
// free_mean have no reference-held vectors (parity unpinned beyond those).
// ---------------------------------------------------------------------
constexpr float kSurfaceTolerance = 1.0e-5f;   // geometry_constraints.slang:28
constexpr float kRestOffset = 1.0e-3f;         // :29
constexpr uint32_t kRealtimeProfile = 1u;      // :32
constexpr uint32_t kRealtimeSweeps = 8u;       // :33
constexpr float kRealtimeDamping = 0.05f;      // :34

inline Vec6 jac(V3 l, V3 a) { return {{l.x, l.y, l.z, a.x, a.y, a.z}}; }
inline V3 trans(const float *d) { return {d[0], d[1], d[2]}; }
inline V3 rot6(const float *d) { return {d[3], d[4], d[5]}; }
inline Vec6 disp(const PmoAvbdBody &b) { Vec6 v; std::memcpy(v.v, b.displacement, sizeof v.v); return v; }

inline void basis(V3 n, V3 &t0, V3 &t1) {                        // :424-430
    t0 = normalized_or(cross(n, std::fabs(n.x) < 0.57735f ? V3{1, 0, 0} : V3{0, 1, 0}), {0, 0, 1});
    t1 = cross(n, t0);
}
inline float comp(V3 v, uint32_t a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
inline float pair_stiffness(const PmoParams &a, const PmoParams &b, float dt) {  // :432-436
    return 1.0f / (fmax_(a.inverse_mass + b.inverse_mass, 1.0e-6f) * dt * dt);
}
constexpr float kGuidedRestOffset = 1.0e-4f;     // geometry_constraints.slang:31
// rigid_avbd_contact_skin, :438-445: a guided contact rests at 0.1 mm.
inline float contact_skin(const PmoContact &c, const PmoParams &a, const PmoParams &b) {
    const float offset = guided_contact(c) ? fmin_(a.collision_margin + b.collision_margin, kGuidedRestOffset)
                                           : fmin_(a.collision_margin + b.collision_margin, kRestOffset);
    return offset + 2.0f * kSurfaceTolerance;
}

PmoState avbd_pose(const PmoAvbdBody &b) {                       // :374-384
    PmoState s = b.target;
    s.position = add(s.position, trans(b.displacement));
    Q d = b.orientation_delta;
    if (d.x != 0.0f || d.y != 0.0f || d.z != 0.0f)
        s.orientation = qnormalize(qmul(d, s.orientation));
    return s;
}

void apply_pose_update(PmoAvbdBody &b, const Vec6 &u) {          // :386-408
    for (int i = 0; i < 3; ++i) b.displacement[i] += u.v[i];
    V3 r = {u.v[3], u.v[4], u.v[5]};
    if (r.x == 0.0f && r.y == 0.0f && r.z == 0.0f) return;
    Q rel = b.orientation_delta;
    Q der = qmul({r.x, r.y, r.z, 0.0f}, rel);
    rel = qnormalize({rel.x + 0.5f * der.x, rel.y + 0.5f * der.y, rel.z + 0.5f * der.z, rel.w + 0.5f * der.w});
    if (rel.w < 0.0f) rel = {-rel.x, -rel.y, -rel.z, -rel.w};
    b.orientation_delta = rel;
    b.displacement[3] = 2.0f * rel.x;
    b.displacement[4] = 2.0f * rel.y;
    b.displacement[5] = 2.0f * rel.z;
}

// rigid_avbd_solve_contact_rows, geometry_constraints.slang:506-529:
// normal arm from the world contact point, tangent arms from the anchors.
void solve_contact_rows(const PmoContact &c, const PmoAvbdBody &A, const PmoAvbdBody &B,
                        Vec6 ja[3], Vec6 jb[3], float err[3]) {
    V3 t0, t1;
    basis(c.normal, t0, t1);
    V3 axes[3] = {c.normal, t0, t1};
    V3 armA = sub(c.point, A.target.position), armB = sub(c.point, B.target.position);
    Vec6 da = disp(A), db = disp(B);
    for (uint32_t ax = 0; ax < 3; ++ax) {
        V3 ra = ax == 0 ? armA : rotate(A.target.orientation, c.avbd_anchor_a);
        V3 rb = ax == 0 ? armB : rotate(B.target.orientation, c.avbd_anchor_b);
        ja[ax] = jac(mul(axes[ax], -1.0f), mul(cross(ra, axes[ax]), -1.0f));
        jb[ax] = jac(axes[ax], cross(rb, axes[ax]));
        err[ax] = comp(c.avbd_error, ax) + dot6(ja[ax], da) + dot6(jb[ax], db);
    }
}
float solve_contact_normal_row(const PmoContact &c, const PmoAvbdBody &A, const PmoAvbdBody &B,
                               Vec6 &ja, Vec6 &jb) {              // :531-544
    V3 armA = sub(c.point, A.target.position), armB = sub(c.point, B.target.position);
    ja = jac(mul(c.normal, -1.0f), mul(cross(armA, c.normal), -1.0f));
    jb = jac(c.normal, cross(armB, c.normal));
    Vec6 da = disp(A), db = disp(B);
    return c.avbd_error.x + dot6(ja, da) + dot6(jb, db);
}

constexpr uint32_t kResidualSlots = 64;

struct Ctx {
    const PmoParams *P;
    const PmoState *prev;
    PmoState *states;
    uint32_t n;
    PmoManifold *M;
    const uint32_t *pairs;
    uint32_t np;
    PmoAvbdBody *B;
    PmoSchedule *S;
    float dt;
    uint32_t lanes;
    PmoJoint *J = nullptr;
    uint32_t nj = 0;
    // Constraint residual after each dual update, per phase and iteration
    // (kResidualSlots each; NaN = not run), and the tolerance of the
    // to-residual mode (0: fixed budget).
    float *residual = nullptr;           // RMS per phase and iteration
    float *residual_max = nullptr;       // max per phase and iteration
    float tolerance = 0.0f;
    float last_residual = 0.0f;
    uint32_t executed[2] = {0u, 0u};     // iterations run per phase
    std::vector<uint8_t> stick{};        // per point of the last dual update
    // Mistakes planted by tests to show that a pin discriminates (never set
    // otherwise): bit 0 skips the free-motion mean, bit 1 the crossing
    // safeguard (bit 2, read by the narrowphase, drops the hinge contact
    // frames).
    uint32_t planted = 0u;
};

inline bool dyn(const Ctx &x, uint32_t i) { return x.P[i].motion == kDynamic; }

// rigid_avbd_prepare_contact_impact, geometry_constraints.slang:983-1005.
bool contact_impact(const Ctx &x, uint32_t active) {
    uint32_t a = x.pairs[active] / x.n, b = x.pairs[active] % x.n;
    const PmoManifold &m = x.M[active];
    bool impact = false;
    for (uint32_t p = 0; p < m.count; ++p) {
        const PmoContact &c = m.contacts[p];
        float skin = contact_skin(c, x.P[a], x.P[b]);
        bool fresh = (c.avbd_cached == 0u || c.avbd_dual[0].lambda <= 1.0e-6f ||
                      (c.impact_fraction > 0.0f && c.impact_fraction < 1.0f)) &&
                     c.initial_normal_speed < -0.1f && c.penetration >= -kSurfaceTolerance;
        float prev_pen = c.penetration + c.initial_normal_speed * x.dt;
        impact = impact || fresh || prev_pen > skin + 1.0e-4f;
    }
    return impact;
}

// ---------------------------------------------------------------------
// Joints: rigid_avbd_visit_joint(_row), geometry_constraints.slang:
// 289-445, 626-926.  One routine builds every row of a joint for the four
// visits (initialise in prepare, accumulate in the colour sweep, advance
// in the dual update, the velocity-phase rows).  Pinned by closed forms
// (a pendulum's anchor drift and period, a hinge's free axis, a fixed
// weld, limits and breaking) in the tests; no reference-held vectors.
// ---------------------------------------------------------------------
enum JointType : uint32_t { kFixed = 0, kPoint = 1, kHinge = 2, kSlider = 3, kPiston = 4, kGeneric = 5,
                            kGenericSpring = 6, kMotor = 7 };
enum JointVisit : uint32_t { kInitialize = 0, kAccumulate = 1, kAdvance = 2, kImpact = 3 };
constexpr float kLimitStabilization = 0.8f;                   // geometry_constraints.slang:32

inline V3 basis_axis(uint32_t a) { return a == 0 ? V3{1, 0, 0} : (a == 1 ? V3{0, 1, 0} : V3{0, 0, 1}); }
inline Vec6 jac6(V3 l, V3 a) { return {{l.x, l.y, l.z, a.x, a.y, a.z}}; }
inline V3 relative_rotation_vector(Q a, Q b) {                 // :312-316
    return quaternion_rotation_vector(relative_quaternion(a, b, true));
}
// rigid_avbd_log_axis, :401-422: the derivative of the rotation
// logarithm along a body axis (series below one radian).
V3 log_axis(Q frame, V3 lg, uint32_t axis) {
    float sq = len2(lg);
    float coef = 1.0f / 12.0f + sq * (1.0f / 720.0f + sq * (1.0f / 30240.0f + sq * (1.0f / 1209600.0f +
                 sq / 47900160.0f)));
    if (sq > 1.0f) {
        float angle = std::sqrt(sq);
        float half = 0.5f * angle;
        coef = (1.0f - half * std::cos(half) / std::sin(half)) / sq;
    }
    V3 unit = basis_axis(axis);
    V3 local = add(add(unit, mul(cross(lg, unit), 0.5f)), mul(cross(lg, cross(lg, unit)), coef));
    return rotate(frame, local);
}

struct JointRowArgs {
    Vec6 ja, jb;
    float current, previous, penalty;
    bool angular;
    float stiffness, damping, lower, upper;
    bool motor;
    float motor_speed;
    V3 curvature_a, curvature_b;
    float stabilization;
};

// rigid_avbd_visit_joint_row, :626-713.
void visit_joint_row(Ctx &x, PmoJoint &j, Block &blk, uint32_t &index, float &impulse, bool first,
                     const JointRowArgs &r, uint32_t visit) {
    if (index >= 18u) return;
    const uint32_t ri = index++;
    Row row = row_from(reinterpret_cast<const float *>(&j.rows[ri]));
    const uint32_t a = j.dense_a, b = j.dense_b;
    const Vec6 da = disp(x.B[a]), db = disp(x.B[b]);
    const float linearized = dot6(r.ja, da) + dot6(r.jb, db);
    row.a = r.ja;
    row.b = r.jb;
    row.stiffness = r.stiffness;
    row.damping = r.damping;
    if (visit == kInitialize) {
        if (!(row.dual.penalty > kMinPenalty)) row.dual.penalty = r.penalty;
        row.dual = warm_start(row.dual, r.stiffness);
        if (r.stiffness < kMaxPenalty) row.dual.lambda = 0.0f;
        row.beta = r.penalty * (r.angular ? 10.0f : 100.0f);
        row.reference_force = 0.0f;
    }
    row.error = r.current - (r.stiffness < kMaxPenalty ? 0.0f : r.stabilization * r.previous) - linearized;
    row.velocity = r.motor ? r.motor_speed : (r.current - r.previous - linearized) / x.dt;
    row.lower = r.lower;
    row.upper = r.upper;
    if (visit == kAccumulate) {
        row_accumulate(blk, row, first, da, db, x.dt);
        if (!r.angular) {
            // Geometric stiffness of a linear row on the angular diagonal.
            V3 axis = {r.jb.v[0], r.jb.v[1], r.jb.v[2]};
            V3 arm = first ? r.curvature_a : r.curvature_b;
            float force = row_force(row, da, db, x.dt);
            for (uint32_t c = 0; c < 3; ++c) {
                V3 value = sub(mul(add(mul(axis, comp(arm, c)), mul(arm, comp(axis, c))), 0.5f),
                               mul(basis_axis(c), dot(axis, arm)));
                blk.h[bidx(3 + c, 3 + c)] += std::fabs(force) * std::sqrt(len2(value));
            }
        }
        return;
    }
    if (visit == kAdvance) {
        impulse += std::fabs(row_force(row, da, db, x.dt)) * x.dt;
        row_advance(row, da, db);
    } else if (visit == kImpact) {
        if (r.stiffness < kMaxPenalty) {
            row.reference_force = row_force(row, da, db, x.dt);
        } else {
            float gap = r.lower == 0.0f ? fmin_(0.0f, r.current) : (r.upper == 0.0f ? fmax_(0.0f, r.current) : 0.0f);
            Vec6 va = jac6(mul(x.B[a].target.linear_velocity, x.dt), mul(x.B[a].target.angular_velocity, x.dt));
            Vec6 vb = jac6(mul(x.B[b].target.linear_velocity, x.dt), mul(x.B[b].target.angular_velocity, x.dt));
            row.reference_force = row.dual.lambda;
            row.error = dot6(r.ja, va) + dot6(r.jb, vb) + gap - (r.motor ? r.motor_speed * x.dt : 0.0f);
        }
    }
    row_to(row, reinterpret_cast<float *>(&j.rows[ri]));
}

// rigid_avbd_visit_joint, :715-926.  Returns the row impulses of an
// advance visit.
float visit_joint(Ctx &x, PmoJoint &j, Block &blk, bool first, uint32_t visit) {
    uint32_t index = 0;
    float impulse = 0.0f;
    const PmoJointOptions &o = j.options;
    const uint32_t a = j.dense_a, b = j.dense_b;
    const PmoState sa = avbd_pose(x.B[a]), sb = avbd_pose(x.B[b]);
    const V3 armA = rotate(sa.orientation, o.local_anchor_a), armB = rotate(sb.orientation, o.local_anchor_b);
    const V3 error = sub(add(sb.position, armB), add(sa.position, armA));
    const V3 old_error = sub(add(x.prev[b].position, rotate(x.prev[b].orientation, o.local_anchor_b)),
                             add(x.prev[a].position, rotate(x.prev[a].orientation, o.local_anchor_a)));
    const Q fA = qnormalize(qmul(sa.orientation, o.local_orientation_a));
    const Q fB = qnormalize(qmul(sb.orientation, o.local_orientation_b));
    const Q ofA = qnormalize(qmul(x.prev[a].orientation, o.local_orientation_a));
    const Q ofB = qnormalize(qmul(x.prev[b].orientation, o.local_orientation_b));
    const V3 twist = relative_rotation_vector(fA, fB);
    const V3 old_twist = relative_rotation_vector(ofA, ofB);
    const uint32_t type = o.type;
    const bool swing = type == kHinge || type == kPiston || (type == kMotor && o.motor.angular_enabled != 0u);
    const V3 free_axis = type == kHinge ? V3{0.0f, 0.0f, 1.0f} : V3{1.0f, 0.0f, 0.0f};
    const V3 swing_axis = inv_rotate(fA, rotate(fB, free_axis));
    const V3 old_swing_axis = inv_rotate(ofA, rotate(ofB, free_axis));
    const V3 rotation = swing ? cross(free_axis, swing_axis) : twist;
    const V3 old_rotation = swing ? cross(free_axis, old_swing_axis) : old_twist;
    const float linear_penalty = pair_stiffness(x.P[a], x.P[b], x.dt);
    const bool generic = type == kGeneric || type == kGenericSpring;
    const bool motor_type = type == kMotor;
    for (uint32_t axis = 0; axis < 3; ++axis) {
        const V3 direction = rotate(fA, basis_axis(axis));
        const V3 old_direction = rotate(ofA, basis_axis(axis));
        for (uint32_t av = 0; av < 2; ++av) {
            const bool angular = av != 0;
            const bool lock = angular
                ? type == kFixed || type == kSlider || (type == kHinge && axis != 2u) ||
                      (type == kPiston && axis != 0u) ||
                      (motor_type && (axis != 0u || o.motor.angular_enabled == 0u))
                : type == kFixed || type == kPoint || type == kHinge ||
                      ((type == kSlider || type == kPiston) && axis != 0u) ||
                      (motor_type && (axis != 0u || o.motor.linear_enabled == 0u));
            const PmoJointLimits &limits = angular ? o.angular_limits : o.linear_limits;
            const PmoJointSprings &springs = angular ? o.angular_springs : o.linear_springs;
            const float current = angular ? comp(rotation, axis) : dot(error, direction);
            const float old_current = angular ? comp(old_rotation, axis) : dot(old_error, old_direction);
            const V3 angular_axis = swing
                ? rotate(fA, sub(mul(basis_axis(axis), dot(free_axis, swing_axis)),
                                 mul(free_axis, comp(swing_axis, axis))))
                : log_axis(fA, twist, axis);
            const V3 zero = {0.0f, 0.0f, 0.0f};
            const Vec6 ja = angular ? jac6(zero, mul(angular_axis, -1.0f))
                                    : jac6(mul(direction, -1.0f), add(mul(cross(armA, direction), -1.0f),
                                                                     cross(direction, error)));
            const Vec6 jb = angular ? jac6(zero, angular_axis) : jac6(direction, cross(armB, direction));
            const float penalty = angular
                ? 1.0f / (fmax_(dot(direction, add(inverse_inertia_world(x.P[a], sa, direction),
                                                   inverse_inertia_world(x.P[b], sb, direction))),
                                1.0e-6f) * x.dt * x.dt)
                : linear_penalty;
            const float lock_penalty = angular
                ? fmax_(penalty, 2.0f * linear_penalty * fmax_(len2(cross(direction, armA)), len2(cross(direction, armB))))
                : penalty;
            const V3 curv_a = add(error, armA);
            if (lock) {
                if (visit == kInitialize && angular && index < 18u)
                    j.rows[index].penalty = fmax_(j.rows[index].penalty, lock_penalty / kGamma);
                visit_joint_row(x, j, blk, index, impulse, first,
                                {ja, jb, current, old_current, lock_penalty, angular, kMaxPenalty, 0.0f,
                                 -kMaxPenalty, kMaxPenalty, false, 0.0f, curv_a, armB, kAlpha},
                                visit);
            }
            const bool free_limit = angular ? (type == kHinge && axis == 2u) || (type == kPiston && axis == 0u)
                                            : (type == kSlider || type == kPiston) && axis == 0u;
            if ((generic || free_limit) && (limits.axes & (1u << axis)) != 0u) {
                const float limit_current = angular ? comp(twist, axis) : current;
                const float old_limit_current = angular ? comp(old_twist, axis) : old_current;
                const V3 limit_axis = log_axis(fA, twist, axis);
                const Vec6 la = angular ? jac6(zero, mul(limit_axis, -1.0f)) : ja;
                const Vec6 lb = angular ? jac6(zero, limit_axis) : jb;
                const float lower = comp(limits.lower, axis), upper = comp(limits.upper, axis);
                if (lower == upper) {
                    const float equal_penalty = angular && generic ? lock_penalty : penalty;
                    if (visit == kInitialize && angular && generic && index < 18u)
                        j.rows[index].penalty = fmax_(j.rows[index].penalty, equal_penalty / kGamma);
                    visit_joint_row(x, j, blk, index, impulse, first,
                                    {la, lb, limit_current - lower, old_limit_current - lower, equal_penalty, angular,
                                     kMaxPenalty, 0.0f, -kMaxPenalty, kMaxPenalty, false, 0.0f, curv_a, armB, kAlpha},
                                    visit);
                } else if (lower < upper) {
                    visit_joint_row(x, j, blk, index, impulse, first,
                                    {la, lb, limit_current - upper, fmax_(old_limit_current - upper, 0.0f), penalty,
                                     angular, kMaxPenalty, 0.0f, 0.0f, kMaxPenalty, false, 0.0f, curv_a, armB,
                                     kLimitStabilization},
                                    visit);
                    visit_joint_row(x, j, blk, index, impulse, first,
                                    {la, lb, limit_current - lower, fmin_(old_limit_current - lower, 0.0f), penalty,
                                     angular, kMaxPenalty, 0.0f, -kMaxPenalty, 0.0f, false, 0.0f, curv_a, armB,
                                     kLimitStabilization},
                                    visit);
                }
            }
            if (type == kGenericSpring && (springs.axes & (1u << axis)) != 0u)
                visit_joint_row(x, j, blk, index, impulse, first,
                                {ja, jb, current, old_current, penalty, angular, comp(springs.stiffness, axis),
                                 comp(springs.damping, axis), -kMaxPenalty, kMaxPenalty, false, 0.0f, curv_a, armB,
                                 kAlpha},
                                visit);
            if (motor_type && axis == 0u && (angular ? o.motor.angular_enabled != 0u : o.motor.linear_enabled != 0u)) {
                const float speed = angular ? o.motor.angular_target_velocity : o.motor.linear_target_velocity;
                const float bound = (angular ? o.motor.angular_maximum_impulse : o.motor.linear_maximum_impulse) / x.dt;
                const Q rel = relative_quaternion(fA, fB, true);
                const Q old_rel = relative_quaternion(ofA, ofB, true);
                const V3 motor_rotation = relative_rotation_vector({0.0f, 0.0f, 0.0f, 1.0f},
                                                                   relative_quaternion(old_rel, rel, false));
                const V3 motor_axis = log_axis(fA, motor_rotation, axis);
                const Vec6 ma = angular ? jac6(zero, mul(motor_axis, -1.0f)) : ja;
                const Vec6 mb = angular ? jac6(zero, motor_axis) : jb;
                const float travel = angular ? comp(motor_rotation, axis) : current - old_current;
                visit_joint_row(x, j, blk, index, impulse, first,
                                {ma, mb, travel - speed * x.dt, 0.0f, penalty, angular, kMaxPenalty, 0.0f, -bound,
                                 bound, true, speed, curv_a, armB, kAlpha},
                                visit);
            }
        }
    }
    if (visit == kInitialize) j.avbd_count = index;
    return impulse;
}

// Joints incident to body i in list order (push-front: descending index).
template <class F>
void for_each_joint(const Ctx &x, uint32_t i, F f) {
    for (uint32_t k = x.B[i].joint_head; k != kInvalid;) {
        const bool first = i == x.J[k].dense_a;
        f(k, first);
        k = first ? x.J[k].avbd_next_a : x.J[k].avbd_next_b;
    }
}

void prepare(Ctx &x, V3 gravity, uint32_t base_iterations, uint32_t requested, uint32_t profile) {
    const bool realtime = profile == kRealtimeProfile;
    const float dt = x.dt;
    // Bodies, geometry_constraints.slang:1235-1295.
    for (uint32_t i = 0; i < x.n; ++i) {
        V3 old_v = x.B[i].previous_velocity;
        bool hist = x.B[i].history_valid != 0u;
        PmoAvbdBody body;
        std::memset(&body, 0, sizeof body);
        body.parent = i;
        body.color = kInvalid;
        body.contact_head = kInvalid;
        body.joint_head = kInvalid;
        body.target = x.states[i];
        body.orientation_delta = {0, 0, 0, 1};
        if (x.P[i].motion != kDynamic) { x.B[i] = body; continue; }
        float g2 = len2(gravity);
        V3 acc = mul(sub(x.prev[i].linear_velocity, old_v), 1.0f / dt);
        float w = hist && g2 > 0.0f ? clampf(dot(acc, gravity) / g2, 0.0f, 1.0f) : 0.0f;
        if (!std::isfinite(w)) w = 0.0f;
        V3 corr = mul(gravity, -(1.0f - w) * dt * dt);
        body.displacement[0] = corr.x;
        body.displacement[1] = corr.y;
        body.displacement[2] = corr.z;
        body.previous_velocity = x.prev[i].linear_velocity;
        body.history_valid = 1u;
        float mass = 1.0f / fmax_(x.P[i].inverse_mass, 1.0e-12f);
        for (uint32_t ax = 0; ax < 3; ++ax) {
            body.inertia_h[bidx(ax, ax)] = mass / (dt * dt);
            V3 unit = ax == 0 ? V3{1, 0, 0} : (ax == 1 ? V3{0, 1, 0} : V3{0, 0, 1});
            V3 la = inv_rotate(x.states[i].orientation, unit);
            V3 col = rotate(x.states[i].orientation,
                {la.x / fmax_(x.P[i].inverse_inertia_local.x, 1.0e-12f),
                 la.y / fmax_(x.P[i].inverse_inertia_local.y, 1.0e-12f),
                 la.z / fmax_(x.P[i].inverse_inertia_local.z, 1.0e-12f)});
            for (uint32_t r = ax; r < 3; ++r)
                body.inertia_h[bidx(3 + r, 3 + ax)] = comp(col, r) / (dt * dt);
        }
        x.B[i] = body;
    }
    // Contacts, geometry_constraints.slang:1297-1375.
    for (uint32_t active = 0; active < x.np; ++active) {
        uint32_t a = x.pairs[active] / x.n, b = x.pairs[active] % x.n;
        float penalty = pair_stiffness(x.P[a], x.P[b], dt);
        PmoManifold &m = x.M[active];
        for (uint32_t p = 0; p < m.count; ++p) {
            PmoContact &c = m.contacts[p];
            if (c.avbd_cached == 0u) {
                c.avbd_anchor_a = inv_rotate(x.prev[a].orientation, sub(c.point, x.prev[a].position));
                c.avbd_anchor_b = inv_rotate(x.prev[b].orientation, sub(c.point, x.prev[b].position));
            }
            V3 ma = sub(x.states[a].position, x.prev[a].position);
            V3 mb = sub(x.states[b].position, x.prev[b].position);
            V3 wa = mul(quaternion_delta_velocity(x.prev[a].orientation, x.states[a].orientation, dt), dt);
            V3 wb = mul(quaternion_delta_velocity(x.prev[b].orientation, x.states[b].orientation, dt), dt);
            V3 armA = sub(c.point, x.states[a].position), armB = sub(c.point, x.states[b].position);
            float closing = dot(c.normal, sub(add(mb, cross(wb, armB)), add(ma, cross(wa, armA))));
            float old_pen = c.penetration - closing;
            V3 t0, t1;
            basis(c.normal, t0, t1);
            V3 anchor_motion = add(sub(mb, ma),
                sub(rotation_displacement(x.prev[b].orientation, x.states[b].orientation, c.avbd_anchor_b),
                    rotation_displacement(x.prev[a].orientation, x.states[a].orientation, c.avbd_anchor_a)));
            V3 old_delta = c.avbd_cached == 0u ? V3{0, 0, 0} : add(
                sub(x.prev[b].position, x.prev[a].position),
                sub(rotate(x.prev[b].orientation, c.avbd_anchor_b), rotate(x.prev[a].orientation, c.avbd_anchor_a)));
            V3 terr = add(anchor_motion, mul(old_delta, 1.0f - kAlpha));
            c.initial_normal_speed = -closing / dt;
            float skin = contact_skin(c, x.P[a], x.P[b]);
            float impact_disp = realtime && c.avbd_cached == 0u && c.initial_normal_speed < -0.1f
                ? -(1.0f + fmin_(x.P[a].restitution, x.P[b].restitution)) * c.initial_normal_speed * dt
                : 0.0f;
            c.avbd_error = {c.penetration - clampf(old_pen, 0.0f, skin) + impact_disp,
                            dot(terr, t0), dot(terr, t1)};
            for (uint32_t ax = 0; ax < 3; ++ax) {
                Dual d = {c.avbd_dual[ax].lambda, c.avbd_dual[ax].penalty};
                if (c.avbd_cached == 0u) d = {0.0f, penalty};
                d = warm_start(d, kMaxPenalty);
                c.avbd_dual[ax] = {d.lambda, d.penalty};
            }
        }
    }
    // Joints, geometry_constraints.slang:1391-1417 (body ids are
    // {dense index, 1}: rigid_avbd_find_dense over this world's ids).
    for (uint32_t k = 0; k < x.nj; ++k) {
        PmoJoint &j = x.J[k];
        j.avbd_count = 0u;
        if (j.alive == 0u) continue;
        j.state_enabled = j.options.enabled != 0u && j.state_broken == 0u ? 1u : 0u;
        if (j.state_enabled == 0u) continue;
        j.applied_impulse = 0.0f;
        auto find = [&](PmoBodyId id) { return id.index < x.n && id.generation == 1u ? id.index : kInvalid; };
        j.dense_a = find(j.options.body_a);
        j.dense_b = find(j.options.body_b);
        if (j.dense_a != kInvalid && j.dense_b != kInvalid) {
            Block scratch = zero_block();
            visit_joint(x, j, scratch, true, kInitialize);
        }
    }
    // Schedule and graph (lane 0), geometry_constraints.slang:1432-1560.
    PmoSchedule s;
    std::memset(&s, 0, sizeof s);
    s.budget = realtime ? kRealtimeSweeps : (requested != 0u ? requested : base_iterations);
    s.offsets[0] = profile;
    s.offsets[1] = base_iterations;
    s.counts[2] = 0u;
    s.candidates = x.np;
    std::vector<uint32_t> parent(x.n);
    for (uint32_t i = 0; i < x.n; ++i) parent[i] = i;
    auto root = [&](uint32_t v) { while (v != parent[v]) v = parent[v]; return v; };
    auto unite = [&](uint32_t a, uint32_t b) {
        if (!dyn(x, a) || !dyn(x, b)) return;
        a = root(a); b = root(b);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
    };
    for (uint32_t active = 0; active < x.np; ++active) {
        if (x.M[active].count == 0u) continue;
        s.contacts++;
        uint32_t a = x.pairs[active] / x.n, b = x.pairs[active] % x.n;
        unite(a, b);
        x.M[active].avbd_next_a = x.B[a].contact_head;
        x.B[a].contact_head = active;
        x.M[active].avbd_next_b = x.B[b].contact_head;
        x.B[b].contact_head = active;
    }
    for (uint32_t k = 0; k < x.nj; ++k) {
        PmoJoint &j = x.J[k];
        if (j.avbd_count == 0u) continue;
        uint32_t a = j.dense_a, b = j.dense_b;
        if (!realtime && requested == 0u) s.budget = std::max(s.budget, j.options.solver_iterations);
        unite(a, b);
        j.avbd_next_a = x.B[a].joint_head;
        x.B[a].joint_head = k;
        j.avbd_next_b = x.B[b].joint_head;
        x.B[b].joint_head = k;
    }
    for (uint32_t i = 0; i < x.n; ++i) parent[i] = root(i);
    std::vector<uint32_t> impact(x.n, 0u);
    for (uint32_t active = 0; active < x.np; ++active) {
        uint32_t a = x.pairs[active] / x.n, b = x.pairs[active] % x.n;
        if (contact_impact(x, active)) {
            if (dyn(x, a)) impact[parent[a]] = 1u;
            if (dyn(x, b)) impact[parent[b]] = 1u;
            s.counts[0] = 1u;
        }
    }
    for (uint32_t i = 0; i < x.n; ++i) {
        if (!dyn(x, i)) continue;
        impact[i] = impact[parent[i]];
        if (parent[i] == i && (x.B[i].contact_head != kInvalid || x.B[i].joint_head != kInvalid)) s.island_count++;
    }
    for (uint32_t i = 0; i < x.n; ++i) {
        if (!dyn(x, i)) continue;
        uint32_t color = 0;
        for (;; ++color) {
            bool conflict = false;
            for (uint32_t active = x.B[i].contact_head; active != kInvalid;) {
                uint32_t a = x.pairs[active] / x.n, b = x.pairs[active] % x.n;
                bool first = i == a;
                conflict = conflict || x.B[first ? b : a].color == color;
                active = first ? x.M[active].avbd_next_a : x.M[active].avbd_next_b;
            }
            for_each_joint(x, i, [&](uint32_t k, bool first) {
                conflict = conflict || x.B[first ? x.J[k].dense_b : x.J[k].dense_a].color == color;
            });
            if (!conflict) break;
        }
        x.B[i].color = color;
        s.remaining = std::max(s.remaining, color + 1u);
    }
    for (uint32_t i = 0; i < x.n; ++i) {
        x.B[i].parent = parent[i];
        x.B[i].impact = dyn(x, i) ? impact[i] : 0u;
    }
    s.blocks = 1u;
    *x.S = s;
}

// Butterfly reduction of L partial blocks, matching a wavefront's xor
// shuffle tree: level s combines partial[l] and partial[l ^ s].
Block tree_reduce(std::vector<Block> &parts) {
    const size_t L = parts.size();
    for (size_t s = 1; s < L; s <<= 1) {
        std::vector<Block> next(L);
        for (size_t l = 0; l < L; ++l) {
            const Block &u = parts[l], &v = parts[l ^ s];
            Block r;
            for (int i = 0; i < 21; ++i) r.h[i] = u.h[i] + v.h[i];
            for (int i = 0; i < 6; ++i) r.g[i] = u.g[i] + v.g[i];
            next[l] = r;
        }
        parts.swap(next);
    }
    return parts[0];
}

struct PointRef { uint32_t active, point; };

// Contact points incident to a body in list order (reverse active order).
void body_points(const Ctx &x, uint32_t i, std::vector<PointRef> &out) {
    out.clear();
    for (uint32_t active = x.B[i].contact_head; active != kInvalid;) {
        uint32_t a = x.pairs[active] / x.n;
        bool first = i == a;
        for (uint32_t p = 0; p < x.M[active].count; ++p) out.push_back({active, p});
        active = first ? x.M[active].avbd_next_a : x.M[active].avbd_next_b;
    }
}

// One contact point's three rows for body i (solve_color inner body,
// geometry_constraints_solve.slang:356-397).
void contact_point_rows(const Ctx &x, uint32_t i, PointRef r, uint32_t phase, Block &blk) {
    uint32_t a = x.pairs[r.active] / x.n, b = x.pairs[r.active] % x.n;
    bool first = i == a;
    float mu = std::sqrt(x.P[a].friction * x.P[b].friction);
    const PmoContact &c = x.M[r.active].contacts[r.point];
    Vec6 ja[3], jb[3];
    float err[3];
    solve_contact_rows(c, x.B[a], x.B[b], ja, jb, err);
    Force3 trial = {c.avbd_dual[0].lambda + c.avbd_dual[0].penalty * err[0],
                    c.avbd_dual[1].lambda + c.avbd_dual[1].penalty * err[1],
                    c.avbd_dual[2].lambda + c.avbd_dual[2].penalty * err[2]};
    Force3 f = project_contact(trial, mu);
    Force3 sc = stiffness_scales(trial, mu);
    float stiff[3] = {sc.normal, sc.t0, sc.t1};
    V3 t0, t1;
    basis(c.normal, t0, t1);
    float forces[3] = {
        f.normal - (phase != 0u ? c.accumulated_normal_impulse / x.dt : 0.0f),
        f.t0 - (phase != 0u ? dot(c.accumulated_friction_impulse, t0) / x.dt : 0.0f),
        f.t1 - (phase != 0u ? dot(c.accumulated_friction_impulse, t1) / x.dt : 0.0f)};
    for (uint32_t ax = 0; ax < 3; ++ax)
        add_row(blk, first ? ja[ax] : jb[ax], c.avbd_dual[ax].penalty * stiff[ax], forces[ax]);
}

// Crossing safeguard for one point (geometry_constraints_solve.slang:
// 433-465); returns true when the row was added.
bool crossing_row(const Ctx &x, uint32_t i, PointRef r, const Vec6 &update, Block &blk) {
    uint32_t a = x.pairs[r.active] / x.n, b = x.pairs[r.active] % x.n;
    bool first = i == a;
    const PmoContact &c = x.M[r.active].contacts[r.point];
    Vec6 ja, jb;
    float e = solve_contact_normal_row(c, x.B[a], x.B[b], ja, jb);
    const Vec6 &nr = first ? ja : jb;
    float trial = c.avbd_dual[0].lambda + c.avbd_dual[0].penalty * e;
    if (trial <= 0.0f && trial + c.avbd_dual[0].penalty * dot6(nr, update) >
                             c.avbd_dual[0].penalty * 1.0e-6f) {
        add_row(blk, nr, c.avbd_dual[0].penalty, trial);
        return true;
    }
    return false;
}

void solve_body(Ctx &x, uint32_t i, uint32_t phase, std::vector<PointRef> &pts) {
    Block base;
    std::memcpy(base.h, x.B[i].inertia_h, sizeof base.h);
    std::memcpy(base.g, x.B[i].inertia_g, sizeof base.g);
    for (uint32_t r = 0; r < 6; ++r)
        for (uint32_t c = 0; c < 6; ++c) base.g[r] += base.h[bidx(r, c)] * x.B[i].displacement[c];
    body_points(x, i, pts);
    // Joint rows (solve.slang:401-430) follow the contact rows: recomputed
    // in the pose phase, stored hard rows in the velocity phase. In lane
    // order the k-th enabled joint of the body's list goes to partial k mod L.
    auto joint_rows = [&](auto &&target) {
        uint32_t pos = 0;
        for_each_joint(x, i, [&](uint32_t k, bool first) {
            PmoJoint &j = x.J[k];
            if (j.state_enabled == 0u) return;
            Block &b = target(pos++);
            if (phase == 0u) {
                visit_joint(x, j, b, first, kAccumulate);
                return;
            }
            const Vec6 da = disp(x.B[j.dense_a]), db = disp(x.B[j.dense_b]);
            for (uint32_t r = 0; r < j.avbd_count; ++r) {
                Row row = row_from(reinterpret_cast<const float *>(&j.rows[r]));
                if (row.stiffness >= kMaxPenalty) row_accumulate(b, row, first, da, db, x.dt);
            }
        });
    };
    Block blk;
    if (x.lanes == 0u) {
        blk = base;
        for (auto &r : pts) contact_point_rows(x, i, r, phase, blk);
        joint_rows([&](uint32_t) -> Block & { return blk; });
    } else {
        std::vector<Block> parts(x.lanes, zero_block());
        parts[0] = base;
        for (size_t k = 0; k < pts.size(); ++k) contact_point_rows(x, i, pts[k], phase, parts[k % x.lanes]);
        joint_rows([&](uint32_t pos) -> Block & { return parts[pos % x.lanes]; });
        blk = tree_reduce(parts);
    }
    Vec6 update = zero6();
    if (!block_solve(blk, update)) return;
    bool crossing = false;
    if (x.planted & 2u) {
        // planted mistake: no crossing safeguard
    } else if (x.lanes == 0u) {
        for (auto &r : pts) crossing = crossing_row(x, i, r, update, blk) || crossing;
    } else {
        std::vector<Block> parts(x.lanes, zero_block());
        for (size_t k = 0; k < pts.size(); ++k)
            crossing = crossing_row(x, i, pts[k], update, parts[k % x.lanes]) || crossing;
        if (crossing) {
            Block c = tree_reduce(parts);
            for (int q = 0; q < 21; ++q) blk.h[q] += c.h[q];
            for (int q = 0; q < 6; ++q) blk.g[q] += c.g[q];
        }
    }
    if (crossing) {
        Vec6 safe = zero6();
        if (block_solve(blk, safe)) update = safe;
    }
    if (phase == 0u) apply_pose_update(x.B[i], update);
    else for (int q = 0; q < 6; ++q) x.B[i].displacement[q] += update.v[q];
}

// advance_duals, geometry_constraints_solve.slang:492-566.
// Residual of a joint's hard equality rows (locked axes, equal limits):
// |C| at the current displacements.
float joint_residual(const Ctx &x, const PmoJoint &j) {
    if (j.state_enabled == 0u) return 0.0f;
    const Vec6 da = disp(x.B[j.dense_a]), db = disp(x.B[j.dense_b]);
    float r = 0.0f;
    for (uint32_t k = 0; k < j.avbd_count; ++k) {
        Row row = row_from(reinterpret_cast<const float *>(&j.rows[k]));
        if (row.stiffness < kMaxPenalty || row.lower != -kMaxPenalty || row.upper != kMaxPenalty) continue;
        r = fmax_(r, std::fabs(row.error + dot6(row.a, da) + dot6(row.b, db)));
    }
    return r;
}

// The impulses published after the last dual update and the re-anchoring
// of sliding contacts (solve.slang:540-562), from the stored forces and
// stick flags of that update.
void publish_impulses(Ctx &x, uint32_t phase) {
    size_t slot = 0;
    for (uint32_t active = 0; active < x.np; ++active) {
        uint32_t a = x.pairs[active] / x.n, b = x.pairs[active] % x.n;
        PmoManifold &m = x.M[active];
        const bool skip = phase != 0u && x.B[a].impact == 0u && x.B[b].impact == 0u;
        for (uint32_t p = 0; p < m.count; ++p, ++slot) {
            if (skip) continue;
            PmoContact &c = m.contacts[p];
            V3 t0, t1;
            basis(c.normal, t0, t1);
            c.accumulated_normal_impulse = c.avbd_dual[0].lambda * x.dt;
            c.reported_normal_impulse = c.accumulated_normal_impulse;
            c.accumulated_friction_impulse = mul(add(mul(t0, c.avbd_dual[1].lambda), mul(t1, c.avbd_dual[2].lambda)), x.dt);
            c.reported_friction_impulse = c.accumulated_friction_impulse;
            if (!x.stick[slot]) {
                PmoState pa = phase != 0u ? x.B[a].target : avbd_pose(x.B[a]);
                PmoState pb = phase != 0u ? x.B[b].target : avbd_pose(x.B[b]);
                c.avbd_anchor_a = inv_rotate(pa.orientation, sub(c.point, pa.position));
                c.avbd_anchor_b = inv_rotate(pb.orientation, sub(c.point, pb.position));
            }
        }
    }
}

// Constraint residual of one contact point at the primal state the dual
// update sees (no reference counterpart): an active normal row (force > 0)
// must hold exactly, an inactive one must not penetrate, and a sticking
// friction row must not slip.
inline float contact_residual(const float err[3], float normal_force, bool stick) {
    float r = normal_force > 0.0f ? std::fabs(err[0]) : fmax_(err[0], 0.0f);
    if (stick) r = fmax_(r, fmax_(std::fabs(err[1]), std::fabs(err[2])));
    return r;
}

// The residual of an update is the RMS of the per-point and per-joint
// values (the tolerance test) and their maximum (reported).  Squares are
// summed in 2^-40 fixed point (1e-6 m resolves), each clamped to 2^44
// units (r = 4 m), so the sum is exact and order-free; a million rows at
// the clamp still fit 64 bits.
inline uint64_t residual_square(float r) {
    const double q = double(r) * double(r) * 1099511627776.0 + 0.5;
    return !(q < 17592186044416.0) ? uint64_t(17592186044416ull) : uint64_t(q);
}
inline float residual_rms(uint64_t sum, uint32_t count) {
    return count ? float(std::sqrt(double(sum) / 1099511627776.0 / double(count))) : 0.0f;
}
struct ResidualSum {
    uint64_t sum = 0;
    uint32_t count = 0;
    float max = 0.0f;
    void add(float r) { sum += residual_square(r); ++count; max = fmax_(max, r); }
};

void advance_duals(Ctx &x, uint32_t phase, uint32_t iteration) {
    // Points in manifold order; the stick flags feed publish_impulses.
    size_t slot = 0;
    ResidualSum residual;
    const bool fixed_last = x.tolerance <= 0.0f && iteration + 1u == x.S->budget;
    x.stick.clear();
    for (uint32_t active = 0; active < x.np; ++active) x.stick.resize(x.stick.size() + x.M[active].count, 1u);
    for (uint32_t active = 0; active < x.np; ++active) {
        uint32_t a = x.pairs[active] / x.n, b = x.pairs[active] % x.n;
        if (phase != 0u && x.B[a].impact == 0u && x.B[b].impact == 0u) {
            slot += x.M[active].count;
            continue;
        }
        float mu = std::sqrt(x.P[a].friction * x.P[b].friction);
        float beta = pair_stiffness(x.P[a], x.P[b], x.dt) * 100.0f;
        PmoManifold &m = x.M[active];
        for (uint32_t p = 0; p < m.count; ++p) {
            PmoContact &c = m.contacts[p];
            Vec6 ja[3], jb[3];
            float err[3];
            solve_contact_rows(c, x.B[a], x.B[b], ja, jb, err);
            Force3 trial = {c.avbd_dual[0].lambda + c.avbd_dual[0].penalty * err[0],
                            c.avbd_dual[1].lambda + c.avbd_dual[1].penalty * err[1],
                            c.avbd_dual[2].lambda + c.avbd_dual[2].penalty * err[2]};
            Force3 f = project_contact(trial, mu);
            float tx = c.avbd_dual[1].lambda + c.avbd_dual[1].penalty * err[1];
            float ty = c.avbd_dual[2].lambda + c.avbd_dual[2].penalty * err[2];
            float fb = mu * f.normal;
            bool stick = tx * tx + ty * ty <= fb * fb;
            c.avbd_dual[0].lambda = f.normal;
            c.avbd_dual[1].lambda = f.t0;
            c.avbd_dual[2].lambda = f.t1;
            for (uint32_t ax = 0; ax < 3; ++ax)
                if (ax == 0 ? f.normal > 0.0f : stick)
                    c.avbd_dual[ax].penalty = fmin_(kMaxPenalty, c.avbd_dual[ax].penalty + beta * std::fabs(err[ax]));
            residual.add(contact_residual(err, f.normal, stick));
            x.stick[slot + p] = stick ? 1u : 0u;
        }
        slot += m.count;
    }
    if (fixed_last) publish_impulses(x, phase);
    // Joints (solve.slang:566-614): dual updates, applied impulse, breaking.
    for (uint32_t k = 0; k < x.nj; ++k) {
        PmoJoint &j = x.J[k];
        if (j.avbd_count == 0u || j.state_enabled == 0u) continue;
        const uint32_t a = j.dense_a, b = j.dense_b;
        if (phase != 0u && x.B[a].impact == 0u && x.B[b].impact == 0u) continue;
        float impulse = 0.0f;
        if (phase == 0u) {
            Block scratch = zero_block();
            impulse = visit_joint(x, j, scratch, true, kAdvance);
        } else {
            const Vec6 da = disp(x.B[a]), db = disp(x.B[b]);
            for (uint32_t r = 0; r < j.avbd_count; ++r) {
                Row row = row_from(reinterpret_cast<const float *>(&j.rows[r]));
                impulse += std::fabs(row.stiffness < kMaxPenalty ? row.reference_force : row_force(row, da, db, x.dt)) * x.dt;
                if (row.stiffness >= kMaxPenalty) {
                    row_advance(row, da, db);
                    row_to(row, reinterpret_cast<float *>(&j.rows[r]));
                }
            }
        }
        j.applied_impulse = impulse;
        if (j.options.breaking_impulse_threshold > 0.0f && impulse > j.options.breaking_impulse_threshold) {
            j.state_broken = 1u;
            j.state_enabled = 0u;
        }
        residual.add(joint_residual(x, j));
    }
    x.last_residual = residual_rms(residual.sum, residual.count);
    if (iteration < kResidualSlots && x.residual) {
        x.residual[phase * kResidualSlots + iteration] = x.last_residual;
        x.residual_max[phase * kResidualSlots + iteration] = residual.max;
    }
}

// free_mean / free_subtract / free_clear, :617-698, with each island's
// members reduced in ascending index order.
void free_motion(Ctx &x, uint32_t phase) {
    const bool vel = phase != 0u;
    std::vector<std::vector<uint32_t>> members(x.n);
    for (uint32_t i = 0; i < x.n; ++i)
        if (dyn(x, i)) members[x.B[i].parent].push_back(i);
    std::vector<V3> mean(x.n, V3{0, 0, 0});
    for (uint32_t r = 0; r < x.n; ++r) {
        if (!dyn(x, r) || x.B[r].parent != r) continue;
        x.B[r].inertia_g[0] = x.B[r].inertia_g[1] = x.B[r].inertia_g[2] = 0.0f;
        if ((vel && x.B[r].impact == 0u) || (x.B[r].contact_head == kInvalid && x.B[r].joint_head == kInvalid)) continue;
        V3 weighted{0, 0, 0};
        float total = 0.0f;
        bool anchored = false;
        for (uint32_t i : members[r]) {
            if (anchored) break;
            float mass = 1.0f / fmax_(x.P[i].inverse_mass, 1.0e-12f);
            weighted = add(weighted, mul(trans(x.B[i].displacement), mass));
            total += mass;
            for (uint32_t active = x.B[i].contact_head; active != kInvalid;) {
                uint32_t a = x.pairs[active] / x.n, b = x.pairs[active] % x.n;
                bool first = i == a;
                anchored = anchored || x.P[first ? b : a].motion != kDynamic;
                active = first ? x.M[active].avbd_next_a : x.M[active].avbd_next_b;
            }
            for_each_joint(x, i, [&](uint32_t k, bool first) {
                anchored = anchored || x.P[first ? x.J[k].dense_b : x.J[k].dense_a].motion != kDynamic;
            });
        }
        if (!anchored && total > 0.0f) {
            weighted = mul(weighted, 1.0f / total);
            x.B[r].inertia_g[0] = weighted.x;
            x.B[r].inertia_g[1] = weighted.y;
            x.B[r].inertia_g[2] = weighted.z;
        }
    }
    for (uint32_t i = 0; i < x.n; ++i) {
        if (!dyn(x, i) || (vel && x.B[i].impact == 0u)) continue;
        const PmoAvbdBody &root = x.B[x.B[i].parent];
        for (int ax = 0; ax < 3; ++ax) x.B[i].displacement[ax] -= root.inertia_g[ax];
    }
    for (uint32_t i = 0; i < x.n; ++i) x.B[i].inertia_g[0] = x.B[i].inertia_g[1] = x.B[i].inertia_g[2] = 0.0f;
}

// apply_states, :699-733.
void apply_states(Ctx &x, uint32_t phase, bool realtime) {
    for (uint32_t i = 0; i < x.n; ++i) {
        if (!dyn(x, i)) continue;
        if (phase != 0u && x.B[i].impact == 0u) continue;
        PmoState s = x.B[i].target;
        const float *d = x.B[i].displacement;
        if (phase != 0u) {
            s.linear_velocity = add(s.linear_velocity, mul(trans(d), 1.0f / x.dt));
            s.angular_velocity = add(s.angular_velocity, mul(rot6(d), 1.0f / x.dt));
            x.states[i] = s;
            continue;
        }
        s.position = add(s.position, trans(d));
        s.orientation = avbd_pose(x.B[i]).orientation;
        s.linear_velocity = add(s.linear_velocity, mul(trans(d), 1.0f / x.dt));
        s.angular_velocity = add(s.angular_velocity, mul(quaternion_rotation_vector(x.B[i].orientation_delta), 1.0f / x.dt));
        if (realtime) {
            s.linear_velocity = mul(s.linear_velocity, 1.0f - kRealtimeDamping);
            s.angular_velocity = mul(s.angular_velocity, 1.0f - kRealtimeDamping);
        }
        x.states[i] = s;
    }
}

// Impact (velocity) phase set-up, :232-329.
void impact_stages(Ctx &x) {
    for (uint32_t active = 0; active < x.np; ++active) {                // impact_gaps
        uint32_t a = x.pairs[active] / x.n, b = x.pairs[active] % x.n;
        if (x.B[a].impact == 0u && x.B[b].impact == 0u) continue;
        PmoManifold &m = x.M[active];
        for (uint32_t p = 0; p < m.count; ++p) {
            PmoContact &c = m.contacts[p];
            V3 armA = sub(c.point, x.B[a].target.position), armB = sub(c.point, x.B[b].target.position);
            Vec6 ja = jac(mul(c.normal, -1.0f), mul(cross(armA, c.normal), -1.0f));
            Vec6 jb = jac(c.normal, cross(armB, c.normal));
            float gap = fmax_(0.0f, -(c.penetration + dot6(ja, disp(x.B[a])) + dot6(jb, disp(x.B[b]))));
            c.avbd_error.x = gap;
        }
    }
    for (uint32_t i = 0; i < x.n; ++i) {                                // impact_bodies
        if (x.B[i].impact != 0u || x.P[i].motion != kDynamic) {
            x.B[i].target = x.states[i];
            std::memset(x.B[i].displacement, 0, sizeof x.B[i].displacement);
            x.B[i].orientation_delta = {0, 0, 0, 1};
        }
    }
    for (uint32_t k = 0; k < x.nj; ++k) {                               // impact_constraints: joints
        PmoJoint &j = x.J[k];
        if (j.avbd_count == 0u || j.state_enabled == 0u) continue;
        if (x.B[j.dense_a].impact == 0u && x.B[j.dense_b].impact == 0u) continue;
        Block scratch = zero_block();
        visit_joint(x, j, scratch, true, kImpact);
    }
    for (uint32_t active = 0; active < x.np; ++active) {                // impact_constraints: contacts
        uint32_t a = x.pairs[active] / x.n, b = x.pairs[active] % x.n;
        if (x.B[a].impact == 0u && x.B[b].impact == 0u) continue;
        Vec6 va = jac(mul(x.states[a].linear_velocity, x.dt), mul(x.states[a].angular_velocity, x.dt));
        Vec6 vb = jac(mul(x.states[b].linear_velocity, x.dt), mul(x.states[b].angular_velocity, x.dt));
        PmoManifold &m = x.M[active];
        for (uint32_t p = 0; p < m.count; ++p) {
            PmoContact &c = m.contacts[p];
            float gap = c.avbd_error.x;
            // rigid_avbd_contact_rows (:447-471): same Jacobians as the
            // solve rows, anchors from the contact.
            Vec6 ja[3], jb[3];
            float err[3];
            solve_contact_rows(c, x.B[a], x.B[b], ja, jb, err);
            float restitution = c.avbd_cached == 0u && c.initial_normal_speed < -1.0f
                ? fmin_(x.P[a].restitution, x.P[b].restitution) : 0.0f;
            c.avbd_error = {dot6(ja[0], va) + dot6(jb[0], vb) - gap - restitution * c.initial_normal_speed * x.dt,
                            dot6(ja[1], va) + dot6(jb[1], vb),
                            dot6(ja[2], va) + dot6(jb[2], vb)};
        }
    }
}

// rigid_avbd_project_hinge_anchors (solve.slang:174-209): after the pose
// phase, enabled hinges move their bodies' anchors together, split by
// inverse mass, in joint index order.
void project_hinges(Ctx &x) {
    for (uint32_t k = 0; k < x.nj; ++k) {
        const PmoJoint &j = x.J[k];
        if (j.avbd_count == 0u || j.state_enabled == 0u || j.options.type != kHinge) continue;
        const uint32_t a = j.dense_a, b = j.dense_b;
        const float ima = x.P[a].motion == kDynamic ? x.P[a].inverse_mass : 0.0f;
        const float imb = x.P[b].motion == kDynamic ? x.P[b].inverse_mass : 0.0f;
        const float sum = ima + imb;
        if (sum <= kEpsilon) continue;
        const PmoState sa = x.states[a], sb = x.states[b];
        const V3 pa = add(sa.position, rotate(sa.orientation, j.options.local_anchor_a));
        const V3 pb = add(sb.position, rotate(sb.orientation, j.options.local_anchor_b));
        const V3 e = sub(pb, pa);
        x.states[a].position = add(x.states[a].position, mul(e, ima / sum));
        x.states[b].position = sub(x.states[b].position, mul(e, imb / sum));
    }
}

struct SolveClock {
    double prepare_ms = 0.0, solve_ms = 0.0;
    uint64_t body_iterations = 0;
};

int prepare_solve(Ctx &x, V3 gravity, uint32_t base_iterations, uint32_t requested, uint32_t profile,
                  SolveClock *clock = nullptr) {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    prepare(x, gravity, base_iterations, requested, profile);
    const auto t1 = clk::now();
    const bool realtime = x.S->offsets[0] == kRealtimeProfile;
    const uint32_t phases = realtime ? 1u : 2u;
    std::vector<uint32_t> by_color;
    std::vector<PointRef> pts;
    for (uint32_t phase = 0; phase < phases; ++phase) {
        if (phase != 0u && x.S->counts[0] == 0u) break;
        if (phase != 0u) impact_stages(x);
        // Fixed budget, or (to-residual mode) until the residual after a
        // dual update is at most the tolerance, the budget being the cap;
        // the impulses are then published from that update.
        uint32_t executed = 0;
        for (uint32_t it = 0; it < x.S->budget; ++it) {
            for (uint32_t color = 0; color < x.S->remaining; ++color) {
                for (uint32_t i = 0; i < x.n; ++i) {
                    if (x.B[i].color != color) continue;
                    if (phase != 0u && x.B[i].impact == 0u) continue;
                    solve_body(x, i, phase, pts);
                }
            }
            advance_duals(x, phase, it);
            executed = it + 1u;
            if (x.tolerance > 0.0f && (x.last_residual <= x.tolerance || executed == x.S->budget)) {
                publish_impulses(x, phase);
                break;
            }
        }
        x.executed[phase] = executed;
        if (!(x.planted & 1u)) free_motion(x, phase);
        apply_states(x, phase, realtime);
        if (phase == 0u) project_hinges(x);
        x.S->counts[phase == 0u ? 6u : 7u] = executed;
        x.S->maximum_passes += executed;
        if (clock) {
            uint64_t bodies = 0;
            for (uint32_t i = 0; i < x.n; ++i)
                bodies += dyn(x, i) && (phase == 0u || x.B[i].impact != 0u);
            clock->body_iterations += bodies * executed;
        }
    }
    if (clock) {
        const auto t2 = clk::now();
        clock->prepare_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
        clock->solve_ms += std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
    return 0;
}

// ---------------------------------------------------------------------
// World: the production step order of emit_rigid_solve
// (production_realtime_backend.cu.inc:7664-7861) preceded by the predictor
// and proxy phases (:2642-2687).
// ---------------------------------------------------------------------
struct Proxy { V3 mn, mx; };

struct CachedPair {
    uint32_t count;
    std::array<V3, 8> local_point, normal, anchor_a, anchor_b;
    std::array<float, 8> normal_impulse;
    std::array<V3, 8> friction_impulse;
    std::array<std::array<PmoDual, 3>, 8> dual;
};

struct World {
    PmoWorldDesc desc;
    std::vector<Shape> shapes;
    std::vector<PmoParams> params;
    std::vector<PmoState> states;
    std::vector<PmoAvbdBody> bodies;
    std::vector<uint32_t> pairs;
    std::vector<PmoManifold> manifolds;
    std::vector<Proxy> proxies;
    PmoSchedule schedule{};
    PmoWorldStats stats{};
    SolveClock clock;
    double step_ms = 0.0;
    std::map<uint32_t, CachedPair> cache;
    uint32_t cache_bodies = 0;
    uint64_t epoch = 0;
    // One-step inputs (BodyAccumulator, KinematicTarget), cleared by every
    // step as pm_clear_rigid_inputs does (geometry.slang:2877-2886).
    std::vector<PmoJoint> joints;
    std::vector<float> residual = std::vector<float>(2 * kResidualSlots, std::nanf(""));
    std::vector<float> residual_max = std::vector<float>(2 * kResidualSlots, std::nanf(""));
    uint32_t executed[2] = {0u, 0u};
    std::vector<std::array<V3, 4>> accumulators;   // force, torque, impulse, angular impulse
    std::vector<PmoState> targets;
    std::vector<uint8_t> target_active;
    uint32_t substep = 0;                          // substep of the current frame
};

// pm_integrate_rigid_bodies, geometry.slang:524-612: one substep of a
// frame, with remaining_substeps substeps left (the kinematic target moves
// 1/remaining of the way) and the impulses applied in the frame's first.
PmoState integrate(const PmoParams &b, const PmoState &prev, const std::array<V3, 4> &acc, bool target_active,
                   const PmoState &target, V3 g, float dt, uint32_t remaining, bool apply_impulses) {
    PmoState next = prev;
    if (b.motion == kStatic) {
        next.linear_velocity = {0, 0, 0};
        next.angular_velocity = {0, 0, 0};
        return next;
    }
    if (b.motion == kKinematic) {
        if (target_active) {
            const float fraction = 1.0f / float(remaining);
            next.position = add(prev.position, mul(sub(target.position, prev.position), fraction));
            Q t = target.orientation;
            const Q o = prev.orientation;
            if (o.x * t.x + o.y * t.y + o.z * t.z + o.w * t.w < 0.0f) t = {-t.x, -t.y, -t.z, -t.w};
            next.orientation = qnormalize({o.x + (t.x - o.x) * fraction, o.y + (t.y - o.y) * fraction,
                                           o.z + (t.z - o.z) * fraction, o.w + (t.w - o.w) * fraction});
            next.linear_velocity = mul(sub(next.position, prev.position), 1.0f / dt);
            next.angular_velocity = quaternion_delta_velocity(prev.orientation, next.orientation, dt);
        } else {
            next.linear_velocity = {0, 0, 0};
            next.angular_velocity = {0, 0, 0};
        }
        return next;
    }
    const V3 force = acc[0], torque = acc[1], imp = acc[2], aimp = acc[3];
    next.linear_velocity = add(next.linear_velocity, mul(add(g, mul(force, b.inverse_mass)), dt));
    next.angular_velocity = add(next.angular_velocity, mul(inverse_inertia_world(b, next, torque), dt));
    if (apply_impulses) {
        next.linear_velocity = add(next.linear_velocity, mul(imp, b.inverse_mass));
        next.angular_velocity = add(next.angular_velocity, inverse_inertia_world(b, next, aimp));
    }
    next.linear_velocity = mul(next.linear_velocity, 1.0f / (1.0f + b.linear_damping * dt));
    next.angular_velocity = mul(next.angular_velocity, 1.0f / (1.0f + b.angular_damping * dt));
    next.linear_velocity = clamp_length(next.linear_velocity, b.maximum_linear_speed);
    next.angular_velocity = clamp_length(next.angular_velocity, b.maximum_angular_speed);
    next.position = add(next.position, mul(next.linear_velocity, dt));
    Q w = {next.angular_velocity.x, next.angular_velocity.y, next.angular_velocity.z, 0.0f};
    Q d = qmul(w, next.orientation);
    next.orientation = qnormalize({next.orientation.x + 0.5f * d.x * dt, next.orientation.y + 0.5f * d.y * dt,
                                   next.orientation.z + 0.5f * d.z * dt, next.orientation.w + 0.5f * d.w * dt});
    return next;
}

// pm_emit_rigid_spatial_proxies, geometry.slang:716-827, with the
// production arguments path_deviation 0, correction_budget 1, epsilon 1e-5
// (production_realtime_backend.cu.inc:2663-2684).
Proxy make_proxy(const PmoParams &p, const Shape &s, const PmoState &prev, const PmoState &pred) {
    const float eps = 1.0e-5f;
    float reach = len(s.center) + s.radius + p.collision_margin + eps;
    float cr = correction_reach(p, s);
    V3 bc = {fmin_(1.0f, cr), fmin_(1.0f, cr), fmin_(1.0f, cr)};
    V3 e = add(add(V3{0, 0, 0}, bc), V3{reach, reach, reach});
    return {sub(cmin(prev.position, pred.position), e), add(cmax(prev.position, pred.position), e)};
}

inline bool overlap(const Proxy &a, const Proxy &b) {          // spatial_build.slang:631-635
    return a.mn.x <= b.mx.x && a.mx.x >= b.mn.x && a.mn.y <= b.mx.y && a.mx.y >= b.mn.y &&
           a.mn.z <= b.mx.z && a.mx.z >= b.mn.z;
}

// Candidate pairs (H3-H6): every proxy overlap whose first body is dynamic,
// filtered as pm_filter_rigid_spatial_candidates (geometry.slang:897-918)
// with identity order = dense order (ids {index = dense, generation 1}),
// no fixed groups and no constraints; ascending pair code.
void find_pairs(World &w, const std::vector<PmoState> &pred) {
    const uint32_t n = uint32_t(w.params.size());
    w.pairs.clear();
    // Collision exclusions of joints (geometry.slang:494-511, 3108-3136):
    // enabled, unbroken joints with disable_collisions exclude their pair;
    // fixed ones also join their bodies' weld group, whose members never
    // collide.
    std::vector<uint32_t> weld(n);
    for (uint32_t i = 0; i < n; ++i) weld[i] = i;
    auto weld_root = [&](uint32_t v) { while (weld[v] != v) v = weld[v]; return v; };
    std::vector<uint64_t> disabled;
    for (const PmoJoint &j : w.joints) {
        const PmoJointOptions &o = j.options;
        if (j.alive == 0u || o.enabled == 0u || j.state_broken != 0u || o.disable_collisions == 0u) continue;
        const uint32_t a = o.body_a.index < n && o.body_a.generation == 1u ? o.body_a.index : kInvalid;
        const uint32_t b = o.body_b.index < n && o.body_b.generation == 1u ? o.body_b.index : kInvalid;
        if (a == kInvalid || b == kInvalid) continue;
        disabled.push_back((uint64_t(std::min(a, b)) << 32) | std::max(a, b));
        if (o.type != kFixed) continue;
        const uint32_t ra = weld_root(a), rb = weld_root(b);
        if (ra != rb) weld[std::max(ra, rb)] = std::min(ra, rb);
    }
    for (uint32_t i = 0; i < n; ++i) weld[i] = weld_root(i);
    std::sort(disabled.begin(), disabled.end());
    auto keep = [&](uint32_t a, uint32_t b) {
        if (w.params[a].motion != kDynamic || a == b) return false;
        if (w.params[b].motion == kDynamic && !(a < b)) return false;
        if (weld[a] == weld[b]) return false;
        if (!disabled.empty() &&
            std::binary_search(disabled.begin(), disabled.end(), (uint64_t(std::min(a, b)) << 32) | std::max(a, b)))
            return false;
        return true;
    };
    if (w.desc.flags & 1u) {
        for (uint32_t a = 0; a < n; ++a)
            for (uint32_t b = 0; b < n; ++b)
                if (a != b && keep(a, b) && overlap(w.proxies[a], w.proxies[b])) w.pairs.push_back(a * n + b);
        return;
    }
    // Sort-and-sweep on x: same inclusive comparisons, so the same set.
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return w.proxies[a].mn.x < w.proxies[b].mn.x; });
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t a = order[k];
        for (uint32_t j = k + 1; j < n; ++j) {
            uint32_t b = order[j];
            if (w.proxies[b].mn.x > w.proxies[a].mx.x) break;
            if (!overlap(w.proxies[a], w.proxies[b])) continue;
            if (keep(a, b)) w.pairs.push_back(a * n + b);
            if (keep(b, a)) w.pairs.push_back(b * n + a);
        }
    }
    std::sort(w.pairs.begin(), w.pairs.end());
    (void)pred;
}

uint32_t slot_capacity(const World &w);

// rigid_fixed_hinge_inverse_moment, geometry.slang:384-402: the inverse
// moment about a guide axis through its anchor (a piston's turn).
float fixed_hinge_inverse_moment(const PmoParams &p, const PmoState &s, const PmoHingeFrame &f) {
    if ((f.fixed == 0u && f.axial_rotation == 0u) || p.inverse_mass <= kEpsilon) return 0.0f;
    V3 la = inv_rotate(s.orientation, f.axis);
    float center = la.x * la.x / fmax_(p.inverse_inertia_local.x, kEpsilon) +
                   la.y * la.y / fmax_(p.inverse_inertia_local.y, kEpsilon) +
                   la.z * la.z / fmax_(p.inverse_inertia_local.z, kEpsilon);
    V3 arm = sub(s.position, f.anchor);
    V3 perpendicular = sub(arm, mul(f.axis, dot(arm, f.axis)));
    float pivot = center + len2(perpendicular) / p.inverse_mass;
    return pivot > kEpsilon ? 1.0f / pivot : 0.0f;
}

// rigid_hinge_contact_frame, geometry.slang:404-491: over the body's
// enabled, unbroken joints in index order, the first hinge (or slider or
// piston held by a static body with no breaking threshold) gives the
// frame: its anchor and axis, from the static body when it holds the
// joint; a slider or piston counts only as the body's single joint.  Fixed
// and point members are flagged.  `reference` is the anchor of a
// non-axial frame, else the body's position.  Body ids are {dense index, 1}.
PmoHingeFrame hinge_frame(const World &w, uint32_t dense, const std::vector<PmoState> &states, const PmoState &state,
                          V3 &reference) {
    PmoHingeFrame r;
    std::memset(&r, 0, sizeof r);
    const uint32_t n = uint32_t(w.params.size());
    const PmoParams &pd = w.params[dense];
    r.static_body = pd.motion == kStatic ? 1u : 0u;
    if (pd.inverse_mass <= 0.0f) {
        reference = state.position;
        return r;
    }
    if ((w.desc.flags >> 1) & 4u) {          // planted mistake: joints shape no contact
        reference = state.position;
        return r;
    }
    auto same = [&](const PmoBodyId &id, uint32_t d) { return id.index == d && id.generation == 1u; };
    auto find = [&](const PmoBodyId &id) { return id.index < n && id.generation == 1u ? id.index : kInvalid; };
    uint32_t incident = 0;
    for (const PmoJoint &j : w.joints) {
        const PmoJointOptions &o = j.options;
        if (j.alive == 0u || o.enabled == 0u || j.state_broken != 0u) continue;
        const bool is_a = same(o.body_a, dense), is_b = same(o.body_b, dense);
        if (!is_a && !is_b) continue;
        ++incident;
        if (o.type == kFixed) r.fixed_member = 1u;
        if (o.type == kPoint) r.point_member = 1u;
        const bool axial = o.type == kPiston || o.type == kSlider;
        const uint32_t other = find(is_a ? o.body_b : o.body_a);
        const bool static_anchor = other != kInvalid && w.params[other].motion == kStatic;
        if (r.present != 0u || (o.type != kHinge && !(axial && static_anchor && o.breaking_impulse_threshold <= 0.0f)))
            continue;
        const V3 local_axis = axial ? V3{1.0f, 0.0f, 0.0f} : V3{0.0f, 0.0f, 1.0f};
        r.local_anchor = is_a ? o.local_anchor_a : o.local_anchor_b;
        const Q local_orientation = is_a ? o.local_orientation_a : o.local_orientation_b;
        r.anchor = add(state.position, rotate(state.orientation, r.local_anchor));
        r.axis = normalized_or(rotate(qmul(state.orientation, local_orientation), local_axis), local_axis);
        r.present = 1u;
        r.fixed = static_anchor && !axial ? 1u : 0u;
        r.axial = static_anchor && axial ? 1u : 0u;
        r.axial_rotation = r.axial != 0u && o.type == kPiston ? 1u : 0u;
        if (static_anchor) {
            const V3 other_anchor = is_a ? o.local_anchor_b : o.local_anchor_a;
            const Q other_orientation = is_a ? o.local_orientation_b : o.local_orientation_a;
            const PmoState &os = states[other];
            r.anchor = add(os.position, rotate(os.orientation, other_anchor));
            r.axis = normalized_or(rotate(qmul(os.orientation, other_orientation), local_axis), r.axis);
        }
    }
    if (r.axial != 0u && incident != 1u) {
        r.axial = 0u;
        r.axial_rotation = 0u;
        r.present = 0u;
    }
    if (r.axial != 0u) {
        r.inverse_mass = pd.inverse_mass;
        r.inverse_moment = fixed_hinge_inverse_moment(pd, state, r);
    }
    reference = r.present != 0u && r.axial == 0u ? r.anchor : state.position;
    return r;
}

// Per pair narrowphase, in the reference's dispatch order
// (production_realtime_backend.cu.inc:7710-7776): leaf-pair candidates
// (pm_generate_rigid_leaf_pairs, geometry.slang:2303-2492), the convex face
// path (pm_prepare_rigid_face_manifolds, :2743-2803) which claims the pairs
// it handles, the triangle path over the remaining pairs' leaf candidates
// in candidate order (pm_evaluate_rigid_leaf_pairs, :2525-2706), and
// pm_finalize_rigid_contact_manifolds (:2807-2873), which runs the BVH
// traversal for pairs without a candidate slot or with too many
// candidates.  Round shapes (no reference counterpart) take the analytic
// path instead.
// The timestep of one pass of the pipeline: the frame's dt over its
// substeps.  Every substep runs the whole pipeline (integration,
// narrowphase, contact cache, solve) at this timestep.
inline float substep_dt(const World &w) {
    const uint32_t substeps = w.desc.substeps ? w.desc.substeps : 1u;
    return substeps == 1u ? w.desc.dt : w.desc.dt / float(substeps);
}

void narrowphase(World &w, const std::vector<PmoState> &prev, const std::vector<PmoState> &pred) {
    const uint32_t n = uint32_t(w.params.size());
    const uint32_t slots = slot_capacity(w);
    const uint32_t capacity = w.desc.body_capacity ? w.desc.body_capacity : n;
    const uint32_t per_slot = capacity <= 8u ? 4096u : 512u;          // geometry_storage.hpp:165-166
    w.manifolds.assign(w.pairs.size(), PmoManifold{});
    w.stats.unsupported_pairs = 0;
    w.stats.triangle_pairs = 0;
    w.stats.fallback_pairs = 0;
    std::vector<std::pair<uint32_t, uint32_t>> candidates;
    for (size_t active = 0; active < w.pairs.size(); ++active) {
        uint32_t a = w.pairs[active] / n, b = w.pairs[active] % n;
        const Shape &sa = w.shapes[w.params[a].mesh_index], &sb = w.shapes[w.params[b].mesh_index];
        float margin = w.params[a].collision_margin + w.params[b].collision_margin;
        float cr = correction_reach(w.params[a], sa) + correction_reach(w.params[b], sb);
        // Frames at the predicted states (both bodies) and the body's at
        // its previous state, for its reference point (geometry.slang:
        // 2561-2578, 2771-2779, 2834-2846).
        V3 ref_prev, ref, collider_ref;
        (void)hinge_frame(w, a, prev, prev[a], ref_prev);
        const PmoHingeFrame ha = hinge_frame(w, a, pred, pred[a], ref);
        const PmoHingeFrame hb = hinge_frame(w, b, pred, pred[b], collider_ref);
        const bool guided = guided_pair(ha, hb);
        PmoManifold m;
        clear_manifold(m);
        if (sa.kind != kMesh || sb.kind != kMesh) {
            if (!analytic_manifold(pred[a], sa, pred[b], sb, margin, margin + cr, m)) {
                clear_manifold(m);
                w.stats.unsupported_pairs++;
            }
        } else {
            // The leaf-pair candidates use the motion test alone; the
            // triangle ranges of a guided pair always take the swept pass
            // (geometry.slang:2350, 2574-2579, 2167-2170).
            bool swept = requires_swept(prev[a], pred[a], sa, prev[b], pred[b], sb, margin);
            const bool swept_ranges = guided || swept;
            bool face = false;
            if (convex_face_shape(sa) && convex_face_shape(sb) && !guided) {
                if (swept) face = swept_convex_face_manifold(prev[a], pred[a], sa, prev[b], pred[b], sb, margin, m);
                if (!face) {
                    face = convex_face_manifold(pred[a], sa, pred[b], sb, margin, cr, m);
                    if (swept && m.count == 0u) face = false;
                }
            }
            if (!face) {
                clear_manifold(m);
                TrianglePair P{&sa, &sb, prev[a], pred[a], prev[b], pred[b], ha, hb, margin, substep_dt(w), ref_prev, ref};
                w.stats.triangle_pairs++;
                if (active < slots && leaf_candidates(P, swept, per_slot, candidates)) {
                    const float spacing = fmax_(margin * 2.0f, 1.0e-4f);
                    PmoManifold local;
                    for (auto &c : candidates) {
                        clear_manifold(local);
                        collide_leaf_ranges(P, swept_ranges, sa.nodes[c.first], sb.nodes[c.second], local);
                        for (uint32_t p = 0; p < local.count; ++p) add_contact(m, local.contacts[p], spacing);
                    }
                } else {
                    w.stats.fallback_pairs++;
                    collide_meshes(P, swept_ranges, m);
                }
            }
        }
        for (uint32_t p = 0; p < m.count; ++p) {
            m.contacts[p].body_hinge = ha;
            m.contacts[p].collider_hinge = hb;
        }
        bool persistent = (sa.kind != kMesh || (sa.has_planes && sa.index_count() <= 96u)) &&
            (w.params[b].motion != kDynamic || sb.kind != kMesh || (sb.has_planes && sb.index_count() <= 96u));
        m.initial_relative_position = sub(pred[a].position, pred[b].position);
        for (uint32_t p = 0; p < m.count; ++p) {
            // geometry.slang:2855-2862: guided contacts persist; contacts of
            // a hinged or fixed member do not.
            const PmoContact &c = m.contacts[p];
            m.contacts[p].persistent = guided_contact(c) ||
                (persistent && c.body_hinge.present == 0u && c.collider_hinge.present == 0u &&
                 c.body_hinge.fixed_member == 0u && c.collider_hinge.fixed_member == 0u) ? 1u : 0u;
            m.contacts[p].initial_normal_speed = contact_normal_speed(pred[a], pred[b], m.contacts[p].point, m.contacts[p].normal);
            m.contacts[p].reported_normal_impulse = m.contacts[p].accumulated_normal_impulse;
            m.contacts[p].reported_friction_impulse = m.contacts[p].accumulated_friction_impulse;
        }
        w.manifolds[active] = m;
    }
}

uint32_t slot_capacity(const World &w) {                          // resource_plan.hpp:151-156
    uint64_t cap = w.desc.body_capacity ? w.desc.body_capacity : w.params.size();
    uint64_t man = cap * (cap - 1) / 2;
    return uint32_t(std::min<uint64_t>(man, std::max<uint64_t>(4096, cap * 8)));
}

// Response slots (pm_prepare_parallel_contact_events, geometry.slang:
// 2902-2939) and pm_load_rigid_contact_cache (:2941-3009).
void load_cache(World &w, const std::vector<PmoState> &pred) {
    const uint32_t n = uint32_t(w.params.size());
    uint32_t slot = 0;
    for (auto &m : w.manifolds) m.response_slot = m.count != 0u ? slot++ : kInvalid;
    w.stats.cache_hits = 0;
    for (size_t active = 0; active < w.pairs.size(); ++active) {
        auto it = w.cache.find(w.pairs[active]);
        if (it == w.cache.end()) continue;
        const CachedPair &saved = it->second;
        uint32_t body = w.pairs[active] / n;
        PmoManifold &m = w.manifolds[active];
        uint32_t used = 0;
        for (uint32_t p = 0; p < m.count; ++p) {
            PmoContact &c = m.contacts[p];
            V3 local = inv_rotate(pred[body].orientation, sub(c.point, pred[body].position));
            float nearest = 0.02f * 0.02f;
            uint32_t match = 8;
            for (uint32_t q = 0; q < saved.count; ++q) {
                if ((used & (1u << q)) != 0u || dot(c.normal, saved.normal[q]) < 0.99f) continue;
                float d = len2(sub(local, saved.local_point[q]));
                if (d < nearest) { nearest = d; match = q; }
            }
            if (match == 8u) continue;
            used |= 1u << match;
            for (int ax = 0; ax < 3; ++ax) c.avbd_dual[ax] = saved.dual[match][ax];
            V3 t0 = normalized_or(cross(c.normal, std::fabs(c.normal.x) < 0.57735f ? V3{1, 0, 0} : V3{0, 1, 0}), {0, 0, 1});
            V3 t1 = cross(c.normal, t0);
            c.avbd_dual[1].lambda = dot(saved.friction_impulse[match], t0) / substep_dt(w);
            c.avbd_dual[2].lambda = dot(saved.friction_impulse[match], t1) / substep_dt(w);
            c.avbd_anchor_a = saved.anchor_a[match];
            c.avbd_anchor_b = saved.anchor_b[match];
            c.avbd_cached = 1u;
            c.accumulated_normal_impulse = saved.normal_impulse[match];
            c.accumulated_friction_impulse = sub(saved.friction_impulse[match],
                mul(c.normal, dot(saved.friction_impulse[match], c.normal)));
            w.stats.cache_hits++;
        }
    }
}

// pm_save_rigid_contact_cache, geometry.slang:3011-3070.
void save_cache(World &w) {
    const uint32_t n = uint32_t(w.params.size());
    const uint32_t cap = slot_capacity(w);
    w.cache.clear();
    for (size_t active = 0; active < w.pairs.size(); ++active) {
        const PmoManifold &m = w.manifolds[active];
        if (m.response_slot >= cap) continue;
        uint32_t body = w.pairs[active] / n;
        CachedPair s{};
        s.count = m.count;
        for (uint32_t p = 0; p < m.count; ++p) {
            const PmoContact &c = m.contacts[p];
            s.local_point[p] = inv_rotate(w.states[body].orientation, sub(c.point, w.states[body].position));
            s.normal[p] = c.normal;
            s.normal_impulse[p] = c.accumulated_normal_impulse;
            s.friction_impulse[p] = c.accumulated_friction_impulse;
            for (int ax = 0; ax < 3; ++ax) s.dual[p][ax] = c.avbd_dual[ax];
            s.anchor_a[p] = c.avbd_anchor_a;
            s.anchor_b[p] = c.avbd_anchor_b;
        }
        w.cache[w.pairs[active]] = s;
    }
}

void world_step(World &w) {
    const uint32_t n = uint32_t(w.params.size());
    const V3 g = {w.desc.gravity[0], w.desc.gravity[1], w.desc.gravity[2]};
    const uint32_t substeps = w.desc.substeps ? w.desc.substeps : 1u;
    const float dt = substep_dt(w);
    if (w.cache_bodies != n) { w.cache.clear(); w.cache_bodies = n; }
    std::vector<PmoState> prev = w.states, pred(n);
    w.accumulators.resize(n, std::array<V3, 4>{});
    w.targets.resize(n, PmoState{});
    w.target_active.resize(n, 0);
    for (uint32_t i = 0; i < n; ++i)
        pred[i] = integrate(w.params[i], prev[i], w.accumulators[i], w.target_active[i] != 0, w.targets[i], g, dt,
                            substeps - w.substep, w.substep == 0u);
    // pm_clear_rigid_inputs after the frame's last substep.
    if (w.substep + 1u == substeps) {
        std::fill(w.accumulators.begin(), w.accumulators.end(), std::array<V3, 4>{});
        std::fill(w.target_active.begin(), w.target_active.end(), 0);
    }
    w.substep = (w.substep + 1u) % substeps;
    w.proxies.resize(n);
    for (uint32_t i = 0; i < n; ++i)
        w.proxies[i] = make_proxy(w.params[i], w.shapes[w.params[i].mesh_index], prev[i], pred[i]);
    find_pairs(w, pred);
    narrowphase(w, prev, pred);
    load_cache(w, pred);
    if (w.bodies.size() != n) w.bodies.assign(n, PmoAvbdBody{});
    Ctx x{w.params.data(), prev.data(), pred.data(), n, w.manifolds.data(), w.pairs.data(),
          uint32_t(w.pairs.size()), w.bodies.data(), &w.schedule, dt, w.desc.lanes};
    x.J = w.joints.data();
    x.nj = uint32_t(w.joints.size());
    x.planted = (w.desc.flags >> 1) & 3u;
    std::fill(w.residual.begin(), w.residual.end(), std::nanf(""));
    std::fill(w.residual_max.begin(), w.residual_max.end(), std::nanf(""));
    x.residual = w.residual.data();
    x.residual_max = w.residual_max.data();
    x.tolerance = w.desc.residual_tolerance;
    prepare_solve(x, g, w.desc.base_iterations, w.desc.requested_iterations, w.desc.profile_version, &w.clock);
    w.executed[0] = x.executed[0];
    w.executed[1] = x.executed[1];
    for (uint32_t i = 0; i < n; ++i) {                               // pm_clamp_rigid_speeds, :3086-3100
        if (w.params[i].motion != kDynamic) continue;
        pred[i].linear_velocity = clamp_length(pred[i].linear_velocity, w.params[i].maximum_linear_speed);
        pred[i].angular_velocity = clamp_length(pred[i].angular_velocity, w.params[i].maximum_angular_speed);
    }
    w.states = pred;
    save_cache(w);
    ++w.epoch;
    w.stats.pairs = uint32_t(w.pairs.size());
    w.stats.manifolds = 0;
    w.stats.points = 0;
    for (auto &m : w.manifolds) { w.stats.manifolds += m.count != 0u; w.stats.points += m.count; }
    w.stats.colors = w.schedule.remaining;
    w.stats.islands = w.schedule.island_count;
    w.stats.impact = w.schedule.counts[0];
    w.stats.budget = w.schedule.budget;
    w.stats.phases = w.schedule.counts[7] != 0u ? 2u : 1u;
}

}  // namespace

// ---------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------
extern "C" {

void pmo_warm_start(const float d[2], float s, float out[2]) {
    Dual r = warm_start({d[0], d[1]}, s);
    out[0] = r.lambda; out[1] = r.penalty;
}
float pmo_force(const float d[2], float e, float lo, float hi) { return dual_force({d[0], d[1]}, e, lo, hi); }
void pmo_update_dual(const float d[2], float e, float beta, float lo, float hi, float s, float out[2]) {
    Dual r = update_dual({d[0], d[1]}, e, beta, lo, hi, s);
    out[0] = r.lambda; out[1] = r.penalty;
}
void pmo_project_contact(const float f[3], float mu, float out[3]) {
    Force3 r = project_contact({f[0], f[1], f[2]}, mu);
    out[0] = r.normal; out[1] = r.t0; out[2] = r.t1;
}
void pmo_stiffness_scales(const float f[3], float mu, float out[3]) {
    Force3 r = stiffness_scales({f[0], f[1], f[2]}, mu);
    out[0] = r.normal; out[1] = r.t0; out[2] = r.t1;
}
void pmo_block_add_row(float h[21], float g[6], const float j[6], float k, float f) {
    Block b;
    std::memcpy(b.h, h, sizeof b.h); std::memcpy(b.g, g, sizeof b.g);
    Vec6 v; std::memcpy(v.v, j, sizeof v.v);
    add_row(b, v, k, f);
    std::memcpy(h, b.h, sizeof b.h); std::memcpy(g, b.g, sizeof b.g);
}
int pmo_block_solve(const float h[21], const float g[6], float delta[6]) {
    Block b;
    std::memcpy(b.h, h, sizeof b.h); std::memcpy(b.g, g, sizeof b.g);
    Vec6 d;
    bool ok = block_solve(b, d);
    std::memcpy(delta, d.v, sizeof d.v);
    return ok ? 1 : 0;
}
static Vec6 v6(const float *p) { Vec6 v; std::memcpy(v.v, p, sizeof v.v); return v; }
float pmo_row_error(const float row[22], const float a[6], const float b[6]) { return row_error(row_from(row), v6(a), v6(b)); }
float pmo_row_trial(const float row[22], const float a[6], const float b[6], float dt) { return row_trial(row_from(row), v6(a), v6(b), dt); }
float pmo_row_force(const float row[22], const float a[6], const float b[6], float dt) { return row_force(row_from(row), v6(a), v6(b), dt); }
void pmo_row_accumulate(float h[21], float g[6], const float row[22], int first, const float a[6], const float b[6], float dt) {
    Block blk;
    std::memcpy(blk.h, h, sizeof blk.h); std::memcpy(blk.g, g, sizeof blk.g);
    row_accumulate(blk, row_from(row), first != 0, v6(a), v6(b), dt);
    std::memcpy(h, blk.h, sizeof blk.h); std::memcpy(g, blk.g, sizeof blk.g);
}
void pmo_row_advance(float row[22], const float a[6], const float b[6]) {
    Row r = row_from(row);
    row_advance(r, v6(a), v6(b));
    row_to(r, row);
}

int pmo_avbd_prepare_solve(const PmoParams *parameters, const PmoState *previous, PmoState *states,
                           uint32_t count, PmoManifold *manifolds, const uint32_t *pairs,
                           uint32_t pair_count, PmoAvbdBody *bodies, PmoSchedule *schedule,
                           float dt, const float gravity[3], uint32_t base_iterations,
                           uint32_t requested_iterations, uint32_t profile_version, uint32_t lanes) {
    if (lanes & (lanes - 1u)) return -1;
    Ctx x{parameters, previous, states, count, manifolds, pairs, pair_count, bodies, schedule, dt, lanes};
    return prepare_solve(x, {gravity[0], gravity[1], gravity[2]}, base_iterations, requested_iterations, profile_version);
}

// The seam with joints (prepare_geometry_constraints' constraints,
// constraint_capacity and ids, geometry_constraints_dispatch.hpp:15-24):
// the joints' body ids resolve through `ids` (rigid_avbd_find_dense,
// geometry_constraints.slang:318-325); the solve runs on a copy whose ids
// are {dense index, 1}, and the solver-owned fields go back to `joints`.
int pmo_avbd_prepare_solve_joints(const PmoParams *parameters, const PmoState *previous, PmoState *states,
                                  uint32_t count, PmoManifold *manifolds, const uint32_t *pairs,
                                  uint32_t pair_count, PmoJoint *joints, uint32_t joint_count,
                                  const PmoBodyId *ids, PmoAvbdBody *bodies, PmoSchedule *schedule, float dt,
                                  const float gravity[3], uint32_t base_iterations, uint32_t requested_iterations,
                                  uint32_t profile_version, uint32_t lanes) {
    if (lanes & (lanes - 1u)) return -1;
    std::vector<PmoJoint> copy(joints, joints + joint_count);
    for (PmoJoint &j : copy)
        for (PmoBodyId *id : {&j.options.body_a, &j.options.body_b}) {
            uint32_t found = kInvalid;
            for (uint32_t i = 0; i < count; ++i)
                if (ids[i].index == id->index && ids[i].generation == id->generation) { found = i; break; }
            *id = found == kInvalid ? PmoBodyId{kInvalid, 0u} : PmoBodyId{found, 1u};
        }
    Ctx x{parameters, previous, states, count, manifolds, pairs, pair_count, bodies, schedule, dt, lanes};
    x.J = copy.data();
    x.nj = joint_count;
    const int r = prepare_solve(x, {gravity[0], gravity[1], gravity[2]}, base_iterations, requested_iterations,
                                profile_version);
    for (uint32_t k = 0; k < joint_count; ++k) {
        PmoJoint &d = joints[k];
        const PmoJoint &s = copy[k];
        d.state_enabled = s.state_enabled;
        d.state_broken = s.state_broken;
        d.applied_impulse = s.applied_impulse;
        d.dense_a = s.dense_a;
        d.dense_b = s.dense_b;
        std::memcpy(d.rows, s.rows, sizeof d.rows);
        d.avbd_count = s.avbd_count;
        d.avbd_next_a = s.avbd_next_a;
        d.avbd_next_b = s.avbd_next_b;
    }
    return r;
}

// The realtime contract: a 1/60 s step within one float ulp and one
// substep (rigid_realtime.hpp:16-26; pmCmdStepSimulation rejects anything
// else, src/api/resources.cpp:1240-1241).
static bool realtime_step_valid(float dt, uint32_t substeps) {
    const float r = 1.0f / 60.0f;
    return std::isfinite(dt) && dt >= std::nextafter(r, 0.0f) && dt <= std::nextafter(r, INFINITY) && substeps <= 1u;
}

void *pmo_world_create(const PmoWorldDesc *desc) {
    if (!desc || !(desc->dt > 0.0f) || (desc->lanes & (desc->lanes - 1u))) return nullptr;
    if (desc->profile_version == 1 && !realtime_step_valid(desc->dt, desc->substeps)) return nullptr;
    World *w = new World();
    w->desc = *desc;
    return w;
}
void pmo_world_destroy(void *world) { delete static_cast<World *>(world); }
int pmo_world_add_mesh(void *world, const float *v, uint32_t nv, const uint32_t *idx, uint32_t ni) {
    World &w = *static_cast<World *>(world);
    Shape s;
    if (!prepare_mesh(s, v, nv, idx, ni)) return -1;
    w.shapes.push_back(std::move(s));
    return int(w.shapes.size() - 1);
}
int pmo_world_add_sphere(void *world, float r) {
    World &w = *static_cast<World *>(world);
    if (!(r > 0.0f)) return -1;
    Shape s;
    s.kind = kSphere;
    s.sphere_radius = r;
    s.minimum = {-r, -r, -r}; s.maximum = {r, r, r};
    s.center = {0, 0, 0}; s.radius = r;
    float i = 0.4f * r * r;
    s.unit_inertia = {i, i, i};
    w.shapes.push_back(s);
    return int(w.shapes.size() - 1);
}
int pmo_world_add_capsule(void *world, float r, float h) {
    World &w = *static_cast<World *>(world);
    if (!(r > 0.0f) || !(h >= 0.0f)) return -1;
    Shape s;
    s.kind = kCapsule;
    s.sphere_radius = r;
    s.half_length = h;
    s.minimum = {-r, -(h + r), -r}; s.maximum = {r, h + r, r};
    s.center = {0, 0, 0}; s.radius = h + r;
    V3 half = {r, h + r, r};
    s.unit_inertia = {(half.y * half.y + half.z * half.z) / 3.0f, (half.x * half.x + half.z * half.z) / 3.0f,
                      (half.x * half.x + half.y * half.y) / 3.0f};
    w.shapes.push_back(s);
    return int(w.shapes.size() - 1);
}
int pmo_world_add_bodies(void *world, uint32_t count, const PmoParams *p, const PmoState *s) {
    World &w = *static_cast<World *>(world);
    for (uint32_t i = 0; i < count; ++i)
        if (p[i].mesh_index >= w.shapes.size() || p[i].motion > kDynamic) return -1;
    w.params.insert(w.params.end(), p, p + count);
    w.states.insert(w.states.end(), s, s + count);
    return 0;
}
// apply_rigid_input (runtime_rigid.cu.inc:265-289): inputs accumulate on
// dynamic bodies until the next step.
int pmo_world_apply_inputs(void *world, uint32_t count, const PmoRigidInput *inputs) {
    World &w = *static_cast<World *>(world);
    const uint32_t n = uint32_t(w.params.size());
    w.accumulators.resize(n, std::array<V3, 4>{});
    for (uint32_t k = 0; k < count; ++k) {
        const PmoRigidInput &in = inputs[k];
        if (in.body_index >= n || w.params[in.body_index].motion != kDynamic) return -1;
        for (int c = 0; c < 3; ++c)
            if (!std::isfinite(in.force[c]) || !std::isfinite(in.torque[c]) || !std::isfinite(in.linear_impulse[c]) ||
                !std::isfinite(in.angular_impulse[c]))
                return -1;
        auto &a = w.accumulators[in.body_index];
        a[0] = add(a[0], V3{in.force[0], in.force[1], in.force[2]});
        a[1] = add(a[1], V3{in.torque[0], in.torque[1], in.torque[2]});
        a[2] = add(a[2], V3{in.linear_impulse[0], in.linear_impulse[1], in.linear_impulse[2]});
        a[3] = add(a[3], V3{in.angular_impulse[0], in.angular_impulse[1], in.angular_impulse[2]});
    }
    return 0;
}

// set_kinematic_target (runtime_rigid.cu.inc:197-230): kinematic bodies
// only, finite, non-zero orientation, stored normalised, for the next step.
int pmo_world_set_kinematic_targets(void *world, uint32_t first, uint32_t count, const PmoState *targets) {
    World &w = *static_cast<World *>(world);
    const uint32_t n = uint32_t(w.params.size());
    if (uint64_t(first) + count > n) return -1;
    w.targets.resize(n, PmoState{});
    w.target_active.resize(n, 0);
    for (uint32_t k = 0; k < count; ++k) {
        PmoState t = targets[k];
        const uint32_t i = first + k;
        if (w.params[i].motion != kKinematic) return -1;
        const float *f = reinterpret_cast<const float *>(&t);
        for (size_t c = 0; c < sizeof(PmoState) / sizeof(float); ++c)
            if (!std::isfinite(f[c])) return -1;
        const Q o = t.orientation;
        if (o.x * o.x + o.y * o.y + o.z * o.z + o.w * o.w <= kEpsilon * kEpsilon) return -1;
        t.orientation = qnormalize(o);
        w.targets[i] = t;
        w.target_active[i] = 1;
    }
    return 0;
}

int pmo_world_add_joints(void *world, uint32_t count, const PmoJointOptions *options) {
    World &w = *static_cast<World *>(world);
    for (uint32_t k = 0; k < count; ++k)
        if (options[k].type > kMotor) return -1;
    for (uint32_t k = 0; k < count; ++k) {
        PmoJoint j;
        std::memset(&j, 0, sizeof j);
        j.options = options[k];
        j.options.local_orientation_a = qnormalize(j.options.local_orientation_a);
        j.options.local_orientation_b = qnormalize(j.options.local_orientation_b);
        j.generation = 1u;
        j.alive = 1u;
        j.state_enabled = j.options.enabled;            // add_rigid_constraint, runtime_rigid.cu.inc:588-589
        j.dense_a = j.dense_b = kInvalid;
        j.avbd_next_a = j.avbd_next_b = kInvalid;
        w.joints.push_back(j);
    }
    return 0;
}
// validate_constraint_options (runtime.cpp:400-456, the checks a joint
// record can fail here) plus add_rigid_constraint's body checks (:552-567).
static bool joint_options_valid(const World &w, const PmoJointOptions &o) {
    const uint32_t n = uint32_t(w.params.size());
    if (o.type > kMotor || o.body_a.index >= n || o.body_b.index >= n) return false;
    if (w.params[o.body_a.index].motion != kDynamic && w.params[o.body_b.index].motion != kDynamic) return false;
    if (!std::isfinite(o.breaking_impulse_threshold) || o.breaking_impulse_threshold < 0.0f) return false;
    return o.solver_iterations >= 1u && o.solver_iterations <= 64u;
}

int pmo_world_write_joints(void *world, uint32_t first, uint32_t count, const PmoJointOptions *options,
                           const uint32_t *active) {
    World &w = *static_cast<World *>(world);
    if (uint64_t(first) + count > w.joints.size() || (count && (!options || !active))) return -1;
    for (uint32_t k = 0; k < count; ++k)
        if (active[k] && !joint_options_valid(w, options[k])) return -1;
    for (uint32_t k = 0; k < count; ++k) {
        PmoJoint &j = w.joints[first + k];
        if (j.alive) {                                  // remove_rigid_constraint
            j.alive = 0u;
            j.state_enabled = j.state_broken = 0u;
            j.applied_impulse = 0.0f;
            j.generation = j.generation + 1u ? j.generation + 1u : 1u;
        }
        if (!active[k]) continue;
        const uint32_t generation = j.generation ? j.generation : 1u;
        std::memset(&j, 0, sizeof j);                   // add_rigid_constraint
        j.options = options[k];
        j.options.local_orientation_a = qnormalize(j.options.local_orientation_a);
        j.options.local_orientation_b = qnormalize(j.options.local_orientation_b);
        j.generation = generation;
        j.alive = 1u;
        j.state_enabled = j.options.enabled;
        j.dense_a = j.dense_b = kInvalid;
        j.avbd_next_a = j.avbd_next_b = kInvalid;
    }
    return 0;
}

int pmo_world_set_rigid_states(void *world, uint32_t first, uint32_t count, const PmoState *states) {
    World &w = *static_cast<World *>(world);
    const uint32_t n = uint32_t(w.params.size());
    if (uint64_t(first) + count > n || (count && !states)) return -1;
    for (uint32_t k = 0; k < count; ++k) {
        const float *f = reinterpret_cast<const float *>(&states[k]);
        for (size_t c = 0; c < sizeof(PmoState) / sizeof(float); ++c)
            if (!std::isfinite(f[c])) return -1;
        const Q o = states[k].orientation;
        if (o.x * o.x + o.y * o.y + o.z * o.z + o.w * o.w <= kEpsilon * kEpsilon) return -1;
    }
    w.accumulators.resize(n, std::array<V3, 4>{});
    w.target_active.resize(n, 0);
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t i = first + k;
        PmoState s = states[k];
        s.orientation = qnormalize(s.orientation);
        w.states[i] = s;
        if (i < w.bodies.size()) w.bodies[i] = PmoAvbdBody{};
        w.accumulators[i] = std::array<V3, 4>{};
        w.target_active[i] = 0;
        for (PmoJoint &j : w.joints)
            if (j.alive && (j.options.body_a.index == i || j.options.body_b.index == i))
                std::memset(j.rows, 0, sizeof j.rows);
    }
    return 0;
}

uint32_t pmo_world_contact_events(void *world, PmoContactEvent *events, uint32_t capacity) {
    World &w = *static_cast<World *>(world);
    const uint32_t n = uint32_t(w.params.size());
    uint32_t cursor = 0;
    for (size_t m = 0; m < w.pairs.size(); ++m) {
        const PmoManifold &mf = w.manifolds[m];
        for (uint32_t k = 0; k < mf.count; ++k, ++cursor) {
            if (cursor >= capacity) continue;
            const PmoContact &c = mf.contacts[k];
            PmoContactEvent &e = events[cursor];
            e.type = 0u;
            e.first = w.pairs[m] / n;
            e.second = w.pairs[m] % n;
            e.position = c.point;
            e.normal = c.normal;
            e.penetration = std::max(0.0f, c.penetration);
            e.normal_impulse = c.persistent ? c.accumulated_normal_impulse : c.reported_normal_impulse;
            e.friction_impulse = c.persistent ? c.accumulated_friction_impulse : c.reported_friction_impulse;
        }
    }
    return cursor;
}

void pmo_world_residuals(void *world, float *out, float *maximum, uint32_t *executed) {
    World &w = *static_cast<World *>(world);
    std::copy(w.residual.begin(), w.residual.end(), out);
    std::copy(w.residual_max.begin(), w.residual_max.end(), maximum);
    executed[0] = w.executed[0];
    executed[1] = w.executed[1];
}
uint32_t pmo_world_joint_count(void *world) { return uint32_t(static_cast<World *>(world)->joints.size()); }
void pmo_world_get_joints(void *world, PmoJoint *out) {
    World &w = *static_cast<World *>(world);
    std::copy(w.joints.begin(), w.joints.end(), out);
}

int pmo_world_step(void *world, uint32_t steps) {
    World &w = *static_cast<World *>(world);
    uint64_t n = w.params.size();
    if (n * n > 0xffffffffull) return -2;  // pair codes a*n+b must fit 32 bits
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t substeps = w.desc.substeps ? w.desc.substeps : 1u;
    for (uint64_t s = 0; s < uint64_t(steps) * substeps; ++s) world_step(w);
    w.step_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}
void pmo_world_timing(void *world, double out[4], int reset) {
    World &w = *static_cast<World *>(world);
    out[0] = w.step_ms;
    out[1] = w.clock.prepare_ms;
    out[2] = w.clock.solve_ms;
    out[3] = double(w.clock.body_iterations);
    if (reset) { w.step_ms = 0.0; w.clock = SolveClock{}; }
}
uint32_t pmo_world_body_count(void *world) { return uint32_t(static_cast<World *>(world)->params.size()); }
void pmo_world_get_states(void *world, PmoState *out) {
    World &w = *static_cast<World *>(world);
    std::copy(w.states.begin(), w.states.end(), out);
}
void pmo_world_set_states(void *world, const PmoState *s) {
    World &w = *static_cast<World *>(world);
    std::copy(s, s + w.states.size(), w.states.begin());
    w.cache.clear();
    for (auto &b : w.bodies) b = PmoAvbdBody{};
}
void pmo_world_get_stats(void *world, PmoWorldStats *out) { *out = static_cast<World *>(world)->stats; }
uint32_t pmo_world_pair_count(void *world) { return uint32_t(static_cast<World *>(world)->pairs.size()); }
void pmo_world_get_pairs(void *world, uint32_t *out) {
    World &w = *static_cast<World *>(world);
    std::copy(w.pairs.begin(), w.pairs.end(), out);
}
void pmo_world_get_manifolds(void *world, PmoManifold *out) {
    World &w = *static_cast<World *>(world);
    std::copy(w.manifolds.begin(), w.manifolds.end(), out);
}
void pmo_world_get_bodies(void *world, PmoAvbdBody *out) {
    World &w = *static_cast<World *>(world);
    std::copy(w.bodies.begin(), w.bodies.end(), out);
}
void pmo_world_get_schedule(void *world, PmoSchedule *out) { *out = static_cast<World *>(world)->schedule; }
void pmo_world_get_proxies(void *world, float *out) {
    World &w = *static_cast<World *>(world);
    for (size_t i = 0; i < w.proxies.size(); ++i) {
        const Proxy &p = w.proxies[i];
        float v[6] = {p.mn.x, p.mn.y, p.mn.z, p.mx.x, p.mx.y, p.mx.z};
        std::memcpy(out + 6 * i, v, sizeof v);
    }
}
void pmo_world_get_shape(void *world, uint32_t shape, float out[16]) {
    World &w = *static_cast<World *>(world);
    const Shape &s = w.shapes[shape];
    float v[16] = {s.minimum.x, s.minimum.y, s.minimum.z, s.maximum.x, s.maximum.y, s.maximum.z,
                   s.center.x, s.center.y, s.center.z, s.radius, s.unit_inertia.x, s.unit_inertia.y,
                   s.unit_inertia.z, float(s.kind), float(s.index_count()), float(s.unique_planes)};
    std::memcpy(out, v, sizeof v);
}
void pmo_world_get_mesh_indices(void *world, uint32_t shape, uint32_t *out) {
    World &w = *static_cast<World *>(world);
    std::copy(w.shapes[shape].indices.begin(), w.shapes[shape].indices.end(), out);
}
void pmo_world_get_mesh_planes(void *world, uint32_t shape, float *out) {
    World &w = *static_cast<World *>(world);
    const Shape &s = w.shapes[shape];
    for (size_t i = 0; i < s.planes.size(); ++i) {
        out[4 * i] = s.planes[i].normal.x; out[4 * i + 1] = s.planes[i].normal.y;
        out[4 * i + 2] = s.planes[i].normal.z; out[4 * i + 3] = s.planes[i].offset;
    }
}


uint32_t pmo_world_get_mesh_bvh(void *world, uint32_t shape, uint32_t *nodes, float *bounds) {
    auto &w = *static_cast<World *>(world);
    if (shape >= w.shapes.size()) return 0;
    const Shape &s = w.shapes[shape];
    for (size_t k = 0; k < s.nodes.size(); ++k) {
        const BvhNode &nd = s.nodes[k];
        if (nodes) {
            nodes[4 * k] = nd.left;
            nodes[4 * k + 1] = nd.right;
            nodes[4 * k + 2] = nd.first;
            nodes[4 * k + 3] = nd.count;
        }
        if (bounds) {
            const float b[6] = {nd.minimum.x, nd.minimum.y, nd.minimum.z, nd.maximum.x, nd.maximum.y, nd.maximum.z};
            std::memcpy(bounds + 6 * k, b, sizeof b);
        }
    }
    return uint32_t(s.nodes.size());
}

uint32_t pmo_world_get_mesh_shell(void *world, uint32_t shape, float *out) {
    auto &w = *static_cast<World *>(world);
    if (shape >= w.shapes.size() || w.shapes[shape].shell.empty()) return 0;
    const Shape &s = w.shapes[shape];
    if (out)
        for (size_t t = 0; t < s.shell.size(); ++t) {
            out[3 * t] = s.shell[t].x;
            out[3 * t + 1] = s.shell[t].y;
            out[3 * t + 2] = s.shell[t].z;
        }
    return uint32_t(s.shell.size());
}

static V3 v3_in(const float *f) { return {f[0], f[1], f[2]}; }
static void v3_out(V3 v, float *f) { f[0] = v.x; f[1] = v.y; f[2] = v.z; }

void pmo_closest_segments(const float p0[3], const float p1[3], const float q0[3], const float q1[3],
                          float on_p[3], float on_q[3]) {
    V3 u, w;
    closest_segments(v3_in(p0), v3_in(p1), v3_in(q0), v3_in(q1), u, w);
    v3_out(u, on_p);
    v3_out(w, on_q);
}

void pmo_closest_on_triangle(const float p[3], const float tri[9], float out[3]) {
    v3_out(closest_on_triangle(v3_in(p), v3_in(tri), v3_in(tri + 3), v3_in(tri + 6)), out);
}

void pmo_closest_triangle_pair(const float a[9], const float b[9], float on_a[3], float on_b[3]) {
    V3 A[3] = {v3_in(a), v3_in(a + 3), v3_in(a + 6)}, B[3] = {v3_in(b), v3_in(b + 3), v3_in(b + 6)};
    V3 u, w;
    closest_triangle_pair(A, B, u, w);
    v3_out(u, on_a);
    v3_out(w, on_b);
}

void pmo_triangle_feature_normal(const float a[9], const float b[9], const float on_a[3], const float on_b[3],
                                 const float fallback[3], float out[3]) {
    V3 A[3] = {v3_in(a), v3_in(a + 3), v3_in(a + 6)}, B[3] = {v3_in(b), v3_in(b + 3), v3_in(b + 6)};
    v3_out(triangle_feature_normal(A, B, v3_in(on_a), v3_in(on_b), v3_in(fallback)), out);
}

}  // extern "C"
