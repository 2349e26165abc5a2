// SPDX-License-Identifier: MIT
// Predictor, proxies, broadphase, narrowphase and the contact cache on
// gfx950, and the world step that chains them with the AVBD solve.
//
// Broadphase: the reference filters every proxy overlap (inclusive float
// AABB test) with the rules of pm_filter_rigid_spatial_candidates and
// orders pairs by (body, identity) (geometry.slang:829-947).  Here the
// proxies of bodies no larger than the largest dynamic proxy go into a
// uniform hash grid (cell = 1.0625 x that extent, keyed by the cell of the
// proxy's minimum corner, so any overlapping pair is found among the 27
// neighbouring cells); larger static proxies (floors, walls) form a short
// list every query scans.  Each dynamic body sorts its kept candidates, and
// an exclusive scan over the per-body counts emits the pairs in ascending
// code a*n+b.  The overlap test is the same comparison on the same floats,
// so the pair set and order equal the reference's bit for bit.
//
// Narrowphase: one thread per pair; convex meshes with <= 96 indices take
// the reference's SAT + face clipping path (rigid_convex_face_manifold and
// its swept variant, geometry.slang:1606-1853); analytic spheres/capsules
// use closed forms.  Mesh pairs the face path declines take the
// triangle-pair path (leaf-pair candidates or BVH traversal) in a second
// kernel over a compact list of those pairs.
#include "pmh_internal.hpp"
#include "pmh_world.hpp"

#include <algorithm>

namespace pmh {

__device__ __forceinline__ V3 sv(const PmhVec3 &v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ Q sq(const PmhQuat &q) { return {q.x, q.y, q.z, q.w}; }
__device__ __forceinline__ PmhVec3 pv(V3 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ PmhQuat pq(Q q) { return {q.x, q.y, q.z, q.w}; }
__device__ __forceinline__ V3 v3(float4 f) { return {f.x, f.y, f.z}; }
__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ V3 sarr(const float *a) { return {a[0], a[1], a[2]}; }
__device__ __forceinline__ V3 transform_point(const PmhRigidBodyState &s, V3 p) {
    return add(sv(s.position), rotate(sq(s.orientation), p));
}

// ---------------------------------------------------------------------
// Predictor + proxies (geometry.slang:524-612, 716-827).
// ---------------------------------------------------------------------
__device__ __forceinline__ float correction_reach(const PmhBodyParameters &p, const ShapeDev &s) {
    if (p.motion != kDynamic) return 0.0f;
    const V3 e = sub(sarr(s.mx), sarr(s.mn));
    return smin(0.5f, 0.25f * smax(0.0f, smin(e.x, smin(e.y, e.z))));
}

__device__ __forceinline__ V3 inverse_inertia_world(const PmhBodyParameters &p, Q o, V3 w) {
    const V3 l = inv_rotate(o, w);
    return rotate(o, {l.x * p.inverse_inertia_local.x, l.y * p.inverse_inertia_local.y,
                      l.z * p.inverse_inertia_local.z});
}

__global__ void k_integrate_proxy(uint32_t n, const PmhBodyParameters *P, const ShapeDev *shapes,
                                  const PmhRigidBodyState *cur, PmhRigidBodyState *prev,
                                  PmhRigidBodyState *pred, float4 *pmn, float4 *pmx, V3 g, float dt,
                                  float4 *inputs, const PmhRigidBodyState *targets, uint32_t *target_flags,
                                  uint32_t remaining, uint32_t first_substep, uint32_t last_substep,
                                  StepCounters *ctr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PmhBodyParameters b = P[i];
    const PmhRigidBodyState s0 = cur[i];
    prev[i] = s0;
    PmhRigidBodyState s = s0;
    const uint32_t target_active = target_flags[i];
    if (b.motion == kStatic) {
        s.linear_velocity = {0.0f, 0.0f, 0.0f};
        s.angular_velocity = {0.0f, 0.0f, 0.0f};
    } else if (b.motion == kKinematic) {
        if (target_active) {
            // The target moves 1/remaining of the way each substep and is
            // reached in the frame's last.
            const PmhRigidBodyState t = targets[i];
            const float fraction = 1.0f / float(remaining);
            const V3 p0 = sv(s0.position);
            const V3 p1 = add(p0, mul(sub(sv(t.position), p0), fraction));
            const Q o = sq(s0.orientation);
            Q q = sq(t.orientation);
            if (o.x * q.x + o.y * q.y + o.z * q.z + o.w * q.w < 0.0f) q = {-q.x, -q.y, -q.z, -q.w};
            const Q no = qnormalize({o.x + (q.x - o.x) * fraction, o.y + (q.y - o.y) * fraction,
                                     o.z + (q.z - o.z) * fraction, o.w + (q.w - o.w) * fraction});
            s.position = pv(p1);
            s.orientation = pq(no);
            s.linear_velocity = pv(mul(sub(p1, p0), 1.0f / dt));
            s.angular_velocity = pv(quaternion_delta_velocity(o, no, dt));
        } else {
            s.linear_velocity = {0.0f, 0.0f, 0.0f};
            s.angular_velocity = {0.0f, 0.0f, 0.0f};
        }
    } else {
        const V3 force = v3(inputs[4 * i]), torque = v3(inputs[4 * i + 1]);
        const V3 imp = v3(inputs[4 * i + 2]), aimp = v3(inputs[4 * i + 3]);
        V3 v = add(sv(s.linear_velocity), mul(add(g, mul(force, b.inverse_mass)), dt));
        V3 w = add(sv(s.angular_velocity), mul(inverse_inertia_world(b, sq(s.orientation), torque), dt));
        if (first_substep) {
            v = add(v, mul(imp, b.inverse_mass));
            w = add(w, inverse_inertia_world(b, sq(s.orientation), aimp));
        }
        v = mul(v, 1.0f / (1.0f + b.linear_damping * dt));
        w = mul(w, 1.0f / (1.0f + b.angular_damping * dt));
        v = clamp_length(v, b.maximum_linear_speed);
        w = clamp_length(w, b.maximum_angular_speed);
        const V3 p = add(sv(s.position), mul(v, dt));
        const Q o = sq(s.orientation);
        const Q d = qmul({w.x, w.y, w.z, 0.0f}, o);
        const Q no = qnormalize({o.x + 0.5f * d.x * dt, o.y + 0.5f * d.y * dt, o.z + 0.5f * d.z * dt,
                                 o.w + 0.5f * d.w * dt});
        s.position = pv(p);
        s.orientation = pq(no);
        s.linear_velocity = pv(v);
        s.angular_velocity = pv(w);
    }
    // pm_clear_rigid_inputs (geometry.slang:2877-2886) after the frame's
    // last substep: inputs and targets last one frame.
    if (last_substep) {
        const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        inputs[4 * i] = z;
        inputs[4 * i + 1] = z;
        inputs[4 * i + 2] = z;
        inputs[4 * i + 3] = z;
        if (target_active) target_flags[i] = 0u;
    }
    pred[i] = s;
    // Proxy: production arguments path_deviation 0, correction_budget 1,
    // epsilon 1e-5 (production_realtime_backend.cu.inc:2663-2684).
    const ShapeDev sh = shapes[b.mesh_index];
    const float eps = 1.0e-5f;
    const float reach = len(sarr(sh.center)) + sh.bradius + b.collision_margin + eps;
    const float cr = correction_reach(b, sh);
    const V3 bc = {smin(1.0f, cr), smin(1.0f, cr), smin(1.0f, cr)};
    const V3 e = add(add(V3{0.0f, 0.0f, 0.0f}, bc), V3{reach, reach, reach});
    const V3 mn = sub(cmin(sv(s0.position), sv(s.position)), e);
    const V3 mx = add(cmax(sv(s0.position), sv(s.position)), e);
    pmn[i] = f4(mn, 0.0f);
    pmx[i] = f4(mx, 0.0f);
    (void)ctr;
}

// ---------------------------------------------------------------------
// Broadphase.
// ---------------------------------------------------------------------
__device__ __forceinline__ uint32_t cell_hash(int x, int y, int z) {
    uint32_t h = uint32_t(x) * 0x8da6b343u ^ uint32_t(y) * 0xd8163841u ^ uint32_t(z) * 0xcb1ab31fu;
    h ^= h >> 15;
    h *= 0x2c1b3c6du;
    h ^= h >> 12;
    return h;
}

constexpr int kLargeCell = 0x7fffffff;

// Proxies no wider than `cell` go into the hash grid under the cell of
// their minimum corner; wider ones (floors, walls, fast movers) into the
// large list.  Any overlapping gridded pair then lies in neighbouring cells.
__global__ void k_bp_cells(uint32_t n, const float4 *pmn, const float4 *pmx, int4 *cell_of, uint32_t *bucket_count,
                           uint32_t table_mask, uint32_t *large, StepCounters *ctr, float cell) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 mn = pmn[i], mx = pmx[i];
    const float ext = smax(mx.x - mn.x, smax(mx.y - mn.y, mx.z - mn.z));
    if (!(ext <= cell) || !isfinite(mn.x) || !isfinite(mn.y) || !isfinite(mn.z)) {
        large[atomicAdd(&ctr->large_count, 1u)] = i;
        cell_of[i] = make_int4(kLargeCell, kLargeCell, kLargeCell, 0);
        return;
    }
    const int cx = int(floorf(mn.x / cell)), cy = int(floorf(mn.y / cell)), cz = int(floorf(mn.z / cell));
    const uint32_t b = cell_hash(cx, cy, cz) & table_mask;
    cell_of[i] = make_int4(cx, cy, cz, int(b));
    atomicAdd(&bucket_count[b], 1u);
}

// Places each body in its bucket and copies, in entry order, what a query
// tests of a candidate (cell, dynamic flag, box), so the query reads those
// records by entry index instead of gathering them by body index.
__global__ void k_bp_scatter(uint32_t n, const PmhBodyParameters *P, const float4 *pmn, const float4 *pmx,
                             const int4 *cell_of, const uint32_t *bucket_start, uint32_t *cursor, int4 *entry_cell,
                             float4 *entry_mn, float4 *entry_mx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 c = cell_of[i];
    if (c.x == kLargeCell) return;
    const uint32_t e = bucket_start[c.w] + atomicAdd(&cursor[c.w], 1u);
    const float4 mn = pmn[i];
    entry_cell[e] = make_int4(c.x, c.y, c.z, P[i].motion == kDynamic ? 1 : 0);
    entry_mn[e] = make_float4(mn.x, mn.y, mn.z, __uint_as_float(i));
    entry_mx[e] = pmx[i];
}

__device__ __forceinline__ bool overlap(float4 amn, float4 amx, float4 bmn, float4 bmx) {
    return amn.x <= bmx.x && amx.x >= bmn.x && amn.y <= bmx.y && amx.y >= bmn.y && amn.z <= bmx.z &&
           amx.z >= bmn.z;
}

// Collision exclusions of joints (geometry.slang:494-511, 3108-3136): bodies
// of one weld group never collide, and a pair joined by a joint with
// disable_collisions does not while that joint is alive, enabled and
// unbroken.  A null `weld` or `excl_start` turns that test off.
// A batch of independent scenes adds the scene filter: bodies of
// different scenes never pair (null `scene`: one scene).
struct JointFilter {
    const uint32_t *weld = nullptr;
    const uint32_t *excl_start = nullptr;
    const uint2 *excl_list = nullptr;   // (partner, joint)
    const PmhRigidConstraint *joints = nullptr;
    const uint32_t *scene = nullptr;    // scene of each body
    // Batch: each scene's large bodies (a segment of scene_large per scene)
    // and each scene's body range.
    const uint32_t *scene_large_start = nullptr, *scene_large = nullptr;
    const uint32_t *scene_first = nullptr, *scene_count = nullptr;
};

__device__ __forceinline__ bool joint_excluded(const JointFilter &f, uint32_t a, uint32_t b) {
    if (f.scene && f.scene[a] != f.scene[b]) return true;
    if (f.weld && f.weld[a] == f.weld[b]) return true;
    if (f.excl_start) {
        for (uint32_t e = f.excl_start[a]; e < f.excl_start[a + 1]; ++e) {
            const uint2 x = f.excl_list[e];
            if (x.x == b && f.joints[x.y].state_broken == 0u) return true;
        }
    }
    return false;
}

__device__ uint32_t weld_find(const uint32_t *weld, uint32_t v) {
    uint32_t p = __hip_atomic_load(&weld[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (p != v) {
        v = p;
        p = __hip_atomic_load(&weld[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return v;
}

__global__ void k_weld_init(uint32_t n, uint32_t *weld) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) weld[i] = i;
}

// Fixed joints that disable collisions join their bodies' weld group; the
// larger root hooks under the smaller, so each group's root is its
// smallest index whatever the order of the hooks.
__global__ void k_weld_hook(uint32_t n, uint32_t nj, const PmhRigidConstraint *joints, uint32_t *weld) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nj) return;
    const PmhRigidConstraint &j = joints[k];
    const PmhRigidConstraintOptions &o = j.options;
    if (j.alive == 0u || o.enabled == 0u || j.state_broken != 0u || o.disable_collisions == 0u ||
        o.type != kJointFixed)
        return;
    if (o.body_a.index >= n || o.body_a.generation != 1u || o.body_b.index >= n || o.body_b.generation != 1u)
        return;
    uint32_t u = weld_find(weld, o.body_a.index), v = weld_find(weld, o.body_b.index);
    while (u != v) {
        if (u > v) { const uint32_t t = u; u = v; v = t; }
        const uint32_t old = atomicCAS(&weld[v], v, u);
        if (old == v) break;
        u = weld_find(weld, u);
        v = weld_find(weld, old);
    }
}

__global__ void k_weld_compress(uint32_t n, uint32_t *weld) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) weld[i] = weld_find(weld, i);
}

// A batch's large bodies by scene: counts, then (after a scan over the
// scenes) placement.  The order within a scene's segment is free: every
// query sorts its candidates.
__global__ void k_scene_large(const StepCounters *ctr, const uint32_t *large, const uint32_t *scene,
                              uint32_t *count, const uint32_t *start, uint32_t *out) {
    const uint32_t nl = ctr->large_count;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nl; k += gridDim.x * blockDim.x) {
        const uint32_t b = large[k];
        const uint32_t slot = atomicAdd(&count[scene[b]], 1u);
        if (start) out[start[scene[b]] + slot] = b;
    }
}

// Kept candidates of dynamic body a, found by a group of 32 lanes: lane
// q < 27 walks neighbour cell q of the 3x3x3 block around the body's cell,
// and every lane takes every 32nd entry of the large-body list.  Pass 0
// counts them; pass 1 writes each lane's candidates at the body's offset
// (from an exclusive scan of the counts) plus the lane's prefix within the
// group, then one lane sorts the segment ascending (identity order =
// index), so a body may own any number of pairs.
constexpr uint32_t kQueryLanes = 32;

template <int kPass>
__global__ void __launch_bounds__(256) k_bp_query(uint32_t n, const PmhBodyParameters *P, const float4 *pmn,
                                                  const float4 *pmx, const int4 *cell_of, const uint32_t *bucket_start,
                                                  const int4 *entry_cell, const float4 *entry_mn,
                                                  const float4 *entry_mx, uint32_t table_mask, const uint32_t *large,
                                                  uint32_t *cand_count, const uint32_t *offset, uint32_t *code,
                                                  uint2 *ab, const StepCounters *ctr, JointFilter jf) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t a = uint32_t(t / kQueryLanes), q = uint32_t(t % kQueryLanes);
    // Every exit above the shuffles is uniform over the body's group.
    if (a >= n) return;
    if (P[a].motion != kDynamic) {
        if (kPass == 0 && q == 0) cand_count[a] = 0;
        return;
    }
    if (kPass == 1 && ctr->pairs == 0) return;   // pair pool overflow: nothing is emitted
    const int4 c = cell_of[a];
    if (c.x == kLargeCell) return;   // k_bp_large owns this body
    const float4 amn = pmn[a], amx = pmx[a];
    uint32_t count = 0;
    auto walk = [&](auto &&emit) {
        if (q < 27u) {
            const int x = c.x + int(q % 3u) - 1, y = c.y + int((q / 3u) % 3u) - 1, z = c.z + int(q / 9u) - 1;
            const uint32_t bk = cell_hash(x, y, z) & table_mask;
            for (uint32_t e = bucket_start[bk]; e < bucket_start[bk + 1]; ++e) {
                const int4 cb = entry_cell[e];
                if (cb.x != x || cb.y != y || cb.z != z) continue;
                const float4 bmn = entry_mn[e];
                const uint32_t b = __float_as_uint(bmn.w);
                if (b == a || (cb.w != 0 && !(a < b))) continue;
                if (!overlap(amn, amx, bmn, entry_mx[e])) continue;
                if (joint_excluded(jf, a, b)) continue;
                emit(b);
            }
        }
        auto consider = [&](uint32_t b) {
            if (b == a) return;
            if (P[b].motion == kDynamic && !(a < b)) return;
            if (!overlap(amn, amx, pmn[b], pmx[b])) return;
            if (joint_excluded(jf, a, b)) return;
            emit(b);
        };
        if (jf.scene_large) {
            const uint32_t sc = jf.scene[a];
            for (uint32_t k = jf.scene_large_start[sc] + q; k < jf.scene_large_start[sc + 1]; k += kQueryLanes)
                consider(jf.scene_large[k]);
        } else {
            const uint32_t nl = ctr->large_count;
            for (uint32_t k = q; k < nl; k += kQueryLanes) consider(large[k]);
        }
    };
    walk([&](uint32_t) { ++count; });   // pass 1 counts again: its lanes need their own prefixes
    // Inclusive scan of the lane counts within the group.
    uint32_t incl = count;
    for (uint32_t d = 1; d < kQueryLanes; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, kQueryLanes);
        if (q >= d) incl += y;
    }
    const uint32_t total = __shfl(incl, kQueryLanes - 1, kQueryLanes);
    if (kPass == 0) {
        if (q == 0) cand_count[a] = total;
        return;
    }
    const uint32_t base = offset[a];
    uint32_t *out = code + base;
    uint32_t at = incl - count;
    walk([&](uint32_t b) { out[at++] = b; });
    // The segment's writes of every lane are visible to the whole group.
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
    // Rank sort of the segment over the group's lanes: element k goes to
    // the count of elements below it, ties in segment order (the order a
    // stable insertion sort leaves).  The pairs are written first, then the
    // codes from them, since the codes' slots hold the unsorted segment.
    for (uint32_t k = q; k < total; k += kQueryLanes) {
        const uint32_t v = out[k];
        uint32_t rank = 0;
        for (uint32_t i = 0; i < total; ++i) {
            const uint32_t w = out[i];
            rank += (w < v || (w == v && i < k)) ? 1u : 0u;
        }
        ab[base + rank] = make_uint2(a, v);
    }
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
    for (uint32_t k = q; k < total; k += kQueryLanes) out[k] = a * n + ab[base + k].y;
}

__global__ void k_set_pairs(StepCounters *ctr, const uint32_t *offset, uint32_t n, uint32_t pair_cap) {
    if (threadIdx.x != 0) return;
    const uint32_t M = offset[n];
    if (M > pair_cap) {
        atomicOr(&ctr->status, kStatusPairOverflow);
        ctr->pairs = 0;
    } else {
        ctr->pairs = M;
    }
}

// Large bodies: one block scans every body in index order; a block-wide
// prefix sum places the hits, so the segment comes out ascending.  Blocks
// beyond the number of large bodies return at once.
constexpr uint32_t kLargeBlocks = 1024;
template <int kPass>
__global__ void __launch_bounds__(256) k_bp_large(uint32_t n, const PmhBodyParameters *P, const float4 *pmn,
                                                  const float4 *pmx, const uint32_t *large, uint32_t *cand_count,
                                                  const uint32_t *offset, uint32_t *code, uint2 *ab,
                                                  const StepCounters *ctr, JointFilter jf) {
    __shared__ uint32_t warp_tot[4];
    const uint32_t nl = ctr->large_count;
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (uint32_t k = blockIdx.x; k < nl; k += gridDim.x) {
        const uint32_t a = large[k];
        if (P[a].motion != kDynamic) {
            if (kPass == 0 && threadIdx.x == 0) cand_count[a] = 0;
            continue;
        }
        if (kPass == 1 && ctr->pairs == 0) continue;
        const float4 amn = pmn[a], amx = pmx[a];
        uint32_t running = kPass == 1 ? offset[a] : 0u;
        // In a batch only the body's own scene can pair with it.
        const uint32_t first = jf.scene_first ? jf.scene_first[jf.scene[a]] : 0u;
        const uint32_t last = jf.scene_first ? first + jf.scene_count[jf.scene[a]] : n;
        for (uint32_t base = first; base < last; base += 256) {
            const uint32_t b = base + threadIdx.x;
            bool hit = false;
            if (b < last && b != a && !(P[b].motion == kDynamic && !(a < b)))
                hit = overlap(amn, amx, pmn[b], pmx[b]) && !joint_excluded(jf, a, b);
            const uint64_t bal = __ballot(hit);
            if (lane == 0) warp_tot[wid] = __popcll(bal);
            __syncthreads();
            uint32_t before = 0, total = 0;
            for (uint32_t w = 0; w < 4; ++w) {
                if (w < wid) before += warp_tot[w];
                total += warp_tot[w];
            }
            if (kPass == 1 && hit) {
                const uint32_t pos = running + before + __popcll(bal & ((1ull << lane) - 1ull));
                code[pos] = a * n + b;
                ab[pos] = make_uint2(a, b);
            }
            running += total;
            __syncthreads();
        }
        if (kPass == 0 && threadIdx.x == 0) cand_count[a] = running;
    }
}

// ---------------------------------------------------------------------
// Narrowphase.
// ---------------------------------------------------------------------
struct Contact {
    V3 normal, point;
    float pen, impact_fraction;
};
struct Manifold {
    Contact c[8];
    uint32_t count;
    uint32_t face_patch;
};

struct ShapeView {
    ShapeDev s;
    const V3 *v;
    const uint32_t *idx;
    const float4 *planes;
    const BvhDev *nodes = nullptr;
    const uint32_t *leaves = nullptr;
    const float4 *shell = nullptr;
    __device__ float4 plane(uint32_t f) const { return planes[s.plane_off + f]; }
    __device__ V3 vert(uint32_t k) const { return v[s.vtx_off + k]; }
    __device__ uint32_t index(uint32_t k) const { return idx[s.idx_off + k]; }
    __device__ const BvhDev &node(uint32_t k) const { return nodes[s.node_off + k]; }
    __device__ uint32_t leaf(uint32_t k) const { return leaves[s.leaf_off + k]; }
};

// rigid_add_contact, geometry.slang:1129-1220 (unguided, no swept-witness
// canonicalisation on the convex face path).
__device__ void add_contact(Manifold &m, const Contact &c, float spacing) {
    const float sp2 = spacing * spacing;
    for (uint32_t i = 0; i < m.count; ++i) {
        if (len2(sub(m.c[i].point, c.point)) < sp2) {
            if (c.pen - m.c[i].pen > 0.0f) m.c[i] = c;
            return;
        }
    }
    if (m.count < 8u) { m.c[m.count++] = c; return; }
    uint32_t sh = 0;
    for (uint32_t i = 1; i < m.count; ++i) if (m.c[i].pen < m.c[sh].pen) sh = i;
    if (c.pen > m.c[sh].pen) m.c[sh] = c;
}

// ---- hinge contact frames (joints that shape contacts) ------------------
using HingeFrame = PmhHingeContactFrame;

__device__ __forceinline__ bool guided_pair(const HingeFrame &a, const HingeFrame &b) {   // geometry.slang:373-377
    return (a.axial != 0u && b.static_body != 0u) || (b.axial != 0u && a.static_body != 0u);
}

// rigid_guide_direction, geometry.slang:1118-1127 (`g` is the pair's axial frame).
__device__ __forceinline__ V3 guide_direction(const Contact &c, const HingeFrame &g) {
    return {dot(c.normal, sv(g.axis)) * sqrtf(g.inverse_mass),
            g.axial_rotation != 0u ? dot(sv(g.axis), cross(sub(c.point, sv(g.anchor)), c.normal)) * sqrtf(g.inverse_moment)
                                   : 0.0f,
            0.0f};
}

// rigid_add_contact's guided branch, geometry.slang:1129-1168: the first
// impact keeps its rows (later impacts are dropped, earlier ones replace
// the swept rows), and a contact whose guide direction matches a kept one
// replaces it when it is deeper per unit of direction.  All contacts of a
// pair share its frames, so `g` serves for both.
__device__ void add_contact_guided(Manifold &m, const Contact &c, const HingeFrame &g) {
    const V3 cand_dir = guide_direction(c, g);
    const float cand_len = len(cand_dir);
    if (cand_len * cand_len <= kEpsilon) return;
    if (c.impact_fraction > 0.0f) {
        float first_impact = 1.0f;
        for (uint32_t r = 0; r < m.count; ++r)
            if (m.c[r].impact_fraction > 0.0f) first_impact = smin(first_impact, m.c[r].impact_fraction);
        if (c.impact_fraction > first_impact + 1.0e-4f) return;
        if (c.impact_fraction < first_impact - 1.0e-4f) {
            uint32_t kept = 0;
            for (uint32_t r = 0; r < m.count; ++r)
                if (m.c[r].impact_fraction == 0.0f) m.c[kept++] = m.c[r];
            m.count = kept;
        }
    }
    for (uint32_t i = 0; i < m.count; ++i) {
        const V3 dir = guide_direction(m.c[i], g);
        const float size = len(dir);
        if (dot(cand_dir, dir) > 0.9999f * cand_len * size) {
            if (c.pen / cand_len > m.c[i].pen / size) m.c[i] = c;
            return;
        }
    }
    if (m.count < 8u) { m.c[m.count++] = c; return; }
    uint32_t sh = 0;
    for (uint32_t i = 1; i < m.count; ++i) if (m.c[i].pen < m.c[sh].pen) sh = i;
    if (c.pen > m.c[sh].pen) m.c[sh] = c;
}

// rigid_fixed_hinge_inverse_moment, geometry.slang:384-402.
__device__ float fixed_hinge_inverse_moment(const PmhBodyParameters &p, const PmhRigidBodyState &s,
                                            const HingeFrame &f) {
    if ((f.fixed == 0u && f.axial_rotation == 0u) || p.inverse_mass <= kEpsilon) return 0.0f;
    const V3 la = inv_rotate(sq(s.orientation), sv(f.axis));
    const float center = la.x * la.x / smax(p.inverse_inertia_local.x, kEpsilon) +
                         la.y * la.y / smax(p.inverse_inertia_local.y, kEpsilon) +
                         la.z * la.z / smax(p.inverse_inertia_local.z, kEpsilon);
    const V3 arm = sub(sv(s.position), sv(f.anchor));
    const V3 perpendicular = sub(arm, mul(sv(f.axis), dot(arm, sv(f.axis))));
    const float pivot = center + len2(perpendicular) / p.inverse_mass;
    return pivot > kEpsilon ? 1.0f / pivot : 0.0f;
}

// The world's joints as the narrowphase sees them: the records and each
// body's candidate joints (alive, both ids valid) in descending index.
struct JointView {
    const PmhRigidConstraint *joints = nullptr;
    const uint32_t *list_start = nullptr, *list = nullptr;
};

// rigid_hinge_contact_frame, geometry.slang:404-491: over the body's
// enabled, unbroken joints in ascending index, the first hinge (or slider
// or piston held by a static body with no breaking threshold) gives the
// frame, anchored on the static body when that holds it; a slider or
// piston counts only as the body's single joint; fixed and point members
// are flagged.  `reference` is the anchor of a non-axial frame, else the
// body's position.  Body ids are {dense index, 1}.
__device__ HingeFrame hinge_frame(const JointView &jv, const PmhBodyParameters *P, uint32_t n,
                                  const PmhRigidBodyState *states, uint32_t dense, const PmhRigidBodyState &state,
                                  V3 &reference) {
    HingeFrame r;
    memset(&r, 0, sizeof r);
    const PmhBodyParameters pd = P[dense];
    r.static_body = pd.motion == kStatic ? 1u : 0u;
    if (pd.inverse_mass <= 0.0f || !jv.joints) {
        reference = sv(state.position);
        return r;
    }
    uint32_t incident = 0;
    for (uint32_t e = jv.list_start[dense + 1]; e > jv.list_start[dense]; --e) {
        const PmhRigidConstraint &j = jv.joints[jv.list[e - 1]];
        const PmhRigidConstraintOptions &o = j.options;
        if (j.alive == 0u || o.enabled == 0u || j.state_broken != 0u) continue;
        const bool is_a = o.body_a.index == dense && o.body_a.generation == 1u;
        const bool is_b = o.body_b.index == dense && o.body_b.generation == 1u;
        if (!is_a && !is_b) continue;
        ++incident;
        if (o.type == kJointFixed) r.fixed_member = 1u;
        if (o.type == kJointPoint) r.point_member = 1u;
        const bool axial = o.type == kJointPiston || o.type == kJointSlider;
        const PmhBodyId oid = is_a ? o.body_b : o.body_a;
        const uint32_t other = oid.index < n && oid.generation == 1u ? oid.index : kInvalid;
        const bool static_anchor = other != kInvalid && P[other].motion == kStatic;
        if (r.present != 0u ||
            (o.type != kJointHinge && !(axial && static_anchor && o.breaking_impulse_threshold <= 0.0f)))
            continue;
        const V3 local_axis = axial ? V3{1.0f, 0.0f, 0.0f} : V3{0.0f, 0.0f, 1.0f};
        r.local_anchor = is_a ? o.local_anchor_a : o.local_anchor_b;
        const Q lo = sq(is_a ? o.local_orientation_a : o.local_orientation_b);
        const Q so = sq(state.orientation);
        V3 anchor = add(sv(state.position), rotate(so, sv(r.local_anchor)));
        V3 axis = normalized_or(rotate(qmul(so, lo), local_axis), local_axis);
        r.present = 1u;
        r.fixed = static_anchor && !axial ? 1u : 0u;
        r.axial = static_anchor && axial ? 1u : 0u;
        r.axial_rotation = r.axial != 0u && o.type == kJointPiston ? 1u : 0u;
        if (static_anchor) {
            const PmhRigidBodyState os = states[other];
            const Q oo = sq(os.orientation);
            anchor = add(sv(os.position), rotate(oo, sv(is_a ? o.local_anchor_b : o.local_anchor_a)));
            axis = normalized_or(rotate(qmul(oo, sq(is_a ? o.local_orientation_b : o.local_orientation_a)), local_axis),
                                 axis);
        }
        r.anchor = pv(anchor);
        r.axis = pv(axis);
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
    reference = r.present != 0u && r.axial == 0u ? sv(r.anchor) : sv(state.position);
    return r;
}

// Per body: its frame at the predicted state and its reference point at the
// previous and the predicted state (.xyz of ref[2i], ref[2i + 1]).  Worlds
// without joints skip this: every frame then carries only the static flag.
__global__ void k_hinge_frames(uint32_t n, const PmhBodyParameters *P, const PmhRigidBodyState *prev,
                               const PmhRigidBodyState *pred, JointView jv, HingeFrame *frames, float4 *ref) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V3 rp, rc;
    (void)hinge_frame(jv, P, n, prev, i, prev[i], rp);
    frames[i] = hinge_frame(jv, P, n, pred, i, pred[i], rc);
    ref[2 * i] = f4(rp, 0.0f);
    ref[2 * i + 1] = f4(rc, 0.0f);
}

// The frames of one pair, and the body's reference points.
struct PairFrames {
    HingeFrame ha, hb;
    V3 ref_prev, ref;
};
__device__ __forceinline__ PairFrames pair_frames(const HingeFrame *frames, const float4 *ref, const PmhBodyParameters &pa,
                                                  const PmhBodyParameters &pb, uint2 p, const PmhRigidBodyState &A0,
                                                  const PmhRigidBodyState &A) {
    PairFrames f;
    if (frames) {
        f.ha = frames[p.x];
        f.hb = frames[p.y];
        f.ref_prev = v3(ref[2 * p.x]);
        f.ref = v3(ref[2 * p.x + 1]);
    } else {
        memset(&f.ha, 0, sizeof f.ha);
        memset(&f.hb, 0, sizeof f.hb);
        f.ha.static_body = pa.motion == kStatic ? 1u : 0u;
        f.hb.static_body = pb.motion == kStatic ? 1u : 0u;
        f.ref_prev = sv(A0.position);
        f.ref = sv(A.position);
    }
    return f;
}

// rigid_reduce_face_patch, geometry.slang:1228-1270.
__device__ void reduce_face_patch(Manifold &m, float tol) {
    const float t2 = tol * tol;
    bool removed = true;
    while (removed) {
        removed = false;
        for (uint32_t cand = 0; cand < m.count && !removed; ++cand) {
            const V3 p = m.c[cand].point;
            for (uint32_t f = 0; f < m.count && !removed; ++f) {
                if (f == cand) continue;
                for (uint32_t s = f + 1; s < m.count; ++s) {
                    if (s == cand) continue;
                    const V3 e = sub(m.c[s].point, m.c[f].point);
                    const float e2 = len2(e);
                    if (e2 <= t2) continue;
                    const float t = dot(sub(p, m.c[f].point), e) / e2;
                    if (t <= 1.0e-4f || t >= 0.9999f) continue;
                    const V3 cl = add(m.c[f].point, mul(e, t));
                    if (len2(sub(p, cl)) > t2) continue;
                    for (uint32_t k = cand + 1; k < m.count; ++k) m.c[k - 1] = m.c[k];
                    --m.count;
                    removed = true;
                    break;
                }
            }
        }
    }
}

__device__ float rotational_motion_bound(const PmhRigidBodyState &p, const PmhRigidBodyState &c,
                                         const ShapeDev &s) {
    const Q d = relative_quaternion(sq(p.orientation), sq(c.orientation), false);
    const float sh = len({d.x, d.y, d.z});
    const V3 ma = {smax(fabsf(s.mn[0]), fabsf(s.mx[0])), smax(fabsf(s.mn[1]), fabsf(s.mx[1])),
                   smax(fabsf(s.mn[2]), fabsf(s.mx[2]))};
    return 2.0f * len(ma) * sh;
}

// rigid_convex_face_manifold, geometry.slang:1606-1745.
__device__ bool convex_face_manifold(const PmhRigidBodyState &bs, const ShapeView &bm,
                                     const PmhRigidBodyState &cs, const ShapeView &cm, float margin,
                                     float reach_corr, Manifold &out) {
    out.count = 0;
    out.face_patch = 0;
    const float rest = smin(margin, 1.0e-3f);
    const float reach = margin + reach_corr;
    float best = -3.402823466e+38f;
    uint32_t ref_face = 0;
    bool ref_is_body = false;
    for (uint32_t side = 0; side < 2; ++side) {
        const ShapeView &rm = side == 0 ? cm : bm;
        const PmhRigidBodyState &rs = side == 0 ? cs : bs;
        const ShapeView &im = side == 0 ? bm : cm;
        const PmhRigidBodyState &is = side == 0 ? bs : cs;
        const Q rq = sq(rs.orientation), iq = sq(is.orientation);
        const V3 rel = sub(sv(is.position), sv(rs.position));
        for (uint32_t f = 0; f < rm.s.idx_count / 3; ++f) {
            const float4 pl = rm.plane(f);
            const V3 nrm = rotate(rq, v3(pl));
            const V3 inn = inv_rotate(iq, nrm);
            float support = 3.402823466e+38f;
            for (uint32_t k = 0; k < im.s.vtx_count; ++k) support = smin(support, dot(inn, im.vert(k)));
            const float sep = support + dot(nrm, rel) - pl.w;
            if (sep > best) { best = sep; ref_face = f; ref_is_body = side != 0; }
        }
    }
    if (best > reach) return true;
    const ShapeView &rm = ref_is_body ? bm : cm;
    const PmhRigidBodyState &rs = ref_is_body ? bs : cs;
    const ShapeView &im = ref_is_body ? cm : bm;
    const PmhRigidBodyState &is = ref_is_body ? cs : bs;
    const Q rq = sq(rs.orientation);
    const float4 reference = rm.plane(ref_face);
    const V3 rn = v3(reference);
    const V3 outward = rotate(rq, rn);
    const V3 iaxis = inv_rotate(sq(is.orientation), outward);
    float alignment = 1.0f;
    uint32_t inc_face = 0;
    for (uint32_t f = 0; f < im.s.idx_count / 3; ++f) {
        const float val = dot(iaxis, v3(im.plane(f)));
        if (val < alignment) { alignment = val; inc_face = f; }
    }
    const V3 normal = mul(outward, ref_is_body ? -1.0f : 1.0f);
    const float4 incident = im.plane(inc_face);
    const uint32_t rfaces = rm.s.idx_count / 3;
    for (uint32_t t = 0; t < im.s.idx_count / 3; ++t) {
        const float4 face = im.plane(t);
        if (dot(v3(face), v3(incident)) < 0.99999f || fabsf(face.w - incident.w) > 1.0e-5f) continue;
        // The polygon is clipped back and forth between two buffers.
        V3 buf[2][36];
        V3 *poly = buf[0], *clipped = buf[1];
        uint32_t count = 3;
        for (uint32_t c = 0; c < 3; ++c)
            poly[c] = inv_rotate(rq, sub(transform_point(is, im.vert(im.index(t * 3 + c))), sv(rs.position)));
        for (uint32_t pi = 0; pi < rfaces && count != 0; ++pi) {
            const float4 pl = rm.plane(pi);
            const V3 pn = v3(pl);
            const float off = pl.w + (dot(pn, rn) > 0.99999f ? reach : 0.0f);
            uint32_t cc = 0;
            V3 prev = poly[count - 1];
            float pd = dot(pn, prev) - off;
            for (uint32_t k = 0; k < count; ++k) {
                const V3 cur = poly[k];
                const float d = dot(pn, cur) - off;
                if ((d <= 0.0f) != (pd <= 0.0f) && cc < 36)
                    clipped[cc++] = add(prev, mul(sub(cur, prev), pd / (pd - d)));
                if (d <= 0.0f && cc < 36) clipped[cc++] = cur;
                prev = cur;
                pd = d;
            }
            count = cc;
            V3 *const t_ = poly;
            poly = clipped;
            clipped = t_;
        }
        for (uint32_t k = 0; k < count; ++k) {
            const float d = dot(rn, poly[k]) - reference.w;
            if (d > reach) continue;
            Contact c;
            c.normal = normal;
            c.point = transform_point(rs, sub(poly[k], mul(rn, d * 0.5f)));
            c.pen = rest - d + 1.0e-5f;
            c.impact_fraction = 1.0f;
            add_contact(out, c, smax(margin * 2.0f, 1.0e-4f));
        }
    }
    if (out.count == 0u) return false;
    reduce_face_patch(out, smax(1.0e-6f, margin * 1.0e-3f));
    out.face_patch = 1u;
    return true;
}

// rigid_swept_convex_face_manifold, geometry.slang:1751-1853.
__device__ bool swept_convex_face_manifold(const PmhRigidBodyState &pb, const PmhRigidBodyState &b,
                                           const ShapeView &bm, const PmhRigidBodyState &pc,
                                           const PmhRigidBodyState &c, const ShapeView &cm, float margin,
                                           Manifold &out) {
    out.count = 0;
    out.face_patch = 0;
    const Q bd = relative_quaternion(sq(pb.orientation), sq(b.orientation), false);
    const Q cd = relative_quaternion(sq(pc.orientation), sq(c.orientation), false);
    if (bd.x != 0.0f || bd.y != 0.0f || bd.z != 0.0f || cd.x != 0.0f || cd.y != 0.0f || cd.z != 0.0f)
        return false;
    const V3 bmo = sub(sv(b.position), sv(pb.position)), cmo = sub(sv(c.position), sv(pc.position));
    const float rest = smin(margin, 1.0e-3f);
    float entry = 0.0f, exit_time = 1.0f;
    V3 entry_normal = {0.0f, 0.0f, 0.0f};
    for (uint32_t side = 0; side < 2; ++side) {
        const ShapeView &rm = side == 0 ? cm : bm;
        const ShapeView &im = side == 0 ? bm : cm;
        const PmhRigidBodyState &rs = side == 0 ? pc : pb;
        const PmhRigidBodyState &is = side == 0 ? pb : pc;
        const V3 rel = side == 0 ? sub(bmo, cmo) : sub(cmo, bmo);
        for (uint32_t f = 0; f < rm.s.idx_count / 3; ++f) {
            const float4 pl = rm.plane(f);
            const V3 nrm = rotate(sq(rs.orientation), v3(pl));
            const V3 inn = inv_rotate(sq(is.orientation), nrm);
            float support = 3.402823466e+38f;
            for (uint32_t k = 0; k < im.s.vtx_count; ++k) support = smin(support, dot(inn, im.vert(k)));
            const float gap = support + dot(nrm, sub(sv(is.position), sv(rs.position))) - pl.w;
            const float speed = dot(nrm, rel);
            if (speed == 0.0f) {
                if (gap > rest) return true;
            } else {
                const float crossing = (rest - gap) / speed;
                if (speed < 0.0f && crossing > entry) {
                    entry = crossing;
                    entry_normal = mul(nrm, side == 0 ? 1.0f : -1.0f);
                }
                if (speed > 0.0f) exit_time = smin(exit_time, crossing);
            }
        }
    }
    if (entry > exit_time || entry > 1.0f || exit_time < 0.0f) return true;
    if (entry == 0.0f) {
        const bool clipped = convex_face_manifold(b, bm, c, cm, margin, 0.0f, out);
        if (clipped && out.count != 0u) return true;
        if (!convex_face_manifold(pb, bm, pc, cm, margin, 0.0f, out) || out.count == 0u) return false;
        for (uint32_t p = 0; p < out.count; ++p) {
            const float closing = -dot(sub(bmo, cmo), out.c[p].normal);
            if (closing <= 0.0f) return false;
            out.c[p].pen += closing;
            out.c[p].impact_fraction = 0.0f;
        }
        return true;
    }
    PmhRigidBodyState ib = pb, ic = pc;
    ib.position = pv(add(sv(pb.position), mul(bmo, entry)));
    ic.position = pv(add(sv(pc.position), mul(cmo, entry)));
    if (!convex_face_manifold(ib, bm, ic, cm, rest + 1.0e-5f, 0.0f, out) || out.count == 0u) return false;
    for (uint32_t p = 0; p < out.count; ++p) {
        Contact &ct = out.c[p];
        if (dot(ct.normal, entry_normal) < 0.99999f) return false;
        const float remaining = -dot(sub(bmo, cmo), ct.normal) * (1.0f - entry);
        ct.pen = smax(0.0f, remaining + ct.pen) + rest + 1.0e-5f;
        ct.impact_fraction = entry;
    }
    return true;
}

__device__ bool requires_swept(const PmhRigidBodyState &pa, const PmhRigidBodyState &a, const ShapeDev &ma,
                               const PmhRigidBodyState &pb, const PmhRigidBodyState &b, const ShapeDev &mb,
                               float threshold) {
    const bool planes_a = ma.kind != kShapeMesh || ma.has_planes;
    const bool planes_b = mb.kind != kShapeMesh || mb.has_planes;
    if (!planes_a || !planes_b) threshold = smin(threshold, 1.0e-3f);
    const V3 rel = sub(sub(sv(a.position), sv(pa.position)), sub(sv(b.position), sv(pb.position)));
    return len(rel) + rotational_motion_bound(pa, a, ma) + rotational_motion_bound(pb, b, mb) > threshold;
}

// Analytic round shapes (no reference counterpart).  The normal points
// from the collider towards the body; the point is the midpoint of the two
// surface points; penetration = rest offset - separation + 1e-5.
__device__ __forceinline__ V3 closest_on_segment(V3 p0, V3 p1, V3 q) {
    const V3 d = sub(p1, p0);
    const float dd = len2(d);
    const float t = dd > 0.0f ? clampf(dot(sub(q, p0), d) / dd, 0.0f, 1.0f) : 0.0f;
    return add(p0, mul(d, t));
}

__device__ bool point_shape_distance(V3 c, const PmhRigidBodyState &ss, const ShapeDev &sh, V3 &normal,
                                     V3 &surface, float &distance) {
    const Q o = sq(ss.orientation);
    const V3 sp = sv(ss.position);
    if (sh.kind == kShapeSphere) {
        const V3 d = sub(c, sp);
        normal = normalized_or(d, {0.0f, 1.0f, 0.0f});
        distance = len(d) - sh.sphere_radius;
        surface = add(sp, mul(normal, sh.sphere_radius));
        return true;
    }
    if (sh.kind == kShapeCapsule) {
        const V3 axis = rotate(o, {0.0f, sh.half_length, 0.0f});
        const V3 p = closest_on_segment(sub(sp, axis), add(sp, axis), c);
        const V3 d = sub(c, p);
        normal = normalized_or(d, rotate(o, {1.0f, 0.0f, 0.0f}));
        distance = len(d) - sh.sphere_radius;
        surface = add(p, mul(normal, sh.sphere_radius));
        return true;
    }
    if (!sh.is_box) return false;
    const V3 half = sarr(sh.box_half);
    const V3 l = inv_rotate(o, sub(c, sp));
    V3 q = {clampf(l.x, -half.x, half.x), clampf(l.y, -half.y, half.y), clampf(l.z, -half.z, half.z)};
    const bool inside = q.x == l.x && q.y == l.y && q.z == l.z;
    V3 ln;
    float dist;
    if (!inside) {
        const V3 d = sub(l, q);
        const float dl = len(d);
        ln = mul(d, 1.0f / dl);
        dist = dl;
    } else {
        const float dx = half.x - fabsf(l.x), dy = half.y - fabsf(l.y), dz = half.z - fabsf(l.z);
        if (dx <= dy && dx <= dz) { ln = {l.x < 0.0f ? -1.0f : 1.0f, 0.0f, 0.0f}; q.x = l.x < 0.0f ? -half.x : half.x; dist = -dx; }
        else if (dy <= dz) { ln = {0.0f, l.y < 0.0f ? -1.0f : 1.0f, 0.0f}; q.y = l.y < 0.0f ? -half.y : half.y; dist = -dy; }
        else { ln = {0.0f, 0.0f, l.z < 0.0f ? -1.0f : 1.0f}; q.z = l.z < 0.0f ? -half.z : half.z; dist = -dz; }
    }
    normal = rotate(o, ln);
    surface = transform_point(ss, q);
    distance = dist;
    return true;
}

__device__ void push_round_contact(Manifold &m, V3 normal, V3 surface_collider, float gap, float margin,
                                   float reach) {
    if (gap > reach) return;
    const float rest = smin(margin, 1.0e-3f);
    Contact c;
    c.normal = normal;
    c.point = add(surface_collider, mul(normal, gap * 0.5f));
    c.pen = rest - gap + 1.0e-5f;
    c.impact_fraction = 1.0f;
    add_contact(m, c, smax(margin * 2.0f, 1.0e-4f));
}

__device__ bool analytic_manifold(const PmhRigidBodyState &as, const ShapeDev &am, const PmhRigidBodyState &bs,
                                  const ShapeDev &bm, float margin, float reach, Manifold &out) {
    out.count = 0;
    out.face_patch = 0;
    const bool a_round = am.kind == kShapeSphere || am.kind == kShapeCapsule;
    const bool b_round = bm.kind == kShapeSphere || bm.kind == kShapeCapsule;
    if (!a_round && !b_round) return false;
    V3 n, s;
    float d;
    if (am.kind == kShapeSphere) {
        if (!point_shape_distance(sv(as.position), bs, bm, n, s, d)) return false;
        push_round_contact(out, n, s, d - am.sphere_radius, margin, reach);
        return true;
    }
    if (bm.kind == kShapeSphere) {
        if (!point_shape_distance(sv(bs.position), as, am, n, s, d)) return false;
        const V3 nb = mul(n, -1.0f);
        const V3 sb = sub(sv(bs.position), mul(n, bm.sphere_radius));
        push_round_contact(out, nb, sb, d - bm.sphere_radius, margin, reach);
        return true;
    }
    const bool a_cap = am.kind == kShapeCapsule;
    const ShapeDev &cap = a_cap ? am : bm;
    const PmhRigidBodyState &cs = a_cap ? as : bs;
    const ShapeDev &other = a_cap ? bm : am;
    const PmhRigidBodyState &os = a_cap ? bs : as;
    const V3 axis = rotate(sq(cs.orientation), {0.0f, cap.half_length, 0.0f});
    V3 ends[3] = {sub(sv(cs.position), axis), add(sv(cs.position), axis), sv(cs.position)};
    uint32_t samples = 2;
    if (other.kind == kShapeCapsule) {
        const V3 oaxis = rotate(sq(os.orientation), {0.0f, other.half_length, 0.0f});
        ends[2] = closest_on_segment(ends[0], ends[1],
                                     closest_on_segment(sub(sv(os.position), oaxis), add(sv(os.position), oaxis),
                                                        sv(cs.position)));
        samples = 3;
    }
    for (uint32_t k = 0; k < samples; ++k) {
        if (!point_shape_distance(ends[k], os, other, n, s, d)) return false;
        const float gap = d - cap.sphere_radius;
        if (a_cap) push_round_contact(out, n, s, gap, margin, reach);
        else push_round_contact(out, mul(n, -1.0f), sub(ends[k], mul(n, cap.sphere_radius)), gap, margin, reach);
    }
    return true;
}

// ---------------------------------------------------------------------
// Triangle-pair narrowphase: mesh pairs the convex face path does not take
// (edge-on-edge contacts of convex meshes, non-convex or large meshes).
// Same steps and float operations as rigid_collide_triangle_ranges and its
// swept variant (geometry.slang:949-1117, 1272-1584, 1855-2154), the
// leaf-pair candidate order of pm_generate_rigid_leaf_pairs (:2303-2492),
// the per-candidate merge of pm_evaluate_rigid_leaf_pairs (:2525-2706)
// and the BVH traversal of rigid_collide_meshes (:2156-2285).  One thread
// per listed pair.
// ---------------------------------------------------------------------
__device__ __forceinline__ float normal_speed(const PmhRigidBodyState &b, const PmhRigidBodyState &c, V3 p, V3 n) {
    const V3 bv = add(sv(b.linear_velocity), cross(sv(b.angular_velocity), sub(p, sv(b.position))));
    const V3 cv = add(sv(c.linear_velocity), cross(sv(c.angular_velocity), sub(p, sv(c.position))));
    return dot(sub(bv, cv), n);
}

__device__ void closest_segments(V3 p0, V3 p1, V3 q0, V3 q1, V3 &on_p, V3 &on_q) {
    const V3 d0 = sub(p1, p0), d1 = sub(q1, q0), r = sub(p0, q0);
    const float a = dot(d0, d0), e = dot(d1, d1);
    float s = 0.0f, t = 0.0f;
    if (a <= 1.0e-12f && e <= 1.0e-12f) { on_p = p0; on_q = q0; return; }
    if (a <= 1.0e-12f) {
        t = clampf(dot(d1, r) / e, 0.0f, 1.0f);
    } else {
        const float c = dot(d0, r);
        if (e <= 1.0e-12f) {
            s = clampf(-c / a, 0.0f, 1.0f);
        } else {
            const float b = dot(d0, d1), f = dot(d1, r);
            const float den = a * e - b * b;
            if (fabsf(den) > kEpsilon * a * e) s = clampf((b * f - c * e) / den, 0.0f, 1.0f);
            t = (b * s + f) / e;
            if (t < 0.0f) { t = 0.0f; s = clampf(-c / a, 0.0f, 1.0f); }
            else if (t > 1.0f) { t = 1.0f; s = clampf((b - c) / a, 0.0f, 1.0f); }
        }
    }
    on_p = add(p0, mul(d0, s));
    on_q = add(q0, mul(d1, t));
}

__device__ V3 closest_on_triangle(V3 p, V3 a, V3 b, V3 c) {
    const V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    const V3 bp = sub(p, b);
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return add(a, mul(ab, d1 / (d1 - d3)));
    const V3 cp = sub(p, c);
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return add(a, mul(ac, d2 / (d2 - d6)));
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && d4 - d3 >= 0.0f && d5 - d6 >= 0.0f)
        return add(b, mul(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    const V3 n = cross(ab, ac);
    return sub(p, mul(n, dot(ap, n) / len2(n)));
}

__device__ bool segment_hits_triangle(V3 p, V3 q, V3 a, V3 b, V3 c, V3 &hit) {
    hit = {0.0f, 0.0f, 0.0f};
    const V3 n = cross(sub(b, a), sub(c, a));
    const V3 dir = sub(q, p);
    const float den = dot(n, dir);
    if (len2(n) <= kEpsilon * kEpsilon || fabsf(den) <= kEpsilon) return false;
    const float amount = dot(n, sub(a, p)) / den;
    if (amount < 0.0f || amount > 1.0f) return false;
    hit = add(p, mul(dir, amount));
    const float tol = -1.0e-5f * len2(n);
    return dot(cross(sub(b, a), sub(hit, a)), n) >= tol && dot(cross(sub(c, b), sub(hit, b)), n) >= tol &&
           dot(cross(sub(a, c), sub(hit, c)), n) >= tol;
}

__device__ __forceinline__ void keep_closer(V3 a, V3 b, float &best, V3 &best_a, V3 &best_b) {
    const float d = len2(sub(a, b));
    if (d < best) { best = d; best_a = a; best_b = b; }
}

__device__ __noinline__ void closest_triangle_pair(const V3 *A, const V3 *B, V3 &pa, V3 &pb) {
    pa = {0.0f, 0.0f, 0.0f};
    pb = {0.0f, 0.0f, 0.0f};
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

__device__ __forceinline__ bool boxes_touch(V3 amin, V3 amax, V3 bmin, V3 bmax) {
    return amin.x <= bmax.x && amax.x >= bmin.x && amin.y <= bmax.y && amax.y >= bmin.y && amin.z <= bmax.z &&
           amax.z >= bmin.z;
}

__device__ V3 advect_triangle_point(V3 q, const V3 *from, const V3 *to) {
    const V3 ab = sub(from[1], from[0]), ac = sub(from[2], from[0]);
    const V3 n = cross(ab, ac);
    const float sq2 = len2(n);
    if (sq2 <= 1.0e-20f) return to[0];
    const V3 rel = sub(q, from[0]);
    const float v = dot(cross(rel, ac), n) / sq2;
    const float w = dot(cross(ab, rel), n) / sq2;
    return add(to[0], add(mul(sub(to[1], to[0]), v), mul(sub(to[2], to[0]), w)));
}

// Certified outward direction of a (possibly swept) world triangle: unit
// for exterior faces, a quarter for two-sided interior features, zero when
// uncertified (rigid_shell_outward, geometry.slang:1308-1322).
__device__ V3 shell_outward(const ShapeView &m, uint32_t t, const V3 *W) {
    if (!m.s.has_shell) return {0.0f, 0.0f, 0.0f};
    const V3 o = v3(m.shell[m.s.shell_off + t]);
    const float sq2 = len2(o);
    if (sq2 < 0.01f) return {0.0f, 0.0f, 0.0f};
    const V3 la = m.vert(m.index(3 * t)), lb = m.vert(m.index(3 * t + 1)), lc = m.vert(m.index(3 * t + 2));
    const float sign = dot(cross(sub(lb, la), sub(lc, la)), o) >= 0.0f ? 1.0f : -1.0f;
    return mul(normalized_or(cross(sub(W[1], W[0]), sub(W[2], W[0])), {0.0f, 0.0f, 0.0f}),
               sign * (sq2 > 0.5f ? 1.0f : 0.25f));
}

__device__ __forceinline__ bool shell_facing(V3 n, V3 oa, V3 ob) {
    const float sa = len2(oa), sb = len2(ob);
    return (sa < 0.01f || (sa > 0.5f ? dot(n, oa) < -1.0e-3f : fabsf(dot(n, oa)) > 2.5e-4f)) &&
           (sb < 0.01f || (sb > 0.5f ? dot(n, ob) > 1.0e-3f : fabsf(dot(n, ob)) > 2.5e-4f));
}

__device__ bool point_inside_convex(const ShapeView &m, const PmhRigidBodyState &st, V3 world) {
    if (!m.s.has_planes) return false;
    const V3 l = inv_rotate(sq(st.orientation), sub(world, sv(st.position)));
    const uint32_t tris = m.s.idx_count / 3u;
    const uint32_t unique = m.s.unique_planes;
    const uint32_t count = unique != 0u ? unique : tris;
    const uint32_t first = unique != 0u ? tris : 0u;
    for (uint32_t f = 0; f < count; ++f) {
        const float4 pl = m.plane(first + f);
        if (dot(v3(pl), l) > pl.w + 1.0e-5f) return false;
    }
    return true;
}

__device__ __noinline__ V3 triangle_feature_normal(const V3 *A, const V3 *B, V3 pa, V3 pb, V3 fallback) {
    const V3 delta = sub(pa, pb);
    V3 normal = normalized_or(delta, fallback);
    const V3 faces[2] = {normalized_or(cross(sub(A[1], A[0]), sub(A[2], A[0])), normal),
                         normalized_or(cross(sub(B[1], B[0]), sub(B[2], B[0])), normal)};
    float align_best = -1.0f;
    V3 interior_normal = {0.0f, 0.0f, 0.0f};
    for (uint32_t side = 0; side < 2; ++side) {
        const V3 *T = side == 0 ? A : B;
        const V3 pt = side == 0 ? pa : pb;
        const V3 raw = cross(sub(T[1], T[0]), sub(T[2], T[0]));
        const float tol = 1.0e-5f * len2(raw);
        bool interior = tol > 0.0f;
        for (uint32_t e = 0; e < 3; ++e)
            interior = interior && dot(cross(sub(T[(e + 1) % 3], T[e]), sub(pt, T[e])), raw) > tol;
        const float alignment = fabsf(dot(faces[side], normal));
        if (interior && alignment > align_best) {
            align_best = alignment;
            const float proj = dot(delta, faces[side]);
            const float dir = fabsf(proj) > kEpsilon ? proj : dot(fallback, faces[side]);
            interior_normal = dir >= 0.0f ? faces[side] : mul(faces[side], -1.0f);
        }
    }
    if (align_best >= 0.0f) return interior_normal;
    float best = 4.0f * kEpsilon * kEpsilon;
    if (len2(delta) > kEpsilon * kEpsilon) {
        for (uint32_t side = 0; side < 2; ++side) {
            const float proj = dot(delta, faces[side]);
            const float err = len2(sub(delta, mul(faces[side], proj)));
            if (err < best) {
                best = err;
                normal = proj >= 0.0f ? faces[side] : mul(faces[side], -1.0f);
            }
        }
    }
    return normal;
}

// Inputs of one pair on the triangle path: shapes, previous and predicted
// states, the pair's hinge contact frames and the body's reference points
// (rigid_hinge_contact_frame, geometry.slang:404-491).
struct TrianglePair {
    ShapeView am, bm;
    PmhRigidBodyState ap, a, bp, b;   // previous and predicted states
    PairFrames f;
    float margin, dt;
};

__device__ __forceinline__ void world_triangle(const ShapeView &m, const PmhRigidBodyState &st, uint32_t t, V3 *W) {
    for (uint32_t c = 0; c < 3; ++c) W[c] = transform_point(st, m.vert(m.index(3 * t + c)));
}

// The discrete pass for one triangle pair (the body of the triangle loops
// of rigid_collide_triangle_ranges, geometry.slang:1359-1584): at most one
// contact, independent of every other triangle pair.
__device__ __noinline__ bool triangle_pair_contact(const TrianglePair &P, uint32_t at, uint32_t bt, bool inside_a,
                                                   bool inside_b, Contact &c) {
    const float margin = P.margin;
    const float rest = smin(margin, 1.0e-3f);
    const V3 body_ref = P.f.ref;
    V3 A[3], B[3];
    world_triangle(P.am, P.a, at, A);
    world_triangle(P.bm, P.b, bt, B);
    {
        const V3 e = {margin, margin, margin};
        const V3 amin = sub(cmin(A[0], cmin(A[1], A[2])), e), amax = add(cmax(A[0], cmax(A[1], A[2])), e);
        const V3 bmin = cmin(B[0], cmin(B[1], B[2])), bmax = cmax(B[0], cmax(B[1], B[2]));
        if (!boxes_touch(amin, amax, bmin, bmax)) return false;
    }
    const V3 oa = shell_outward(P.am, at, A), ob = shell_outward(P.bm, bt, B);
    const V3 coa = P.f.ha.static_body != 0u && inside_a ? V3{0.0f, 0.0f, 0.0f} : oa;
    const V3 cob = P.f.hb.static_body != 0u && inside_b ? V3{0.0f, 0.0f, 0.0f} : ob;
    if (dot(oa, ob) > 0.9999f) return false;
    V3 pa, pb;
    closest_triangle_pair(A, B, pa, pb);
    const V3 delta = sub(pa, pb);
    const float sq2 = len2(delta);
    if (sq2 > margin * margin) return false;
    const V3 cn = normalized_or(cross(sub(B[1], B[0]), sub(B[2], B[0])), {0.0f, 1.0f, 0.0f});
    V3 fallback = dot(cn, sub(body_ref, pb)) >= 0.0f ? cn : mul(cn, -1.0f);
    if (len2(cob) > 0.5f) fallback = cob;
    const float distance = sqrtf(smax(sq2, 0.0f));
    const V3 point = mul(add(pa, pb), 0.5f);
    const V3 proj = sub(pa, mul(cn, dot(sub(pa, B[0]), cn)));
    const bool on_face = len2(sub(proj, closest_on_triangle(proj, B[0], B[1], B[2]))) <= 1.0e-10f;
    const V3 normal = (distance <= 1.0e-5f || on_face) ? fallback : normalized_or(delta, fallback);
    if (on_face && len2(oa) < 0.01f && len2(ob) > 0.5f && P.bm.s.has_planes) {
        const V3 face_a = normalized_or(cross(sub(A[1], A[0]), sub(A[2], A[0])), {0.0f, 0.0f, 0.0f});
        if (fabsf(dot(face_a, normal)) < 1.0e-3f) {
            const float depth = smax(0.0f, -smin(dot(sub(A[0], B[0]), normal),
                                                 smin(dot(sub(A[1], B[0]), normal), dot(sub(A[2], B[0]), normal))));
            const V3 axis = inv_rotate(sq(P.b.orientation), face_a);
            const float offset = dot(sub(sv(P.b.position), A[0]), face_a);
            float lo = 3.402823466e+38f, hi = -3.402823466e+38f;
            for (uint32_t k = 0; k < P.bm.s.vtx_count; ++k) {
                const float pr = dot(P.bm.vert(k), axis) + offset;
                lo = smin(lo, pr);
                hi = smax(hi, pr);
            }
            const float alternate = smax(0.0f, smin(-lo, hi));
            if (alternate + 1.0e-5f < depth) return false;
        }
    }
    const float separation = on_face && len2(cob) > 0.5f ? dot(delta, normal) : distance;
    // rigid_contact_reaches_rest_offset (:1282-1290)
    if (!(separation <= rest + 1.0e-5f)) {
        const float speed = normal_speed(P.a, P.b, point, normal);
        if (speed > 1.0e-5f || !(-speed * P.dt + rest + 1.0e-5f >= separation)) return false;
    }
    c.normal = normal;
    c.point = point;
    c.pen = rest - separation + 1.0e-5f;
    bool recovered = false;
    if (!P.am.s.has_planes && !P.bm.s.has_planes && distance <= 1.0e-5f) {
        V3 oldA[3], oldB[3];
        world_triangle(P.am, P.ap, at, oldA);
        world_triangle(P.bm, P.bp, bt, oldB);
        V3 opa, opb;
        closest_triangle_pair(oldA, oldB, opa, opb);
        V3 od = sub(opa, opb);
        if (len2(od) > 1.0e-10f) {
            V3 IA[3], IB[3];
            float lo = 0.0f, hi = 1.0f;
            const float gap = smin(rest, 0.5f * sqrtf(len2(od)));
            for (uint32_t k = 0; k < 10u; ++k) {
                const float time = 0.5f * (lo + hi);
                for (uint32_t q = 0; q < 3; ++q) {
                    IA[q] = add(oldA[q], mul(sub(A[q], oldA[q]), time));
                    IB[q] = add(oldB[q], mul(sub(B[q], oldB[q]), time));
                }
                V3 u, w;
                closest_triangle_pair(IA, IB, u, w);
                if (len2(sub(u, w)) > gap * gap) lo = time;
                else hi = time;
            }
            for (uint32_t q = 0; q < 3; ++q) {
                IA[q] = add(oldA[q], mul(sub(A[q], oldA[q]), lo));
                IB[q] = add(oldB[q], mul(sub(B[q], oldB[q]), lo));
            }
            closest_triangle_pair(IA, IB, opa, opb);
            od = sub(opa, opb);
            c.normal = triangle_feature_normal(IA, IB, opa, opb, normalized_or(od, fallback));
            const V3 ca = advect_triangle_point(opa, IA, A), cb = advect_triangle_point(opb, IB, B);
            c.point = mul(add(ca, cb), 0.5f);
            c.pen = rest - dot(sub(ca, cb), c.normal) + 1.0e-5f;
            recovered = true;
        }
    }
    if (distance <= 1.0e-5f) {
        if (!recovered && (P.f.ha.fixed_member != 0u || P.f.hb.fixed_member != 0u || on_face)) {
            const float depth = smax(0.0f, -smin(dot(sub(A[0], pb), normal),
                                                 smin(dot(sub(A[1], pb), normal), dot(sub(A[2], pb), normal))));
            c.pen = smin(margin, depth + rest) + 1.0e-5f;
        } else if (!recovered) {
            c.pen = margin + 1.0e-5f;
        }
    }
    if (!shell_facing(c.normal, coa, cob)) return false;
    c.impact_fraction = 1.0f;
    return true;
}

// rigid_triangle_pair_face_contact, geometry.slang:1910-1919.
__device__ __forceinline__ bool triangle_pair_face_contact(const V3 *A, const V3 *B, V3 separation) {
    const V3 n = normalized_or(separation, {0.0f, 0.0f, 0.0f});
    const V3 fa = normalized_or(cross(sub(A[1], A[0]), sub(A[2], A[0])), {0.0f, 0.0f, 0.0f});
    const V3 fb = normalized_or(cross(sub(B[1], B[0]), sub(B[2], B[0])), {0.0f, 0.0f, 0.0f});
    return fabsf(dot(n, fa)) > 0.9999f || fabsf(dot(n, fb)) > 0.9999f;
}

// The swept pass for one triangle pair (conservative advancement, the body
// of rigid_collide_triangle_ranges_swept, geometry.slang:1921-2154): at
// most one contact.  A guided pair rests at 0.1 mm, takes edge contacts that
// both slide along and turn about the guide axis at zero offset, and
// advances by the faces' closing speed along the feature normal.
__device__ __noinline__ bool triangle_pair_swept_contact(const TrianglePair &P, uint32_t at, uint32_t bt,
                                                         bool inside_a, bool inside_b, Contact &c) {
    const float margin = P.margin;
    const bool guided = guided_pair(P.f.ha, P.f.hb);
    const float rest = guided ? smin(margin, 1.0e-4f) : smin(margin, 1.0e-3f);
    V3 pA[3], cA[3], pB[3], cB[3];
    world_triangle(P.am, P.ap, at, pA);
    world_triangle(P.am, P.a, at, cA);
    world_triangle(P.bm, P.bp, bt, pB);
    world_triangle(P.bm, P.b, bt, cB);
    V3 amin = cmin(pA[0], cA[0]), amax = cmax(pA[0], cA[0]);
    for (uint32_t q = 1; q < 3; ++q) {
        amin = cmin(amin, cmin(pA[q], cA[q]));
        amax = cmax(amax, cmax(pA[q], cA[q]));
    }
    V3 bmin = cmin(pB[0], cB[0]), bmax = cmax(pB[0], cB[0]);
    for (uint32_t q = 1; q < 3; ++q) {
        bmin = cmin(bmin, cmin(pB[q], cB[q]));
        bmax = cmax(bmax, cmax(pB[q], cB[q]));
    }
    const V3 e = {margin, margin, margin};
    if (!boxes_touch(sub(amin, e), add(amax, e), bmin, bmax)) return false;
    V3 dA[3], dB[3];
    float speed_abs = 0.0f;
    for (uint32_t q = 0; q < 3; ++q) {
        dA[q] = sub(cA[q], pA[q]);
        dB[q] = sub(cB[q], pB[q]);
        speed_abs = smax(speed_abs, len(dA[q]));
    }
    float speed_b = 0.0f;
    for (uint32_t q = 0; q < 3; ++q) speed_b = smax(speed_b, len(dB[q]));
    speed_abs += speed_b;
    const V3 common = dB[0];
    float rel_a = 0.0f, rel_b = 0.0f;
    for (uint32_t q = 0; q < 3; ++q) {
        rel_a = smax(rel_a, len(sub(dA[q], common)));
        rel_b = smax(rel_b, len(sub(dB[q], common)));
    }
    const float bound = smin(speed_abs, rel_a + rel_b);
    if (bound <= kEpsilon) return false;
    float time = 0.0f;
    bool starts_near = false;
    V3 sep_normal = {0.0f, 0.0f, 0.0f};
    bool have_sep = false;
    for (uint32_t it = 0; it < 32u; ++it) {
        V3 A[3], B[3];
        for (uint32_t q = 0; q < 3; ++q) {
            A[q] = add(pA[q], mul(dA[q], time));
            B[q] = add(pB[q], mul(dB[q], time));
        }
        V3 pa, pb;
        closest_triangle_pair(A, B, pa, pb);
        const V3 delta = sub(pa, pb);
        const float distance = sqrtf(smax(0.0f, len2(delta)));
        float offset = rest;
        if (guided && !triangle_pair_face_contact(A, B, delta)) {
            const HingeFrame &guide = P.f.ha.axial != 0u ? P.f.ha : P.f.hb;
            const V3 direction = normalized_or(delta, {0.0f, 0.0f, 0.0f});
            const float axial = dot(direction, sv(guide.axis));
            const float angular = dot(sv(guide.axis), cross(sub(pa, sv(guide.anchor)), direction));
            if (fabsf(axial) > 1.0e-3f && fabsf(angular) > 1.0e-3f) offset = 0.0f;
        }
        if (it == 0u) starts_near = distance <= offset + 5.0e-5f;
        if (distance <= offset + 1.0e-5f) {
            const V3 cn = normalized_or(cross(sub(B[1], B[0]), sub(B[2], B[0])), {0.0f, 1.0f, 0.0f});
            const V3 center = add(sv(P.ap.position), mul(sub(sv(P.a.position), sv(P.ap.position)), time));
            V3 fallback = dot(cn, sub(center, pb)) >= 0.0f ? cn : mul(cn, -1.0f);
            if (P.f.ha.present != 0u) {
                const V3 reference = add(P.f.ref_prev, mul(sub(P.f.ref, P.f.ref_prev), time));
                fallback = dot(cn, sub(reference, pb)) >= 0.0f ? cn : mul(cn, -1.0f);
            }
            const V3 oa = shell_outward(P.am, at, A), ob = shell_outward(P.bm, bt, B);
            const V3 coa = P.f.ha.static_body != 0u && inside_a ? V3{0.0f, 0.0f, 0.0f} : oa;
            const V3 cob = P.f.hb.static_body != 0u && inside_b ? V3{0.0f, 0.0f, 0.0f} : ob;
            if (dot(oa, ob) > 0.9999f) return false;
            if (len2(cob) > 0.5f) fallback = cob;
            else if (len2(coa) > 0.5f) fallback = mul(coa, -1.0f);
            const V3 normal = triangle_feature_normal(A, B, pa, pb, have_sep ? sep_normal : fallback);
            if (!shell_facing(normal, coa, cob)) return false;
            const V3 point = mul(add(pa, pb), 0.5f);
            const float vn = normal_speed(P.a, P.b, point, normal);
            if (vn > 1.0e-5f) return false;
            const float remaining = -vn * P.dt * (1.0f - time) - distance;
            c.normal = normal;
            c.point = point;
            c.pen = smax(0.0f, remaining + offset) + 1.0e-5f;
            c.impact_fraction = starts_near ? 0.0f : time;
            return true;
        }
        sep_normal = triangle_feature_normal(A, B, pa, pb, normalized_or(delta, {0.0f, 1.0f, 0.0f}));
        have_sep = true;
        float advance = (distance - offset) / (bound + kEpsilon) * 0.9f;
        if (guided) {
            const V3 plane = triangle_feature_normal(A, B, pa, pb, sep_normal);
            float min_a = 3.402823466e+38f, max_b = -3.402823466e+38f;
            float min_speed_a = 3.402823466e+38f, max_speed_b = -3.402823466e+38f;
            for (uint32_t q = 0; q < 3; ++q) {
                min_a = smin(min_a, dot(sub(A[q], pb), plane));
                max_b = smax(max_b, dot(sub(B[q], pb), plane));
                min_speed_a = smin(min_speed_a, dot(dA[q], plane));
                max_speed_b = smax(max_speed_b, dot(dB[q], plane));
            }
            const float gap = min_a - max_b - offset;
            const float closing = max_speed_b - min_speed_a;
            if (gap > 1.0e-5f && closing <= 0.0f) return false;
            if (gap > 0.0f && closing > kEpsilon) advance = smax(advance, 0.9f * gap / closing);
        }
        time += smax(advance, 1.0e-5f);
        if (time > 1.0f) return false;
    }
    return false;
}

__device__ __forceinline__ void inside_flags(const TrianglePair &P, bool &inside_a, bool &inside_b) {
    inside_a = P.f.ha.static_body != 0u && point_inside_convex(P.am, P.ap, sv(P.bp.position));
    inside_b = P.f.hb.static_body != 0u && point_inside_convex(P.bm, P.bp, sv(P.ap.position));
}

// The merge rule of the pair: guided pairs merge by guide direction.
__device__ __forceinline__ void merge_contact(const TrianglePair &P, Manifold &m, const Contact &c, float spacing) {
    if (guided_pair(P.f.ha, P.f.hb)) add_contact_guided(m, c, P.f.ha.axial != 0u ? P.f.ha : P.f.hb);
    else add_contact(m, c, spacing);
}

// rigid_collide_triangle_ranges: triangles in order, body-major.
__device__ void collide_triangle_ranges(const TrianglePair &P, uint32_t a_first, uint32_t a_count,
                                        uint32_t b_first, uint32_t b_count, Manifold &m) {
    if (guided_pair(P.f.ha, P.f.hb)) return;   // geometry.slang:1371
    bool inside_a, inside_b;
    inside_flags(P, inside_a, inside_b);
    const float spacing = smax(2.0f * P.margin, 1.0e-4f);
    for (uint32_t ai = 0; ai < a_count; ++ai) {
        const uint32_t at = a_first + ai;
        if (3u * at + 2u >= P.am.s.idx_count) break;
        for (uint32_t bi = 0; bi < b_count; ++bi) {
            const uint32_t bt = b_first + bi;
            if (3u * bt + 2u >= P.bm.s.idx_count) break;
            Contact c;
            if (triangle_pair_contact(P, at, bt, inside_a, inside_b, c)) add_contact(m, c, spacing);
        }
    }
}

__device__ void collide_triangle_ranges_swept(const TrianglePair &P, uint32_t a_first, uint32_t a_count,
                                              uint32_t b_first, uint32_t b_count, Manifold &m) {
    bool inside_a, inside_b;
    inside_flags(P, inside_a, inside_b);
    const float spacing = smax(P.margin * 2.0f, 1.0e-4f);
    for (uint32_t ai = 0; ai < a_count; ++ai) {
        const uint32_t at = a_first + ai;
        if (3u * at + 2u >= P.am.s.idx_count) break;
        for (uint32_t bi = 0; bi < b_count; ++bi) {
            const uint32_t bt = b_first + bi;
            if (3u * bt + 2u >= P.bm.s.idx_count) break;
            Contact c;
            if (triangle_pair_swept_contact(P, at, bt, inside_a, inside_b, c)) merge_contact(P, m, c, spacing);
        }
    }
}

// BoundsTransform and transformed_(motion_)bounds, math.slang:142-190.
struct BoundsFrame {
    V3 position, ax, ay, az;
};
__device__ __forceinline__ BoundsFrame bounds_frame(const PmhRigidBodyState &s) {
    const Q q = sq(s.orientation);
    return {sv(s.position), rotate(q, {1.0f, 0.0f, 0.0f}), rotate(q, {0.0f, 1.0f, 0.0f}), rotate(q, {0.0f, 0.0f, 1.0f})};
}
__device__ __forceinline__ void frame_bounds(V3 lmin, V3 lmax, const BoundsFrame &f, float margin, V3 &mn, V3 &mx) {
    const V3 c = mul(add(lmin, lmax), 0.5f), h = mul(sub(lmax, lmin), 0.5f);
    const V3 wc = add(f.position, add(mul(f.ax, c.x), add(mul(f.ay, c.y), mul(f.az, c.z))));
    const V3 wh = {fabsf(f.ax.x) * h.x + fabsf(f.ay.x) * h.y + fabsf(f.az.x) * h.z,
                   fabsf(f.ax.y) * h.x + fabsf(f.ay.y) * h.y + fabsf(f.az.y) * h.z,
                   fabsf(f.ax.z) * h.x + fabsf(f.ay.z) * h.y + fabsf(f.az.z) * h.z};
    const V3 e = {margin, margin, margin};
    mn = sub(sub(wc, wh), e);
    mx = add(add(wc, wh), e);
}
__device__ __forceinline__ void motion_bounds(V3 lmin, V3 lmax, const BoundsFrame &prev, const BoundsFrame &cur,
                                              bool swept, float margin, V3 &mn, V3 &mx) {
    frame_bounds(lmin, lmax, cur, margin, mn, mx);
    if (!swept) return;
    V3 pmn, pmx;
    frame_bounds(lmin, lmax, prev, margin, pmn, pmx);
    mn = cmin(mn, pmn);
    mx = cmax(mx, pmx);
}
__device__ __forceinline__ void node_bounds(const BvhDev &nd, const BoundsFrame &prev, const BoundsFrame &cur,
                                            bool swept, float margin, V3 &mn, V3 &mx) {
    motion_bounds(sarr(nd.mn), sarr(nd.mx), prev, cur, swept, margin, mn, mx);
}

__device__ __forceinline__ void collide_leaf_ranges(const TrianglePair &P, bool swept, const BvhDev &na,
                                                    const BvhDev &nb, Manifold &m) {
    collide_triangle_ranges(P, na.first, na.count, nb.first, nb.count, m);
    if (swept) collide_triangle_ranges_swept(P, na.first, na.count, nb.first, nb.count, m);
}

// rigid_collide_meshes: simultaneous descent with a 64-entry stack, all
// triangles when the stack would overflow.
__device__ __noinline__ void collide_meshes(const TrianglePair &P, bool swept, Manifold &m) {
    m.count = 0;
    m.face_patch = 0;
    const BoundsFrame fa = bounds_frame(P.a), fb = bounds_frame(P.b);
    const BoundsFrame pfa = swept ? bounds_frame(P.ap) : BoundsFrame{};
    const BoundsFrame pfb = swept ? bounds_frame(P.bp) : BoundsFrame{};
    V3 amn, amx, bmn, bmx;
    motion_bounds(sarr(P.am.s.mn), sarr(P.am.s.mx), pfa, fa, swept, P.margin, amn, amx);
    motion_bounds(sarr(P.bm.s.mn), sarr(P.bm.s.mx), pfb, fb, swept, 0.0f, bmn, bmx);
    if (!boxes_touch(amn, amx, bmn, bmx)) return;
    constexpr uint32_t kStack = 64;
    uint32_t sa[kStack], sb[kStack];
    sa[0] = 0;
    sb[0] = 0;
    uint32_t top = 1;
    bool overflow = P.am.s.node_count == 0u || P.bm.s.node_count == 0u;
    while (top > 0u && !overflow) {
        --top;
        const uint32_t ia = sa[top], ib = sb[top];
        const BvhDev na = P.am.node(ia), nb = P.bm.node(ib);
        node_bounds(na, pfa, fa, swept, P.margin, amn, amx);
        node_bounds(nb, pfb, fb, swept, 0.0f, bmn, bmx);
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
        m.count = 0;
        m.face_patch = 0;
        collide_triangle_ranges(P, 0, P.am.s.idx_count / 3u, 0, P.bm.s.idx_count / 3u, m);
        if (swept) collide_triangle_ranges_swept(P, 0, P.am.s.idx_count / 3u, 0, P.bm.s.idx_count / 3u, m);
    }
}

// Number of (body leaf, collider leaf) pairs whose motion bounds overlap.
__device__ uint32_t count_leaf_candidates(const TrianglePair &P, bool swept, const BoundsFrame &fa,
                                          const BoundsFrame &fb, const BoundsFrame &pfa, const BoundsFrame &pfb) {
    const uint32_t la = P.am.s.leaf_count, lb = P.bm.s.leaf_count;
    uint32_t total = 0;
    for (uint32_t i = 0; i < la; ++i) {
        V3 amn, amx;
        node_bounds(P.am.node(P.am.leaf(i)), pfa, fa, swept, P.margin, amn, amx);
        for (uint32_t j = 0; j < lb; ++j) {
            V3 bmn, bmx;
            node_bounds(P.bm.node(P.bm.leaf(j)), pfb, fb, swept, 0.0f, bmn, bmx);
            total += boxes_touch(amn, amx, bmn, bmx) ? 1u : 0u;
        }
    }
    return total;
}

// Narrowphase outputs: per-pair records and the point arrays, into which
// each non-empty manifold writes a contiguous range.
struct NarrowOut {
    uint32_t *m_count, *m_start, *m_face;
    float4 *m_init_rel;
    float4 *c_np, *c_pt, *c_aa, *c_ab, *c_err, *c_lam, *c_pen, *c_acc, *c_rep;
    uint32_t *c_flags, *c_pair;
    uint32_t point_cap;
    StepCounters *ctr;
};

// Reserves `count` (< 16) consecutive slots of *counter per active lane
// with one atomic per wavefront: each lane's offset is the sum of the
// counts of the active lanes below it, built from one ballot per bit.
// Works in divergent code (ballots and the broadcast see active lanes only).
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *counter, uint32_t count) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t active = __ballot(1);
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t offset = 0, total = 0;
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b) {
        const uint64_t bits = __ballot((count >> b) & 1u);
        offset += uint32_t(__popcll(bits & below)) << b;
        total += uint32_t(__popcll(bits)) << b;
    }
    const uint32_t leader = __builtin_ctzll(active);
    uint32_t base = 0;
    if (lane == leader && total) base = atomicAdd(counter, total);
    base = __shfl(base, leader, 64);
    return base + offset;
}

// pm_finalize_rigid_contact_manifolds (geometry.slang:2805-2873): the
// initial relative position, persistence and the initial normal speed, and
// the point range.  Where the range lies does not affect any result.
// Guided contacts persist; contacts of a hinged or fixed member do not
// (:2855-2862).  Point flags: 2 persistent, 8 guided (the solver's rest
// offset).
constexpr uint32_t kPointGuided = 8u;
__device__ void emit_manifold(const NarrowOut &o, uint32_t m, Manifold &man, const PmhRigidBodyState &A,
                              const PmhRigidBodyState &B, bool persistent_shapes, const HingeFrame &ha,
                              const HingeFrame &hb) {
    const bool guided = guided_pair(ha, hb);
    const bool persistent = guided || (persistent_shapes && ha.present == 0u && hb.present == 0u &&
                                       ha.fixed_member == 0u && hb.fixed_member == 0u);
    uint32_t start = wave_reserve(&o.ctr->point_alloc, man.count);
    if (man.count) {
        if (start + man.count > o.point_cap) {
            atomicOr(&o.ctr->status, kStatusPointOverflow);
            man.count = 0;
        }
    }
    o.m_count[m] = man.count;
    o.m_start[m] = start;
    o.m_face[m] = man.face_patch;
    o.m_init_rel[m] = f4(sub(sv(A.position), sv(B.position)), 0.0f);
    for (uint32_t k = 0; k < man.count; ++k) {
        const Contact &c = man.c[k];
        const float speed = normal_speed(A, B, c.point, c.normal);
        const uint32_t i = start + k;
        const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        o.c_np[i] = f4(c.normal, c.pen);
        o.c_pt[i] = f4(c.point, speed);
        o.c_aa[i] = make_float4(0.0f, 0.0f, 0.0f, c.impact_fraction);   // w: impact fraction
        o.c_ab[i] = z;
        o.c_err[i] = z;
        o.c_lam[i] = z;
        o.c_pen[i] = z;
        o.c_acc[i] = z;
        o.c_rep[i] = z;
        o.c_flags[i] = (persistent ? 2u : 0u) | (guided ? kPointGuided : 0u);
        o.c_pair[i] = m;
    }
}

__device__ __forceinline__ bool persistent_pair(const ShapeDev &a, const ShapeDev &b, const PmhBodyParameters &pb) {
    const auto sup = [](const ShapeDev &s) { return s.kind != kShapeMesh || (s.has_planes && s.idx_count <= 96u); };
    return sup(a) && (pb.motion != kDynamic || sup(b));
}

// One thread per pair: the convex face path (and its swept variant) for
// convex meshes, closed forms for round shapes.  Mesh pairs the face path
// declines are queued for the triangle-pair kernel.
// The narrowphase runs in two kernels so that no wavefront mixes the cheap
// analytic pairs with the convex face path's separating-axis search:
// k_narrowphase solves every pair with an analytic shape and queues the
// pairs of two meshes; k_face_pairs runs the face path over that queue and
// passes what it declines on to the triangle-pair path.  Each pair's output
// sits at its own index, so the queues' order changes nothing.
__global__ void __launch_bounds__(128) k_narrowphase(const uint32_t *npairs, const PmhBodyParameters *P,
                                                     const ShapeDev *shapes, const PmhRigidBodyState *prev,
                                                     const PmhRigidBodyState *pred, const uint2 *ab, NarrowOut o,
                                                     uint32_t *face_list, const HingeFrame *frames,
                                                     const float4 *frame_ref) {
    const uint32_t M = *npairs;
    uint32_t unsupported = 0;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint2 p = ab[m];
        const PmhBodyParameters pa = P[p.x], pb = P[p.y];
        const ShapeDev sa = shapes[pa.mesh_index], sb = shapes[pb.mesh_index];
        if (sa.kind == kShapeMesh && sb.kind == kShapeMesh) {
            face_list[wave_reserve(&o.ctr->face_pairs, 1u)] = m;
            continue;
        }
        const PmhRigidBodyState A0 = prev[p.x], A = pred[p.x], B = pred[p.y];
        const float margin = pa.collision_margin + pb.collision_margin;
        const float cr = correction_reach(pa, sa) + correction_reach(pb, sb);
        const PairFrames fr = pair_frames(frames, frame_ref, pa, pb, p, A0, A);
        Manifold man;
        man.count = 0;
        man.face_patch = 0;
        if (!analytic_manifold(A, sa, B, sb, margin, margin + cr, man)) {
            man.count = 0;
            man.face_patch = 0;
            ++unsupported;
        }
        emit_manifold(o, m, man, A, B, persistent_pair(sa, sb, pb), fr.ha, fr.hb);
    }
    if (unsupported) atomicAdd(&o.ctr->unsupported, unsupported);
}

__global__ void __launch_bounds__(128) k_face_pairs(const uint32_t *face_list, const PmhBodyParameters *P,
                                                    const ShapeDev *shapes, const V3 *verts, const uint32_t *idx,
                                                    const float4 *planes, const PmhRigidBodyState *prev,
                                                    const PmhRigidBodyState *pred, const uint2 *ab, NarrowOut o,
                                                    uint32_t *tri_list, const HingeFrame *frames,
                                                    const float4 *frame_ref) {
    const uint32_t F = o.ctr->face_pairs;
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < F; f += gridDim.x * blockDim.x) {
        const uint32_t m = face_list[f];
        const uint2 p = ab[m];
        const PmhBodyParameters pa = P[p.x], pb = P[p.y];
        ShapeView sa{shapes[pa.mesh_index], verts, idx, planes};
        ShapeView sb{shapes[pb.mesh_index], verts, idx, planes};
        const PmhRigidBodyState A0 = prev[p.x], A = pred[p.x], B0 = prev[p.y], B = pred[p.y];
        const float margin = pa.collision_margin + pb.collision_margin;
        const float cr = correction_reach(pa, sa.s) + correction_reach(pb, sb.s);
        const PairFrames fr = pair_frames(frames, frame_ref, pa, pb, p, A0, A);
        Manifold man;
        man.count = 0;
        man.face_patch = 0;
        bool face = false;
        // A guided pair leaves the face path (geometry.slang:2779).
        if (sa.s.convex_face && sb.s.convex_face && !guided_pair(fr.ha, fr.hb)) {
            const bool swept = requires_swept(A0, A, sa.s, B0, B, sb.s, margin);
            if (swept) face = swept_convex_face_manifold(A0, A, sa, B0, B, sb, margin, man);
            if (!face) {
                face = convex_face_manifold(A, sa, B, sb, margin, cr, man);
                if (swept && man.count == 0u) face = false;
            }
        }
        if (!face) {
            tri_list[wave_reserve(&o.ctr->tri_pairs, 1u)] = m;
            continue;
        }
        emit_manifold(o, m, man, A, B, persistent_pair(sa.s, sb.s, pb), fr.ha, fr.hb);
    }
}

// Scenes of a batch: each body's scene, each scene's body count and the
// index of its first pair (pairs are sorted by body, scenes are contiguous
// body ranges).
struct SceneRules {
    const uint32_t *scene = nullptr;
    const uint32_t *count = nullptr;
    const uint32_t *pair_start = nullptr;
};

// Candidate slots of a world of n bodies: the contact-cache slot rule,
// min(n(n-1)/2, max(4096, 8n)) (resource_plan.hpp:151-164).
__host__ __device__ inline uint32_t leaf_slot_rule(uint64_t n) {
    const uint64_t pairs = n * (n - 1) / 2;
    const uint64_t cap = 8 * n > 4096 ? 8 * n : 4096;
    return uint32_t(pairs < cap ? pairs : cap);
}

__global__ void k_scene_pairs(const uint32_t *npairs, const uint2 *ab, const uint32_t *scene, uint32_t *pair_start) {
    const uint32_t M = *npairs;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint32_t s = scene[ab[m].x];
        if (m == 0 || scene[ab[m - 1].x] != s) pair_start[s] = m;
    }
}

// ---- triangle-pair kernel: one wavefront per queued pair ----------------
constexpr uint32_t kTriWaves = 4;          // wavefronts per block
constexpr uint32_t kLeafListCap = 512;     // ordered leaf pairs held in LDS per wavefront
constexpr uint32_t kListOverflow = 0xffffffffu;

struct TriWaveShared {
    uint2 list[kLeafListCap];              // (body leaf node, collider leaf node) in evaluation order
    float4 disc[64][2], swept[64][2];      // per lane: normal+pen, point+impact fraction
    Manifold man, local;
    uint32_t count;
    // The wavefront's triangle pair (one per wavefront, read by every lane):
    // held here rather than copied into each lane's private memory.
    alignas(16) unsigned char pair[sizeof(TrianglePair)];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ uint32_t wave_exclusive_scan(uint32_t v, uint32_t lane) {
    uint32_t x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= uint32_t(d)) x += y;
    }
    return x - v;
}

__device__ __forceinline__ bool leaf_cell_overlaps(const TrianglePair &tp, bool swept, const BoundsFrame &fa,
                                                   const BoundsFrame &fb, const BoundsFrame &pfa,
                                                   const BoundsFrame &pfb, uint64_t f, uint2 &nodes) {
    const uint32_t lb = tp.bm.s.leaf_count;
    nodes = make_uint2(tp.am.leaf(uint32_t(f / lb)), tp.bm.leaf(uint32_t(f % lb)));
    V3 amn, amx, bmn, bmx;
    node_bounds(tp.am.node(nodes.x), pfa, fa, swept, tp.margin, amn, amx);
    node_bounds(tp.bm.node(nodes.y), pfb, fb, swept, 0.0f, bmn, bmx);
    return boxes_touch(amn, amx, bmn, bmx);
}

// The leaf pairs rigid_collide_meshes visits, in visiting order, or
// kListOverflow when its stack would overflow (the reference then runs
// every triangle pair) or the list does not fit.
__device__ uint32_t traversal_list(const TrianglePair &P, bool swept, uint2 *list) {
    const BoundsFrame fa = bounds_frame(P.a), fb = bounds_frame(P.b);
    const BoundsFrame pfa = swept ? bounds_frame(P.ap) : BoundsFrame{};
    const BoundsFrame pfb = swept ? bounds_frame(P.bp) : BoundsFrame{};
    V3 amn, amx, bmn, bmx;
    motion_bounds(sarr(P.am.s.mn), sarr(P.am.s.mx), pfa, fa, swept, P.margin, amn, amx);
    motion_bounds(sarr(P.bm.s.mn), sarr(P.bm.s.mx), pfb, fb, swept, 0.0f, bmn, bmx);
    if (!boxes_touch(amn, amx, bmn, bmx)) return 0;
    if (P.am.s.node_count == 0u || P.bm.s.node_count == 0u) return kListOverflow;
    constexpr uint32_t kStack = 64;
    uint32_t sa[kStack], sb[kStack];
    sa[0] = 0;
    sb[0] = 0;
    uint32_t top = 1, count = 0;
    while (top > 0u) {
        --top;
        const uint32_t ia = sa[top], ib = sb[top];
        const BvhDev na = P.am.node(ia), nb = P.bm.node(ib);
        node_bounds(na, pfa, fa, swept, P.margin, amn, amx);
        node_bounds(nb, pfb, fb, swept, 0.0f, bmn, bmx);
        if (!boxes_touch(amn, amx, bmn, bmx)) continue;
        const bool la = na.count != 0u, lb = nb.count != 0u;
        if (la && lb) {
            if (count == kLeafListCap) return kListOverflow;
            list[count++] = make_uint2(ia, ib);
            continue;
        }
        const uint32_t need = la || lb ? 2u : 4u;
        if (top + need > kStack) return kListOverflow;
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
    return count;
}

// Serial evaluation for the rare pairs whose leaf list does not fit in
// LDS: the same candidate order and merge rule on one lane.  Leaf
// candidates come from the motion test alone (`swept`); the ranges of a
// guided pair always take the swept pass (`swept_eval`).
__device__ void triangle_pair_serial(const TrianglePair &tp, bool swept, bool swept_eval, bool leaf_mode,
                                     Manifold &man, Manifold &local) {
    man.count = 0;
    man.face_patch = 0;
    if (!leaf_mode) {
        collide_meshes(tp, swept_eval, man);
        return;
    }
    const BoundsFrame fa = bounds_frame(tp.a), fb = bounds_frame(tp.b);
    const BoundsFrame pfa = swept ? bounds_frame(tp.ap) : BoundsFrame{};
    const BoundsFrame pfb = swept ? bounds_frame(tp.bp) : BoundsFrame{};
    const uint64_t cells = uint64_t(tp.am.s.leaf_count) * tp.bm.s.leaf_count;
    const float spacing = smax(tp.margin * 2.0f, 1.0e-4f);
    for (uint32_t w = 0; w < 128u && w < cells; ++w)
        for (uint64_t f = w; f < cells; f += 128u) {
            uint2 nodes;
            if (!leaf_cell_overlaps(tp, swept, fa, fb, pfa, pfb, f, nodes)) continue;
            local.count = 0;
            local.face_patch = 0;
            collide_leaf_ranges(tp, swept_eval, tp.am.node(nodes.x), tp.bm.node(nodes.y), local);
            for (uint32_t k = 0; k < local.count; ++k) merge_contact(tp, man, local.c[k], spacing);
        }
}

// The triangle-pair path for the queued pairs, one wavefront per pair.
// A pair with a candidate slot (active index below `slots`) and at most
// `limit` overlapping leaf pairs takes its candidates in the order 128
// workers emit them (worker w takes cells w, w + 128, ... of the row-major
// leaf-pair grid) and merges each candidate's own manifold in that order;
// any other pair takes the leaf pairs of the BVH traversal in visiting
// order and merges contacts directly.  The lanes evaluate up to 4 leaf
// pairs x 16 triangle pairs at once (leaves hold at most 4 triangles); one
// lane then replays the contacts through the merge rule in order: per leaf
// pair the discrete contacts body-triangle-major, then the swept ones.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_triangle_pairs(const uint32_t *tri_list, const PmhBodyParameters *P,
                                                        const ShapeDev *shapes, const V3 *verts, const uint32_t *idx,
                                                        const float4 *planes, const BvhDev *nodes,
                                                        const uint32_t *leaves, const float4 *shell,
                                                        const PmhRigidBodyState *prev,
                                                        const PmhRigidBodyState *pred, const uint2 *ab,
                                                        NarrowOut o, uint32_t slots, uint32_t limit, float dt,
                                                        SceneRules scenes, const HingeFrame *frames,
                                                        const float4 *frame_ref) {
    __shared__ TriWaveShared shared_waves[kTriWaves];
    const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    TriWaveShared &S = shared_waves[wv];
    const uint32_t T = o.ctr->tri_pairs;
    const uint32_t nwaves = gridDim.x * kTriWaves;
    uint32_t fallback = 0;
    for (uint32_t t = blockIdx.x * kTriWaves + wv; t < T; t += nwaves) {
        const uint32_t m = tri_list[t];
        const uint2 p = ab[m];
        const PmhBodyParameters pa = P[p.x], pb = P[p.y];
        TrianglePair &shared_pair = *reinterpret_cast<TrianglePair *>(S.pair);
        wave_sync();   // every lane is done with the previous pair
        if (lane == 0)
            shared_pair = TrianglePair{ShapeView{shapes[pa.mesh_index], verts, idx, planes, nodes, leaves, shell},
                                       ShapeView{shapes[pb.mesh_index], verts, idx, planes, nodes, leaves, shell},
                                       prev[p.x], pred[p.x], prev[p.y], pred[p.y],
                                       pair_frames(frames, frame_ref, pa, pb, p, prev[p.x], pred[p.x]),
                                       pa.collision_margin + pb.collision_margin, dt};
        wave_sync();
        const TrianglePair &tp = shared_pair;
        // Leaf candidates from the motion test alone; the ranges of a guided
        // pair always take the swept pass (geometry.slang:2350, 2574-2579,
        // 2167-2170), and its discrete pass finds nothing (:1371).
        const bool guided = guided_pair(tp.f.ha, tp.f.hb);
        const bool swept = requires_swept(tp.ap, tp.a, tp.am.s, tp.bp, tp.b, tp.bm.s, tp.margin);
        const bool swept_eval = guided || swept;
        const BoundsFrame fa = bounds_frame(tp.a), fb = bounds_frame(tp.b);
        const BoundsFrame pfa = swept ? bounds_frame(tp.ap) : BoundsFrame{};
        const BoundsFrame pfb = swept ? bounds_frame(tp.bp) : BoundsFrame{};
        bool inside_a, inside_b;
        inside_flags(tp, inside_a, inside_b);
        const uint64_t cells = uint64_t(tp.am.s.leaf_count) * tp.bm.s.leaf_count;
        // 1. Leaf pairs in evaluation order.  In a batch the slot and
        // limit rules are those of the pair's own scene.
        bool leaf_mode;
        if (scenes.scene) {
            const uint32_t sc = scenes.scene[p.x];
            const uint64_t ns = scenes.count[sc];
            const uint32_t local = m - scenes.pair_start[sc];
            leaf_mode = local < leaf_slot_rule(ns);
            limit = ns <= 8u ? 4096u : 512u;
        } else {
            leaf_mode = m < slots;
        }
        // Each lane counts the cells f = lane (mod 128) and f = lane + 64
        // (mod 128) apart, so the same counts give the total and the list
        // offsets.
        uint32_t total = 0, c0 = 0, c1 = 0;
        if (leaf_mode) {
            uint2 nd;
            for (uint64_t f = lane; f < cells; f += 128u) c0 += leaf_cell_overlaps(tp, swept, fa, fb, pfa, pfb, f, nd) ? 1u : 0u;
            for (uint64_t f = lane + 64u; f < cells; f += 128u) c1 += leaf_cell_overlaps(tp, swept, fa, fb, pfa, pfb, f, nd) ? 1u : 0u;
            total = wave_sum(c0 + c1);
            leaf_mode = total <= limit;
        }
        uint32_t nlist = 0;
        if (leaf_mode) {
            if (total > kLeafListCap) {
                nlist = kListOverflow;
            } else {
                uint2 nd;
                uint32_t at0 = wave_exclusive_scan(c0, lane);
                uint32_t at1 = wave_sum(c0) + wave_exclusive_scan(c1, lane);
                for (uint64_t f = lane; f < cells; f += 128u)
                    if (leaf_cell_overlaps(tp, swept, fa, fb, pfa, pfb, f, nd)) S.list[at0++] = nd;
                for (uint64_t f = lane + 64u; f < cells; f += 128u)
                    if (leaf_cell_overlaps(tp, swept, fa, fb, pfa, pfb, f, nd)) S.list[at1++] = nd;
                nlist = total;
            }
        } else {
            if (lane == 0) {
                ++fallback;
                S.count = traversal_list(tp, swept_eval, S.list);
            }
            wave_sync();
            nlist = S.count;
        }
        wave_sync();
        // 2. Contacts.
        if (nlist == kListOverflow) {
            if (lane == 0) triangle_pair_serial(tp, swept, swept_eval, leaf_mode, S.man, S.local);
        } else {
            if (lane == 0) {
                S.man.count = 0;
                S.man.face_patch = 0;
            }
            const float spacing = smax(tp.margin * 2.0f, 1.0e-4f);
            for (uint32_t c0 = 0; c0 < nlist; c0 += 4u) {
                const uint32_t j = c0 + lane / 16u, q = lane % 16u;
                bool hd = false, hs = false;
                Contact cd, cs;
                if (j < nlist) {
                    const uint2 e = S.list[j];
                    const BvhDev na = tp.am.node(e.x), nb = tp.bm.node(e.y);
                    const uint32_t ai = q / 4u, bi = q % 4u;
                    if (ai < na.count && bi < nb.count) {
                        const uint32_t at = na.first + ai, bt = nb.first + bi;
                        if (3u * at + 2u < tp.am.s.idx_count && 3u * bt + 2u < tp.bm.s.idx_count) {
                            hd = !guided && triangle_pair_contact(tp, at, bt, inside_a, inside_b, cd);
                            if (swept_eval) hs = triangle_pair_swept_contact(tp, at, bt, inside_a, inside_b, cs);
                        }
                    }
                }
                if (hd) {
                    S.disc[lane][0] = f4(cd.normal, cd.pen);
                    S.disc[lane][1] = f4(cd.point, cd.impact_fraction);
                }
                if (hs) {
                    S.swept[lane][0] = f4(cs.normal, cs.pen);
                    S.swept[lane][1] = f4(cs.point, cs.impact_fraction);
                }
                const uint64_t md = __ballot(hd), ms = __ballot(hs);
                wave_sync();
                if (lane == 0) {
                    for (uint32_t g = 0; g < 4u && c0 + g < nlist; ++g) {
                        Manifold &dst = leaf_mode ? S.local : S.man;
                        if (leaf_mode) {
                            S.local.count = 0;
                            S.local.face_patch = 0;
                        }
                        for (uint32_t k = 16u * g; k < 16u * g + 16u; ++k) {
                            if (!((md >> k) & 1ull)) continue;
                            const Contact c{v3(S.disc[k][0]), v3(S.disc[k][1]), S.disc[k][0].w, S.disc[k][1].w};
                            merge_contact(tp, dst, c, spacing);
                        }
                        for (uint32_t k = 16u * g; k < 16u * g + 16u; ++k) {
                            if (!((ms >> k) & 1ull)) continue;
                            const Contact c{v3(S.swept[k][0]), v3(S.swept[k][1]), S.swept[k][0].w, S.swept[k][1].w};
                            merge_contact(tp, dst, c, spacing);
                        }
                        if (leaf_mode)
                            for (uint32_t k = 0; k < S.local.count; ++k) merge_contact(tp, S.man, S.local.c[k], spacing);
                    }
                }
                wave_sync();
            }
        }
        // 3. Emit.
        if (lane == 0) {
            Manifold man = S.man;
            emit_manifold(o, m, man, tp.a, tp.b, persistent_pair(tp.am.s, tp.bm.s, pb), tp.f.ha, tp.f.hb);
        }
        wave_sync();
    }
    if (lane == 0 && fallback) atomicAdd(&o.ctr->fallback_pairs, fallback);
}

__global__ void k_nonempty(const uint32_t *npairs, const uint32_t *m_count, uint32_t *flag) {
    const uint32_t M = *npairs;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x)
        flag[m] = m_count[m] != 0u ? 1u : 0u;
}

// An overflowing step solves nothing (its manifolds are incomplete) and is
// reported as PM_ERROR_CAPACITY_EXCEEDED.
__global__ void k_guard_points(StepCounters *ctr, uint32_t point_cap) {
    if (threadIdx.x != 0) return;
    const uint32_t P = ctr->point_alloc;
    if (P > point_cap) {
        atomicOr(&ctr->status, kStatusPointOverflow);
        ctr->points = 0;
        ctr->pairs = 0;
    } else {
        ctr->points = P;
    }
}

// pm_load_rigid_contact_cache (geometry.slang:2941-3009) against the
// previous step's saved pairs, found by binary search on their sorted codes.
// Cache key of a pair: (a, b) as one 64-bit number, ordered like the
// reference's pair code a*n+b (ascending a, then b) for any body count.
__device__ __forceinline__ unsigned long long pair_key(uint2 p) {
    return ((unsigned long long)p.x << 32) | p.y;
}

__global__ void k_cache_load(const uint32_t *npairs, const uint32_t *code, const uint2 *ab, const uint32_t *m_count,
                             const uint32_t *m_start, const PmhRigidBodyState *pred, const unsigned long long *saved_code,
                             const uint32_t *saved_count, const float4 *saved, const uint32_t *n_saved_ptr,
                             float4 *c_np, float4 *c_pt, float4 *c_aa, float4 *c_ab, float4 *c_lam,
                             float4 *c_pen, float4 *c_acc, uint32_t *c_flags, float dt, StepCounters *ctr) {
    const uint32_t M = *npairs;
    const uint32_t ns = *n_saved_ptr;
    uint32_t hits = 0;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        if (m_count[m] == 0) continue;
        const unsigned long long key = pair_key(ab[m]);
        uint32_t lo = 0, hi = ns;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (saved_code[mid] < key) lo = mid + 1; else hi = mid;
        }
        if (lo >= ns || saved_code[lo] != key) continue;
        const uint32_t sc = saved_count[lo];
        const float4 *rec = saved + (size_t)lo * 8 * kCacheFloat4PerPoint;
        const PmhRigidBodyState body = pred[ab[m].x];
        uint32_t used = 0;
        for (uint32_t k = 0; k < m_count[m]; ++k) {
            const uint32_t c = m_start[m] + k;
            const V3 n = v3(c_np[c]), point = v3(c_pt[c]);
            const V3 local = inv_rotate(sq(body.orientation), sub(point, sv(body.position)));
            float nearest = 0.02f * 0.02f;
            uint32_t match = 8;
            for (uint32_t q = 0; q < sc; ++q) {
                const float4 *r = rec + q * kCacheFloat4PerPoint;
                if ((used & (1u << q)) != 0u || dot(n, v3(r[1])) < 0.99f) continue;
                const float d = len2(sub(local, v3(r[0])));
                if (d < nearest) { nearest = d; match = q; }
            }
            if (match == 8u) continue;
            used |= 1u << match;
            const float4 *r = rec + match * kCacheFloat4PerPoint;
            // r0: local point, normal impulse; r1: normal; r2: friction
            // impulse; r3: lambdas; r4: penalties; r5: anchor a; anchor b in
            // r1.w/r2.w/r4.w
            V3 t0, t1;
            contact_basis(n, t0, t1);
            const V3 fi = v3(r[2]);
            const float4 lam = r[3], pen = r[4];
            c_lam[c] = make_float4(lam.x, dot(fi, t0) / dt, dot(fi, t1) / dt, 0.0f);
            c_pen[c] = make_float4(pen.x, pen.y, pen.z, 0.0f);
            c_aa[c] = f4(v3(r[5]), c_aa[c].w);
            c_ab[c] = make_float4(r[1].w, r[2].w, r[4].w, 0.0f);
            c_flags[c] |= 1u;
            const V3 af = sub(fi, mul(n, dot(fi, n)));
            c_acc[c] = make_float4(r[0].w, af.x, af.y, af.z);
            ++hits;
        }
    }
    // One atomic per wavefront (one per thread queued ~10^5 atomics on one
    // address).
    hits = wave_sum(hits);
    if ((threadIdx.x & 63) == 0 && hits) atomicAdd(&ctr->cache_hits, hits);
}

// pm_save_rigid_contact_cache (geometry.slang:3011-3070): the pair of
// response slot s (its rank among non-empty manifolds) is saved at s.
__global__ void k_cache_save(const uint32_t *npairs, const uint32_t *code, const uint2 *ab, const uint32_t *m_count,
                             const uint32_t *m_start, const uint32_t *m_slot, const PmhRigidBodyState *states,
                             const float4 *c_np, const float4 *c_pt, const float4 *c_aa, const float4 *c_ab,
                             const float4 *c_lam, const float4 *c_pen, const float4 *c_acc, unsigned long long *saved_code,
                             uint32_t *saved_count, float4 *saved, uint32_t *n_saved, uint32_t cap,
                             const StepCounters *ctr) {
    const uint32_t M = *npairs;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        if (m_count[m] == 0) continue;
        const uint32_t s = m_slot[m];
        if (s >= cap) continue;
        saved_code[s] = pair_key(ab[m]);
        saved_count[s] = m_count[m];
        const PmhRigidBodyState body = states[ab[m].x];
        float4 *rec = saved + (size_t)s * 8 * kCacheFloat4PerPoint;
        for (uint32_t k = 0; k < m_count[m]; ++k) {
            const uint32_t c = m_start[m] + k;
            const V3 local = inv_rotate(sq(body.orientation), sub(v3(c_pt[c]), sv(body.position)));
            const float4 acc = c_acc[c], bb = c_ab[c];
            float4 *r = rec + k * kCacheFloat4PerPoint;
            r[0] = f4(local, acc.x);
            r[1] = f4(v3(c_np[c]), bb.x);
            r[2] = make_float4(acc.y, acc.z, acc.w, bb.y);
            r[3] = c_lam[c];
            r[4] = make_float4(c_pen[c].x, c_pen[c].y, c_pen[c].z, bb.z);
            r[5] = f4(v3(c_aa[c]), 0.0f);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *n_saved = std::min(ctr->manifolds, cap);
}

// A step whose dataflow sweep failed keeps its states: the pre-step
// states replace the result.
__global__ void k_clamp_speeds(uint32_t n, const PmhBodyParameters *P, PmhRigidBodyState *s,
                               const PmhRigidBodyState *before, const StepCounters *ctr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (ctr->color_abort != 0u) {
        s[i] = before[i];
        return;
    }
    if (P[i].motion != kDynamic) return;
    s[i].linear_velocity = pv(clamp_length(sv(s[i].linear_velocity), P[i].maximum_linear_speed));
    s[i].angular_velocity = pv(clamp_length(sv(s[i].angular_velocity), P[i].maximum_angular_speed));
}

__global__ void k_step_begin(StepCounters *ctr) {
    if (threadIdx.x != 0) return;
    ctr->status = 0;   // per substep: the host checks it before the solve and after the last substep
    ctr->color_abort = 0;
    ctr->pairs = 0;
    ctr->points = 0;
    ctr->manifolds = 0;
    ctr->colors = 0;
    ctr->islands = 0;
    ctr->impact = 0;
    ctr->unsupported = 0;
    ctr->cache_hits = 0;
    ctr->large_count = 0;
    ctr->color_ticket = 0;
    ctr->point_alloc = 0;
    ctr->tri_pairs = 0;
    ctr->face_pairs = 0;
    ctr->fallback_pairs = 0;
    ctr->cell_size_bits = 0.0f;
}

static uint32_t blocks_for(uint32_t items, uint32_t block = 256, uint32_t cap = 4096) {
    uint32_t g = (items + block - 1) / block;
    return std::max<uint32_t>(1, std::min(g, cap));
}

PmhResult World::step_once(uint32_t k_step, bool warm_infer) {
    hipStream_t s = stream;
    const uint32_t n = body_count;
    Solver &v = solver;
    StepCounters *ctr = v.counters.p;
    // Grids of one thread per body or joint are never capped: those kernels
    // take one item per thread.
    const uint32_t bg = blocks_for(n, 256, 0xffffffffu);
    const uint32_t pg = blocks_for(v.pair_cap, 256, 2048);
    auto mark = [&](int k) { if (timing_enabled) (void)hipEventRecord(events[k], s); };
    mark(0);
    k_step_begin<<<1, 64, 0, s>>>(ctr);
    PmhRigidBodyState *cur = state_cur.p, *prev = state_prev.p, *pred = state_next.p;
    k_integrate_proxy<<<bg, 256, 0, s>>>(n, params.p, shapes_dev.p, cur, prev, pred, proxy_mn.p, proxy_mx.p,
                                         gravity, dt, inputs_dev.p, targets_dev.p, target_flags_dev.p,
                                         substeps - substep, substep == 0 ? 1u : 0u,
                                         substep + 1 == substeps ? 1u : 0u, ctr);
    mark(1);
    JointFilter jf;
    SceneRules rules;
    if (scene_first.size() > 1) {
        jf.scene = scene_of.p;
        rules.scene = scene_of.p;
        rules.count = scene_count.p;
        rules.pair_start = scene_pair_start.p;
    }
    if (solver.nj && weld_joints) {
        k_weld_init<<<bg, 256, 0, s>>>(n, weld.p);
        k_weld_hook<<<blocks_for(solver.nj, 256, 0xffffffffu), 256, 0, s>>>(n, solver.nj, joints_dev.p, weld.p);
        k_weld_compress<<<bg, 256, 0, s>>>(n, weld.p);
        jf.weld = weld.p;
    }
    if (solver.nj && excl_count) {
        jf.excl_start = excl_start.p;
        jf.excl_list = excl_list.p;
        jf.joints = joints_dev.p;
    }
    // Broadphase.
    PMH_CHECK(hipMemsetAsync(bucket_count.p, 0, sizeof(uint32_t) * (table_size + 1), s));
    k_bp_cells<<<bg, 256, 0, s>>>(n, proxy_mn.p, proxy_mx.p, cell_of.p, bucket_count.p, table_size - 1,
                                  large.p, ctr, grid_cell);
    PmhResult r = exclusive_scan(s, bucket_count.p, bucket_start.p, table_size, nullptr, nullptr, v.scan_tmp);
    if (r != PMH_SUCCESS) return r;
    PMH_CHECK(hipMemsetAsync(bucket_count.p, 0, sizeof(uint32_t) * table_size, s));
    k_bp_scatter<<<bg, 256, 0, s>>>(n, params.p, proxy_mn.p, proxy_mx.p, cell_of.p, bucket_start.p, bucket_count.p,
                                    entry_cell.p, entry_mn.p, entry_mx.p);
    if (jf.scene) {
        const uint32_t ns = uint32_t(scene_first.size());
        PMH_CHECK(hipMemsetAsync(scene_large_count.p, 0, sizeof(uint32_t) * (ns + 1), s));
        k_scene_large<<<64, 256, 0, s>>>(ctr, large.p, scene_of.p, scene_large_count.p, nullptr, nullptr);
        r = exclusive_scan(s, scene_large_count.p, scene_large_start.p, ns, nullptr, nullptr, v.scan_tmp);
        if (r != PMH_SUCCESS) return r;
        PMH_CHECK(hipMemsetAsync(scene_large_count.p, 0, sizeof(uint32_t) * (ns + 1), s));
        k_scene_large<<<64, 256, 0, s>>>(ctr, large.p, scene_of.p, scene_large_count.p, scene_large_start.p,
                                         scene_large.p);
        jf.scene_large_start = scene_large_start.p;
        jf.scene_large = scene_large.p;
        jf.scene_first = scene_first_dev.p;
        jf.scene_count = scene_count.p;
    }
    const uint32_t query_blocks = uint32_t((uint64_t(n) * kQueryLanes + 255) / 256);
    k_bp_query<0><<<query_blocks, 256, 0, s>>>(n, params.p, proxy_mn.p, proxy_mx.p, cell_of.p,
                                                  bucket_start.p, entry_cell.p, entry_mn.p, entry_mx.p, table_size - 1, large.p,
                                                  cand_count.p, nullptr, nullptr, nullptr, ctr, jf);
    k_bp_large<0><<<kLargeBlocks, 256, 0, s>>>(n, params.p, proxy_mn.p, proxy_mx.p, large.p, cand_count.p, nullptr, nullptr,
                                     nullptr, ctr, jf);
    r = exclusive_scan(s, cand_count.p, cand_offset.p, n, nullptr, nullptr, v.scan_tmp);
    if (r != PMH_SUCCESS) return r;
    k_set_pairs<<<1, 64, 0, s>>>(ctr, cand_offset.p, n, v.pair_cap);
    k_bp_large<1><<<kLargeBlocks, 256, 0, s>>>(n, params.p, proxy_mn.p, proxy_mx.p, large.p, cand_count.p, cand_offset.p,
                                     v.code.p, v.ab.p, ctr, jf);
    k_bp_query<1><<<query_blocks, 256, 0, s>>>(n, params.p, proxy_mn.p, proxy_mx.p, cell_of.p,
                                                  bucket_start.p, entry_cell.p, entry_mn.p, entry_mx.p, table_size - 1, large.p,
                                                  cand_count.p, cand_offset.p, v.code.p, v.ab.p, ctr, jf);
    mark(2);
    if (rules.scene) k_scene_pairs<<<pg, 256, 0, s>>>(&ctr->pairs, v.ab.p, scene_of.p, scene_pair_start.p);
    // Narrowphase + point compaction.
    const NarrowOut out{v.m_count.p, v.m_start.p, v.m_face.p, v.m_init_rel.p, v.c_np.p, v.c_pt.p, v.c_aa.p,
                        v.c_ab.p, v.c_err.p, v.c_lam.p, v.c_pen.p, v.c_acc.p, v.c_rep.p, v.c_flags.p,
                        v.c_pair.p, v.point_cap, ctr};
    // Hinge contact frames of every body (worlds with joints only).
    HingeFrame *frames = nullptr;
    float4 *frame_ref = nullptr;
    if (v.nj) {
        if ((r = hinge_frames.reserve(n)) != PMH_SUCCESS || (r = hinge_ref.reserve(2 * size_t(n))) != PMH_SUCCESS)
            return r;
        frames = hinge_frames.p;
        frame_ref = hinge_ref.p;
        JointView jv;
        jv.joints = v.joints;
        jv.list_start = v.jlist_start;
        jv.list = v.jlist;
        k_hinge_frames<<<bg, 256, 0, s>>>(n, params.p, prev, pred, jv, frames, frame_ref);
    }
    k_narrowphase<<<blocks_for(v.pair_cap, 128, 4096), 128, 0, s>>>(&ctr->pairs, params.p, shapes_dev.p, prev, pred,
                                                                    v.ab.p, out, v.face_list.p, frames, frame_ref);
    k_face_pairs<<<blocks_for(v.pair_cap, 128, 4096), 128, 0, s>>>(
        v.face_list.p, params.p, shapes_dev.p, verts_dev.p, idx_dev.p, planes_dev.p, prev, pred, v.ab.p, out,
        v.tri_list.p, frames, frame_ref);
    k_triangle_pairs<<<2048, 256, 0, s>>>(
        v.tri_list.p, params.p, shapes_dev.p, verts_dev.p, idx_dev.p, planes_dev.p, nodes_dev.p, leaves_dev.p,
        shell_dev.p, prev, pred, v.ab.p, out, leaf_slots, leaf_limit, dt, rules, frames, frame_ref);
    k_guard_points<<<1, 64, 0, s>>>(ctr, v.point_cap);
    mark(3);
    // Response slots and the contact cache.
    k_nonempty<<<pg, 256, 0, s>>>(&ctr->pairs, v.m_count.p, v.m_impact.p);
    r = exclusive_scan(s, v.m_impact.p, v.m_slot.p, v.pair_cap, &ctr->pairs, nullptr, v.scan_tmp);
    if (r != PMH_SUCCESS) return r;
    const int rd = cache_parity, wr = cache_parity ^ 1;
    k_cache_load<<<pg, 256, 0, s>>>(&ctr->pairs, v.code.p, v.ab.p, v.m_count.p, v.m_start.p, pred,
                                    cache_code[rd].p, cache_count[rd].p, cache_data[rd].p, &cache_n.p[rd],
                                    v.c_np.p, v.c_pt.p, v.c_aa.p, v.c_ab.p, v.c_lam.p, v.c_pen.p, v.c_acc.p,
                                    v.c_flags.p, dt, ctr);
    mark(4);
    SolveParams sp{dt, gravity, base_iterations, requested_iterations, profile, cooperative ? 1u : 0u, solve_lanes,
                   residual_tolerance, warm_infer && warm_enabled, color_wait_limit};
    sp.sweep_wait_limit = sweep_wait_limit;
    sp.dataflow = !per_color_sweep;
    SolveTiming st;
    st.kernels = &kernel_timer;
    if (timing_enabled) {
        st.start = solve_events[2 * k_step];
        st.stop = solve_events[2 * k_step + 1];
    }
    r = avbd_prepare_solve(s, v, params.p, prev, pred, sp, timing_enabled ? &st : nullptr);
    solve_timing.launches += st.launches;
    if (r != PMH_SUCCESS) return r;
    mark(5);
    k_clamp_speeds<<<bg, 256, 0, s>>>(n, params.p, pred, cur, ctr);
    k_cache_save<<<pg, 256, 0, s>>>(&ctr->pairs, v.code.p, v.ab.p, v.m_count.p, v.m_start.p, v.m_slot.p, pred,
                                    v.c_np.p, v.c_pt.p, v.c_aa.p, v.c_ab.p, v.c_lam.p, v.c_pen.p, v.c_acc.p,
                                    cache_code[wr].p, cache_count[wr].p, cache_data[wr].p, &cache_n.p[wr],
                                    cache_capacity, ctr);
    mark(6);
    cache_parity = wr;
    std::swap(state_cur.p, state_next.p);
    substep = (substep + 1) % substeps;
    PMH_CHECK(hipGetLastError());
    return PMH_SUCCESS;
}

}  // namespace pmh
