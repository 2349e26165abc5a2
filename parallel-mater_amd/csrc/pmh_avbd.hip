// SPDX-License-Identifier: MIT
// AVBD prepare and solve on gfx950.
//
// Prepare (prepare_avbd_kernel, geometry_constraints.slang:1213-1591) runs
// as data-parallel kernels: per-body inertia and targets, per-contact
// anchors/errors/duals, adjacency lists in the reference's list order
// (reverse active order), islands by lock-free hooking onto the smaller
// root (same min-index roots as the serial union-find), and first-fit
// colouring in index order as a dataflow over index-ordered tickets: a
// body takes the smallest colour its lower-index neighbours do not use,
// exactly the reference's serial greedy result.
//
// Solve (solve_avbd_kernel, geometry_constraints_solve.slang:771-865) is
// one cooperative launch: budget x (colours + dual pass) separated by grid
// barriers.  Each body of the active colour is owned by a group of L lanes
// of a wavefront; the lanes gather the body's contact rows from a
// colour-ordered CSR list, each accumulating a partial 6x6 block, and the
// partial blocks meet in an xor-butterfly over the group.  The block is
// then solved in registers by an unrolled LDL^T.  Per-point Jacobians are
// computed once per phase (they only depend on the target poses).
#include "pmh_internal.hpp"
#include "pmh_warmstart.hpp"

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>

namespace cg = cooperative_groups;

namespace pmh {

constexpr uint32_t kRealtimeProfile = 1u;
constexpr uint32_t kRealtimeSweeps = 8u;
constexpr float kRealtimeDamping = 0.05f;
constexpr float kSurfaceTolerance = 1.0e-5f;
constexpr float kRestOffset = 1.0e-3f;
constexpr uint32_t kSideFirst = 0x80000000u;
// Colour-ordered entry, second word: the other body, and this flag when it
// is dynamic and of a lower colour (the dataflow sweep waits for it).
constexpr uint32_t kEntryLower = 0x80000000u;
constexpr uint32_t kEntryBody = 0x7fffffffu;

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Agent-coherent 64-bit words: relaxed agent-scope atomics, issued as
// global_load/global_store with the sc1 bit.  The dataflow sweep hands
// displacements between the wavefronts of one launch with them: inside a
// kernel the L2 caches of the eight XCDs are not coherent with each other
// for ordinary accesses.
__device__ __forceinline__ uint64_t ld64_co(const float *p) {
    return __hip_atomic_load(reinterpret_cast<const uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st64_co(float *p, float x, float y) {
    const uint64_t v = uint64_t(__float_as_uint(x)) | (uint64_t(__float_as_uint(y)) << 32);
    __hip_atomic_store(reinterpret_cast<uint64_t *>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float lo_f(uint64_t v) { return __uint_as_float(uint32_t(v)); }
__device__ __forceinline__ float hi_f(uint64_t v) { return __uint_as_float(uint32_t(v >> 32)); }
__device__ __forceinline__ float4 ld4_co(const float4 *p) {
    const uint64_t a = ld64_co(&p->x), b = ld64_co(&p->z);
    return make_float4(lo_f(a), hi_f(a), lo_f(b), hi_f(b));
}
__device__ __forceinline__ void st4_co(float4 *p, float4 v) {
    st64_co(&p->x, v.x, v.y);
    st64_co(&p->z, v.z, v.w);
}

// Sum over a converged wavefront (every lane active).
__device__ __forceinline__ uint32_t wave_total(uint32_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Adds each lane's value to one counter with one atomic per wavefront
// (every lane active).  Same-address atomics from every lane serialise in
// one L2 channel; the counters here are read only after the kernel.
__device__ __forceinline__ void wave_counter_add(uint32_t *counter, uint32_t v) {
    const uint32_t t = wave_total(v);
    if ((threadIdx.x & 63) == 0 && t) atomicAdd(counter, t);
}

// Adds each thread's K values to K counters with one atomic per counter and
// block (the block's wavefronts meet in LDS; blocks of at most 1024
// threads, every thread calls it).  Atomics on one address serialise: one
// per wavefront of a 670k-pair grid kept such a kernel at 0.2 ms.
template <int K>
__device__ __forceinline__ void block_counters_add(uint32_t *const (&counter)[K], const uint32_t (&v)[K]) {
    __shared__ uint32_t part[K][16];
    const uint32_t wave = threadIdx.x >> 6, waves = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t t = wave_total(v[k]);
        if ((threadIdx.x & 63) == 0) part[k][wave] = t;
    }
    __syncthreads();
    if (threadIdx.x < uint32_t(K)) {
        uint32_t sum = 0;
        for (uint32_t w = 0; w < waves; ++w) sum += part[threadIdx.x][w];
        if (sum) atomicAdd(counter[threadIdx.x], sum);
    }
}

__device__ __forceinline__ V3 v3(float4 f) { return {f.x, f.y, f.z}; }
__device__ __forceinline__ Q q4(float4 f) { return {f.x, f.y, f.z, f.w}; }
__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ float4 f4q(Q q) { return make_float4(q.x, q.y, q.z, q.w); }

__device__ __forceinline__ V3 sv(const PmhVec3 &v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ Q sq(const PmhQuat &q) { return {q.x, q.y, q.z, q.w}; }
__device__ __forceinline__ PmhVec3 pv(V3 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ PmhQuat pq(Q q) { return {q.x, q.y, q.z, q.w}; }

__device__ __forceinline__ float pair_stiffness(const PmhBodyParameters &a, const PmhBodyParameters &b, float dt) {
    return 1.0f / (smax(a.inverse_mass + b.inverse_mass, 1.0e-6f) * dt * dt);
}
// rigid_avbd_contact_skin (geometry_constraints.slang:438-445): a guided
// contact (point flag 8, set by the narrowphase) rests at 0.1 mm.
constexpr float kGuidedRestOffset = 1.0e-4f;
constexpr uint32_t kPointGuided = 8u;
__device__ __forceinline__ float contact_skin(const PmhBodyParameters &a, const PmhBodyParameters &b, bool guided) {
    return smin(a.collision_margin + b.collision_margin, guided ? kGuidedRestOffset : kRestOffset) +
           2.0f * kSurfaceTolerance;
}

// ---------------------------------------------------------------------
// Device-wide exclusive scan (three passes, 1024 elements per block).
// ---------------------------------------------------------------------
constexpr uint32_t kScanBlock = 256, kScanItems = 4, kScanTile = kScanBlock * kScanItems;

__device__ uint32_t block_exclusive(uint32_t v, uint32_t *total) {
    __shared__ uint32_t warp_sums[kScanBlock / 64];
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) warp_sums[wid] = x;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (uint32_t w = 0; w < kScanBlock / 64; ++w) {
        if (w < wid) base += warp_sums[w];
        all += warp_sums[w];
    }
    __syncthreads();
    if (total) *total = all;
    return base + x - v;
}

__global__ void k_scan_reduce(const uint32_t *in, uint32_t count, const uint32_t *count_ptr, uint32_t *sums) {
    const uint32_t n = count_ptr ? *count_ptr : count;
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (uint32_t k = 0; k < kScanItems; ++k) {
        uint32_t i = base + k * kScanBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t total;
    block_exclusive(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void k_scan_sums(uint32_t *sums, uint32_t blocks, uint32_t *total_out) {
    // One block scans up to blocks <= 1024*? sums sequentially in tiles.
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t t = 0; t < blocks; t += kScanBlock) {
        uint32_t i = t + threadIdx.x;
        uint32_t v = i < blocks ? sums[i] : 0;
        uint32_t total;
        uint32_t ex = block_exclusive(v, &total);
        if (i < blocks) sums[i] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ void k_scan_apply(const uint32_t *in, uint32_t *out, uint32_t count, const uint32_t *count_ptr,
                             const uint32_t *sums) {
    const uint32_t n = count_ptr ? *count_ptr : count;
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems], s = 0;
    for (uint32_t k = 0; k < kScanItems; ++k) {
        uint32_t i = base + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    uint32_t total;
    uint32_t ex = block_exclusive(s, &total) + sums[blockIdx.x];
    for (uint32_t k = 0; k < kScanItems; ++k) {
        uint32_t i = base + k;
        if (i < n) out[i] = ex;
        ex += v[k];
    }
}

__global__ void k_scan_tail(const uint32_t *in, uint32_t *out, uint32_t count, const uint32_t *count_ptr) {
    const uint32_t n = count_ptr ? *count_ptr : count;
    out[n] = n ? out[n - 1] + in[n - 1] : 0;
}

PmhResult exclusive_scan(hipStream_t s, const uint32_t *in, uint32_t *out, uint32_t count,
                         const uint32_t *count_ptr, uint32_t *total, DBuf<uint32_t> &tmp) {
    // `count` is the capacity bound when count_ptr is given.
    const uint32_t blocks = std::max<uint32_t>(1, (count + kScanTile - 1) / kScanTile);
    if (tmp.reserve(blocks) != PMH_SUCCESS) return PMH_ERROR_OUT_OF_DEVICE_MEMORY;
    k_scan_reduce<<<blocks, kScanBlock, 0, s>>>(in, count, count_ptr, tmp.p);
    k_scan_sums<<<1, kScanBlock, 0, s>>>(tmp.p, blocks, total);
    k_scan_apply<<<blocks, kScanBlock, 0, s>>>(in, out, count, count_ptr, tmp.p);
    k_scan_tail<<<1, 1, 0, s>>>(in, out, count, count_ptr);
    PMH_CHECK(hipGetLastError());
    return PMH_SUCCESS;
}

// ---------------------------------------------------------------------
// Prepare: bodies (geometry_constraints.slang:1235-1295).
// ---------------------------------------------------------------------
__global__ void k_body_init(uint32_t n, const PmhBodyParameters *P, const PmhRigidBodyState *prev,
                            const PmhRigidBodyState *st, float dt, V3 g, float4 *tpos, float4 *tq,
                            float4 *tv, float4 *tw, float4 *disp, float4 *dq, float4 *inertia,
                            float4 *prev_vel, uint32_t *color, uint32_t *parent, uint32_t *impact,
                            uint32_t *deg, uint32_t *anchored, uint32_t *adj_cursor, uint32_t *members,
                            float4 *mean, uint32_t *jdeg, uint32_t *jcursor) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    members[i] = 0;
    jdeg[i] = 0;
    jcursor[i] = 0;
    mean[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const PmhRigidBodyState s = st[i];
    tpos[i] = f4(sv(s.position), 0.0f);
    tq[i] = f4q(sq(s.orientation));
    tv[i] = f4(sv(s.linear_velocity), 0.0f);
    tw[i] = f4(sv(s.angular_velocity), 0.0f);
    dq[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    color[i] = kInvalid;
    parent[i] = i;
    impact[i] = 0;
    deg[i] = 0;
    anchored[i] = 0;
    adj_cursor[i] = 0;
    float4 old = prev_vel[i];
    if (P[i].motion != kDynamic) {
        disp[2 * i] = make_float4(0, 0, 0, 0);
        disp[2 * i + 1] = make_float4(0, 0, 0, 0);
        inertia[2 * i] = make_float4(0, 0, 0, 0);
        inertia[2 * i + 1] = make_float4(0, 0, 0, 0);
        prev_vel[i] = make_float4(0, 0, 0, 0);
        return;
    }
    const bool hist = old.w != 0.0f;
    const float g2 = len2(g);
    const V3 pv_ = sv(prev[i].linear_velocity);
    const V3 acc = mul(sub(pv_, v3(old)), 1.0f / dt);
    float w = hist && g2 > 0.0f ? clampf(dot(acc, g) / g2, 0.0f, 1.0f) : 0.0f;
    if (!isfinite(w)) w = 0.0f;
    const V3 corr = mul(g, -(1.0f - w) * dt * dt);
    disp[2 * i] = f4(corr, 0.0f);
    disp[2 * i + 1] = make_float4(0, 0, 0, 0);
    prev_vel[i] = f4(pv_, 1.0f);
    const float mass = 1.0f / smax(P[i].inverse_mass, 1.0e-12f);
    const Q o = sq(s.orientation);
    float rot[6];   // lower triangle of the rotational block: 33 43 44 53 54 55
    for (uint32_t ax = 0; ax < 3; ++ax) {
        V3 la = inv_rotate(o, basis_axis(ax));
        V3 col = rotate(o, {la.x / smax(P[i].inverse_inertia_local.x, 1.0e-12f),
                            la.y / smax(P[i].inverse_inertia_local.y, 1.0e-12f),
                            la.z / smax(P[i].inverse_inertia_local.z, 1.0e-12f)});
        for (uint32_t r = ax; r < 3; ++r) {
            // element (3+r, 3+ax) of the packed lower triangle
            const uint32_t k = r * (r + 1) / 2 + ax;
            rot[k] = comp(col, r) / (dt * dt);
        }
    }
    inertia[2 * i] = make_float4(mass / (dt * dt), rot[0], rot[1], rot[2]);
    inertia[2 * i + 1] = make_float4(rot[3], rot[4], rot[5], 0.0f);
}

// Contacts (geometry_constraints.slang:1297-1375) and the per-manifold
// impact flag (rigid_avbd_prepare_contact_impact, :983-1005).
__global__ void k_contact_init(const uint32_t *npairs, const PmhBodyParameters *P, const PmhRigidBodyState *prev,
                               const PmhRigidBodyState *st, const uint2 *ab, const uint32_t *m_count,
                               const uint32_t *m_start, float4 *c_np, float4 *c_pt, float4 *c_aa, float4 *c_ab,
                               float4 *c_err, float4 *c_lam, float4 *c_pen, const uint32_t *c_flags,
                               uint32_t *m_impact, uint32_t *c_entry, float dt, uint32_t profile) {
    const uint32_t M = *npairs;
    const bool realtime = profile == kRealtimeProfile;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint2 p = ab[m];
        const PmhBodyParameters pa = P[p.x], pb = P[p.y];
        const float penalty = pair_stiffness(pa, pb, dt);
        const PmhRigidBodyState A0 = prev[p.x], B0 = prev[p.y], A = st[p.x], B = st[p.y];
        bool impact = false;
        const uint32_t start = m_start[m], count = m_count[m];
        for (uint32_t k = 0; k < count; ++k) {
            const uint32_t c = start + k;
            // Colour-ordered slots of the point's endpoints: written once
            // the bodies are placed by colour; a body the colouring gave up
            // on has none.
            c_entry[2 * c] = c_entry[2 * c + 1] = kInvalid;
            const bool cached = (c_flags[c] & 1u) != 0u;
            const float skin = contact_skin(pa, pb, (c_flags[c] & kPointGuided) != 0u);
            float4 np = c_np[c], pt = c_pt[c];
            const V3 normal = v3(np), point = v3(pt);
            V3 aa = v3(c_aa[c]), bb = v3(c_ab[c]);
            if (!cached) {
                aa = inv_rotate(sq(A0.orientation), sub(point, sv(A0.position)));
                bb = inv_rotate(sq(B0.orientation), sub(point, sv(B0.position)));
            }
            const V3 ma = sub(sv(A.position), sv(A0.position));
            const V3 mb = sub(sv(B.position), sv(B0.position));
            const V3 wa = mul(quaternion_delta_velocity(sq(A0.orientation), sq(A.orientation), dt), dt);
            const V3 wb = mul(quaternion_delta_velocity(sq(B0.orientation), sq(B.orientation), dt), dt);
            const V3 armA = sub(point, sv(A.position)), armB = sub(point, sv(B.position));
            const float closing = dot(normal, sub(add(mb, cross(wb, armB)), add(ma, cross(wa, armA))));
            const float old_pen = np.w - closing;
            V3 t0, t1;
            contact_basis(normal, t0, t1);
            const V3 motion = add(sub(mb, ma), sub(rotation_displacement(sq(B0.orientation), sq(B.orientation), bb),
                                                   rotation_displacement(sq(A0.orientation), sq(A.orientation), aa)));
            const V3 old_delta = !cached ? V3{0.0f, 0.0f, 0.0f}
                : add(sub(sv(B0.position), sv(A0.position)),
                      sub(rotate(sq(B0.orientation), bb), rotate(sq(A0.orientation), aa)));
            const V3 terr = add(motion, mul(old_delta, 1.0f - kAlpha));
            const float speed = -closing / dt;
            const float impact_disp = realtime && !cached && speed < -0.1f
                ? -(1.0f + smin(pa.restitution, pb.restitution)) * speed * dt : 0.0f;
            const V3 err = {np.w - clampf(old_pen, 0.0f, skin) + impact_disp, dot(terr, t0), dot(terr, t1)};
            float4 lam = c_lam[c], pen = c_pen[c];
            float l[3] = {lam.x, lam.y, lam.z}, k3[3] = {pen.x, pen.y, pen.z};
            for (int ax = 0; ax < 3; ++ax) {
                if (!cached) { l[ax] = 0.0f; k3[ax] = penalty; }
                warm_start(l[ax], k3[ax], kMaxPenalty);
            }
            c_lam[c] = make_float4(l[0], l[1], l[2], 0.0f);
            c_pen[c] = make_float4(k3[0], k3[1], k3[2], 0.0f);
            c_aa[c] = f4(aa, c_aa[c].w);
            c_ab[c] = f4(bb, 0.0f);
            c_err[c] = f4(err, 0.0f);
            c_pt[c] = f4(point, speed);
            // Impact test on the updated contact.
            const float impact_fraction = c_aa[c].w;
            const bool fresh = (!cached || l[0] <= 1.0e-6f || (impact_fraction > 0.0f && impact_fraction < 1.0f)) &&
                               speed < -0.1f && np.w >= -kSurfaceTolerance;
            const float prev_pen = np.w + speed * dt;
            impact = impact || fresh || prev_pen > skin + 1.0e-4f;
        }
        m_impact[m] = impact ? 1u : 0u;
    }
}

// ---------------------------------------------------------------------
// Graph: adjacency (dynamic endpoints only), islands, impact, colouring.
// ---------------------------------------------------------------------
__global__ void k_degree(const uint32_t *npairs, const PmhBodyParameters *P, const uint2 *ab,
                         const uint32_t *m_count, uint32_t *deg, uint32_t *anchored_edge, StepCounters *ctr) {
    const uint32_t M = *npairs;
    uint32_t local = 0, m1 = 0, m2 = 0, p1 = 0, p2 = 0;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        if (m_count[m] == 0) continue;
        ++local;
        const uint2 p = ab[m];
        const bool da = P[p.x].motion == kDynamic, db = P[p.y].motion == kDynamic;
        if (da) atomicAdd(&deg[p.x], 1u);
        if (db) atomicAdd(&deg[p.y], 1u);
        if (da && db) { ++m2; p2 += m_count[m]; } else { ++m1; p1 += m_count[m]; }
    }
    uint32_t *const counters[5] = {&ctr->manifolds, &ctr->m1, &ctr->m2, &ctr->p1, &ctr->p2};
    const uint32_t values[5] = {local, m1, m2, p1, p2};
    block_counters_add<5>(counters, values);
    (void)anchored_edge;
}

__global__ void k_adj_fill(const uint32_t *npairs, const PmhBodyParameters *P, const uint2 *ab,
                           const uint32_t *m_count, const uint32_t *adj_start, uint32_t *cursor, uint32_t *adj) {
    const uint32_t M = *npairs;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        if (m_count[m] == 0) continue;
        const uint2 p = ab[m];
        if (P[p.x].motion == kDynamic) adj[adj_start[p.x] + atomicAdd(&cursor[p.x], 1u)] = m;
        if (P[p.y].motion == kDynamic) adj[adj_start[p.y] + atomicAdd(&cursor[p.y], 1u)] = m;
    }
}

// Reference list order: contacts are pushed to the front in active order,
// so each body's list is in descending active index.
__global__ void k_adj_sort(uint32_t n, const uint32_t *adj_start, uint32_t *adj) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = adj_start[i], e = adj_start[i + 1];
    for (uint32_t k = b + 1; k < e; ++k) {
        uint32_t v = adj[k];
        uint32_t j = k;
        while (j > b && adj[j - 1] < v) { adj[j] = adj[j - 1]; --j; }
        adj[j] = v;
    }
}

__device__ uint32_t uf_find(const uint32_t *parent, uint32_t v) {
    uint32_t p = ld_relaxed(&parent[v]);
    while (p != v) { v = p; p = ld_relaxed(&parent[v]); }
    return v;
}

// Joins the trees of u and v: the larger root is hooked under the smaller
// (parents only decrease), retrying when another hook moved a root first.
__device__ void uf_union(uint32_t *parent, uint32_t u, uint32_t v) {
    u = uf_find(parent, u);
    v = uf_find(parent, v);
    while (u != v) {
        if (u > v) { const uint32_t t = u; u = v; v = t; }
        const uint32_t old = atomicCAS(&parent[v], v, u);
        if (old == v) break;
        u = uf_find(parent, u);
        v = uf_find(parent, old);
    }
}

// First hooking pass over one edge per body: each dynamic body joins its
// smallest-index dynamic neighbour.  After a compression the trees are
// flat and cover most of every island, so the pass over all edges
// (k_uf_hook) mostly finds both ends under one root at once.  The final
// roots do not depend on the order of the hooks: each island's root is
// its smallest index.
__global__ void k_uf_sample(uint32_t n, const PmhBodyParameters *P, const uint2 *ab, const uint32_t *adj_start,
                            const uint32_t *adj, uint32_t *parent) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || P[i].motion != kDynamic) return;
    uint32_t best = kInvalid;
    for (uint32_t e = adj_start[i]; e < adj_start[i + 1]; ++e) {
        const uint2 p = ab[adj[e]];
        const uint32_t j = p.x == i ? p.y : p.x;
        if (j < best && P[j].motion == kDynamic) best = j;
    }
    if (best < i) uf_union(parent, i, best);
}

// Hook the larger root under the smaller one (parents only decrease), so
// every component's final root is its minimum index, as the serial
// union-find of geometry_constraints.slang:1178-1189 produces.
__global__ void k_uf_hook(const uint32_t *npairs, const PmhBodyParameters *P, const uint2 *ab,
                          const uint32_t *m_count, uint32_t *parent) {
    const uint32_t M = *npairs;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        if (m_count[m] == 0) continue;
        const uint2 p = ab[m];
        if (P[p.x].motion != kDynamic || P[p.y].motion != kDynamic) continue;
        // After the sampled hooks and a compression most pairs already
        // share a parent: nothing to join, and no walk to the root.
        if (ld_relaxed(&parent[p.x]) == ld_relaxed(&parent[p.y])) continue;
        uf_union(parent, p.x, p.y);
    }
}

__global__ void k_uf_compress(uint32_t n, uint32_t *parent) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t r = uf_find(parent, i);
    st_relaxed(&parent[i], r);
}

__global__ void k_impact_roots(const uint32_t *npairs, const PmhBodyParameters *P, const uint2 *ab,
                               const uint32_t *m_impact, const uint32_t *m_count, const uint32_t *parent,
                               uint32_t *impact, uint32_t *anchored, StepCounters *ctr) {
    const uint32_t M = *npairs;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint2 p = ab[m];
        const bool da = P[p.x].motion == kDynamic, db = P[p.y].motion == kDynamic;
        // Flags are read before the atomic: most islands are flagged by
        // their first pairs, and a pile's one large island would otherwise
        // take one same-address atomic per pair.
        if (m_impact[m]) {
            if (da && ld_relaxed(&impact[parent[p.x]]) == 0u) atomicOr(&impact[parent[p.x]], 1u);
            if (db && ld_relaxed(&impact[parent[p.y]]) == 0u) atomicOr(&impact[parent[p.y]], 1u);
            ctr->impact = 1u;
        }
        // An island touching any non-dynamic body keeps its momentum
        // correction (free_mean's `anchored`, solve.slang:641-652).
        if (m_count[m] != 0) {
            if (da && !db && ld_relaxed(&anchored[parent[p.x]]) == 0u) atomicOr(&anchored[parent[p.x]], 1u);
            if (db && !da && ld_relaxed(&anchored[parent[p.y]]) == 0u) atomicOr(&anchored[parent[p.y]], 1u);
        }
    }
}

// Also counts each island's dynamic members and lists the roots whose
// islands take the free-motion mean correction (unanchored, with contacts).
__global__ void k_impact_bodies(uint32_t n, const PmhBodyParameters *P, const uint32_t *parent, uint32_t *impact,
                                const uint32_t *deg, const uint32_t *jdeg, const uint32_t *anchored,
                                uint32_t *members, uint32_t *mean_list, StepCounters *ctr) {
    // Every lane stays active to the end: the counters are wave-aggregated.
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    const bool dyn = i < n && P[i].motion == kDynamic;
    if (i < n && !dyn) impact[i] = 0;
    uint32_t r = kInvalid, flag = 0;
    bool root_listed = false, mean_root = false;
    if (dyn) {
        r = parent[i];
        flag = ld_relaxed(&impact[r]);
        if (r != i) impact[i] = flag;
        root_listed = r == i && (deg[i] != 0 || jdeg[i] != 0);
        mean_root = root_listed && anchored[i] == 0;
    }
    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t roots = __ballot(root_listed), means = __ballot(mean_root);
    const uint64_t dyns = __ballot(dyn), flagged = __ballot(dyn && flag != 0);
    uint32_t mean_base = 0;
    if (lane == 0) {
        if (roots) atomicAdd(&ctr->islands, uint32_t(__popcll(roots)));
        if (means) mean_base = atomicAdd(&ctr->mean_roots, uint32_t(__popcll(means)));
        if (dyns) atomicAdd(&ctr->dyn_bodies, uint32_t(__popcll(dyns)));
        if (flagged) atomicAdd(&ctr->impact_bodies, uint32_t(__popcll(flagged)));
    }
    mean_base = __shfl(mean_base, 0, 64);
    if (mean_root) mean_list[mean_base + __popcll(means & below)] = i;
    // Island member counts: lanes sharing the first dynamic lane's root
    // add together (a large island covers whole wavefronts).
    if (dyns) {
        const uint32_t lead = __builtin_ctzll(dyns);
        const uint32_t r0 = __shfl(r, lead, 64);
        const uint64_t same = __ballot(dyn && r == r0);
        if (lane == lead) atomicAdd(&members[r0], uint32_t(__popcll(same)));
        if (dyn && r != r0) atomicAdd(&members[r], 1u);
    }
}

// Velocity-phase work: non-empty manifolds touching an impact body.
__global__ void k_count_impact_pairs(const uint32_t *npairs, const PmhBodyParameters *P, const uint2 *ab,
                                     const uint32_t *m_count, const uint32_t *impact, StepCounters *ctr) {
    const uint32_t M = *npairs;
    uint32_t m1 = 0, m2 = 0, p1 = 0, p2 = 0;
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        if (m_count[m] == 0) continue;
        const uint2 p = ab[m];
        if (impact[p.x] == 0 && impact[p.y] == 0) continue;
        const bool da = P[p.x].motion == kDynamic, db = P[p.y].motion == kDynamic;
        if (da && db) { ++m2; p2 += m_count[m]; } else { ++m1; p1 += m_count[m]; }
    }
    uint32_t *const counters[4] = {&ctr->m1i, &ctr->m2i, &ctr->p1i, &ctr->p2i};
    const uint32_t values[4] = {m1, m2, p1, p2};
    block_counters_add<4>(counters, values);
}

// Algorithmic bytes of one AVBD iteration: per dynamic body 136 B
// (target pose, displacement, rotation delta and inertia read; displacement
// and rotation delta written); per contact point 84 B read by each dynamic
// endpoint in the primal and once in the dual, plus 24 B of dual written;
// per manifold 12 B of pair record plus 52 B of the other body per dynamic
// endpoint and 104 B for both bodies in the dual.
__device__ __forceinline__ double iteration_bytes(uint32_t d, uint32_t m1, uint32_t m2, uint32_t p1, uint32_t p2) {
    return 136.0 * d + (84.0 * 2 + 24.0) * p1 + (84.0 * 3 + 24.0) * p2 + (64.0 + 104.0) * m1 +
           (128.0 + 104.0) * m2;
}

// First-fit colouring in index order, as a dataflow: a body's colour is
// the smallest not taken by its lower-index neighbours (contacts and
// joints), so any schedule that colours a body after those neighbours gives
// the sequential result.  Each lane takes one body ticket (a wavefront 64
// consecutive ones, issued in index order) and walks its lower-index
// neighbours with a cursor, waiting at the first one not yet coloured;
// the wavefront loops until all its lanes have published.  Every lower
// ticket is held by a resident wavefront whose lanes progress on their own,
// so the waits end.  Every wait is bounded all the same: a wait that runs
// past `limit` polls raises ctr->color_abort and the timeout status, every
// other wait gives up as soon as it sees the flag, and a body that gave up
// publishes kColorAborted, so no body waits on it.  The whole colouring then
// ends within about one limit's worth of polls, and the step returns an
// error before its solve.
constexpr float kLaneColorNeighbours = 6.0f;   // below this many neighbours per body: per-lane tickets

__device__ __forceinline__ bool color_wait(const uint32_t *color, uint32_t j, uint64_t &u0, uint64_t &u1,
                                           uint64_t &u2, uint64_t &u3) {
    const uint32_t c = ld_relaxed(&color[j]);
    if (c == kInvalid) return false;
    if (c < kMaxColors) {
        const uint64_t bit = 1ull << (c & 63u);
        const uint32_t w = c >> 6;
        u0 |= w == 0u ? bit : 0ull;
        u1 |= w == 1u ? bit : 0ull;
        u2 |= w == 2u ? bit : 0ull;
        u3 |= w == 3u ? bit : 0ull;
    }
    return true;
}

__global__ void __launch_bounds__(256) k_color_lane(uint32_t n, const PmhBodyParameters *P, const uint2 *ab,
                                               const uint32_t *adj_start, const uint32_t *adj, uint32_t *color,
                                               const uint32_t *jstart, const uint32_t *jadj,
                                               const PmhRigidConstraint *joints, StepCounters *ctr,
                                               uint32_t limit) {
    const uint32_t lane = threadIdx.x & 63;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&ctr->color_ticket, 64u);
        base = __shfl(base, 0, 64);
        if (base >= n) break;
        const uint32_t i = base + lane;
        bool pending = i < n && P[i].motion == kDynamic;
        uint32_t k = 0, ke = 0, jk = 0, je = 0;
        if (pending) {
            k = adj_start[i];
            ke = adj_start[i + 1];
            if (joints) {
                jk = jstart[i];
                je = jstart[i + 1];
            }
        }
        uint64_t u0 = 0, u1 = 0, u2 = 0, u3 = 0;
        uint32_t top = 0;   // largest colour published by this lane, plus one
        for (uint32_t polls = 0;; ++polls) {
            if (pending) {
                bool blocked = false;
                for (; k < ke; ++k) {
                    const uint2 p = ab[adj[k]];
                    const uint32_t j = p.x == i ? p.y : p.x;
                    if (j >= i || P[j].motion != kDynamic) continue;
                    if (!color_wait(color, j, u0, u1, u2, u3)) { blocked = true; break; }
                }
                if (!blocked)
                    for (; jk < je; ++jk) {
                        const PmhRigidConstraint &jt = joints[jadj[jk]];
                        const uint32_t j = jt.dense_a == i ? jt.dense_b : jt.dense_a;
                        if (j >= i || P[j].motion != kDynamic) continue;
                        if (!color_wait(color, j, u0, u1, u2, u3)) { blocked = true; break; }
                    }
                if (!blocked) {
                    uint32_t chosen = ~u0 ? __builtin_ctzll(~u0)
                                    : ~u1 ? 64u + __builtin_ctzll(~u1)
                                    : ~u2 ? 128u + __builtin_ctzll(~u2)
                                    : ~u3 ? 192u + __builtin_ctzll(~u3) : kMaxColors;
                    if (chosen >= kMaxColors) {
                        atomicOr(&ctr->status, kStatusColorOverflow);
                        chosen = kMaxColors - 1;
                    }
                    st_relaxed(&color[i], chosen);
                    top = chosen + 1;
                    pending = false;
                }
            }
            if (__ballot(pending) == 0ull) break;
            if (polls >= limit && lane == 0) st_relaxed(&ctr->color_abort, 1u);
            // The abort flag every 256th poll only (one hot address).
            if (polls >= limit || ((polls & 255u) == 255u && ld_relaxed(&ctr->color_abort) != 0u)) {
                if (pending) {
                    st_relaxed(&color[i], kColorAborted);
                    atomicOr(&ctr->status, kStatusColorTimeout);
                }
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) top = max(top, uint32_t(__shfl_xor(top, d, 64)));
        if (lane == 0 && top) atomicMax(&ctr->colors, top);
    }
}

// The same colouring with one wavefront per body, its lanes sharing the
// body's neighbour scan: the faster form when bodies have many neighbours.
// Blocks take tickets of 16 consecutive bodies in index order (one atomic
// per block: same-address atomics serialise), so every lower-index
// neighbour was taken earlier by a resident wavefront and eventually
// publishes its colour: the wait chain always ends.
constexpr uint32_t kColorWaveBlock = 16;   // wavefronts (bodies) per block ticket
__device__ __forceinline__ uint32_t color_wait_bounded(const uint32_t *color, uint32_t j, StepCounters *ctr,
                                                       uint32_t limit) {
    uint32_t c, polls = 0;
    while ((c = ld_relaxed(&color[j])) == kInvalid) {
        if (++polls > limit) {
            st_relaxed(&ctr->color_abort, 1u);
            break;
        }
        // The abort flag every 256th poll only (one hot address).
        if ((polls & 255u) == 0u && ld_relaxed(&ctr->color_abort) != 0u) break;
        __builtin_amdgcn_s_sleep(1);
    }
    return c;
}

__global__ void __launch_bounds__(64 * kColorWaveBlock) k_color_wave(uint32_t n, const PmhBodyParameters *P,
                                                                   const uint2 *ab, const uint32_t *adj_start,
                                                                   const uint32_t *adj, uint32_t *color,
                                                                   const uint32_t *jstart, const uint32_t *jadj,
                                                                   const PmhRigidConstraint *joints,
                                                                   StepCounters *ctr, uint32_t limit) {
    __shared__ uint32_t s_ticket[2];
    __shared__ uint32_t s_top[kColorWaveBlock];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // The colour count (largest colour + 1) is gathered per wavefront and
    // added with one atomic per block at the end: one atomic per body on
    // that one address queued behind each other.
    uint32_t top = 0;
    for (uint32_t round = 0;; ++round) {
        // Double-buffered: the slot written in round r + 1 was last read
        // before every wavefront passed round r's barrier.
        if (threadIdx.x == 0) s_ticket[round & 1u] = atomicAdd(&ctr->color_ticket, 1u);
        __syncthreads();
        const uint32_t first = s_ticket[round & 1u] * kColorWaveBlock;
        if (first >= n) break;               // the same for the whole block
        const uint32_t i = first + wave;
        if (i >= n || P[i].motion != kDynamic) continue;
        uint64_t used[kMaxColors / 64] = {0, 0, 0, 0};
        bool gave_up = false;
        const uint32_t b = adj_start[i], e = adj_start[i + 1];
        for (uint32_t k = b + lane; k < e; k += 64) {
            const uint2 p = ab[adj[k]];
            const uint32_t j = p.x == i ? p.y : p.x;
            if (j >= i || P[j].motion != kDynamic) continue;
            const uint32_t c = color_wait_bounded(color, j, ctr, limit);
            if (c == kInvalid) { gave_up = true; break; }
            if (c < kMaxColors) used[c >> 6] |= 1ull << (c & 63);
        }
        if (joints && !gave_up) {
            for (uint32_t k = jstart[i] + lane; k < jstart[i + 1]; k += 64) {
                const PmhRigidConstraint &jt = joints[jadj[k]];
                const uint32_t j = jt.dense_a == i ? jt.dense_b : jt.dense_a;
                if (j >= i || P[j].motion != kDynamic) continue;
                const uint32_t c = color_wait_bounded(color, j, ctr, limit);
                if (c == kInvalid) { gave_up = true; break; }
                if (c < kMaxColors) used[c >> 6] |= 1ull << (c & 63);
            }
        }
        const bool abandoned = __ballot(gave_up) != 0ull;
        uint32_t chosen = kMaxColors;
#pragma unroll
        for (uint32_t w = 0; w < kMaxColors / 64; ++w) {
            uint64_t m = used[w];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) m |= __shfl_xor(m, o, 64);
            if (chosen == kMaxColors && m != ~0ull) chosen = w * 64 + __builtin_ctzll(~m);
        }
        if (lane == 0 && abandoned) {
            st_relaxed(&color[i], kColorAborted);
            atomicOr(&ctr->status, kStatusColorTimeout);
        } else if (lane == 0) {
            if (chosen >= kMaxColors) { atomicOr(&ctr->status, kStatusColorOverflow); chosen = kMaxColors - 1; }
            st_relaxed(&color[i], chosen);
            top = max(top, chosen + 1);
        }
    }
    if (lane == 0) s_top[wave] = top;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = 0;
        for (uint32_t w = 0; w < kColorWaveBlock; ++w) m = max(m, s_top[w]);
        if (m) atomicMax(&ctr->colors, m);
    }
}

// Wave-aggregated histogram: one atomic per distinct colour per wavefront
// (a few colours, tens of thousands of bodies).
__global__ void k_color_count(uint32_t n, const uint32_t *color, uint32_t *count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t c = i < n ? color[i] : kInvalid;
    bool pending = c < kMaxColors;
    while (__ballot(pending) != 0ull) {
        const uint64_t live = __ballot(pending);
        const uint32_t leader = __builtin_ctzll(live);
        const uint32_t c0 = __shfl(c, leader, 64);
        const uint64_t same = __ballot(pending && c == c0);
        if (lane == leader) atomicAdd(&count[c0], uint32_t(__popcll(same)));
        if (pending && c == c0) pending = false;
    }
}

// Colour starts at multiples of `groups` (bodies per wavefront in the
// sweep), the gaps filled with kInvalid: a wavefront never holds two
// colours.
__global__ void k_color_scan(uint32_t *count, uint32_t *start, const StepCounters *ctr, uint32_t *bodies,
                             uint32_t groups) {
    if (threadIdx.x != 0) return;
    uint32_t s = 0;
    const uint32_t C = ctr->colors;
    for (uint32_t c = 0; c <= C && c <= kMaxColors; ++c) {
        start[c] = s;
        if (c < C) {
            const uint32_t m = count[c], padded = (m + groups - 1) / groups * groups;
            for (uint32_t k = s + m; k < s + padded; ++k) bodies[k] = kInvalid;
            s += padded;
            count[c] = 0;
        }
    }
}

// Order within a colour does not affect results (its bodies share no
// contact); one atomic per distinct colour per wavefront places them.
__global__ void k_color_place(uint32_t n, const uint32_t *color, const uint32_t *start, uint32_t *cursor,
                              uint32_t *bodies) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63;
    uint32_t c = i < n ? color[i] : kInvalid;
    bool pending = c < kMaxColors;
    while (__ballot(pending) != 0ull) {
        const uint64_t live = __ballot(pending);
        const uint32_t leader = __builtin_ctzll(live);
        const uint32_t c0 = __shfl(c, leader, 64);
        const uint64_t same = __ballot(pending && c == c0);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&cursor[c0], uint32_t(__popcll(same)));
        base = __shfl(base, leader, 64);
        if (pending && c == c0) {
            bodies[start[c0] + base + __popcll(same & ((1ull << lane) - 1ull))] = i;
            pending = false;
        }
    }
}

// Per colour position: the body's contact points in list order.
__global__ void k_bp_count(uint32_t n, const StepCounters *ctr, const uint32_t *start, const uint32_t *bodies,
                           const uint32_t *adj_start, const uint32_t *adj, const uint32_t *m_count,
                           uint32_t *bp_count) {
    const uint32_t D = start[ctr->colors];
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = k < D ? bodies[k] : kInvalid;
    if (i == kInvalid) { bp_count[k] = 0; return; }
    uint32_t s = 0;
    for (uint32_t e = adj_start[i]; e < adj_start[i + 1]; ++e) s += m_count[adj[e]];
    bp_count[k] = s;
}

__global__ void k_bp_fill(const StepCounters *ctr, const uint32_t *start, const uint32_t *bodies,
                          const uint32_t *adj_start, const uint32_t *adj, const uint32_t *m_count,
                          const uint32_t *m_start, const uint2 *ab, const uint32_t *bp_start, uint2 *entries,
                          uint32_t *c_entry, const uint32_t *color) {
    const uint32_t D = start[ctr->colors];
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= D) return;
    const uint32_t i = bodies[k];
    if (i == kInvalid) return;
    const uint32_t ci = color[i];
    uint32_t o = bp_start[k];
    for (uint32_t e = adj_start[i]; e < adj_start[i + 1]; ++e) {
        const uint32_t m = adj[e];
        const uint2 p = ab[m];
        const bool first = p.x == i;
        const uint32_t other = first ? p.y : p.x;
        // Static bodies hold no colour (kInvalid): never lower.
        const uint32_t tag = other | (color[other] < ci ? kEntryLower : 0u);
        for (uint32_t q = 0; q < m_count[m]; ++q) {
            c_entry[2 * (m_start[m] + q) + (first ? 0u : 1u)] = o;
            entries[o++] = make_uint2((m_start[m] + q) | (first ? kSideFirst : 0u), tag);
        }
    }
}

// Stores a point's nine row records in its Jacobian slot and in the
// colour-ordered slot of each dynamic endpoint.
__device__ __forceinline__ void store_rows(float4 *jac, float4 *erec, size_t ecap, const uint32_t *c_entry,
                                           uint32_t c, bool dyn_a, bool dyn_b, const float4 *J) {
    float4 *dst = jac + 9 * (size_t)c;
#pragma unroll
    for (int q = 0; q < 9; ++q) dst[q] = J[q];
    const uint32_t ea = dyn_a ? c_entry[2 * c] : kInvalid, eb = dyn_b ? c_entry[2 * c + 1] : kInvalid;
    if (ea != kInvalid) {
#pragma unroll
        for (int q = 0; q < 9; ++q) erec[q * ecap + ea] = J[q];
    }
    if (eb != kInvalid) {
#pragma unroll
        for (int q = 0; q < 9; ++q) erec[q * ecap + eb] = J[q];
    }
}

// Per-point rows: axes, -(r_a x axis), r_b x axis, initial error, the
// reference forces of the velocity phase, friction and dual stiffness
// (rigid_avbd_solve_contact_rows, geometry_constraints.slang:506-529).
__global__ void k_jacobians(const uint32_t *npts, const PmhBodyParameters *P, const uint32_t *c_pair,
                            const uint2 *ab, const float4 *tpos, const float4 *tq, const float4 *c_np,
                            const float4 *c_pt, const float4 *c_aa, const float4 *c_ab, const float4 *c_err,
                            const float4 *c_acc, float4 *jac, float dt, uint32_t phase, const uint32_t *impact,
                            float4 *erec, size_t ecap, const uint32_t *c_entry) {
    const uint32_t Pn = *npts;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < Pn; c += gridDim.x * blockDim.x) {
        const uint2 p = ab[c_pair[c]];
        if (phase != 0 && impact[p.x] == 0 && impact[p.y] == 0) continue;
        const V3 n = v3(c_np[c]), point = v3(c_pt[c]);
        V3 t0, t1;
        contact_basis(n, t0, t1);
        const V3 axes[3] = {n, t0, t1};
        const Q qa = q4(tq[p.x]), qb = q4(tq[p.y]);
        const V3 armA = sub(point, v3(tpos[p.x])), armB = sub(point, v3(tpos[p.y]));
        const V3 ra_t = rotate(qa, v3(c_aa[c])), rb_t = rotate(qb, v3(c_ab[c]));
        const float4 e = c_err[c];
        const float e0[3] = {e.x, e.y, e.z};
        float ref[3] = {0.0f, 0.0f, 0.0f};
        if (phase != 0) {
            const float4 acc = c_acc[c];
            ref[0] = acc.x / dt;
            ref[1] = dot(v3(make_float4(acc.y, acc.z, acc.w, 0.0f)), t0) / dt;
            ref[2] = dot(v3(make_float4(acc.y, acc.z, acc.w, 0.0f)), t1) / dt;
        }
        const float mu = sqrtf(P[p.x].friction * P[p.y].friction);
        const float beta = pair_stiffness(P[p.x], P[p.y], dt) * 100.0f;
        float4 J[9];
        for (uint32_t ax = 0; ax < 3; ++ax) {
            const V3 ra = ax == 0 ? armA : ra_t;
            const V3 rb = ax == 0 ? armB : rb_t;
            J[ax] = f4(axes[ax], e0[ax]);
            J[3 + ax] = f4(mul(cross(ra, axes[ax]), -1.0f), ref[ax]);
            J[6 + ax] = f4(cross(rb, axes[ax]), ax == 0 ? mu : (ax == 1 ? beta : 0.0f));
        }
        store_rows(jac, erec, ecap, c_entry, c, P[p.x].motion == kDynamic, P[p.y].motion == kDynamic, J);
    }
}

// ---------------------------------------------------------------------
// Solve kernel pieces.
// ---------------------------------------------------------------------
struct SolveArgs {
    uint32_t n;
    const PmhBodyParameters *P;
    PmhRigidBodyState *states;
    float4 *tpos, *tq, *tv, *tw, *disp, *dq;
    const float4 *inertia;
    const uint32_t *color_bodies, *color_start, *bp_start;
    const uint2 *entries;
    const float4 *jac;
    ResidualLog *log;                    // per-iteration constraint residual
    uint32_t *c_flags;
    float tol;                           // > 0: to-residual mode
    const float4 *erec;                  // contact rows in colour order (plane stride ecap)
    size_t ecap;
    const uint32_t *c_entry;
    float4 *c_lam, *c_pen, *c_aa, *c_ab, *c_acc, *c_rep, *c_err;
    const float4 *c_np, *c_pt;
    const uint32_t *c_pair;
    const uint2 *ab;
    const uint32_t *parent, *impact, *deg, *anchored;
    const uint32_t *members, *mean_list;
    float4 *mean;
    PmhContactSchedule *schedule;
    StepCounters *ctr;
    SolveTotals *totals;
    const uint32_t *npts, *npairs;
    float dt;
    uint32_t budget, realtime;
    // joints
    PmhRigidConstraint *joints;
    uint32_t nj;
    const uint32_t *jstart, *jadj;       // per-body incident joints, descending index
    const PmhRigidBodyState *prev;
    // dataflow sweep: colours, published stamps, this sweep's stamp (0: not
    // in a dataflow sweep) and the poll limit of its waits
    const uint32_t *color;
    uint32_t *done;
    uint32_t flow_stamp, flow_limit;
};

__device__ __forceinline__ void load_disp(const float4 *disp, uint32_t i, float *d) {
    const float4 a = disp[2 * i], b = disp[2 * i + 1];
    d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = b.x; d[4] = b.y; d[5] = b.z;
}
// Another body's displacement inside a dataflow sweep (agent-coherent).
__device__ __forceinline__ void load_disp_co(const float4 *disp, uint32_t i, float *d) {
    const float *p = &disp[2 * i].x;
    const uint64_t t = ld64_co(p), tz = ld64_co(p + 2), r = ld64_co(p + 4), rz = ld64_co(p + 6);
    d[0] = lo_f(t); d[1] = hi_f(t); d[2] = lo_f(tz); d[3] = lo_f(r); d[4] = hi_f(r); d[5] = lo_f(rz);
}
__device__ __forceinline__ void load_disp_a(const SolveArgs &a, uint32_t i, float *d) {
    if (a.flow_stamp) load_disp_co(a.disp, i, d);
    else load_disp(a.disp, i, d);
}

// Dataflow sweep: waits until body j has published this sweep's
// displacement.  Bounded: a wait that runs out raises the abort flag, which
// ends every other wait of the step, and marks the status, so the step
// fails instead of hanging.
// A wait that runs out also marks the sweep failed (kept until the host
// reports it): the step then keeps its states and fails.  The abort flag
// is read every 256th poll only: thousands of waiting lanes reading one
// address on every poll would queue on one memory channel, and the waits
// would slow down the bodies they wait for.
__device__ __forceinline__ void wait_published(const SolveArgs &a, uint32_t j) {
    for (uint32_t polls = 1; int32_t(ld_relaxed(&a.done[j]) - a.flow_stamp) < 0; ++polls) {
        if (polls > a.flow_limit) {
            st_relaxed(&a.ctr->color_abort, 1u);
            st_relaxed(&a.ctr->sweep_failed, 1u);
            atomicOr(&a.ctr->status, kStatusColorTimeout);
            break;
        }
        if ((polls & 255u) == 0u && ld_relaxed(&a.ctr->color_abort) != 0u) break;
        __builtin_amdgcn_s_sleep(2);
    }
    asm volatile("" ::: "memory");   // the displacement reads stay after the wait
}
// Dataflow sweep: body i's displacement stores (agent-coherent) have
// completed before the stamp that releases its waiters is stored.
__device__ __forceinline__ void publish_done(const SolveArgs &a, uint32_t i) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    st_relaxed(&a.done[i], a.flow_stamp);
}

// err = e0 + dot6(JA, dA) + dot6(JB, dB) with JA = (-axis, angA),
// JB = (axis, angB): the operation order of the reference.
__device__ __forceinline__ float row_error(float e0, float4 axis, float4 angA, float4 angB, const float *dA,
                                           const float *dB) {
    const float ja[6] = {-axis.x, -axis.y, -axis.z, angA.x, angA.y, angA.z};
    const float jb[6] = {axis.x, axis.y, axis.z, angB.x, angB.y, angB.z};
    return e0 + dot6(ja, dA) + dot6(jb, dB);
}

__device__ void body_joint_rows(const SolveArgs &a, uint32_t i, uint32_t phase, Block &blk, uint32_t sub_lane,
                                uint32_t L);

template <int L>
__device__ __forceinline__ void group_reduce(Block &b) {
#pragma unroll
    for (int s = 1; s < L; s <<= 1) {
#pragma unroll
        for (int k = 0; k < 21; ++k) b.h[k] = b.h[k] + __shfl_xor(b.h[k], s, 64);
#pragma unroll
        for (int k = 0; k < 6; ++k) b.g[k] = b.g[k] + __shfl_xor(b.g[k], s, 64);
    }
}

// apply_pose_update, geometry_constraints.slang:386-408: the translation
// always, the rotation delta and its quaternion only for a turn (returns
// whether it turned).
__device__ __forceinline__ bool pose_update(float4 &t, float4 &r, float4 &dqv, const float *u) {
    t.x += u[0]; t.y += u[1]; t.z += u[2];
    const V3 rot = {u[3], u[4], u[5]};
    if (rot.x == 0.0f && rot.y == 0.0f && rot.z == 0.0f) return false;
    Q rel = q4(dqv);
    const Q der = qmul({rot.x, rot.y, rot.z, 0.0f}, rel);
    rel = qnormalize({rel.x + 0.5f * der.x, rel.y + 0.5f * der.y, rel.z + 0.5f * der.z, rel.w + 0.5f * der.w});
    if (rel.w < 0.0f) rel = {-rel.x, -rel.y, -rel.z, -rel.w};
    dqv = f4q(rel);
    r.x = 2.0f * rel.x; r.y = 2.0f * rel.y; r.z = 2.0f * rel.z;
    return true;
}

// A lane's share of one update's residual: maximum, fixed-point sum of
// squares and count of its points or joints.
struct ResidualAcc {
    float max = 0.0f;
    unsigned long long sum = 0;
    uint32_t count = 0;
    __device__ void add(float r) { max = smax(max, r); sum += residual_square(r); ++count; }
};

// Combines the lanes of a converged wavefront into the log's slot with one
// atomic per field (non-negative floats order like their bit patterns).
__device__ __forceinline__ void residual_commit(ResidualLog *log, uint32_t slot, const ResidualAcc &acc) {
    uint32_t bits = __float_as_uint(acc.max), count = acc.count;
    unsigned long long sum = acc.sum;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        bits = max(bits, uint32_t(__shfl_xor(bits, d, 64)));
        count += uint32_t(__shfl_xor(count, d, 64));
        sum += (unsigned long long)__shfl_xor((long long)sum, d, 64);
    }
    if ((threadIdx.x & 63) == 0 && count) {
        const uint32_t bank = blockIdx.x & (kResidualBanks - 1);
        atomicMax(&log->bits[slot][bank], bits);
        atomicAdd(&log->sum[slot][bank], sum);
        atomicAdd(&log->count[slot][bank], count);
    }
}

// One contact point's dual update (advance_duals, solve.slang:492-566)
// from both ends' displacements and the point's rows; defined with
// advance_duals below.
__device__ void dual_point(const SolveArgs &a, uint32_t c, uint32_t phase, const float *dA, const float *dB,
                           const float4 *Jx, bool last, ResidualAcc &residual);

// The bodies at colour-ordered positions k = base + group of one wavefront
// (solve_color, geometry_constraints_solve.slang:331-491).  Each body is
// owned by L lanes: every lane streams every L-th of the body's contact
// points, the partial blocks meet in a butterfly, the LDLT solve and the
// crossing safeguard run on the whole group.
// kJoints: the world has joints.  Joint-free worlds run the false
// instance, which carries none of the joint code (registers, stack).
// kFlow: inside a dataflow sweep (one launch for every colour): a body
// waits for the lower-colour neighbours it reads, reads other bodies'
// displacements agent-coherently, and publishes its own with a stamp.
template <int L, bool kJoints, bool kFlow>
__device__ __forceinline__ void solve_group(const SolveArgs &a, uint32_t base, uint32_t end, uint32_t phase,
                                            uint32_t it = 0u) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t grp = lane / L, sub_lane = lane % L;
    const uint32_t k = base + grp;
    bool active = k < end;
    uint32_t i = 0;
    if (active) {
        i = a.color_bodies[k];
        if (i == kInvalid || (phase != 0 && a.impact[i] == 0)) active = false;
    }
    Block blk;
    block_zero(blk);
    float dI[6] = {0, 0, 0, 0, 0, 0};
    uint32_t eb = 0, ee = 0;
    if (active) {
        load_disp(a.disp, i, dI);
        eb = a.bp_start[k];
        ee = a.bp_start[k + 1];
        if (sub_lane == 0) {
            const float4 i0 = a.inertia[2 * i], i1 = a.inertia[2 * i + 1];
            // Full packed block of the body: diag(m/dt^2) and the
            // rotational lower triangle; the rest is zero.
            float h[21];
#pragma unroll
            for (int q = 0; q < 21; ++q) h[q] = 0.0f;
            h[0] = i0.x; h[2] = i0.x; h[5] = i0.x;
            h[9] = i0.y; h[13] = i0.z; h[14] = i0.w; h[18] = i1.x; h[19] = i1.y; h[20] = i1.z;
#pragma unroll
            for (int q = 0; q < 21; ++q) blk.h[q] = h[q];
#pragma unroll
            for (int r = 0; r < 6; ++r)
#pragma unroll
                for (int c = 0; c < 6; ++c)
                    blk.g[r] += h[r > c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r] * dI[c];
        }
    }
    // The crossing safeguard revisits the lane's rows: the first one's
    // values stay in registers (a lane holds one row in most bodies).
    float4 keep_axis = make_float4(0.0f, 0.0f, 0.0f, 0.0f), keep_angA = keep_axis, keep_angB = keep_axis;
    float keep_lam = 0.0f, keep_pen = 0.0f, keep_dO[6] = {0, 0, 0, 0, 0, 0};
    bool keep_first = false;
    for (uint32_t e = eb + sub_lane; e < ee; e += L) {
        const uint2 ent = a.entries[e];
        const uint32_t c = ent.x & 0x7fffffffu;
        const bool first = (ent.x & kSideFirst) != 0u;
        const uint32_t other = ent.y & kEntryBody;
        const float4 lam = a.c_lam[c], pen = a.c_pen[c];
        float4 Jx[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) Jx[q] = a.erec[q * a.ecap + e];
        float dO[6];
        if (kFlow) {
            if ((ent.y & kEntryLower) != 0u && (phase == 0 || a.impact[other] != 0u)) wait_published(a, other);
            load_disp_co(a.disp, other, dO);
        } else {
            load_disp(a.disp, other, dO);
        }
        if (e == eb + sub_lane) {
            keep_axis = Jx[0];
            keep_angA = Jx[3];
            keep_angB = Jx[6];
            keep_lam = lam.x;
            keep_pen = pen.x;
            keep_first = first;
#pragma unroll
            for (int q = 0; q < 6; ++q) keep_dO[q] = dO[q];
        }
        const float *dA = first ? dI : dO;
        const float *dB = first ? dO : dI;
        const float err0 = row_error(Jx[0].w, Jx[0], Jx[3], Jx[6], dA, dB);
        const float err1 = row_error(Jx[1].w, Jx[1], Jx[4], Jx[7], dA, dB);
        const float err2 = row_error(Jx[2].w, Jx[2], Jx[5], Jx[8], dA, dB);
        const Force3 trial = {lam.x + pen.x * err0, lam.y + pen.y * err1, lam.z + pen.z * err2};
        const float mu = Jx[6].w;
        const Force3 f = project_contact(trial, mu);
        const Force3 sc = stiffness_scales(trial, mu);
        const float stiff[3] = {pen.x * sc.n, pen.y * sc.t0, pen.z * sc.t1};
        const float force[3] = {f.n - Jx[3].w, f.t0 - Jx[4].w, f.t1 - Jx[5].w};
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const float4 axis = Jx[ax];
            const float4 ang = first ? Jx[3 + ax] : Jx[6 + ax];
            const float row[6] = {first ? -axis.x : axis.x, first ? -axis.y : axis.y,
                                  first ? -axis.z : axis.z, ang.x, ang.y, ang.z};
            add_row(blk, row, stiff[ax], force[ax]);
        }
    }
    // Joint rows follow the contact rows (solve.slang:401-430), each
    // joint in one lane's partial block.
    if (kJoints && active) body_joint_rows(a, i, phase, blk, sub_lane, L);
    group_reduce<L>(blk);
    float update[6];
    const bool solved = active && block_solve(blk, update);
    // Crossing safeguard (solve.slang:433-475).
    Block cross_blk;
    block_zero(cross_blk);
    bool crossing = false;
    if (solved) {
        for (uint32_t e = eb + sub_lane; e < ee; e += L) {
            float4 axis, angA, angB;
            float lam0, k0, dO[6];
            bool first;
            if (e == eb + sub_lane) {
                axis = keep_axis;
                first = keep_first;
                angA = keep_angA;
                angB = keep_angB;
                lam0 = keep_lam;
                k0 = keep_pen;
#pragma unroll
                for (int q = 0; q < 6; ++q) dO[q] = keep_dO[q];
            } else {
                const uint2 ent = a.entries[e];
                const uint32_t c = ent.x & 0x7fffffffu;
                first = (ent.x & kSideFirst) != 0u;
                const uint32_t other = ent.y & kEntryBody;
                if (kFlow) load_disp_co(a.disp, other, dO);
                else load_disp(a.disp, other, dO);
                axis = a.erec[e];
                angA = a.erec[3 * a.ecap + e];
                angB = a.erec[6 * a.ecap + e];
                lam0 = a.c_lam[c].x;
                k0 = a.c_pen[c].x;
            }
            const float *dA = first ? dI : dO;
            const float *dB = first ? dO : dI;
            const float err = row_error(axis.w, axis, angA, angB, dA, dB);
            const float4 ang = first ? angA : angB;
            const float row[6] = {first ? -axis.x : axis.x, first ? -axis.y : axis.y,
                                  first ? -axis.z : axis.z, ang.x, ang.y, ang.z};
            const float trial = lam0 + k0 * err;
            if (trial <= 0.0f && trial + k0 * dot6(row, update) > k0 * 1.0e-6f) {
                add_row(cross_blk, row, k0, trial);
                crossing = true;
            }
        }
    }
    // Group-wide "any crossing" (uniform across the group's lanes).
    uint64_t bal = __ballot(crossing);
    const uint64_t gmask = (L == 64 ? ~0ull : ((1ull << L) - 1ull)) << (grp * L);
    const bool group_crossing = (bal & gmask) != 0ull;
    if (bal != 0ull) {
        group_reduce<L>(cross_blk);
        if (group_crossing && solved) {
#pragma unroll
            for (int q = 0; q < 21; ++q) blk.h[q] += cross_blk.h[q];
#pragma unroll
            for (int q = 0; q < 6; ++q) blk.g[q] += cross_blk.g[q];
            float safe[6];
            if (block_solve(blk, safe))
#pragma unroll
                for (int q = 0; q < 6; ++q) update[q] = safe[q];
        }
    }
    // The body's displacement after this update (the group's first lane
    // writes it).
    float nd[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) nd[q] = dI[q];
    if (active && sub_lane == 0) {
        if (solved) {
            float4 t = a.disp[2 * i], r = a.disp[2 * i + 1];
            if (phase == 0) {
                float4 q = a.dq[i];
                const bool turned = pose_update(t, r, q, update);
                if (kFlow) {
                    st4_co(&a.disp[2 * i], t);
                    if (turned) {
                        st4_co(&a.dq[i], q);
                        st4_co(&a.disp[2 * i + 1], r);
                    }
                } else {
                    a.disp[2 * i] = t;
                    if (turned) {
                        a.dq[i] = q;
                        a.disp[2 * i + 1] = r;
                    }
                }
            } else {
                t.x += update[0]; t.y += update[1]; t.z += update[2];
                r.x += update[3]; r.y += update[4]; r.z += update[5];
                if (kFlow) {
                    st4_co(&a.disp[2 * i], t);
                    st4_co(&a.disp[2 * i + 1], r);
                } else {
                    a.disp[2 * i] = t;
                    a.disp[2 * i + 1] = r;
                }
            }
            nd[0] = t.x; nd[1] = t.y; nd[2] = t.z; nd[3] = r.x; nd[4] = r.y; nd[5] = r.z;
        }
        if (kFlow) publish_done(a, i);
    }
    if (kFlow) {
        // This update's duals (advance_duals) of the points whose other end
        // is solved before this body in the sweep (a lower colour) or not in
        // it (not dynamic, or not an impact body in the velocity phase):
        // both ends' displacements are final here, and the points' duals are
        // read again only by the next update's launch.  Every point has one
        // such end, so each is updated once, as the separate dual pass did.
#pragma unroll
        for (int q = 0; q < 6; ++q) nd[q] = __shfl(nd[q], int(grp * L), 64);
        const bool last = a.tol <= 0.0f && it + 1u == a.budget;
        ResidualAcc residual;
        if (active) {
            for (uint32_t e = eb + sub_lane; e < ee; e += L) {
                const uint2 ent = a.entries[e];
                const uint32_t other = ent.y & kEntryBody;
                const bool other_solved = a.P[other].motion == kDynamic && (phase == 0 || a.impact[other] != 0u);
                if (other_solved && (ent.y & kEntryLower) == 0u) continue;
                const uint32_t c = ent.x & 0x7fffffffu;
                const bool first = (ent.x & kSideFirst) != 0u;
                float dO[6];
                load_disp_co(a.disp, other, dO);
                float4 Jx[9];
#pragma unroll
                for (int q = 0; q < 9; ++q) Jx[q] = a.erec[q * a.ecap + e];
                dual_point(a, c, phase, first ? nd : dO, first ? dO : nd, Jx, last, residual);
            }
        }
        if (it < kResidualSlots) residual_commit(a.log, phase * kResidualSlots + it, residual);
    }
}

// One colour: the positions [begin, end), `gw`/`nw` indexing the wavefront
// within the launch.
template <int L, bool kJoints>
__device__ void solve_color(const SolveArgs &a, uint32_t begin, uint32_t end, uint32_t phase, uint32_t gw,
                            uint32_t nw) {
    constexpr uint32_t groups = 64 / L;
    for (uint32_t base = begin + gw * groups; base < end; base += nw * groups)
        solve_group<L, kJoints, false>(a, base, end, phase);
}

__device__ __forceinline__ PmhRigidBodyState avbd_pose(const SolveArgs &a, uint32_t i) {
    PmhRigidBodyState s;
    const V3 t = v3(a.tpos[i]);
    const float4 d = a.flow_stamp ? ld4_co(&a.disp[2 * i]) : a.disp[2 * i];
    s.position = pv(add(t, {d.x, d.y, d.z}));
    Q o = q4(a.tq[i]);
    const Q dq = q4(a.flow_stamp ? ld4_co(&a.dq[i]) : a.dq[i]);
    if (dq.x != 0.0f || dq.y != 0.0f || dq.z != 0.0f) o = qnormalize(qmul(dq, o));
    s.orientation = pq(o);
    return s;
}

// ---------------------------------------------------------------------
// Joints: rigid_avbd_visit_joint(_row), geometry_constraints.slang:
// 289-445, 626-926.  One routine builds a joint's rows for the four visits
// (initialise in prepare, accumulate in the colour sweep, advance in the
// dual update, the velocity-phase rows).  Same operations and order as the
// reference.
// ---------------------------------------------------------------------
enum : uint32_t { kVisitInitialize = 0, kVisitAccumulate = 1, kVisitAdvance = 2, kVisitImpact = 3 };
constexpr float kLimitStabilization = 0.8f;

__device__ __forceinline__ float aabs_(float v) { return v < 0.0f ? -v : v; }

__device__ __forceinline__ float jrow_trial(const PmhAvbdRow &r, const float *da, const float *db, float dt) {
    const float d = dot6(r.a, da) + dot6(r.b, db);
    return r.lambda + r.penalty * (r.error + d) + r.damping * (r.velocity + d / dt);
}
__device__ __forceinline__ float jrow_force(const PmhAvbdRow &r, const float *da, const float *db, float dt) {
    return aclamp(jrow_trial(r, da, db, dt), r.lower, r.upper);
}
__device__ __forceinline__ void jrow_accumulate(Block &blk, const PmhAvbdRow &r, bool first, const float *da,
                                                const float *db, float dt) {
    const float trial = jrow_trial(r, da, db, dt);
    const bool two_sided = r.lower < 0.0f && r.upper > 0.0f;
    const float k = two_sided || (trial > r.lower && trial < r.upper) ? r.penalty + r.damping / dt : 0.0f;
    add_row(blk, first ? r.a : r.b, k, aclamp(trial, r.lower, r.upper) - r.reference_force);
}
__device__ __forceinline__ void jrow_advance(PmhAvbdRow &r, const float *da, const float *db) {
    const float e = r.error + dot6(r.a, da) + dot6(r.b, db);
    const float u = r.lambda + r.penalty * e;
    r.lambda = r.stiffness < kMaxPenalty ? 0.0f : aclamp(u, r.lower, r.upper);
    if (u > r.lower && u < r.upper) r.penalty = amin(r.stiffness, amin(kMaxPenalty, r.penalty + r.beta * aabs_(e)));
}

__device__ __forceinline__ V3 iiw(const PmhBodyParameters &p, Q o, V3 w) {
    const V3 l = inv_rotate(o, w);
    return rotate(o, {l.x * p.inverse_inertia_local.x, l.y * p.inverse_inertia_local.y,
                      l.z * p.inverse_inertia_local.z});
}

// rigid_avbd_log_axis, :401-422.
__device__ V3 log_axis(Q frame, V3 lg, uint32_t axis) {
    const float sq2 = len2(lg);
    float coef = 1.0f / 12.0f + sq2 * (1.0f / 720.0f + sq2 * (1.0f / 30240.0f + sq2 * (1.0f / 1209600.0f +
                 sq2 / 47900160.0f)));
    if (sq2 > 1.0f) {
        const float angle = sqrtf(sq2);
        const float half = 0.5f * angle;
        coef = (1.0f - half * cosf(half) / sinf(half)) / sq2;
    }
    const V3 unit = basis_axis(axis);
    const V3 local = add(add(unit, mul(cross(lg, unit), 0.5f)), mul(cross(lg, cross(lg, unit)), coef));
    return rotate(frame, local);
}

__device__ __forceinline__ V3 rel_rotation_vector(Q a, Q b) {
    return quaternion_rotation_vector(relative_quaternion(a, b, true));
}

struct JointRowArgs {
    float ja[6], jb[6];
    float current, previous, penalty;
    bool angular;
    float stiffness, damping, lower, upper;
    bool motor;
    float motor_speed;
    V3 curv_a, curv_b;
    float stabilization;
};

__device__ __forceinline__ void set6(float *d, V3 l, V3 r) {
    d[0] = l.x; d[1] = l.y; d[2] = l.z; d[3] = r.x; d[4] = r.y; d[5] = r.z;
}

// rigid_avbd_visit_joint_row, :626-713.
__device__ void visit_joint_row(const SolveArgs &a, PmhRigidConstraint &j, Block &blk, uint32_t &index,
                                float &impulse, bool first, const JointRowArgs &r, uint32_t visit) {
    if (index >= 18u) return;
    const uint32_t ri = index++;
    PmhAvbdRow row = j.rows[ri];
    const uint32_t ia = j.dense_a, ib = j.dense_b;
    float da[6], db[6];
    load_disp_a(a, ia, da);
    load_disp_a(a, ib, db);
    const float linearized = dot6(r.ja, da) + dot6(r.jb, db);
#pragma unroll
    for (int k = 0; k < 6; ++k) { row.a[k] = r.ja[k]; row.b[k] = r.jb[k]; }
    row.stiffness = r.stiffness;
    row.damping = r.damping;
    if (visit == kVisitInitialize) {
        if (!(row.penalty > kMinPenalty)) row.penalty = r.penalty;
        warm_start(row.lambda, row.penalty, r.stiffness);
        if (r.stiffness < kMaxPenalty) row.lambda = 0.0f;
        row.beta = r.penalty * (r.angular ? 10.0f : 100.0f);
        row.reference_force = 0.0f;
    }
    row.error = r.current - (r.stiffness < kMaxPenalty ? 0.0f : r.stabilization * r.previous) - linearized;
    row.velocity = r.motor ? r.motor_speed : (r.current - r.previous - linearized) / a.dt;
    row.lower = r.lower;
    row.upper = r.upper;
    if (visit == kVisitAccumulate) {
        jrow_accumulate(blk, row, first, da, db, a.dt);
        if (!r.angular) {
            const V3 axis = {r.jb[0], r.jb[1], r.jb[2]};
            const V3 arm = first ? r.curv_a : r.curv_b;
            const float force = jrow_force(row, da, db, a.dt);
            for (uint32_t c = 0; c < 3; ++c) {
                const V3 value = sub(mul(add(mul(axis, comp(arm, c)), mul(arm, comp(axis, c))), 0.5f),
                                     mul(basis_axis(c), dot(axis, arm)));
                const uint32_t d = 3 + c;
                blk.h[d * (d + 1) / 2 + d] += fabsf(force) * sqrtf(len2(value));
            }
        }
        return;
    }
    if (visit == kVisitAdvance) {
        impulse += fabsf(jrow_force(row, da, db, a.dt)) * a.dt;
        jrow_advance(row, da, db);
    } else if (visit == kVisitImpact) {
        if (r.stiffness < kMaxPenalty) {
            row.reference_force = jrow_force(row, da, db, a.dt);
        } else {
            const float gap = r.lower == 0.0f ? smin(0.0f, r.current) : (r.upper == 0.0f ? smax(0.0f, r.current) : 0.0f);
            float va[6], vb[6];
            set6(va, mul(v3(a.tv[ia]), a.dt), mul(v3(a.tw[ia]), a.dt));
            set6(vb, mul(v3(a.tv[ib]), a.dt), mul(v3(a.tw[ib]), a.dt));
            row.reference_force = row.lambda;
            row.error = dot6(r.ja, va) + dot6(r.jb, vb) + gap - (r.motor ? r.motor_speed * a.dt : 0.0f);
        }
    }
    j.rows[ri] = row;
}

// rigid_avbd_visit_joint, :715-926.
__device__ __forceinline__ float visit_joint(const SolveArgs &a, PmhRigidConstraint &j, Block &blk, bool first,
                                          uint32_t visit) {
    uint32_t index = 0;
    float impulse = 0.0f;
    const PmhRigidConstraintOptions &o = j.options;
    const uint32_t ia = j.dense_a, ib = j.dense_b;
    const PmhRigidBodyState sa = avbd_pose(a, ia), sb = avbd_pose(a, ib);
    const PmhRigidBodyState pa = a.prev[ia], pb = a.prev[ib];
    const Q qa = sq(sa.orientation), qb = sq(sb.orientation);
    const V3 armA = rotate(qa, sv(o.local_anchor_a)), armB = rotate(qb, sv(o.local_anchor_b));
    const V3 error = sub(add(sv(sb.position), armB), add(sv(sa.position), armA));
    const V3 old_error = sub(add(sv(pb.position), rotate(sq(pb.orientation), sv(o.local_anchor_b))),
                             add(sv(pa.position), rotate(sq(pa.orientation), sv(o.local_anchor_a))));
    const Q fA = qnormalize(qmul(qa, sq(o.local_orientation_a)));
    const Q fB = qnormalize(qmul(qb, sq(o.local_orientation_b)));
    const Q ofA = qnormalize(qmul(sq(pa.orientation), sq(o.local_orientation_a)));
    const Q ofB = qnormalize(qmul(sq(pb.orientation), sq(o.local_orientation_b)));
    const V3 twist = rel_rotation_vector(fA, fB);
    const V3 old_twist = rel_rotation_vector(ofA, ofB);
    const uint32_t type = o.type;
    const bool swing = type == kJointHinge || type == kJointPiston ||
                       (type == kJointMotor && o.motor.angular_enabled != 0u);
    const V3 free_axis = type == kJointHinge ? V3{0.0f, 0.0f, 1.0f} : V3{1.0f, 0.0f, 0.0f};
    const V3 swing_axis = inv_rotate(fA, rotate(fB, free_axis));
    const V3 old_swing_axis = inv_rotate(ofA, rotate(ofB, free_axis));
    const V3 rotation = swing ? cross(free_axis, swing_axis) : twist;
    const V3 old_rotation = swing ? cross(free_axis, old_swing_axis) : old_twist;
    const PmhBodyParameters Pa = a.P[ia], Pb = a.P[ib];
    const float linear_penalty = pair_stiffness(Pa, Pb, a.dt);
    const bool generic = type == kJointGeneric || type == kJointGenericSpring;
    const bool motor_type = type == kJointMotor;
    const V3 zero = {0.0f, 0.0f, 0.0f};
    for (uint32_t axis = 0; axis < 3; ++axis) {
        const V3 direction = rotate(fA, basis_axis(axis));
        const V3 old_direction = rotate(ofA, basis_axis(axis));
        for (uint32_t av = 0; av < 2; ++av) {
            const bool angular = av != 0;
            const bool lock = angular
                ? type == kJointFixed || type == kJointSlider || (type == kJointHinge && axis != 2u) ||
                      (type == kJointPiston && axis != 0u) ||
                      (motor_type && (axis != 0u || o.motor.angular_enabled == 0u))
                : type == kJointFixed || type == kJointPoint || type == kJointHinge ||
                      ((type == kJointSlider || type == kJointPiston) && axis != 0u) ||
                      (motor_type && (axis != 0u || o.motor.linear_enabled == 0u));
            const PmhJointLimits &limits = angular ? o.angular_limits : o.linear_limits;
            const PmhJointSprings &springs = angular ? o.angular_springs : o.linear_springs;
            const float current = angular ? comp(rotation, axis) : dot(error, direction);
            const float old_current = angular ? comp(old_rotation, axis) : dot(old_error, old_direction);
            const V3 angular_axis = swing
                ? rotate(fA, sub(mul(basis_axis(axis), dot(free_axis, swing_axis)),
                                 mul(free_axis, comp(swing_axis, axis))))
                : log_axis(fA, twist, axis);
            JointRowArgs r;
            if (angular) {
                set6(r.ja, zero, mul(angular_axis, -1.0f));
                set6(r.jb, zero, angular_axis);
            } else {
                set6(r.ja, mul(direction, -1.0f), add(mul(cross(armA, direction), -1.0f), cross(direction, error)));
                set6(r.jb, direction, cross(armB, direction));
            }
            const float penalty = angular
                ? 1.0f / (smax(dot(direction, add(iiw(Pa, qa, direction), iiw(Pb, qb, direction))), 1.0e-6f) *
                          a.dt * a.dt)
                : linear_penalty;
            const float lock_penalty = angular
                ? smax(penalty, 2.0f * linear_penalty * smax(len2(cross(direction, armA)), len2(cross(direction, armB))))
                : penalty;
            r.angular = angular;
            r.motor = false;
            r.motor_speed = 0.0f;
            r.curv_a = add(error, armA);
            r.curv_b = armB;
            if (lock) {
                if (visit == kVisitInitialize && angular && index < 18u)
                    j.rows[index].penalty = smax(j.rows[index].penalty, lock_penalty / kGamma);
                r.current = current; r.previous = old_current; r.penalty = lock_penalty;
                r.stiffness = kMaxPenalty; r.damping = 0.0f; r.lower = -kMaxPenalty; r.upper = kMaxPenalty;
                r.stabilization = kAlpha;
                visit_joint_row(a, j, blk, index, impulse, first, r, visit);
            }
            const bool free_limit = angular ? (type == kJointHinge && axis == 2u) || (type == kJointPiston && axis == 0u)
                                            : (type == kJointSlider || type == kJointPiston) && axis == 0u;
            if ((generic || free_limit) && (limits.axes & (1u << axis)) != 0u) {
                JointRowArgs l = r;
                const float limit_current = angular ? comp(twist, axis) : current;
                const float old_limit_current = angular ? comp(old_twist, axis) : old_current;
                if (angular) {
                    const V3 limit_axis = log_axis(fA, twist, axis);
                    set6(l.ja, zero, mul(limit_axis, -1.0f));
                    set6(l.jb, zero, limit_axis);
                }
                const float lower = comp(sv(limits.lower), axis), upper = comp(sv(limits.upper), axis);
                l.stiffness = kMaxPenalty;
                l.damping = 0.0f;
                if (lower == upper) {
                    const float equal_penalty = angular && generic ? lock_penalty : penalty;
                    if (visit == kVisitInitialize && angular && generic && index < 18u)
                        j.rows[index].penalty = smax(j.rows[index].penalty, equal_penalty / kGamma);
                    l.current = limit_current - lower; l.previous = old_limit_current - lower;
                    l.penalty = equal_penalty; l.lower = -kMaxPenalty; l.upper = kMaxPenalty;
                    l.stabilization = kAlpha;
                    visit_joint_row(a, j, blk, index, impulse, first, l, visit);
                } else if (lower < upper) {
                    l.current = limit_current - upper; l.previous = smax(old_limit_current - upper, 0.0f);
                    l.penalty = penalty; l.lower = 0.0f; l.upper = kMaxPenalty;
                    l.stabilization = kLimitStabilization;
                    visit_joint_row(a, j, blk, index, impulse, first, l, visit);
                    l.current = limit_current - lower; l.previous = smin(old_limit_current - lower, 0.0f);
                    l.lower = -kMaxPenalty; l.upper = 0.0f;
                    visit_joint_row(a, j, blk, index, impulse, first, l, visit);
                }
            }
            if (type == kJointGenericSpring && (springs.axes & (1u << axis)) != 0u) {
                JointRowArgs sr = r;
                sr.current = current; sr.previous = old_current; sr.penalty = penalty;
                sr.stiffness = comp(sv(springs.stiffness), axis); sr.damping = comp(sv(springs.damping), axis);
                sr.lower = -kMaxPenalty; sr.upper = kMaxPenalty; sr.stabilization = kAlpha;
                visit_joint_row(a, j, blk, index, impulse, first, sr, visit);
            }
            if (motor_type && axis == 0u && (angular ? o.motor.angular_enabled != 0u : o.motor.linear_enabled != 0u)) {
                const float speed = angular ? o.motor.angular_target_velocity : o.motor.linear_target_velocity;
                const float bound = (angular ? o.motor.angular_maximum_impulse : o.motor.linear_maximum_impulse) / a.dt;
                const Q rel = relative_quaternion(fA, fB, true);
                const Q old_rel = relative_quaternion(ofA, ofB, true);
                const V3 motor_rotation = rel_rotation_vector({0.0f, 0.0f, 0.0f, 1.0f},
                                                              relative_quaternion(old_rel, rel, false));
                JointRowArgs mr = r;
                if (angular) {
                    const V3 motor_axis = log_axis(fA, motor_rotation, axis);
                    set6(mr.ja, zero, mul(motor_axis, -1.0f));
                    set6(mr.jb, zero, motor_axis);
                }
                const float travel = angular ? comp(motor_rotation, axis) : current - old_current;
                mr.current = travel - speed * a.dt; mr.previous = 0.0f; mr.penalty = penalty;
                mr.stiffness = kMaxPenalty; mr.damping = 0.0f; mr.lower = -bound; mr.upper = bound;
                mr.motor = true; mr.motor_speed = speed; mr.stabilization = kAlpha;
                visit_joint_row(a, j, blk, index, impulse, first, mr, visit);
            }
        }
    }
    if (visit == kVisitInitialize) j.avbd_count = index;
    return impulse;
}

// The joint rows of body i in its list order, added to its block: the
// pose phase recomputes the rows, the velocity phase adds the stored hard
// rows (solve.slang:401-430).
// The body's enabled joints are dealt round-robin over the L lanes of its
// group (the k-th enabled joint of the list to lane k mod L) and added to
// that lane's partial block before the group reduction.
__device__ void body_joint_rows(const SolveArgs &a, uint32_t i, uint32_t phase, Block &blk, uint32_t sub_lane,
                                uint32_t L) {
    uint32_t pos = 0;
    for (uint32_t e = a.jstart[i]; e < a.jstart[i + 1]; ++e) {
        PmhRigidConstraint &j = a.joints[a.jadj[e]];
        if (j.state_enabled == 0u) continue;
        if (pos++ % L != sub_lane) continue;
        const bool first = i == j.dense_a;
        if (a.flow_stamp) {
            const uint32_t other = first ? j.dense_b : j.dense_a;
            if (a.color[other] < a.color[i] && (phase == 0u || a.impact[other] != 0u)) wait_published(a, other);
        }
        if (phase == 0u) {
            visit_joint(a, j, blk, first, kVisitAccumulate);
            continue;
        }
        float da[6], db[6];
        load_disp_a(a, j.dense_a, da);
        load_disp_a(a, j.dense_b, db);
        for (uint32_t r = 0; r < j.avbd_count; ++r) {
            const PmhAvbdRow row = j.rows[r];
            if (row.stiffness >= kMaxPenalty) jrow_accumulate(blk, row, first, da, db, a.dt);
        }
    }
}

// Joint part of advance_duals (solve.slang:566-614): dual updates, the
// applied impulse and breaking, one thread per joint.
__device__ void advance_joint(const SolveArgs &a, uint32_t k, uint32_t phase) {
    PmhRigidConstraint &j = a.joints[k];
    if (j.avbd_count == 0u || j.state_enabled == 0u) return;
    const uint32_t ia = j.dense_a, ib = j.dense_b;
    if (phase != 0u && a.impact[ia] == 0u && a.impact[ib] == 0u) return;
    float impulse = 0.0f;
    if (phase == 0u) {
        Block scratch;
        block_zero(scratch);
        impulse = visit_joint(a, j, scratch, true, kVisitAdvance);
    } else {
        float da[6], db[6];
        load_disp(a.disp, ia, da);
        load_disp(a.disp, ib, db);
        for (uint32_t r = 0; r < j.avbd_count; ++r) {
            PmhAvbdRow row = j.rows[r];
            impulse += fabsf(row.stiffness < kMaxPenalty ? row.reference_force : jrow_force(row, da, db, a.dt)) * a.dt;
            if (row.stiffness >= kMaxPenalty) {
                jrow_advance(row, da, db);
                j.rows[r] = row;
            }
        }
    }
    j.applied_impulse = impulse;
    if (j.options.breaking_impulse_threshold > 0.0f && impulse > j.options.breaking_impulse_threshold) {
        j.state_broken = 1u;
        j.state_enabled = 0u;
    }
}

// Residual of a joint's hard equality rows: |C| at the current
// displacements.
__device__ float joint_residual(const SolveArgs &a, const PmhRigidConstraint &j) {
    if (j.avbd_count == 0u || j.state_enabled == 0u) return 0.0f;
    float da[6], db[6];
    load_disp(a.disp, j.dense_a, da);
    load_disp(a.disp, j.dense_b, db);
    float r = 0.0f;
    for (uint32_t k = 0; k < j.avbd_count; ++k) {
        const PmhAvbdRow &row = j.rows[k];
        if (row.stiffness < kMaxPenalty || row.lower != -kMaxPenalty || row.upper != kMaxPenalty) continue;
        r = smax(r, fabsf(row.error + dot6(row.a, da) + dot6(row.b, db)));
    }
    return r;
}

// Point flag: the point slid at the last dual update (to-residual mode,
// where the impulses are published after the loop).
constexpr uint32_t kFlagSlipping = 4u;

// The impulses of the last dual update and the re-anchoring of a sliding
// contact (solve.slang:540-562).
__device__ void publish_point(const SolveArgs &a, uint32_t c, uint32_t phase, uint2 p, float fn, float ft0,
                              float ft1, bool stick, V3 t0, V3 t1) {
    const float an = fn * a.dt;
    const V3 af = mul(add(mul(t0, ft0), mul(t1, ft1)), a.dt);
    a.c_acc[c] = make_float4(an, af.x, af.y, af.z);
    a.c_rep[c] = make_float4(an, af.x, af.y, af.z);
    if (!stick) {
        PmhRigidBodyState pa, pb;
        if (phase != 0) {
            pa.position = pv(v3(a.tpos[p.x])); pa.orientation = pq(q4(a.tq[p.x]));
            pb.position = pv(v3(a.tpos[p.y])); pb.orientation = pq(q4(a.tq[p.y]));
        } else {
            pa = avbd_pose(a, p.x);
            pb = avbd_pose(a, p.y);
        }
        const V3 point = v3(a.c_pt[c]);
        const V3 na = inv_rotate(sq(pa.orientation), sub(point, sv(pa.position)));
        const V3 nb = inv_rotate(sq(pb.orientation), sub(point, sv(pb.position)));
        a.c_aa[c] = f4(na, a.c_aa[c].w);
        a.c_ab[c] = f4(nb, 0.0f);
    }
}

// Constraint residual of a contact point (no reference counterpart): an
// active normal row must hold, an inactive one must not penetrate, and a
// sticking friction row must not slip.
__device__ __forceinline__ float contact_residual(const float *err, float fn, bool stick) {
    float r = fn > 0.0f ? fabsf(err[0]) : smax(err[0], 0.0f);
    if (stick) r = smax(r, smax(fabsf(err[1]), fabsf(err[2])));
    return r;
}

__device__ void dual_point(const SolveArgs &a, uint32_t c, uint32_t phase, const float *dA, const float *dB,
                           const float4 *Jx, bool last, ResidualAcc &residual) {
    float4 lam = a.c_lam[c], pen = a.c_pen[c];
    const float err[3] = {row_error(Jx[0].w, Jx[0], Jx[3], Jx[6], dA, dB),
                          row_error(Jx[1].w, Jx[1], Jx[4], Jx[7], dA, dB),
                          row_error(Jx[2].w, Jx[2], Jx[5], Jx[8], dA, dB)};
    const float mu = Jx[6].w, beta = Jx[7].w;
    const Force3 trial = {lam.x + pen.x * err[0], lam.y + pen.y * err[1], lam.z + pen.z * err[2]};
    const Force3 f = project_contact(trial, mu);
    const float tx = lam.y + pen.y * err[1];
    const float ty = lam.z + pen.z * err[2];
    const float fb = mu * f.n;
    const bool stick = tx * tx + ty * ty <= fb * fb;
    lam = make_float4(f.n, f.t0, f.t1, 0.0f);
    if (f.n > 0.0f) pen.x = smin(kMaxPenalty, pen.x + beta * fabsf(err[0]));
    if (stick) {
        pen.y = smin(kMaxPenalty, pen.y + beta * fabsf(err[1]));
        pen.z = smin(kMaxPenalty, pen.z + beta * fabsf(err[2]));
    }
    a.c_lam[c] = lam;
    a.c_pen[c] = pen;
    residual.add(contact_residual(err, f.n, stick));
    if (last) publish_point(a, c, phase, a.ab[a.c_pair[c]], f.n, f.t0, f.t1, stick, v3(Jx[1]), v3(Jx[2]));
    else if (a.tol > 0.0f) a.c_flags[c] = (a.c_flags[c] & ~kFlagSlipping) | (stick ? 0u : kFlagSlipping);
}

// advance_duals for contacts (solve.slang:492-566): one thread per point.
__device__ void advance_duals(const SolveArgs &a, uint32_t phase, uint32_t iteration, uint32_t tid, uint32_t nthreads) {
    const uint32_t Pn = *a.npts;
    const bool last = a.tol <= 0.0f && iteration + 1u == a.budget;
    ResidualAcc residual;
    for (uint32_t c = tid; c < Pn; c += nthreads) {
        const uint2 p = a.ab[a.c_pair[c]];
        if (phase != 0 && a.impact[p.x] == 0 && a.impact[p.y] == 0) continue;
        float dA[6], dB[6];
        load_disp(a.disp, p.x, dA);
        load_disp(a.disp, p.y, dB);
        const float4 *J = a.jac + 9 * (size_t)c;
        float4 Jx[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) Jx[q] = J[q];
        dual_point(a, c, phase, dA, dB, Jx, last, residual);
    }
    if (iteration < kResidualSlots) residual_commit(a.log, phase * kResidualSlots + iteration, residual);
}

// free_mean (solve.slang:617-675): the mass-weighted mean translation of
// each unanchored island with contacts, its members summed in ascending
// index order.  One wavefront per listed root scans 64 indices at a time
// from the root (the smallest index of its island) and stops once it has
// met every member; the sum itself stays sequential in index order.
__device__ void free_mean(const SolveArgs &a, uint32_t phase, uint32_t gw, uint32_t nw) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t roots = a.ctr->mean_roots;
    for (uint32_t k = gw; k < roots; k += nw) {
        const uint32_t r = a.mean_list[k];
        float4 mean = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (!(phase != 0 && a.impact[r] == 0)) {
            const uint32_t want = a.members[r];
            uint32_t found = 0;
            V3 w = {0.0f, 0.0f, 0.0f};
            float total = 0.0f;
            for (uint32_t base = r; found < want && base < a.n; base += 64) {
                const uint32_t i = base + lane;
                const bool member = i < a.n && a.parent[i] == r && a.P[i].motion == kDynamic;
                V3 prod = {0.0f, 0.0f, 0.0f};
                float mass = 0.0f;
                if (member) {
                    mass = 1.0f / smax(a.P[i].inverse_mass, 1.0e-12f);
                    const float4 d = a.disp[2 * i];
                    prod = mul({d.x, d.y, d.z}, mass);
                }
                uint64_t bal = __ballot(member);
                while (bal) {
                    const int b = __builtin_ctzll(bal);
                    bal &= bal - 1ull;
                    const V3 p = {__shfl(prod.x, b, 64), __shfl(prod.y, b, 64), __shfl(prod.z, b, 64)};
                    w = add(w, p);
                    total += __shfl(mass, b, 64);
                    ++found;
                }
            }
            if (total > 0.0f) {
                w = mul(w, 1.0f / total);
                mean = make_float4(w.x, w.y, w.z, 0.0f);
            }
        }
        if (lane == 0) a.mean[r] = mean;
    }
}

__device__ void free_subtract_apply(const SolveArgs &a, uint32_t phase, uint32_t tid, uint32_t nthreads) {
    const bool vel = phase != 0;
    for (uint32_t i = tid; i < a.n; i += nthreads) {
        if (a.P[i].motion != kDynamic) continue;
        if (vel && a.impact[i] == 0) continue;
        const float4 m = a.mean[a.parent[i]];
        float4 t = a.disp[2 * i];
        t.x -= m.x; t.y -= m.y; t.z -= m.z;
        a.disp[2 * i] = t;
        // apply_states (solve.slang:699-733)
        PmhRigidBodyState s;
        s.position = pv(v3(a.tpos[i]));
        s.orientation = pq(q4(a.tq[i]));
        V3 lv = v3(a.tv[i]), av = v3(a.tw[i]);
        const float4 r = a.disp[2 * i + 1];
        const V3 dt_ = {t.x, t.y, t.z};
        if (phase != 0) {
            lv = add(lv, mul(dt_, 1.0f / a.dt));
            av = add(av, mul({r.x, r.y, r.z}, 1.0f / a.dt));
            s.linear_velocity = pv(lv);
            s.angular_velocity = pv(av);
            a.states[i] = s;
            continue;
        }
        s.position = pv(add(sv(s.position), dt_));
        s.orientation = avbd_pose(a, i).orientation;
        lv = add(lv, mul(dt_, 1.0f / a.dt));
        av = add(av, mul(quaternion_rotation_vector(q4(a.dq[i])), 1.0f / a.dt));
        if (a.realtime) {
            lv = mul(lv, 1.0f - kRealtimeDamping);
            av = mul(av, 1.0f - kRealtimeDamping);
        }
        s.linear_velocity = pv(lv);
        s.angular_velocity = pv(av);
        a.states[i] = s;
    }
}

// Velocity-phase set-up (impact_gaps / impact_bodies /
// impact_constraints, solve.slang:232-329).
__device__ void impact_gaps(const SolveArgs &a, uint32_t tid, uint32_t nthreads) {
    const uint32_t Pn = *a.npts;
    for (uint32_t c = tid; c < Pn; c += nthreads) {
        const uint2 p = a.ab[a.c_pair[c]];
        if (a.impact[p.x] == 0 && a.impact[p.y] == 0) continue;
        float dA[6], dB[6];
        load_disp(a.disp, p.x, dA);
        load_disp(a.disp, p.y, dB);
        const float4 *J = a.jac + 9 * (size_t)c;
        const float pen = a.c_np[c].w;
        // normal row with the contact penetration as the constant term
        const float e = row_error(pen, J[0], J[3], J[6], dA, dB);
        float4 err = a.c_err[c];
        err.x = smax(0.0f, -e);
        a.c_err[c] = err;
    }
}

__device__ void impact_bodies(const SolveArgs &a, uint32_t tid, uint32_t nthreads) {
    for (uint32_t i = tid; i < a.n; i += nthreads) {
        if (a.impact[i] != 0 || a.P[i].motion != kDynamic) {
            const PmhRigidBodyState s = a.states[i];
            a.tpos[i] = f4(sv(s.position), 0.0f);
            a.tq[i] = f4q(sq(s.orientation));
            a.tv[i] = f4(sv(s.linear_velocity), 0.0f);
            a.tw[i] = f4(sv(s.angular_velocity), 0.0f);
            a.disp[2 * i] = make_float4(0, 0, 0, 0);
            a.disp[2 * i + 1] = make_float4(0, 0, 0, 0);
            a.dq[i] = make_float4(0, 0, 0, 1);
        }
    }
}

// Velocity-level errors from the reset targets, with restitution for cold
// contacts closing faster than 1 m/s (impact_constraints, solve.slang:
// 293-329; rows as rigid_avbd_contact_rows).
__device__ void velocity_errors(const SolveArgs &a, const uint32_t *c_flags, uint32_t tid, uint32_t nthreads) {
    const uint32_t Pn = *a.npts;
    for (uint32_t c = tid; c < Pn; c += nthreads) {
        const uint2 p = a.ab[a.c_pair[c]];
        if (a.impact[p.x] == 0 && a.impact[p.y] == 0) continue;
        const PmhRigidBodyState A = a.states[p.x], B = a.states[p.y];
        const float dt = a.dt;
        const float va[6] = {A.linear_velocity.x * dt, A.linear_velocity.y * dt, A.linear_velocity.z * dt,
                             A.angular_velocity.x * dt, A.angular_velocity.y * dt, A.angular_velocity.z * dt};
        const float vb[6] = {B.linear_velocity.x * dt, B.linear_velocity.y * dt, B.linear_velocity.z * dt,
                             B.angular_velocity.x * dt, B.angular_velocity.y * dt, B.angular_velocity.z * dt};
        const V3 n = v3(a.c_np[c]), point = v3(a.c_pt[c]);
        V3 t0, t1;
        contact_basis(n, t0, t1);
        const V3 axes[3] = {n, t0, t1};
        const V3 armA = sub(point, v3(a.tpos[p.x])), armB = sub(point, v3(a.tpos[p.y]));
        const V3 ra_t = rotate(q4(a.tq[p.x]), v3(a.c_aa[c])), rb_t = rotate(q4(a.tq[p.y]), v3(a.c_ab[c]));
        float out[3];
        for (uint32_t ax = 0; ax < 3; ++ax) {
            const V3 ra = ax == 0 ? armA : ra_t, rb = ax == 0 ? armB : rb_t;
            const V3 la = mul(axes[ax], -1.0f), aa = mul(cross(ra, axes[ax]), -1.0f), ba = cross(rb, axes[ax]);
            const float ja[6] = {la.x, la.y, la.z, aa.x, aa.y, aa.z};
            const float jb[6] = {axes[ax].x, axes[ax].y, axes[ax].z, ba.x, ba.y, ba.z};
            out[ax] = dot6(ja, va) + dot6(jb, vb);
        }
        const float gap = a.c_err[c].x;
        const bool cached = (c_flags[c] & 1u) != 0u;
        const float speed = a.c_pt[c].w;
        const float rest = !cached && speed < -1.0f ? smin(a.P[p.x].restitution, a.P[p.y].restitution) : 0.0f;
        a.c_err[c] = make_float4(out[0] - gap - rest * speed * dt, out[1], out[2], 0.0f);
    }
}

// Jacobians of the velocity phase from the reset targets, with the
// accumulated impulses of the pose phase as reference forces.
__device__ void velocity_jacobians(const SolveArgs &a, uint32_t tid, uint32_t nthreads) {
    const uint32_t Pn = *a.npts;
    for (uint32_t c = tid; c < Pn; c += nthreads) {
        const uint2 p = a.ab[a.c_pair[c]];
        if (a.impact[p.x] == 0 && a.impact[p.y] == 0) continue;
        const V3 n = v3(a.c_np[c]), point = v3(a.c_pt[c]);
        V3 t0, t1;
        contact_basis(n, t0, t1);
        const V3 axes[3] = {n, t0, t1};
        const V3 armA = sub(point, v3(a.tpos[p.x])), armB = sub(point, v3(a.tpos[p.y]));
        const V3 ra_t = rotate(q4(a.tq[p.x]), v3(a.c_aa[c])), rb_t = rotate(q4(a.tq[p.y]), v3(a.c_ab[c]));
        const float4 e = a.c_err[c];
        const float e0[3] = {e.x, e.y, e.z};
        const float4 acc = a.c_acc[c];
        const V3 af = {acc.y, acc.z, acc.w};
        const float ref[3] = {acc.x / a.dt, dot(af, t0) / a.dt, dot(af, t1) / a.dt};
        const float4 *old = a.jac + 9 * (size_t)c;
        const float mu = old[6].w, beta = old[7].w;
        float4 J[9];
        for (uint32_t ax = 0; ax < 3; ++ax) {
            const V3 ra = ax == 0 ? armA : ra_t, rb = ax == 0 ? armB : rb_t;
            J[ax] = f4(axes[ax], e0[ax]);
            J[3 + ax] = f4(mul(cross(ra, axes[ax]), -1.0f), ref[ax]);
            J[6 + ax] = f4(cross(rb, axes[ax]), ax == 0 ? mu : (ax == 1 ? beta : 0.0f));
        }
        store_rows(const_cast<float4 *>(a.jac), const_cast<float4 *>(a.erec), a.ecap, a.c_entry, c,
                   a.P[p.x].motion == kDynamic, a.P[p.y].motion == kDynamic, J);
    }
}

// Iterations a phase ran: the budget, or in to-residual mode the update
// that met the tolerance.
__device__ __forceinline__ uint32_t phase_iterations(const SolveArgs &a, uint32_t phase) {
    return a.tol > 0.0f && a.log->executed[phase] != 0u ? a.log->executed[phase] : a.budget;
}

__device__ void account_phase(const SolveArgs &a, uint32_t phase) {
    const StepCounters *c = a.ctr;
    SolveTotals *t = a.totals;
    const uint32_t its = phase_iterations(a, phase);
    if (phase == 0) {
        t->steps += 1;
        t->body_iterations += (unsigned long long)c->dyn_bodies * its;
        t->algorithmic_bytes += its * iteration_bytes(c->dyn_bodies, c->m1, c->m2, c->p1, c->p2);
    } else {
        t->velocity_phases += 1;
        t->body_iterations += (unsigned long long)c->impact_bodies * its;
        t->algorithmic_bytes += its * iteration_bytes(c->impact_bodies, c->m1i, c->m2i, c->p1i, c->p2i);
    }
}

__device__ __forceinline__ uint32_t global_wave() { return (blockIdx.x * blockDim.x + threadIdx.x) >> 6; }

// The whole solve in one cooperative launch.
template <int L>
__global__ void __launch_bounds__(256) k_solve_all(SolveArgs a, const uint32_t *c_flags) {
    cg::grid_group grid = cg::this_grid();
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    const uint32_t gw = tid >> 6, nw = nthreads >> 6;
    const uint32_t colors = a.ctr->colors;
    const uint32_t phases = a.realtime ? 1u : 2u;
    for (uint32_t phase = 0; phase < phases; ++phase) {
        if (phase != 0) {
            if (a.ctr->impact == 0) break;
            impact_gaps(a, tid, nthreads);
            grid.sync();
            impact_bodies(a, tid, nthreads);
            grid.sync();
            velocity_errors(a, c_flags, tid, nthreads);
            grid.sync();
            velocity_jacobians(a, tid, nthreads);
            grid.sync();
        }
        for (uint32_t it = 0; it < a.budget; ++it) {
            for (uint32_t c = 0; c < colors; ++c) {
                solve_color<L, false>(a, a.color_start[c], a.color_start[c + 1], phase, gw, nw);
                grid.sync();
            }
            advance_duals(a, phase, it, tid, nthreads);
            grid.sync();
        }
        free_mean(a, phase, gw, nw);
        grid.sync();
        free_subtract_apply(a, phase, tid, nthreads);
        grid.sync();
        if (tid == 0) {
            a.log->executed[phase] = a.budget;
            a.schedule->counts[phase == 0 ? 6 : 7] = a.budget;
            a.schedule->maximum_passes += a.budget;
            account_phase(a, phase);
        }
    }
}

// Staged variant: one launch per stage (no grid barrier).
template <int L, bool kJoints>
__global__ void __launch_bounds__(256) k_stage_color(SolveArgs a, uint32_t begin, uint32_t end, uint32_t phase) {
    if (a.tol > 0.0f && a.log->executed[phase] != 0u) return;   // converged earlier
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    solve_color<L, kJoints>(a, begin, end, phase, tid >> 6, (gridDim.x * blockDim.x) >> 6);
}
// Dataflow sweep: every colour in one launch per iteration.  Blocks take
// tickets of one block's worth of wavefront slots (each slot one
// wavefront's worth of colour-ordered positions; a colour starts at a slot
// boundary, so a wavefront holds one colour), in order, one atomic per
// block.  A body waits only for lower-colour neighbours: their positions
// were ticketed earlier, to blocks that are running, so every wait ends
// (and each is bounded besides).  Colour c + 1 starts wherever colour c is
// done, not after all of colour c.
constexpr uint32_t kSweepWaves = 8;          // wavefronts per block without joints (512 threads)
constexpr uint32_t kSweepJointWaves = 4;     // with joints (more registers per lane)
// Without joints the sweep is held to 128 VGPRs (four wavefronts per SIMD,
// two blocks per CU): at 137 only one block fit per CU, and twice the
// bodies in flight cut the sweep from 175 to 149 µs on the settled pile
// for 4 spilled registers.
template <int L, bool kJoints>
__global__ void __launch_bounds__(64 * (kJoints ? kSweepJointWaves : kSweepWaves), kJoints ? 1 : 4)
k_sweep(SolveArgs a, uint32_t phase, uint32_t it) {
    __shared__ uint32_t s_ticket[2];
    uint32_t *tickets = a.ctr->sweep_ticket;
    // The other counter is the next sweep's: it starts at zero.
    if (blockIdx.x == 0 && threadIdx.x == 0) st_relaxed(&tickets[(a.flow_stamp + 1u) & 1u], 0u);
    if (a.tol > 0.0f && a.log->executed[phase] != 0u) return;   // converged earlier
    constexpr uint32_t groups = 64 / L;
    constexpr uint32_t waves = kJoints ? kSweepJointWaves : kSweepWaves;
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t end = a.color_start[a.ctr->colors];
    const uint32_t wave_slots = (end + groups - 1) / groups;
    uint32_t *ticket = &tickets[a.flow_stamp & 1u];
    for (uint32_t round = 0;; ++round) {
        // Double-buffered: the slot written in round r + 1 was last read
        // before every wavefront passed round r's barrier.
        if (threadIdx.x == 0) s_ticket[round & 1u] = atomicAdd(ticket, 1u);
        __syncthreads();
        const uint32_t first = s_ticket[round & 1u] * waves;
        if (first >= wave_slots) break;      // the same for the whole block
        if (first + wave < wave_slots) solve_group<L, kJoints, true>(a, (first + wave) * groups, end, phase, it);
    }
}
template <bool kJoints>
__global__ void __launch_bounds__(256) k_stage_duals(SolveArgs a, uint32_t phase, uint32_t it) {
    if (a.tol > 0.0f && a.log->executed[phase] != 0u) return;   // converged earlier
    advance_duals(a, phase, it, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
    if (kJoints) {
        ResidualAcc r;
        for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < a.nj; k += gridDim.x * blockDim.x) {
            const PmhRigidConstraint &j = a.joints[k];
            if (j.avbd_count == 0u || j.state_enabled == 0u) continue;
            if (phase != 0u && a.impact[j.dense_a] == 0u && a.impact[j.dense_b] == 0u) continue;
            advance_joint(a, k, phase);
            r.add(joint_residual(a, j));
        }
        if (it < kResidualSlots) residual_commit(a.log, phase * kResidualSlots + it, r);
    }
}

// The joint half of k_stage_duals, after a dataflow sweep (which updates
// the contact points' duals itself).
__global__ void __launch_bounds__(256) k_stage_joint_duals(SolveArgs a, uint32_t phase, uint32_t it) {
    if (a.tol > 0.0f && a.log->executed[phase] != 0u) return;   // converged earlier
    ResidualAcc r;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < a.nj; k += gridDim.x * blockDim.x) {
        const PmhRigidConstraint &j = a.joints[k];
        if (j.avbd_count == 0u || j.state_enabled == 0u) continue;
        if (phase != 0u && a.impact[j.dense_a] == 0u && a.impact[j.dense_b] == 0u) continue;
        advance_joint(a, k, phase);
        r.add(joint_residual(a, j));
    }
    if (it < kResidualSlots) residual_commit(a.log, phase * kResidualSlots + it, r);
}

// To-residual mode: the phase stops at the first update whose residual is
// at most the tolerance, or at the budget.
// One wavefront: lane b folds bank b of the slot.
__global__ void k_stage_check(SolveArgs a, uint32_t phase, uint32_t it) {
    if (a.log->executed[phase] != 0u) return;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t slot = phase * kResidualSlots + it;
    unsigned long long sum = it < kResidualSlots ? a.log->sum[slot][lane] : 0ull;
    uint32_t count = it < kResidualSlots ? a.log->count[slot][lane] : 0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        sum += (unsigned long long)__shfl_xor((long long)sum, d, 64);
        count += uint32_t(__shfl_xor(count, d, 64));
    }
    if (lane != 0) return;
    const float r = it < kResidualSlots ? residual_rms(sum, count) : 0.0f;
    if (r <= a.tol || it + 1u == a.budget) a.log->executed[phase] = it + 1u;
}

// To-residual mode: impulses and re-anchoring from the last update.
__global__ void k_stage_publish(SolveArgs a, uint32_t phase) {
    const uint32_t Pn = *a.npts;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < Pn; c += gridDim.x * blockDim.x) {
        const uint2 p = a.ab[a.c_pair[c]];
        if (phase != 0 && a.impact[p.x] == 0 && a.impact[p.y] == 0) continue;
        const float4 lam = a.c_lam[c];
        const float4 *J = a.jac + 9 * (size_t)c;
        publish_point(a, c, phase, p, lam.x, lam.y, lam.z, (a.c_flags[c] & kFlagSlipping) == 0u, v3(J[1]), v3(J[2]));
    }
}
__global__ void k_stage_impact_gaps(SolveArgs a) {
    impact_gaps(a, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}
__global__ void k_stage_impact_bodies(SolveArgs a) {
    impact_bodies(a, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}
template <bool kJoints>
__global__ void k_stage_velocity_errors(SolveArgs a, const uint32_t *c_flags) {
    velocity_errors(a, c_flags, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
    // Joint rows of the velocity phase (impact_constraints, solve.slang:277-291).
    if (kJoints) for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < a.nj; k += gridDim.x * blockDim.x) {
        PmhRigidConstraint &j = a.joints[k];
        if (j.avbd_count == 0u || j.state_enabled == 0u) continue;
        if (a.impact[j.dense_a] == 0u && a.impact[j.dense_b] == 0u) continue;
        Block scratch;
        block_zero(scratch);
        visit_joint(a, j, scratch, true, kVisitImpact);
    }
}
__global__ void k_stage_velocity_jacobians(SolveArgs a) {
    velocity_jacobians(a, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}
__global__ void k_stage_mean(SolveArgs a, uint32_t phase) {
    free_mean(a, phase, global_wave(), (gridDim.x * blockDim.x) >> 6);
}
__global__ void k_stage_apply(SolveArgs a, uint32_t phase) {
    free_subtract_apply(a, phase, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t its = phase_iterations(a, phase);
        if (a.log->executed[phase] == 0u) a.log->executed[phase] = its;
        a.schedule->counts[phase == 0 ? 6 : 7] = its;
        a.schedule->maximum_passes += its;
        account_phase(a, phase);
    }
}

// ---- learned warm start: features of cold points and the contact log -----
// 16 features of a cold contact point (no reference counterpart): depth,
// closing motion over the step, normal, initial normal error, log pair
// stiffness, log effective mass, friction, both-dynamic flag, support
// against gravity, manifold size, face-patch flag, impact fraction,
// persistence, and 1.  Lengths are in centimetres.
__global__ void k_warm_features(const uint32_t *npts, const PmhBodyParameters *P, const uint32_t *c_pair,
                                const uint2 *ab, const uint32_t *m_count, const uint32_t *m_face,
                                const float4 *c_np, const float4 *c_pt, const float4 *c_err, const float4 *c_aa,
                                const uint32_t *c_flags, V3 gravity, float dt, uint32_t cap, uint32_t *list,
                                float *feat, float *scale, StepCounters *ctr) {
    const uint32_t Pn = *npts;
    const uint32_t lane = threadIdx.x & 63;
    const float g = sqrtf(dot(gravity, gravity));
    const V3 up = g > 0.0f ? mul(gravity, -1.0f / g) : V3{0.0f, 0.0f, 0.0f};
    for (uint32_t base = blockIdx.x * blockDim.x; base < Pn; base += gridDim.x * blockDim.x) {
        const uint32_t c = base + threadIdx.x;
        const bool cold = c < Pn && (c_flags[c] & 1u) == 0u;
        const uint64_t mask = __ballot(cold);
        uint32_t first = 0;
        if (lane == 0 && mask) first = atomicAdd(&ctr->cold_points, uint32_t(__popcll(mask)));
        first = __shfl(first, 0, 64);
        if (!cold) continue;
        const uint32_t idx = first + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
        if (idx >= cap) continue;
        const uint32_t m = c_pair[c];
        const uint2 p = ab[m];
        const PmhBodyParameters pa = P[p.x], pb = P[p.y];
        const float ia = pa.motion == kDynamic ? pa.inverse_mass : 0.0f;
        const float ib = pb.motion == kDynamic ? pb.inverse_mass : 0.0f;
        const float meff = ia + ib > 0.0f ? 1.0f / (ia + ib) : 1.0f;
        const float k0 = pair_stiffness(pa, pb, dt);
        const float4 np = c_np[c];
        const float f[kMlpIn] = {np.w * 100.0f, c_pt[c].w * dt * 100.0f, np.x, np.y, np.z, c_err[c].x * 100.0f,
                                 log10f(k0) * 0.1f, log10f(meff) * 0.25f, sqrtf(pa.friction * pb.friction),
                                 pa.motion == kDynamic && pb.motion == kDynamic ? 1.0f : 0.0f,
                                 dot(V3{np.x, np.y, np.z}, up), float(m_count[m]) * 0.125f,
                                 m_face[m] ? 1.0f : 0.0f, c_aa[c].w, (c_flags[c] & 2u) ? 1.0f : 0.0f, 1.0f};
        list[idx] = c;
#pragma unroll
        for (int q = 0; q < kMlpIn; ++q) feat[size_t(idx) * kMlpIn + q] = f[q];
        scale[2 * size_t(idx)] = g > 0.0f ? meff * g : meff;
        scale[2 * size_t(idx) + 1] = k0;
    }
}

// Appends each listed cold point's features and its duals at the end of
// the step (log10(1 + lambda / (m_eff |g|)), log10(k / k0)) to the log.
__global__ void k_warm_log(StepCounters *ctr, uint32_t cap_list, const uint32_t *list, const float *feat,
                           const float *scale, const float4 *c_lam, const float4 *c_pen, float *log,
                           uint32_t log_cap) {
    const uint32_t n = min(ctr->cold_points, cap_list);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t slot = atomicAdd(&ctr->log_count, 1u);
        if (slot >= log_cap) return;
        const uint32_t c = list[k];
        float *row = log + size_t(slot) * kMlpSample;
        for (int q = 0; q < kMlpIn; ++q) row[q] = feat[size_t(k) * kMlpIn + q];
        row[kMlpIn] = log10f(1.0f + fmaxf(c_lam[c].x, 0.0f) / scale[2 * size_t(k)]);
        row[kMlpIn + 1] = log10f(fmaxf(c_pen[c].x, 1.0e-30f) / scale[2 * size_t(k) + 1]);
    }
}

// ---- joint preparation (geometry_constraints.slang:1391-1417, 1462-1475) --
// State, dense ids and initial rows; one thread per joint.  Body ids are
// {dense index, 1}.
__global__ void k_joint_init(SolveArgs a) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nj) return;
    PmhRigidConstraint &j = a.joints[k];
    // Links stay as they were for a joint without rows this step (the
    // reference rewrites them only for listed joints).
    j.avbd_count = 0u;
    if (j.alive == 0u) return;
    j.state_enabled = j.options.enabled != 0u && j.state_broken == 0u ? 1u : 0u;
    if (j.state_enabled == 0u) return;
    j.applied_impulse = 0.0f;
    const PmhBodyId ia = j.options.body_a, ib = j.options.body_b;
    j.dense_a = ia.index < a.n && ia.generation == 1u ? ia.index : kInvalid;
    j.dense_b = ib.index < a.n && ib.generation == 1u ? ib.index : kInvalid;
    if (j.dense_a == kInvalid || j.dense_b == kInvalid) return;
    Block scratch;
    block_zero(scratch);
    visit_joint(a, j, scratch, true, kVisitInitialize);
}

// The joints' iteration floor: the legacy profile without a requested
// count raises the scene budget to the largest joint iteration count
// (:1466-1470); one atomic per wavefront.
__global__ void k_joint_budget(SolveArgs a, StepCounters *ctr) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t want = 0;
    if (k < a.nj && a.joints[k].avbd_count != 0u) want = a.joints[k].options.solver_iterations;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) want = max(want, uint32_t(__shfl_xor(want, d, 64)));
    if ((threadIdx.x & 63) == 0 && want) {
        atomicMax(&ctr->joint_budget, want);
        atomicMax(&a.schedule->budget, want);
    }
}

// Per-body lists of the joints with rows (geometry_constraints.slang:
// 1462-1475): each body's candidate list, already in descending joint
// index, filtered in order.  Pass 0 counts, pass 1 writes at the scanned
// offsets.  One thread per body; bodies with long lists go to the
// wavefront kernel below.
template <int kPass>
__global__ void k_joint_lists(SolveArgs a, const uint32_t *jlist_start, const uint32_t *jlist, uint32_t *jdeg,
                              uint32_t *jadj) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t b = jlist_start[i], e = jlist_start[i + 1];
    if (e - b > 64u) return;
    uint32_t o = kPass == 1 ? a.jstart[i] : 0u;
    for (uint32_t p = b; p < e; ++p) {
        const uint32_t k = jlist[p];
        if (a.joints[k].avbd_count == 0u) continue;
        if (kPass == 1) jadj[o] = k;
        ++o;
    }
    if (kPass == 0) jdeg[i] = o;
}

template <int kPass>
__global__ void __launch_bounds__(64) k_joint_lists_heavy(SolveArgs a, const uint32_t *heavy, uint32_t nheavy,
                                                          const uint32_t *jlist_start, const uint32_t *jlist,
                                                          uint32_t *jdeg, uint32_t *jadj) {
    const uint32_t lane = threadIdx.x;
    for (uint32_t h = blockIdx.x; h < nheavy; h += gridDim.x) {
        const uint32_t i = heavy[h];
        const uint32_t b = jlist_start[i], e = jlist_start[i + 1];
        uint32_t o = kPass == 1 ? a.jstart[i] : 0u;
        for (uint32_t base = b; base < e; base += 64) {
            const uint32_t p = base + lane;
            const bool keep = p < e && a.joints[jlist[p]].avbd_count != 0u;
            const uint64_t mask = __ballot(keep);
            if (kPass == 1 && keep) jadj[o + __popcll(mask & ((1ull << lane) - 1ull))] = jlist[p];
            o += uint32_t(__popcll(mask));
        }
        if (kPass == 0 && lane == 0) jdeg[i] = o;
    }
}

// List links of the reference layout (push-front in joint order, so each
// body's list runs in descending joint index) from the sorted segments.
__global__ void k_joint_links(SolveArgs a, uint32_t *joint_head) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t b = a.jstart[i], e = a.jstart[i + 1];
    joint_head[i] = b < e ? a.jadj[b] : kInvalid;
    for (uint32_t p = b; p < e; ++p) {
        PmhRigidConstraint &j = a.joints[a.jadj[p]];
        const uint32_t next = p + 1 < e ? a.jadj[p + 1] : kInvalid;
        if (j.dense_a == i) j.avbd_next_a = next;
        else j.avbd_next_b = next;
    }
}

__global__ void k_joint_hook(SolveArgs a, uint32_t *parent) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nj) return;
    const PmhRigidConstraint &j = a.joints[k];
    if (j.avbd_count == 0u) return;
    if (a.P[j.dense_a].motion != kDynamic || a.P[j.dense_b].motion != kDynamic) return;
    if (ld_relaxed(&parent[j.dense_a]) == ld_relaxed(&parent[j.dense_b])) return;   // already joined
    uf_union(parent, j.dense_a, j.dense_b);
}

__global__ void k_joint_anchor(SolveArgs a, const uint32_t *parent, uint32_t *anchored) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nj) return;
    const PmhRigidConstraint &j = a.joints[k];
    if (j.avbd_count == 0u) return;
    const bool da = a.P[j.dense_a].motion == kDynamic, db = a.P[j.dense_b].motion == kDynamic;
    if (da && !db) atomicOr(&anchored[parent[j.dense_a]], 1u);
    if (db && !da) atomicOr(&anchored[parent[j.dense_b]], 1u);
}

// rigid_avbd_project_hinge_anchors (solve.slang:174-209): after the pose
// phase, each enabled hinge moves its anchors together, split by inverse
// mass, one joint after the other in index order.  Here the hinges run by
// levels of a schedule built at upload (below).
__device__ void project_hinge(const SolveArgs &a, const PmhRigidConstraint &j) {
    const uint32_t ia = j.dense_a, ib = j.dense_b;
    const float ima = a.P[ia].motion == kDynamic ? a.P[ia].inverse_mass : 0.0f;
    const float imb = a.P[ib].motion == kDynamic ? a.P[ib].inverse_mass : 0.0f;
    const float sum = ima + imb;
    if (sum <= kEpsilon) return;
    const PmhRigidBodyState sa = a.states[ia], sb = a.states[ib];
    const V3 pa = add(sv(sa.position), rotate(sq(sa.orientation), sv(j.options.local_anchor_a)));
    const V3 pb = add(sv(sb.position), rotate(sq(sb.orientation), sv(j.options.local_anchor_b)));
    const V3 e = sub(pb, pa);
    a.states[ia].position = pv(add(sv(sa.position), mul(e, ima / sum)));
    a.states[ib].position = pv(sub(sv(sb.position), mul(e, imb / sum)));
}

constexpr uint32_t kHingeLevelLaunches = 64;   // deeper schedules take the one-block walk

// One thread per hinge of one schedule level: the hinges of a level share
// no body, so their projections commute; the levels run in order, which
// gives every hinge the bodies as the serial pass in index order leaves them.
__global__ void k_project_hinge_level(SolveArgs a, const uint32_t *list, uint32_t begin, uint32_t end) {
    const uint32_t t = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= end) return;
    const PmhRigidConstraint &j = a.joints[list[t]];
    if (j.avbd_count == 0u || j.state_enabled == 0u) return;
    project_hinge(a, j);
}

// Deep schedules: one block walks the levels, with an agent-scope fence
// (writes done, stale cache lines dropped) and a block barrier between them.
__global__ void __launch_bounds__(1024) k_project_hinge_levels(SolveArgs a, const uint32_t *list,
                                                               const uint32_t *level_start, uint32_t levels) {
    for (uint32_t l = 0; l < levels; ++l) {
        for (uint32_t t = level_start[l] + threadIdx.x; t < level_start[l + 1]; t += blockDim.x) {
            const PmhRigidConstraint &j = a.joints[list[t]];
            if (j.avbd_count != 0u && j.state_enabled != 0u) project_hinge(a, j);
        }
        __threadfence();
        __syncthreads();
    }
}

__global__ void k_residual_reset(ResidualLog *log) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < 2 * kResidualSlots * kResidualBanks;
         k += gridDim.x * blockDim.x) {
        const uint32_t slot = k / kResidualBanks, bank = k % kResidualBanks;
        log->bits[slot][bank] = 0u;
        log->sum[slot][bank] = 0ull;
        log->count[slot][bank] = 0u;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) log->executed[0] = log->executed[1] = 0u;
}

__global__ void k_schedule_init(PmhContactSchedule *s, StepCounters *ctr, uint32_t budget, uint32_t profile,
                                uint32_t base_iterations, uint32_t reset_status) {
    if (threadIdx.x != 0) return;
    if (reset_status) ctr->status = 0;
    ctr->color_abort = 0;
    PmhContactSchedule v;
    memset(&v, 0, sizeof v);
    v.budget = budget;
    v.offsets[0] = profile;
    v.offsets[1] = base_iterations;
    *s = v;
    ctr->manifolds = 0;
    ctr->colors = 0;
    ctr->islands = 0;
    ctr->impact = 0;
    ctr->color_ticket = 0;
    ctr->sweep_ticket[0] = ctr->sweep_ticket[1] = 0;
    ctr->dyn_bodies = ctr->impact_bodies = 0;
    ctr->mean_roots = 0;
    ctr->joint_budget = 0;
    ctr->cold_points = 0;
    ctr->m1 = ctr->m2 = ctr->p1 = ctr->p2 = 0;
    ctr->m1i = ctr->m2i = ctr->p1i = ctr->p2i = 0;
}

__global__ void k_schedule_finish(PmhContactSchedule *s, const StepCounters *ctr, uint32_t blocks) {
    if (threadIdx.x != 0) return;
    s->remaining = ctr->colors;
    s->island_count = ctr->islands;
    s->counts[0] = ctr->impact;
    s->candidates = ctr->pairs;
    s->contacts = ctr->manifolds;
    s->blocks = blocks;
}

PmhResult report_sweep_failure(hipStream_t s, StepCounters *ctr) {
    static const uint32_t zero = 0u;
    PMH_CHECK(hipMemcpyAsync(&ctr->sweep_failed, &zero, sizeof zero, hipMemcpyHostToDevice, s));
    PmhResult r = sync_stream(s);
    return r != PMH_SUCCESS ? r : PMH_ERROR_INVALID_STATE;
}

void build_joint_lists(const PmhRigidConstraint *joints, uint32_t nj, uint32_t n, JointLists &out) {
    auto dense = [&](PmhBodyId id) { return id.index < n && id.generation == 1u ? id.index : kInvalid; };
    out.start.assign(n + 1, 0u);
    for (uint32_t k = 0; k < nj; ++k) {
        const PmhRigidConstraint &j = joints[k];
        const uint32_t a = dense(j.options.body_a), b = dense(j.options.body_b);
        if (j.alive == 0u || a == kInvalid || b == kInvalid) continue;
        out.start[a]++;
        out.start[b]++;
    }
    uint32_t total = 0;
    for (uint32_t i = 0; i <= n; ++i) { const uint32_t c = out.start[i]; out.start[i] = total; total += c; }
    out.list.assign(std::max<uint32_t>(1, total), 0u);
    std::vector<uint32_t> cursor(out.start.begin(), out.start.end());
    for (uint32_t k = nj; k-- > 0;) {
        const PmhRigidConstraint &j = joints[k];
        const uint32_t a = dense(j.options.body_a), b = dense(j.options.body_b);
        if (j.alive == 0u || a == kInvalid || b == kInvalid) continue;
        out.list[cursor[a]++] = k;
        out.list[cursor[b]++] = k;
    }
    out.heavy.clear();
    for (uint32_t i = 0; i < n; ++i)
        if (out.start[i + 1] - out.start[i] > 64u) out.heavy.push_back(i);
    std::vector<uint32_t> body_level(n, 0u), hinge_level, hinge_ids;
    uint32_t levels = 0;
    for (uint32_t k = 0; k < nj; ++k) {
        const PmhRigidConstraint &j = joints[k];
        const uint32_t a = dense(j.options.body_a), b = dense(j.options.body_b);
        if (j.alive == 0u || j.options.type != kJointHinge || a == kInvalid || b == kInvalid) continue;
        const uint32_t level = std::max(body_level[a], body_level[b]) + 1u;
        body_level[a] = body_level[b] = level;
        hinge_ids.push_back(k);
        hinge_level.push_back(level);
        levels = std::max(levels, level);
    }
    std::vector<uint32_t> level_start(levels + 2u, 0u);
    out.hinges.assign(hinge_ids.size(), 0u);
    for (uint32_t level : hinge_level) level_start[level + 1u]++;
    for (uint32_t l = 1; l <= levels + 1u; ++l) level_start[l] += level_start[l - 1];
    std::vector<uint32_t> cur(level_start.begin(), level_start.end());
    for (size_t h = 0; h < hinge_ids.size(); ++h) out.hinges[cur[hinge_level[h]]++] = hinge_ids[h];
    out.level_start.assign(level_start.begin() + 1, level_start.end());
}

PmhResult status_result(uint32_t status) {
    if (status & (kStatusPairOverflow | kStatusPointOverflow | kStatusColorOverflow)) return PMH_ERROR_CAPACITY_EXCEEDED;
    if (status & kStatusColorTimeout) return PMH_ERROR_INVALID_STATE;
    return PMH_SUCCESS;
}

PmhResult sync_stream(hipStream_t s) {
    static const double limit = [] {
        const char *e = getenv("PMH_SYNC_TIMEOUT_S");
        const double v = e ? atof(e) : 0.0;
        return v > 0.0 ? v : 120.0;
    }();
    const auto start = std::chrono::steady_clock::now();
    for (uint32_t polls = 0;; ++polls) {
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) return PMH_SUCCESS;
        if (e != hipErrorNotReady) return hip_result(e, "hipStreamQuery");
        if (polls >= 64) {
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count() > limit) {
                (void)hip_result(hipErrorNotReady, "device did not finish the step within PMH_SYNC_TIMEOUT_S");
                return PMH_ERROR_DEVICE_LOST;
            }
            std::this_thread::yield();
        }
    }
}

PmhResult sync_event(hipEvent_t ev) {
    static const double limit = [] {
        const char *e = getenv("PMH_SYNC_TIMEOUT_S");
        const double v = e ? atof(e) : 0.0;
        return v > 0.0 ? v : 120.0;
    }();
    const auto start = std::chrono::steady_clock::now();
    for (uint32_t polls = 0;; ++polls) {
        const hipError_t e = hipEventQuery(ev);
        if (e == hipSuccess) return PMH_SUCCESS;
        if (e != hipErrorNotReady) return hip_result(e, "hipEventQuery");
        if (polls >= 64) {
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count() > limit) {
                (void)hip_result(hipErrorNotReady, "device did not finish the step within PMH_SYNC_TIMEOUT_S");
                return PMH_ERROR_DEVICE_LOST;
            }
            std::this_thread::yield();
        }
    }
}

static int g_cus = 0;
int device_compute_units() {
    if (!g_cus) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) g_cus = prop.multiProcessorCount;
        if (!g_cus) g_cus = 256;
    }
    return g_cus;
}

static uint32_t grid_for(uint32_t items, uint32_t block = 256, uint32_t cap = 4096) {
    uint32_t g = (items + block - 1) / block;
    return std::max<uint32_t>(1, std::min(g, cap));
}

PmhResult KernelTimer::collect() {
    for (uint32_t j = 0; j < pending; ++j) {
        float t = 0.0f;
        PMH_CHECK(hipEventElapsedTime(&t, events[2 * j], events[2 * j + 1]));
        ms += t;
    }
    launches += pending;
    if (pending) bytes += pending_bytes;
    pending = 0;
    pending_bytes = 0.0;
    return PMH_SUCCESS;
}

void KernelTimer::release() {
    for (hipEvent_t e : events) if (e) (void)hipEventDestroy(e);
    events.clear();
    pending = 0;
}

PmhResult Solver::reserve_joints(uint32_t count) {
    PmhResult r = PMH_SUCCESS;
    if ((r = jadj.reserve(2 * size_t(std::max<uint32_t>(1, count)))) != PMH_SUCCESS) return r;
    return PMH_SUCCESS;
}

// Every device buffer of the solver with its element count for `bodies`
// bodies, `pairs` pairs and `points` contact points.
template <class F>
PmhResult Solver::each_buffer(uint32_t bodies, uint32_t pairs, uint32_t points, F &&f) {
    PmhResult r = PMH_SUCCESS;
    // Colour-ordered positions: the bodies and at most 7 gap slots per
    // colour (8 bodies per wavefront at 8 lanes each).
    const uint32_t positions = bodies + kMaxColors * 7u;
#define R(buf, cnt) if ((r = f(buf, size_t(cnt))) != PMH_SUCCESS) return r
    R(tpos, bodies); R(tq, bodies); R(tv, bodies); R(tw, bodies); R(disp, 2 * (size_t)bodies);
    R(dq, bodies); R(inertia, 2 * (size_t)bodies); R(color, bodies); R(parent, bodies);
    R(impact, bodies); R(deg, positions + 1); R(adj_start, bodies + 1); R(adj_cursor, bodies);
    R(color_bodies, positions); R(bp_start, positions + 1); R(anchored, bodies); R(mean, bodies);
    R(members, bodies); R(mean_list, bodies);
    R(jdeg, bodies + 1); R(jstart, bodies + 1); R(jcursor, bodies); R(joint_head, bodies);
    R(prev_vel, bodies); R(done, bodies);
    R(code, pairs); R(ab, pairs); R(m_count, pairs + 1); R(m_start, pairs + 1); R(m_slot, pairs + 1);
    R(m_impact, pairs); R(m_face, pairs); R(m_init_rel, pairs); R(adj, 2 * (size_t)pairs);
    R(tri_list, pairs); R(face_list, pairs);
    R(c_np, points); R(c_pt, points); R(c_aa, points); R(c_ab, points); R(c_err, points);
    R(c_lam, points); R(c_pen, points); R(c_acc, points); R(c_rep, points); R(c_flags, points);
    R(c_pair, points); R(jac, 9 * (size_t)points); R(bp_entries, 4 * (size_t)points);
    R(erec, 18 * (size_t)points); R(c_entry, 2 * (size_t)points); R(residual, 1);
    R(schedule, 1); R(color_count, kMaxColors + 1); R(color_start, kMaxColors + 1);
    R(counters, 1); R(totals, 1);
#undef R
    return PMH_SUCCESS;
}

PmhResult Solver::reserve(uint32_t bodies, uint32_t pairs, uint32_t points) {
    // Zeroed when first allocated: the velocity history, the published
    // sweep stamps, the counters and the running totals.
    const bool fresh_prev = prev_vel.n < bodies, fresh_done = done.n < bodies;
    const bool fresh_ctr = !counters.p, fresh_totals = !totals.p;
    PmhResult r = each_buffer(bodies, pairs, points, [](auto &buf, size_t count) { return buf.reserve(count); });
    if (r != PMH_SUCCESS) return r;
    if (fresh_prev) PMH_CHECK(hipMemset(prev_vel.p, 0, sizeof(float4) * bodies));
    if (fresh_done) PMH_CHECK(hipMemset(done.p, 0, sizeof(uint32_t) * bodies));
    if (fresh_ctr) PMH_CHECK(hipMemset(counters.p, 0, sizeof(StepCounters)));
    if (fresh_totals) PMH_CHECK(hipMemset(totals.p, 0, sizeof(SolveTotals)));
    ecap = 2 * (size_t)points;
    n = bodies;
    slots = bodies + kMaxColors * 7u;
    pair_cap = pairs;
    point_cap = points;
    return PMH_SUCCESS;
}

// Scan partials of the largest exclusive scan over these counts.
static uint32_t scan_blocks_for(uint32_t bodies, uint32_t pairs) {
    const uint32_t count = std::max(bodies + kMaxColors * 7u, pairs);
    return (count + kScanTile - 1) / kScanTile + 1;
}

size_t Solver::carve_bytes(uint32_t bodies, uint32_t pairs, uint32_t points) {
    Solver probe;
    size_t used = 0;
    auto count = [&](auto &buf, size_t cnt) {
        using T = std::remove_reference_t<decltype(*buf.p)>;
        used += (std::max<size_t>(cnt, 1) * sizeof(T) + 255) & ~size_t(255);
        return PMH_SUCCESS;
    };
    (void)probe.each_buffer(bodies, pairs, points, count);
    (void)count(probe.scan_tmp, scan_blocks_for(bodies, pairs));
    return used;
}

PmhResult Solver::carve(uint8_t *base, size_t bytes, uint32_t bodies, uint32_t pairs, uint32_t points,
                        StepCounters *host_counters, uint32_t *host_starts) {
    size_t used = 0;
    auto take = [&](auto &buf, size_t cnt) {
        using T = std::remove_reference_t<decltype(*buf.p)>;
        const size_t need = (std::max<size_t>(cnt, 1) * sizeof(T) + 255) & ~size_t(255);
        if (used + need > bytes) return PMH_ERROR_CAPACITY_EXCEEDED;
        buf.borrow(reinterpret_cast<T *>(base + used), cnt);
        used += need;
        return PMH_SUCCESS;
    };
    PmhResult r = each_buffer(bodies, pairs, points, take);
    if (r == PMH_SUCCESS) r = take(scan_tmp, scan_blocks_for(bodies, pairs));
    if (r != PMH_SUCCESS) return r;
    host_ctr = host_counters;
    host_color_start = host_starts;
    host_owned = false;
    ecap = 2 * (size_t)points;
    return use(bodies, pairs, points);
}

PmhResult Solver::use(uint32_t bodies, uint32_t pairs, uint32_t points) {
    if (size_t(bodies) > tpos.n || size_t(pairs) > code.n || size_t(points) > c_np.n)
        return PMH_ERROR_CAPACITY_EXCEEDED;
    n = bodies;
    slots = bodies + kMaxColors * 7u;
    pair_cap = pairs;
    point_cap = points;
    return PMH_SUCCESS;
}

void Solver::release() {
    tpos.release(); tq.release(); tv.release(); tw.release(); disp.release(); dq.release();
    inertia.release(); prev_vel.release(); color.release(); parent.release(); impact.release();
    deg.release(); adj_start.release(); adj_cursor.release(); color_bodies.release();
    bp_start.release(); anchored.release(); mean.release(); members.release(); mean_list.release(); done.release();
    jdeg.release(); jstart.release(); jcursor.release(); joint_head.release(); jadj.release(); code.release(); ab.release();
    m_count.release(); m_start.release(); m_slot.release(); m_impact.release(); m_face.release();
    m_init_rel.release(); adj.release(); tri_list.release(); face_list.release(); c_np.release(); c_pt.release(); c_aa.release();
    c_ab.release(); c_err.release(); c_lam.release(); c_pen.release(); c_acc.release();
    c_rep.release(); c_flags.release(); c_pair.release(); jac.release(); bp_entries.release();
    erec.release(); c_entry.release(); residual.release();
    cold_list.release(); cold_feat.release(); cold_scale.release();
    schedule.release(); color_count.release(); color_start.release(); counters.release(); totals.release();
    scan_tmp.release();
    if (host_owned && host_ctr) (void)hipHostFree(host_ctr);
    if (host_owned && host_color_start) (void)hipHostFree(host_color_start);
    if (host_exec) (void)hipHostFree(host_exec);
    for (hipEvent_t &e : exec_events) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    host_exec = nullptr;
    host_ctr = nullptr;
    host_color_start = nullptr;
    host_owned = true;
}

static uint32_t g_coop_blocks = 0;

// Resident blocks of a dataflow-sweep kernel over the device (at most 8 per
// CU), found once per kernel.
static uint32_t sweep_blocks(const void *kernel, uint32_t threads) {
    static std::mutex lock;
    static std::vector<std::pair<const void *, uint32_t>> known;
    std::lock_guard<std::mutex> guard(lock);
    for (const auto &k : known) if (k.first == kernel) return k.second;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, int(threads), 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const uint32_t blocks = uint32_t(std::min(per_cu, 8)) * uint32_t(device_compute_units());
    known.emplace_back(kernel, blocks);
    return blocks;
}

PmhResult avbd_prepare_solve(hipStream_t s, Solver &v, const PmhBodyParameters *P,
                             const PmhRigidBodyState *previous, PmhRigidBodyState *states,
                             const SolveParams &sp, SolveTiming *timing) {
    const uint32_t n = v.n;
    const float dt = sp.dt;
    const bool realtime = sp.profile == kRealtimeProfile;
    const uint32_t budget = realtime ? kRealtimeSweeps
                                     : (sp.requested_iterations ? sp.requested_iterations : sp.base_iterations);
    StepCounters *ctr = v.counters.p;
    const uint32_t *npairs = &ctr->pairs;
    const uint32_t *npts = &ctr->points;
    // One thread per body or joint: these grids are never capped.
    const uint32_t bg = grid_for(n, 256, 0xffffffffu);
    const uint32_t pg = grid_for(v.pair_cap, 256, 2048);
    const uint32_t cg_ = grid_for(v.point_cap, 256, 4096);

    SolveArgs a;
    a.n = n;
    a.P = P;
    a.states = states;
    a.tpos = v.tpos.p; a.tq = v.tq.p; a.tv = v.tv.p; a.tw = v.tw.p; a.disp = v.disp.p; a.dq = v.dq.p;
    a.inertia = v.inertia.p;
    a.color_bodies = v.color_bodies.p; a.color_start = v.color_start.p; a.bp_start = v.bp_start.p;
    a.entries = reinterpret_cast<const uint2 *>(v.bp_entries.p);
    a.jac = v.jac.p;
    a.erec = v.erec.p;
    a.ecap = v.ecap;
    a.c_entry = v.c_entry.p;
    a.c_lam = v.c_lam.p; a.c_pen = v.c_pen.p; a.c_aa = v.c_aa.p; a.c_ab = v.c_ab.p; a.c_acc = v.c_acc.p;
    a.c_rep = v.c_rep.p; a.c_err = v.c_err.p; a.c_np = v.c_np.p; a.c_pt = v.c_pt.p;
    a.c_pair = v.c_pair.p; a.ab = v.ab.p;
    a.parent = v.parent.p; a.impact = v.impact.p; a.deg = v.adj_cursor.p; a.anchored = v.anchored.p;
    a.members = v.members.p; a.mean_list = v.mean_list.p;
    a.mean = v.mean.p;
    a.schedule = v.schedule.p;
    a.ctr = ctr;
    a.totals = v.totals.p;
    a.npts = npts; a.npairs = npairs;
    a.dt = dt;
    a.budget = budget;
    a.realtime = realtime ? 1u : 0u;
    a.joints = v.joints;
    a.nj = v.nj;
    a.jstart = v.jstart.p;
    a.jadj = v.jadj.p;
    a.prev = previous;
    a.color = v.color.p;
    a.done = v.done.p;
    a.flow_stamp = 0;
    a.flow_limit = sp.sweep_wait_limit;
    a.log = v.residual.p;
    a.tol = sp.tolerance;
    a.c_flags = v.c_flags.p;
    const uint32_t jg = grid_for(std::max<uint32_t>(1, v.nj), 256, 0xffffffffu);
    k_schedule_init<<<1, 64, 0, s>>>(v.schedule.p, ctr, budget, sp.profile, sp.base_iterations,
                                     sp.reset_status ? 1u : 0u);
    k_residual_reset<<<32, 256, 0, s>>>(v.residual.p);
    k_body_init<<<bg, 256, 0, s>>>(n, P, previous, states, dt, sp.gravity, v.tpos.p, v.tq.p, v.tv.p, v.tw.p,
                                   v.disp.p, v.dq.p, v.inertia.p, v.prev_vel.p, v.color.p, v.parent.p,
                                   v.impact.p, v.deg.p, v.anchored.p, v.adj_cursor.p, v.members.p, v.mean.p,
                                   v.jdeg.p, v.jcursor.p);
    k_contact_init<<<pg, 256, 0, s>>>(npairs, P, previous, states, v.ab.p, v.m_count.p, v.m_start.p, v.c_np.p,
                                      v.c_pt.p, v.c_aa.p, v.c_ab.p, v.c_err.p, v.c_lam.p, v.c_pen.p, v.c_flags.p,
                                      v.m_impact.p, v.c_entry.p, dt, sp.profile);
    const uint32_t cold_cap = uint32_t(v.cold_list.n);
    const bool warm = (v.mlp && sp.warm_infer) || v.log;
    if (warm && cold_cap) {
        k_warm_features<<<cg_, 256, 0, s>>>(npts, P, v.c_pair.p, v.ab.p, v.m_count.p, v.m_face.p, v.c_np.p, v.c_pt.p,
                                            v.c_err.p, v.c_aa.p, v.c_flags.p, sp.gravity, dt, cold_cap,
                                            v.cold_list.p, v.cold_feat.p, v.cold_scale.p, ctr);
        if (v.mlp && sp.warm_infer) {
            PmhResult rr = mlp_infer(s, &ctr->cold_points, v.cold_list.p, v.cold_feat.p, v.cold_scale.p, v.mlp,
                                     v.c_lam.p, v.c_pen.p, cold_cap, v.mlp_policy);
            if (rr != PMH_SUCCESS) return rr;
        }
    }
    if (v.nj) k_joint_init<<<jg, 256, 0, s>>>(a);
    // Counting passes over the pairs: a smaller grid, so fewer blocks add
    // to the step's counters.
    const uint32_t count_g = std::min<uint32_t>(pg, 2u * uint32_t(device_compute_units()));
    k_degree<<<count_g, 256, 0, s>>>(npairs, P, v.ab.p, v.m_count.p, v.deg.p, nullptr, ctr);
    PmhResult r = exclusive_scan(s, v.deg.p, v.adj_start.p, n, nullptr, nullptr, v.scan_tmp);
    if (r != PMH_SUCCESS) return r;
    k_adj_fill<<<pg, 256, 0, s>>>(npairs, P, v.ab.p, v.m_count.p, v.adj_start.p, v.adj_cursor.p, v.adj.p);
    k_adj_sort<<<bg, 256, 0, s>>>(n, v.adj_start.p, v.adj.p);
    if (v.nj) {
        // Joint lists, geometry_constraints.slang:1462-1475.
        if (!realtime && sp.requested_iterations == 0u) k_joint_budget<<<jg, 256, 0, s>>>(a, ctr);
        const uint32_t hg = std::max<uint32_t>(1, std::min<uint32_t>(v.nheavy, 4096));
        k_joint_lists<0><<<bg, 256, 0, s>>>(a, v.jlist_start, v.jlist, v.jdeg.p, v.jadj.p);
        if (v.nheavy)
            k_joint_lists_heavy<0><<<hg, 64, 0, s>>>(a, v.jheavy, v.nheavy, v.jlist_start, v.jlist, v.jdeg.p,
                                                     v.jadj.p);
        r = exclusive_scan(s, v.jdeg.p, v.jstart.p, n, nullptr, nullptr, v.scan_tmp);
        if (r != PMH_SUCCESS) return r;
        k_joint_lists<1><<<bg, 256, 0, s>>>(a, v.jlist_start, v.jlist, v.jdeg.p, v.jadj.p);
        if (v.nheavy)
            k_joint_lists_heavy<1><<<hg, 64, 0, s>>>(a, v.jheavy, v.nheavy, v.jlist_start, v.jlist, v.jdeg.p,
                                                     v.jadj.p);
        k_joint_links<<<bg, 256, 0, s>>>(a, v.joint_head.p);
    }
    k_uf_sample<<<bg, 256, 0, s>>>(n, P, v.ab.p, v.adj_start.p, v.adj.p, v.parent.p);
    k_uf_compress<<<bg, 256, 0, s>>>(n, v.parent.p);
    k_uf_hook<<<pg, 256, 0, s>>>(npairs, P, v.ab.p, v.m_count.p, v.parent.p);
    if (v.nj) k_joint_hook<<<jg, 256, 0, s>>>(a, v.parent.p);
    k_uf_compress<<<bg, 256, 0, s>>>(n, v.parent.p);
    k_impact_roots<<<pg, 256, 0, s>>>(npairs, P, v.ab.p, v.m_impact.p, v.m_count.p, v.parent.p, v.impact.p,
                                      v.anchored.p, ctr);
    if (v.nj) k_joint_anchor<<<jg, 256, 0, s>>>(a, v.parent.p, v.anchored.p);
    k_impact_bodies<<<bg, 256, 0, s>>>(n, P, v.parent.p, v.impact.p, v.deg.p, v.jdeg.p, v.anchored.p, v.members.p,
                                       v.mean_list.p, ctr);
    if (!realtime)
        k_count_impact_pairs<<<count_g, 256, 0, s>>>(npairs, P, v.ab.p, v.m_count.p, v.impact.p, ctr);
    // Per-lane tickets when bodies have few neighbours (the previous step's
    // contact and joint ends per body), one wavefront per body otherwise.
    if (v.neighbour_hint < kLaneColorNeighbours) {
        const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((n + 255) / 256, 4 * device_compute_units()));
        k_color_lane<<<blocks, 256, 0, s>>>(n, P, v.ab.p, v.adj_start.p, v.adj.p, v.color.p, v.jstart.p, v.jadj.p,
                                            v.nj ? v.joints : nullptr, ctr, sp.color_wait_limit);
    } else {
        const uint32_t blocks = std::max<uint32_t>(
            1, std::min<uint32_t>((n + kColorWaveBlock - 1) / kColorWaveBlock, 2 * device_compute_units()));
        k_color_wave<<<blocks, 64 * kColorWaveBlock, 0, s>>>(n, P, v.ab.p, v.adj_start.p, v.adj.p, v.color.p, v.jstart.p, v.jadj.p,
                                            v.nj ? v.joints : nullptr, ctr, sp.color_wait_limit);
    }
    PMH_CHECK(hipMemsetAsync(v.color_count.p, 0, sizeof(uint32_t) * (kMaxColors + 1), s));
    k_color_count<<<bg, 256, 0, s>>>(n, v.color.p, v.color_count.p);
    k_color_scan<<<1, 64, 0, s>>>(v.color_count.p, v.color_start.p, ctr, v.color_bodies.p, 64u / sp.lanes);
    k_color_place<<<bg, 256, 0, s>>>(n, v.color.p, v.color_start.p, v.color_count.p, v.color_bodies.p);
    // colour-ordered point lists (deg array reused as the per-position count)
    // The degree array is free again once the adjacency is built (the
    // fill cursor keeps each body's degree): it holds the point counts.
    const uint32_t sg = grid_for(v.slots, 256, 0xffffffffu);
    k_bp_count<<<sg, 256, 0, s>>>(v.slots, ctr, v.color_start.p, v.color_bodies.p, v.adj_start.p, v.adj.p,
                                  v.m_count.p, v.deg.p);
    r = exclusive_scan(s, v.deg.p, v.bp_start.p, v.slots, nullptr, nullptr, v.scan_tmp);
    if (r != PMH_SUCCESS) return r;
    k_bp_fill<<<sg, 256, 0, s>>>(ctr, v.color_start.p, v.color_bodies.p, v.adj_start.p, v.adj.p, v.m_count.p,
                                 v.m_start.p, v.ab.p, v.bp_start.p, reinterpret_cast<uint2 *>(v.bp_entries.p),
                                 v.c_entry.p, v.color.p);
    k_jacobians<<<cg_, 256, 0, s>>>(npts, P, v.c_pair.p, v.ab.p, v.tpos.p, v.tq.p, v.c_np.p, v.c_pt.p, v.c_aa.p,
                                    v.c_ab.p, v.c_err.p, v.c_acc.p, v.jac.p, dt, 0, v.impact.p, v.erec.p, v.ecap,
                                    v.c_entry.p);
    PMH_CHECK(hipGetLastError());


    if (sp.cooperative && v.nj == 0 && sp.tolerance <= 0.0f) {
        if (timing && timing->start) PMH_CHECK(hipEventRecord(timing->start, s));
        const void *kernel = sp.lanes == 64 ? reinterpret_cast<const void *>(k_solve_all<64>)
                           : sp.lanes == 32 ? reinterpret_cast<const void *>(k_solve_all<32>)
                           : sp.lanes == 8  ? reinterpret_cast<const void *>(k_solve_all<8>)
                                            : reinterpret_cast<const void *>(k_solve_all<16>);
        if (!g_coop_blocks) {
            int per_cu = 0;
            PMH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_solve_all<64>, 256, 0));
            // Two resident blocks per CU at most: more blocks only add
            // grid-barrier arrivals.
            g_coop_blocks = std::max(1, std::min(per_cu, 2)) * device_compute_units();
        }
        uint32_t blocks = g_coop_blocks;
        const uint32_t *cf = v.c_flags.p;
        void *args[] = {&a, &cf};
        PMH_CHECK(hipLaunchCooperativeKernel(kernel, dim3(blocks), dim3(256), args, 0, s));
        if (timing) timing->launches += 1;
        k_schedule_finish<<<1, 64, 0, s>>>(v.schedule.p, ctr, blocks);
    } else {
        // One launch per stage: a kernel boundary on the stream is a far
        // cheaper grid-wide barrier than an in-kernel one.  The colour
        // schedule comes back to the host once per step.
        if (!v.host_ctr) {
            PMH_CHECK(hipHostMalloc(reinterpret_cast<void **>(&v.host_ctr), sizeof(StepCounters)));
            PMH_CHECK(hipHostMalloc(reinterpret_cast<void **>(&v.host_color_start), sizeof(uint32_t) * (kMaxColors + 1)));
        }
        // Iterating to a residual, the host stops enqueueing iterations
        // once a phase has converged (every iteration after that exits at
        // once on the device, so the results are those of enqueueing the
        // whole budget).  A solver without allocation enqueues the budget.
        const bool chunked = sp.tolerance > 0.0f && v.host_owned;
        if (chunked && !v.host_exec) {
            PMH_CHECK(hipHostMalloc(reinterpret_cast<void **>(&v.host_exec), 2 * sizeof(uint32_t)));
            for (hipEvent_t &e : v.exec_events) PMH_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        PMH_CHECK(hipMemcpyAsync(v.host_ctr, ctr, sizeof(StepCounters), hipMemcpyDeviceToHost, s));
        PMH_CHECK(hipMemcpyAsync(v.host_color_start, v.color_start.p, sizeof(uint32_t) * (kMaxColors + 1),
                                 hipMemcpyDeviceToHost, s));
        PmhResult sr = sync_stream(s);
        if (sr != PMH_SUCCESS) return sr;
        const StepCounters host = *v.host_ctr;
        // Overflowed pools or a colouring that gave up: no solve on a
        // partial schedule; the caller keeps the pre-step states.
        if (host.status & kStatusFatal) return status_result(host.status);
        // A dataflow sweep of an earlier step that ran out of its wait
        // (that step kept its states): reported once.
        if (host.sweep_failed) return report_sweep_failure(s, ctr);
        v.neighbour_hint = 2.0f * float(uint64_t(host.manifolds) + v.nj) / float(std::max<uint32_t>(1, n));
        const uint32_t lanes = sp.lanes;
        if (!realtime && sp.requested_iterations == 0u) a.budget = std::max(budget, host.joint_budget);
        const uint32_t budget = a.budget;
        KernelTimer *kt = timing ? timing->kernels : nullptr;
        bool time_colors = false;
        if (kt && kt->stride) {
            PmhResult r = kt->collect();   // the previous step's launches are complete
            if (r != PMH_SUCCESS) return r;
            time_colors = (kt->step++ % kt->stride) == 0;
        }
        uint32_t timed = 0;
        if (timing && timing->start) PMH_CHECK(hipEventRecord(timing->start, s));
        const uint32_t colors = std::min(host.colors, kMaxColors);
        const uint32_t pts_blocks = std::max<uint32_t>(1, (std::max(host.points, v.nj) + 255) / 256);
        const uint32_t joint_blocks = std::max<uint32_t>(1, (v.nj + 255) / 256);
        const uint32_t phases = realtime ? 1u : (host.impact ? 2u : 1u);
        const bool joints = v.nj != 0;
        // A timed launch goes through hipExtLaunchKernel, which writes the
        // dispatch's own start and end timestamps into the two events: no
        // separate event commands on the stream around the kernel.
        auto launch_color = [&](uint32_t blocks, uint32_t c, uint32_t phase, hipEvent_t start,
                                hipEvent_t stop) -> hipError_t {
            uint32_t b = v.host_color_start[c], e = v.host_color_start[c + 1];
            const void *kernel;
            if (joints) {
                kernel = lanes == 64   ? reinterpret_cast<const void *>(&k_stage_color<64, true>)
                         : lanes == 32 ? reinterpret_cast<const void *>(&k_stage_color<32, true>)
                         : lanes == 8  ? reinterpret_cast<const void *>(&k_stage_color<8, true>)
                                       : reinterpret_cast<const void *>(&k_stage_color<16, true>);
            } else {
                kernel = lanes == 64   ? reinterpret_cast<const void *>(&k_stage_color<64, false>)
                         : lanes == 32 ? reinterpret_cast<const void *>(&k_stage_color<32, false>)
                         : lanes == 8  ? reinterpret_cast<const void *>(&k_stage_color<8, false>)
                                       : reinterpret_cast<const void *>(&k_stage_color<16, false>);
            }
            void *args[] = {&a, &b, &e, &phase};
            if (start) return hipExtLaunchKernel(kernel, dim3(blocks), dim3(256), args, 0, s, start, stop, 0);
            return hipLaunchKernel(kernel, dim3(blocks), dim3(256), args, 0, s);
        };
        // Dataflow sweep: one launch per iteration, as many resident
        // wavefronts as fit (tickets hand out the positions).
        const bool flow = sp.dataflow;
        const uint32_t positions = v.host_color_start[colors];
        auto launch_sweep = [&](SolveArgs &af, uint32_t phase, uint32_t it, hipEvent_t start,
                                hipEvent_t stop) -> hipError_t {
            const void *kernel;
            if (joints) {
                kernel = lanes == 64   ? reinterpret_cast<const void *>(&k_sweep<64, true>)
                         : lanes == 32 ? reinterpret_cast<const void *>(&k_sweep<32, true>)
                         : lanes == 8  ? reinterpret_cast<const void *>(&k_sweep<8, true>)
                                       : reinterpret_cast<const void *>(&k_sweep<16, true>);
            } else {
                kernel = lanes == 64   ? reinterpret_cast<const void *>(&k_sweep<64, false>)
                         : lanes == 32 ? reinterpret_cast<const void *>(&k_sweep<32, false>)
                         : lanes == 8  ? reinterpret_cast<const void *>(&k_sweep<8, false>)
                                       : reinterpret_cast<const void *>(&k_sweep<16, false>);
            }
            const uint32_t block_waves = joints ? kSweepJointWaves : kSweepWaves;
            const uint32_t slots = (positions + (64 / lanes) - 1) / (64 / lanes);
            const uint32_t blocks = std::max<uint32_t>(
                1, std::min<uint32_t>((slots + block_waves - 1) / block_waves, sweep_blocks(kernel, 64 * block_waves)));
            void *args[] = {&af, &phase, &it};
            const dim3 block(64 * block_waves);
            if (start) return hipExtLaunchKernel(kernel, dim3(blocks), block, args, 0, s, start, stop, 0);
            return hipLaunchKernel(kernel, dim3(blocks), block, args, 0, s);
        };
        uint32_t launches = 0;
        // Every kCheckChunk iterations the phase's executed count is copied
        // to pinned memory; the host then waits for the previous chunk's
        // copy (the device works on this chunk meanwhile) and stops once
        // the phase has converged: at most two chunks run empty.
        constexpr uint32_t kCheckChunk = 4;
        auto check_chunk = [&](uint32_t phase, uint32_t it, uint32_t &chunk, bool &done) -> PmhResult {
            done = false;
            if (!chunked || (it + 1) % kCheckChunk != 0 || it + 1 >= budget) return PMH_SUCCESS;
            const uint32_t slot = chunk & 1u;
            PMH_CHECK(hipMemcpyAsync(&v.host_exec[slot], &v.residual.p->executed[phase], sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, s));
            PMH_CHECK(hipEventRecord(v.exec_events[slot], s));
            if (chunk++ > 0) {
                PmhResult r = sync_event(v.exec_events[slot ^ 1u]);
                if (r != PMH_SUCCESS) return r;
                done = v.host_exec[slot ^ 1u] != 0u;
            }
            return PMH_SUCCESS;
        };
        for (uint32_t phase = 0; phase < phases; ++phase) {
            uint32_t chunk = 0;
            if (phase != 0) {
                k_stage_impact_gaps<<<pts_blocks, 256, 0, s>>>(a);
                k_stage_impact_bodies<<<bg, 256, 0, s>>>(a);
                if (joints) k_stage_velocity_errors<true><<<pts_blocks, 256, 0, s>>>(a, v.c_flags.p);
                else k_stage_velocity_errors<false><<<pts_blocks, 256, 0, s>>>(a, v.c_flags.p);
                k_stage_velocity_jacobians<<<pts_blocks, 256, 0, s>>>(a);
                launches += 4;
            }
            for (uint32_t it = 0; it < budget; ++it) {
                if (flow) {
                    SolveArgs af = a;
                    if (++v.sweep_stamp == 0u) {   // stamps wrapped: no stale stamp may match
                        PMH_CHECK(hipMemsetAsync(v.done.p, 0, sizeof(uint32_t) * n, s));
                        v.sweep_stamp = 1u;
                    }
                    af.flow_stamp = v.sweep_stamp;
                    hipEvent_t e0 = nullptr, e1 = nullptr;
                    if (time_colors) {
                        if (kt->events.size() < 2 * size_t(timed + 1)) {
                            const size_t old = kt->events.size();
                            kt->events.resize(2 * size_t(timed + 1) + 64, nullptr);
                            for (size_t e = old; e < kt->events.size(); ++e) PMH_CHECK(hipEventCreate(&kt->events[e]));
                        }
                        e0 = kt->events[2 * timed];
                        e1 = kt->events[2 * timed + 1];
                        ++timed;
                    }
                    PMH_CHECK(launch_sweep(af, phase, it, e0, e1));
                    launches += 1;
                    if (joints) {
                        k_stage_joint_duals<<<joint_blocks, 256, 0, s>>>(a, phase, it);
                        launches += 1;
                    }
                    if (sp.tolerance > 0.0f) {
                        k_stage_check<<<1, 64, 0, s>>>(a, phase, it);
                        launches += 1;
                    }
                    bool converged = false;
                    PmhResult cr = check_chunk(phase, it, chunk, converged);
                    if (cr != PMH_SUCCESS) return cr;
                    if (converged) break;
                    continue;
                }
                for (uint32_t c = 0; c < colors; ++c) {
                    const uint32_t bodies = v.host_color_start[c + 1] - v.host_color_start[c];
                    const uint32_t per_wave = 64 / lanes;
                    const uint32_t waves = (bodies + per_wave - 1) / per_wave;
                    const uint32_t blocks = std::max<uint32_t>(1, (waves + 3) / 4);
                    if (time_colors) {
                        if (kt->events.size() < 2 * size_t(timed + 1)) {
                            const size_t old = kt->events.size();
                            kt->events.resize(2 * size_t(timed + 1) + 64, nullptr);
                            for (size_t e = old; e < kt->events.size(); ++e) PMH_CHECK(hipEventCreate(&kt->events[e]));
                        }
                        PMH_CHECK(launch_color(blocks, c, phase, kt->events[2 * timed], kt->events[2 * timed + 1]));
                        ++timed;
                    } else {
                        PMH_CHECK(launch_color(blocks, c, phase, nullptr, nullptr));
                    }
                }
                if (joints) k_stage_duals<true><<<pts_blocks, 256, 0, s>>>(a, phase, it);
                else k_stage_duals<false><<<pts_blocks, 256, 0, s>>>(a, phase, it);
                launches += colors + 1;
                if (sp.tolerance > 0.0f) {
                    k_stage_check<<<1, 64, 0, s>>>(a, phase, it);
                    launches += 1;
                }
                bool converged = false;
                PmhResult cr = check_chunk(phase, it, chunk, converged);
                if (cr != PMH_SUCCESS) return cr;
                if (converged) break;
            }
            if (sp.tolerance > 0.0f) {
                k_stage_publish<<<pts_blocks, 256, 0, s>>>(a, phase);
                launches += 1;
            }
            if (host.mean_roots) {
                k_stage_mean<<<(host.mean_roots + 3) / 4, 256, 0, s>>>(a, phase);
                launches += 1;
            }
            k_stage_apply<<<bg, 256, 0, s>>>(a, phase);
            launches += 1;
            if (phase == 0 && v.hinges) {
                const std::vector<uint32_t> &ls = v.hinge_level_start;
                const uint32_t levels = uint32_t(ls.size()) - 1u;
                if (levels <= kHingeLevelLaunches) {
                    for (uint32_t l = 0; l < levels; ++l) {
                        const uint32_t count = ls[l + 1] - ls[l];
                        if (count) k_project_hinge_level<<<(count + 255) / 256, 256, 0, s>>>(a, v.hinge_list, ls[l], ls[l + 1]);
                    }
                    launches += levels;
                } else {
                    k_project_hinge_levels<<<1, 1024, 0, s>>>(a, v.hinge_list, v.hinge_level_dev, levels);
                    launches += 1;
                }
            }
        }
        if (timing) timing->launches += launches;
        if (time_colors) {
            // Algorithmic bytes of the colour sweeps: per iteration 136 B per
            // body, 84 B per contact point and dynamic endpoint, 64 B per
            // manifold and dynamic endpoint (pair record + other body); the
            // dataflow sweep adds the dual update's (iteration_bytes).
            auto sweep = [flow](double d, double m1, double m2, double p1, double p2) {
                if (flow)   // the dataflow sweep also updates the duals: the whole iteration
                    return 136.0 * d + (84.0 * 2 + 24.0) * p1 + (84.0 * 3 + 24.0) * p2 + (64.0 + 104.0) * m1 +
                           (128.0 + 104.0) * m2;
                return 136.0 * d + 84.0 * (p1 + 2.0 * p2) + 64.0 * (m1 + 2.0 * m2);
            };
            double bytes = budget * sweep(host.dyn_bodies, host.m1, host.m2, host.p1, host.p2);
            if (phases > 1) bytes += budget * sweep(host.impact_bodies, host.m1i, host.m2i, host.p1i, host.p2i);
            kt->pending = timed;
            kt->pending_bytes = bytes;
        }
        k_schedule_finish<<<1, 64, 0, s>>>(v.schedule.p, ctr, 0);
    }
    if (v.log && cold_cap)
        k_warm_log<<<grid_for(cold_cap), 256, 0, s>>>(ctr, cold_cap, v.cold_list.p, v.cold_feat.p, v.cold_scale.p,
                                                      v.c_lam.p, v.c_pen.p, v.log, v.log_cap);
    if (timing && timing->stop) PMH_CHECK(hipEventRecord(timing->stop, s));
    PMH_CHECK(hipGetLastError());
    return PMH_SUCCESS;
}

}  // namespace pmh
