// SPDX-License-Identifier: MIT
// Kernel seam: prepare_geometry_constraints + solve_geometry_constraints
// (src/simulation/geometry_constraints_dispatch.hpp:15-43) over the
// reference's ContactManifold / AvbdBody / ContactSchedule records in
// device memory.  The records are unpacked into the solver's
// structure-of-arrays form, solved, and the solver-owned fields are written
// back (duals, errors, anchors, impulses, list links; the AvbdBody record).
#include "pmh_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace pmh {

__global__ void k_seam_counts(uint32_t M, const PmhContactManifold *man, uint32_t *m_count) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m < M) m_count[m] = std::min(man[m].count, 8u);
}

__global__ void k_seam_in(uint32_t n, uint32_t M, const uint32_t *pairs, const PmhContactManifold *man,
                          const uint32_t *m_start, uint32_t *code, uint2 *ab, const uint32_t *m_count,
                          float4 *m_init_rel, uint32_t *m_face, float4 *c_np, float4 *c_pt, float4 *c_aa,
                          float4 *c_ab, float4 *c_err, float4 *c_lam, float4 *c_pen, float4 *c_acc, float4 *c_rep,
                          uint32_t *c_flags, uint32_t *c_pair, StepCounters *ctr) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m == 0) {
        ctr->pairs = M;
        ctr->points = m_start[M];
    }
    if (m >= M) return;
    const uint32_t p = pairs[m];
    code[m] = p;
    ab[m] = make_uint2(p / n, p % n);
    const PmhContactManifold &mm = man[m];
    m_init_rel[m] = make_float4(mm.initial_relative_position.x, mm.initial_relative_position.y,
                                mm.initial_relative_position.z, 0.0f);
    m_face[m] = mm.face_patch;
    for (uint32_t k = 0; k < m_count[m]; ++k) {
        const PmhContact &t = mm.contacts[k];
        const uint32_t c = m_start[m] + k;
        c_np[c] = make_float4(t.normal.x, t.normal.y, t.normal.z, t.penetration);
        c_pt[c] = make_float4(t.point.x, t.point.y, t.point.z, t.initial_normal_speed);
        c_aa[c] = make_float4(t.avbd_anchor_a.x, t.avbd_anchor_a.y, t.avbd_anchor_a.z, t.impact_fraction);
        c_ab[c] = make_float4(t.avbd_anchor_b.x, t.avbd_anchor_b.y, t.avbd_anchor_b.z, 0.0f);
        c_err[c] = make_float4(t.avbd_error.x, t.avbd_error.y, t.avbd_error.z, 0.0f);
        c_lam[c] = make_float4(t.avbd_dual[0].lambda, t.avbd_dual[1].lambda, t.avbd_dual[2].lambda, 0.0f);
        c_pen[c] = make_float4(t.avbd_dual[0].penalty, t.avbd_dual[1].penalty, t.avbd_dual[2].penalty, 0.0f);
        c_acc[c] = make_float4(t.accumulated_normal_impulse, t.accumulated_friction_impulse.x,
                               t.accumulated_friction_impulse.y, t.accumulated_friction_impulse.z);
        c_rep[c] = make_float4(t.reported_normal_impulse, t.reported_friction_impulse.x,
                               t.reported_friction_impulse.y, t.reported_friction_impulse.z);
        // Point flags: 1 cached, 2 persistent, 8 guided (an axially guided
        // member against a static body, rigid_avbd_guided_contact,
        // geometry_constraints.slang:327-333).
        const bool guided = (t.body_hinge.axial != 0u && t.collider_hinge.static_body != 0u) ||
                            (t.collider_hinge.axial != 0u && t.body_hinge.static_body != 0u);
        c_flags[c] = (t.avbd_cached ? 1u : 0u) | (t.persistent ? 2u : 0u) | (guided ? 8u : 0u);
        c_pair[c] = m;
    }
}

__global__ void k_seam_history_in(uint32_t n, const PmhAvbdBody *bodies, float4 *prev_vel) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const PmhAvbdBody &b = bodies[i];
    prev_vel[i] = make_float4(b.previous_velocity.x, b.previous_velocity.y, b.previous_velocity.z,
                              b.history_valid ? 1.0f : 0.0f);
}

__global__ void k_seam_out(uint32_t M, PmhContactManifold *man, const uint32_t *m_count, const uint32_t *m_start,
                           const float4 *c_pt, const float4 *c_aa, const float4 *c_ab, const float4 *c_err,
                           const float4 *c_lam, const float4 *c_pen, const float4 *c_acc, const float4 *c_rep) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    PmhContactManifold &mm = man[m];
    for (uint32_t k = 0; k < m_count[m]; ++k) {
        const uint32_t c = m_start[m] + k;
        PmhContact &t = mm.contacts[k];
        t.initial_normal_speed = c_pt[c].w;
        t.avbd_anchor_a = {c_aa[c].x, c_aa[c].y, c_aa[c].z};
        t.avbd_anchor_b = {c_ab[c].x, c_ab[c].y, c_ab[c].z};
        t.avbd_error = {c_err[c].x, c_err[c].y, c_err[c].z};
        t.avbd_dual[0] = {c_lam[c].x, c_pen[c].x};
        t.avbd_dual[1] = {c_lam[c].y, c_pen[c].y};
        t.avbd_dual[2] = {c_lam[c].z, c_pen[c].z};
        t.accumulated_normal_impulse = c_acc[c].x;
        t.accumulated_friction_impulse = {c_acc[c].y, c_acc[c].z, c_acc[c].w};
        t.reported_normal_impulse = c_rep[c].x;
        t.reported_friction_impulse = {c_rep[c].y, c_rep[c].z, c_rep[c].w};
    }
}

// Bodies and the contact list links (heads and next pointers of dynamic
// endpoints, in the reference's descending active order).
__global__ void k_seam_bodies_out(uint32_t n, PmhAvbdBody *bodies, PmhContactManifold *man, const uint2 *ab,
                                  const float4 *tpos, const float4 *tq, const float4 *tv, const float4 *tw,
                                  const float4 *disp, const float4 *dq, const float4 *inertia, const float4 *prev_vel,
                                  const uint32_t *color, const uint32_t *parent, const uint32_t *impact,
                                  const uint32_t *adj_start, const uint32_t *adj, const uint32_t *joint_head) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PmhAvbdBody b;
    memset(&b, 0, sizeof b);
    b.target.position = {tpos[i].x, tpos[i].y, tpos[i].z};
    b.target.orientation = {tq[i].x, tq[i].y, tq[i].z, tq[i].w};
    b.target.linear_velocity = {tv[i].x, tv[i].y, tv[i].z};
    b.target.angular_velocity = {tw[i].x, tw[i].y, tw[i].z};
    b.orientation_delta = {dq[i].x, dq[i].y, dq[i].z, dq[i].w};
    b.displacement[0] = disp[2 * i].x; b.displacement[1] = disp[2 * i].y; b.displacement[2] = disp[2 * i].z;
    b.displacement[3] = disp[2 * i + 1].x; b.displacement[4] = disp[2 * i + 1].y; b.displacement[5] = disp[2 * i + 1].z;
    const float4 i0 = inertia[2 * i], i1 = inertia[2 * i + 1];
    b.inertia_h[0] = b.inertia_h[2] = b.inertia_h[5] = i0.x;
    b.inertia_h[9] = i0.y; b.inertia_h[13] = i0.z; b.inertia_h[14] = i0.w;
    b.inertia_h[18] = i1.x; b.inertia_h[19] = i1.y; b.inertia_h[20] = i1.z;
    b.joint_head = joint_head ? joint_head[i] : kInvalid;
    b.color = color[i];
    b.parent = parent[i];
    b.impact = impact[i];
    b.previous_velocity = {prev_vel[i].x, prev_vel[i].y, prev_vel[i].z};
    b.history_valid = prev_vel[i].w != 0.0f ? 1u : 0u;
    const uint32_t s = adj_start[i], e = adj_start[i + 1];
    b.contact_head = s < e ? adj[s] : kInvalid;
    for (uint32_t k = s; k < e; ++k) {
        const uint32_t m = adj[k];
        const uint32_t next = k + 1 < e ? adj[k + 1] : kInvalid;
        if (ab[m].x == i) man[m].avbd_next_a = next;
        else man[m].avbd_next_b = next;
    }
    bodies[i] = b;
}

// The seam over a solver whose buffers hold at least `body_count` bodies,
// `M` pairs and 8 M points.
static PmhResult seam_run(Solver &v, hipStream_t s, const PmhBodyParameters *parameters,
                          const PmhRigidBodyState *previous, PmhRigidBodyState *states, uint32_t body_count,
                          PmhContactManifold *manifolds, const uint32_t *active_pairs, uint32_t M,
                          PmhAvbdBody *bodies, PmhContactSchedule *schedule, float timestep, const float gravity[3],
                          uint32_t base_iterations, uint32_t requested_iterations, uint32_t profile_version) {
    const uint32_t mg = std::max<uint32_t>(1, (M + 255) / 256), bg = (body_count + 255) / 256;
    if (M) k_seam_counts<<<mg, 256, 0, s>>>(M, manifolds, v.m_count.p);
    PmhResult r = exclusive_scan(s, v.m_count.p, v.m_start.p, M, nullptr, nullptr, v.scan_tmp);
    if (r != PMH_SUCCESS) return r;
    k_seam_in<<<mg, 256, 0, s>>>(body_count, M, active_pairs, manifolds, v.m_start.p, v.code.p, v.ab.p, v.m_count.p,
                                 v.m_init_rel.p, v.m_face.p, v.c_np.p, v.c_pt.p, v.c_aa.p, v.c_ab.p, v.c_err.p,
                                 v.c_lam.p, v.c_pen.p, v.c_acc.p, v.c_rep.p, v.c_flags.p, v.c_pair.p, v.counters.p);
    k_seam_history_in<<<bg, 256, 0, s>>>(body_count, bodies, v.prev_vel.p);
    SolveParams sp{timestep, {gravity[0], gravity[1], gravity[2]}, base_iterations, requested_iterations,
                   profile_version, 0u, kDefaultLanes};
    sp.reset_status = true;
    r = avbd_prepare_solve(s, v, parameters, previous, states, sp, nullptr);
    if (r != PMH_SUCCESS) return r;
    if (M)
        k_seam_out<<<mg, 256, 0, s>>>(M, manifolds, v.m_count.p, v.m_start.p, v.c_pt.p, v.c_aa.p, v.c_ab.p,
                                      v.c_err.p, v.c_lam.p, v.c_pen.p, v.c_acc.p, v.c_rep.p);
    k_seam_bodies_out<<<bg, 256, 0, s>>>(body_count, bodies, manifolds, v.ab.p, v.tpos.p, v.tq.p, v.tv.p, v.tw.p,
                                         v.disp.p, v.dq.p, v.inertia.p, v.prev_vel.p, v.color.p, v.parent.p,
                                         v.impact.p, v.adj_start.p, v.adj.p, v.nj ? v.joint_head.p : nullptr);
    PMH_CHECK(hipMemcpyAsync(schedule, v.schedule.p, sizeof(PmhContactSchedule), hipMemcpyDeviceToDevice, s));
    PMH_CHECK(hipStreamSynchronize(s));
    PMH_CHECK(hipGetLastError());
    return PMH_SUCCESS;
}

static bool seam_args_valid(const PmhBodyParameters *parameters, const PmhRigidBodyState *previous,
                            PmhRigidBodyState *states, uint32_t body_count, PmhAvbdBody *bodies,
                            PmhContactSchedule *schedule, float timestep, const float gravity[3],
                            uint32_t profile_version, const uint32_t *active_pair_count) {
    return parameters && previous && states && bodies && schedule && gravity && timestep > 0.0f && body_count != 0 &&
           profile_version <= 1 && active_pair_count;
}

// ---- joints at the seam -----------------------------------------------------
// The caller's RigidConstraintResource records name their bodies by id;
// `ids` maps dense index -> id (rigid_avbd_find_dense,
// geometry_constraints.slang:318-325).  The solve runs on a copy whose body
// ids are rewritten to {dense index, 1}; afterwards the solver-owned fields
// (state, applied impulse, dense ids, rows, list links) go back to the
// caller's records, whose options stay as they were.
struct SeamJoints {
    DBuf<PmhRigidConstraint> copy;
    DBuf<uint32_t> list_start, list, heavy, hinges, levels;
    std::vector<PmhRigidConstraint> host;
    std::vector<PmhBodyId> host_ids;
};

__global__ void k_seam_joints_out(uint32_t nj, PmhRigidConstraint *dst, const PmhRigidConstraint *src) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nj) return;
    PmhRigidConstraint &d = dst[k];
    const PmhRigidConstraint &s = src[k];
    d.state_enabled = s.state_enabled;
    d.state_broken = s.state_broken;
    d.applied_impulse = s.applied_impulse;
    d.dense_a = s.dense_a;
    d.dense_b = s.dense_b;
    for (uint32_t r = 0; r < 18u; ++r) d.rows[r] = s.rows[r];
    d.avbd_count = s.avbd_count;
    d.avbd_next_a = s.avbd_next_a;
    d.avbd_next_b = s.avbd_next_b;
}

// Uploads the translated copy and its schedule and points the solver at them.
static PmhResult seam_joints_in(Solver &v, SeamJoints &sj, hipStream_t s, const PmhRigidConstraint *constraints,
                                uint32_t capacity, const PmhBodyId *ids, uint32_t n) {
    sj.host.resize(capacity);
    sj.host_ids.resize(n);
    PMH_CHECK(hipMemcpyAsync(sj.host.data(), constraints, sizeof(PmhRigidConstraint) * capacity,
                             hipMemcpyDeviceToHost, s));
    PMH_CHECK(hipMemcpyAsync(sj.host_ids.data(), ids, sizeof(PmhBodyId) * n, hipMemcpyDeviceToHost, s));
    PMH_CHECK(hipStreamSynchronize(s));
    std::unordered_map<uint64_t, uint32_t> dense;
    dense.reserve(2 * size_t(n));
    for (uint32_t i = 0; i < n; ++i)
        dense.emplace((uint64_t(sj.host_ids[i].index) << 32) | sj.host_ids[i].generation, i);
    auto resolve = [&](PmhBodyId &id) {
        const auto it = dense.find((uint64_t(id.index) << 32) | id.generation);
        id = it == dense.end() ? PmhBodyId{kInvalid, 0u} : PmhBodyId{it->second, 1u};
    };
    for (PmhRigidConstraint &j : sj.host) {
        resolve(j.options.body_a);
        resolve(j.options.body_b);
    }
    JointLists jl;
    build_joint_lists(sj.host.data(), capacity, n, jl);
    PmhResult r;
    if ((r = sj.copy.reserve(std::max<uint32_t>(1, capacity))) || (r = sj.list_start.reserve(n + 1)) ||
        (r = sj.list.reserve(jl.list.size())) || (r = sj.heavy.reserve(std::max<size_t>(1, jl.heavy.size()))) ||
        (r = sj.hinges.reserve(std::max<size_t>(1, jl.hinges.size()))) ||
        (r = sj.levels.reserve(std::max<size_t>(1, jl.level_start.size()))) ||
        (r = v.jadj.reserve(std::max<size_t>(1, jl.list.size()))))
        return r;
    PMH_CHECK(hipMemcpyAsync(sj.copy.p, sj.host.data(), sizeof(PmhRigidConstraint) * capacity, hipMemcpyHostToDevice, s));
    PMH_CHECK(hipMemcpyAsync(sj.list_start.p, jl.start.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice, s));
    PMH_CHECK(hipMemcpyAsync(sj.list.p, jl.list.data(), sizeof(uint32_t) * jl.list.size(), hipMemcpyHostToDevice, s));
    if (!jl.heavy.empty())
        PMH_CHECK(hipMemcpyAsync(sj.heavy.p, jl.heavy.data(), sizeof(uint32_t) * jl.heavy.size(), hipMemcpyHostToDevice, s));
    if (!jl.hinges.empty())
        PMH_CHECK(hipMemcpyAsync(sj.hinges.p, jl.hinges.data(), sizeof(uint32_t) * jl.hinges.size(), hipMemcpyHostToDevice, s));
    if (!jl.level_start.empty())
        PMH_CHECK(hipMemcpyAsync(sj.levels.p, jl.level_start.data(), sizeof(uint32_t) * jl.level_start.size(),
                                 hipMemcpyHostToDevice, s));
    PMH_CHECK(hipStreamSynchronize(s));   // the host vectors stay valid until the copies are done
    v.joints = sj.copy.p;
    v.nj = capacity;
    v.jlist_start = sj.list_start.p;
    v.jlist = sj.list.p;
    v.jheavy = sj.heavy.p;
    v.nheavy = uint32_t(jl.heavy.size());
    v.hinge_list = sj.hinges.p;
    v.hinge_level_start = jl.level_start;
    v.hinge_level_dev = sj.levels.p;
    v.hinges = !jl.hinges.empty();
    return PMH_SUCCESS;
}

static void seam_joints_clear(Solver &v) {
    v.joints = nullptr;
    v.nj = 0;
    v.jlist_start = v.jlist = v.jheavy = nullptr;
    v.nheavy = 0;
    v.hinge_list = v.hinge_level_dev = nullptr;
    v.hinge_level_start.clear();
    v.hinges = false;
}

static std::mutex g_seam_mutex;
static Solver *g_seam_solver = nullptr;
static SeamJoints *g_seam_joints = nullptr;

// ---- the seam without allocation ------------------------------------------
// Every buffer of the solve is a slice of one file-scope device array per
// GPU, sized for kStaticBodies bodies, kStaticPairs manifolds and 8 points
// per manifold; the host read-back records are file-scope host arrays.
constexpr uint32_t kStaticBodies = 4096, kStaticPairs = 4096, kStaticPoints = 8 * kStaticPairs;
constexpr size_t kStaticScratchBytes = size_t(24) << 20;
constexpr int kStaticDevices = 64;
static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

// Clears the carved counters and totals (a fill kernel, so the call records
// no memset node where a caller captures the stream).
__global__ void k_seam_clear(uint32_t *words, uint32_t count) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) words[k] = 0u;
}

struct StaticSeam {
    uint8_t *base = nullptr;             // the device's address of `scratch`
    bool ready = false;
    Solver solver;
    StepCounters host_counters;
    uint32_t host_starts[kMaxColors + 1];
};
static StaticSeam g_static_seam[kStaticDevices];

}  // namespace pmh

extern "C" PmhResult pmhPrepareSolveGeometryConstraints(
    void *stream, const PmhBodyParameters *parameters, const PmhRigidBodyState *previous, PmhRigidBodyState *states,
    uint32_t body_count, PmhContactManifold *manifolds, const uint32_t *active_pairs,
    const uint32_t *active_pair_count, PmhAvbdBody *bodies, PmhContactSchedule *schedule, float timestep,
    const float gravity[3], uint32_t base_iterations, uint32_t requested_iterations, uint32_t profile_version) {
    using namespace pmh;
    if (!seam_args_valid(parameters, previous, states, body_count, bodies, schedule, timestep, gravity,
                         profile_version, active_pair_count))
        return PMH_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(g_seam_mutex);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    uint32_t M = 0;
    PMH_CHECK(hipMemcpyAsync(&M, active_pair_count, sizeof M, hipMemcpyDeviceToHost, s));
    PMH_CHECK(hipStreamSynchronize(s));
    if (M && (!manifolds || !active_pairs)) return PMH_ERROR_INVALID_ARGUMENT;
    if (!g_seam_solver) g_seam_solver = new Solver();
    Solver &v = *g_seam_solver;
    PmhResult r = v.reserve(body_count, std::max<uint32_t>(M, 1), std::max<uint32_t>(8 * M, 8));
    if (r != PMH_SUCCESS) return r;
    return seam_run(v, s, parameters, previous, states, body_count, manifolds, active_pairs, M, bodies, schedule,
                    timestep, gravity, base_iterations, requested_iterations, profile_version);
}

extern "C" PmhResult pmhPrepareSolveGeometryConstraintsJoints(
    void *stream, const PmhBodyParameters *parameters, const PmhRigidBodyState *previous, PmhRigidBodyState *states,
    uint32_t body_count, PmhContactManifold *manifolds, const uint32_t *active_pairs,
    const uint32_t *active_pair_count, PmhRigidConstraint *constraints, uint32_t constraint_capacity,
    const PmhBodyId *ids, PmhAvbdBody *bodies, PmhContactSchedule *schedule, float timestep,
    const float gravity[3], uint32_t base_iterations, uint32_t requested_iterations, uint32_t profile_version) {
    using namespace pmh;
    if (!seam_args_valid(parameters, previous, states, body_count, bodies, schedule, timestep, gravity,
                         profile_version, active_pair_count) ||
        (constraint_capacity && (!constraints || !ids)))
        return PMH_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(g_seam_mutex);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    uint32_t M = 0;
    PMH_CHECK(hipMemcpyAsync(&M, active_pair_count, sizeof M, hipMemcpyDeviceToHost, s));
    PMH_CHECK(hipStreamSynchronize(s));
    if (M && (!manifolds || !active_pairs)) return PMH_ERROR_INVALID_ARGUMENT;
    if (!g_seam_solver) g_seam_solver = new Solver();
    if (!g_seam_joints) g_seam_joints = new SeamJoints();
    Solver &v = *g_seam_solver;
    PmhResult r = v.reserve(body_count, std::max<uint32_t>(M, 1), std::max<uint32_t>(8 * M, 8));
    if (r != PMH_SUCCESS) return r;
    if (constraint_capacity) {
        r = seam_joints_in(v, *g_seam_joints, s, constraints, constraint_capacity, ids, body_count);
        if (r != PMH_SUCCESS) {
            seam_joints_clear(v);
            return r;
        }
    }
    r = seam_run(v, s, parameters, previous, states, body_count, manifolds, active_pairs, M, bodies, schedule,
                 timestep, gravity, base_iterations, requested_iterations, profile_version);
    if (r == PMH_SUCCESS && constraint_capacity) {
        k_seam_joints_out<<<(constraint_capacity + 255) / 256, 256, 0, s>>>(constraint_capacity, constraints,
                                                                             g_seam_joints->copy.p);
        const hipError_t e = hipGetLastError();
        r = e != hipSuccess ? hip_result(e, "k_seam_joints_out") : sync_stream(s);
    }
    seam_joints_clear(v);
    return r;
}

extern "C" PmhResult pmhPrepareSolveGeometryConstraintsNoAlloc(
    void *stream, const PmhBodyParameters *parameters, const PmhRigidBodyState *previous, PmhRigidBodyState *states,
    uint32_t body_count, PmhContactManifold *manifolds, const uint32_t *active_pairs,
    const uint32_t *active_pair_count, PmhAvbdBody *bodies, PmhContactSchedule *schedule, float timestep,
    const float gravity[3], uint32_t base_iterations, uint32_t requested_iterations, uint32_t profile_version) {
    using namespace pmh;
    if (!seam_args_valid(parameters, previous, states, body_count, bodies, schedule, timestep, gravity,
                         profile_version, active_pair_count))
        return PMH_ERROR_INVALID_ARGUMENT;
    if (body_count > kStaticBodies) {
        std::fprintf(stderr, "pmhPrepareSolveGeometryConstraintsNoAlloc: %u bodies, at most %u\n", body_count,
                     kStaticBodies);
        return PMH_ERROR_CAPACITY_EXCEEDED;
    }
    std::lock_guard<std::mutex> lock(g_seam_mutex);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    int dev = 0;
    PMH_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kStaticDevices) return PMH_ERROR_INVALID_ARGUMENT;
    StaticSeam &g = g_static_seam[dev];
    if (!g.base) {
        // The array's address on this device (host code never uses the
        // symbol itself as a device address).
        void *ptr = nullptr;
        PMH_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return PMH_ERROR_INVALID_STATE;
        g.base = static_cast<uint8_t *>(ptr);
    }
    if (!g.ready) {
        if (Solver::carve_bytes(kStaticBodies, kStaticPairs, kStaticPoints) > kStaticScratchBytes)
            return PMH_ERROR_CAPACITY_EXCEEDED;
        PmhResult r = g.solver.carve(g.base, kStaticScratchBytes, kStaticBodies, kStaticPairs, kStaticPoints,
                                     &g.host_counters, g.host_starts);
        if (r != PMH_SUCCESS) return r;
        // The published sweep stamps start below every stamp to come.
        k_seam_clear<<<std::max<uint32_t>(1, (kStaticBodies + 255) / 256), 256, 0, s>>>(g.solver.done.p,
                                                                                      kStaticBodies);
        g.ready = true;
    }
    uint32_t M = 0;
    PMH_CHECK(hipMemcpyAsync(&M, active_pair_count, sizeof M, hipMemcpyDeviceToHost, s));
    PMH_CHECK(hipStreamSynchronize(s));
    if (M && (!manifolds || !active_pairs)) return PMH_ERROR_INVALID_ARGUMENT;
    if (M > kStaticPairs) {
        std::fprintf(stderr, "pmhPrepareSolveGeometryConstraintsNoAlloc: %u active pairs, at most %u\n", M,
                     kStaticPairs);
        return PMH_ERROR_CAPACITY_EXCEEDED;
    }
    Solver &v = g.solver;
    PmhResult r = v.use(body_count, std::max<uint32_t>(M, 1), std::max<uint32_t>(8 * M, 8));
    if (r != PMH_SUCCESS) return r;
    // The array keeps the last call's contents: the counters and totals
    // start from zero.
    k_seam_clear<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t *>(v.counters.p), sizeof(StepCounters) / 4);
    k_seam_clear<<<1, 64, 0, s>>>(reinterpret_cast<uint32_t *>(v.totals.p), sizeof(SolveTotals) / 4);
    return seam_run(v, s, parameters, previous, states, body_count, manifolds, active_pairs, M, bodies, schedule,
                    timestep, gravity, base_iterations, requested_iterations, profile_version);
}
