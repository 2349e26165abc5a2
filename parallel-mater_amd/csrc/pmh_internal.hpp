// SPDX-License-Identifier: MIT
// Internal layout of the rigid step on one GPU.
//
// Storage is O(n + pairs + points): per-body arrays, per-active-pair
// arrays and per-contact-point arrays, all structure-of-arrays with 16-byte
// records so a wavefront's gathers touch whole cache lines.  Nothing is
// indexed by pair code, so 32k-body scenes need tens of MB, not the
// n(n-1)/2 manifold table of the reference (resource_plan.hpp:146-215).
#pragma once

#include "pmh_math.hpp"
#include "../../include/pm_hip_rigid.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace pmh {

#define PMH_CHECK(expr)                                                  \
    do {                                                                 \
        hipError_t pmh_e_ = (expr);                                      \
        if (pmh_e_ != hipSuccess) return pmh::hip_result(pmh_e_, #expr); \
    } while (0)

PmhResult hip_result(hipError_t e, const char *what);
const std::string &last_error();

// Shape kinds: triangle mesh (boxes and the reference's polyhedral
// sphere), analytic sphere and capsule.
enum : uint32_t { kShapeMesh = 0, kShapeSphere = 1, kShapeCapsule = 2 };

struct ShapeDev {
    uint32_t kind, vtx_off, vtx_count, idx_off;
    uint32_t idx_count, plane_off, has_planes, convex_face;  // convex_face: planes && <=96 indices
    float mn[3], bradius;
    float mx[3], sphere_radius;
    float center[3], half_length;
    float box_half[3];
    uint32_t is_box;   // centred axis-aligned 8-vertex box
    uint32_t node_off, node_count, leaf_off, leaf_count;   // BVH nodes and leaf-node list
    uint32_t shell_off, has_shell;                         // certified outward normal per triangle
    uint32_t unique_planes;                                // deduplicated planes after the per-triangle ones
};

// Median-split BVH node over the reordered triangles: leaves (count != 0)
// hold triangles [first, first + count); inner nodes their two children.
struct BvhDev {
    float mn[3];
    uint32_t first;
    float mx[3];
    uint32_t count;
    uint32_t left, right, pad0, pad1;
};

struct HostShape {
    ShapeDev info{};
    std::vector<V3> vertices;
    std::vector<uint32_t> indices;
    std::vector<float> planes;   // 4 per triangle plane, then the unique tail
    std::vector<BvhDev> nodes;   // preorder
    std::vector<uint32_t> leaves;
    std::vector<V3> shell;       // empty when nothing is certified
    uint32_t unique_planes = 0;
    float unit_inertia[3]{};
};

// Device status bits (accumulated across steps, read after a batch).
enum : uint32_t {
    kStatusPairOverflow = 1u << 0,
    kStatusPointOverflow = 1u << 1,
    kStatusColorOverflow = 1u << 2,
    kStatusInvalidShape = 1u << 3,
    kStatusColorTimeout = 1u << 4,
};

// Device-side counters of one step (one record, read back with stats).
struct StepCounters {
    uint32_t pairs;            // active pairs M
    uint32_t points;           // contact points P
    uint32_t manifolds;        // non-empty manifolds
    uint32_t colors;           // colour count
    uint32_t islands;
    uint32_t impact;
    uint32_t unsupported;      // round-shape pairs with no analytic path
    uint32_t cache_hits;
    uint32_t status;
    uint32_t large_count;      // bodies outside the hash grid
    uint32_t cache_count;      // saved pairs for the next step
    uint32_t color_ticket;     // work ticket of the colouring kernel
    float cell_size_bits;      // broadphase cell (as float)
    uint32_t barrier_count, barrier_gen;
    // Work of the step's solve: dynamic bodies, impact bodies, non-empty
    // manifolds and their points by number of dynamic endpoints (1 or 2),
    // over all pairs (phase 0) and over pairs touching an impact body
    // (velocity phase).
    uint32_t dyn_bodies, impact_bodies;
    uint32_t m1, m2, p1, p2, m1i, m2i, p1i, p2i;
    uint32_t point_alloc;      // narrowphase point-range allocator
    uint32_t tri_pairs;        // mesh pairs on the triangle-pair path
    uint32_t fallback_pairs;   // of those, solved by BVH traversal
    uint32_t mean_roots;       // unanchored islands with contacts (free-motion mean)
    uint32_t joint_budget;     // largest joint iteration count (legacy budget floor)
    uint32_t cold_points;      // cold contact points listed for the learned warm start
    uint32_t log_count;        // contact samples logged so far (learned warm start)
    uint32_t color_abort;      // set when a bounded wait (colouring or sweep) ran out: every other wait gives up
    uint32_t sweep_ticket[2];  // work tickets of the dataflow sweep, alternating per sweep
    uint32_t sweep_failed;     // a dataflow-sweep wait ran out; stays set until the host reports it
    uint32_t face_pairs;       // mesh pairs queued for the convex face path
};

// Running totals over every step since the last reset.
struct SolveTotals {
    unsigned long long steps, body_iterations, velocity_phases;
    double algorithmic_bytes;
};

constexpr uint32_t kMaxColors = 256;
// A body the colouring gave up on (a wait ran out) holds this value: not a
// colour, so no solve stage sees it, and published, so nobody waits on it.
constexpr uint32_t kColorAborted = kMaxColors;
constexpr uint32_t kColorWaitLimit = 1u << 22;
// Status bits that end a step with an error before its solve.
constexpr uint32_t kStatusFatal = kStatusPairOverflow | kStatusPointOverflow | kStatusColorOverflow |
                                  kStatusColorTimeout;
PmhResult status_result(uint32_t status);
// Clears the sweep-failure mark on the device and returns
// PMH_ERROR_INVALID_STATE (or the stream's own error).
PmhResult report_sweep_failure(hipStream_t s, StepCounters *ctr);
// Waits for the stream with a deadline (PMH_SYNC_TIMEOUT_S, default 120 s):
// a device that stops answering gives PMH_ERROR_DEVICE_LOST, not a caller
// blocked for ever.
PmhResult sync_stream(hipStream_t s);
// Waits for an event with the same deadline.
PmhResult sync_event(hipEvent_t e);
constexpr uint32_t kDefaultLanes = 32;   // lanes per body in the colour sweep
// Joint types (RigidConstraintType, src/simulation/types.hpp:662).
enum : uint32_t { kJointFixed = 0, kJointPoint = 1, kJointHinge = 2, kJointSlider = 3, kJointPiston = 4,
                  kJointGeneric = 5, kJointGenericSpring = 6, kJointMotor = 7 };
constexpr uint32_t kCacheFloat4PerPoint = 6;

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    bool borrowed = false;               // a slice of memory the buffer does not own (never freed or grown)
    PmhResult reserve(size_t count) {
        if (count <= n && p) return PMH_SUCCESS;
        if (borrowed) return PMH_ERROR_CAPACITY_EXCEEDED;
        if (p) { (void)hipFree(p); p = nullptr; n = 0; }
        size_t bytes = (count ? count : 1) * sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return hip_result(e, "hipMalloc");
        n = count;
        return PMH_SUCCESS;
    }
    void borrow(T *slice, size_t count) {
        release();
        p = slice;
        n = count;
        borrowed = true;
    }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr;
        n = 0;
        borrowed = false;
    }
};

// Constraint residual after each dual update of the step, 64 slots per
// phase: the maximum (float bits), the sum of squares in 2^-40 fixed point
// and the number of contributing points and joints; and the iterations
// each phase ran.
constexpr uint32_t kResidualSlots = 64;
// Per update slot, kResidualBanks partial sums: a wavefront commits to the
// bank of its block (blockIdx & 63), so the atomics of one update spread
// over 64 addresses per field instead of meeting at one; readers fold the
// banks (integer sums and maxima, so the fold is exact in any order).
constexpr uint32_t kResidualBanks = 64;
struct ResidualLog {
    unsigned long long sum[2 * kResidualSlots][kResidualBanks];
    uint32_t bits[2 * kResidualSlots][kResidualBanks];
    uint32_t count[2 * kResidualSlots][kResidualBanks];
    uint32_t executed[2];
    uint32_t pad[2];
};

// Folded totals of one slot.
struct ResidualTotal {
    unsigned long long sum = 0;
    uint32_t bits = 0, count = 0;
};
inline ResidualTotal residual_fold(const ResidualLog &log, uint32_t slot) {
    ResidualTotal t;
    for (uint32_t b = 0; b < kResidualBanks; ++b) {
        t.sum += log.sum[slot][b];
        t.count += log.count[slot][b];
        t.bits = log.bits[slot][b] > t.bits ? log.bits[slot][b] : t.bits;
    }
    return t;
}

// Square of a residual in 2^-40 fixed point, clamped at 2^44 (r = 4 m).
__host__ __device__ inline unsigned long long residual_square(float r) {
    const double q = double(r) * double(r) * 1099511627776.0 + 0.5;
    return !(q < 17592186044416.0) ? 17592186044416ull : (unsigned long long)q;
}
__host__ __device__ inline float residual_rms(unsigned long long sum, uint32_t count) {
    return count ? float(sqrt(double(sum) / 1099511627776.0 / double(count))) : 0.0f;
}

// Everything AVBD needs after the narrowphase: per-body, per-pair and
// per-point solver state.  Shared by the world step and the kernel seam.
struct Solver {
    uint32_t n = 0, pair_cap = 0, point_cap = 0;
    // bodies
    DBuf<float4> tpos, tq, tv, tw;       // AVBD target (RigidBodyState)
    DBuf<float4> disp;                   // 2 per body: translation, rotation
    DBuf<float4> dq;                     // orientation_delta
    DBuf<float4> inertia;                // 2 per body: m/dt^2, I/dt^2 lower tri
    DBuf<float4> prev_vel;               // xyz + history_valid
    DBuf<uint32_t> color, parent, impact, deg, adj_start, adj_cursor;
    DBuf<uint32_t> color_bodies, bp_start, anchored;
    DBuf<float4> mean;                   // free-motion mean per root
    DBuf<uint32_t> members, mean_list;   // dynamic members per root; roots taking the mean
    // joints (records owned by the caller: the world or the kernel seam)
    PmhRigidConstraint *joints = nullptr;
    uint32_t nj = 0;
    bool hinges = false;                 // some joint is a hinge (anchor projection)
    DBuf<uint32_t> jdeg, jstart, jcursor, jadj, joint_head;
    // Every body's candidate joints (alive, both ids valid) in descending
    // joint index, built by the owner at upload; each step keeps the joints
    // that initialised rows.  Bodies with more than 64 are listed apart
    // (one wavefront each).
    const uint32_t *jlist_start = nullptr, *jlist = nullptr, *jheavy = nullptr;
    uint32_t nheavy = 0;
    // Hinge anchor projection schedule: alive hinges by (level, index); the
    // hinges of one level share no body, and every hinge's level is above
    // those of the earlier hinges on its bodies.
    const uint32_t *hinge_list = nullptr, *hinge_level_dev = nullptr;
    // Neighbours per body (2 x (manifolds + joints) / bodies) of the last
    // staged solve: picks the colouring kernel's form; infinite before one.
    float neighbour_hint = 1.0e30f;
    std::vector<uint32_t> hinge_level_start;   // host copy: levels + 1 offsets
    // Contact rows in colour order: plane q (q = 0..8) of entry e at
    // erec[q * ecap + e], the same nine float4 as the point's Jacobian record,
    // so the lanes of a body read consecutive slots.  c_entry holds each
    // point's entry on its first and second body.
    DBuf<float4> erec;
    DBuf<uint32_t> c_entry;
    size_t ecap = 0;
    DBuf<ResidualLog> residual;
    // Learned warm start: cold points of the step, their 16 features and
    // (lambda scale, k0); network parameters (null: plain warm start); the
    // contact log (features and end-of-step duals of cold points).
    DBuf<uint32_t> cold_list;
    DBuf<float> cold_feat, cold_scale;
    const float *mlp = nullptr;
    uint32_t mlp_policy = 1u;            // predictions applied (penalty; lambda)
    float *log = nullptr;
    uint32_t log_cap = 0;
    // pairs
    DBuf<uint32_t> code;                 // a*n+b
    DBuf<uint2> ab;
    DBuf<uint32_t> m_count, m_start, m_slot, m_impact, m_face;
    DBuf<float4> m_init_rel;
    DBuf<uint32_t> adj;                  // 2 per pair
    DBuf<uint32_t> tri_list;             // pairs for the triangle-pair path
    DBuf<uint32_t> face_list;            // mesh pairs for the convex face path
    // points
    DBuf<float4> c_np, c_pt, c_aa, c_ab, c_err, c_lam, c_pen, c_acc, c_rep;
    DBuf<uint32_t> c_flags, c_pair;
    DBuf<float4> jac;                    // 9 per point
    DBuf<uint32_t> bp_entries;           // 2 per point (one per endpoint)
    // schedule
    DBuf<PmhContactSchedule> schedule;
    DBuf<uint32_t> color_count, color_start;
    // Dataflow sweep: each colour's positions start at a multiple of the
    // bodies per wavefront (64 / lanes), the gaps holding kInvalid, so one
    // wavefront never holds two colours; `done[i]` is the stamp of the last
    // sweep that published body i's displacement.
    uint32_t slots = 0;                  // colour-ordered positions (bodies plus padding)
    DBuf<uint32_t> done;
    uint32_t sweep_stamp = 0;
    DBuf<StepCounters> counters;
    DBuf<SolveTotals> totals;
    DBuf<uint32_t> scan_tmp;
    StepCounters *host_ctr = nullptr;        // pinned read-back of the schedule
    uint32_t *host_color_start = nullptr;
    bool host_owned = true;                  // the read-back records were allocated here
    // Iterating to a residual: pinned copies of a phase's executed-iteration
    // count, taken after every few iterations, and their events (allocating
    // solvers only).
    uint32_t *host_exec = nullptr;
    hipEvent_t exec_events[2] = {nullptr, nullptr};
    PmhResult reserve(uint32_t bodies, uint32_t pairs, uint32_t points);
    PmhResult reserve_joints(uint32_t count);
    // Every buffer as a slice of `bytes` bytes at `base` (device memory the
    // caller owns), sized for at most the given counts: no allocation
    // afterwards.  The host read-back records are the caller's too.
    PmhResult carve(uint8_t *base, size_t bytes, uint32_t bodies, uint32_t pairs, uint32_t points,
                    StepCounters *host_counters, uint32_t *host_starts);
    // Bytes carve() takes for these counts.
    static size_t carve_bytes(uint32_t bodies, uint32_t pairs, uint32_t points);
    // Sizes of this call within the carved capacity.
    PmhResult use(uint32_t bodies, uint32_t pairs, uint32_t points);
    template <class F> PmhResult each_buffer(uint32_t bodies, uint32_t pairs, uint32_t points, F &&f);
    void release();
};

// Host-built joint schedule: each body's candidate joints (alive, both ids
// valid) in descending joint index (the reference pushes a joint to the
// front of its bodies' lists in index order), the bodies with more than 64
// (one wavefront each), and the hinge anchor projection schedule: alive
// hinges by level, a hinge's level one above the highest already on either
// of its bodies (in joint order); level l spans level_start[l - 1] ..
// level_start[l].  Body ids are {dense index, generation 1}.
struct JointLists {
    std::vector<uint32_t> start, list, heavy, hinges, level_start;
};
void build_joint_lists(const PmhRigidConstraint *joints, uint32_t nj, uint32_t n, JointLists &out);

struct SolveParams {
    float dt;
    V3 gravity;
    uint32_t base_iterations, requested_iterations, profile;
    uint32_t cooperative;                // one cooperative launch with grid barriers instead of per-stage launches
    uint32_t lanes;                      // lanes of a wavefront per body in the colour sweep: 8, 16, 32, 64
    float tolerance = 0.0f;              // > 0: to-residual mode, the budget caps each phase
    bool warm_infer = false;             // predict cold points' normal duals
    uint32_t color_wait_limit = kColorWaitLimit;   // polls of one colour wait before it gives up
    uint32_t sweep_wait_limit = kColorWaitLimit;   // polls of one dataflow-sweep wait before it fails the step
    bool dataflow = true;                // one dataflow sweep per iteration instead of one launch per colour
    bool reset_status = false;           // clear the status word first (the kernel seam; a world step clears it itself)
};


// Per-launch HIP-event timing of the colour-sweep kernel on sampled steps.
// A step's events are read once the stream has passed them (at the next
// step's schedule read-back or at the end of a batch).
struct KernelTimer {
    uint32_t stride = 0;                  // 0: off
    uint32_t step = 0;
    std::vector<hipEvent_t> events;       // start/stop pairs
    uint32_t pending = 0;                 // recorded pairs not yet read
    double pending_bytes = 0.0;
    double ms = 0.0, bytes = 0.0;
    unsigned long long launches = 0;
    PmhResult collect();
    void release();
};

struct SolveTiming {
    hipEvent_t start = nullptr, stop = nullptr;   // recorded around the solve launch
    float total_ms = 0.0f;
    uint32_t launches = 0;
    KernelTimer *kernels = nullptr;
};

// Device-wide exclusive scan of `count_ptr[0]` (or `count` when the
// pointer is null) uint32 values; writes the total to `total` when given.
PmhResult exclusive_scan(hipStream_t s, const uint32_t *in, uint32_t *out, uint32_t count,
                         const uint32_t *count_ptr, uint32_t *total, DBuf<uint32_t> &tmp);

// AVBD prepare + solve on the solver state. Bodies/pairs/points must be
// filled: pair arrays (code, ab, m_count, m_start, m_init_rel), point
// arrays (c_*), counters.pairs/points, and the states.
PmhResult avbd_prepare_solve(hipStream_t s, Solver &v, const PmhBodyParameters *params,
                             const PmhRigidBodyState *previous, PmhRigidBodyState *states,
                             const SolveParams &sp, SolveTiming *timing);

int device_compute_units();

}  // namespace pmh
