// SPDX-License-Identifier: MIT
#pragma once

#include "pmh_internal.hpp"

namespace pmh {

struct World {
    // configuration
    V3 gravity{0.0f, -9.81f, 0.0f};
    float dt = 1.0f / 60.0f;               // substep length: the frame's timestep / substeps
    uint32_t substeps = 1, substep = 0;    // substeps per frame; substep of the current frame
    uint32_t profile = 1, base_iterations = 10, requested_iterations = 0;
    uint32_t body_capacity = 0, pairs_per_body = 96, points_per_body = 256;
    bool cooperative = false;
    bool per_color_sweep = false;     // one launch per colour instead of the dataflow sweep
    uint32_t solve_lanes = kDefaultLanes;
    float residual_tolerance = 0.0f;
    uint32_t color_wait_limit = kColorWaitLimit;
    uint32_t sweep_wait_limit = kColorWaitLimit;
    int device = 0;
    hipStream_t stream = nullptr;

    // host-side scene
    std::vector<HostShape> shapes;
    std::vector<PmhBodyParameters> host_params;
    std::vector<PmhRigidBodyState> host_states;
    bool shapes_dirty = true, bodies_dirty = true;
    // One-step inputs: 4 float4 per body (force, torque, linear and angular
    // impulse) and kinematic targets, uploaded before the next step.
    std::vector<float4> host_inputs;
    std::vector<PmhRigidBodyState> host_targets;
    std::vector<uint32_t> host_target_flags;
    bool inputs_dirty = false;
    uint32_t body_count = 0;
    // Scenes: contiguous body ranges starting at scene_first[s]; pairs never
    // join bodies of different scenes.
    std::vector<uint32_t> scene_first{0u};
    DBuf<uint32_t> scene_of, scene_count, scene_pair_start, scene_first_dev;
    DBuf<uint32_t> scene_large_count, scene_large_start, scene_large;
    // Joints: the host copy is authoritative until the first upload; after
    // a step the device copy is (breaking, rows and warm-start duals).
    std::vector<PmhRigidConstraint> host_joints;
    bool joints_dirty = false, joints_on_device = false;

    // device scene
    DBuf<ShapeDev> shapes_dev;
    DBuf<V3> verts_dev;
    DBuf<uint32_t> idx_dev;
    DBuf<float4> planes_dev;
    DBuf<BvhDev> nodes_dev;
    DBuf<uint32_t> leaves_dev;
    DBuf<float4> shell_dev;
    DBuf<PmhBodyParameters> params;
    DBuf<PmhRigidBodyState> state_cur, state_prev, state_next;
    DBuf<float4> inputs_dev;
    DBuf<PmhRigidBodyState> targets_dev;
    DBuf<uint32_t> target_flags_dev;
    DBuf<PmhRigidConstraint> joints_dev;
    // Collision exclusions of joints: weld roots (fixed joints that disable
    // collisions) and, per body, its disable_collisions partners with the
    // joint that excludes them.
    DBuf<uint32_t> jlist_start, jlist, jheavy;   // candidate joint lists (descending index)
    // Per body, at the last step's predicted states: its hinge contact frame,
    // and its reference point at the previous and the predicted state.
    DBuf<PmhHingeContactFrame> hinge_frames;
    DBuf<float4> hinge_ref;
    DBuf<uint32_t> hinge_list, hinge_levels;     // hinge projection schedule and its level offsets
    DBuf<uint32_t> weld, excl_start;
    DBuf<uint2> excl_list;
    uint32_t weld_joints = 0, excl_count = 0;

    // broadphase
    DBuf<float4> proxy_mn, proxy_mx;
    DBuf<int4> cell_of;
    uint32_t table_size = 0;
    float grid_cell = 1.0f;
    DBuf<uint32_t> bucket_count, bucket_start, large, cand_count, cand_offset;
    // Per bucket entry, in entry order: the body's cell (w: 1 if dynamic), its
    // box minimum (w: the body index's bits) and maximum.
    DBuf<int4> entry_cell;
    DBuf<float4> entry_mn, entry_mx;

    // contact cache, double buffered
    uint32_t cache_capacity = 0;
    uint32_t leaf_slots = 0, leaf_limit = 512;   // triangle-path candidate slots and per-pair limit
    int cache_parity = 0;
    DBuf<unsigned long long> cache_code[2];   // pair keys (a << 32 | b), ascending
    DBuf<uint32_t> cache_count[2];
    DBuf<float4> cache_data[2];
    DBuf<uint32_t> cache_n;

    Solver solver;

    // learned warm start
    DBuf<float> mlp_dev, log_dev;
    bool warm_enabled = false;
    uint32_t warm_policy = 1u;
    uint32_t log_cap = 0;
    uint64_t warm_steps = 0, warm_fallbacks = 0;
    DBuf<float4> snap_vel, snap_inputs;
    DBuf<uint32_t> snap_flags;
    DBuf<PmhRigidConstraint> snap_joints;

    // timing
    bool timing_enabled = false;
    hipEvent_t events[8] = {};
    SolveTiming solve_timing;
    KernelTimer kernel_timer;
    std::vector<hipEvent_t> solve_events;   // start/stop per step of a batch
    float last_solve_ms = 0.0f;
    bool has_stepped = false;

    PmhResult ensure_device();
    PmhResult upload_joints();
    PmhResult pull_joints();
    PmhResult step_once(uint32_t k, bool warm_infer = true);
    PmhResult step_with_fallback(uint32_t k);
    PmhResult reserve_warm();
    void release();
};

}  // namespace pmh
