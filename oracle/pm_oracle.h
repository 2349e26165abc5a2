/* SPDX-License-Identifier: MIT
 *
 * CPU oracle for the ParallelMater rigid-body step (broadphase, convex
 * narrowphase, AVBD prepare and solve).  TEST INFRASTRUCTURE ONLY: tests,
 * the smoke check and the CPU-baseline leg of the benchmark call it as the
 * checker.  The product never links, loads or calls anything here.
 *
 * The record layouts below restate the reference's storage layouts
 * byte-for-byte because the kernel seam exchanges them
 * (src/simulation/geometry_storage.hpp:18-213, src/slang/types.slang:81-101,
 * src/slang/geometry.slang:143-238).  Sizes are checked by static_assert in
 * pm_oracle.cpp.
 */
#ifndef PM_ORACLE_H
#define PM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct PmoVec3 { float x, y, z; } PmoVec3;
typedef struct PmoQuat { float x, y, z, w; } PmoQuat;

/* RigidBodyState, 52 bytes (types.slang:97-103). */
typedef struct PmoState {
    PmoVec3 position;
    PmoQuat orientation;
    PmoVec3 linear_velocity;
    PmoVec3 angular_velocity;
} PmoState;

/* BodyParameters, 64 bytes (types.slang:81-95). motion: 0 static,
 * 1 kinematic, 2 dynamic. mesh_index selects the shape. */
typedef struct PmoParams {
    uint32_t motion;
    uint32_t mesh_index;
    uint32_t mesh_generation;
    float inverse_mass;
    PmoVec3 inverse_inertia_local;
    float friction;
    float restitution;
    float linear_damping;
    float angular_damping;
    float maximum_linear_speed;
    float maximum_angular_speed;
    float collision_margin;
    uint64_t user_data;
} PmoParams;

/* HingeContactFrame, 72 bytes (geometry.slang:143-157). */
typedef struct PmoHingeFrame {
    PmoVec3 anchor, axis, local_anchor;
    float inverse_mass, inverse_moment;
    uint32_t present, fixed, axial, axial_rotation, fixed_member,
        point_member, static_body;
} PmoHingeFrame;

typedef struct PmoDual { float lambda, penalty; } PmoDual;

/* Contact, 284 bytes (geometry.slang:159-179). */
typedef struct PmoContact {
    PmoVec3 normal;
    PmoVec3 point;
    float penetration;
    PmoHingeFrame body_hinge;
    PmoHingeFrame collider_hinge;
    float accumulated_normal_impulse;
    PmoVec3 accumulated_friction_impulse;
    float initial_normal_speed;
    uint32_t warm_started;
    uint32_t persistent;
    float impact_fraction;
    float reported_normal_impulse;
    PmoVec3 reported_friction_impulse;
    PmoDual avbd_dual[3];
    PmoVec3 avbd_error;
    PmoVec3 avbd_anchor_a;
    PmoVec3 avbd_anchor_b;
    uint32_t avbd_cached;
} PmoContact;

/* ContactManifold, 2304 bytes (geometry.slang:181-190). */
typedef struct PmoManifold {
    PmoContact contacts[8];
    uint32_t count;
    PmoVec3 initial_relative_position;
    uint32_t face_patch;
    uint32_t response_slot;
    uint32_t avbd_next_a;
    uint32_t avbd_next_b;
} PmoManifold;

/* AvbdBody / GeometryConstraintBody, 236 bytes
 * (geometry_constraints.slang:42-55). */
typedef struct PmoAvbdBody {
    PmoState target;
    PmoQuat orientation_delta;
    float displacement[6];
    float inertia_h[21];
    float inertia_g[6];
    uint32_t contact_head, joint_head, color, parent, impact;
    PmoVec3 previous_velocity;
    uint32_t history_valid;
} PmoAvbdBody;

/* ContactSchedule, 264 bytes (geometry.slang:226-238). */
typedef struct PmoSchedule {
    uint32_t counts[29];
    uint32_t offsets[29];
    uint32_t budget, remaining, island_count, early_exit_count,
        maximum_passes, blocks, candidates, contacts;
} PmoSchedule;

/* ---- joints (geometry.slang:24-141) ----------------------------------
 * RigidAvbdRow 88 B, RigidConstraintOptions 228 B, RigidConstraintResource
 * 1852 B.  Joint types: 0 fixed, 1 point, 2 hinge, 3 slider, 4 piston,
 * 5 generic, 6 generic spring, 7 motor.  Body ids are {dense index, 1}. */
typedef struct PmoRow {
    float a[6], b[6];
    float lambda, penalty;
    float error, velocity, lower, upper, stiffness, damping, beta,
        reference_force;
} PmoRow;
typedef struct PmoBodyId { uint32_t index, generation; } PmoBodyId;
typedef struct PmoJointLimits { uint32_t axes; PmoVec3 lower, upper; } PmoJointLimits;
typedef struct PmoJointSprings { uint32_t axes; PmoVec3 stiffness, damping; } PmoJointSprings;
typedef struct PmoJointMotor {
    uint32_t linear_enabled, angular_enabled;
    float linear_target_velocity, linear_maximum_impulse,
        angular_target_velocity, angular_maximum_impulse;
} PmoJointMotor;
typedef struct PmoJointOptions {
    uint32_t type;
    PmoBodyId body_a, body_b;
    PmoVec3 local_anchor_a, local_anchor_b;
    PmoQuat local_orientation_a, local_orientation_b;
    PmoJointLimits linear_limits, angular_limits;
    PmoJointSprings linear_springs, angular_springs;
    PmoJointMotor motor;
    uint32_t enabled, disable_collisions;
    float breaking_impulse_threshold;
    uint32_t solver_iterations;
} PmoJointOptions;
typedef struct PmoJoint {
    PmoJointOptions options;
    uint32_t state_enabled, state_broken;
    float applied_impulse;
    uint32_t generation, alive;
    uint32_t dense_a, dense_b;
    PmoRow rows[18];
    uint32_t avbd_count, avbd_next_a, avbd_next_b;
} PmoJoint;

/* ---- numerical core (constraint_math.slang:16-345) ------------------- */
/* Row record, 22 floats: a[6] b[6] lambda penalty error velocity lower
 * upper stiffness damping beta reference_force. */
void pmo_warm_start(const float dual[2], float stiffness, float out[2]);
float pmo_force(const float dual[2], float error, float lower, float upper);
void pmo_update_dual(const float dual[2], float error, float beta,
                     float lower, float upper, float stiffness,
                     float out[2]);
void pmo_project_contact(const float force[3], float friction,
                         float out[3]);
void pmo_stiffness_scales(const float trial[3], float friction,
                          float out[3]);
void pmo_block_add_row(float h[21], float g[6], const float jacobian[6],
                       float stiffness, float row_force);
int pmo_block_solve(const float h[21], const float g[6], float delta[6]);
float pmo_row_error(const float row[22], const float a[6], const float b[6]);
float pmo_row_trial(const float row[22], const float a[6], const float b[6],
                    float dt);
float pmo_row_force(const float row[22], const float a[6], const float b[6],
                    float dt);
void pmo_row_accumulate(float h[21], float g[6], const float row[22],
                        int first, const float a[6], const float b[6],
                        float dt);
void pmo_row_advance(float row[22], const float a[6], const float b[6]);

/* ---- kernel seam: prepare + solve on caller-built manifolds ----------
 * Restates prepare_avbd_kernel (geometry_constraints.slang:1213-1591)
 * followed by solve_avbd_kernel (geometry_constraints_solve.slang:771-865)
 * with the reference's serial CPU-driver semantics. `lanes` = 0 keeps the
 * reference accumulation order; lanes = L > 0 sums each body's contact rows
 * as L interleaved partial blocks combined by a butterfly tree. */
int pmo_avbd_prepare_solve(const PmoParams *parameters,
                           const PmoState *previous, PmoState *states,
                           uint32_t count, PmoManifold *manifolds,
                           const uint32_t *pairs, uint32_t pair_count,
                           PmoAvbdBody *bodies, PmoSchedule *schedule,
                           float dt, const float gravity[3],
                           uint32_t base_iterations,
                           uint32_t requested_iterations,
                           uint32_t profile_version, uint32_t lanes);
/* The seam with joints: the joints' body ids resolve through ids[dense];
 * their solver-owned fields are written back. */
int pmo_avbd_prepare_solve_joints(const PmoParams *parameters, const PmoState *previous, PmoState *states,
                                  uint32_t count, PmoManifold *manifolds, const uint32_t *pairs,
                                  uint32_t pair_count, PmoJoint *joints, uint32_t joint_count,
                                  const PmoBodyId *ids, PmoAvbdBody *bodies, PmoSchedule *schedule, float dt,
                                  const float gravity[3], uint32_t base_iterations, uint32_t requested_iterations,
                                  uint32_t profile_version, uint32_t lanes);

/* ---- world ------------------------------------------------------------ */
typedef struct PmoWorldDesc {
    float gravity[3];
    float dt;
    uint32_t profile_version;      /* 1 realtime (8 sweeps), 0 legacy */
    uint32_t base_iterations;
    uint32_t requested_iterations;
    uint32_t body_capacity;        /* 0: the body count at first step */
    uint32_t lanes;                /* accumulation order, see above */
    uint32_t flags;                /* bit 0: brute-force broadphase; bits 1-3: mistakes planted by tests
                                      (1: no free-motion mean, 2: no crossing safeguard, 4: no hinge
                                      contact frames) */
    float residual_tolerance;      /* > 0: iterate each phase until the constraint residual
                                      is at most this (the budget caps); 0: fixed budget */
    uint32_t substeps;             /* substeps per step of dt / substeps (0 = 1), as
                                      pm_integrate_rigid_bodies' remaining_substeps */
} PmoWorldDesc;

typedef struct PmoWorldStats {
    uint32_t pairs;                /* active pairs after filtering */
    uint32_t manifolds;            /* non-empty manifolds */
    uint32_t points;               /* contact points */
    uint32_t colors;
    uint32_t islands;
    uint32_t impact;               /* schedule counts[0] */
    uint32_t budget;
    uint32_t unsupported_pairs;    /* round-shape pairs with no analytic path */
    uint32_t cache_hits;           /* contact points restored from cache */
    uint32_t phases;
    uint32_t triangle_pairs;       /* mesh pairs on the triangle-pair path */
    uint32_t fallback_pairs;       /* of those, solved by BVH traversal */
} PmoWorldStats;

void *pmo_world_create(const PmoWorldDesc *desc);
/* Constraint residual after each dual update of the last step: RMS in out[phase * 64 +
 * iteration] and maximum in maximum[...] (NaN where no update ran), and the iterations
 * run per phase. */
void pmo_world_residuals(void *world, float *out, float *maximum, uint32_t *executed);
void pmo_world_destroy(void *world);
/* Returns the shape index or -1. Meshes run the reference's upload
 * preparation (runtime_deformables.cu.inc:4-304). */
int pmo_world_add_mesh(void *world, const float *vertices, uint32_t vertex_count,
                       const uint32_t *indices, uint32_t index_count);
/* Analytic sphere (radius) and capsule (radius, half segment length along
 * local y): shapes with no reference counterpart. */
int pmo_world_add_sphere(void *world, float radius);
int pmo_world_add_capsule(void *world, float radius, float half_length);
int pmo_world_add_bodies(void *world, uint32_t count,
                         const PmoParams *parameters, const PmoState *states);
/* PmRigidInput (include/parallel_mater/parallel_mater.h:421-427). */
typedef struct PmoRigidInput {
    uint32_t body_index;
    float force[3], torque[3], linear_impulse[3], angular_impulse[3];
} PmoRigidInput;
/* Forces, torques and impulses for the next step on dynamic bodies
 * (accumulating); kinematic targets for the next step on kinematic
 * bodies.  Both return 0 or -1 (invalid body, motion type or values). */
int pmo_world_apply_inputs(void *world, uint32_t count, const PmoRigidInput *inputs);
int pmo_world_set_kinematic_targets(void *world, uint32_t first, uint32_t count, const PmoState *targets);
/* Joints between bodies of the world (alive, generation 1); returns 0 or
 * -1 for an invalid type or body. */
int pmo_world_add_joints(void *world, uint32_t count, const PmoJointOptions *options);
/* pmCmdWriteRigidConstraints (resources.cpp:998-1060): each slot first..
 * first+count-1 is removed (remove_rigid_constraint, runtime_rigid.cu.inc:
 * 639-654) and, where active[k] != 0, added again with options[k] and a
 * fresh state (add_rigid_constraint, :552-596). 0, or -1 on bad input. */
int pmo_world_write_joints(void *world, uint32_t first, uint32_t count, const PmoJointOptions *options,
                           const uint32_t *active);
/* set_rigid_body_state (runtime_rigid.cu.inc:146-195) for bodies first..
 * first+count-1: the state (orientation normalised) becomes current and
 * previous, the body's solver history, inputs and kinematic target are
 * cleared and the rows of every alive joint on the body are zeroed. */
int pmo_world_set_rigid_states(void *world, uint32_t first, uint32_t count, const PmoState *states);
/* The last step's rigid contact events (pm_prepare_parallel_contact_events
 * and pm_save_rigid_contact_cache, geometry.slang:2904-2936, 3023-3037):
 * one per manifold point in active-pair order, PmContactEvent layout
 * (parallel_mater.h:959-968). Writes min(total, capacity); returns total. */
typedef struct PmoContactEvent {
    uint32_t type, first, second;
    PmoVec3 position, normal;
    float penetration, normal_impulse;
    PmoVec3 friction_impulse;
} PmoContactEvent;
uint32_t pmo_world_contact_events(void *world, PmoContactEvent *events, uint32_t capacity);
uint32_t pmo_world_joint_count(void *world);
void pmo_world_get_joints(void *world, PmoJoint *out);
int pmo_world_step(void *world, uint32_t steps);
uint32_t pmo_world_body_count(void *world);
/* Cumulative wall time of world steps, of AVBD prepare and of the AVBD
 * solve (milliseconds), and body iterations solved; optionally reset. */
void pmo_world_timing(void *world, double out[4], int reset);

/* The learned warm start's network (pm_oracle_mlp.cpp): forward pass and
 * mean-squared-error gradients, with (round_bf16 != 0) or without the
 * device path's bf16 rounding points. */
void pmo_mlp_forward(const float *params, const float *x, uint32_t n, float *y, int round_bf16);
float pmo_mlp_gradients(const float *params, const float *samples, uint32_t count, float *grads, int round_bf16);
void pmo_world_get_states(void *world, PmoState *out);
void pmo_world_set_states(void *world, const PmoState *states);
void pmo_world_get_stats(void *world, PmoWorldStats *out);
/* Results of the last step: active pair codes a*n+b (ascending), their
 * manifolds, the AVBD bodies (colour, parent, impact) and schedule. */
uint32_t pmo_world_pair_count(void *world);
void pmo_world_get_pairs(void *world, uint32_t *out);
void pmo_world_get_manifolds(void *world, PmoManifold *out);
void pmo_world_get_bodies(void *world, PmoAvbdBody *out);
void pmo_world_get_schedule(void *world, PmoSchedule *out);
/* Proxy AABBs of the last step, 6 floats per body (min xyz, max xyz). */
void pmo_world_get_proxies(void *world, float *out);
/* Shape metadata after upload preparation: 16 floats
 * (min xyz, max xyz, bounding centre xyz, radius, unit inertia xyz,
 * kind, index count, unique plane count). */
void pmo_world_get_shape(void *world, uint32_t shape, float out[16]);
/* Mesh indices after BVH reordering and per-triangle planes (4 floats). */
void pmo_world_get_mesh_indices(void *world, uint32_t shape, uint32_t *out);
void pmo_world_get_mesh_planes(void *world, uint32_t shape, float *out);

/* Mesh BVH after upload preparation: per node 4 uints (left, right, first
 * triangle, triangle count; leaves have count != 0) and 6 floats of local
 * bounds.  Returns the node count; either output may be NULL. */
uint32_t pmo_world_get_mesh_bvh(void *world, uint32_t shape, uint32_t *nodes, float *bounds);
/* Certified shell normals, 3 floats per triangle in the reordered index
 * order; returns 0 (and writes nothing) when the mesh has none. */
uint32_t pmo_world_get_mesh_shell(void *world, uint32_t shape, float *out);

/* ---- geometric kernels of the triangle-pair path (test hooks) -------- */
void pmo_closest_segments(const float p0[3], const float p1[3], const float q0[3],
                          const float q1[3], float on_p[3], float on_q[3]);
void pmo_closest_on_triangle(const float p[3], const float tri[9], float out[3]);
void pmo_closest_triangle_pair(const float a[9], const float b[9], float on_a[3], float on_b[3]);
void pmo_triangle_feature_normal(const float a[9], const float b[9], const float on_a[3],
                                 const float on_b[3], const float fallback[3], float out[3]);

#ifdef __cplusplus
}
#endif
#endif
