/* SPDX-License-Identifier: MIT
 *
 * MI355X (gfx950) HIP driver for the ParallelMater rigid-body step:
 * broadphase, convex/analytic narrowphase and the Augmented Vertex Block
 * Descent (AVBD) contact solver.
 *
 * C ABI only: plain pointers, sizes and the reference's record layouts.
 * Two levels are exported:
 *   - the world API, which is what the reference's rigid pm* command path
 *     drives (pmCmdBuildGeometry, pmCmdWriteRigidBodies,
 *     pmCmdStepSimulation, pmCmdCopyRigidStates ...);
 *   - the kernel seam, which stands where the reference's dispatch
 *     wrappers stand (src/simulation/geometry_constraints_dispatch.hpp:15-43)
 *     and takes device pointers to the reference's own record layouts.
 * Every call returns a PmhResult whose values equal the reference's
 * PmResult (include/parallel_mater/parallel_mater.h:93-110).
 */
#ifndef PM_HIP_RIGID_H
#define PM_HIP_RIGID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t PmhResult;
#define PMH_SUCCESS 0
#define PMH_ERROR_INVALID_ARGUMENT (-1)
#define PMH_ERROR_INVALID_STATE (-2)
#define PMH_ERROR_OUT_OF_HOST_MEMORY (-3)
#define PMH_ERROR_OUT_OF_DEVICE_MEMORY (-4)
#define PMH_ERROR_FEATURE_NOT_PRESENT (-5)
#define PMH_ERROR_CAPACITY_EXCEEDED (-7)
#define PMH_ERROR_DEVICE_LOST (-8)

/* ---- records (reference layouts, byte for byte) ---------------------- */
typedef struct PmhVec3 { float x, y, z; } PmhVec3;
typedef struct PmhQuat { float x, y, z, w; } PmhQuat;

/* RigidBodyState (src/slang/types.slang:97-103), 52 bytes. */
typedef struct PmhRigidBodyState {
    PmhVec3 position;
    PmhQuat orientation;
    PmhVec3 linear_velocity;
    PmhVec3 angular_velocity;
} PmhRigidBodyState;

#define PMH_MOTION_STATIC 0u
#define PMH_MOTION_KINEMATIC 1u
#define PMH_MOTION_DYNAMIC 2u

/* BodyParameters (types.slang:81-95), 64 bytes. mesh_index selects the
 * shape returned by pmhWorldAdd{Mesh,Sphere,Capsule}. */
typedef struct PmhBodyParameters {
    uint32_t motion;
    uint32_t mesh_index;
    uint32_t mesh_generation;
    float inverse_mass;
    PmhVec3 inverse_inertia_local;
    float friction;
    float restitution;
    float linear_damping;
    float angular_damping;
    float maximum_linear_speed;
    float maximum_angular_speed;
    float collision_margin;
    uint64_t user_data;
} PmhBodyParameters;

/* HingeContactFrame (geometry.slang:143-157), 72 bytes. */
typedef struct PmhHingeContactFrame {
    PmhVec3 anchor, axis, local_anchor;
    float inverse_mass, inverse_moment;
    uint32_t present, fixed, axial, axial_rotation, fixed_member,
        point_member, static_body;
} PmhHingeContactFrame;

typedef struct PmhAvbdDual { float lambda, penalty; } PmhAvbdDual;

/* Contact (geometry.slang:159-179), 284 bytes. */
typedef struct PmhContact {
    PmhVec3 normal;
    PmhVec3 point;
    float penetration;
    PmhHingeContactFrame body_hinge;
    PmhHingeContactFrame collider_hinge;
    float accumulated_normal_impulse;
    PmhVec3 accumulated_friction_impulse;
    float initial_normal_speed;
    uint32_t warm_started;
    uint32_t persistent;
    float impact_fraction;
    float reported_normal_impulse;
    PmhVec3 reported_friction_impulse;
    PmhAvbdDual avbd_dual[3];
    PmhVec3 avbd_error;
    PmhVec3 avbd_anchor_a;
    PmhVec3 avbd_anchor_b;
    uint32_t avbd_cached;
} PmhContact;

/* ContactManifold (geometry.slang:181-190), 2304 bytes. */
typedef struct PmhContactManifold {
    PmhContact contacts[8];
    uint32_t count;
    PmhVec3 initial_relative_position;
    uint32_t face_patch;
    uint32_t response_slot;
    uint32_t avbd_next_a;
    uint32_t avbd_next_b;
} PmhContactManifold;

/* AvbdBody / GeometryConstraintBody (geometry_constraints.slang:42-55),
 * 236 bytes. */
typedef struct PmhAvbdBody {
    PmhRigidBodyState target;
    PmhQuat orientation_delta;
    float displacement[6];
    float inertia_h[21];
    float inertia_g[6];
    uint32_t contact_head, joint_head, color, parent, impact;
    PmhVec3 previous_velocity;
    uint32_t history_valid;
} PmhAvbdBody;

/* ContactSchedule (geometry.slang:226-238), 264 bytes. */
typedef struct PmhContactSchedule {
    uint32_t counts[29];
    uint32_t offsets[29];
    uint32_t budget, remaining, island_count, early_exit_count,
        maximum_passes, blocks, candidates, contacts;
} PmhContactSchedule;

/* ---- world API ---------------------------------------------------------
 * One world = one PmSimulation's rigid domain on one GPU.  Bodies are
 * dense; identities are {index = dense, generation = 1}. */
typedef struct PmhWorld_T *PmhWorld;

typedef struct PmhWorldDesc {
    float gravity[3];
    float timestep;                  /* dt; realtime v1 requires 1/60 */
    uint32_t profile_version;        /* 1: realtime v1 (8 sweeps); 0: legacy */
    uint32_t base_iterations;        /* legacy profile iteration budget */
    uint32_t requested_iterations;   /* legacy override, 0 = base */
    uint32_t body_capacity;          /* contact-cache slot rule; 0 = bodies */
    uint32_t pairs_per_body;         /* pair pool per body (average), 0 = 96 */
    uint32_t points_per_body;        /* contact-point pool per body, 0 = 256 */
    uint32_t device;                 /* HIP device ordinal */
    uint32_t flags;                  /* PMH_WORLD_* */
    uint32_t solve_lanes;            /* lanes per body in the colour sweep: 8, 16, 32, 64; 0 = 32.
                                        Sets the order in which a body's contact rows are summed. */
    uint32_t kernel_timing_stride;   /* every Nth step brackets each colour-sweep launch with HIP
                                        events (pmhWorldKernelTiming); 0 = off */
    float residual_tolerance;        /* > 0: each solve phase iterates until the constraint residual
                                        (pmhWorldGetResiduals) is at most this, the iteration budget
                                        being the cap; 0 = the fixed budget of the profile */
    uint32_t substeps;               /* substeps per pmhWorldStep step, 0 = 1, at most 64: each of
                                        timestep / substeps; forces act in every substep, impulses in
                                        the first, a kinematic target moves 1/remaining of the way per
                                        substep (pm_integrate_rigid_bodies' remaining_substeps,
                                        geometry.slang:524-612) and inputs clear after the last */
} PmhWorldDesc;

#define PMH_WORLD_COOPERATIVE_SOLVE 1u  /* whole solve in one cooperative launch with grid barriers */
#define PMH_WORLD_PHASE_TIMING 2u       /* HIP events around each phase and the solve of every step
                                           (pmhWorldGetPhaseTimes, pmhWorldSolveTiming, and the per-launch
                                           colour-sweep timing of kernel_timing_stride); off by default */
#define PMH_WORLD_PER_COLOR_SWEEP 4u    /* one launch per colour per iteration instead of the dataflow
                                           sweep (one launch per iteration, each body starting once its
                                           lower-colour neighbours are done); same results */

typedef struct PmhWorldStats {
    uint32_t pairs, manifolds, points, colors, islands, impact, budget,
        unsupported_pairs,       /* round-shape pairs with no analytic path */
        cache_hits, phases;
    uint32_t status;             /* the last substep's status bits: 1 pair pool, 2 point pool,
                                    4 colours (all overflows: PMH_ERROR_CAPACITY_EXCEEDED), 16 a
                                    colouring that gave up (PMH_ERROR_INVALID_STATE) */
    uint32_t pair_capacity, point_capacity, points_allocated;
    uint32_t triangle_pairs;     /* mesh pairs on the triangle-pair path */
    uint32_t fallback_pairs;     /* of those, solved by BVH traversal (no candidate slot or too many) */
} PmhWorldStats;

/* Phase timings of the last step (milliseconds, HIP events on the
 * world's stream): integrate+proxies, broadphase, narrowphase, contact
 * cache, prepare, solve, finish. */
#define PMH_PHASE_COUNT 7
typedef struct PmhPhaseTimes { float ms[PMH_PHASE_COUNT]; } PmhPhaseTimes;

/* pmCreateSimulation + pmBindSimulationMemory (src/api/resources.cpp:535,
 * :700) for the rigid domain. */
PmhResult pmhCreateWorld(const PmhWorldDesc *desc, PmhWorld *world);
void pmhDestroyWorld(PmhWorld world);

/* pmCmdBuildGeometry -> SimulationRuntime::add_triangle_mesh
 * (src/api/resources.cpp:523, src/simulation/runtime_deformables.cu.inc:
 * 4-304): validation, median-split BVH triangle order, convex face
 * planes, bounds, bounding sphere and unit inertia. */
PmhResult pmhWorldAddMesh(PmhWorld world, const float *vertices,
                          uint32_t vertex_count, const uint32_t *indices,
                          uint32_t index_count, uint32_t *shape);
/* Analytic sphere and capsule (radius, half segment length along local y).
 * No reference counterpart: the reference models every shape as a mesh. */
PmhResult pmhWorldAddSphere(PmhWorld world, float radius, uint32_t *shape);
PmhResult pmhWorldAddCapsule(PmhWorld world, float radius, float half_length,
                             uint32_t *shape);
/* Shape metadata: min xyz, max xyz, bounding centre xyz, bounding radius,
 * unit inertia xyz, kind, index count, unique plane count. */
PmhResult pmhWorldGetShapeInfo(PmhWorld world, uint32_t shape, float out[16]);

/* pmCmdWriteRigidBodies -> add_rigid_body (src/api/resources.cpp:821,
 * src/simulation/runtime_rigid.cu.inc:4-68), host arrays. */
PmhResult pmhWorldAddBodies(PmhWorld world, uint32_t count,
                            const PmhBodyParameters *parameters,
                            const PmhRigidBodyState *states);
uint32_t pmhWorldBodyCount(PmhWorld world);

/* Batched scenes (no counterpart in the reference, whose simulations each
 * own one scene): appends a scene of `count` bodies as a new contiguous body
 * range and returns its index.  Bodies of different scenes never pair, so
 * every scene steps as it would alone, with its own body count for the
 * triangle-path slot and limit rules.  Plain pmhWorldAddBodies appends to
 * the last scene. */
PmhResult pmhWorldAddScene(PmhWorld world, uint32_t count, const PmhBodyParameters *parameters,
                           const PmhRigidBodyState *states, uint32_t *scene);
uint32_t pmhWorldSceneCount(PmhWorld world);

/* Learned warm start (no counterpart in the reference; its cold contacts
 * start from lambda = 0 and the pair stiffness, geometry_constraints.slang:
 * 1366-1373).  A 16 -> 64 -> 64 -> 2 network (ReLU; BF16 weights and
 * activations on MFMA, FP32 accumulation) predicts the normal row's
 * (lambda, penalty) of every cold contact point from 16 features.  The
 * 5380 parameters are packed W1[64][16], b1[64], W2[64][64], b2[64],
 * W3[2][64], b3[2], then the largest training target of each output
 * (set by pmhWarmstartTrain; predictions are clamped to [0, it]).
 *   pmhWorldSetWarmstart: use `params` from the next step (NULL: plain),
 *     applying the predicted penalty (PMH_WARMSTART_PENALTY) and/or the
 *     predicted lambda (PMH_WARMSTART_LAMBDA, only on cold points whose
 *     initial normal error is positive: those active from the first
 *     iteration anyway).  In
 *     to-residual mode (residual_tolerance > 0) a step whose residual ends
 *     above the tolerance is taken again with the plain warm start;
 *     pmhWorldWarmstartStats counts steps and such fallbacks.
 *   pmhWorldSetContactLog: log every cold point's 16 features and its duals
 *     at the end of its step (log10(1 + lambda / (m_eff |g|)), log10(k / k0)) as 18
 *     floats, up to `capacity` samples (0: off); pmhWorldGetContactLog reads
 *     them.
 *   pmhWarmstartTrain: Adam over minibatches of 64 on the GPU, mean squared
 *     error, `epochs` passes in seeded orders; params in/out.
 *   pmhWarmstartForward / pmhWarmstartGradients: the network's outputs and
 *     one minibatch's loss and gradients (for checks). */
#define PMH_WARMSTART_PENALTY 1u
#define PMH_WARMSTART_LAMBDA 2u
PmhResult pmhWorldSetWarmstart(PmhWorld world, const float *params, uint32_t policy);
PmhResult pmhWorldSetContactLog(PmhWorld world, uint32_t capacity);
PmhResult pmhWorldGetContactLog(PmhWorld world, float *samples, uint32_t capacity, uint32_t *count);
PmhResult pmhWorldWarmstartStats(PmhWorld world, uint64_t *steps, uint64_t *fallbacks);
void pmhWarmstartInit(uint32_t seed, float *params);
PmhResult pmhWarmstartTrain(const float *samples, uint32_t n, uint32_t epochs, float learning_rate, uint32_t seed,
                            float *params, float *epoch_loss);
PmhResult pmhWarmstartForward(const float *params, const float *features, uint32_t n, float *outputs);
PmhResult pmhWarmstartGradients(const float *samples, uint32_t count, const float *params, float *grads,
                                float *loss);

/* Lanes per body of the colour sweep for the following steps (8, 16, 32 or
 * 64): the tile the autotuner picks.  It sets the order in which a body's
 * contact rows are summed. */
PmhResult pmhWorldSetSolveLanes(PmhWorld world, uint32_t lanes);

/* pmCmdStepSimulation (src/api/resources.cpp:1230-1288) -> step_async ->
 * the production rigid order: predict, proxies, broadphase, narrowphase,
 * contact cache, AVBD prepare and solve, speed clamp, cache save
 * (production_realtime_backend.cu.inc:2642-2687, 7664-7861).  Enqueues
 * `steps` steps and waits for them. */
/* ---- joints ---------------------------------------------------------------
 * RigidAvbdRow, RigidConstraintOptions and RigidConstraintResource
 * (src/slang/geometry.slang:24-141; 88, 228 and 1852 bytes).  Types:
 * 0 fixed, 1 point, 2 hinge, 3 slider, 4 piston, 5 generic, 6 generic
 * spring, 7 motor.  A body id is {dense index, 1} in a world. */
typedef struct PmhAvbdRow {
    float a[6], b[6];
    float lambda, penalty;
    float error, velocity, lower, upper, stiffness, damping, beta, reference_force;
} PmhAvbdRow;
typedef struct PmhBodyId { uint32_t index, generation; } PmhBodyId;
typedef struct PmhJointLimits { uint32_t axes; PmhVec3 lower, upper; } PmhJointLimits;
typedef struct PmhJointSprings { uint32_t axes; PmhVec3 stiffness, damping; } PmhJointSprings;
typedef struct PmhJointMotor {
    uint32_t linear_enabled, angular_enabled;
    float linear_target_velocity, linear_maximum_impulse;
    float angular_target_velocity, angular_maximum_impulse;
} PmhJointMotor;
typedef struct PmhRigidConstraintOptions {
    uint32_t type;
    PmhBodyId body_a, body_b;
    PmhVec3 local_anchor_a, local_anchor_b;
    PmhQuat local_orientation_a, local_orientation_b;
    PmhJointLimits linear_limits, angular_limits;
    PmhJointSprings linear_springs, angular_springs;
    PmhJointMotor motor;
    uint32_t enabled, disable_collisions;
    float breaking_impulse_threshold;
    uint32_t solver_iterations;
} PmhRigidConstraintOptions;
typedef struct PmhRigidConstraint {
    PmhRigidConstraintOptions options;
    uint32_t state_enabled, state_broken;
    float applied_impulse;
    uint32_t generation, alive;
    uint32_t dense_a, dense_b;
    PmhAvbdRow rows[18];
    uint32_t avbd_count, avbd_next_a, avbd_next_b;
} PmhRigidConstraint;

/* pmCmdWriteRigidConstraints (src/api/resources.cpp:998) -> the runtime's
 * constraint resources: appends joints (state, rows and lists start
 * empty).  PM_ERROR_INVALID_ARGUMENT for an unknown type. */
PmhResult pmhWorldAddJoints(PmhWorld world, uint32_t count, const PmhRigidConstraintOptions *options);
uint32_t pmhWorldJointCount(PmhWorld world);
/* The joint records after the last step (pmCmdCopyConstraintStates,
 * src/api/domains.cpp:1585):
 * enabled/broken state, applied impulse, dense ids, rows and list links. */
PmhResult pmhWorldGetJoints(PmhWorld world, PmhRigidConstraint *joints);
/* pmCmdWriteRigidConstraints (resources.cpp:998-1060): each joint slot
 * first..first+count-1 is removed (remove_rigid_constraint, runtime_rigid.cu.inc:
 * 639-654: not alive, state cleared, generation advanced) and, where
 * active[k] != 0, added again from options[k] with a fresh state and zero rows
 * (add_rigid_constraint, :552-596). The slot is the joint's index; the
 * reference re-adds into its first free slot, which is the same slot whenever
 * removed joints are re-added in index order. Invalid options (type, body ids,
 * no dynamic body, breaking threshold, solver_iterations 1..64) fail the whole
 * call with PMH_ERROR_INVALID_ARGUMENT before anything changes. */
PmhResult pmhWorldWriteJoints(PmhWorld world, uint32_t first, uint32_t count,
                              const PmhRigidConstraintOptions *options, const uint32_t *active);

/* PmRigidInput (include/parallel_mater/parallel_mater.h:421-427). */
typedef struct PmhRigidInput {
    uint32_t body_index;
    float force[3], torque[3], linear_impulse[3], angular_impulse[3];
} PmhRigidInput;

/* pmCmdApplyRigidInputs (src/api/resources.cpp:908) ->
 * apply_rigid_input (src/simulation/runtime_rigid.cu.inc:265-289): forces,
 * torques and impulses accumulate on dynamic bodies and act during the
 * next step only.  PM_ERROR_INVALID_ARGUMENT for a body that is not
 * dynamic or non-finite values (nothing is applied then). */
PmhResult pmhWorldApplyRigidInputs(PmhWorld world, uint32_t count, const PmhRigidInput *inputs);

/* pmCmdSetKinematicTargets (src/api/resources.cpp:967) ->
 * set_kinematic_target (runtime_rigid.cu.inc:197-230): poses that
 * kinematic bodies reach during the next step (velocities follow from the
 * motion); kinematic bodies only, finite, non-zero orientation (stored
 * normalised). */
PmhResult pmhWorldSetKinematicTargets(PmhWorld world, uint32_t first, uint32_t count,
                                      const PmhRigidBodyState *targets);

/* pmCmdStepSimulation (src/api/resources.cpp:1230-1288): `steps` steps
 * of `substeps` substeps each.  Returns at the first substep whose status
 * fails (PMH_ERROR_CAPACITY_EXCEEDED for a full pair or point pool or more
 * than 256 colours, PMH_ERROR_INVALID_STATE for a colouring that gave up);
 * the states are then those before that substep.  Worlds of fewer than four
 * bodies are refused with PMH_ERROR_FEATURE_NOT_PRESENT: the first GPU steps
 * of a two- and a three-body world with no active pair did not return on an
 * earlier build, and no cause has been shown.  PMH_ERROR_DEVICE_LOST when
 * the device does not finish within PMH_SYNC_TIMEOUT_S seconds (default
 * 120). */
PmhResult pmhWorldStep(PmhWorld world, uint32_t steps);

/* PmStepInfo.gravity of pmCmdStepSimulation (src/api/resources.cpp:1262):
 * the gravity of the following steps (finite components). */
PmhResult pmhWorldSetGravity(PmhWorld world, const float gravity[3]);

/* Polls of one colouring wait before it gives up (default 2^22): a wait
 * that runs out ends the whole colouring (every other wait gives up too)
 * and the step returns PMH_ERROR_INVALID_STATE.  No counterpart in the
 * reference; 0 makes the first wait on an uncoloured neighbour give up, so
 * a test can take that path without a stalled device. */
PmhResult pmhWorldSetColorWaitLimit(PmhWorld world, uint32_t polls);

/* Polls of one dataflow-sweep wait (a body waiting for a lower-colour
 * neighbour) before it fails the step (default 2^22).  A wait that runs out
 * ends every other wait of the step; the step keeps its states and
 * pmhWorldStep returns PMH_ERROR_INVALID_STATE.  The ticket order of the
 * sweep rules this out short of a device fault.  No counterpart in the
 * reference; 0 makes the first unmet wait fail, so a test can take that
 * path without a stalled device. */
PmhResult pmhWorldSetSweepWaitLimit(PmhWorld world, uint32_t polls);

/* pmCmdCopyRigidStates / pmCmdSetRigidStates (src/api/resources.cpp:1293;
 * runtime_rigid.cu.inc:147-195: a set-state resets AVBD history and the
 * contact cache). */
PmhResult pmhWorldCopyStates(PmhWorld world, PmhRigidBodyState *states);
PmhResult pmhWorldSetStates(PmhWorld world, const PmhRigidBodyState *states);
/* pmCmdSetRigidStates (resources.cpp:1293-1330 family) -> set_rigid_body_state
 * (runtime_rigid.cu.inc:146-195) for bodies first..first+count-1: each state
 * (finite, non-zero orientation, stored normalised) becomes the current one;
 * the body's solver history, pending inputs and kinematic target are
 * cleared and the rows of every alive joint on the body are zeroed. The
 * contact cache is kept. */
PmhResult pmhWorldSetRigidStates(PmhWorld world, uint32_t first, uint32_t count,
                                 const PmhRigidBodyState *states);

PmhResult pmhWorldGetStats(PmhWorld world, PmhWorldStats *stats);
PmhResult pmhWorldGetPhaseTimes(PmhWorld world, PmhPhaseTimes *times);
/* Kernel-duration probe for the dominant solve kernel: total milliseconds
 * and launch count of the solve phase since the last reset. */
PmhResult pmhWorldSolveTiming(PmhWorld world, float *milliseconds,
                              uint32_t *launches, int reset);

/* Work totals of the AVBD solve since the last reset: steps, body
 * iterations (dynamic bodies x iterations of the pose phase plus impact
 * bodies x iterations of the velocity phase), velocity phases run and
 * algorithmic bytes (per iteration: 136 B per dynamic body, 192 B per
 * contact point with one dynamic endpoint and 276 B with two, 168 B per
 * manifold with one dynamic endpoint and 232 B with two). */
typedef struct PmhWorldTotals {
    uint64_t steps, body_iterations, velocity_phases;
    double algorithmic_bytes;
} PmhWorldTotals;
PmhResult pmhWorldGetTotals(PmhWorld world, PmhWorldTotals *totals, int reset);

/* Per-launch timing of the dominant kernel, the colour sweep of the staged
 * solve (one launch per colour per iteration), on the steps sampled by
 * kernel_timing_stride: summed HIP-event milliseconds, the number of timed
 * launches and their algorithmic bytes (per iteration: 136 B per body,
 * 84 B per contact point and dynamic endpoint, 64 B per manifold and
 * dynamic endpoint). */
typedef struct PmhKernelTiming {
    double milliseconds;
    uint64_t launches;
    double algorithmic_bytes;
} PmhKernelTiming;
/* Constraint residual after each dual update of the last step, per phase:
 * RMS in out[phase * 64 + iteration] and maximum in maximum[...] (NULL to
 * skip; NaN where no update ran), and the iterations each phase ran.  No
 * counterpart in the reference, which runs a fixed budget
 * (geometry_constraints.slang:1441-1444).  Each contact point contributes
 * |C_n| if its normal force is positive, else max(C_n, 0), raised to |C_t|
 * when its friction sticks; each joint the largest |C| of its hard equality
 * rows; all at the displacements the update sees.  Squares are summed in
 * 2^-40 fixed point (exact in any order), each clamped at r = 4 m. */
PmhResult pmhWorldGetResiduals(PmhWorld world, float *out, float *maximum, uint32_t *executed);

PmhResult pmhWorldKernelTiming(PmhWorld world, PmhKernelTiming *timing, int reset);

/* Results of the last step in the reference layouts (for parity checks
 * and pmCmdCopyContactEvents-style consumers). */
uint32_t pmhWorldPairCount(PmhWorld world);
/* Pair codes a*n+b, ascending.  The world itself keys pairs by (a, b) and
 * runs any body count; the 32-bit codes exist while n*n < 2^32, and for a
 * larger world this call returns PMH_ERROR_CAPACITY_EXCEEDED. */
PmhResult pmhWorldGetPairs(PmhWorld world, uint32_t *codes);
PmhResult pmhWorldGetManifolds(PmhWorld world, PmhContactManifold *manifolds);
/* PmContactEvent (parallel_mater.h:959-968), 56 bytes. */
typedef struct PmhContactEvent {
    uint32_t type;               /* 0: rigid-rigid */
    uint32_t first, second;      /* body indices of the pair (a, b) */
    PmhVec3 position, normal;
    float penetration;           /* max(0, penetration) */
    float normal_impulse;        /* accumulated if the contact persisted, else reported */
    PmhVec3 friction_impulse;
} PmhContactEvent;
/* pmCmdCopyContactEvents: the last step's rigid contact events, one per
 * contact point in active-pair order (pm_prepare_parallel_contact_events and
 * pm_save_rigid_contact_cache, geometry.slang:2904-2936, 3023-3037). Writes
 * min(total, capacity) events; *count receives the total. */
PmhResult pmhWorldGetContactEvents(PmhWorld world, PmhContactEvent *events, uint32_t capacity,
                                   uint32_t *count);
PmhResult pmhWorldGetAvbdBodies(PmhWorld world, PmhAvbdBody *bodies);
PmhResult pmhWorldGetSchedule(PmhWorld world, PmhContactSchedule *schedule);
PmhResult pmhWorldGetProxies(PmhWorld world, float *bounds);

/* ---- kernel seam --------------------------------------------------------
 * prepare_geometry_constraints + solve_geometry_constraints
 * (src/simulation/geometry_constraints_dispatch.hpp:15-43): all pointers
 * are device pointers in the reference layouts; `active_pair_count`
 * points to one device uint32.  `bodies` carries the per-body history
 * between calls exactly as the reference's avbd_bodies buffer.  Runs on
 * `stream` (a hipStream_t, NULL for the default stream). */
PmhResult pmhPrepareSolveGeometryConstraints(
    void *stream, const PmhBodyParameters *parameters,
    const PmhRigidBodyState *previous, PmhRigidBodyState *states,
    uint32_t body_count, PmhContactManifold *manifolds,
    const uint32_t *active_pairs, const uint32_t *active_pair_count,
    PmhAvbdBody *bodies, PmhContactSchedule *schedule, float timestep,
    const float gravity[3], uint32_t base_iterations,
    uint32_t requested_iterations, uint32_t profile_version);

/* The seam with the simulation's joints: prepare_geometry_constraints's
 * `constraints` (the reference's RigidConstraintResource records,
 * PmhRigidConstraint, `constraint_capacity` slots, dead slots not alive) and
 * `ids` (the id of each dense body, which the joints' body ids resolve
 * through; rigid_avbd_find_dense, geometry_constraints.slang:318-325), with
 * the same meaning as in src/simulation/geometry_constraints_dispatch.hpp:
 * 15-24.  The joints' state, applied impulse, dense ids, rows and list links
 * are written back; their options are not touched.  No joints
 * (capacity 0) is pmhPrepareSolveGeometryConstraints. */
PmhResult pmhPrepareSolveGeometryConstraintsJoints(
    void *stream, const PmhBodyParameters *parameters,
    const PmhRigidBodyState *previous, PmhRigidBodyState *states,
    uint32_t body_count, PmhContactManifold *manifolds,
    const uint32_t *active_pairs, const uint32_t *active_pair_count,
    PmhRigidConstraint *constraints, uint32_t constraint_capacity,
    const PmhBodyId *ids, PmhAvbdBody *bodies, PmhContactSchedule *schedule,
    float timestep, const float gravity[3], uint32_t base_iterations,
    uint32_t requested_iterations, uint32_t profile_version);

/* The same seam without allocating: no counterpart in the reference; for
 * callers that cannot allocate device or pinned memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors as pmhPrepareSolveGeometryConstraints.
 * Every buffer of the solve is a slice of one static 24 MiB device array per
 * GPU (held until the process ends): at most 4096 bodies and 4096 active
 * pairs (8 contact points each), PMH_ERROR_CAPACITY_EXCEEDED beyond.  Calls
 * on one GPU share that array; they are serialised inside the library, and
 * work that a caller replays from a capture of one call must not overlap
 * another call on the same GPU. */
PmhResult pmhPrepareSolveGeometryConstraintsNoAlloc(
    void *stream, const PmhBodyParameters *parameters,
    const PmhRigidBodyState *previous, PmhRigidBodyState *states,
    uint32_t body_count, PmhContactManifold *manifolds,
    const uint32_t *active_pairs, const uint32_t *active_pair_count,
    PmhAvbdBody *bodies, PmhContactSchedule *schedule, float timestep,
    const float gravity[3], uint32_t base_iterations,
    uint32_t requested_iterations, uint32_t profile_version);

/* Device memory helpers for callers without a HIP runtime of their own. */
PmhResult pmhDeviceAlloc(uint64_t bytes, void **pointer);
PmhResult pmhDeviceFree(void *pointer);
PmhResult pmhCopyToDevice(void *device, const void *host, uint64_t bytes);
PmhResult pmhCopyToHost(void *host, const void *device, uint64_t bytes);
PmhResult pmhDeviceSynchronize(void);
/* gfx target name of the current device ("gfx950") and CU count. */
PmhResult pmhDeviceInfo(char *name, uint32_t name_bytes, uint32_t *compute_units);

#ifdef __cplusplus
}
#endif
#endif
