// SPDX-License-Identifier: MIT
//
// The installable HIP driver of the ParallelMater loader for the rigid
// domain: libpm_hip_driver.so exports the four pm_driver* entry points
// (include/parallel_mater/pm_driver.h:34-41) and nothing else, and resolves
// the rigid part of the pm* command API by name onto the MI355X rigid step
// of libpm_hip_rigid.so (include/pm_hip_rigid.h).
//
// Execution model (the reference's, docs/SLANG.md:37-42 and
// src/api/synchronization.cpp:167-240): commands are recorded on the host
// into command buffers; pmQueueSubmit hands a submission to the queue's own
// worker thread, which runs the submissions in order (waits, commands,
// signals, fence).  Every simulation command runs on that worker, so one
// simulation's world is only ever touched by one thread.
//
// Simulations map onto one PmhWorld each.  Rigid-body slots (the API's body
// indices) become dense world bodies in slot order of the active slots;
// constraint slot k is world joint k.  A write that adds, removes or
// replaces bodies rebuilds the world before the next command that needs it,
// keeping the current states of the bodies that stayed (the contact cache
// and joint warm starts then start cold).  Steps use the realtime v1
// profile, the only one pmCmdStepSimulation accepts (dt 1/60, one substep;
// src/api/resources.cpp:1240-1241).
#include "../../include/pm_api_rigid.h"
#include "../../include/pm_hip_rigid.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <unistd.h>
#include <vector>

#define PMD_EXPORT extern "C" __attribute__((visibility("default")))

struct CommandList;
struct Submission;

struct PmInstance_T;
struct PmPhysicalDevice_T {
    PmInstance_T *instance = nullptr;
    int ordinal = 0;
    hipDeviceProp_t prop{};
};
struct PmInstance_T {
    std::vector<std::unique_ptr<PmPhysicalDevice_T>> devices;
};
struct PmQueue_T {
    PmDevice_T *device = nullptr;
    std::mutex m;
    std::condition_variable work, idle;
    std::deque<std::unique_ptr<Submission>> fifo;
    bool busy = false, stop = false;
    std::thread worker;
};
struct PmDevice_T {
    PmPhysicalDevice_T *physical = nullptr;
    PmQueue_T queue;
    PmPhysicalDeviceFeatures enabled{};
    std::atomic<uint64_t> application_bytes{0};
};
struct PmDeviceMemory_T {
    PmDevice_T *device = nullptr;
    PmDeviceSize size = 0;
    uint32_t type = 0;   // 0 device-local, 1 host-visible and coherent
    void *ptr = nullptr;
    bool mapped = false;
};
struct PmBuffer_T {
    PmDevice_T *device = nullptr;
    PmDeviceSize size = 0;
    PmFlags usage = 0;
    PmDeviceMemory_T *memory = nullptr;
    PmDeviceSize offset = 0;
};
struct PmGeometry_T {
    PmDevice_T *device = nullptr;
    PmGeometryCreateInfo info{};
    bool bound = false, built = false;
    std::vector<float> vertices;
    std::vector<uint32_t> indices;
    uint64_t revision = 0;
};
struct PmPhysicsPipeline_T {
    PmDevice_T *device = nullptr;
    PmRigidSolverState rigid{};
};
struct BodySlot {
    bool active = false, fresh = false;
    PmRigidBodyInfo info{};
};
struct JointSlot {
    bool active = false;
    PmRigidConstraintInfo info{};
};
struct ShapeRef {
    uint32_t shape;
    uint64_t revision;
};
struct PmSimulation_T {
    PmDevice_T *device = nullptr;
    PmSimulationCreateInfo info{};
    std::vector<BodySlot> bodies;
    std::vector<JointSlot> joints;
    bool bound = false;
    // world (touched on the queue worker only)
    PmhWorld world = nullptr;
    bool rebuild = true;
    std::map<PmGeometry_T *, ShapeRef> shapes;
    std::vector<uint32_t> dense;     // slot -> dense body, PM_INVALID_INDEX when inactive
    std::vector<uint32_t> slot_of;   // dense body -> slot
    float gravity[3] = {0.0f, -9.81f, 0.0f};
    std::vector<PmContactEvent> events;
    uint64_t event_revision = 0, steps = 0;
    std::atomic<int> pending{0};
    std::atomic<int> result{PM_SUCCESS};
};
struct CommandList {
    struct Command {
        PmSimulation_T *sim;
        std::function<PmResult()> run;
    };
    std::vector<Command> commands;
    std::vector<PmSimulation_T *> sims;   // simulations the list touches (pending counts)
};
struct PmCommandPool_T {
    PmDevice_T *device = nullptr;
    PmFlags flags = 0;
};
struct PmCommandBuffer_T {
    PmCommandPool_T *pool = nullptr;
    enum State { kInitial, kRecording, kExecutable } state = kInitial;
    std::shared_ptr<CommandList> list;
    PmPhysicsPipeline_T *bound = nullptr;
};
struct PmFence_T {
    PmDevice_T *device = nullptr;
    std::mutex m;
    std::condition_variable cv;
    bool signaled = false;
};
struct PmSemaphore_T {
    PmDevice_T *device = nullptr;
    std::mutex m;
    std::condition_variable cv;
    uint64_t value = 0;
};
struct Submission {
    std::vector<std::shared_ptr<CommandList>> lists;
    std::vector<PmSemaphoreSubmitInfo> waits, signals;
    PmFence_T *fence = nullptr;
};

namespace {

constexpr uint32_t kDriverVersion = PM_MAKE_API_VERSION(0, 3, 1);
constexpr uint32_t kMaxRigidBodies = 1u << 24;
constexpr uint32_t kMaxConstraints = 1u << 24;

PmResult from_pmh(PmhResult r) { return static_cast<PmResult>(r); }

bool finite_state(const PmRigidState &s) {
    const float *f = s.position;
    for (int k = 0; k < 13; ++k)
        if (!std::isfinite(f[k])) return false;
    const float q = s.orientation[0] * s.orientation[0] + s.orientation[1] * s.orientation[1] +
                    s.orientation[2] * s.orientation[2] + s.orientation[3] * s.orientation[3];
    return q > 1.0e-12f;
}

bool range_fits(PmDeviceSize offset, PmDeviceSize size, PmDeviceSize total) {
    return offset <= total && size <= total - offset;
}

uint8_t *buffer_pointer(const PmBuffer_T *b, PmDeviceSize offset) {
    return static_cast<uint8_t *>(b->memory->ptr) + b->offset + offset;
}

PmResult write_range(const PmBufferRange &dst, const void *src, size_t bytes) {
    if (!dst.buffer || !dst.buffer->memory) return PM_ERROR_INVALID_STATE;
    if (bytes > dst.size) return PM_ERROR_CAPACITY_EXCEEDED;
    if (!bytes) return PM_SUCCESS;
    return hipMemcpy(buffer_pointer(dst.buffer, dst.offset), src, bytes, hipMemcpyDefault) == hipSuccess
               ? PM_SUCCESS : PM_ERROR_DEVICE_LOST;
}

PmResult read_view(const PmBufferView &v, size_t element, std::vector<uint8_t> &out) {
    if (!v.range.buffer || !v.range.buffer->memory) return PM_ERROR_INVALID_STATE;
    const size_t stride = v.stride ? v.stride : element;
    const size_t bytes = v.count ? stride * (v.count - 1) + element : 0;
    if (bytes > v.range.size) return PM_ERROR_INVALID_ARGUMENT;
    out.resize(bytes);
    if (bytes && hipMemcpy(out.data(), buffer_pointer(v.range.buffer, v.range.offset), bytes, hipMemcpyDefault) !=
                     hipSuccess)
        return PM_ERROR_DEVICE_LOST;
    return PM_SUCCESS;
}

PmPhysicalDeviceFeatures supported_features() {
    PmPhysicalDeviceFeatures f{};
    f.rigidBodies = PM_TRUE;
    f.rigidConstraints = PM_TRUE;
    f.contactEvents = PM_TRUE;
    f.timelineSemaphores = PM_TRUE;
    f.memoryBudget = PM_TRUE;
    return f;
}

bool features_subset(const PmPhysicalDeviceFeatures &want, const PmPhysicalDeviceFeatures &have) {
    const PmBool32 *a = &want.rigidBodies, *b = &have.rigidBodies;
    for (size_t k = 0; k < sizeof(PmPhysicalDeviceFeatures) / sizeof(PmBool32); ++k)
        if (a[k] && !b[k]) return false;
    return true;
}

// The interactions of the rigid domain (registry/pm.xml:108-109): rigid
// contact, and rigid attachment through constraints.
struct Capability {
    PmInteractionCapability cap;
    bool needs_constraints;
};
const Capability kCapabilities[] = {
    {{{PM_INTERACTION_FAMILY_RIGID_BODY, PM_INTERACTION_FAMILY_RIGID_BODY, PM_INTERACTION_SCOPE_SAME_DOMAIN,
       PM_INTERACTION_MODE_CONTACT, PM_INTERACTION_DIRECTION_BIDIRECTIONAL, PM_INTERACTION_MATERIAL_ENDPOINT_NATIVE},
      PM_INTERACTION_GRAPH_EDGES_NOT_APPLICABLE}, false},
    {{{PM_INTERACTION_FAMILY_RIGID_BODY, PM_INTERACTION_FAMILY_RIGID_BODY, PM_INTERACTION_SCOPE_SAME_DOMAIN,
       PM_INTERACTION_MODE_ATTACHMENT, PM_INTERACTION_DIRECTION_BIDIRECTIONAL, PM_INTERACTION_MATERIAL_ENDPOINT_NATIVE},
      PM_INTERACTION_GRAPH_EDGES_NOT_APPLICABLE}, true},
};

bool valid_key(const PmInteractionCapabilityKey &k) {
    const int a = k.firstFamily, b = k.secondFamily;
    if (a < PM_INTERACTION_FAMILY_FLUID || a > PM_INTERACTION_FAMILY_RIGID_BODY || b < PM_INTERACTION_FAMILY_FLUID ||
        b > PM_INTERACTION_FAMILY_RIGID_BODY || k.scope < PM_INTERACTION_SCOPE_SELF ||
        k.scope > PM_INTERACTION_SCOPE_DISTINCT_FAMILY || k.mode < PM_INTERACTION_MODE_CONTACT ||
        k.mode > PM_INTERACTION_MODE_INTERFACE || k.direction < PM_INTERACTION_DIRECTION_BIDIRECTIONAL ||
        k.direction > PM_INTERACTION_DIRECTION_EXTERNAL_TO_SECOND ||
        k.materialProfile < PM_INTERACTION_MATERIAL_ENDPOINT_NATIVE ||
        k.materialProfile > PM_INTERACTION_MATERIAL_ROPE_CAPSULE)
        return false;
    return k.scope == PM_INTERACTION_SCOPE_DISTINCT_FAMILY ? a < b : a == b;
}

bool same_key(const PmInteractionCapabilityKey &x, const PmInteractionCapabilityKey &y) {
    return std::memcmp(&x, &y, sizeof x) == 0;
}

// ---- simulation <-> world ------------------------------------------------

// make_parameters (src/simulation/runtime.cpp:459-482): inverse mass and
// inverse inertia of dynamic bodies, a zero inertia diagonal meaning the
// mesh's unit inertia times the mass.
PmhBodyParameters body_parameters(const PmRigidBodyInfo &b, uint32_t shape, const float unit_inertia[3]) {
    PmhBodyParameters p{};
    p.motion = uint32_t(b.motionType);
    p.mesh_index = shape;
    p.mesh_generation = 1u;
    const bool dynamic = b.motionType == PM_MOTION_DYNAMIC;
    const bool zero = b.inertiaDiagonal[0] == 0.0f && b.inertiaDiagonal[1] == 0.0f && b.inertiaDiagonal[2] == 0.0f;
    float inertia[3];
    for (int k = 0; k < 3; ++k) inertia[k] = zero ? unit_inertia[k] * b.mass : b.inertiaDiagonal[k];
    p.inverse_mass = dynamic ? 1.0f / b.mass : 0.0f;
    p.inverse_inertia_local = dynamic ? PmhVec3{1.0f / inertia[0], 1.0f / inertia[1], 1.0f / inertia[2]}
                                      : PmhVec3{0.0f, 0.0f, 0.0f};
    p.friction = b.friction;
    p.restitution = b.restitution;
    p.linear_damping = b.linearDamping;
    p.angular_damping = b.angularDamping;
    p.maximum_linear_speed = b.maximumLinearSpeed;
    p.maximum_angular_speed = b.maximumAngularSpeed;
    p.collision_margin = b.collisionMargin;
    p.user_data = b.userData;
    return p;
}

bool valid_body(const PmRigidBodyInfo &b) {
    if (!b.geometry || b.motionType > PM_MOTION_DYNAMIC || !finite_state(b.state)) return false;
    const float f[] = {b.mass, b.inertiaDiagonal[0], b.inertiaDiagonal[1], b.inertiaDiagonal[2], b.friction,
                       b.restitution, b.linearDamping, b.angularDamping, b.maximumLinearSpeed,
                       b.maximumAngularSpeed, b.collisionMargin};
    for (float x : f)
        if (!std::isfinite(x) || x < 0.0f) return false;
    if (b.motionType == PM_MOTION_DYNAMIC && !(b.mass > 0.0f)) return false;
    return true;
}

PmhRigidBodyState world_state(const PmRigidState &s) {
    PmhRigidBodyState w;
    static_assert(sizeof w == sizeof s, "PmRigidState and RigidBodyState share a layout");
    std::memcpy(&w, &s, sizeof w);
    return w;
}

// pmCmdWriteRigidConstraints' conversion (src/api/resources.cpp:1029-1060,
// convert_axes :418-436): frames become local anchors and orientations,
// enabled axes become axis bits, the first motor of each kind drives.
PmhRigidConstraintOptions joint_options(const PmRigidConstraintInfo &c, uint32_t a, uint32_t b) {
    PmhRigidConstraintOptions o{};
    o.type = uint32_t(c.type);
    o.body_a = {a, 1u};
    o.body_b = {b, 1u};
    o.local_anchor_a = {c.frameA.position[0], c.frameA.position[1], c.frameA.position[2]};
    o.local_anchor_b = {c.frameB.position[0], c.frameB.position[1], c.frameB.position[2]};
    o.local_orientation_a = {c.frameA.orientation[0], c.frameA.orientation[1], c.frameA.orientation[2],
                             c.frameA.orientation[3]};
    o.local_orientation_b = {c.frameB.orientation[0], c.frameB.orientation[1], c.frameB.orientation[2],
                             c.frameB.orientation[3]};
    auto limits = [](const PmAxisLimit *src, PmhJointLimits &dst) {
        float *lo = &dst.lower.x, *hi = &dst.upper.x;
        for (uint32_t k = 0; k < 3; ++k) {
            if (src[k].enabled == PM_TRUE) dst.axes |= 1u << k;
            lo[k] = src[k].lower;
            hi[k] = src[k].upper;
        }
    };
    auto springs = [](const PmAxisSpring *src, PmhJointSprings &dst) {
        float *k_ = &dst.stiffness.x, *d = &dst.damping.x;
        for (uint32_t k = 0; k < 3; ++k) {
            if (src[k].enabled == PM_TRUE) dst.axes |= 1u << k;
            k_[k] = src[k].stiffness;
            d[k] = src[k].damping;
        }
    };
    limits(c.linearLimits, o.linear_limits);
    limits(c.angularLimits, o.angular_limits);
    springs(c.linearSprings, o.linear_springs);
    springs(c.angularSprings, o.angular_springs);
    o.motor.linear_enabled = c.linearMotors[0].enabled == PM_TRUE ? 1u : 0u;
    o.motor.linear_target_velocity = c.linearMotors[0].targetVelocity;
    o.motor.linear_maximum_impulse = c.linearMotors[0].maximumForce;
    o.motor.angular_enabled = c.angularMotors[0].enabled == PM_TRUE ? 1u : 0u;
    o.motor.angular_target_velocity = c.angularMotors[0].targetVelocity;
    o.motor.angular_maximum_impulse = c.angularMotors[0].maximumForce;
    o.enabled = 1u;
    o.disable_collisions = c.collideConnected == PM_FALSE ? 1u : 0u;
    o.breaking_impulse_threshold = c.breakingImpulseThreshold;
    o.solver_iterations = c.solverIterations;
    return o;
}

// Options of every constraint slot (inactive slots as placeholders) and the
// active flags, with body slots resolved to dense bodies.
PmResult slot_joints(PmSimulation_T *s, std::vector<PmhRigidConstraintOptions> &opts, std::vector<uint32_t> &active) {
    const size_t cap = s->joints.size();
    opts.assign(cap, PmhRigidConstraintOptions{});
    active.assign(cap, 0u);
    for (size_t k = 0; k < cap; ++k) {
        const JointSlot &j = s->joints[k];
        if (!j.active) continue;
        const uint32_t a = j.info.bodyA < s->dense.size() ? s->dense[j.info.bodyA] : PM_INVALID_INDEX;
        const uint32_t b = j.info.bodyB < s->dense.size() ? s->dense[j.info.bodyB] : PM_INVALID_INDEX;
        if (a == PM_INVALID_INDEX || b == PM_INVALID_INDEX) return PM_ERROR_INVALID_ARGUMENT;
        opts[k] = joint_options(j.info, a, b);
        active[k] = 1u;
    }
    return PM_SUCCESS;
}

PmResult register_shape(PmSimulation_T *s, PmGeometry_T *g, uint32_t &shape, float unit[3]) {
    if (!g->built) return PM_ERROR_INVALID_STATE;
    auto it = s->shapes.find(g);
    if (it == s->shapes.end() || it->second.revision != g->revision) {
        uint32_t out = 0;
        PmResult r = from_pmh(pmhWorldAddMesh(s->world, g->vertices.data(), uint32_t(g->vertices.size() / 3),
                                              g->indices.data(), uint32_t(g->indices.size()), &out));
        if (r != PM_SUCCESS) return r;
        s->shapes[g] = {out, g->revision};
        shape = out;
    } else {
        shape = it->second.shape;
    }
    float info[16];
    PmResult r = from_pmh(pmhWorldGetShapeInfo(s->world, shape, info));
    if (r != PM_SUCCESS) return r;
    unit[0] = info[10];
    unit[1] = info[11];
    unit[2] = info[12];
    return PM_SUCCESS;
}

// Builds (or rebuilds) the simulation's world from its slots; the bodies
// that stayed keep their current states.
PmResult ensure_world(PmSimulation_T *s) {
    bool stale = s->rebuild;
    for (const auto &e : s->shapes)
        if (e.first->revision != e.second.revision) stale = true;
    if (s->world && !stale) return PM_SUCCESS;
    std::vector<PmhRigidBodyState> current;
    std::vector<uint32_t> old_dense = s->dense;
    if (s->world) {
        current.resize(pmhWorldBodyCount(s->world));
        if (!current.empty()) {
            PmResult r = from_pmh(pmhWorldCopyStates(s->world, current.data()));
            if (r != PM_SUCCESS) return r;
        }
        pmhDestroyWorld(s->world);
        s->world = nullptr;
    }
    s->shapes.clear();
    PmhWorldDesc desc{};
    std::memcpy(desc.gravity, s->gravity, sizeof desc.gravity);
    desc.timestep = 1.0f / 60.0f;
    desc.profile_version = 1;
    desc.base_iterations = 10;
    desc.device = uint32_t(s->device->physical->ordinal);
    desc.substeps = 1;
    PmResult r = from_pmh(pmhCreateWorld(&desc, &s->world));
    if (r != PM_SUCCESS) {
        s->world = nullptr;
        return r;
    }
    const size_t cap = s->bodies.size();
    std::vector<PmhBodyParameters> params;
    std::vector<PmhRigidBodyState> states;
    s->dense.assign(cap, PM_INVALID_INDEX);
    s->slot_of.clear();
    for (size_t k = 0; k < cap; ++k) {
        BodySlot &b = s->bodies[k];
        if (!b.active) continue;
        uint32_t shape = 0;
        float unit[3];
        r = register_shape(s, b.info.geometry, shape, unit);
        if (r != PM_SUCCESS) return r;
        params.push_back(body_parameters(b.info, shape, unit));
        const uint32_t od = k < old_dense.size() ? old_dense[k] : PM_INVALID_INDEX;
        states.push_back(!b.fresh && od != PM_INVALID_INDEX && od < current.size() ? current[od]
                                                                                   : world_state(b.info.state));
        s->dense[k] = uint32_t(s->slot_of.size());
        s->slot_of.push_back(uint32_t(k));
        b.fresh = false;
    }
    if (!params.empty()) {
        r = from_pmh(pmhWorldAddBodies(s->world, uint32_t(params.size()), params.data(), states.data()));
        if (r != PM_SUCCESS) return r;
    }
    if (!s->joints.empty()) {
        std::vector<PmhRigidConstraintOptions> opts;
        std::vector<uint32_t> active;
        r = slot_joints(s, opts, active);
        if (r != PM_SUCCESS) return r;
        r = from_pmh(pmhWorldAddJoints(s->world, uint32_t(opts.size()), opts.data()));
        if (r == PM_SUCCESS)
            r = from_pmh(pmhWorldWriteJoints(s->world, 0, uint32_t(opts.size()), opts.data(), active.data()));
        if (r != PM_SUCCESS) return r;
    }
    s->events.clear();
    s->rebuild = false;
    return PM_SUCCESS;
}

// ---- queue ---------------------------------------------------------------

void run_submission(Submission &sub) {
    for (const PmSemaphoreSubmitInfo &w : sub.waits) {
        std::unique_lock<std::mutex> lock(w.semaphore->m);
        w.semaphore->cv.wait(lock, [&] { return w.semaphore->value >= w.value; });
    }
    for (auto &list : sub.lists) {
        for (auto &c : list->commands) {
            const PmResult r = c.run();
            if (r != PM_SUCCESS && c.sim) c.sim->result.store(r);
        }
        for (PmSimulation_T *s : list->sims) s->pending.fetch_sub(1);
    }
    for (const PmSemaphoreSubmitInfo &sg : sub.signals) {
        std::lock_guard<std::mutex> lock(sg.semaphore->m);
        sg.semaphore->value = std::max(sg.semaphore->value, sg.value);
        sg.semaphore->cv.notify_all();
    }
    if (sub.fence) {
        std::lock_guard<std::mutex> lock(sub.fence->m);
        sub.fence->signaled = true;
        sub.fence->cv.notify_all();
    }
}

void worker_main(PmQueue_T *q) {
    (void)hipSetDevice(q->device->physical->ordinal);
    for (;;) {
        std::unique_ptr<Submission> sub;
        {
            std::unique_lock<std::mutex> lock(q->m);
            q->work.wait(lock, [&] { return q->stop || !q->fifo.empty(); });
            if (q->fifo.empty()) return;
            sub = std::move(q->fifo.front());
            q->fifo.pop_front();
            q->busy = true;
        }
        run_submission(*sub);
        {
            std::lock_guard<std::mutex> lock(q->m);
            q->busy = false;
            if (q->fifo.empty()) q->idle.notify_all();
        }
    }
}

PmResult queue_wait_idle(PmQueue_T *q) {
    std::unique_lock<std::mutex> lock(q->m);
    q->idle.wait(lock, [&] { return q->fifo.empty() && !q->busy; });
    return PM_SUCCESS;
}

PmResult recording(PmCommandBuffer_T *cb) {
    if (!cb) return PM_ERROR_INVALID_ARGUMENT;
    return cb->state == PmCommandBuffer_T::kRecording ? PM_SUCCESS : PM_ERROR_INVALID_STATE;
}

PmResult record(PmCommandBuffer_T *cb, PmSimulation_T *sim, std::function<PmResult()> run) {
    try {
        cb->list->commands.push_back({sim, std::move(run)});
        if (sim && std::find(cb->list->sims.begin(), cb->list->sims.end(), sim) == cb->list->sims.end())
            cb->list->sims.push_back(sim);
    } catch (const std::bad_alloc &) {
        return PM_ERROR_OUT_OF_HOST_MEMORY;
    }
    return PM_SUCCESS;
}

bool sim_of(PmCommandBuffer_T *cb, PmSimulation_T *s) { return s && s->device == cb->pool->device; }

bool destination_ok(PmCommandBuffer_T *cb, const PmBufferRange *d, PmDeviceSize minimum) {
    return d && d->buffer && d->buffer->device == cb->pool->device && d->size >= minimum &&
           range_fits(d->offset, d->size, d->buffer->size);
}

}  // namespace

// ---- instance and physical devices ------------------------------------------
namespace pmd {

PmResult EnumerateInstanceVersion(uint32_t *version) {
    if (!version) return PM_ERROR_INVALID_ARGUMENT;
    *version = PM_API_VERSION_0_3;
    return PM_SUCCESS;
}

PmResult CreateInstance(const PmInstanceCreateInfo *info, const PmAllocationCallbacks *, PmInstance *out) {
    if (!out) return PM_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!info || info->sType != PM_STRUCTURE_TYPE_INSTANCE_CREATE_INFO) return PM_ERROR_INVALID_ARGUMENT;
    if (info->apiVersion != 0u && (info->apiVersion >> 22) != 0u) return PM_ERROR_INCOMPATIBLE_VERSION;
    if (info->enabledExtensionCount != 0u) return PM_ERROR_EXTENSION_NOT_PRESENT;
    auto *inst = new (std::nothrow) PmInstance_T();
    if (!inst) return PM_ERROR_OUT_OF_HOST_MEMORY;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
    for (int d = 0; d < count; ++d) {
        auto pd = std::make_unique<PmPhysicalDevice_T>();
        pd->instance = inst;
        pd->ordinal = d;
        if (hipGetDeviceProperties(&pd->prop, d) != hipSuccess) continue;
        inst->devices.push_back(std::move(pd));
    }
    *out = inst;
    return PM_SUCCESS;
}

void DestroyInstance(PmInstance instance) { delete instance; }

PmResult EnumeratePhysicalDevices(PmInstance instance, uint32_t *count, PmPhysicalDevice *devices) {
    if (!instance || !count) return PM_ERROR_INVALID_ARGUMENT;
    const uint32_t available = uint32_t(instance->devices.size());
    if (!devices) {
        *count = available;
        return PM_SUCCESS;
    }
    const uint32_t written = std::min(*count, available);
    for (uint32_t k = 0; k < written; ++k) devices[k] = instance->devices[k].get();
    *count = written;
    return written == available ? PM_SUCCESS : PM_INCOMPLETE;
}

void GetPhysicalDeviceProperties(PmPhysicalDevice pd, PmPhysicalDeviceProperties *p) {
    if (!p) return;
    *p = {};
    if (!pd) return;
    p->apiVersion = PM_API_VERSION_0_3;
    p->driverVersion = kDriverVersion;
    p->vendorId = 0x1002u;
    p->deviceId = uint32_t(pd->prop.pciDeviceID);
    std::memcpy(p->deviceUuid, pd->prop.uuid.bytes, sizeof p->deviceUuid);
    std::snprintf(p->deviceName, sizeof p->deviceName, "%s (%s, HIP rigid driver)", pd->prop.name,
                  pd->prop.gcnArchName);
}

PmResult GetPhysicalDeviceQueueFamilies(PmPhysicalDevice pd, uint32_t *count, PmQueueFamilyProperties *props) {
    if (!pd || !count) return PM_ERROR_INVALID_ARGUMENT;
    if (!props) {
        *count = 1u;
        return PM_SUCCESS;
    }
    if (*count == 0u) return PM_INCOMPLETE;
    props[0] = {PM_QUEUE_PHYSICS_BIT | PM_QUEUE_TRANSFER_BIT, 1u, 64u, 1.0f};
    *count = 1u;
    return PM_SUCCESS;
}

void GetPhysicalDeviceFeatures(PmPhysicalDevice pd, PmPhysicalDeviceFeatures *f) {
    if (!f) return;
    *f = pd ? supported_features() : PmPhysicalDeviceFeatures{};
}

void GetPhysicalDeviceLimits(PmPhysicalDevice pd, PmPhysicalDeviceLimits *l) {
    if (!l) return;
    *l = {};
    if (!pd) return;
    l->maxAllocationSize = pd->prop.totalGlobalMem;
    l->maxBufferSize = pd->prop.totalGlobalMem;
    l->nonCoherentAtomSize = 1u;
    l->maxMemoryAllocations = 1u << 20;
    l->maxRigidBodies = kMaxRigidBodies;
    l->maxRigidConstraints = kMaxConstraints;
    l->maxContacts = 1u << 28;
    l->maxRigidIterations = 64u;
    l->maxSubsteps = 1u;
    l->maxTimelineValueDifference = UINT64_MAX;
}

void GetPhysicalDeviceMemoryProperties(PmPhysicalDevice pd, PmMemoryProperties *m) {
    if (!m) return;
    *m = {};
    if (!pd) return;
    m->heapCount = 2;
    m->heaps[0] = {pd->prop.totalGlobalMem, PM_MEMORY_DEVICE_LOCAL_BIT};
    m->heaps[1] = {PmDeviceSize(sysconf(_SC_PHYS_PAGES)) * PmDeviceSize(sysconf(_SC_PAGE_SIZE)), 0u};
    m->typeCount = 2;
    m->types[0] = {PM_MEMORY_DEVICE_LOCAL_BIT, 0u};
    m->types[1] = {PM_MEMORY_HOST_VISIBLE_BIT | PM_MEMORY_HOST_COHERENT_BIT, 1u};
}

PmResult EnumeratePhysicalDeviceInteractionCapabilities(PmPhysicalDevice pd, uint32_t *count,
                                                        PmInteractionCapability *caps) {
    if (!pd || !count) return PM_ERROR_INVALID_ARGUMENT;
    const uint32_t available = uint32_t(sizeof kCapabilities / sizeof kCapabilities[0]);
    if (!caps) {
        *count = available;
        return PM_SUCCESS;
    }
    const uint32_t written = std::min(*count, available);
    for (uint32_t k = 0; k < written; ++k) caps[k] = kCapabilities[k].cap;
    *count = written;
    return written == available ? PM_SUCCESS : PM_INCOMPLETE;
}

PmResult GetPhysicalDeviceInteractionCapability(PmPhysicalDevice pd, const PmInteractionCapabilityKey *key,
                                                PmInteractionCapability *cap) {
    if (!key || !cap) return PM_ERROR_INVALID_ARGUMENT;
    const PmInteractionCapabilityKey k = *key;
    *cap = {};
    if (!pd || !valid_key(k)) return PM_ERROR_INVALID_ARGUMENT;
    for (const Capability &c : kCapabilities)
        if (same_key(c.cap.key, k)) {
            *cap = c.cap;
            return PM_SUCCESS;
        }
    return PM_ERROR_FEATURE_NOT_PRESENT;
}

PmResult EnumerateDeviceExtensionProperties(PmPhysicalDevice pd, const char *layer, uint32_t *count,
                                            PmExtensionProperties *) {
    if (!pd || !count) return PM_ERROR_INVALID_ARGUMENT;
    if (layer) return PM_ERROR_LAYER_NOT_PRESENT;
    *count = 0;
    return PM_SUCCESS;
}

// ---- device, queue ----------------------------------------------------------

PmResult CreateDevice(PmPhysicalDevice pd, const PmDeviceCreateInfo *info, PmDevice *out) {
    if (!out) return PM_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!pd || !info || info->sType != PM_STRUCTURE_TYPE_DEVICE_CREATE_INFO) return PM_ERROR_INVALID_ARGUMENT;
    if (info->queueCreateInfoCount > 1u || (info->queueCreateInfoCount && !info->pQueueInfos))
        return PM_ERROR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < info->queueCreateInfoCount; ++k)
        if (info->pQueueInfos[k].queueFamilyIndex != 0u || info->pQueueInfos[k].queueCount > 1u)
            return PM_ERROR_INVALID_ARGUMENT;
    if (info->enabledExtensionCount) return PM_ERROR_EXTENSION_NOT_PRESENT;
    if (!features_subset(info->enabledFeatures, supported_features())) return PM_ERROR_FEATURE_NOT_PRESENT;
    auto *d = new (std::nothrow) PmDevice_T();
    if (!d) return PM_ERROR_OUT_OF_HOST_MEMORY;
    d->physical = pd;
    d->enabled = info->enabledFeatures;
    d->queue.device = d;
    d->queue.worker = std::thread(worker_main, &d->queue);
    *out = d;
    return PM_SUCCESS;
}

void DestroyDevice(PmDevice d) {
    if (!d) return;
    {
        std::lock_guard<std::mutex> lock(d->queue.m);
        d->queue.stop = true;
        d->queue.work.notify_all();
    }
    if (d->queue.worker.joinable()) d->queue.worker.join();
    delete d;
}

void GetDeviceQueue(PmDevice d, uint32_t family, uint32_t index, PmQueue *q) {
    if (!q) return;
    *q = (d && family == 0u && index == 0u) ? &d->queue : nullptr;
}

PmResult QueueWaitIdle(PmQueue q) { return q ? queue_wait_idle(q) : PM_ERROR_INVALID_ARGUMENT; }
PmResult DeviceWaitIdle(PmDevice d) { return d ? queue_wait_idle(&d->queue) : PM_ERROR_INVALID_ARGUMENT; }

PmResult QueueSubmit(PmQueue q, const PmSubmitInfo *info, PmFence fence) {
    if (!q || !info || info->sType != PM_STRUCTURE_TYPE_SUBMIT_INFO) return PM_ERROR_INVALID_ARGUMENT;
    if ((info->commandBufferCount && !info->pCommandBuffers) || (info->waitCount && !info->pWaits) ||
        (info->signalCount && !info->pSignals))
        return PM_ERROR_INVALID_ARGUMENT;
    auto sub = std::make_unique<Submission>();
    for (uint32_t k = 0; k < info->commandBufferCount; ++k) {
        PmCommandBuffer_T *cb = info->pCommandBuffers[k];
        if (!cb || cb->pool->device != q->device) return PM_ERROR_INVALID_ARGUMENT;
        if (cb->state != PmCommandBuffer_T::kExecutable) return PM_ERROR_INVALID_STATE;
        sub->lists.push_back(cb->list);
    }
    for (uint32_t k = 0; k < info->waitCount; ++k) {
        if (!info->pWaits[k].semaphore || info->pWaits[k].semaphore->device != q->device)
            return PM_ERROR_INVALID_ARGUMENT;
        sub->waits.push_back(info->pWaits[k]);
    }
    for (uint32_t k = 0; k < info->signalCount; ++k) {
        if (!info->pSignals[k].semaphore || info->pSignals[k].semaphore->device != q->device)
            return PM_ERROR_INVALID_ARGUMENT;
        sub->signals.push_back(info->pSignals[k]);
    }
    if (fence) {
        if (fence->device != q->device) return PM_ERROR_INVALID_ARGUMENT;
        std::lock_guard<std::mutex> lock(fence->m);
        if (fence->signaled) return PM_ERROR_INVALID_STATE;
    }
    sub->fence = fence;
    for (auto &list : sub->lists)
        for (PmSimulation_T *s : list->sims) s->pending.fetch_add(1);
    {
        std::lock_guard<std::mutex> lock(q->m);
        q->fifo.push_back(std::move(sub));
        q->work.notify_one();
    }
    return PM_SUCCESS;
}

// ---- fences, semaphores ------------------------------------------------------

PmResult CreateFence(PmDevice d, PmFlags flags, PmFence *out) {
    if (!out) return PM_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!d || (flags & ~PmFlags(PM_FENCE_SIGNALED_BIT))) return PM_ERROR_INVALID_ARGUMENT;
    auto *f = new (std::nothrow) PmFence_T();
    if (!f) return PM_ERROR_OUT_OF_HOST_MEMORY;
    f->device = d;
    f->signaled = (flags & PM_FENCE_SIGNALED_BIT) != 0u;
    *out = f;
    return PM_SUCCESS;
}

void DestroyFence(PmDevice d, PmFence f) {
    if (!d || !f || f->device != d) return;
    (void)queue_wait_idle(&d->queue);
    delete f;
}

PmResult GetFenceStatus(PmDevice d, PmFence f) {
    if (!d || !f || f->device != d) return PM_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(f->m);
    return f->signaled ? PM_SUCCESS : PM_NOT_READY;
}

PmResult ResetFences(PmDevice d, uint32_t count, const PmFence *fences) {
    if (!d || (count && !fences)) return PM_ERROR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < count; ++k) {
        if (!fences[k] || fences[k]->device != d) return PM_ERROR_INVALID_ARGUMENT;
        std::lock_guard<std::mutex> lock(fences[k]->m);
        fences[k]->signaled = false;
    }
    return PM_SUCCESS;
}

PmResult WaitForFences(PmDevice d, uint32_t count, const PmFence *fences, PmBool32 all, uint64_t timeout_ns) {
    if (!d || !count || !fences) return PM_ERROR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < count; ++k)
        if (!fences[k] || fences[k]->device != d) return PM_ERROR_INVALID_ARGUMENT;
    const auto deadline = timeout_ns == PM_TIMEOUT_INFINITE
                              ? std::chrono::steady_clock::time_point::max()
                              : std::chrono::steady_clock::now() + std::chrono::nanoseconds(timeout_ns);
    for (;;) {
        uint32_t done = 0;
        for (uint32_t k = 0; k < count; ++k) {
            std::lock_guard<std::mutex> lock(fences[k]->m);
            done += fences[k]->signaled ? 1u : 0u;
        }
        if ((all && done == count) || (!all && done > 0)) return PM_SUCCESS;
        if (std::chrono::steady_clock::now() >= deadline) return PM_TIMEOUT;
        // Wait on the first fence not yet signalled, a little at a time.
        for (uint32_t k = 0; k < count; ++k) {
            std::unique_lock<std::mutex> lock(fences[k]->m);
            if (fences[k]->signaled) continue;
            const auto until = std::min(deadline, std::chrono::steady_clock::now() + std::chrono::milliseconds(5));
            fences[k]->cv.wait_until(lock, until, [&] { return fences[k]->signaled; });
            break;
        }
    }
}

PmResult CreateSemaphore(PmDevice d, const PmSemaphoreCreateInfo *info, PmSemaphore *out) {
    if (!out) return PM_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!d || !info || info->sType != PM_STRUCTURE_TYPE_SEMAPHORE_CREATE_INFO) return PM_ERROR_INVALID_ARGUMENT;
    auto *s = new (std::nothrow) PmSemaphore_T();
    if (!s) return PM_ERROR_OUT_OF_HOST_MEMORY;
    s->device = d;
    s->value = info->initialValue;
    *out = s;
    return PM_SUCCESS;
}

void DestroySemaphore(PmDevice d, PmSemaphore s) {
    if (!d || !s || s->device != d) return;
    (void)queue_wait_idle(&d->queue);
    delete s;
}

PmResult GetSemaphoreCounterValue(PmDevice d, PmSemaphore s, uint64_t *value) {
    if (!d || !s || !value || s->device != d) return PM_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(s->m);
    *value = s->value;
    return PM_SUCCESS;
}

PmResult WaitSemaphore(PmDevice d, PmSemaphore s, uint64_t value, uint64_t timeout_ns) {
    if (!d || !s || s->device != d) return PM_ERROR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lock(s->m);
    if (timeout_ns == PM_TIMEOUT_INFINITE) {
        s->cv.wait(lock, [&] { return s->value >= value; });
        return PM_SUCCESS;
    }
    return s->cv.wait_for(lock, std::chrono::nanoseconds(timeout_ns), [&] { return s->value >= value; })
               ? PM_SUCCESS : PM_TIMEOUT;
}

// ---- memory, buffers -----------------------------------------------------------

PmResult AllocateMemory(PmDevice d, const PmMemoryAllocateInfo *info, PmDeviceMemory *out) {
    if (!out) return PM_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!d || !info || info->sType != PM_STRUCTURE_TYPE_MEMORY_ALLOCATE_INFO || info->allocationSize == 0u ||
        info->memoryTypeIndex > 1u)
        return PM_ERROR_INVALID_ARGUMENT;
    auto *m = new (std::nothrow) PmDeviceMemory_T();
    if (!m) return PM_ERROR_OUT_OF_HOST_MEMORY;
    m->device = d;
    m->size = info->allocationSize;
    m->type = info->memoryTypeIndex;
    (void)hipSetDevice(d->physical->ordinal);
    const bool host = m->type == 1u;
    const hipError_t e = host ? hipHostMalloc(&m->ptr, m->size, hipHostMallocDefault) : hipMalloc(&m->ptr, m->size);
    if (e != hipSuccess) {
        delete m;
        return host ? PM_ERROR_OUT_OF_HOST_MEMORY : PM_ERROR_OUT_OF_DEVICE_MEMORY;
    }
    d->application_bytes.fetch_add(m->size);
    *out = m;
    return PM_SUCCESS;
}

void FreeMemory(PmDevice d, PmDeviceMemory m) {
    if (!d || !m || m->device != d) return;
    (void)queue_wait_idle(&d->queue);
    if (m->type == 1u) (void)hipHostFree(m->ptr);
    else (void)hipFree(m->ptr);
    d->application_bytes.fetch_sub(m->size);
    delete m;
}

PmResult MapMemory(PmDevice d, PmDeviceMemory m, PmDeviceSize offset, PmDeviceSize size, void **data) {
    if (!data) return PM_ERROR_INVALID_ARGUMENT;
    *data = nullptr;
    if (!d || !m || m->device != d || m->type != 1u) return PM_ERROR_INVALID_ARGUMENT;
    if (size == PM_WHOLE_SIZE) size = offset <= m->size ? m->size - offset : 0;
    if (!range_fits(offset, size, m->size)) return PM_ERROR_INVALID_ARGUMENT;
    if (m->mapped) return PM_ERROR_INVALID_STATE;
    m->mapped = true;
    *data = static_cast<uint8_t *>(m->ptr) + offset;
    return PM_SUCCESS;
}

void UnmapMemory(PmDevice d, PmDeviceMemory m) {
    if (d && m && m->device == d) m->mapped = false;
}

// Host-visible memory is coherent: flushes and invalidations only check
// their ranges.
PmResult CheckMappedRanges(PmDevice d, uint32_t count, const PmMappedMemoryRange *ranges) {
    if (!d || (count && !ranges)) return PM_ERROR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < count; ++k) {
        const PmMappedMemoryRange &r = ranges[k];
        if (!r.memory || r.memory->device != d || r.memory->type != 1u) return PM_ERROR_INVALID_ARGUMENT;
        const PmDeviceSize size = r.size == PM_WHOLE_SIZE ? (r.offset <= r.memory->size ? r.memory->size - r.offset : 0)
                                                          : r.size;
        if (!range_fits(r.offset, size, r.memory->size)) return PM_ERROR_INVALID_ARGUMENT;
    }
    return PM_SUCCESS;
}

PmResult CreateBuffer(PmDevice d, const PmBufferCreateInfo *info, PmBuffer *out) {
    if (!out) return PM_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!d || !info || info->sType != PM_STRUCTURE_TYPE_BUFFER_CREATE_INFO || info->size == 0u ||
        info->initialQueueFamilyIndex != 0u)
        return PM_ERROR_INVALID_ARGUMENT;
    auto *b = new (std::nothrow) PmBuffer_T();
    if (!b) return PM_ERROR_OUT_OF_HOST_MEMORY;
    b->device = d;
    b->size = info->size;
    b->usage = info->usage;
    *out = b;
    return PM_SUCCESS;
}

void DestroyBuffer(PmDevice d, PmBuffer b) {
    if (!d || !b || b->device != d) return;
    (void)queue_wait_idle(&d->queue);
    delete b;
}

void GetBufferMemoryRequirements(PmDevice d, PmBuffer b, PmMemoryRequirements *req) {
    if (!req) return;
    *req = {};
    if (!d || !b || b->device != d) return;
    req->size = b->size;
    req->alignment = 256u;
    req->memoryTypeBits = 0x3u;
}

PmResult BindBufferMemory(PmDevice d, PmBuffer b, PmDeviceMemory m, PmDeviceSize offset) {
    if (!d || !b || !m || b->device != d || m->device != d || offset % 256u != 0u ||
        !range_fits(offset, b->size, m->size))
        return PM_ERROR_INVALID_ARGUMENT;
    if (b->memory) return PM_ERROR_INVALID_STATE;
    b->memory = m;
    b->offset = offset;
    return PM_SUCCESS;
}

// ---- geometry -----------------------------------------------------------------

PmResult CreateGeometry(PmDevice d, const PmGeometryCreateInfo *info, PmGeometry *out) {
    if (!out) return PM_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!d || !info || info->sType != PM_STRUCTURE_TYPE_GEOMETRY_CREATE_INFO || info->queueFamilyIndex != 0u)
        return PM_ERROR_INVALID_ARGUMENT;
    const PmBufferView &v = info->vertices, &i = info->indices;
    if (!v.range.buffer || !i.range.buffer || v.range.buffer->device != d || i.range.buffer->device != d)
        return PM_ERROR_INVALID_ARGUMENT;
    if (v.format != PM_FORMAT_R32G32B32_SFLOAT || i.format != PM_FORMAT_R32_UINT) return PM_ERROR_FORMAT_NOT_SUPPORTED;
    if (v.count < 3u || i.count < 3u || i.count % 3u != 0u || (v.stride && v.stride < 12u) ||
        (i.stride && i.stride < 4u))
        return PM_ERROR_INVALID_ARGUMENT;
    auto *g = new (std::nothrow) PmGeometry_T();
    if (!g) return PM_ERROR_OUT_OF_HOST_MEMORY;
    g->device = d;
    g->info = *info;
    *out = g;
    return PM_SUCCESS;
}

void DestroyGeometry(PmDevice d, PmGeometry g) {
    if (!d || !g || g->device != d) return;
    (void)queue_wait_idle(&d->queue);
    delete g;
}

// The mesh data live in the application's vertex and index buffers; the
// driver keeps no device memory of its own for a geometry.
void GetGeometryMemoryRequirements(PmDevice d, PmGeometry g, PmMemoryRequirements *req) {
    if (!req) return;
    *req = {};
    if (!d || !g || g->device != d) return;
    req->size = 256u;
    req->alignment = 256u;
    req->memoryTypeBits = 0x1u;
}

PmResult BindGeometryMemory(PmDevice d, PmGeometry g, PmDeviceMemory m, PmDeviceSize offset) {
    if (!d || !g || !m || g->device != d || m->device != d || !range_fits(offset, 256u, m->size))
        return PM_ERROR_INVALID_ARGUMENT;
    g->bound = true;
    return PM_SUCCESS;
}

PmResult CmdBuildGeometry(PmCommandBuffer cb, PmGeometry g) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!g || g->device != cb->pool->device) return PM_ERROR_INVALID_ARGUMENT;
    return record(cb, nullptr, [g]() -> PmResult {
        if (!g->bound) return PM_ERROR_INVALID_STATE;
        std::vector<uint8_t> vb, ib;
        PmResult r = read_view(g->info.vertices, 12u, vb);
        if (r == PM_SUCCESS) r = read_view(g->info.indices, 4u, ib);
        if (r != PM_SUCCESS) return r;
        const PmBufferView &v = g->info.vertices, &i = g->info.indices;
        const size_t vs = v.stride ? v.stride : 12u, is = i.stride ? i.stride : 4u;
        std::vector<float> verts(3 * size_t(v.count));
        std::vector<uint32_t> idx(i.count);
        for (uint32_t k = 0; k < v.count; ++k) std::memcpy(&verts[3 * size_t(k)], vb.data() + vs * k, 12u);
        for (uint32_t k = 0; k < i.count; ++k) {
            std::memcpy(&idx[k], ib.data() + is * k, 4u);
            if (idx[k] >= v.count) return PM_ERROR_INVALID_ARGUMENT;
        }
        for (float x : verts)
            if (!std::isfinite(x)) return PM_ERROR_INVALID_ARGUMENT;
        g->vertices.swap(verts);
        g->indices.swap(idx);
        g->built = true;
        ++g->revision;
        return PM_SUCCESS;
    });
}

// ---- simulation ------------------------------------------------------------------

PmResult CreateSimulation(PmDevice d, const PmSimulationCreateInfo *info, PmSimulation *out) {
    if (!out) return PM_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!d || !info || info->sType != PM_STRUCTURE_TYPE_SIMULATION_CREATE_INFO || info->queueFamilyIndex != 0u)
        return PM_ERROR_INVALID_ARGUMENT;
    if (info->fluidCount || info->clothCount || info->ropeCount || info->softBodyCount || info->smokeCount ||
        info->couplingCapacity || info->emitterCapacity || info->destroyerCapacity || info->paintFieldCount ||
        info->paintRuleCapacity)
        return PM_ERROR_FEATURE_NOT_PRESENT;
    if (info->rigidBodyCapacity > kMaxRigidBodies || info->constraintCapacity > kMaxConstraints)
        return PM_ERROR_CAPACITY_EXCEEDED;
    if (info->constraintCapacity && !d->enabled.rigidConstraints) return PM_ERROR_FEATURE_NOT_PRESENT;
    auto *s = new (std::nothrow) PmSimulation_T();
    if (!s) return PM_ERROR_OUT_OF_HOST_MEMORY;
    try {
        s->bodies.resize(info->rigidBodyCapacity);
        s->joints.resize(info->constraintCapacity);
    } catch (const std::bad_alloc &) {
        delete s;
        return PM_ERROR_OUT_OF_HOST_MEMORY;
    }
    s->device = d;
    s->info = *info;
    s->info.pNext = nullptr;
    *out = s;
    return PM_SUCCESS;
}

void DestroySimulation(PmDevice d, PmSimulation s) {
    if (!d || !s || s->device != d) return;
    (void)queue_wait_idle(&d->queue);
    if (s->world) pmhDestroyWorld(s->world);
    delete s;
}

// The world allocates its own pools (sized from the body count at the first
// step); the application's binding is checked and kept, and the device
// memory budget reports the library's allocations apart.
void GetSimulationMemoryRequirements(PmDevice d, PmSimulation s, PmSimulationMemoryRequirements *req) {
    if (!req) return;
    *req = {};
    if (!d || !s || s->device != d) return;
    req->persistent = {256u, 256u, 0x1u, PM_FALSE};
    req->scratch = {256u, 256u, 0x1u, PM_FALSE};
}

PmResult BindSimulationMemory(PmDevice d, PmSimulation s, const PmSimulationMemoryBinding *b) {
    if (!d || !s || !b || s->device != d || !b->persistentMemory || !b->scratchMemory ||
        b->persistentMemory->device != d || b->scratchMemory->device != d ||
        !range_fits(b->persistentOffset, 256u, b->persistentMemory->size) ||
        !range_fits(b->scratchOffset, 256u, b->scratchMemory->size))
        return PM_ERROR_INVALID_ARGUMENT;
    if (s->bound) return PM_ERROR_INVALID_STATE;
    s->bound = true;
    return PM_SUCCESS;
}

PmResult GetSimulationStatus(PmDevice d, PmSimulation s, PmResult *execution) {
    if (!execution) return PM_ERROR_INVALID_ARGUMENT;
    *execution = PM_ERROR_FEATURE_NOT_PRESENT;
    if (!d || !s || s->device != d) return PM_ERROR_INVALID_ARGUMENT;
    if (s->pending.load() != 0) return PM_NOT_READY;
    *execution = PmResult(s->result.load());
    return PM_SUCCESS;
}

PmResult CreatePhysicsPipeline(PmDevice d, const PmPhysicsPipelineCreateInfo *info, PmPhysicsPipeline *out) {
    if (!out) return PM_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!d || !info || info->sType != PM_STRUCTURE_TYPE_PHYSICS_PIPELINE_CREATE_INFO || !info->pRigidState)
        return PM_ERROR_INVALID_ARGUMENT;
    if (info->fluidMaterialCount || info->couplingCount || info->clothMaterialCount || info->ropeMaterialCount ||
        info->softBodyMaterialCount || info->smokeMaterialCount)
        return PM_ERROR_FEATURE_NOT_PRESENT;
    const PmRigidSolverState &r = *info->pRigidState;
    if (r.rigidIterations == 0u || r.rigidIterations > 64u) return PM_ERROR_INVALID_ARGUMENT;
    auto *p = new (std::nothrow) PmPhysicsPipeline_T();
    if (!p) return PM_ERROR_OUT_OF_HOST_MEMORY;
    p->device = d;
    p->rigid = r;
    *out = p;
    return PM_SUCCESS;
}

void DestroyPhysicsPipeline(PmDevice d, PmPhysicsPipeline p) {
    if (!d || !p || p->device != d) return;
    (void)queue_wait_idle(&d->queue);
    delete p;
}

// ---- command pools and buffers --------------------------------------------------

PmResult CreateCommandPool(PmDevice d, const PmCommandPoolCreateInfo *info, PmCommandPool *out) {
    if (!out) return PM_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!d || !info || info->sType != PM_STRUCTURE_TYPE_COMMAND_POOL_CREATE_INFO || info->queueFamilyIndex != 0u)
        return PM_ERROR_INVALID_ARGUMENT;
    auto *p = new (std::nothrow) PmCommandPool_T();
    if (!p) return PM_ERROR_OUT_OF_HOST_MEMORY;
    p->device = d;
    p->flags = info->flags;
    *out = p;
    return PM_SUCCESS;
}

void DestroyCommandPool(PmDevice d, PmCommandPool p) {
    if (!d || !p || p->device != d) return;
    (void)queue_wait_idle(&d->queue);
    delete p;
}

PmResult ResetCommandPool(PmDevice d, PmCommandPool p) {
    if (!d || !p || p->device != d) return PM_ERROR_INVALID_ARGUMENT;
    return PM_SUCCESS;
}

PmResult AllocateCommandBuffers(PmDevice d, PmCommandPool p, uint32_t count, PmCommandBuffer *out) {
    if (!d || !p || !out || p->device != d) return PM_ERROR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < count; ++k) {
        out[k] = new (std::nothrow) PmCommandBuffer_T();
        if (!out[k]) {
            for (uint32_t j = 0; j < k; ++j) delete out[j];
            return PM_ERROR_OUT_OF_HOST_MEMORY;
        }
        out[k]->pool = p;
    }
    return PM_SUCCESS;
}

void FreeCommandBuffers(PmDevice d, PmCommandPool p, uint32_t count, const PmCommandBuffer *bufs) {
    if (!d || !p || !bufs || p->device != d) return;
    for (uint32_t k = 0; k < count; ++k)
        if (bufs[k] && bufs[k]->pool == p) delete bufs[k];
}

PmResult BeginCommandBuffer(PmCommandBuffer cb, const PmCommandBufferBeginInfo *info) {
    if (!cb || !info || info->sType != PM_STRUCTURE_TYPE_COMMAND_BUFFER_BEGIN_INFO) return PM_ERROR_INVALID_ARGUMENT;
    if (cb->state == PmCommandBuffer_T::kRecording) return PM_ERROR_INVALID_STATE;
    cb->list = std::make_shared<CommandList>();   // a submitted list stays alive with its submission
    cb->bound = nullptr;
    cb->state = PmCommandBuffer_T::kRecording;
    return PM_SUCCESS;
}

PmResult EndCommandBuffer(PmCommandBuffer cb) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    cb->state = PmCommandBuffer_T::kExecutable;
    return PM_SUCCESS;
}

PmResult ResetCommandBuffer(PmCommandBuffer cb) {
    if (!cb) return PM_ERROR_INVALID_ARGUMENT;
    cb->list.reset();
    cb->bound = nullptr;
    cb->state = PmCommandBuffer_T::kInitial;
    return PM_SUCCESS;
}

PmResult CmdBindPhysicsPipeline(PmCommandBuffer cb, PmPhysicsPipeline p) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!p || p->device != cb->pool->device) return PM_ERROR_INVALID_ARGUMENT;
    cb->bound = p;
    return PM_SUCCESS;
}

// Commands run in submission order on one worker and each finishes before
// the next starts, so every barrier is already met.
PmResult CmdPipelineBarrier(PmCommandBuffer cb, uint32_t nm, const PmMemoryBarrier *m, uint32_t nb,
                            const PmBufferMemoryBarrier *b) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if ((nm && !m) || (nb && !b)) return PM_ERROR_INVALID_ARGUMENT;
    return PM_SUCCESS;
}

PmResult CmdCopyBuffer(PmCommandBuffer cb, PmBuffer src, PmBuffer dst, uint32_t count, const PmBufferCopy *regions) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!src || !dst || !count || !regions || src->device != cb->pool->device || dst->device != cb->pool->device)
        return PM_ERROR_INVALID_ARGUMENT;
    std::vector<PmBufferCopy> copied(regions, regions + count);
    for (const PmBufferCopy &c : copied)
        if (!c.size || !range_fits(c.srcOffset, c.size, src->size) || !range_fits(c.dstOffset, c.size, dst->size))
            return PM_ERROR_INVALID_ARGUMENT;
    return record(cb, nullptr, [src, dst, copied]() -> PmResult {
        if (!src->memory || !dst->memory) return PM_ERROR_INVALID_STATE;
        for (const PmBufferCopy &c : copied)
            if (hipMemcpy(buffer_pointer(dst, c.dstOffset), buffer_pointer(src, c.srcOffset), c.size,
                          hipMemcpyDefault) != hipSuccess)
                return PM_ERROR_DEVICE_LOST;
        return PM_SUCCESS;
    });
}

// ---- rigid commands -----------------------------------------------------------------

PmResult CmdWriteRigidBodies(PmCommandBuffer cb, PmSimulation s, uint32_t first, uint32_t count,
                             const PmRigidBodyInfo *bodies) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!sim_of(cb, s) || !count || !bodies || first > s->bodies.size() || count > s->bodies.size() - first)
        return PM_ERROR_INVALID_ARGUMENT;
    std::vector<PmRigidBodyInfo> copied(bodies, bodies + count);
    for (const PmRigidBodyInfo &b : copied) {
        if (b.active == PM_FALSE) continue;
        if (!valid_body(b) || b.geometry->device != cb->pool->device) return PM_ERROR_INVALID_ARGUMENT;
    }
    return record(cb, s, [s, first, copied]() -> PmResult {
        for (uint32_t k = 0; k < copied.size(); ++k) {
            BodySlot &slot = s->bodies[first + k];
            slot.active = copied[k].active != PM_FALSE;
            slot.fresh = slot.active;
            slot.info = copied[k];
        }
        s->rebuild = true;
        return PM_SUCCESS;
    });
}

PmResult CmdApplyRigidInputs(PmCommandBuffer cb, PmSimulation s, uint32_t count, const PmRigidInput *inputs) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!sim_of(cb, s) || !count || !inputs) return PM_ERROR_INVALID_ARGUMENT;
    std::vector<PmRigidInput> copied(inputs, inputs + count);
    for (const PmRigidInput &in : copied)
        if (in.bodyIndex >= s->bodies.size()) return PM_ERROR_INVALID_ARGUMENT;
    return record(cb, s, [s, copied]() -> PmResult {
        PmResult r = ensure_world(s);
        if (r != PM_SUCCESS) return r;
        std::vector<PmhRigidInput> mapped;
        for (const PmRigidInput &in : copied) {
            const uint32_t d = s->dense[in.bodyIndex];
            if (d == PM_INVALID_INDEX) return PM_ERROR_INVALID_STATE;
            PmhRigidInput m;
            static_assert(sizeof m == sizeof in, "PmRigidInput layout");
            std::memcpy(&m, &in, sizeof m);
            m.body_index = d;
            mapped.push_back(m);
        }
        return from_pmh(pmhWorldApplyRigidInputs(s->world, uint32_t(mapped.size()), mapped.data()));
    });
}

// A slot range mapped to dense bodies, applied run by run (consecutive
// dense indices).
template <class F>
PmResult for_dense_runs(PmSimulation_T *s, uint32_t first, uint32_t count, F &&apply) {
    for (uint32_t k = 0; k < count;) {
        const uint32_t d = s->dense[first + k];
        if (d == PM_INVALID_INDEX) return PM_ERROR_INVALID_STATE;
        uint32_t run = 1;
        while (k + run < count && s->dense[first + k + run] == d + run) ++run;
        PmResult r = apply(k, d, run);
        if (r != PM_SUCCESS) return r;
        k += run;
    }
    return PM_SUCCESS;
}

PmResult CmdSetRigidStates(PmCommandBuffer cb, PmSimulation s, uint32_t first, uint32_t count,
                           const PmRigidState *states) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!sim_of(cb, s) || !count || !states || first > s->bodies.size() || count > s->bodies.size() - first)
        return PM_ERROR_INVALID_ARGUMENT;
    std::vector<PmRigidState> copied(states, states + count);
    return record(cb, s, [s, first, count, copied]() -> PmResult {
        PmResult r = ensure_world(s);
        if (r != PM_SUCCESS) return r;
        return for_dense_runs(s, first, count, [&](uint32_t k, uint32_t d, uint32_t run) {
            std::vector<PmhRigidBodyState> w(run);
            for (uint32_t j = 0; j < run; ++j) w[j] = world_state(copied[k + j]);
            return from_pmh(pmhWorldSetRigidStates(s->world, d, run, w.data()));
        });
    });
}

PmResult CmdSetKinematicTargets(PmCommandBuffer cb, PmSimulation s, uint32_t first, uint32_t count,
                                const PmRigidState *targets) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!sim_of(cb, s) || !count || !targets || first > s->bodies.size() || count > s->bodies.size() - first)
        return PM_ERROR_INVALID_ARGUMENT;
    std::vector<PmRigidState> copied(targets, targets + count);
    for (const PmRigidState &t : copied)
        if (!finite_state(t)) return PM_ERROR_INVALID_ARGUMENT;
    return record(cb, s, [s, first, count, copied]() -> PmResult {
        PmResult r = ensure_world(s);
        if (r != PM_SUCCESS) return r;
        return for_dense_runs(s, first, count, [&](uint32_t k, uint32_t d, uint32_t run) {
            std::vector<PmhRigidBodyState> w(run);
            for (uint32_t j = 0; j < run; ++j) w[j] = world_state(copied[k + j]);
            return from_pmh(pmhWorldSetKinematicTargets(s->world, d, run, w.data()));
        });
    });
}

PmResult CmdWriteRigidConstraints(PmCommandBuffer cb, PmSimulation s, uint32_t first, uint32_t count,
                                  const PmRigidConstraintInfo *constraints) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!sim_of(cb, s) || !count || !constraints || first > s->joints.size() || count > s->joints.size() - first)
        return PM_ERROR_INVALID_ARGUMENT;
    std::vector<PmRigidConstraintInfo> copied(constraints, constraints + count);
    for (const PmRigidConstraintInfo &c : copied)
        if (c.active != PM_FALSE && (c.type > PM_CONSTRAINT_MOTOR || !std::isfinite(c.breakingImpulseThreshold) ||
                                     c.breakingImpulseThreshold < 0.0f || c.solverIterations == 0u ||
                                     c.solverIterations > 64u))
            return PM_ERROR_INVALID_ARGUMENT;
    return record(cb, s, [s, first, count, copied]() -> PmResult {
        for (uint32_t k = 0; k < count; ++k) {
            const PmRigidConstraintInfo &c = copied[k];
            if (c.active != PM_FALSE &&
                (c.bodyA >= s->bodies.size() || c.bodyB >= s->bodies.size() || !s->bodies[c.bodyA].active ||
                 !s->bodies[c.bodyB].active))
                return PM_ERROR_INVALID_ARGUMENT;
            s->joints[first + k].active = c.active != PM_FALSE;
            s->joints[first + k].info = c;
        }
        if (!s->world || s->rebuild) return PM_SUCCESS;   // written into the world when it is built
        std::vector<PmhRigidConstraintOptions> opts;
        std::vector<uint32_t> active;
        PmResult r = slot_joints(s, opts, active);
        if (r != PM_SUCCESS) return r;
        return from_pmh(pmhWorldWriteJoints(s->world, first, count, opts.data() + first, active.data() + first));
    });
}

PmResult CmdResetSimulation(PmCommandBuffer cb, PmSimulation s) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!sim_of(cb, s)) return PM_ERROR_INVALID_ARGUMENT;
    return record(cb, s, [s]() -> PmResult {
        if (s->world) pmhDestroyWorld(s->world);
        s->world = nullptr;
        s->rebuild = true;
        for (BodySlot &b : s->bodies) b = BodySlot{};
        for (JointSlot &j : s->joints) j = JointSlot{};
        s->events.clear();
        s->steps = 0;
        s->result.store(PM_SUCCESS);
        return PM_SUCCESS;
    });
}

// pmCmdStepSimulation (src/api/resources.cpp:1230-1288): realtime v1 only,
// the bound pipeline required; contact events are kept when asked for.
PmResult CmdStepSimulation(PmCommandBuffer cb, PmSimulation s, const PmStepInfo *step) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!sim_of(cb, s) || !step || !(step->timestep > 0.0f) || step->substeps == 0u || !cb->bound ||
        (step->flags & ~PmFlags(PM_STEP_COLLECT_CONTACT_EVENTS_BIT)) != 0u)
        return PM_ERROR_INVALID_ARGUMENT;
    const float dt = 1.0f / 60.0f;
    if (!(std::isfinite(step->timestep) && step->timestep >= std::nextafter(dt, 0.0f) &&
          step->timestep <= std::nextafter(dt, INFINITY)) || step->substeps != 1u)
        return PM_ERROR_INVALID_ARGUMENT;
    for (float g : step->gravity)
        if (!std::isfinite(g)) return PM_ERROR_INVALID_ARGUMENT;
    const PmStepInfo copied = *step;
    return record(cb, s, [s, copied]() -> PmResult {
        if (!s->bound) return PM_ERROR_INVALID_STATE;
        std::memcpy(s->gravity, copied.gravity, sizeof s->gravity);
        PmResult r = ensure_world(s);
        if (r == PM_SUCCESS && pmhWorldBodyCount(s->world) != 0u) {
            r = from_pmh(pmhWorldSetGravity(s->world, copied.gravity));
            if (r == PM_SUCCESS) r = from_pmh(pmhWorldStep(s->world, 1));
        }
        s->result.store(r);
        if (r != PM_SUCCESS) return r;
        ++s->steps;
        s->events.clear();
        if (copied.flags & PM_STEP_COLLECT_CONTACT_EVENTS_BIT) {
            uint32_t n = 0;
            r = from_pmh(pmhWorldGetContactEvents(s->world, nullptr, 0, &n));
            if (r != PM_SUCCESS) return r;
            std::vector<PmhContactEvent> e(n);
            if (n) r = from_pmh(pmhWorldGetContactEvents(s->world, e.data(), n, &n));
            if (r != PM_SUCCESS) return r;
            for (const PmhContactEvent &x : e) {
                PmContactEvent o;
                static_assert(sizeof o == sizeof x, "PmContactEvent layout");
                std::memcpy(&o, &x, sizeof o);
                o.type = PM_CONTACT_RIGID_RIGID;
                o.first = s->slot_of[x.first];
                o.second = s->slot_of[x.second];
                s->events.push_back(o);
            }
            s->event_revision = s->steps;
        }
        return PM_SUCCESS;
    });
}

PmResult CmdCopyRigidStates(PmCommandBuffer cb, PmSimulation s, uint32_t first, uint32_t count,
                            const PmBufferRange *dst) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!sim_of(cb, s) || first > s->bodies.size() || count > s->bodies.size() - first ||
        !destination_ok(cb, dst, PmDeviceSize(count) * sizeof(PmRigidState)))
        return PM_ERROR_INVALID_ARGUMENT;
    const PmBufferRange copied = *dst;
    return record(cb, s, [s, first, count, copied]() -> PmResult {
        PmResult r = ensure_world(s);
        if (r != PM_SUCCESS) return r;
        std::vector<PmhRigidBodyState> all(pmhWorldBodyCount(s->world));
        if (!all.empty()) {
            r = from_pmh(pmhWorldCopyStates(s->world, all.data()));
            if (r != PM_SUCCESS) return r;
        }
        std::vector<PmRigidState> out(count);
        for (uint32_t k = 0; k < count; ++k) {
            const uint32_t d = s->dense[first + k];
            if (d == PM_INVALID_INDEX || d >= all.size()) return PM_ERROR_INVALID_STATE;
            std::memcpy(&out[k], &all[d], sizeof(PmRigidState));
        }
        return write_range(copied, out.data(), out.size() * sizeof(PmRigidState));
    });
}

PmResult CmdCopyConstraintStates(PmCommandBuffer cb, PmSimulation s, uint32_t first, uint32_t count,
                                 const PmBufferRange *dst) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!sim_of(cb, s) || !count || first > s->joints.size() || count > s->joints.size() - first ||
        !destination_ok(cb, dst, PmDeviceSize(count) * sizeof(PmConstraintState)))
        return PM_ERROR_INVALID_ARGUMENT;
    const PmBufferRange copied = *dst;
    return record(cb, s, [s, first, count, copied]() -> PmResult {
        PmResult r = ensure_world(s);
        if (r != PM_SUCCESS) return r;
        std::vector<PmhRigidConstraint> j(pmhWorldJointCount(s->world));
        if (!j.empty()) {
            r = from_pmh(pmhWorldGetJoints(s->world, j.data()));
            if (r != PM_SUCCESS) return r;
        }
        std::vector<PmConstraintState> out(count);
        for (uint32_t k = 0; k < count; ++k) {
            const uint32_t slot = first + k;
            if (!s->joints[slot].active || slot >= j.size() || !j[slot].alive) {
                out[k] = {PM_FALSE, PM_FALSE, 0.0f};
                continue;
            }
            out[k] = {j[slot].state_enabled ? PM_TRUE : PM_FALSE, j[slot].state_broken ? PM_TRUE : PM_FALSE,
                      j[slot].applied_impulse};
        }
        return write_range(copied, out.data(), out.size() * sizeof(PmConstraintState));
    });
}

// PmDomainOutputHeader then the records (src/api/domains.cpp:314-339,
// 1692-1744); PM_ERROR_CAPACITY_EXCEEDED when they do not fit.
PmResult CmdCopyContactEvents(PmCommandBuffer cb, PmSimulation s, const PmBufferRange *dst) {
    PmResult r = recording(cb);
    if (r != PM_SUCCESS) return r;
    if (!sim_of(cb, s) || !destination_ok(cb, dst, sizeof(PmDomainOutputHeader))) return PM_ERROR_INVALID_ARGUMENT;
    const PmBufferRange copied = *dst;
    return record(cb, s, [s, copied]() -> PmResult {
        const size_t bytes = sizeof(PmDomainOutputHeader) + s->events.size() * sizeof(PmContactEvent);
        if (bytes > copied.size) return PM_ERROR_CAPACITY_EXCEEDED;
        std::vector<uint8_t> staging(bytes);
        const PmDomainOutputHeader h{uint32_t(s->events.size()), s->info.contactCapacity, s->event_revision};
        std::memcpy(staging.data(), &h, sizeof h);
        if (!s->events.empty())
            std::memcpy(staging.data() + sizeof h, s->events.data(), s->events.size() * sizeof(PmContactEvent));
        return write_range(copied, staging.data(), bytes);
    });
}

PmResult GetDeviceMemoryBudget(PmDevice d, PmMemoryBudget *b) {
    if (!d || !b) return PM_ERROR_INVALID_ARGUMENT;
    *b = {};
    size_t free_b = 0, total_b = 0;
    (void)hipSetDevice(d->physical->ordinal);
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return PM_ERROR_DEVICE_LOST;
    b->heapCount = 2;
    b->budget[0] = total_b;
    b->usage[0] = total_b - free_b;
    b->confidence[0] = PM_MEMORY_BUDGET_CONFIDENCE_DRIVER_REPORTED;
    b->budget[1] = PmDeviceSize(sysconf(_SC_PHYS_PAGES)) * PmDeviceSize(sysconf(_SC_PAGE_SIZE));
    b->usage[1] = b->budget[1] - PmDeviceSize(sysconf(_SC_AVPHYS_PAGES)) * PmDeviceSize(sysconf(_SC_PAGE_SIZE));
    b->confidence[1] = PM_MEMORY_BUDGET_CONFIDENCE_OS_REPORTED;
    b->applicationAllocationBytes = d->application_bytes.load();
    const uint64_t used = total_b - free_b;
    b->libraryDeviceBytes = used > b->applicationAllocationBytes ? used - b->applicationAllocationBytes : 0u;
    return PM_SUCCESS;
}

PFN_pmVoidFunction GetInstanceProcAddr(PmInstance instance, const char *name);
PFN_pmVoidFunction GetDeviceProcAddr(PmDevice device, const char *name);

struct Entry {
    const char *name;
    PFN_pmVoidFunction fn;
};
#define PMD_ENTRY(name, impl) {#name, reinterpret_cast<PFN_pmVoidFunction>(&impl)}
const Entry kEntries[] = {
    PMD_ENTRY(pmGetInstanceProcAddr, GetInstanceProcAddr),
    PMD_ENTRY(pmGetDeviceProcAddr, GetDeviceProcAddr),
    PMD_ENTRY(pmEnumerateInstanceVersion, EnumerateInstanceVersion),
    PMD_ENTRY(pmCreateInstance, CreateInstance),
    PMD_ENTRY(pmDestroyInstance, DestroyInstance),
    PMD_ENTRY(pmEnumeratePhysicalDevices, EnumeratePhysicalDevices),
    PMD_ENTRY(pmGetPhysicalDeviceProperties, GetPhysicalDeviceProperties),
    PMD_ENTRY(pmGetPhysicalDeviceQueueFamilies, GetPhysicalDeviceQueueFamilies),
    PMD_ENTRY(pmGetPhysicalDeviceFeatures, GetPhysicalDeviceFeatures),
    PMD_ENTRY(pmGetPhysicalDeviceLimits, GetPhysicalDeviceLimits),
    PMD_ENTRY(pmGetPhysicalDeviceMemoryProperties, GetPhysicalDeviceMemoryProperties),
    PMD_ENTRY(pmEnumeratePhysicalDeviceInteractionCapabilities, EnumeratePhysicalDeviceInteractionCapabilities),
    PMD_ENTRY(pmGetPhysicalDeviceInteractionCapability, GetPhysicalDeviceInteractionCapability),
    PMD_ENTRY(pmEnumerateDeviceExtensionProperties, EnumerateDeviceExtensionProperties),
    PMD_ENTRY(pmCreateDevice, CreateDevice),
    PMD_ENTRY(pmDestroyDevice, DestroyDevice),
    PMD_ENTRY(pmGetDeviceQueue, GetDeviceQueue),
    PMD_ENTRY(pmQueueSubmit, QueueSubmit),
    PMD_ENTRY(pmQueueWaitIdle, QueueWaitIdle),
    PMD_ENTRY(pmDeviceWaitIdle, DeviceWaitIdle),
    PMD_ENTRY(pmCreateFence, CreateFence),
    PMD_ENTRY(pmDestroyFence, DestroyFence),
    PMD_ENTRY(pmGetFenceStatus, GetFenceStatus),
    PMD_ENTRY(pmResetFences, ResetFences),
    PMD_ENTRY(pmWaitForFences, WaitForFences),
    PMD_ENTRY(pmCreateSemaphore, CreateSemaphore),
    PMD_ENTRY(pmDestroySemaphore, DestroySemaphore),
    PMD_ENTRY(pmGetSemaphoreCounterValue, GetSemaphoreCounterValue),
    PMD_ENTRY(pmWaitSemaphore, WaitSemaphore),
    PMD_ENTRY(pmAllocateMemory, AllocateMemory),
    PMD_ENTRY(pmFreeMemory, FreeMemory),
    PMD_ENTRY(pmMapMemory, MapMemory),
    PMD_ENTRY(pmUnmapMemory, UnmapMemory),
    PMD_ENTRY(pmFlushMappedMemoryRanges, CheckMappedRanges),
    PMD_ENTRY(pmInvalidateMappedMemoryRanges, CheckMappedRanges),
    PMD_ENTRY(pmCreateBuffer, CreateBuffer),
    PMD_ENTRY(pmDestroyBuffer, DestroyBuffer),
    PMD_ENTRY(pmGetBufferMemoryRequirements, GetBufferMemoryRequirements),
    PMD_ENTRY(pmBindBufferMemory, BindBufferMemory),
    PMD_ENTRY(pmCreateGeometry, CreateGeometry),
    PMD_ENTRY(pmDestroyGeometry, DestroyGeometry),
    PMD_ENTRY(pmGetGeometryMemoryRequirements, GetGeometryMemoryRequirements),
    PMD_ENTRY(pmBindGeometryMemory, BindGeometryMemory),
    PMD_ENTRY(pmCmdBuildGeometry, CmdBuildGeometry),
    PMD_ENTRY(pmCreateSimulation, CreateSimulation),
    PMD_ENTRY(pmDestroySimulation, DestroySimulation),
    PMD_ENTRY(pmGetSimulationMemoryRequirements, GetSimulationMemoryRequirements),
    PMD_ENTRY(pmBindSimulationMemory, BindSimulationMemory),
    PMD_ENTRY(pmGetSimulationStatus, GetSimulationStatus),
    PMD_ENTRY(pmCreatePhysicsPipeline, CreatePhysicsPipeline),
    PMD_ENTRY(pmDestroyPhysicsPipeline, DestroyPhysicsPipeline),
    PMD_ENTRY(pmCreateCommandPool, CreateCommandPool),
    PMD_ENTRY(pmDestroyCommandPool, DestroyCommandPool),
    PMD_ENTRY(pmResetCommandPool, ResetCommandPool),
    PMD_ENTRY(pmAllocateCommandBuffers, AllocateCommandBuffers),
    PMD_ENTRY(pmFreeCommandBuffers, FreeCommandBuffers),
    PMD_ENTRY(pmBeginCommandBuffer, BeginCommandBuffer),
    PMD_ENTRY(pmEndCommandBuffer, EndCommandBuffer),
    PMD_ENTRY(pmResetCommandBuffer, ResetCommandBuffer),
    PMD_ENTRY(pmCmdBindPhysicsPipeline, CmdBindPhysicsPipeline),
    PMD_ENTRY(pmCmdPipelineBarrier, CmdPipelineBarrier),
    PMD_ENTRY(pmCmdCopyBuffer, CmdCopyBuffer),
    PMD_ENTRY(pmCmdWriteRigidBodies, CmdWriteRigidBodies),
    PMD_ENTRY(pmCmdApplyRigidInputs, CmdApplyRigidInputs),
    PMD_ENTRY(pmCmdSetRigidStates, CmdSetRigidStates),
    PMD_ENTRY(pmCmdSetKinematicTargets, CmdSetKinematicTargets),
    PMD_ENTRY(pmCmdWriteRigidConstraints, CmdWriteRigidConstraints),
    PMD_ENTRY(pmCmdResetSimulation, CmdResetSimulation),
    PMD_ENTRY(pmCmdStepSimulation, CmdStepSimulation),
    PMD_ENTRY(pmCmdCopyRigidStates, CmdCopyRigidStates),
    PMD_ENTRY(pmCmdCopyConstraintStates, CmdCopyConstraintStates),
    PMD_ENTRY(pmCmdCopyContactEvents, CmdCopyContactEvents),
    PMD_ENTRY(pmGetDeviceMemoryBudget, GetDeviceMemoryBudget),
};
#undef PMD_ENTRY

PFN_pmVoidFunction lookup(const char *name) {
    if (!name) return nullptr;
    for (const Entry &e : kEntries)
        if (std::strcmp(e.name, name) == 0) return e.fn;
    return nullptr;
}

PFN_pmVoidFunction GetInstanceProcAddr(PmInstance instance, const char *name) {
    if (!instance && name && std::strcmp(name, "pmGetInstanceProcAddr") != 0 &&
        std::strcmp(name, "pmEnumerateInstanceVersion") != 0 && std::strcmp(name, "pmCreateInstance") != 0)
        return nullptr;
    return lookup(name);
}

PFN_pmVoidFunction GetDeviceProcAddr(PmDevice device, const char *name) { return device ? lookup(name) : nullptr; }

}  // namespace pmd

// ---- the four exports (src/driver/slang/entry.cpp:26-56) -----------------------

PMD_EXPORT PmResult pm_driverNegotiateLoaderInterfaceVersion(uint32_t *version) {
    if (!version || *version == 0u) return PM_ERROR_INVALID_ARGUMENT;
    *version = std::min(*version, PM_DRIVER_INTERFACE_VERSION);
    return PM_SUCCESS;
}

PMD_EXPORT PFN_pmVoidFunction pm_driverGetInstanceProcAddr(PmInstance instance, const char *name) {
    return pmd::GetInstanceProcAddr(instance, name);
}

PMD_EXPORT PFN_pmVoidFunction pm_driverGetDeviceProcAddr(PmDevice device, const char *name) {
    return pmd::GetDeviceProcAddr(device, name);
}

PMD_EXPORT PFN_pmVoidFunction pm_driverGetPhysicalDeviceProcAddr(PmInstance instance, const char *name) {
    return instance ? pmd::lookup(name) : nullptr;
}
