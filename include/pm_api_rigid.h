/* SPDX-License-Identifier: MIT
 *
 * The part of the ParallelMater 0.3 command API that the rigid HIP driver
 * (libpm_hip_driver.so) implements: handles, results, structures and
 * function types of include/parallel_mater/parallel_mater.h and the driver
 * entry points of include/parallel_mater/pm_driver.h.  The layouts must be
 * those of the reference header, byte for byte (the loader passes the
 * application's structures through), so every record below keeps the
 * reference's member order and types; tests/test_driver.py compares sizes
 * and offsets with the reference header when its checkout is present.
 *
 * Structures of the other domains (fluids, cloth, ropes, soft bodies,
 * smoke, paint, couplings) are not declared: the driver reports those
 * features absent and the loader returns PM_ERROR_FEATURE_NOT_PRESENT for
 * their commands.  Pointer members to such records keep their place as
 * opaque pointers.
 */
#ifndef PM_API_RIGID_H
#define PM_API_RIGID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PM_MAKE_API_VERSION(major, minor, patch) \
    ((((uint32_t)(major)) << 22) | (((uint32_t)(minor)) << 12) | ((uint32_t)(patch)))
#define PM_API_VERSION_0_3 PM_MAKE_API_VERSION(0, 3, 0)
#define PM_WHOLE_SIZE UINT64_MAX
#define PM_TIMEOUT_INFINITE UINT64_MAX
#define PM_INVALID_INDEX UINT32_MAX
#define PM_FALSE 0u
#define PM_TRUE 1u
#define PM_DRIVER_INTERFACE_VERSION 2u

typedef uint32_t PmBool32;
typedef uint32_t PmFlags;
typedef uint64_t PmDeviceSize;
typedef uint64_t PmStageFlags;
typedef uint64_t PmAccessFlags;
typedef struct PmInstance_T *PmInstance;
typedef struct PmPhysicalDevice_T *PmPhysicalDevice;
typedef struct PmDevice_T *PmDevice;
typedef struct PmQueue_T *PmQueue;
typedef struct PmDeviceMemory_T *PmDeviceMemory;
typedef struct PmBuffer_T *PmBuffer;
typedef struct PmGeometry_T *PmGeometry;
typedef struct PmSimulation_T *PmSimulation;
typedef struct PmPhysicsPipeline_T *PmPhysicsPipeline;
typedef struct PmCommandPool_T *PmCommandPool;
typedef struct PmCommandBuffer_T *PmCommandBuffer;
typedef struct PmSemaphore_T *PmSemaphore;
typedef struct PmFence_T *PmFence;

typedef enum PmStructureType {
    PM_STRUCTURE_TYPE_UNDEFINED = 0,
    PM_STRUCTURE_TYPE_INSTANCE_CREATE_INFO = 1,
    PM_STRUCTURE_TYPE_DEVICE_CREATE_INFO = 2,
    PM_STRUCTURE_TYPE_MEMORY_ALLOCATE_INFO = 3,
    PM_STRUCTURE_TYPE_BUFFER_CREATE_INFO = 4,
    PM_STRUCTURE_TYPE_GEOMETRY_CREATE_INFO = 5,
    PM_STRUCTURE_TYPE_SIMULATION_CREATE_INFO = 6,
    PM_STRUCTURE_TYPE_PHYSICS_PIPELINE_CREATE_INFO = 7,
    PM_STRUCTURE_TYPE_COMMAND_POOL_CREATE_INFO = 8,
    PM_STRUCTURE_TYPE_COMMAND_BUFFER_BEGIN_INFO = 9,
    PM_STRUCTURE_TYPE_SEMAPHORE_CREATE_INFO = 10,
    PM_STRUCTURE_TYPE_SUBMIT_INFO = 11
} PmStructureType;

typedef enum PmResult {
    PM_SUCCESS = 0,
    PM_NOT_READY = 1,
    PM_TIMEOUT = 2,
    PM_INCOMPLETE = 3,
    PM_ERROR_INVALID_ARGUMENT = -1,
    PM_ERROR_INVALID_STATE = -2,
    PM_ERROR_OUT_OF_HOST_MEMORY = -3,
    PM_ERROR_OUT_OF_DEVICE_MEMORY = -4,
    PM_ERROR_FEATURE_NOT_PRESENT = -5,
    PM_ERROR_FORMAT_NOT_SUPPORTED = -6,
    PM_ERROR_CAPACITY_EXCEEDED = -7,
    PM_ERROR_DEVICE_LOST = -8,
    PM_ERROR_INCOMPATIBLE_VERSION = -9,
    PM_ERROR_DRIVER_NOT_FOUND = -10,
    PM_ERROR_LAYER_NOT_PRESENT = -11,
    PM_ERROR_EXTENSION_NOT_PRESENT = -12
} PmResult;

typedef void (*PFN_pmVoidFunction)(void);

typedef struct PmExtensionProperties {
    char extensionName[256];
    uint32_t specificationVersion;
} PmExtensionProperties;

typedef struct PmInstanceCreateInfo {
    PmStructureType sType;
    const void *pNext;
    uint32_t apiVersion;
    const char *pApplicationName;
    PmBool32 enableValidation;
    uint32_t enabledLayerCount;
    const char *const *ppEnabledLayerNames;
    uint32_t enabledExtensionCount;
    const char *const *ppEnabledExtensionNames;
} PmInstanceCreateInfo;

typedef void *(*PmAllocateFn)(void *user, size_t size, size_t alignment);
typedef void (*PmFreeFn)(void *user, void *allocation);
typedef struct PmAllocationCallbacks {
    void *pUserData;
    PmAllocateFn pfnAllocate;
    PmFreeFn pfnFree;
} PmAllocationCallbacks;

typedef struct PmPhysicalDeviceProperties {
    uint32_t apiVersion;
    uint32_t driverVersion;
    uint32_t vendorId;
    uint32_t deviceId;
    uint8_t deviceUuid[16];
    char deviceName[256];
} PmPhysicalDeviceProperties;

enum { PM_QUEUE_PHYSICS_BIT = 0x1, PM_QUEUE_TRANSFER_BIT = 0x2 };
typedef struct PmQueueFamilyProperties {
    PmFlags queueFlags;
    uint32_t queueCount;
    uint32_t timestampValidBits;
    float timestampPeriodNanoseconds;
} PmQueueFamilyProperties;

typedef struct PmPhysicalDeviceFeatures {
    PmBool32 rigidBodies;
    PmBool32 rigidConstraints;
    PmBool32 particleLiquids;
    PmBool32 liquidRigidCoupling;
    PmBool32 liquidLiquidCoupling;
    PmBool32 cloth;
    PmBool32 rope;
    PmBool32 softBodies;
    PmBool32 smoke;
    PmBool32 domainCouplings;
    PmBool32 particleEmitters;
    PmBool32 paint;
    PmBool32 contactEvents;
    PmBool32 debugData;
    PmBool32 timelineSemaphores;
    PmBool32 timestampQueries;
    PmBool32 statisticsQueries;
    PmBool32 memoryBudget;
    PmBool32 vulkanInterop;
} PmPhysicalDeviceFeatures;

typedef struct PmDeviceQueueCreateInfo {
    uint32_t queueFamilyIndex;
    uint32_t queueCount;
    const float *pPriorities;
} PmDeviceQueueCreateInfo;
typedef struct PmDeviceCreateInfo {
    PmStructureType sType;
    const void *pNext;
    uint32_t queueCreateInfoCount;
    const PmDeviceQueueCreateInfo *pQueueInfos;
    PmPhysicalDeviceFeatures enabledFeatures;
    uint32_t enabledExtensionCount;
    const char *const *ppEnabledExtensionNames;
} PmDeviceCreateInfo;

enum {
    PM_MEMORY_DEVICE_LOCAL_BIT = 0x1,
    PM_MEMORY_HOST_VISIBLE_BIT = 0x2,
    PM_MEMORY_HOST_COHERENT_BIT = 0x4,
    PM_MEMORY_HOST_CACHED_BIT = 0x8
};
typedef struct PmMemoryHeap { PmDeviceSize size; PmFlags flags; } PmMemoryHeap;
typedef struct PmMemoryType { PmFlags propertyFlags; uint32_t heapIndex; } PmMemoryType;
typedef struct PmMemoryProperties {
    uint32_t heapCount;
    PmMemoryHeap heaps[16];
    uint32_t typeCount;
    PmMemoryType types[32];
} PmMemoryProperties;
typedef struct PmMemoryRequirements {
    PmDeviceSize size;
    PmDeviceSize alignment;
    uint32_t memoryTypeBits;
    PmBool32 requiresDedicatedAllocation;
} PmMemoryRequirements;
typedef struct PmMemoryAllocateInfo {
    PmStructureType sType;
    const void *pNext;
    PmDeviceSize allocationSize;
    uint32_t memoryTypeIndex;
} PmMemoryAllocateInfo;
typedef struct PmMappedMemoryRange {
    PmDeviceMemory memory;
    PmDeviceSize offset;
    PmDeviceSize size;
} PmMappedMemoryRange;

typedef struct PmBufferCreateInfo {
    PmStructureType sType;
    const void *pNext;
    PmDeviceSize size;
    PmFlags usage;
    uint32_t initialQueueFamilyIndex;
} PmBufferCreateInfo;
typedef struct PmBufferRange {
    PmBuffer buffer;
    PmDeviceSize offset;
    PmDeviceSize size;
} PmBufferRange;
typedef enum PmFormat {
    PM_FORMAT_UNDEFINED = 0,
    PM_FORMAT_R32_UINT = 1,
    PM_FORMAT_R32_SFLOAT = 2,
    PM_FORMAT_R32G32B32_SFLOAT = 3,
    PM_FORMAT_R32G32B32A32_SFLOAT = 4,
    PM_FORMAT_PARTICLE_STATE = 5,
    PM_FORMAT_RIGID_STATE = 6,
    PM_FORMAT_FLUID_PARTICLE_OUTPUT = 7
} PmFormat;
typedef struct PmBufferView {
    PmBufferRange range;
    PmFormat format;
    uint32_t count;
    uint32_t stride;
} PmBufferView;
typedef struct PmGeometryCreateInfo {
    PmStructureType sType;
    const void *pNext;
    PmBufferView vertices;
    PmBufferView indices;
    uint32_t queueFamilyIndex;
} PmGeometryCreateInfo;

typedef struct PmSimulationCreateInfo {
    PmStructureType sType;
    const void *pNext;
    uint32_t queueFamilyIndex;
    uint32_t rigidBodyCapacity;
    uint32_t constraintCapacity;
    uint32_t contactCapacity;
    uint32_t fluidCount;
    const void *pFluidCapacities;
    uint32_t clothCount;
    const void *pClothCapacities;
    uint32_t ropeCount;
    const void *pRopeCapacities;
    uint32_t softBodyCount;
    const void *pSoftBodyCapacities;
    uint32_t smokeCount;
    const void *pSmokeCapacities;
    uint32_t couplingCapacity;
    uint32_t emitterCapacity;
    uint32_t destroyerCapacity;
    uint32_t paintFieldCount;
    const void *pPaintFieldCapacities;
    uint32_t paintRuleCapacity;
    uint32_t debugFrameCapacity;
} PmSimulationCreateInfo;
typedef struct PmSimulationMemoryRequirements {
    PmMemoryRequirements persistent;
    PmMemoryRequirements scratch;
} PmSimulationMemoryRequirements;
typedef struct PmSimulationMemoryBinding {
    PmDeviceMemory persistentMemory;
    PmDeviceSize persistentOffset;
    PmDeviceMemory scratchMemory;
    PmDeviceSize scratchOffset;
} PmSimulationMemoryBinding;

typedef struct PmRigidState {
    float position[3];
    float orientation[4];
    float linearVelocity[3];
    float angularVelocity[3];
} PmRigidState;
typedef enum PmMotionType { PM_MOTION_STATIC = 0, PM_MOTION_KINEMATIC = 1, PM_MOTION_DYNAMIC = 2 } PmMotionType;
typedef struct PmRigidBodyInfo {
    PmBool32 active;
    PmMotionType motionType;
    PmGeometry geometry;
    PmRigidState state;
    float mass;
    float inertiaDiagonal[3];
    float friction;
    float restitution;
    float linearDamping;
    float angularDamping;
    float maximumLinearSpeed;
    float maximumAngularSpeed;
    float collisionMargin;
    uint64_t userData;
} PmRigidBodyInfo;
typedef struct PmRigidInput {
    uint32_t bodyIndex;
    float force[3];
    float torque[3];
    float linearImpulse[3];
    float angularImpulse[3];
} PmRigidInput;

typedef enum PmConstraintType {
    PM_CONSTRAINT_FIXED = 0,
    PM_CONSTRAINT_POINT = 1,
    PM_CONSTRAINT_HINGE = 2,
    PM_CONSTRAINT_SLIDER = 3,
    PM_CONSTRAINT_PISTON = 4,
    PM_CONSTRAINT_GENERIC = 5,
    PM_CONSTRAINT_GENERIC_SPRING = 6,
    PM_CONSTRAINT_MOTOR = 7
} PmConstraintType;
typedef struct PmConstraintFrame { float position[3]; float orientation[4]; } PmConstraintFrame;
typedef struct PmAxisLimit { PmBool32 enabled; float lower; float upper; } PmAxisLimit;
typedef struct PmAxisSpring { PmBool32 enabled; float stiffness; float damping; float equilibrium; } PmAxisSpring;
typedef struct PmAxisMotor { PmBool32 enabled; float targetVelocity; float maximumForce; } PmAxisMotor;
typedef struct PmRigidConstraintInfo {
    PmBool32 active;
    PmConstraintType type;
    uint32_t bodyA;
    uint32_t bodyB;
    PmConstraintFrame frameA;
    PmConstraintFrame frameB;
    PmAxisLimit linearLimits[3];
    PmAxisLimit angularLimits[3];
    PmAxisSpring linearSprings[3];
    PmAxisSpring angularSprings[3];
    PmAxisMotor linearMotors[3];
    PmAxisMotor angularMotors[3];
    PmBool32 collideConnected;
    float breakingImpulseThreshold;
    uint32_t solverIterations;
} PmRigidConstraintInfo;
typedef struct PmRigidSolverState {
    uint32_t rigidIterations;
    PmBool32 enableSleeping;
    float sleepLinearThreshold;
    float sleepAngularThreshold;
    float sleepTimeSeconds;
} PmRigidSolverState;

typedef struct PmPhysicsPipelineCreateInfo {
    PmStructureType sType;
    const void *pNext;
    const PmRigidSolverState *pRigidState;
    uint32_t fluidMaterialCount;
    const void *pFluidMaterials;
    uint32_t couplingCount;
    const void *pCouplings;
    uint32_t clothMaterialCount;
    const void *pClothMaterials;
    uint32_t ropeMaterialCount;
    const void *pRopeMaterials;
    uint32_t softBodyMaterialCount;
    const void *pSoftBodyMaterials;
    uint32_t smokeMaterialCount;
    const void *pSmokeMaterials;
} PmPhysicsPipelineCreateInfo;

typedef struct PmFloat3 { float x; float y; float z; } PmFloat3;
typedef struct PmDomainOutputHeader {
    uint32_t activeCount;
    uint32_t capacity;
    uint64_t revision;
} PmDomainOutputHeader;
typedef struct PmConstraintState {
    PmBool32 enabled;
    PmBool32 broken;
    float appliedImpulse;
} PmConstraintState;
typedef enum PmContactEventType { PM_CONTACT_RIGID_RIGID = 0, PM_CONTACT_LIQUID_RIGID = 1 } PmContactEventType;
typedef struct PmContactEvent {
    PmContactEventType type;
    uint32_t first;
    uint32_t second;
    PmFloat3 position;
    PmFloat3 normal;
    float penetration;
    float normalImpulse;
    PmFloat3 frictionImpulse;
} PmContactEvent;

enum { PM_COMMAND_POOL_RESET_BUFFER_BIT = 0x1, PM_COMMAND_BUFFER_ONE_TIME_BIT = 0x1 };
typedef struct PmCommandPoolCreateInfo {
    PmStructureType sType;
    const void *pNext;
    PmFlags flags;
    uint32_t queueFamilyIndex;
} PmCommandPoolCreateInfo;
typedef struct PmCommandBufferBeginInfo {
    PmStructureType sType;
    const void *pNext;
    PmFlags flags;
} PmCommandBufferBeginInfo;

enum { PM_STEP_COLLECT_CONTACT_EVENTS_BIT = 0x1 };
typedef struct PmStepInfo {
    float timestep;
    uint32_t substeps;
    float gravity[3];
    PmFlags flags;
} PmStepInfo;
typedef struct PmBufferCopy {
    PmDeviceSize srcOffset;
    PmDeviceSize dstOffset;
    PmDeviceSize size;
} PmBufferCopy;
typedef struct PmMemoryBarrier {
    PmStageFlags srcStages;
    PmAccessFlags srcAccess;
    PmStageFlags dstStages;
    PmAccessFlags dstAccess;
} PmMemoryBarrier;
typedef struct PmBufferMemoryBarrier {
    PmMemoryBarrier memory;
    PmBufferRange range;
    uint32_t srcQueueFamilyIndex;
    uint32_t dstQueueFamilyIndex;
} PmBufferMemoryBarrier;
typedef struct PmSemaphoreCreateInfo {
    PmStructureType sType;
    const void *pNext;
    uint64_t initialValue;
} PmSemaphoreCreateInfo;
typedef struct PmSemaphoreSubmitInfo {
    PmSemaphore semaphore;
    uint64_t value;
    PmStageFlags stages;
} PmSemaphoreSubmitInfo;
typedef struct PmSubmitInfo {
    PmStructureType sType;
    const void *pNext;
    uint32_t waitCount;
    const PmSemaphoreSubmitInfo *pWaits;
    uint32_t commandBufferCount;
    const PmCommandBuffer *pCommandBuffers;
    uint32_t signalCount;
    const PmSemaphoreSubmitInfo *pSignals;
} PmSubmitInfo;
enum { PM_FENCE_SIGNALED_BIT = 0x1 };

typedef enum PmMemoryBudgetConfidence {
    PM_MEMORY_BUDGET_CONFIDENCE_UNKNOWN = 0,
    PM_MEMORY_BUDGET_CONFIDENCE_IMPLEMENTATION_ESTIMATE = 1,
    PM_MEMORY_BUDGET_CONFIDENCE_DRIVER_REPORTED = 2,
    PM_MEMORY_BUDGET_CONFIDENCE_OS_REPORTED = 3
} PmMemoryBudgetConfidence;
typedef struct PmMemoryBudget {
    uint32_t heapCount;
    PmDeviceSize budget[16];
    PmDeviceSize usage[16];
    PmMemoryBudgetConfidence confidence[16];
    PmDeviceSize applicationAllocationBytes;
    PmDeviceSize libraryDeviceBytes;
    PmDeviceSize libraryHostBytes;
    PmDeviceSize borrowedBufferBytes;
} PmMemoryBudget;

typedef enum PmInteractionFamily {
    PM_INTERACTION_FAMILY_FLUID = 0,
    PM_INTERACTION_FAMILY_CLOTH = 1,
    PM_INTERACTION_FAMILY_SOFT_BODY = 2,
    PM_INTERACTION_FAMILY_ROPE = 3,
    PM_INTERACTION_FAMILY_SMOKE = 4,
    PM_INTERACTION_FAMILY_RIGID_BODY = 5
} PmInteractionFamily;
typedef enum PmInteractionScope {
    PM_INTERACTION_SCOPE_SELF = 0,
    PM_INTERACTION_SCOPE_SAME_DOMAIN = 1,
    PM_INTERACTION_SCOPE_SEPARATE_SYSTEM = 2,
    PM_INTERACTION_SCOPE_DISTINCT_FAMILY = 3
} PmInteractionScope;
typedef enum PmInteractionMode {
    PM_INTERACTION_MODE_CONTACT = 0,
    PM_INTERACTION_MODE_ATTACHMENT = 1,
    PM_INTERACTION_MODE_DRAG = 2,
    PM_INTERACTION_MODE_CONTAINMENT = 3,
    PM_INTERACTION_MODE_PRESSURE = 4,
    PM_INTERACTION_MODE_VISCOSITY = 5,
    PM_INTERACTION_MODE_THERMAL = 6,
    PM_INTERACTION_MODE_PHASE = 7,
    PM_INTERACTION_MODE_INTERFACE = 8
} PmInteractionMode;
typedef enum PmInteractionDirection {
    PM_INTERACTION_DIRECTION_BIDIRECTIONAL = 0,
    PM_INTERACTION_DIRECTION_FIRST_TO_SECOND = 1,
    PM_INTERACTION_DIRECTION_SECOND_TO_FIRST = 2,
    PM_INTERACTION_DIRECTION_FIRST_TO_SECOND_TRANSFER = 3,
    PM_INTERACTION_DIRECTION_SECOND_TO_FIRST_TRANSFER = 4,
    PM_INTERACTION_DIRECTION_EXTERNAL_TO_FIRST = 5,
    PM_INTERACTION_DIRECTION_EXTERNAL_TO_SECOND = 6
} PmInteractionDirection;
typedef enum PmInteractionMaterialProfile {
    PM_INTERACTION_MATERIAL_ENDPOINT_NATIVE = 0,
    PM_INTERACTION_MATERIAL_SMOKE_PARTICLE = 1,
    PM_INTERACTION_MATERIAL_SMOKE_GRID = 2,
    PM_INTERACTION_MATERIAL_EQUAL = 3,
    PM_INTERACTION_MATERIAL_MISCIBLE = 4,
    PM_INTERACTION_MATERIAL_IMMISCIBLE_INTERFACE = 5,
    PM_INTERACTION_MATERIAL_THERMAL_PHASE = 6,
    PM_INTERACTION_MATERIAL_SMOKE_PARTICLE_GRID = 7,
    PM_INTERACTION_MATERIAL_ROPE_CAPSULE = 8
} PmInteractionMaterialProfile;
#define PM_INTERACTION_GRAPH_EDGES_NOT_APPLICABLE 0u
typedef struct PmInteractionCapabilityKey {
    PmInteractionFamily firstFamily;
    PmInteractionFamily secondFamily;
    PmInteractionScope scope;
    PmInteractionMode mode;
    PmInteractionDirection direction;
    PmInteractionMaterialProfile materialProfile;
} PmInteractionCapabilityKey;
typedef struct PmInteractionCapability {
    PmInteractionCapabilityKey key;
    uint32_t maxGraphEdgesPerSimulation;
} PmInteractionCapability;
typedef struct PmPhysicalDeviceLimits {
    PmDeviceSize maxAllocationSize;
    PmDeviceSize maxBufferSize;
    PmDeviceSize nonCoherentAtomSize;
    uint32_t maxMemoryAllocations;
    uint32_t maxRigidBodies;
    uint32_t maxRigidConstraints;
    uint32_t maxContacts;
    uint32_t maxFluidSystems;
    uint32_t maxParticlesPerFluid;
    uint32_t maxNeighborsPerParticle;
    uint32_t maxRigidIterations;
    uint32_t maxLiquidIterations;
    uint32_t maxSubsteps;
    uint32_t maxQueriesPerPool;
    uint64_t maxTimelineValueDifference;
} PmPhysicalDeviceLimits;

/* Driver entry points (include/parallel_mater/pm_driver.h:34-41,
 * src/driver/slang/entry.cpp:18-56): the only symbols libpm_hip_driver.so
 * exports.  The loader negotiates the interface version (at most 2, never
 * 0), then resolves every command by name. */
PmResult pm_driverNegotiateLoaderInterfaceVersion(uint32_t *pVersion);
PFN_pmVoidFunction pm_driverGetInstanceProcAddr(PmInstance instance, const char *pName);
PFN_pmVoidFunction pm_driverGetDeviceProcAddr(PmDevice device, const char *pName);
PFN_pmVoidFunction pm_driverGetPhysicalDeviceProcAddr(PmInstance instance, const char *pName);

#ifdef __cplusplus
}
#endif
#endif
