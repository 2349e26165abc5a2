# SPDX-License-Identifier: MIT
"""ctypes mirror of the pm* records the rigid HIP driver takes
(include/pm_api_rigid.h, the layouts of the reference's
include/parallel_mater/parallel_mater.h), and a minimal loader: reads a
driver manifest, opens the library it names, negotiates the interface
version and resolves commands by name, as the reference's loader does
(src/loader/driver.cpp:100-140)."""
import ctypes
import json
import os

U32, U64, F32, P = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_float, ctypes.c_void_p
I32 = ctypes.c_int32

SUCCESS, NOT_READY, TIMEOUT, INCOMPLETE = 0, 1, 2, 3
INVALID_ARGUMENT, INVALID_STATE, FEATURE_NOT_PRESENT, CAPACITY_EXCEEDED = -1, -2, -5, -7

ST_INSTANCE, ST_DEVICE, ST_MEMORY, ST_BUFFER, ST_GEOMETRY, ST_SIMULATION, ST_PIPELINE = 1, 2, 3, 4, 5, 6, 7
ST_POOL, ST_BEGIN, ST_SEMAPHORE, ST_SUBMIT = 8, 9, 10, 11
FORMAT_R32_UINT, FORMAT_R32G32B32_SFLOAT = 1, 3
STEP_COLLECT_CONTACT_EVENTS = 1
TIMEOUT_INFINITE = 2 ** 64 - 1


class S(ctypes.Structure):
    pass


def struct(name, fields):
    return type(name, (ctypes.Structure,), {"_fields_": fields})


InstanceCreateInfo = struct("InstanceCreateInfo", [
    ("sType", U32), ("pNext", P), ("apiVersion", U32), ("pApplicationName", ctypes.c_char_p),
    ("enableValidation", U32), ("enabledLayerCount", U32), ("ppEnabledLayerNames", P),
    ("enabledExtensionCount", U32), ("ppEnabledExtensionNames", P)])
PhysicalDeviceProperties = struct("PhysicalDeviceProperties", [
    ("apiVersion", U32), ("driverVersion", U32), ("vendorId", U32), ("deviceId", U32),
    ("deviceUuid", ctypes.c_uint8 * 16), ("deviceName", ctypes.c_char * 256)])
QueueFamilyProperties = struct("QueueFamilyProperties", [
    ("queueFlags", U32), ("queueCount", U32), ("timestampValidBits", U32), ("timestampPeriodNanoseconds", F32)])
FEATURE_NAMES = ["rigidBodies", "rigidConstraints", "particleLiquids", "liquidRigidCoupling", "liquidLiquidCoupling",
                 "cloth", "rope", "softBodies", "smoke", "domainCouplings", "particleEmitters", "paint",
                 "contactEvents", "debugData", "timelineSemaphores", "timestampQueries", "statisticsQueries",
                 "memoryBudget", "vulkanInterop"]
PhysicalDeviceFeatures = struct("PhysicalDeviceFeatures", [(n, U32) for n in FEATURE_NAMES])
DeviceQueueCreateInfo = struct("DeviceQueueCreateInfo", [("queueFamilyIndex", U32), ("queueCount", U32),
                                                         ("pPriorities", P)])
DeviceCreateInfo = struct("DeviceCreateInfo", [
    ("sType", U32), ("pNext", P), ("queueCreateInfoCount", U32), ("pQueueInfos", P),
    ("enabledFeatures", PhysicalDeviceFeatures), ("enabledExtensionCount", U32), ("ppEnabledExtensionNames", P)])
MemoryAllocateInfo = struct("MemoryAllocateInfo", [("sType", U32), ("pNext", P), ("allocationSize", U64),
                                                   ("memoryTypeIndex", U32)])
BufferCreateInfo = struct("BufferCreateInfo", [("sType", U32), ("pNext", P), ("size", U64), ("usage", U32),
                                               ("initialQueueFamilyIndex", U32)])
BufferRange = struct("BufferRange", [("buffer", P), ("offset", U64), ("size", U64)])
BufferView = struct("BufferView", [("range", BufferRange), ("format", U32), ("count", U32), ("stride", U32)])
GeometryCreateInfo = struct("GeometryCreateInfo", [("sType", U32), ("pNext", P), ("vertices", BufferView),
                                                   ("indices", BufferView), ("queueFamilyIndex", U32)])
SimulationCreateInfo = struct("SimulationCreateInfo", [
    ("sType", U32), ("pNext", P), ("queueFamilyIndex", U32), ("rigidBodyCapacity", U32),
    ("constraintCapacity", U32), ("contactCapacity", U32), ("fluidCount", U32), ("pFluidCapacities", P),
    ("clothCount", U32), ("pClothCapacities", P), ("ropeCount", U32), ("pRopeCapacities", P),
    ("softBodyCount", U32), ("pSoftBodyCapacities", P), ("smokeCount", U32), ("pSmokeCapacities", P),
    ("couplingCapacity", U32), ("emitterCapacity", U32), ("destroyerCapacity", U32), ("paintFieldCount", U32),
    ("pPaintFieldCapacities", P), ("paintRuleCapacity", U32), ("debugFrameCapacity", U32)])
MemoryRequirements = struct("MemoryRequirements", [("size", U64), ("alignment", U64), ("memoryTypeBits", U32),
                                                   ("requiresDedicatedAllocation", U32)])
SimulationMemoryRequirements = struct("SimulationMemoryRequirements", [("persistent", MemoryRequirements),
                                                                       ("scratch", MemoryRequirements)])
SimulationMemoryBinding = struct("SimulationMemoryBinding", [("persistentMemory", P), ("persistentOffset", U64),
                                                             ("scratchMemory", P), ("scratchOffset", U64)])
RigidState = struct("RigidState", [("position", F32 * 3), ("orientation", F32 * 4), ("linearVelocity", F32 * 3),
                                   ("angularVelocity", F32 * 3)])
RigidBodyInfo = struct("RigidBodyInfo", [
    ("active", U32), ("motionType", U32), ("geometry", P), ("state", RigidState), ("mass", F32),
    ("inertiaDiagonal", F32 * 3), ("friction", F32), ("restitution", F32), ("linearDamping", F32),
    ("angularDamping", F32), ("maximumLinearSpeed", F32), ("maximumAngularSpeed", F32), ("collisionMargin", F32),
    ("userData", U64)])
ConstraintFrame = struct("ConstraintFrame", [("position", F32 * 3), ("orientation", F32 * 4)])
AxisLimit = struct("AxisLimit", [("enabled", U32), ("lower", F32), ("upper", F32)])
AxisSpring = struct("AxisSpring", [("enabled", U32), ("stiffness", F32), ("damping", F32), ("equilibrium", F32)])
AxisMotor = struct("AxisMotor", [("enabled", U32), ("targetVelocity", F32), ("maximumForce", F32)])
RigidConstraintInfo = struct("RigidConstraintInfo", [
    ("active", U32), ("type", U32), ("bodyA", U32), ("bodyB", U32), ("frameA", ConstraintFrame),
    ("frameB", ConstraintFrame), ("linearLimits", AxisLimit * 3), ("angularLimits", AxisLimit * 3),
    ("linearSprings", AxisSpring * 3), ("angularSprings", AxisSpring * 3), ("linearMotors", AxisMotor * 3),
    ("angularMotors", AxisMotor * 3), ("collideConnected", U32), ("breakingImpulseThreshold", F32),
    ("solverIterations", U32)])
RigidSolverState = struct("RigidSolverState", [("rigidIterations", U32), ("enableSleeping", U32),
                                               ("sleepLinearThreshold", F32), ("sleepAngularThreshold", F32),
                                               ("sleepTimeSeconds", F32)])
PhysicsPipelineCreateInfo = struct("PhysicsPipelineCreateInfo", [
    ("sType", U32), ("pNext", P), ("pRigidState", P), ("fluidMaterialCount", U32), ("pFluidMaterials", P),
    ("couplingCount", U32), ("pCouplings", P), ("clothMaterialCount", U32), ("pClothMaterials", P),
    ("ropeMaterialCount", U32), ("pRopeMaterials", P), ("softBodyMaterialCount", U32), ("pSoftBodyMaterials", P),
    ("smokeMaterialCount", U32), ("pSmokeMaterials", P)])
Float3 = struct("Float3", [("x", F32), ("y", F32), ("z", F32)])
DomainOutputHeader = struct("DomainOutputHeader", [("activeCount", U32), ("capacity", U32), ("revision", U64)])
ConstraintState = struct("ConstraintState", [("enabled", U32), ("broken", U32), ("appliedImpulse", F32)])
ContactEvent = struct("ContactEvent", [("type", U32), ("first", U32), ("second", U32), ("position", Float3),
                                       ("normal", Float3), ("penetration", F32), ("normalImpulse", F32),
                                       ("frictionImpulse", Float3)])
CommandPoolCreateInfo = struct("CommandPoolCreateInfo", [("sType", U32), ("pNext", P), ("flags", U32),
                                                         ("queueFamilyIndex", U32)])
CommandBufferBeginInfo = struct("CommandBufferBeginInfo", [("sType", U32), ("pNext", P), ("flags", U32)])
StepInfo = struct("StepInfo", [("timestep", F32), ("substeps", U32), ("gravity", F32 * 3), ("flags", U32)])
SemaphoreSubmitInfo = struct("SemaphoreSubmitInfo", [("semaphore", P), ("value", U64), ("stages", U64)])
SubmitInfo = struct("SubmitInfo", [("sType", U32), ("pNext", P), ("waitCount", U32), ("pWaits", P),
                                   ("commandBufferCount", U32), ("pCommandBuffers", P), ("signalCount", U32),
                                   ("pSignals", P)])
InteractionCapabilityKey = struct("InteractionCapabilityKey", [
    ("firstFamily", U32), ("secondFamily", U32), ("scope", U32), ("mode", U32), ("direction", U32),
    ("materialProfile", U32)])
InteractionCapability = struct("InteractionCapability", [("key", InteractionCapabilityKey),
                                                         ("maxGraphEdgesPerSimulation", U32)])
PhysicalDeviceLimits = struct("PhysicalDeviceLimits", [
    ("maxAllocationSize", U64), ("maxBufferSize", U64), ("nonCoherentAtomSize", U64), ("maxMemoryAllocations", U32),
    ("maxRigidBodies", U32), ("maxRigidConstraints", U32), ("maxContacts", U32), ("maxFluidSystems", U32),
    ("maxParticlesPerFluid", U32), ("maxNeighborsPerParticle", U32), ("maxRigidIterations", U32),
    ("maxLiquidIterations", U32), ("maxSubsteps", U32), ("maxQueriesPerPool", U32),
    ("maxTimelineValueDifference", U64)])
MemoryHeap = struct("MemoryHeap", [("size", U64), ("flags", U32)])
MemoryType = struct("MemoryType", [("propertyFlags", U32), ("heapIndex", U32)])
MemoryProperties = struct("MemoryProperties", [("heapCount", U32), ("heaps", MemoryHeap * 16), ("typeCount", U32),
                                               ("types", MemoryType * 32)])
MemoryBudget = struct("MemoryBudget", [
    ("heapCount", U32), ("budget", U64 * 16), ("usage", U64 * 16), ("confidence", U32 * 16),
    ("applicationAllocationBytes", U64), ("libraryDeviceBytes", U64), ("libraryHostBytes", U64),
    ("borrowedBufferBytes", U64)])

# Records whose sizes and offsets the ABI test checks against the reference header.
RECORDS = {"PmInstanceCreateInfo": InstanceCreateInfo, "PmPhysicalDeviceProperties": PhysicalDeviceProperties,
           "PmQueueFamilyProperties": QueueFamilyProperties, "PmPhysicalDeviceFeatures": PhysicalDeviceFeatures,
           "PmDeviceQueueCreateInfo": DeviceQueueCreateInfo, "PmDeviceCreateInfo": DeviceCreateInfo,
           "PmMemoryAllocateInfo": MemoryAllocateInfo, "PmBufferCreateInfo": BufferCreateInfo,
           "PmBufferRange": BufferRange, "PmBufferView": BufferView, "PmGeometryCreateInfo": GeometryCreateInfo,
           "PmSimulationCreateInfo": SimulationCreateInfo, "PmMemoryRequirements": MemoryRequirements,
           "PmSimulationMemoryRequirements": SimulationMemoryRequirements,
           "PmSimulationMemoryBinding": SimulationMemoryBinding, "PmRigidState": RigidState,
           "PmRigidBodyInfo": RigidBodyInfo, "PmConstraintFrame": ConstraintFrame, "PmAxisLimit": AxisLimit,
           "PmAxisSpring": AxisSpring, "PmAxisMotor": AxisMotor, "PmRigidConstraintInfo": RigidConstraintInfo,
           "PmRigidSolverState": RigidSolverState, "PmPhysicsPipelineCreateInfo": PhysicsPipelineCreateInfo,
           "PmDomainOutputHeader": DomainOutputHeader, "PmConstraintState": ConstraintState,
           "PmContactEvent": ContactEvent, "PmCommandPoolCreateInfo": CommandPoolCreateInfo,
           "PmCommandBufferBeginInfo": CommandBufferBeginInfo, "PmStepInfo": StepInfo,
           "PmSemaphoreSubmitInfo": SemaphoreSubmitInfo, "PmSubmitInfo": SubmitInfo,
           "PmInteractionCapabilityKey": InteractionCapabilityKey, "PmInteractionCapability": InteractionCapability,
           "PmPhysicalDeviceLimits": PhysicalDeviceLimits, "PmMemoryProperties": MemoryProperties,
           "PmMemoryBudget": MemoryBudget}

VOIDFN = ctypes.CFUNCTYPE(None)


class Driver:
    """A driver opened through its manifest, with its commands resolved by
    name (instance-level first, then through the instance)."""

    def __init__(self, manifest):
        with open(manifest) as f:
            m = json.load(f)
        assert m["file_format_version"] == "1.0.0"
        path = m["PMD"]["library_path"]
        if not os.path.isabs(path):
            path = os.path.join(os.path.dirname(os.path.abspath(manifest)), path)
        self.manifest = m
        self.lib = ctypes.CDLL(path)
        self.lib.pm_driverNegotiateLoaderInterfaceVersion.restype = I32
        self.lib.pm_driverNegotiateLoaderInterfaceVersion.argtypes = [ctypes.POINTER(U32)]
        for n in ("pm_driverGetInstanceProcAddr", "pm_driverGetDeviceProcAddr", "pm_driverGetPhysicalDeviceProcAddr"):
            getattr(self.lib, n).restype = P
            getattr(self.lib, n).argtypes = [P, ctypes.c_char_p]

    def negotiate(self, version):
        v = U32(version)
        r = self.lib.pm_driverNegotiateLoaderInterfaceVersion(ctypes.byref(v))
        return r, v.value

    def proc(self, name, restype, *argtypes, instance=None, device=None):
        if device is not None:
            addr = self.lib.pm_driverGetDeviceProcAddr(device, name.encode())
        elif instance is not None:
            addr = self.lib.pm_driverGetPhysicalDeviceProcAddr(instance, name.encode())
        else:
            addr = self.lib.pm_driverGetInstanceProcAddr(None, name.encode())
        if not addr:
            return None
        return ctypes.CFUNCTYPE(restype, *argtypes)(addr)
