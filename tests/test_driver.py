# SPDX-License-Identifier: MIT
"""The installable HIP driver (parallel-mater_amd/lib/libpm_hip_driver.so)
behind its manifest, driven the way the reference's loader drives a driver
(src/loader/driver.cpp, src/driver/slang/entry.cpp).

CPU: the manifest's format, exactly four exported symbols, the interface
negotiation, name resolution (instance-level names before an instance,
no commands of absent domains), and the layouts of include/pm_api_rigid.h
against the reference's include/parallel_mater/parallel_mater.h (compiled
here when the reference checkout is present).

GPU: instance -> physical device (features, interaction capabilities,
limits) -> device -> geometry -> simulation -> pipeline -> command buffer
-> queue submit -> fence, stepping a box pyramid with a hinge through
pmCmdStepSimulation and reading states, contact events and constraint
states back, checked against the oracle world stepping the same scene.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import pm_api as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "parallel-mater_amd", "lib")
MANIFEST = os.path.join(LIB, "pm_hip_driver.json")
REFERENCE_HEADER = "/root/reference/include/parallel_mater/parallel_mater.h"
U32, P, I32 = ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int32


def driver():
    if not os.path.exists(MANIFEST):
        pytest.skip("the driver is not built")
    return A.Driver(MANIFEST)


def test_manifest_names_the_driver_and_it_exports_four_symbols():
    d = driver()
    assert d.manifest["PMD"]["api_version"] == "0.3.0" and d.manifest["PMD"]["library_arch"] == "64"
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIB, "libpm_hip_driver.so")],
                         capture_output=True, text=True, check=True).stdout
    names = sorted(line.split()[-1] for line in out.splitlines()
                   if line.split()[1] in ("T", "W") and not line.split()[-1].startswith("_"))
    assert names == sorted(["pm_driverNegotiateLoaderInterfaceVersion", "pm_driverGetInstanceProcAddr",
                            "pm_driverGetDeviceProcAddr", "pm_driverGetPhysicalDeviceProcAddr"])


def test_interface_negotiation():
    d = driver()
    assert d.negotiate(0)[0] == A.INVALID_ARGUMENT
    assert d.negotiate(1) == (A.SUCCESS, 1)
    assert d.negotiate(2) == (A.SUCCESS, 2)
    assert d.negotiate(7) == (A.SUCCESS, 2)


def test_name_resolution():
    d = driver()
    for name in ("pmCreateInstance", "pmEnumerateInstanceVersion", "pmGetInstanceProcAddr"):
        assert d.lib.pm_driverGetInstanceProcAddr(None, name.encode()), name
    for name in ("pmCreateDevice", "pmCmdStepSimulation", "pmQueueSubmit"):
        assert not d.lib.pm_driverGetInstanceProcAddr(None, name.encode()), name
    assert not d.lib.pm_driverGetDeviceProcAddr(None, b"pmCmdStepSimulation")
    assert not d.lib.pm_driverGetPhysicalDeviceProcAddr(None, b"pmCreateDevice")
    version = d.proc("pmEnumerateInstanceVersion", I32, ctypes.POINTER(U32))
    v = U32(0)
    assert version(ctypes.byref(v)) == A.SUCCESS and v.value == (0 << 22) | (3 << 12)
    create = d.proc("pmCreateInstance", I32, P, P, ctypes.POINTER(P))
    info = A.InstanceCreateInfo(sType=A.ST_INSTANCE, apiVersion=(3 << 12))
    inst = P()
    assert create(ctypes.byref(info), None, ctypes.byref(inst)) == A.SUCCESS and inst.value
    for name in ("pmCreateDevice", "pmCmdWriteRigidBodies", "pmCmdStepSimulation", "pmCmdCopyContactEvents",
                 "pmGetPhysicalDeviceFeatures", "pmEnumeratePhysicalDeviceInteractionCapabilities"):
        assert d.lib.pm_driverGetPhysicalDeviceProcAddr(inst, name.encode()), name
    # commands of the domains this driver does not carry resolve to nothing
    for name in ("pmCmdWriteCloths", "pmCmdSetFluidParticles", "pmCmdWriteSmoke", "pmCreateQueryPool"):
        assert not d.lib.pm_driverGetPhysicalDeviceProcAddr(inst, name.encode()), name
    d.proc("pmDestroyInstance", None, P, instance=inst)(inst)


def _offsets_program(header, define_driver):
    lines = ["#include <cstdio>", "#include <cstddef>"]
    if define_driver:
        lines.append("#define PARALLEL_MATER_DRIVER_IMPLEMENTATION 1")
    lines += [f'#include "{header}"', "int main() {"]
    for name, rec in A.RECORDS.items():
        lines.append(f'    std::printf("{name} %zu\\n", sizeof({name}));')
        for field, _ in rec._fields_:
            lines.append(f'    std::printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    lines.append("    return 0;\n}")
    return "\n".join(lines)


def _layouts(header, tmp_path, tag, define_driver=False):
    src = tmp_path / f"{tag}.cpp"
    exe = tmp_path / tag
    src.write_text(_offsets_program(header, define_driver))
    subprocess.run(["g++", "-std=c++17", "-o", str(exe), str(src)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return dict(line.rsplit(" ", 1) for line in out.splitlines())


def test_record_layouts_match_the_reference_header(tmp_path):
    ours = _layouts(os.path.join(ROOT, "include", "pm_api_rigid.h"), tmp_path, "ours")
    # the ctypes mirror agrees with the header the driver is compiled against
    for name, rec in A.RECORDS.items():
        assert int(ours[name]) == ctypes.sizeof(rec), name
        for field, _ in rec._fields_:
            assert int(ours[f"{name}.{field}"]) == getattr(rec, field).offset, f"{name}.{field}"
    if not os.path.exists(REFERENCE_HEADER):
        pytest.skip("the reference checkout is not present")
    ref = _layouts(REFERENCE_HEADER, tmp_path, "ref", define_driver=True)
    assert ref == ours


def _has_gpu():
    try:
        import pmamd
        pmamd.device_info()
        return True
    except Exception:
        return False


def test_instance_without_a_gpu_has_no_physical_device():
    if _has_gpu():
        pytest.skip("a GPU is present")
    d = driver()
    create = d.proc("pmCreateInstance", I32, P, P, ctypes.POINTER(P))
    info = A.InstanceCreateInfo(sType=A.ST_INSTANCE)
    inst = P()
    assert create(ctypes.byref(info), None, ctypes.byref(inst)) == A.SUCCESS
    enum = d.proc("pmEnumeratePhysicalDevices", I32, P, ctypes.POINTER(U32), P, instance=inst)
    n = U32(7)
    assert enum(inst, ctypes.byref(n), None) == A.SUCCESS and n.value == 0
    d.proc("pmDestroyInstance", None, P, instance=inst)(inst)


# ---------------------------------------------------------------------------
# GPU: the whole command path.

class Session:
    """One device of the driver with helpers for the command path."""

    def __init__(self):
        self.d = driver()
        f = self.d.proc
        self.instance = P()
        info = A.InstanceCreateInfo(sType=A.ST_INSTANCE, apiVersion=(3 << 12))
        assert f("pmCreateInstance", I32, P, P, ctypes.POINTER(P))(ctypes.byref(info), None,
                                                                    ctypes.byref(self.instance)) == A.SUCCESS
        self.i = self.instance

        def fn(name, restype, *args):
            g = f(name, restype, *args, instance=self.i)
            assert g is not None, name
            return g
        self.fn = fn
        n = U32(0)
        assert fn("pmEnumeratePhysicalDevices", I32, P, ctypes.POINTER(U32), P)(self.i, ctypes.byref(n), None) == 0
        assert n.value >= 1
        devs = (P * n.value)()
        assert fn("pmEnumeratePhysicalDevices", I32, P, ctypes.POINTER(U32), P)(self.i, ctypes.byref(n), devs) == 0
        self.physical = P(devs[0])

    def close(self):
        self.fn("pmDestroyInstance", None, P)(self.i)


def box_mesh(h):
    import scenarios as S
    return S.box_mesh(*h)


@pytest.mark.gpu
def test_physical_device_reports_the_rigid_domain_only():
    s = Session()
    feats = A.PhysicalDeviceFeatures()
    s.fn("pmGetPhysicalDeviceFeatures", None, P, P)(s.physical, ctypes.byref(feats))
    on = {n for n in A.FEATURE_NAMES if getattr(feats, n)}
    assert {"rigidBodies", "rigidConstraints", "contactEvents"} <= on
    assert not on & {"particleLiquids", "cloth", "rope", "softBodies", "smoke", "paint", "domainCouplings"}
    n = U32(0)
    caps_fn = s.fn("pmEnumeratePhysicalDeviceInteractionCapabilities", I32, P, ctypes.POINTER(U32), P)
    assert caps_fn(s.physical, ctypes.byref(n), None) == A.SUCCESS and n.value == 2
    caps = (A.InteractionCapability * 2)()
    assert caps_fn(s.physical, ctypes.byref(n), caps) == A.SUCCESS
    assert sorted(c.key.mode for c in caps) == [0, 1]                     # contact, attachment
    assert all(c.key.firstFamily == c.key.secondFamily == 5 for c in caps)   # rigid bodies
    one = U32(1)
    assert caps_fn(s.physical, ctypes.byref(one), caps) == A.INCOMPLETE
    get = s.fn("pmGetPhysicalDeviceInteractionCapability", I32, P, P, P)
    out = A.InteractionCapability()
    key = A.InteractionCapabilityKey(5, 5, 1, 0, 0, 0)
    assert get(s.physical, ctypes.byref(key), ctypes.byref(out)) == A.SUCCESS and out.key.mode == 0
    fluid = A.InteractionCapabilityKey(0, 5, 3, 0, 0, 0)                    # liquid-rigid contact
    assert get(s.physical, ctypes.byref(fluid), ctypes.byref(out)) == A.FEATURE_NOT_PRESENT
    bad = A.InteractionCapabilityKey(5, 0, 3, 0, 0, 0)                      # distinct families out of order
    assert get(s.physical, ctypes.byref(bad), ctypes.byref(out)) == A.INVALID_ARGUMENT
    # a device asking for an absent feature is refused
    dinfo = A.DeviceCreateInfo(sType=A.ST_DEVICE)
    dinfo.enabledFeatures.rigidBodies = 1
    dinfo.enabledFeatures.cloth = 1
    dev = P()
    assert s.fn("pmCreateDevice", I32, P, P, ctypes.POINTER(P))(s.physical, ctypes.byref(dinfo),
                                                                 ctypes.byref(dev)) == A.FEATURE_NOT_PRESENT
    s.close()


def _pyramid(rows):
    """Boxes of half 0.5 in rows on a static ground box (half 20, 0.5, 20 at
    y = -0.5), the reference's default options, slot 0 the ground."""
    pos = [(0.0, -0.5, 0.0)]
    for r in range(rows):
        for i in range(rows - r):
            pos.append((i - (rows - r - 1) / 2.0, 0.5 + r, 0.0))
    return np.array(pos, np.float32)


@pytest.mark.gpu
def test_step_through_the_driver_matches_the_oracle():
    import _oracle as O
    import scenarios as S
    from pmamd import layouts as L
    s = Session()
    fn = s.fn
    dinfo = A.DeviceCreateInfo(sType=A.ST_DEVICE)
    dinfo.enabledFeatures.rigidBodies = 1
    dinfo.enabledFeatures.rigidConstraints = 1
    dinfo.enabledFeatures.contactEvents = 1
    dev = P()
    assert fn("pmCreateDevice", I32, P, P, ctypes.POINTER(P))(s.physical, ctypes.byref(dinfo), ctypes.byref(dev)) == 0
    dfn = lambda name, restype, *args: s.d.proc(name, restype, *args, device=dev)  # noqa: E731
    queue = P()
    dfn("pmGetDeviceQueue", None, P, U32, U32, ctypes.POINTER(P))(dev, 0, 0, ctypes.byref(queue))
    assert queue.value

    def alloc(size, host=True):
        m = P()
        info = A.MemoryAllocateInfo(sType=A.ST_MEMORY, allocationSize=size, memoryTypeIndex=1 if host else 0)
        assert dfn("pmAllocateMemory", I32, P, P, ctypes.POINTER(P))(dev, ctypes.byref(info), ctypes.byref(m)) == 0
        return m

    def buffer(size, memory, offset=0):
        b = P()
        info = A.BufferCreateInfo(sType=A.ST_BUFFER, size=size)
        assert dfn("pmCreateBuffer", I32, P, P, ctypes.POINTER(P))(dev, ctypes.byref(info), ctypes.byref(b)) == 0
        assert dfn("pmBindBufferMemory", I32, P, P, P, ctypes.c_uint64)(dev, b, memory, offset) == 0
        return b

    def mapped(memory, size):
        ptr = P()
        assert dfn("pmMapMemory", I32, P, P, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(P))(
            dev, memory, 0, size, ctypes.byref(ptr)) == 0
        return ptr.value

    # Geometry: a unit box and the ground box, vertex and index data in
    # host-visible buffers.
    geoms = []
    for half in ((0.5, 0.5, 0.5), (20.0, 0.5, 20.0)):
        v, idx = S.box_mesh(*half)
        v = np.ascontiguousarray(v, np.float32)
        idx = np.ascontiguousarray(idx, np.uint32)
        mem = alloc(4096)
        host = mapped(mem, 4096)
        ctypes.memmove(host, v.ctypes.data, v.nbytes)
        ctypes.memmove(host + 1024, idx.ctypes.data, idx.nbytes)
        vb, ib = buffer(1024, mem, 0), buffer(1024, mem, 1024)
        g = P()
        ginfo = A.GeometryCreateInfo(sType=A.ST_GEOMETRY)
        ginfo.vertices = A.BufferView(A.BufferRange(vb.value, 0, v.nbytes), A.FORMAT_R32G32B32_SFLOAT, len(v), 12)
        ginfo.indices = A.BufferView(A.BufferRange(ib.value, 0, idx.nbytes), A.FORMAT_R32_UINT, len(idx), 4)
        assert dfn("pmCreateGeometry", I32, P, P, ctypes.POINTER(P))(dev, ctypes.byref(ginfo), ctypes.byref(g)) == 0
        gm = alloc(4096, host=False)
        assert dfn("pmBindGeometryMemory", I32, P, P, P, ctypes.c_uint64)(dev, g, gm, 0) == 0
        geoms.append(g)

    pos = _pyramid(5)
    n = len(pos)
    sinfo = A.SimulationCreateInfo(sType=A.ST_SIMULATION, rigidBodyCapacity=n + 2, constraintCapacity=2,
                                   contactCapacity=4096)
    sim = P()
    assert dfn("pmCreateSimulation", I32, P, P, ctypes.POINTER(P))(dev, ctypes.byref(sinfo), ctypes.byref(sim)) == 0
    req = A.SimulationMemoryRequirements()
    dfn("pmGetSimulationMemoryRequirements", None, P, P, P)(dev, sim, ctypes.byref(req))
    pm = alloc(max(req.persistent.size, 256), host=False)
    sm = alloc(max(req.scratch.size, 256), host=False)
    bind = A.SimulationMemoryBinding(pm.value, 0, sm.value, 0)
    assert dfn("pmBindSimulationMemory", I32, P, P, P)(dev, sim, ctypes.byref(bind)) == 0
    rigid = A.RigidSolverState(rigidIterations=10)
    pinfo = A.PhysicsPipelineCreateInfo(sType=A.ST_PIPELINE, pRigidState=ctypes.cast(ctypes.pointer(rigid), P).value)
    pipe = P()
    assert dfn("pmCreatePhysicsPipeline", I32, P, P, ctypes.POINTER(P))(dev, ctypes.byref(pinfo),
                                                                        ctypes.byref(pipe)) == 0

    # Bodies: slot 0 the ground, slots 1..n-1 the boxes, two empty slots;
    # mass 1000 kg, inertia from the mesh (zero diagonal), friction 0.5,
    # margin 5 mm, damping 0.05, speed limits 100.
    bodies = (A.RigidBodyInfo * (n + 2))()
    for k in range(n):
        b = bodies[k]
        b.active = 1
        b.motionType = 0 if k == 0 else 2
        b.geometry = geoms[1 if k == 0 else 0].value
        b.state.position[:] = pos[k].tolist()
        b.state.orientation[:] = [0.0, 0.0, 0.0, 1.0]
        b.mass = 1000.0
        b.friction, b.restitution = 0.5, 0.0
        b.linearDamping = b.angularDamping = 0.05
        b.maximumLinearSpeed = b.maximumAngularSpeed = 100.0
        b.collisionMargin = 0.005
    # a hinge between the two boxes of the fourth row (slots 13 and 14, side
    # by side): anchors on the faces they share; slot 1 stays empty
    joints = (A.RigidConstraintInfo * 2)()
    j = joints[0]
    j.active, j.type, j.bodyA, j.bodyB = 1, 2, 13, 14
    j.frameA.position[:] = [0.5, 0.0, 0.0]
    j.frameB.position[:] = [-0.5, 0.0, 0.0]
    j.frameA.orientation[:] = j.frameB.orientation[:] = [0.0, 0.0, 0.0, 1.0]
    j.collideConnected = 0
    j.breakingImpulseThreshold = 0.0
    j.solverIterations = 10

    pool = P()
    cinfo = A.CommandPoolCreateInfo(sType=A.ST_POOL)
    assert dfn("pmCreateCommandPool", I32, P, P, ctypes.POINTER(P))(dev, ctypes.byref(cinfo), ctypes.byref(pool)) == 0
    cb = P()
    assert dfn("pmAllocateCommandBuffers", I32, P, P, U32, ctypes.POINTER(P))(dev, pool, 1, ctypes.byref(cb)) == 0
    begin = A.CommandBufferBeginInfo(sType=A.ST_BEGIN)
    out_mem = alloc(1 << 16)
    out_host = mapped(out_mem, 1 << 16)
    states_buf = buffer(52 * n, out_mem, 0)
    events_buf = buffer(16 + 56 * 1024, out_mem, 4096)
    joints_buf = buffer(12 * 2, out_mem, 61440)
    steps = 20
    step = A.StepInfo(timestep=np.float32(1.0 / 60.0), substeps=1, flags=A.STEP_COLLECT_CONTACT_EVENTS)
    step.gravity[:] = [0.0, -9.81, 0.0]
    fence = P()
    assert dfn("pmCreateFence", I32, P, U32, ctypes.POINTER(P))(dev, 0, ctypes.byref(fence)) == 0

    assert dfn("pmBeginCommandBuffer", I32, P, P)(cb, ctypes.byref(begin)) == 0
    for g in geoms:
        assert dfn("pmCmdBuildGeometry", I32, P, P)(cb, g) == 0
    assert dfn("pmCmdWriteRigidBodies", I32, P, P, U32, U32, P)(cb, sim, 0, n + 2, bodies) == 0
    assert dfn("pmCmdWriteRigidConstraints", I32, P, P, U32, U32, P)(cb, sim, 0, 2, joints) == 0
    step_fn = dfn("pmCmdStepSimulation", I32, P, P, P)
    assert step_fn(cb, sim, ctypes.byref(step)) == A.INVALID_ARGUMENT   # no pipeline bound yet
    assert dfn("pmCmdBindPhysicsPipeline", I32, P, P)(cb, pipe) == 0
    bad = A.StepInfo(timestep=np.float32(1.0 / 120.0), substeps=1)
    assert step_fn(cb, sim, ctypes.byref(bad)) == A.INVALID_ARGUMENT    # realtime contract: 1/60, one substep
    for _ in range(steps):
        assert step_fn(cb, sim, ctypes.byref(step)) == 0
    rng = A.BufferRange(states_buf.value, 0, 52 * n)
    assert dfn("pmCmdCopyRigidStates", I32, P, P, U32, U32, P)(cb, sim, 0, n, ctypes.byref(rng)) == 0
    erng = A.BufferRange(events_buf.value, 0, 16 + 56 * 1024)
    assert dfn("pmCmdCopyContactEvents", I32, P, P, P)(cb, sim, ctypes.byref(erng)) == 0
    jrng = A.BufferRange(joints_buf.value, 0, 24)
    assert dfn("pmCmdCopyConstraintStates", I32, P, P, U32, U32, P)(cb, sim, 0, 2, ctypes.byref(jrng)) == 0
    assert dfn("pmEndCommandBuffer", I32, P)(cb) == 0
    cbs = (P * 1)(cb.value)
    submit = A.SubmitInfo(sType=A.ST_SUBMIT, commandBufferCount=1, pCommandBuffers=ctypes.cast(cbs, P).value)
    assert dfn("pmQueueSubmit", I32, P, P, P)(queue, ctypes.byref(submit), fence) == 0
    fences = (P * 1)(fence.value)
    assert dfn("pmWaitForFences", I32, P, U32, P, U32, ctypes.c_uint64)(dev, 1, fences, 1, A.TIMEOUT_INFINITE) == 0
    status = I32(99)
    assert dfn("pmGetSimulationStatus", I32, P, P, ctypes.POINTER(I32))(dev, sim, ctypes.byref(status)) == 0
    assert status.value == A.SUCCESS

    got = np.frombuffer(ctypes.string_at(out_host, 52 * n), dtype=L.STATE).copy()
    header = A.DomainOutputHeader.from_buffer_copy(ctypes.string_at(out_host + 4096, 16))
    events = (A.ContactEvent * header.activeCount).from_buffer_copy(
        ctypes.string_at(out_host + 4096 + 16, 56 * header.activeCount))
    jstate = (A.ConstraintState * 2).from_buffer_copy(ctypes.string_at(out_host + 61440, 24))

    # The oracle (realtime profile, the library's 32-lane summation order)
    # steps the same scene.
    o = O.OracleWorld(profile=1, lanes=32)
    box = o.add_mesh(*S.box_mesh(0.5, 0.5, 0.5))
    ground = o.add_mesh(*S.box_mesh(20.0, 0.5, 20.0))
    unit = o.shape(box)[10:13]
    motion = np.full(n, L.MOTION_DYNAMIC, np.uint32)
    motion[0] = L.MOTION_STATIC
    params = S.params(n, motion=motion, mass=1000.0, inertia=tuple(float(np.float32(u) * np.float32(1000.0))
                                                                    for u in unit),
                      friction=0.5, margin=0.005, shape=np.where(motion == 0, ground, box), damping=0.05,
                      max_linear=100.0, max_angular=100.0)
    o.add_bodies(params, S.states(n, pos))
    opts = L.joint_options(2, L.JOINT_HINGE, 13, 14, anchor_a=(0.5, 0.0, 0.0), anchor_b=(-0.5, 0.0, 0.0),
                           disable_collisions=1, iterations=10)
    o.add_joints(opts)
    o.write_joints(0, opts, np.array([1, 0], np.uint32))
    o.step(steps)
    want = o.states()
    for f in ("position", "orientation", "linear_velocity", "angular_velocity"):
        np.testing.assert_allclose(got[f], want[f], rtol=1e-5, atol=1e-5, err_msg=f)
    oe = o.contact_events()
    assert header.activeCount == len(oe) > 0 and header.revision == steps
    assert all(e.type == 0 for e in events)
    assert sorted((e.first, e.second) for e in events) == sorted((int(e["first"]), int(e["second"])) for e in oe)
    oj = o.joints()
    assert jstate[0].enabled == 1 and jstate[0].broken == 0
    assert abs(jstate[0].appliedImpulse - float(oj["applied_impulse"][0])) <= 1e-4 + 1e-4 * abs(
        float(oj["applied_impulse"][0]))
    assert (jstate[1].enabled, jstate[1].broken, jstate[1].appliedImpulse) == (0, 0, 0.0)

    dfn("pmDestroyFence", None, P, P)(dev, fence)
    dfn("pmDestroySimulation", None, P, P)(dev, sim)
    dfn("pmDestroyDevice", None, P)(dev)
    s.close()
