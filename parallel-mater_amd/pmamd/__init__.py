# SPDX-License-Identifier: MIT
"""Host-side mirror of the ParallelMater rigid step over the MI355X HIP
driver (libpm_hip_rigid.so, C-ABI in include/pm_hip_rigid.h).

`World` follows the reference's rigid command path: add meshes
(pmCmdBuildGeometry), write bodies (pmCmdWriteRigidBodies), step
(pmCmdStepSimulation), copy states (pmCmdCopyRigidStates).
`prepare_solve` drives the kernel seam that stands where
prepare_geometry_constraints/solve_geometry_constraints stand.
There is no CPU fallback: every entry point raises when the HIP library or
a GPU is missing.
"""
import ctypes
import os

import numpy as np

from . import layouts  # noqa: F401
from .layouts import (AVBD_BODY, CONTACT_EVENT, JOINT, JOINT_OPTIONS, MANIFOLD, PARAMS, RIGID_INPUT, SCHEDULE,
                      STATE)

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(HERE, "lib", "libpm_hip_rigid.so")

PMH_SUCCESS = 0
RESULT_NAMES = {0: "PM_SUCCESS", -1: "PM_ERROR_INVALID_ARGUMENT", -2: "PM_ERROR_INVALID_STATE",
                -3: "PM_ERROR_OUT_OF_HOST_MEMORY", -4: "PM_ERROR_OUT_OF_DEVICE_MEMORY",
                -5: "PM_ERROR_FEATURE_NOT_PRESENT", -7: "PM_ERROR_CAPACITY_EXCEEDED",
                -8: "PM_ERROR_DEVICE_LOST"}
COOPERATIVE_SOLVE = 1
PHASE_TIMING = 2
PER_COLOR_SWEEP = 4
PHASES = ("integrate", "broadphase", "narrowphase", "contact_cache", "prepare", "solve", "finish")

_P = ctypes.c_void_p
_U = ctypes.c_uint32
_F = ctypes.c_float


class PmhError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {RESULT_NAMES.get(code, code)}")
        self.code = code


class WorldDesc(ctypes.Structure):
    _fields_ = [("gravity", _F * 3), ("timestep", _F), ("profile_version", _U),
                ("base_iterations", _U), ("requested_iterations", _U),
                ("body_capacity", _U), ("pairs_per_body", _U), ("points_per_body", _U),
                ("device", _U), ("flags", _U), ("solve_lanes", _U),
                ("kernel_timing_stride", _U), ("residual_tolerance", _F), ("substeps", _U)]


STATS_FIELDS = ["pairs", "manifolds", "points", "colors", "islands", "impact", "budget",
                "unsupported_pairs", "cache_hits", "phases", "status", "pair_capacity",
                "point_capacity", "points_allocated", "triangle_pairs", "fallback_pairs"]


class WorldStats(ctypes.Structure):
    _fields_ = [(n, _U) for n in STATS_FIELDS]


class KernelTiming(ctypes.Structure):
    _fields_ = [("milliseconds", ctypes.c_double), ("launches", ctypes.c_uint64),
                ("algorithmic_bytes", ctypes.c_double)]


class WorldTotals(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_uint64), ("body_iterations", ctypes.c_uint64),
                ("velocity_phases", ctypes.c_uint64), ("algorithmic_bytes", ctypes.c_double)]


_lib = None


def library_path():
    return LIB_PATH


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python parallel-mater_amd/build.py`")
    L = ctypes.CDLL(LIB_PATH)
    i32 = ctypes.c_int32
    sig = {
        "pmhCreateWorld": (i32, [_P, _P]),
        "pmhDestroyWorld": (None, [_P]),
        "pmhWorldAddMesh": (i32, [_P, _P, _U, _P, _U, _P]),
        "pmhWorldAddSphere": (i32, [_P, _F, _P]),
        "pmhWorldAddCapsule": (i32, [_P, _F, _F, _P]),
        "pmhWorldGetShapeInfo": (i32, [_P, _U, _P]),
        "pmhWorldAddBodies": (i32, [_P, _U, _P, _P]),
        "pmhWorldBodyCount": (_U, [_P]),
        "pmhWorldAddScene": (i32, [_P, _U, _P, _P, _P]),
        "pmhWorldSceneCount": (_U, [_P]),
        "pmhWorldSetSolveLanes": (i32, [_P, _U]),
        "pmhWorldStep": (i32, [_P, _U]),
        "pmhWorldSetColorWaitLimit": (i32, [_P, _U]),
        "pmhWorldSetSweepWaitLimit": (i32, [_P, _U]),
        "pmhWorldApplyRigidInputs": (i32, [_P, _U, _P]),
        "pmhWorldSetKinematicTargets": (i32, [_P, _U, _U, _P]),
        "pmhWorldAddJoints": (i32, [_P, _U, _P]),
        "pmhWorldJointCount": (_U, [_P]),
        "pmhWorldGetJoints": (i32, [_P, _P]),
        "pmhWorldWriteJoints": (i32, [_P, _U, _U, _P, _P]),
        "pmhWorldSetRigidStates": (i32, [_P, _U, _U, _P]),
        "pmhWorldGetContactEvents": (i32, [_P, _P, _U, _P]),
        "pmhWorldCopyStates": (i32, [_P, _P]),
        "pmhWorldSetStates": (i32, [_P, _P]),
        "pmhWorldGetStats": (i32, [_P, _P]),
        "pmhWorldGetPhaseTimes": (i32, [_P, _P]),
        "pmhWorldSolveTiming": (i32, [_P, _P, _P, ctypes.c_int]),
        "pmhWorldGetTotals": (i32, [_P, _P, ctypes.c_int]),
        "pmhWorldKernelTiming": (i32, [_P, _P, ctypes.c_int]),
        "pmhWorldGetResiduals": (i32, [_P, _P, _P, _P]),
        "pmhWorldPairCount": (_U, [_P]),
        "pmhWorldGetPairs": (i32, [_P, _P]),
        "pmhWorldGetManifolds": (i32, [_P, _P]),
        "pmhWorldGetAvbdBodies": (i32, [_P, _P]),
        "pmhWorldGetSchedule": (i32, [_P, _P]),
        "pmhWorldGetProxies": (i32, [_P, _P]),
        "pmhPrepareSolveGeometryConstraints": (
            i32, [_P, _P, _P, _P, _U, _P, _P, _P, _P, _P, _F, _P, _U, _U, _U]),
        "pmhPrepareSolveGeometryConstraintsJoints": (
            i32, [_P, _P, _P, _P, _U, _P, _P, _P, _P, _U, _P, _P, _P, _F, _P, _U, _U, _U]),
        "pmhPrepareSolveGeometryConstraintsNoAlloc": (
            i32, [_P, _P, _P, _P, _U, _P, _P, _P, _P, _P, _F, _P, _U, _U, _U]),
        "pmhWorldSetWarmstart": (i32, [_P, _P, _U]),
        "pmhWorldSetContactLog": (i32, [_P, _U]),
        "pmhWorldGetContactLog": (i32, [_P, _P, _U, _P]),
        "pmhWorldWarmstartStats": (i32, [_P, _P, _P]),
        "pmhWarmstartForward": (i32, [_P, _P, _U, _P]),
        "pmhWarmstartGradients": (i32, [_P, _U, _P, _P, _P]),
        "pmhWarmstartTrain": (i32, [_P, _U, _U, _F, _U, _P, _P]),
        "pmhDeviceAlloc": (i32, [ctypes.c_uint64, _P]),
        "pmhDeviceFree": (i32, [_P]),
        "pmhCopyToDevice": (i32, [_P, _P, ctypes.c_uint64]),
        "pmhCopyToHost": (i32, [_P, _P, ctypes.c_uint64]),
        "pmhDeviceSynchronize": (i32, []),
        "pmhDeviceInfo": (i32, [_P, _U, _P]),
    }
    L.pmhWarmstartInit.restype = None
    L.pmhWarmstartInit.argtypes = [_U, _P]
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _check(code, what):
    if code != PMH_SUCCESS:
        raise PmhError(code, what)


def _ptr(a):
    return a.ctypes.data_as(_P)


def device_info():
    name = ctypes.create_string_buffer(64)
    cus = _U()
    _check(lib().pmhDeviceInfo(name, 64, ctypes.byref(cus)), "pmhDeviceInfo")
    return name.value.decode(), cus.value


class World:
    """One rigid simulation on one GPU (realtime v1 or legacy profile)."""

    def __init__(self, dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), profile=1,
                 base_iterations=10, requested_iterations=0, body_capacity=0,
                 pairs_per_body=0, points_per_body=0, device=0, cooperative=False,
                 solve_lanes=0, kernel_timing_stride=0, residual_tolerance=0.0, substeps=1, timing=True,
                 per_color_sweep=False):
        flags = ((COOPERATIVE_SOLVE if cooperative else 0) | (PHASE_TIMING if timing else 0) |
                 (PER_COLOR_SWEEP if per_color_sweep else 0))
        desc = WorldDesc((_F * 3)(*gravity), dt, profile, base_iterations,
                         requested_iterations, body_capacity, pairs_per_body,
                         points_per_body, device, flags,
                         solve_lanes, kernel_timing_stride, residual_tolerance, substeps)
        h = _P()
        _check(lib().pmhCreateWorld(ctypes.byref(desc), ctypes.byref(h)), "pmhCreateWorld")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().pmhDestroyWorld(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add_mesh(self, vertices, indices):
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        i = np.ascontiguousarray(indices, dtype=np.uint32)
        out = _U()
        _check(lib().pmhWorldAddMesh(self._h, _ptr(v), len(v), _ptr(i), len(i),
                                     ctypes.byref(out)), "pmhWorldAddMesh")
        return out.value

    def add_sphere(self, radius):
        out = _U()
        _check(lib().pmhWorldAddSphere(self._h, radius, ctypes.byref(out)), "pmhWorldAddSphere")
        return out.value

    def add_capsule(self, radius, half_length):
        out = _U()
        _check(lib().pmhWorldAddCapsule(self._h, radius, half_length, ctypes.byref(out)),
               "pmhWorldAddCapsule")
        return out.value

    def shape(self, index):
        out = np.zeros(16, dtype=np.float32)
        _check(lib().pmhWorldGetShapeInfo(self._h, index, _ptr(out)), "pmhWorldGetShapeInfo")
        return out

    def add_bodies(self, params, states):
        p = np.ascontiguousarray(params, dtype=PARAMS)
        s = np.ascontiguousarray(states, dtype=STATE)
        _check(lib().pmhWorldAddBodies(self._h, len(p), _ptr(p), _ptr(s)), "pmhWorldAddBodies")

    def add_scene(self, params, states):
        """Appends an independent scene (its bodies never pair with other
        scenes'); returns its index."""
        p = np.ascontiguousarray(params, dtype=PARAMS)
        st = np.ascontiguousarray(states, dtype=STATE)
        out = _U()
        _check(lib().pmhWorldAddScene(self._h, len(p), _ptr(p), _ptr(st), ctypes.byref(out)), "pmhWorldAddScene")
        return out.value

    @property
    def scene_count(self):
        return lib().pmhWorldSceneCount(self._h)

    def set_solve_lanes(self, lanes):
        _check(lib().pmhWorldSetSolveLanes(self._h, lanes), "pmhWorldSetSolveLanes")

    def set_warmstart(self, params, penalty=True, lam=False):
        """Learned warm start for cold contacts (None: the plain one):
        the predicted penalty and/or the predicted lambda (the latter only
        where a contact starts in violation)."""
        if params is None:
            _check(lib().pmhWorldSetWarmstart(self._h, None, 0), "pmhWorldSetWarmstart")
            return
        self._warm = np.ascontiguousarray(params, dtype=np.float32)
        policy = (WARMSTART_PENALTY if penalty else 0) | (WARMSTART_LAMBDA if lam else 0)
        _check(lib().pmhWorldSetWarmstart(self._h, _ptr(self._warm), policy), "pmhWorldSetWarmstart")

    def set_contact_log(self, capacity):
        _check(lib().pmhWorldSetContactLog(self._h, capacity), "pmhWorldSetContactLog")

    def contact_log(self, capacity=None):
        """Logged cold-contact samples: 16 features,
        log10(1 + lambda / (m_eff |g|)), log10(k / k0).

        The device appends rows through atomic slots, so their order varies
        from run to run; the rows come back sorted by a 64-bit FNV-1a hash
        of their bits, an order that depends only on the rows, so that a
        seeded sample of the log is reproducible."""
        count = _U()
        _check(lib().pmhWorldGetContactLog(self._h, None, 0, ctypes.byref(count)), "pmhWorldGetContactLog")
        n = count.value if capacity is None else min(count.value, capacity)
        out = np.zeros((n, 18), dtype=np.float32)
        if n:
            _check(lib().pmhWorldGetContactLog(self._h, _ptr(out), n, ctypes.byref(count)), "pmhWorldGetContactLog")
            bits = out.view(np.uint32)
            h = np.full(n, 0xCBF29CE484222325, dtype=np.uint64)
            for k in range(out.shape[1]):
                h = (h ^ bits[:, k].astype(np.uint64)) * np.uint64(0x100000001B3)
            out = out[np.argsort(h, kind="stable")]
        return out

    def warmstart_stats(self):
        steps, falls = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().pmhWorldWarmstartStats(self._h, ctypes.byref(steps), ctypes.byref(falls)),
               "pmhWorldWarmstartStats")
        return steps.value, falls.value

    @property
    def body_count(self):
        return lib().pmhWorldBodyCount(self._h)

    def apply_inputs(self, inputs):
        """Forces, torques and impulses (layouts.RIGID_INPUT records) for
        the next step."""
        a = np.ascontiguousarray(inputs, dtype=RIGID_INPUT)
        _check(lib().pmhWorldApplyRigidInputs(self._h, len(a), _ptr(a)), "pmhWorldApplyRigidInputs")

    def set_kinematic_targets(self, first, targets):
        a = np.ascontiguousarray(targets, dtype=STATE)
        _check(lib().pmhWorldSetKinematicTargets(self._h, first, len(a), _ptr(a)),
               "pmhWorldSetKinematicTargets")

    def add_joints(self, options):
        """Appends joints (layouts.JOINT_OPTIONS records)."""
        a = np.ascontiguousarray(options, dtype=JOINT_OPTIONS)
        _check(lib().pmhWorldAddJoints(self._h, len(a), _ptr(a)), "pmhWorldAddJoints")

    @property
    def joint_count(self):
        return lib().pmhWorldJointCount(self._h)

    def joints(self):
        """The joint records (layouts.JOINT) after the last step."""
        out = np.zeros(self.joint_count, dtype=JOINT)
        if len(out):
            _check(lib().pmhWorldGetJoints(self._h, _ptr(out)), "pmhWorldGetJoints")
        return out

    def write_joints(self, first, options, active):
        """Removes joint slots first.. and re-adds those with active != 0
        from `options` (pmCmdWriteRigidConstraints)."""
        o = np.ascontiguousarray(options, dtype=JOINT_OPTIONS)
        a = np.ascontiguousarray(active, dtype=np.uint32)
        _check(lib().pmhWorldWriteJoints(self._h, first, len(o), _ptr(o), _ptr(a)), "pmhWorldWriteJoints")

    def set_rigid_states(self, first, states):
        """Replaces the states of bodies first.. (pmCmdSetRigidStates)."""
        a = np.ascontiguousarray(states, dtype=STATE)
        _check(lib().pmhWorldSetRigidStates(self._h, first, len(a), _ptr(a)), "pmhWorldSetRigidStates")

    def contact_events(self):
        """The last step's rigid contact events (layouts.CONTACT_EVENT)."""
        count = ctypes.c_uint32()
        _check(lib().pmhWorldGetContactEvents(self._h, None, 0, ctypes.byref(count)), "pmhWorldGetContactEvents")
        out = np.zeros(count.value, dtype=CONTACT_EVENT)
        if len(out):
            _check(lib().pmhWorldGetContactEvents(self._h, _ptr(out), len(out), ctypes.byref(count)),
                   "pmhWorldGetContactEvents")
        return out

    def step(self, steps=1):
        _check(lib().pmhWorldStep(self._h, steps), "pmhWorldStep")

    def set_color_wait_limit(self, polls):
        """Polls of one colouring wait before it gives up (default 2**22);
        0 gives up at the first wait on an uncoloured neighbour."""
        _check(lib().pmhWorldSetColorWaitLimit(self._h, polls), "pmhWorldSetColorWaitLimit")

    def set_sweep_wait_limit(self, polls):
        """Polls of one dataflow-sweep wait before it fails the step
        (default 2**22); 0 fails at the first unmet wait."""
        _check(lib().pmhWorldSetSweepWaitLimit(self._h, polls), "pmhWorldSetSweepWaitLimit")

    def residuals(self, maximum=False):
        """(per-iteration RMS constraint residuals [2, 64], NaN where no
        update ran; iterations run per phase) of the last step; with
        maximum=True also the per-iteration maxima."""
        out = np.zeros(128, dtype=np.float32)
        mx = np.zeros(128, dtype=np.float32)
        ex = np.zeros(2, dtype=np.uint32)
        _check(lib().pmhWorldGetResiduals(self._h, _ptr(out), _ptr(mx), _ptr(ex)), "pmhWorldGetResiduals")
        if maximum:
            return out.reshape(2, 64), ex, mx.reshape(2, 64)
        return out.reshape(2, 64), ex

    def states(self):
        out = np.zeros(self.body_count, dtype=STATE)
        _check(lib().pmhWorldCopyStates(self._h, _ptr(out)), "pmhWorldCopyStates")
        return out

    def set_states(self, states):
        s = np.ascontiguousarray(states, dtype=STATE)
        _check(lib().pmhWorldSetStates(self._h, _ptr(s)), "pmhWorldSetStates")

    def stats(self):
        s = WorldStats()
        _check(lib().pmhWorldGetStats(self._h, ctypes.byref(s)), "pmhWorldGetStats")
        return {n: getattr(s, n) for n in STATS_FIELDS}

    def phase_times(self):
        out = np.zeros(len(PHASES), dtype=np.float32)
        _check(lib().pmhWorldGetPhaseTimes(self._h, _ptr(out)), "pmhWorldGetPhaseTimes")
        return dict(zip(PHASES, map(float, out)))

    def solve_timing(self, reset=False):
        ms = _F()
        launches = _U()
        _check(lib().pmhWorldSolveTiming(self._h, ctypes.byref(ms), ctypes.byref(launches),
                                         1 if reset else 0), "pmhWorldSolveTiming")
        return ms.value, launches.value

    def kernel_timing(self, reset=False):
        """Colour-sweep launches timed with HIP events on the sampled steps:
        milliseconds, launches, algorithmic bytes."""
        t = KernelTiming()
        _check(lib().pmhWorldKernelTiming(self._h, ctypes.byref(t), 1 if reset else 0),
               "pmhWorldKernelTiming")
        return {f: getattr(t, f) for f, _ in KernelTiming._fields_}

    def totals(self, reset=False):
        t = WorldTotals()
        _check(lib().pmhWorldGetTotals(self._h, ctypes.byref(t), 1 if reset else 0),
               "pmhWorldGetTotals")
        return {f: getattr(t, f) for f, _ in WorldTotals._fields_}

    def pairs(self):
        out = np.zeros(lib().pmhWorldPairCount(self._h), dtype=np.uint32)
        _check(lib().pmhWorldGetPairs(self._h, _ptr(out)), "pmhWorldGetPairs")
        return out

    def manifolds(self):
        out = np.zeros(lib().pmhWorldPairCount(self._h), dtype=MANIFOLD)
        _check(lib().pmhWorldGetManifolds(self._h, _ptr(out)), "pmhWorldGetManifolds")
        return out

    def bodies(self):
        out = np.zeros(self.body_count, dtype=AVBD_BODY)
        _check(lib().pmhWorldGetAvbdBodies(self._h, _ptr(out)), "pmhWorldGetAvbdBodies")
        return out

    def schedule(self):
        out = np.zeros(1, dtype=SCHEDULE)
        _check(lib().pmhWorldGetSchedule(self._h, _ptr(out)), "pmhWorldGetSchedule")
        return out[0]

    def proxies(self):
        out = np.zeros((self.body_count, 6), dtype=np.float32)
        _check(lib().pmhWorldGetProxies(self._h, _ptr(out)), "pmhWorldGetProxies")
        return out


class DeviceArray:
    """A device copy of a numpy array (for the kernel seam)."""

    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.nbytes = max(1, self.host.nbytes)
        p = _P()
        _check(lib().pmhDeviceAlloc(self.nbytes, ctypes.byref(p)), "pmhDeviceAlloc")
        self.ptr = p
        if self.host.nbytes:
            _check(lib().pmhCopyToDevice(self.ptr, _ptr(self.host), self.host.nbytes), "pmhCopyToDevice")

    def read(self):
        out = np.empty_like(self.host)
        if out.nbytes:
            _check(lib().pmhCopyToHost(_ptr(out), self.ptr, out.nbytes), "pmhCopyToHost")
        return out

    def free(self):
        if self.ptr:
            lib().pmhDeviceFree(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


def prepare_solve(params, previous, states, manifolds, pairs, bodies=None,
                  dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), base=10, requested=0,
                  profile=1, entry="pmhPrepareSolveGeometryConstraints", joints=None, ids=None):
    """Kernel seam on the GPU with host arrays: copies in, runs `entry`
    (pmhPrepareSolveGeometryConstraints, or its variant that allocates
    nothing), copies back in place (states, manifolds, bodies) and returns
    the schedule record."""
    n = len(params)
    if bodies is None:
        bodies = np.zeros(n, dtype=AVBD_BODY)
    d_p = DeviceArray(np.ascontiguousarray(params, dtype=PARAMS))
    d_prev = DeviceArray(np.ascontiguousarray(previous, dtype=STATE))
    d_s = DeviceArray(np.ascontiguousarray(states, dtype=STATE))
    d_m = DeviceArray(np.ascontiguousarray(manifolds, dtype=MANIFOLD))
    d_pairs = DeviceArray(np.ascontiguousarray(pairs, dtype=np.uint32))
    d_count = DeviceArray(np.array([len(pairs)], dtype=np.uint32))
    d_b = DeviceArray(np.ascontiguousarray(bodies, dtype=AVBD_BODY))
    d_sched = DeviceArray(np.zeros(1, dtype=SCHEDULE))
    g = np.asarray(gravity, dtype=np.float32)
    if joints is not None:
        # pmhPrepareSolveGeometryConstraintsJoints: the joints (JOINT records)
        # and the id of each dense body; the joints' solver fields come back.
        d_j = DeviceArray(np.ascontiguousarray(joints, dtype=JOINT))
        d_ids = DeviceArray(np.ascontiguousarray(ids, dtype=np.uint32).reshape(n, 2))
        _check(lib().pmhPrepareSolveGeometryConstraintsJoints(
            None, d_p.ptr, d_prev.ptr, d_s.ptr, n, d_m.ptr, d_pairs.ptr, d_count.ptr, d_j.ptr, len(joints),
            d_ids.ptr, d_b.ptr, d_sched.ptr, dt, _ptr(g), base, requested, profile),
            "pmhPrepareSolveGeometryConstraintsJoints")
        joints[...] = d_j.read()
    else:
        _check(getattr(lib(), entry)(
            None, d_p.ptr, d_prev.ptr, d_s.ptr, n, d_m.ptr, d_pairs.ptr, d_count.ptr, d_b.ptr,
            d_sched.ptr, dt, _ptr(g), base, requested, profile), entry)
    states[...] = d_s.read()
    manifolds[...] = d_m.read()
    bodies[...] = d_b.read()
    return d_sched.read()[0]


def autotune_lanes(world, steps=8, candidates=(8, 16, 32, 64)):
    """Times `steps` steps of `world` with each lane tile of the colour
    sweep (the world keeps simulating meanwhile) and leaves it on the
    fastest.  Returns (best, {lanes: ms per step})."""
    import time
    times = {}
    for lanes in candidates:
        world.set_solve_lanes(lanes)
        world.step(1)
        lib().pmhDeviceSynchronize()
        t0 = time.perf_counter()
        world.step(steps)
        lib().pmhDeviceSynchronize()
        times[lanes] = (time.perf_counter() - t0) * 1e3 / steps
    best = min(times, key=times.get)
    world.set_solve_lanes(best)
    return best, times


# ---------------------------------------------------------------------------
# Learned warm start: a 16 -> 64 -> 64 -> 2 network on MFMA (BF16 operands,
# FP32 accumulation); parameters packed W1[64][16], b1, W2[64][64], b2,
# W3[2][64], b3, then the largest training target of each output, which
# bounds the predictions (5380 floats).
# ---------------------------------------------------------------------------
MLP_PARAMS = 5380
WARMSTART_PENALTY, WARMSTART_LAMBDA = 1, 2
MLP_FEATURES = 16
MLP_SAMPLE = 18


def warmstart_init(seed=0):
    p = np.zeros(MLP_PARAMS, dtype=np.float32)
    lib().pmhWarmstartInit(seed, _ptr(p))
    return p


def warmstart_forward(params, features):
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, MLP_FEATURES)
    y = np.zeros((len(x), 2), dtype=np.float32)
    _check(lib().pmhWarmstartForward(_ptr(p), _ptr(x), len(x), _ptr(y)), "pmhWarmstartForward")
    return y


def warmstart_gradients(params, samples):
    """Mean-squared-error loss and gradients of at most 64 samples."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    s = np.ascontiguousarray(samples, dtype=np.float32).reshape(-1, MLP_SAMPLE)
    g = np.zeros(MLP_PARAMS, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    _check(lib().pmhWarmstartGradients(_ptr(s), len(s), _ptr(p), _ptr(g), _ptr(loss)), "pmhWarmstartGradients")
    return float(loss[0]), g


def warmstart_train(samples, epochs=4, learning_rate=1e-3, seed=0, params=None):
    """Adam over minibatches of 64 on the GPU; returns (params, loss per epoch)."""
    s = np.ascontiguousarray(samples, dtype=np.float32).reshape(-1, MLP_SAMPLE)
    p = warmstart_init(seed) if params is None else np.array(params, dtype=np.float32)
    loss = np.zeros(epochs, dtype=np.float32)
    _check(lib().pmhWarmstartTrain(_ptr(s), len(s), epochs, learning_rate, seed, _ptr(p), _ptr(loss)),
           "pmhWarmstartTrain")
    return p, loss

