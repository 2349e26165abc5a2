# SPDX-License-Identifier: MIT
"""ctypes bindings to the CPU oracle (oracle/lib/libpm_oracle.so).

Test infrastructure: only tests, the smoke check and the benchmark's CPU
baseline leg load the oracle, and only as the checker.
"""
import ctypes
import os

import numpy as np

from pmamd import layouts as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "lib", "libpm_oracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libpm_ref_cases.so")

_P = ctypes.c_void_p
_F = ctypes.c_float
_U = ctypes.c_uint32


def _ptr(a):
    return a.ctypes.data_as(_P)


class WorldDesc(ctypes.Structure):
    _fields_ = [("gravity", _F * 3), ("dt", _F), ("profile_version", _U),
                ("base_iterations", _U), ("requested_iterations", _U),
                ("body_capacity", _U), ("lanes", _U), ("flags", _U),
                ("residual_tolerance", _F), ("substeps", _U)]


STATS_FIELDS = ["pairs", "manifolds", "points", "colors", "islands", "impact",
                "budget", "unsupported_pairs", "cache_hits", "phases", "triangle_pairs",
                "fallback_pairs"]


class WorldStats(ctypes.Structure):
    _fields_ = [(name, _U) for name in STATS_FIELDS]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise FileNotFoundError(
                f"{ORACLE_LIB} is missing: run `make -C oracle`")
        _lib = ctypes.CDLL(ORACLE_LIB)
        _lib.pmo_world_create.restype = _P
        _lib.pmo_world_create.argtypes = [ctypes.POINTER(WorldDesc)]
        for name in ("pmo_world_destroy",):
            getattr(_lib, name).argtypes = [_P]
        _lib.pmo_world_add_mesh.argtypes = [_P, _P, _U, _P, _U]
        _lib.pmo_world_add_sphere.argtypes = [_P, _F]
        _lib.pmo_world_add_capsule.argtypes = [_P, _F, _F]
        _lib.pmo_world_add_bodies.argtypes = [_P, _U, _P, _P]
        _lib.pmo_world_step.argtypes = [_P, _U]
        _lib.pmo_world_apply_inputs.argtypes = [_P, _U, _P]
        _lib.pmo_world_add_joints.argtypes = [_P, _U, _P]
        _lib.pmo_world_residuals.argtypes = [_P, _P, _P, _P]
        _lib.pmo_world_write_joints.argtypes = [_P, _U, _U, _P, _P]
        _lib.pmo_world_set_rigid_states.argtypes = [_P, _U, _U, _P]
        _lib.pmo_world_contact_events.argtypes = [_P, _P, _U]
        _lib.pmo_world_contact_events.restype = _U
        _lib.pmo_mlp_forward.argtypes = [_P, _P, _U, _P, ctypes.c_int]
        _lib.pmo_mlp_forward.restype = None
        _lib.pmo_mlp_gradients.argtypes = [_P, _P, _U, _P, ctypes.c_int]
        _lib.pmo_mlp_gradients.restype = ctypes.c_float
        _lib.pmo_world_joint_count.argtypes = [_P]
        _lib.pmo_world_joint_count.restype = _U
        _lib.pmo_world_get_joints.argtypes = [_P, _P]
        _lib.pmo_world_set_kinematic_targets.argtypes = [_P, _U, _U, _P]
        _lib.pmo_world_body_count.argtypes = [_P]
        _lib.pmo_world_body_count.restype = _U
        _lib.pmo_world_timing.argtypes = [_P, _P, ctypes.c_int]
        _lib.pmo_world_pair_count.argtypes = [_P]
        _lib.pmo_world_pair_count.restype = _U
        for name in ("pmo_world_get_states", "pmo_world_set_states",
                     "pmo_world_get_pairs", "pmo_world_get_manifolds",
                     "pmo_world_get_bodies", "pmo_world_get_schedule",
                     "pmo_world_get_proxies", "pmo_world_get_stats"):
            getattr(_lib, name).argtypes = [_P, _P]
        _lib.pmo_world_get_shape.argtypes = [_P, _U, _P]
        _lib.pmo_world_get_mesh_indices.argtypes = [_P, _U, _P]
        _lib.pmo_world_get_mesh_planes.argtypes = [_P, _U, _P]
        _lib.pmo_avbd_prepare_solve.argtypes = [
            _P, _P, _P, _U, _P, _P, _U, _P, _P, _F, _P, _U, _U, _U, _U]
        _lib.pmo_avbd_prepare_solve_joints.argtypes = [
            _P, _P, _P, _U, _P, _P, _U, _P, _U, _P, _P, _P, _F, _P, _U, _U, _U, _U]
        for name in ("pmo_row_error", "pmo_row_trial", "pmo_row_force",
                     "pmo_force"):
            getattr(_lib, name).restype = _F
        _lib.pmo_row_error.argtypes = [_P, _P, _P]
        _lib.pmo_row_trial.argtypes = [_P, _P, _P, _F]
        _lib.pmo_row_force.argtypes = [_P, _P, _P, _F]
        _lib.pmo_row_accumulate.argtypes = [_P, _P, _P, ctypes.c_int, _P, _P, _F]
        _lib.pmo_row_advance.argtypes = [_P, _P, _P]
        _lib.pmo_block_solve.argtypes = [_P, _P, _P]
        _lib.pmo_block_add_row.argtypes = [_P, _P, _P, _F, _F]
        _lib.pmo_warm_start.argtypes = [_P, _F, _P]
        _lib.pmo_force.argtypes = [_P, _F, _F, _F]
        _lib.pmo_update_dual.argtypes = [_P, _F, _F, _F, _F, _F, _P]
        _lib.pmo_project_contact.argtypes = [_P, _F, _P]
        _lib.pmo_stiffness_scales.argtypes = [_P, _F, _P]
        _lib.pmo_world_get_mesh_bvh.argtypes = [_P, _U, _P, _P]
        _lib.pmo_world_get_mesh_bvh.restype = _U
        _lib.pmo_world_get_mesh_shell.argtypes = [_P, _U, _P]
        _lib.pmo_world_get_mesh_shell.restype = _U
        _lib.pmo_closest_segments.argtypes = [_P] * 6
        _lib.pmo_closest_on_triangle.argtypes = [_P] * 3
        _lib.pmo_closest_triangle_pair.argtypes = [_P] * 4
        _lib.pmo_triangle_feature_normal.argtypes = [_P] * 6
    return _lib


def f32(values, n=None):
    a = np.zeros(n if n is not None else len(values), dtype=np.float32)
    a[:len(values)] = values
    return a


class OracleWorld:
    """The oracle's world with the same surface as pmamd.World."""

    def __init__(self, dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), profile=1,
                 base_iterations=10, requested_iterations=0,
                 body_capacity=0, lanes=0, brute_force=False, residual_tolerance=0.0, substeps=1, planted=0):
        """planted: mistakes planted by tests (1: no free-motion mean,
        2: no crossing safeguard)."""
        desc = WorldDesc((_F * 3)(*gravity), dt, profile, base_iterations,
                         requested_iterations, body_capacity, lanes,
                         (1 if brute_force else 0) | (planted << 1), residual_tolerance, substeps)
        self._h = lib().pmo_world_create(ctypes.byref(desc))
        if not self._h:
            raise ValueError("invalid world description")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pmo_world_destroy(self._h)
            self._h = None

    def add_mesh(self, vertices, indices):
        v = np.ascontiguousarray(vertices, dtype=np.float32)
        i = np.ascontiguousarray(indices, dtype=np.uint32)
        r = lib().pmo_world_add_mesh(self._h, _ptr(v), len(v), _ptr(i), len(i))
        if r < 0:
            raise ValueError("mesh rejected")
        return r

    def add_sphere(self, radius):
        return lib().pmo_world_add_sphere(self._h, radius)

    def add_capsule(self, radius, half_length):
        return lib().pmo_world_add_capsule(self._h, radius, half_length)

    def add_bodies(self, params, states):
        p = np.ascontiguousarray(params, dtype=L.PARAMS)
        s = np.ascontiguousarray(states, dtype=L.STATE)
        if lib().pmo_world_add_bodies(self._h, len(p), _ptr(p), _ptr(s)) != 0:
            raise ValueError("bodies rejected")

    def add_joints(self, options):
        a = np.ascontiguousarray(options, dtype=L.JOINT_OPTIONS)
        if lib().pmo_world_add_joints(self._h, len(a), _ptr(a)) != 0:
            raise ValueError("invalid joint options")

    def residuals(self, maximum=False):
        """(per-iteration RMS residuals [2, 64] with NaN where no update
        ran, iterations run per phase) of the last step; with maximum=True
        also the per-iteration maxima."""
        out = np.zeros(128, dtype=np.float32)
        mx = np.zeros(128, dtype=np.float32)
        ex = np.zeros(2, dtype=np.uint32)
        lib().pmo_world_residuals(self._h, _ptr(out), _ptr(mx), _ptr(ex))
        if maximum:
            return out.reshape(2, 64), ex, mx.reshape(2, 64)
        return out.reshape(2, 64), ex

    def joints(self):
        out = np.zeros(lib().pmo_world_joint_count(self._h), dtype=L.JOINT)
        if len(out):
            lib().pmo_world_get_joints(self._h, _ptr(out))
        return out

    def apply_inputs(self, inputs):
        a = np.ascontiguousarray(inputs, dtype=L.RIGID_INPUT)
        if lib().pmo_world_apply_inputs(self._h, len(a), _ptr(a)) != 0:
            raise ValueError("invalid rigid inputs")

    def write_joints(self, first, options, active):
        o = np.ascontiguousarray(options, dtype=L.JOINT_OPTIONS)
        a = np.ascontiguousarray(active, dtype=np.uint32)
        if lib().pmo_world_write_joints(self._h, first, len(o), _ptr(o), _ptr(a)) != 0:
            raise ValueError("invalid joint write")

    def set_rigid_states(self, first, states):
        a = np.ascontiguousarray(states, dtype=L.STATE)
        if lib().pmo_world_set_rigid_states(self._h, first, len(a), _ptr(a)) != 0:
            raise ValueError("invalid rigid states")

    def contact_events(self):
        total = lib().pmo_world_contact_events(self._h, None, 0)
        out = np.zeros(total, dtype=L.CONTACT_EVENT)
        if total:
            lib().pmo_world_contact_events(self._h, _ptr(out), total)
        return out

    def set_kinematic_targets(self, first, targets):
        a = np.ascontiguousarray(targets, dtype=L.STATE)
        if lib().pmo_world_set_kinematic_targets(self._h, first, len(a), _ptr(a)) != 0:
            raise ValueError("invalid kinematic targets")

    def step(self, steps=1):
        r = lib().pmo_world_step(self._h, steps)
        if r != 0:
            raise RuntimeError(f"oracle step failed ({r})")

    @property
    def body_count(self):
        return lib().pmo_world_body_count(self._h)

    def states(self):
        out = np.zeros(self.body_count, dtype=L.STATE)
        lib().pmo_world_get_states(self._h, _ptr(out))
        return out

    def timing(self, reset=False):
        out = np.zeros(4, dtype=np.float64)
        lib().pmo_world_timing(self._h, _ptr(out), 1 if reset else 0)
        return {"step_ms": out[0], "prepare_ms": out[1], "solve_ms": out[2],
                "body_iterations": int(out[3])}

    def set_states(self, states):
        s = np.ascontiguousarray(states, dtype=L.STATE)
        lib().pmo_world_set_states(self._h, _ptr(s))

    def pairs(self):
        out = np.zeros(lib().pmo_world_pair_count(self._h), dtype=np.uint32)
        lib().pmo_world_get_pairs(self._h, _ptr(out))
        return out

    def manifolds(self):
        out = np.zeros(lib().pmo_world_pair_count(self._h), dtype=L.MANIFOLD)
        lib().pmo_world_get_manifolds(self._h, _ptr(out))
        return out

    def bodies(self):
        out = np.zeros(self.body_count, dtype=L.AVBD_BODY)
        lib().pmo_world_get_bodies(self._h, _ptr(out))
        return out

    def schedule(self):
        out = np.zeros(1, dtype=L.SCHEDULE)
        lib().pmo_world_get_schedule(self._h, _ptr(out))
        return out[0]

    def proxies(self):
        out = np.zeros((self.body_count, 6), dtype=np.float32)
        lib().pmo_world_get_proxies(self._h, _ptr(out))
        return out

    def stats(self):
        s = WorldStats()
        lib().pmo_world_get_stats(self._h, ctypes.byref(s))
        return {name: getattr(s, name) for name in STATS_FIELDS}

    def shape(self, index):
        out = np.zeros(16, dtype=np.float32)
        lib().pmo_world_get_shape(self._h, index, _ptr(out))
        return out

    def mesh_indices(self, index, count):
        out = np.zeros(count, dtype=np.uint32)
        lib().pmo_world_get_mesh_indices(self._h, index, _ptr(out))
        return out

    def mesh_bvh(self, index):
        n = lib().pmo_world_get_mesh_bvh(self._h, index, None, None)
        nodes = np.zeros((n, 4), np.uint32)
        bounds = np.zeros((n, 6), np.float32)
        lib().pmo_world_get_mesh_bvh(self._h, index, _ptr(nodes), _ptr(bounds))
        return nodes, bounds

    def mesh_shell(self, index):
        n = lib().pmo_world_get_mesh_shell(self._h, index, None)
        out = np.zeros((n, 3), np.float32)
        if n:
            lib().pmo_world_get_mesh_shell(self._h, index, _ptr(out))
        return out

    def mesh_planes(self, index, count):
        out = np.zeros((count, 4), dtype=np.float32)
        lib().pmo_world_get_mesh_planes(self._h, index, _ptr(out))
        return out


def closest_segments(p0, p1, q0, q1):
    on_p, on_q = np.zeros(3, np.float32), np.zeros(3, np.float32)
    args = [np.ascontiguousarray(v, dtype=np.float32) for v in (p0, p1, q0, q1)]
    lib().pmo_closest_segments(*[_ptr(a) for a in args], _ptr(on_p), _ptr(on_q))
    return on_p, on_q


def closest_on_triangle(p, tri):
    out = np.zeros(3, np.float32)
    pp = np.ascontiguousarray(p, dtype=np.float32)
    t = np.ascontiguousarray(tri, dtype=np.float32).reshape(9)
    lib().pmo_closest_on_triangle(_ptr(pp), _ptr(t), _ptr(out))
    return out


def closest_triangle_pair(a, b):
    on_a, on_b = np.zeros(3, np.float32), np.zeros(3, np.float32)
    ta = np.ascontiguousarray(a, dtype=np.float32).reshape(9)
    tb = np.ascontiguousarray(b, dtype=np.float32).reshape(9)
    lib().pmo_closest_triangle_pair(_ptr(ta), _ptr(tb), _ptr(on_a), _ptr(on_b))
    return on_a, on_b


def triangle_feature_normal(a, b, on_a, on_b, fallback):
    out = np.zeros(3, np.float32)
    args = [np.ascontiguousarray(a, dtype=np.float32).reshape(9),
            np.ascontiguousarray(b, dtype=np.float32).reshape(9)]
    args += [np.ascontiguousarray(v, dtype=np.float32) for v in (on_a, on_b, fallback)]
    lib().pmo_triangle_feature_normal(*[_ptr(x) for x in args], _ptr(out))
    return out


def prepare_solve(params, previous, states, manifolds, pairs, bodies=None,
                  dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), base=10,
                  requested=0, profile=1, lanes=0, joints=None, ids=None):
    """Kernel-seam restatement: prepare_avbd_kernel + solve_avbd_kernel on
    caller-built manifolds. States, manifolds and bodies are updated in
    place; the schedule record is returned.  With `joints` (JOINT records)
    and `ids` (the id of each dense body), the joints take part and their
    solver-owned fields are updated in place."""
    n = len(params)
    if bodies is None:
        bodies = np.zeros(n, dtype=L.AVBD_BODY)
    schedule = np.zeros(1, dtype=L.SCHEDULE)
    g = f32(gravity)
    if joints is not None:
        ids = np.ascontiguousarray(ids, dtype=np.uint32).reshape(n, 2)
        r = lib().pmo_avbd_prepare_solve_joints(
            _ptr(params), _ptr(previous), _ptr(states), n, _ptr(manifolds),
            _ptr(pairs), len(pairs), _ptr(joints), len(joints), _ptr(ids), _ptr(bodies), _ptr(schedule), dt,
            _ptr(g), base, requested, profile, lanes)
        if r != 0:
            raise RuntimeError("prepare_solve failed")
        return schedule[0]
    r = lib().pmo_avbd_prepare_solve(
        _ptr(params), _ptr(previous), _ptr(states), n, _ptr(manifolds),
        _ptr(pairs), len(pairs), _ptr(bodies), _ptr(schedule), dt, _ptr(g),
        base, requested, profile, lanes)
    if r != 0:
        raise RuntimeError("prepare_solve failed")
    return schedule[0]


def mlp_forward(params, features, round_bf16=True):
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, 16)
    y = np.zeros((len(x), 2), dtype=np.float32)
    lib().pmo_mlp_forward(_ptr(p), _ptr(x), len(x), _ptr(y), 1 if round_bf16 else 0)
    return y


def mlp_gradients(params, samples, round_bf16=True):
    p = np.ascontiguousarray(params, dtype=np.float32)
    s = np.ascontiguousarray(samples, dtype=np.float32).reshape(-1, 18)
    g = np.zeros(len(p), dtype=np.float32)
    loss = lib().pmo_mlp_gradients(_ptr(p), _ptr(s), len(s), _ptr(g), 1 if round_bf16 else 0)
    return float(loss), g

