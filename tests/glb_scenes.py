# SPDX-License-Identifier: MIT
"""The reference's integrated (GLB) rigid conformance scenes, built on a
world with the oracle's or the library's surface from the fixtures under
tests/golden/glb (written by tests/golden/make_glb_fixtures.py).

What is restated from the reference, in float32 as there:
- load_glb_scene (examples/support/scene.cpp:636-1000) for the rigid part:
  collision proxies (:664-710), rigid bodies with their metadata
  (read_metadata, :156-217; defaults of RigidBodyOptions,
  include/parallel_mater_scene/authoring.hpp:127-143), primitives scaled by
  the node's scale (append_primitive, :260-360), render meshes shared by
  nodes of one glTF mesh and scale (:820-846), each body's geometry centred
  on its render bounds and its position moved by the rotated centre
  (finalize_body_geometry, :484-525), and rigid constraints (:905-1087) with
  the Blender basis fix (the frame's rotation times a -90 degree turn about
  x) and each body's local frame from its centred initial state;
- the command scene's geometry (examples/gallery/scene_builder_create.cpp.inc
  :117-140 and scene_builder.cpp:393-420): one shape per distinct list of
  meshes, the collision proxy's meshes when a body has one, else its render
  meshes, concatenated;
- make_parameters (src/simulation/runtime.cpp:459-482): a zero inertia
  diagonal becomes the shape's unit inertia times the mass;
- the record names (conformance/v1/src/command_runner.cpp:1074-1090):
  "rigid/<source name>/<index>" and "constraint/<name>/<index>".
"""
import json
import pathlib

import numpy as np

import scenarios as S
from pmamd import layouts as L

GLB = pathlib.Path(__file__).resolve().parent / "golden" / "glb"
CASES = ("constraint-point", "constraint-fixed", "constraint-hinge", "constraint-slider", "constraint-piston",
         "constraint-generic", "constraint-generic-spring", "constraint-motor", "passive-active")
MOTION = {"static": L.MOTION_STATIC, "kinematic": L.MOTION_KINEMATIC, "dynamic": L.MOTION_DYNAMIC}
JOINT_TYPE = {"fixed": L.JOINT_FIXED, "point": L.JOINT_POINT, "hinge": L.JOINT_HINGE, "slider": L.JOINT_SLIDER,
              "piston": L.JOINT_PISTON, "generic": L.JOINT_GENERIC, "generic_spring": L.JOINT_GENERIC_SPRING,
              "motor": L.JOINT_MOTOR}
F = np.float32


def _cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], dtype=F)


def rotate(q, v):
    """v + 2w (q x v) + q x 2(q x v), scene.cpp:136-141."""
    q = np.asarray(q, dtype=F)
    v = np.asarray(v, dtype=F)
    t = _cross(q[:3], v) * F(2.0)
    return v + (t * q[3] + _cross(q[:3], t))


def conjugate(q):
    return np.array([-q[0], -q[1], -q[2], q[3]], dtype=F)


def qmul(a, b):
    a = np.asarray(a, dtype=F)
    b = np.asarray(b, dtype=F)
    return np.array([a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1],
                     a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0],
                     a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3],
                     a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2]], dtype=F)


def _rotation_x(radians):
    half = F(radians) * F(0.5)
    return np.array([np.sin(half), 0.0, 0.0, np.cos(half)], dtype=F)


class Scene:
    """A loaded GLB scene: bodies (dicts with name, source_name, options,
    state, render and collision mesh lists), constraints (dicts with name
    and JOINT_OPTIONS fields) and meshes (float32 positions, uint32 indices)."""

    def __init__(self, name):
        self.name = name
        self.fixture = json.loads((GLB / f"{name}.json").read_text())
        self.arrays = np.load(GLB / f"{name}.npz")
        self.meshes, self.collision_meshes = [], []
        self.bodies, self.constraints = [], []
        self._load()

    def _primitives(self, node, scale, into):
        out = []
        for p, prim in enumerate(self.fixture["meshes"][str(node["mesh"])]):
            assert prim["mode"] == 4, "only triangle primitives are supported"
            pos = self.arrays[f"mesh{node['mesh']}_prim{p}_positions"].astype(F) * scale
            idx = self.arrays[f"mesh{node['mesh']}_prim{p}_indices"].astype(np.uint32)
            out.append(len(into))
            into.append([pos, idx])
        return out

    @staticmethod
    def _state(node):
        st = S.states(1)
        if node["translation"] is not None:
            st["position"][0] = node["translation"]
        if node["rotation"] is not None:
            st["orientation"][0] = node["rotation"]
        return st

    @staticmethod
    def _scale(node):
        return np.asarray(node["scale"] if node["scale"] is not None else (1.0, 1.0, 1.0), dtype=F)

    def _load(self):
        proxies = {}
        for node in self.fixture["nodes"]:                             # scene.cpp:664-710
            e = node["extras"]
            if e.get("pm_system") != "collision_mesh":
                continue
            name = e.get("pm_name", node["name"])
            proxies[name] = {"state": self._state(node),
                             "meshes": self._primitives(node, self._scale(node), self.collision_meshes)}
        shared = {}
        for node in self.fixture["nodes"]:                             # :750-900
            e = node["extras"]
            if e.get("pm_system") not in ("rigid_body", "sphere_cluster"):
                continue
            assert e["pm_schema"] == 2 and not node["has_parent"] and not node["has_matrix"]
            body = {"name": e.get("pm_name", node["name"])}
            body["source_name"] = e.get("pm_source_name", body["name"])
            body["motion"] = MOTION[e["pm_motion"]]
            body["mass"] = F(e.get("pm_mass", 1.0))
            body["friction"] = F(e.get("pm_friction", 0.5))
            body["restitution"] = F(e.get("pm_restitution", 0.0))
            body["linear_damping"] = F(e.get("pm_linear_damping", 0.05))
            body["angular_damping"] = F(e.get("pm_angular_damping", 0.05))
            body["margin"] = F(e.get("pm_collision_margin", 0.005))
            scale = self._scale(node)
            st = self._state(node)
            st["linear_velocity"][0] = [e.get(f"pm_initial_velocity_{a}", 0.0) for a in "xyz"]
            proxy = e.get("pm_collision_proxy")
            key = node["mesh"]
            reuse = proxy is None and key in shared and (shared[key]["scale"] == scale).all()
            body["render"] = shared[key]["meshes"] if reuse else self._primitives(node, scale, self.meshes)
            body["collision"] = []
            if proxy is not None:
                body["collision"] = proxies[proxy]["meshes"]
            if reuse:
                st["position"][0] = st["position"][0] + rotate(st["orientation"][0], shared[key]["center"])
            else:
                center = self._finalize(body, st)
                if proxy is None:
                    shared[key] = {"scale": scale, "meshes": body["render"], "center": center}
            body["state"] = st
            assert e["pm_system"] == "rigid_body", "sphere clusters are not rigid bodies of the scene"
            self.bodies.append(body)
        for node in self.fixture["nodes"]:                             # :905-1087
            e = node["extras"]
            if e.get("pm_system") != "rigid_constraint":
                continue
            self.constraints.append(self._constraint(node, e))

    def _finalize(self, body, st):
        """finalize_body_geometry, scene.cpp:484-525."""
        pts = np.concatenate([self.meshes[m][0] for m in body["render"]])
        lo, hi = pts.min(0), pts.max(0)
        center = (lo + hi) * F(0.5)
        for m in body["render"]:
            self.meshes[m][0] = self.meshes[m][0] - center
        for m in body["collision"]:
            self.collision_meshes[m][0] = self.collision_meshes[m][0] - center
        st["position"][0] = st["position"][0] + rotate(st["orientation"][0], center)
        return center

    def _find(self, target):
        found = [k for k, b in enumerate(self.bodies) if b["name"] == target or b["source_name"] == target]
        assert len(found) == 1, f"constraint body {target!r} is missing or ambiguous"
        return found[0]

    def _constraint(self, node, e):
        o = np.zeros(1, dtype=L.JOINT_OPTIONS)
        o["type"] = JOINT_TYPE[e["pm_constraint_type"]]
        a, b = self._find(e.get("pm_body_a", "")), self._find(e.get("pm_body_b", ""))
        assert a != b
        frame = self._state(node)
        orientation = qmul(frame["orientation"][0], _rotation_x(-np.pi / 2))
        for side, k in (("a", a), ("b", b)):
            bs = self.bodies[k]["state"]
            inverse = conjugate(bs["orientation"][0])
            o[f"body_{side}"][0] = (k, 1)
            o[f"local_anchor_{side}"][0] = rotate(inverse, frame["position"][0] - bs["position"][0])
            o[f"local_orientation_{side}"][0] = qmul(inverse, orientation)
        o["enabled"] = int(e.get("pm_enabled", True))
        o["disable_collisions"] = int(e.get("pm_disable_collisions", True))
        o["breaking_impulse_threshold"] = e.get("pm_breaking_impulse_threshold", 0.0)
        o["solver_iterations"] = int(e.get("pm_solver_iterations", 8))
        for axis, s in enumerate("xyz"):
            bit = 1 << axis
            for field, use, lower, upper in (("linear_limits", "pm_use_limit_lin_", "pm_limit_lin_{}_lower",
                                              "pm_limit_lin_{}_upper"),
                                             ("angular_limits", "pm_use_limit_ang_", "pm_limit_ang_{}_lower",
                                              "pm_limit_ang_{}_upper")):
                if e.get(use + s, False):
                    o[field]["axes"] |= bit
                o[field]["lower"][0, axis] = e.get(lower.format(s), 0.0)
                o[field]["upper"][0, axis] = e.get(upper.format(s), 0.0)
            for field, use, stiff, damp in (("linear_springs", "pm_use_spring_", "pm_spring_stiffness_",
                                             "pm_spring_damping_"),
                                            ("angular_springs", "pm_use_spring_ang_", "pm_spring_stiffness_ang_",
                                             "pm_spring_damping_ang_")):
                if e.get(use + s, False):
                    o[field]["axes"] |= bit
                o[field]["stiffness"][0, axis] = e.get(stiff + s, 0.0)
                o[field]["damping"][0, axis] = e.get(damp + s, 0.0)
        m = o["motor"]
        m["linear_enabled"] = int(e.get("pm_use_motor_lin", False))
        m["linear_target_velocity"] = e.get("pm_motor_lin_target_velocity", 0.0)
        m["linear_maximum_impulse"] = e.get("pm_motor_lin_max_impulse", 1.0)
        m["angular_enabled"] = int(e.get("pm_use_motor_ang", False))
        m["angular_target_velocity"] = e.get("pm_motor_ang_target_velocity", 0.0)
        m["angular_maximum_impulse"] = e.get("pm_motor_ang_max_impulse", 1.0)
        o["motor"] = m
        return {"name": e.get("pm_name", node["name"]), "options": o}

    def build(self, world):
        """Adds shapes, bodies and joints to `world`; returns the record
        names of the bodies and constraints and the joint options."""
        shapes, index = {}, []
        for body in self.bodies:
            collision = bool(body["collision"])
            meshes = body["collision"] if collision else body["render"]
            key = (collision, tuple(meshes))
            if key not in shapes:
                source = self.collision_meshes if collision else self.meshes
                verts, idx, base = [], [], 0
                for m in meshes:
                    verts.append(source[m][0])
                    idx.append(source[m][1] + np.uint32(base))
                    base += len(source[m][0])
                shapes[key] = world.add_mesh(np.concatenate(verts), np.concatenate(idx))
            index.append(shapes[key])
        params = []
        for body, shape in zip(self.bodies, index):
            unit = world.shape(shape)[10:13].astype(F)
            inertia = unit * body["mass"]
            p = S.params(1, motion=body["motion"], mass=body["mass"], inertia=inertia, friction=body["friction"],
                         restitution=body["restitution"], margin=body["margin"], shape=shape,
                         max_linear=100.0, max_angular=100.0)
            p["linear_damping"] = body["linear_damping"]
            p["angular_damping"] = body["angular_damping"]
            params.append(p)
        world.add_bodies(np.concatenate(params), np.concatenate([b["state"] for b in self.bodies]))
        joints = np.concatenate([c["options"] for c in self.constraints]) if self.constraints else None
        if joints is not None:
            world.add_joints(joints)
        bodies = [f"rigid/{b['source_name']}/{k}" for k, b in enumerate(self.bodies)]
        constraints = [f"constraint/{c['name']}/{k}" for k, c in enumerate(self.constraints)]
        return bodies, constraints, joints
