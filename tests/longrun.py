# SPDX-License-Identifier: MIT
"""Long-run aggregates of a world: what a long run must keep, as the
reference's own realtime gates phrase it (maximum penetration, energy
growth, drift; tests/solver/rigid_realtime_cuda_tests.cu:365-370), taken
over the configurations' scenes:

- the 55-box pyramid (10 rows), 300 steps;
- a 4096-body slice of the 32k pile (16^3 bodies, same generator), 300 steps;
- 16 chains of 64 capsule links (ball and hinge joints in turn), 300 steps.

Every `every` steps a sample records the deepest contact of the step's
manifolds (the reported penetration, which includes the rest offset), the
deepest reach of any dynamic body below the floor's top face (y = 0 in all
three scenes: the lowest vertex of a mesh, the lowest point of a sphere or
capsule), the total mechanical energy (linear and angular kinetic, and
m g y), the mean height of the dynamic bodies, the contact-point count and,
for joints, the largest and mean distance between each joint's two anchor
points in world space.  tests/golden/make_longrun_fixtures.py writes these
series for the oracle in the reference's summation order (lanes 0) and in
the GPU's (lanes 32); the GPU's run is compared with both.
"""
import numpy as np

import scenarios as S

GRAVITY = 9.81
STEPS = 300
SCENES = {
    "pyramid": dict(make=lambda: S.pyramid(), every=10),
    "pile4096": dict(make=lambda: S.pile(side=16), every=20),
    "chains16x64": dict(make=lambda: S.chains(count=16, links=64), every=20),
}


def _rotate(q, v):
    """Rotates rows of v by the unit quaternions q (x, y, z, w)."""
    u, w = q[:, :3], q[:, 3:4]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def energy(params, states):
    """Total mechanical energy of the dynamic bodies (float64)."""
    dyn = params["inverse_mass"] > 0
    m = 1.0 / params["inverse_mass"][dyn].astype(np.float64)
    inertia = 1.0 / params["inverse_inertia_local"][dyn].astype(np.float64)
    s = states[dyn]
    q = s["orientation"].astype(np.float64)
    v = s["linear_velocity"].astype(np.float64)
    w = s["angular_velocity"].astype(np.float64)
    conj = q * np.array([-1.0, -1.0, -1.0, 1.0])
    w_local = _rotate(conj, w)
    kinetic = 0.5 * (m * (v * v).sum(1)).sum() + 0.5 * (inertia * w_local * w_local).sum()
    potential = (m * GRAVITY * s["position"][:, 1].astype(np.float64)).sum()
    return float(kinetic + potential)


def joint_gaps(joints, states):
    """World distance between each joint's two anchor points."""
    a = joints["body_a"][:, 0]
    b = joints["body_b"][:, 0]
    pa = states["position"][a].astype(np.float64)
    pb = states["position"][b].astype(np.float64)
    qa = states["orientation"][a].astype(np.float64)
    qb = states["orientation"][b].astype(np.float64)
    wa = pa + _rotate(qa, joints["local_anchor_a"].astype(np.float64))
    wb = pb + _rotate(qb, joints["local_anchor_b"].astype(np.float64))
    return np.linalg.norm(wa - wb, axis=1)


class Lowest:
    """Lowest world point of each dynamic body of a scene (floor at y = 0)."""

    def __init__(self, scene):
        dyn = np.flatnonzero(scene.params["inverse_mass"] > 0)
        self.dyn = dyn
        shapes = [scene.shapes[k] for k in scene.params["mesh_index"][dyn]]
        width = max([len(sh[1]) for sh in shapes if sh[0] == "mesh"], default=1)
        self.verts = np.zeros((len(dyn), width, 3))
        self.radius = np.zeros(len(dyn))
        for i, sh in enumerate(shapes):
            if sh[0] == "mesh":
                v = np.asarray(sh[1], np.float64)
                self.verts[i, :len(v)] = v
                self.verts[i, len(v):] = v[0]
            elif sh[0] == "sphere":
                self.radius[i] = sh[1]
            else:                                   # capsule: radius, half segment along local y
                self.radius[i] = sh[1]
                self.verts[i, 0] = (0.0, sh[2], 0.0)
                self.verts[i, 1:] = (0.0, -sh[2], 0.0)

    def depth(self, states):
        """Deepest reach below y = 0 over the dynamic bodies (negative when
        every body is clear of the floor)."""
        s = states[self.dyn]
        q = np.repeat(s["orientation"].astype(np.float64), self.verts.shape[1], axis=0)
        v = _rotate(q, self.verts.reshape(-1, 3)).reshape(self.verts.shape)
        low = (v[:, :, 1].min(1) + s["position"][:, 1].astype(np.float64)) - self.radius
        return float(-low.min())


def deepest(manifolds):
    if len(manifolds) == 0:
        return 0.0
    count = manifolds["count"]
    pen = manifolds["contacts"]["penetration"]
    mask = np.arange(pen.shape[1])[None, :] < count[:, None]
    return float(pen[mask].max()) if mask.any() else 0.0


def run(name, world, steps=STEPS):
    """Builds the named scene on `world`, steps it and returns the sampled
    series (lists of floats, one entry per sample)."""
    cfg = SCENES[name]
    scene = cfg["make"]()
    scene.build(world)
    dyn = scene.params["inverse_mass"] > 0
    low = Lowest(scene)
    out = {"step": [], "deepest": [], "floor_depth": [], "energy": [], "mean_height": [], "points": [],
           "joint_gap_max": [], "joint_gap_mean": []}
    for s in range(1, steps + 1):
        world.step(1)
        if s % cfg["every"]:
            continue
        st = world.states()
        out["step"].append(s)
        out["deepest"].append(deepest(world.manifolds()))
        out["floor_depth"].append(low.depth(st))
        out["energy"].append(energy(scene.params, st))
        out["mean_height"].append(float(st["position"][dyn, 1].astype(np.float64).mean()))
        out["points"].append(int(world.stats()["points"]))
        if scene.joints is not None:
            g = joint_gaps(scene.joints, st)
            out["joint_gap_max"].append(float(g.max()))
            out["joint_gap_mean"].append(float(g.mean()))
    return out
