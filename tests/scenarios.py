# SPDX-License-Identifier: MIT
"""Seeded scene builders shared by the oracle and GPU tests and the bench.

Meshes follow the reference's builders: the box of
conformance/v1/src/command_runner.cpp:230-242 and the cube-projected
sphere of examples/support/scene.cpp:599-632.  Kernel-level fixtures follow
tests/solver/rigid_realtime_cuda_tests.cu:43-155 and :172-211.
"""
import numpy as np

from pmamd import layouts as L

DT = np.float32(1.0) / np.float32(60.0)
BOX_INDICES = np.array([0, 1, 2, 0, 2, 3, 4, 6, 5, 4, 7, 6, 0, 4, 5, 0, 5, 1,
                        1, 5, 6, 1, 6, 2, 2, 6, 7, 2, 7, 3, 3, 7, 4, 3, 4, 0],
                       dtype=np.uint32)


def box_mesh(hx, hy, hz):
    v = np.array([[-hx, -hy, -hz], [hx, -hy, -hz], [hx, hy, -hz], [-hx, hy, -hz],
                  [-hx, -hy, hz], [hx, -hy, hz], [hx, hy, hz], [-hx, hy, hz]],
                 dtype=np.float32)
    return v, BOX_INDICES.copy()


def cube_sphere_mesh(radius):
    cube = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                     [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
                     [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1],
                     [-1, 0, 0], [1, 0, 0]], dtype=np.float32)
    norm = np.sqrt((cube * cube).sum(axis=1, dtype=np.float32)).astype(np.float32)
    v = (cube / norm[:, None] * np.float32(radius)).astype(np.float32)
    idx = []
    for center, corners in ((8, (0, 1, 5, 4)), (9, (3, 7, 6, 2)), (10, (0, 3, 2, 1)),
                            (11, (4, 5, 6, 7)), (12, (0, 4, 7, 3)), (13, (1, 2, 6, 5))):
        for k in range(4):
            idx += [center, corners[k], corners[(k + 1) % 4]]
    return v, np.array(idx, dtype=np.uint32)


def params(n, motion=L.MOTION_DYNAMIC, mass=1.0, inertia=(1.0, 1.0, 1.0),
           friction=0.5, restitution=0.0, margin=0.01, shape=0,
           max_linear=1000.0, max_angular=1000.0, damping=0.0):
    p = np.zeros(n, dtype=L.PARAMS)
    p["motion"] = motion
    p["mesh_index"] = shape
    p["mesh_generation"] = 1
    dyn = np.asarray(motion) == L.MOTION_DYNAMIC
    mass = np.broadcast_to(np.asarray(mass, dtype=np.float32), (n,))
    inertia = np.broadcast_to(np.asarray(inertia, dtype=np.float32), (n, 3))
    p["inverse_mass"] = np.where(dyn, np.float32(1.0) / mass, 0.0)
    p["inverse_inertia_local"] = np.where(np.broadcast_to(dyn, (n,))[:, None],
                                          np.float32(1.0) / inertia, 0.0)
    p["friction"] = friction
    p["restitution"] = restitution
    p["collision_margin"] = margin
    p["linear_damping"] = damping
    p["angular_damping"] = damping
    p["maximum_linear_speed"] = max_linear
    p["maximum_angular_speed"] = max_angular
    return p


def states(n, positions=None):
    s = np.zeros(n, dtype=L.STATE)
    s["orientation"][:, 3] = 1.0
    if positions is not None:
        s["position"] = positions
    return s


class Scene:
    """Shapes (("mesh", v, i) | ("sphere", r) | ("capsule", r, h)), body
    parameters and initial states."""

    def __init__(self, shapes, params, states, name="", joints=None):
        self.shapes, self.params, self.states, self.name = shapes, params, states, name
        self.joints = joints

    @property
    def n(self):
        return len(self.params)

    def build(self, world):
        for shape in self.shapes:
            if shape[0] == "mesh":
                world.add_mesh(shape[1], shape[2])
            elif shape[0] == "sphere":
                world.add_sphere(shape[1])
            else:
                world.add_capsule(shape[1], shape[2])
        world.add_bodies(self.params, self.states)
        if self.joints is not None:
            world.add_joints(self.joints)
        return world


def box_inertia(mass, half):
    hx, hy, hz = (np.float32(h) for h in half)
    m = np.float32(mass)
    return (m * (hy * hy + hz * hz) / np.float32(3), m * (hx * hx + hz * hz) / np.float32(3),
            m * (hx * hx + hy * hy) / np.float32(3))


# The reference's default rigid-body options (RigidBodyOptions,
# src/simulation/types.hpp:643-660): collision margin 5 mm, linear and
# angular damping 0.05, speed limits 100 m/s and 100 rad/s.  Scenes that
# state no other value use these.
REF_MARGIN, REF_DAMPING, REF_MAX_SPEED = 0.005, 0.05, 100.0


def pyramid(rows=10, seed=1, jitter=1e-3, margin=REF_MARGIN, friction=0.5, damping=REF_DAMPING,
            max_speed=REF_MAX_SPEED):
    """Pyramid of 55 unit boxes (10 rows), density 1000, friction 0.5 on a
    static ground box (half 50, 0.5, 50) at y = -0.5, +-1 mm seeded
    position jitter, reference default body options.  Body 0 is the
    ground."""
    rng = np.random.default_rng(seed)
    shapes = [("mesh", *box_mesh(50.0, 0.5, 50.0)), ("mesh", *box_mesh(0.5, 0.5, 0.5))]
    pos = [(0.0, -0.5, 0.0)]
    for r in range(rows):
        count = rows - r
        for i in range(count):
            pos.append(((i - (count - 1) / 2.0), 0.5 + r, 0.0))
    pos = np.array(pos, dtype=np.float32)
    pos[1:] += rng.uniform(-jitter, jitter, size=(len(pos) - 1, 3)).astype(np.float32)
    n = len(pos)
    motion = np.full(n, L.MOTION_DYNAMIC, dtype=np.uint32)
    motion[0] = L.MOTION_STATIC
    p = params(n, motion=motion, mass=1000.0, inertia=box_inertia(1000.0, (0.5, 0.5, 0.5)),
               friction=friction, margin=margin, shape=np.where(motion == 0, 0, 1), damping=damping,
               max_linear=max_speed, max_angular=max_speed)
    return Scene(shapes, p, states(n, pos), f"pyramid{rows}")


def pile(side=32, seed=2, spacing=0.8, jitter=0.05, analytic_spheres=True,
         margin=REF_MARGIN, height=0.6, damping=REF_DAMPING, max_speed=REF_MAX_SPEED):
    """Pile of side**3 bodies: half boxes (edges U[0.2, 0.6] m) and half
    spheres (r U[0.1, 0.3] m) on a jittered grid of `spacing`, density
    log-uniform [200, 5000] kg/m^3, friction U[0.3, 0.8], reference default
    body options, above a static floor and inside four static walls.
    Bodies 0-4 are the bin (floor first).  Spheres are analytic or the
    reference's polyhedral mesh."""
    rng = np.random.default_rng(seed)
    count = side ** 3
    extent = side * spacing
    half_bin = extent / 2.0 + 0.4
    wall_h = extent + 2.0
    shapes = [("mesh", *box_mesh(half_bin + 0.5, 0.5, half_bin + 0.5)),
              ("mesh", *box_mesh(0.5, wall_h / 2, half_bin + 0.5)),
              ("mesh", *box_mesh(half_bin + 0.5, wall_h / 2, 0.5))]
    bin_pos = [(0.0, -0.5, 0.0), (-half_bin - 0.5, wall_h / 2, 0.0),
               (half_bin + 0.5, wall_h / 2, 0.0), (0.0, wall_h / 2, -half_bin - 0.5),
               (0.0, wall_h / 2, half_bin + 0.5)]
    bin_shape = [0, 1, 1, 2, 2]
    g = np.stack(np.meshgrid(np.arange(side), np.arange(side), np.arange(side),
                             indexing="ij"), -1).reshape(-1, 3)[:, ::-1]  # x fastest
    pos = ((g - (side - 1) / 2.0) * spacing).astype(np.float32)
    pos[:, 1] += np.float32(extent / 2.0 + height)
    pos += rng.uniform(-jitter, jitter, size=pos.shape).astype(np.float32)
    is_box = rng.permutation(np.arange(count) % 2 == 0)
    density = np.exp(rng.uniform(np.log(200.0), np.log(5000.0), size=count)).astype(np.float32)
    friction = rng.uniform(0.3, 0.8, size=count).astype(np.float32)
    edges = rng.uniform(0.2, 0.6, size=(count, 3)).astype(np.float32)
    radii = rng.uniform(0.1, 0.3, size=count).astype(np.float32)
    body_shape = np.zeros(count, dtype=np.uint32)
    mass = np.zeros(count, dtype=np.float32)
    inertia = np.zeros((count, 3), dtype=np.float32)
    sphere_mesh_cache = {}
    for i in range(count):
        if is_box[i]:
            h = edges[i] / np.float32(2)
            shapes.append(("mesh", *box_mesh(*h)))
            m = density[i] * edges[i].prod(dtype=np.float32)
            mass[i] = m
            inertia[i] = box_inertia(m, h)
        else:
            r = radii[i]
            if analytic_spheres:
                shapes.append(("sphere", float(r)))
            else:
                shapes.append(("mesh", *cube_sphere_mesh(r)))
            m = density[i] * np.float32(4.0 / 3.0 * np.pi) * r * r * r
            mass[i] = m
            inertia[i] = np.float32(0.4) * m * r * r
        body_shape[i] = len(shapes) - 1
    n = count + 5
    motion = np.full(n, L.MOTION_DYNAMIC, dtype=np.uint32)
    motion[:5] = L.MOTION_STATIC
    all_mass = np.concatenate([np.ones(5, np.float32), mass])
    all_inertia = np.concatenate([np.ones((5, 3), np.float32), inertia])
    all_mu = np.concatenate([np.full(5, 0.5, np.float32), friction])
    p = params(n, motion=motion, mass=all_mass, inertia=all_inertia, friction=all_mu,
               margin=margin, shape=np.concatenate([np.array(bin_shape, np.uint32), body_shape]),
               damping=damping, max_linear=max_speed, max_angular=max_speed)
    s = states(n, np.concatenate([np.array(bin_pos, np.float32), pos]))
    return Scene(shapes, p, s, f"pile{side}")


# ---------------------------------------------------------------------------
# Kernel-level fixtures of tests/solver/rigid_realtime_cuda_tests.cu
# ---------------------------------------------------------------------------
def realtime_fixture():
    """:43-155 without the fixed joint (joints are not on this build's
    path): static floor 0; resting penetration 1; cold impact 2; 100:1 mass
    ratio 3-4; free body 5; fast rotation 6; teleport 7."""
    n = 8
    p = np.zeros(n, dtype=L.PARAMS)
    prev = states(n)
    for i in range(n):
        p[i]["motion"] = L.MOTION_STATIC if i == 0 else L.MOTION_DYNAMIC
        p[i]["inverse_mass"] = 0.0 if i == 0 else 1.0
        p[i]["inverse_inertia_local"] = (0, 0, 0) if i == 0 else (1, 1, 1)
        p[i]["friction"] = 0.5
        p[i]["restitution"] = 0.5
        p[i]["collision_margin"] = 0.001
    cur = prev.copy()
    m = np.zeros(3, dtype=L.MANIFOLD)
    pairs = np.zeros(3, dtype=np.uint32)
    prev[1]["position"][1] = 0.04
    cur[1] = prev[1]
    cur[1]["position"][1] = 0.0395
    cur[1]["linear_velocity"][1] = -0.03
    pairs[0] = 1 * n
    m[0]["count"] = 1
    m[0]["contacts"][0]["normal"] = (0, 1, 0)
    m[0]["contacts"][0]["penetration"] = 0.01
    prev[2]["position"][1] = 0.1
    prev[2]["linear_velocity"][1] = -3.0
    cur[2] = prev[2]
    cur[2]["position"][1] = 0.05
    pairs[1] = 2 * n
    m[1]["count"] = 1
    m[1]["contacts"][0]["normal"] = (0, 1, 0)
    m[1]["contacts"][0]["penetration"] = 0.002
    p[4]["inverse_mass"] = 0.01
    p[4]["inverse_inertia_local"] = (0.01, 0.01, 0.01)
    prev[3]["position"] = (-0.002, 1.0, 0.0)
    prev[4]["position"] = (0.002, 1.0, 0.0)
    cur[3], cur[4] = prev[3], prev[4]
    pairs[2] = 3 * n + 4
    m[2]["count"] = 1
    m[2]["contacts"][0]["normal"] = (-1, 0, 0)
    m[2]["contacts"][0]["point"] = (0, 1, 0)
    m[2]["contacts"][0]["penetration"] = 0.01
    cur[5] = prev[5]
    cur[5]["position"][0] = 0.02
    half = np.float32(0.05)
    cur[5]["orientation"] = (0, 0, np.sin(half), np.cos(half))
    half = np.float32(0.25)
    prev[6]["angular_velocity"][2] = 30.0
    cur[6] = prev[6]
    cur[6]["orientation"] = (0, 0, np.sin(half), np.cos(half))
    prev[7]["position"][0] = -5.0
    prev[7]["linear_velocity"] = (3, 2, 1)
    prev[7]["angular_velocity"] = (1, 2, 3)
    cur[7] = prev[7]
    cur[7]["position"][0] = 5.0
    cur[7]["linear_velocity"] = 0
    cur[7]["angular_velocity"] = 0
    return p, prev, cur, m, pairs


def check_realtime_fixture(cur, start, schedule):
    """Gates of rigid_realtime_cuda_tests.cu:424-477 (joint gates omitted).
    `start` is the state array before the solve."""
    assert np.all(np.isfinite(cur.view(np.float32)))
    assert schedule["offsets"][0] == 1 and schedule["budget"] == 8
    assert schedule["counts"][6] == 8 and schedule["counts"][7] == 0
    assert schedule["maximum_passes"] == 8
    correction = cur[1]["position"][1] - start[1]["position"][1]
    assert max(0.0, 0.01 - correction) <= 0.005
    assert abs(correction) <= 0.01
    assert cur[2]["linear_velocity"][1] >= -0.05
    impact_energy = 0.5 * 3.0 * 3.0
    assert 0.5 * cur[2]["linear_velocity"][1] ** 2 <= 1.05 * impact_energy
    rel = ((cur[3]["position"][0] - start[3]["position"][0]) -
           (cur[4]["position"][0] - start[4]["position"][0]))
    assert max(0.0, 0.01 + rel) <= 0.005
    assert abs(cur[6]["angular_velocity"][2]) <= 30.0
    assert abs(cur[7]["position"][0] - 5.0) < 1e-6
    assert np.all(cur[7]["linear_velocity"] == 0) and np.all(cur[7]["angular_velocity"] == 0)


def stack_window(prepare_solve, ticks_cold=120, ticks_settled=600):
    """:172-371: a two-body stack on a static floor for 720 ticks with
    externally integrated gravity and hand-built manifolds.  Returns the
    measured metrics; `prepare_solve(params, previous, states, manifolds,
    pairs, bodies)` runs one AVBD step in place and returns the schedule."""
    n = 3
    p = np.zeros(n, dtype=L.PARAMS)
    s = states(n)
    for i in range(n):
        p[i]["motion"] = L.MOTION_STATIC if i == 0 else L.MOTION_DYNAMIC
        p[i]["inverse_mass"] = 0.0 if i == 0 else 1.0
        p[i]["inverse_inertia_local"] = (0, 0, 0) if i == 0 else (1, 1, 1)
        p[i]["collision_margin"] = 0.001
    s[1]["position"][1] = 0.05
    s[2]["position"][1] = 0.149
    m = np.zeros(2, dtype=L.MANIFOLD)
    pairs = np.array([n, 2 * n + 1], dtype=np.uint32)
    bodies = np.zeros(n, dtype=L.AVBD_BODY)
    half, g, dt = np.float32(0.05), np.float32(-9.81), DT
    out = dict(max_pen=0.0, drift=0.0, rot=0.0, energy_growth=0.0, tunnel=0)
    y1 = [None, None]
    y2 = [None, None]
    settled = peak = None

    def set_contact(k, pen, point):
        m[k]["count"] = 1
        c = m[k]["contacts"][0]
        c["normal"] = (0, 1, 0)
        c["point"] = point
        c["penetration"] = max(np.float32(0), pen)

    for tick in range(ticks_cold + ticks_settled):
        prev = s.copy()
        for b in (1, 2):
            s[b]["linear_velocity"][1] = np.float32(s[b]["linear_velocity"][1] + g * dt)
            s[b]["position"][1] = np.float32(s[b]["position"][1] + s[b]["linear_velocity"][1] * dt)
        set_contact(0, half - s[1]["position"][1], (0, 0, 0))
        y_mid = np.float32(0.5) * (s[1]["position"][1] + half + s[2]["position"][1] - half)
        set_contact(1, s[1]["position"][1] + np.float32(2) * half - s[2]["position"][1], (0, y_mid, 0))
        sched = prepare_solve(p, prev, s, m, pairs, bodies)
        assert sched["budget"] == 8 and sched["counts"][6] == 8 and sched["counts"][7] == 0
        m["contacts"][:, 0]["avbd_cached"] = 1
        pen = max(0.0, half - s[1]["position"][1],
                  s[1]["position"][1] + 2 * half - s[2]["position"][1])
        out["max_pen"] = max(out["max_pen"], float(pen))
        for b in (1, 2):
            q = s[b]["orientation"]
            ang = 2.0 * np.arctan2(np.sqrt(q[0] ** 2 + q[1] ** 2 + q[2] ** 2), abs(q[3]))
            out["rot"] = max(out["rot"], float(ang))
        if s[1]["position"][1] < -half or s[2]["position"][1] <= s[1]["position"][1]:
            out["tunnel"] += 1
        if tick >= ticks_cold:
            e = 0.5 * (s[1]["linear_velocity"][1] ** 2 + s[2]["linear_velocity"][1] ** 2) - \
                g * (s[1]["position"][1] + s[2]["position"][1])
            a, b_ = float(s[1]["position"][1]), float(s[2]["position"][1])
            if tick == ticks_cold:
                y1 = [a, a]
                y2 = [b_, b_]
                settled = peak = float(e)
            else:
                y1 = [min(y1[0], a), max(y1[1], a)]
                y2 = [min(y2[0], b_), max(y2[1], b_)]
                peak = max(peak, float(e))
    out["drift"] = max(y1[1] - y1[0], y2[1] - y2[0])
    out["energy_growth"] = peak / settled - 1.0 if settled > 0 else 0.0
    return out


def check_stack_window(out):
    """Gates of rigid_realtime_cuda_tests.cu:365-370."""
    assert out["max_pen"] <= 0.005
    assert out["drift"] <= 0.01
    assert out["rot"] <= 0.035
    assert out["energy_growth"] <= 0.05
    assert out["tunnel"] == 0


def l_prism(flip=False):
    """Closed concave L-shaped prism: cross-section (0,0) (1,0) (2,0) (2,1)
    (1,1) (1,2) (0,2) in xy, extruded over z in [-1, 1], outward wound
    (inward when `flip`)."""
    poly = np.array([[0, 0], [1, 0], [2, 0], [2, 1], [1, 1], [1, 2], [0, 2]], np.float32)
    m = len(poly)
    v = np.zeros((2 * m, 3), np.float32)
    v[:m, :2] = poly
    v[:m, 2] = 1.0
    v[m:, :2] = poly
    v[m:, 2] = -1.0
    cap = [(0, 1, 4), (0, 4, 6), (4, 5, 6), (1, 2, 3), (1, 3, 4)]
    idx = []
    for t in cap:
        idx += list(t)                                   # front, counter-clockwise from +z
        idx += [m + t[0], m + t[2], m + t[1]]            # back
    for i in range(m):
        j = (i + 1) % m
        idx += [i, m + i, m + j, i, m + j, j]
    idx = np.array(idx, np.uint32).reshape(-1, 3)
    if flip:
        idx = idx[:, ::-1]
    return v, np.ascontiguousarray(idx.reshape(-1))


def quat(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    s = np.sin(angle / 2)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, np.cos(angle / 2)], np.float32)


def concave_drop(count=12, seed=5, speed=8.0):
    """Boxes (half 0.1-0.25 m) at random orientations thrown down at up to
    `speed` m/s onto a static concave L-shaped prism (l_prism) and into its
    inner corner: every pair takes the triangle-pair path, fast ones its
    swept pass.  Body 0 is the prism, body 1 a static floor box under it
    (top face at y = 0)."""
    rng = np.random.default_rng(seed)
    shapes = [("mesh", *l_prism()), ("mesh", *box_mesh(6.0, 0.5, 6.0))]
    halves = rng.uniform(0.1, 0.25, size=(count, 3)).astype(np.float32)
    for h in halves:
        shapes.append(("mesh", *box_mesh(*h)))
    n = count + 2
    motion = np.full(n, L.MOTION_DYNAMIC, dtype=np.uint32)
    motion[:2] = L.MOTION_STATIC
    mass = np.ones(n, np.float32)
    inertia = np.ones((n, 3), np.float32)
    for k, h in enumerate(halves):
        mass[k + 2] = 8.0 * h[0] * h[1] * h[2] * 1000.0
        inertia[k + 2] = box_inertia(mass[k + 2], h)
    p = params(n, motion=motion, mass=mass, inertia=inertia, margin=0.01,
               shape=np.arange(n))
    pos = np.zeros((n, 3), np.float32)
    pos[1] = (0.0, -0.5, 0.0)
    pos[2:, 0] = rng.uniform(0.2, 1.8, size=count)
    pos[2:, 1] = rng.uniform(2.4, 3.6, size=count)
    pos[2:, 2] = rng.uniform(-0.8, 0.8, size=count)
    st = states(n, pos)
    for k in range(2, n):
        st["orientation"][k] = quat(rng.normal(size=3), rng.uniform(0, np.pi))
    st["linear_velocity"][2:, 1] = -rng.uniform(0, speed, size=count).astype(np.float32)
    st["angular_velocity"][2:] = rng.uniform(-3, 3, size=(count, 3)).astype(np.float32)
    return Scene(shapes, p, st, "concave_drop")


def crossing_edges(gap=5e-4, half=0.5, far=0):
    """Two boxes meeting edge on edge (the convex face path declines this
    pair): the static one turned 45 degrees about z, the dynamic one 45
    degrees about x and `gap` above it.  `far` more static boxes stand 50 m
    away along x, out of reach of both (they only raise the body count)."""
    h = np.float32(half)
    top = h * np.float32(np.sqrt(2.0))
    shapes = [("mesh", *box_mesh(h, h, h))]
    n = 2 + far
    motion = np.full(n, L.MOTION_STATIC, np.uint32)
    motion[1] = L.MOTION_DYNAMIC
    p = params(n, motion=motion, mass=1000.0, inertia=box_inertia(1000.0, (h, h, h)), shape=np.zeros(n, np.int64))
    pos = [[0, 0, 0], [0, 2 * top + np.float32(gap), 0]] + [[50.0 * (k + 1), 0, 0] for k in range(far)]
    st = states(n, np.array(pos, np.float32))
    st["orientation"][0] = quat((0, 0, 1), np.pi / 4)
    st["orientation"][1] = quat((1, 0, 0), np.pi / 4)
    return Scene(shapes, p, st, "crossing_edges")


def pusher(rows=5, seed=1):
    """A pyramid of unit boxes (rows) with a kinematic unit box beside it
    (last body) that kinematic targets drive into the stack."""
    base = pyramid(rows=rows, seed=seed)
    n = base.n + 1
    p = np.concatenate([base.params, params(1, motion=np.array([L.MOTION_KINEMATIC], np.uint32), shape=1,
                                            margin=REF_MARGIN, damping=REF_DAMPING, max_linear=REF_MAX_SPEED,
                                            max_angular=REF_MAX_SPEED)])
    st = np.concatenate([base.states, states(1, np.array([[-rows / 2.0 - 1.5, 0.5, 0.0]], np.float32))])
    return Scene(base.shapes, p, st, "pusher")


def joint_rig():
    """Every joint type in one scene over a static floor: point and hinge
    pendulums, a four-link chain, a welded pair under a falling cube, a
    slider with a travel limit, a piston, a sprung cube, a motor and a
    joint that breaks under its load.  Jointed pairs do not collide
    (disable_collisions); the weld also joins a weld group."""
    shapes = [("mesh",) + box_mesh(0.1, 0.1, 0.1), ("mesh",) + box_mesh(0.1, 0.5, 0.1),
              ("mesh",) + box_mesh(8.0, 0.5, 8.0), ("mesh",) + box_mesh(0.25, 0.25, 0.25)]
    motion, shape, mass, inertia, pos, rot = [], [], [], [], [], []

    def body(m, sh, ms, p, q=(0, 0, 0, 1)):
        motion.append(m)
        shape.append(sh)
        mass.append(ms)
        half = {0: (0.1, 0.1, 0.1), 1: (0.1, 0.5, 0.1), 2: (8.0, 0.5, 8.0), 3: (0.25, 0.25, 0.25)}[sh]
        inertia.append(box_inertia(ms, half))
        pos.append(p)
        rot.append(q)
        return len(pos) - 1

    S_, D_ = L.MOTION_STATIC, L.MOTION_DYNAMIC
    floor = body(S_, 2, 1.0, (0, -0.5, 0))
    anchor = body(S_, 0, 1.0, (0, 4.0, 0))
    joints = []

    def joint(kind, a, b, aa, ab, **kw):
        o = L.joint_options(1, kind, a, b, anchor_a=aa, anchor_b=ab)
        for k, v in kw.items():
            o[k] = v
        joints.append(o)
        return o

    def hanging_rod(x, angle, m=5.0):
        c = (x + np.sin(angle) * 0.5, 4.0 - np.cos(angle) * 0.5, 0.0)
        return body(D_, 1, m, c, quat((0, 0, 1), angle))

    # pendulums
    r = hanging_rod(-3.0, 0.6)
    joint(L.JOINT_POINT, anchor, r, (-3.0, 0, 0), (0, 0.5, 0))
    r = hanging_rod(-2.0, 0.4)
    joint(L.JOINT_HINGE, anchor, r, (-2.0, 0, 0), (0, 0.5, 0))
    # a four-link chain
    prev, top = anchor, (0.0, 0.0, 0.0)
    for k in range(4):
        ang = 0.3
        cx = 1.0 + k * np.sin(ang) * 1.0 + np.sin(ang) * 0.5
        cy = 4.0 - k * np.cos(ang) * 1.0 - np.cos(ang) * 0.5
        link = body(D_, 1, 2.0, (cx, cy, 0.0), quat((0, 0, 1), ang))
        joint(L.JOINT_POINT, prev, link, (1.0, 0, 0) if k == 0 else (0, -0.5, 0), (0, 0.5, 0))
        prev = link
    # a welded pair resting on the floor, a cube dropping onto it
    wa = body(D_, 3, 2.0, (-1.0, 0.25, 2.0))
    wb = body(D_, 3, 2.0, (-0.5, 0.25, 2.0))
    joint(L.JOINT_FIXED, wa, wb, (0.25, 0, 0), (-0.25, 0, 0))
    body(D_, 3, 1.0, (-0.75, 1.0, 2.0), quat((1, 0, 1), 0.3))
    # a slider along the floor's x axis with a travel limit
    sl = body(D_, 3, 1.0, (2.0, 0.3, 2.0))
    lim = np.zeros((), L.JOINT_OPTIONS)["linear_limits"]
    lim["axes"] = 1
    lim["lower"] = (-0.3, 0, 0)
    lim["upper"] = (0.3, 0, 0)
    joint(L.JOINT_SLIDER, floor, sl, (2.0, 0.8, 2.0), (0, 0, 0), linear_limits=lim)
    # a piston hanging from the anchor: free travel and twist about x
    pi = hanging_rod(3.0, 0.0, m=3.0)
    joint(L.JOINT_PISTON, anchor, pi, (3.0, 0, 0), (0, 0.5, 0))
    # a sprung cube under the anchor
    sp = body(D_, 0, 1.0, (4.0, 3.0, 0.0))
    spr = np.zeros((), L.JOINT_OPTIONS)["linear_springs"]
    spr["axes"] = 7
    spr["stiffness"] = (200.0, 200.0, 200.0)
    spr["damping"] = (2.0, 2.0, 2.0)
    joint(L.JOINT_GENERIC_SPRING, anchor, sp, (4.0, -0.8, 0), (0, 0, 0), linear_springs=spr)
    # a motor turning a rod about its x axis
    mo = hanging_rod(5.0, 0.0, m=2.0)
    mot = np.zeros((), L.JOINT_OPTIONS)["motor"]
    mot["angular_enabled"] = 1
    mot["angular_target_velocity"] = 2.0
    mot["angular_maximum_impulse"] = 5.0
    joint(L.JOINT_MOTOR, anchor, mo, (5.0, 0, 0), (0, 0.5, 0), motor=mot)
    # an overloaded joint: a heavy cube that tears loose
    hv = body(D_, 3, 50.0, (-4.5, 3.4, 0.0))
    joint(L.JOINT_POINT, anchor, hv, (-4.5, -0.35, 0), (0, 0.25, 0), breaking_impulse_threshold=2.0)
    n = len(pos)
    p = params(n, motion=np.array(motion, np.uint32), mass=np.array(mass, np.float32),
               inertia=np.array(inertia, np.float32), shape=np.array(shape), margin=REF_MARGIN,
               damping=REF_DAMPING, max_linear=REF_MAX_SPEED, max_angular=REF_MAX_SPEED)
    st = states(n, np.array(pos, np.float32))
    st["orientation"] = np.array(rot, np.float32)
    st["linear_velocity"][sl] = (2.0, 0.0, 0.0)     # runs into the slider's limit
    return Scene(shapes, p, st, name="joint_rig", joints=np.concatenate(joints))


def hinge_contacts(planted=False):
    """Joints that shape contacts (rigid_hinge_contact_frame,
    geometry.slang:404-491), over a static floor:
    - a slider held by a static post, sent along its axis into a static
      wall at 3 m/s (an axially guided member against a static body: the
      triangle path's guided swept pass, 0.1 mm rest offset);
    - a piston held the same way, spinning about its axis, into a second
      wall (the guide also turns);
    - a box hinged to a static post along its lower left edge, lying on
      the floor (a hinged member: its contacts do not persist);
    - two welded boxes on the lower step of a concave L prism (fixed
      members on the triangle path)."""
    prism_v, prism_i = l_prism()
    shapes = [("mesh",) + box_mesh(0.1, 0.1, 0.1), ("mesh",) + box_mesh(0.25, 0.25, 0.25),
              ("mesh",) + box_mesh(8.0, 0.5, 8.0), ("mesh",) + box_mesh(0.1, 0.5, 0.5),
              ("mesh", prism_v, prism_i), ("mesh",) + box_mesh(0.2, 0.2, 0.2)]
    motion, shape, mass, inertia, pos = [], [], [], [], []
    halves = {0: (0.1, 0.1, 0.1), 1: (0.25, 0.25, 0.25), 2: (8.0, 0.5, 8.0), 3: (0.1, 0.5, 0.5),
              4: (1.0, 1.0, 1.0), 5: (0.2, 0.2, 0.2)}

    def body(m, sh, ms, p):
        motion.append(m)
        shape.append(sh)
        mass.append(ms)
        inertia.append(box_inertia(ms, halves[sh]))
        pos.append(p)
        return len(pos) - 1

    S_, D_ = L.MOTION_STATIC, L.MOTION_DYNAMIC
    body(S_, 2, 1.0, (0, -0.5, 0))                     # floor
    joints = []
    # slider and piston on static posts, each running into a static wall
    post_s = body(S_, 0, 1.0, (0.0, 3.0, 0.0))
    slider = body(D_, 1, 2.0, (0.0, 1.0, 0.0))
    body(S_, 3, 1.0, (1.0, 1.0, 0.0))
    joints.append(L.joint_options(1, L.JOINT_SLIDER, post_s, slider, (0.0, -2.0, 0.0), (0, 0, 0)))
    post_p = body(S_, 0, 1.0, (0.0, 3.0, 2.0))
    piston = body(D_, 1, 2.0, (0.0, 1.0, 2.0))
    body(S_, 3, 1.0, (1.0, 1.0, 2.0))
    joints.append(L.joint_options(1, L.JOINT_PISTON, post_p, piston, (0.0, -2.0, 0.0), (0, 0, 0)))
    # a box hinged along its lower left edge (hinge axis z) to a static post
    post_h = body(S_, 0, 1.0, (-2.5, 0.1, 0.0))
    lid = body(D_, 1, 1.0, (-2.0, 0.2505, 0.0))
    joints.append(L.joint_options(1, L.JOINT_HINGE, post_h, lid, (0.25, -0.1, 0.0), (-0.25, -0.25, 0.0)))
    # two welded boxes on the prism's lower step (its top at y = 1 over x in [1, 2])
    prism = body(S_, 4, 1.0, (3.0, 0.0, -3.0))
    wa = body(D_, 5, 1.0, (4.5, 1.2005, -3.4))
    wb = body(D_, 5, 1.0, (4.5, 1.2005, -2.6))
    joints.append(L.joint_options(1, L.JOINT_FIXED, wa, wb, (0.0, 0.0, 0.4), (0.0, 0.0, -0.4)))
    n = len(pos)
    p = params(n, motion=np.array(motion, np.uint32), mass=np.array(mass, np.float32),
               inertia=np.array(inertia, np.float32), shape=np.array(shape), margin=REF_MARGIN,
               damping=REF_DAMPING, max_linear=REF_MAX_SPEED, max_angular=REF_MAX_SPEED)
    st = states(n, np.array(pos, np.float32))
    st["linear_velocity"][slider] = (3.0, 0.0, 0.0)
    st["linear_velocity"][piston] = (3.0, 0.0, 0.0)
    st["angular_velocity"][piston] = (2.0, 0.0, 0.0)
    return Scene(shapes, p, st, name="hinge_contacts", joints=np.concatenate(joints))


def capsule_inertia(mass, radius, half_length):
    """Solid capsule along local y (cylinder of length 2h plus two
    hemispheres), principal moments (x, y, z)."""
    r, h = float(radius), float(half_length)
    vc = np.pi * r * r * 2.0 * h
    vs = 4.0 / 3.0 * np.pi * r ** 3
    mc = mass * vc / (vc + vs)
    ms = mass - mc
    iy = mc * r * r / 2.0 + ms * 2.0 * r * r / 5.0
    ix = mc * (3.0 * r * r + 4.0 * h * h) / 12.0 + ms * (2.0 * r * r / 5.0 + h * h + 3.0 * h * r / 4.0)
    return (np.float32(ix), np.float32(iy), np.float32(ix))


def chains(count=4, links=64, spacing=0.45, height=None, seed=4, radius=0.05, length=0.3,
           heavy_every=8, heavy_ratio=100.0, density=1000.0, max_tilt=0.35, hinges_only=False):
    """`count` chains (a square grid of anchors, `spacing` apart) of
    `links` capsule links (radius 0.05 m, total length 0.3 m) joined end to
    end by ball (point) and hinge joints in turn, the first link held by a
    ball joint to a static anchor body high above; every `heavy_every`-th
    link is `heavy_ratio` times heavier.  Each chain starts straight but
    tilted by up to `max_tilt` rad in a seeded direction, so the chains
    swing into each other and, hanging down, reach a static ground box
    (the anchors sit at 0.99 of a chain's length).  Reference
    default body options.  Bodies: 0 ground, 1 anchor, then the links chain
    by chain; joint j of chain c joins link j-1 (or the anchor) and link j.
    With `hinges_only` every joint is a hinge."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(count)))
    if height is None:
        height = 0.99 * links * length              # a chain swung down reaches the ground
    h = length / 2.0 - radius                       # capsule half segment
    half_link = length / 2.0
    extent = side * spacing
    shapes = [("mesh",) + box_mesh(extent / 2.0 + 25.0, 0.5, extent / 2.0 + 25.0),
              ("mesh",) + box_mesh(0.05, 0.05, 0.05), ("capsule", radius, h)]
    anchor_pos = np.array([0.0, height + 5.0, 0.0], np.float32)
    n = 2 + count * links
    pos = np.zeros((n, 3), np.float64)
    rot = np.zeros((n, 4), np.float32)
    rot[:, 3] = 1.0
    pos[0] = (0.0, -0.5, 0.0)
    pos[1] = anchor_pos
    motion = np.full(n, L.MOTION_DYNAMIC, np.uint32)
    motion[:2] = L.MOTION_STATIC
    shape = np.full(n, 2)
    shape[:2] = (0, 1)
    base_mass = density * (np.pi * radius ** 2 * 2.0 * h + 4.0 / 3.0 * np.pi * radius ** 3)
    k = np.arange(links)
    link_mass = np.where((k + 1) % heavy_every == 0, base_mass * heavy_ratio, base_mass)
    mass = np.ones(n, np.float64)
    mass[2:] = np.tile(link_mass, count)
    inertia = np.ones((n, 3), np.float32)
    light, heavy = capsule_inertia(base_mass, radius, h), capsule_inertia(base_mass * heavy_ratio, radius, h)
    inertia[2:] = np.tile(np.where(((k + 1) % heavy_every == 0)[:, None], heavy, light), (count, 1))
    kinds = np.full(links, L.JOINT_HINGE) if hinges_only else np.where(k % 2 == 0, L.JOINT_POINT, L.JOINT_HINGE)
    jo = L.joint_options(count * links, np.tile(kinds, count),
                         0, 0, anchor_a=(0, -half_link, 0), anchor_b=(0, half_link, 0))
    for c in range(count):
        gx, gz = c % side, c // side
        top = np.array([(gx - (side - 1) / 2.0) * spacing, height, (gz - (side - 1) / 2.0) * spacing])
        tilt = rng.uniform(0.0, max_tilt)
        heading = rng.uniform(0.0, 2.0 * np.pi)
        q = quat(np.array([np.cos(heading), 0.0, np.sin(heading)]), tilt)   # tilts the chain's down direction
        down = _rotate_np(q, np.array([0.0, -1.0, 0.0]))
        first = 2 + c * links
        pos[first:first + links] = top[None, :] + down[None, :] * ((2.0 * k + 1.0) * half_link)[:, None]
        rot[first:first + links] = q
        j = slice(c * links, (c + 1) * links)
        jo["body_a"][j, 0] = np.concatenate([[1], first + k[:-1]])
        jo["body_b"][j, 0] = first + k
        # the first joint holds the top link to the anchor body, whose frame
        # is turned with the chain so that the rest pose is the joint's zero
        jo["local_anchor_a"][c * links] = top - anchor_pos
        jo["local_orientation_a"][c * links] = q
    p = params(n, motion=motion, mass=mass.astype(np.float32), inertia=inertia, shape=shape, margin=REF_MARGIN,
               damping=REF_DAMPING, max_linear=REF_MAX_SPEED, max_angular=REF_MAX_SPEED)
    st = states(n, pos.astype(np.float32))
    st["orientation"] = rot
    return Scene(shapes, p, st, name=f"chains{count}x{links}", joints=jo)


def _rotate_np(q, v):
    u = np.asarray(q[:3], np.float64)
    t = 2.0 * np.cross(u, v)
    return v + q[3] * t + np.cross(u, t)


class Batch:
    """Independent scenes sharing one shape palette: `scenes[s]` is a
    (params, states) pair whose shape indices point into `shapes`."""

    joints = None

    def __init__(self, shapes, scenes, name=""):
        self.shapes, self.scenes, self.name = shapes, scenes, name

    @property
    def n(self):
        return sum(len(p) for p, _ in self.scenes)

    def add_shapes(self, world):
        for shape in self.shapes:
            if shape[0] == "mesh":
                world.add_mesh(shape[1], shape[2])
            elif shape[0] == "sphere":
                world.add_sphere(shape[1])
            else:
                world.add_capsule(shape[1], shape[2])

    def build(self, world):
        """All scenes in one world (pmhWorldAddScene each)."""
        self.add_shapes(world)
        for p, st in self.scenes:
            world.add_scene(p, st)
        return world

    def build_one(self, world, s):
        """Scene `s` alone, with the same shape palette."""
        self.add_shapes(world)
        p, st = self.scenes[s]
        world.add_bodies(p, st)
        return world


def batch(count=4096, seed=7, min_bodies=8, max_bodies=512, palette=64, region=40.0):
    """`count` independent scenes for one batched launch.  Per scene: a body
    count log-uniform in [min_bodies, max_bodies]; half boxes (edges
    U[0.2, 0.6] m, drawn from a palette of `palette` sizes) and half analytic
    spheres (r U[0.1, 0.3] m, same palette rule), density log-uniform
    [200, 5000] kg/m^3, friction U[0.3, 0.8], reference default body
    options; either a drop (a jittered grid 0.8 m apart above the ground) or
    stacks (columns of bodies resting on each other).  Each scene has its
    own static ground box and is laid out in its own `region`-metre square,
    so the batch also stays apart in space.  Body 0 of a scene is its
    ground."""
    rng = np.random.default_rng(seed)
    box_h = rng.uniform(0.1, 0.3, size=(palette, 3)).astype(np.float32)
    radii = rng.uniform(0.1, 0.3, size=palette).astype(np.float32)
    shapes = [("mesh",) + box_mesh(region * 0.45, 0.5, region * 0.45)]
    shapes += [("mesh",) + box_mesh(*h) for h in box_h]
    shapes += [("sphere", float(r)) for r in radii]
    grid = int(np.ceil(np.sqrt(count)))
    scenes = []
    for sc in range(count):
        nb = int(np.exp(rng.uniform(np.log(min_bodies), np.log(max_bodies + 1))))
        nb = max(min_bodies, min(max_bodies, nb))
        ox = (sc % grid - grid / 2.0) * region
        oz = (sc // grid - grid / 2.0) * region
        is_box = rng.permutation(np.arange(nb) % 2 == 0)
        pick = rng.integers(0, palette, size=nb)
        density = np.exp(rng.uniform(np.log(200.0), np.log(5000.0), size=nb)).astype(np.float32)
        friction = rng.uniform(0.3, 0.8, size=nb).astype(np.float32)
        shape = np.where(is_box, 1 + pick, 1 + palette + pick)
        mass = np.zeros(nb, np.float32)
        inertia = np.zeros((nb, 3), np.float32)
        size = np.zeros(nb, np.float32)                  # vertical half extent
        for i in range(nb):
            if is_box[i]:
                h = box_h[pick[i]]
                mass[i] = density[i] * np.float32(8.0) * h.prod(dtype=np.float32)
                inertia[i] = box_inertia(mass[i], h)
                size[i] = h[1]
            else:
                r = radii[pick[i]]
                mass[i] = density[i] * np.float32(4.0 / 3.0 * np.pi) * r * r * r
                inertia[i] = np.float32(0.4) * mass[i] * r * r
                size[i] = r
        pos = np.zeros((nb, 3), np.float32)
        if rng.random() < 0.5:                           # drop
            side = int(np.ceil(nb ** (1.0 / 3.0)))
            g = np.stack(np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij"),
                         -1).reshape(-1, 3)[:nb, ::-1]
            pos[:] = (g - (side - 1) / 2.0) * 0.8
            pos[:, 1] += (side - 1) * 0.4 + 1.0
            pos += rng.uniform(-0.05, 0.05, size=pos.shape).astype(np.float32)
        else:                                            # stacks of up to 8
            cols = int(np.ceil(nb / 8.0))
            cside = int(np.ceil(np.sqrt(cols)))
            height = np.zeros(cols, np.float32)
            for i in range(nb):
                c = i % cols
                pos[i] = ((c % cside - cside / 2.0) * 0.8, height[c] + size[i] + 0.002, (c // cside - cside / 2.0) * 0.8)
                height[c] += 2.0 * size[i] + 0.004
            pos[:, :2] += rng.uniform(-0.01, 0.01, size=(nb, 2)).astype(np.float32)
        pos[:, 0] += ox
        pos[:, 2] += oz
        n = nb + 1
        motion = np.full(n, L.MOTION_DYNAMIC, np.uint32)
        motion[0] = L.MOTION_STATIC
        p = params(n, motion=motion, mass=np.concatenate([[1.0], mass]).astype(np.float32),
                   inertia=np.concatenate([np.ones((1, 3), np.float32), inertia]),
                   friction=np.concatenate([[0.5], friction]).astype(np.float32),
                   shape=np.concatenate([[0], shape]), margin=REF_MARGIN, damping=REF_DAMPING,
                   max_linear=REF_MAX_SPEED, max_angular=REF_MAX_SPEED)
        st = states(n, np.concatenate([np.array([[ox, -0.5, oz]], np.float32), pos]))
        scenes.append((p, st))
    return Batch(shapes, scenes, f"batch{count}")


def joint_seam_inputs(steps=5, id_offset=1000, id_generation=3):
    """Kernel-seam inputs with joints from the joint rig stepped by the
    oracle: parameters, the states of step `steps` as the previous states
    and of step `steps + 1` as the predicted ones, that step's manifolds,
    pair codes, AvbdBody history and joint records.  Dense body i gets the id
    (i + id_offset, id_generation), and the joints name their bodies by those
    ids, so the seam has to resolve them."""
    import _oracle as O
    scene = joint_rig()
    w = scene.build(O.OracleWorld(profile=0, base_iterations=10, lanes=32))
    w.step(steps)
    previous = w.states()
    w.step(1)
    joints = w.joints()
    n = scene.n
    ids = np.zeros((n, 2), np.uint32)
    ids[:, 0] = np.arange(n) + id_offset
    ids[:, 1] = id_generation
    for side in ("body_a", "body_b"):
        dense = joints["options"][side][:, 0].copy()
        joints["options"][side][:, 0] = dense + id_offset
        joints["options"][side][:, 1] = id_generation
    return dict(params=scene.params.copy(), previous=previous, states=w.states(), manifolds=w.manifolds(),
                pairs=w.pairs(), bodies=w.bodies(), joints=joints, ids=ids)
