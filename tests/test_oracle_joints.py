# SPDX-License-Identifier: MIT
"""Joints on the oracle (rigid_avbd_visit_joint(_row),
geometry_constraints.slang:626-926; graph, budget and dual stages
:1391-1560, geometry_constraints_solve.slang:174-209, 401-430, 566-614;
collision exclusions geometry.slang:494-511, 3108-3136).

No joint vectors are held by the reference; the joint rows are pinned by
what the mechanics fixes: a pendulum keeps its anchor and swings with the
period of a physical pendulum, a hinge keeps its locked axes, a weld moves
as one body, a limit stops the free axis, an overloaded joint breaks and
its body then falls freely, and the graph rules (colours, islands, budget,
collision exclusions) follow their definitions."""
import numpy as np
import pytest

import _oracle as O
import scenarios as S
from pmamd import layouts as L


def rotate(q, v):
    u = q[:3].astype(np.float64)
    t = 2.0 * np.cross(u, v)
    return v + q[3] * t + np.cross(u, t)


def rod_world(kind, angle=0.5, mass=10.0, profile=0, gravity=(0, -9.81, 0), **joint):
    """A static anchor body at the origin and a rod (0.2 x 1.0 x 0.2 m)
    hanging from it by its top end, turned `angle` about z."""
    w = O.OracleWorld(profile=profile, base_iterations=10, gravity=gravity)
    w.add_mesh(*S.box_mesh(0.1, 0.1, 0.1))
    w.add_mesh(*S.box_mesh(0.1, 0.5, 0.1))
    p = S.params(2, motion=np.array([L.MOTION_STATIC, L.MOTION_DYNAMIC], np.uint32),
                 mass=np.array([1, mass], np.float32),
                 inertia=np.array([[1, 1, 1], S.box_inertia(mass, (0.1, 0.5, 0.1))], np.float32),
                 shape=np.array([0, 1]))
    c = np.array([np.sin(angle) * 0.5, -np.cos(angle) * 0.5, 0.0], np.float32)
    st = S.states(2, np.array([[0, 0, 0], c], np.float32))
    st["orientation"][1] = S.quat((0, 0, 1), angle)
    w.add_bodies(p, st)
    o = L.joint_options(1, kind, 0, 1, anchor_a=(0, 0, 0), anchor_b=(0, 0.5, 0))
    for k, v in joint.items():
        o[k] = v
    w.add_joints(o)
    return w


def anchor_gap(s):
    return float(np.linalg.norm(s["position"] + rotate(s["orientation"], np.array([0, 0.5, 0]))))


@pytest.mark.parametrize("kind", [L.JOINT_POINT, L.JOINT_HINGE])
def test_pendulum_keeps_its_anchor_and_period(kind):
    w = rod_world(kind)
    xs, drift = [], 0.0
    for _ in range(300):
        w.step(1)
        s = w.states()[1]
        drift = max(drift, anchor_gap(s))
        xs.append(float(s["position"][0]))
    xs = np.array(xs)
    down = np.nonzero((xs[:-1] > 0) & (xs[1:] <= 0))[0]
    period = np.diff(down).mean() / 60.0
    m = 10.0
    inertia = m * (0.1 ** 2 + 0.5 ** 2) / 3.0 + m * 0.25          # about the pivot
    t0 = 2 * np.pi * np.sqrt(inertia / (m * 9.81 * 0.5))
    assert drift < 1e-4, drift
    assert abs(period - t0) / t0 < 0.03, (period, t0)
    j = w.joints()[0]
    assert j["avbd_count"] == (3 if kind == L.JOINT_POINT else 5)
    assert j["enabled"] == 1 and j["broken"] == 0


def test_hinge_locks_the_swing_axes():
    w = rod_world(L.JOINT_HINGE, angle=0.0, gravity=(0, 0, 0))
    kick = np.zeros(1, L.RIGID_INPUT)
    kick["body_index"] = 1
    kick["angular_impulse"] = (2.0, 0.0, 2.0)
    w.apply_inputs(kick)
    for _ in range(30):
        w.step(1)
    s = w.states()[1]
    axis = rotate(s["orientation"], np.array([0.0, 0.0, 1.0]))
    assert abs(axis[2]) > 0.9999                      # the hinge axis stays along z
    assert abs(s["angular_velocity"][2]) > 0.1        # free rotation about z
    assert anchor_gap(s) < 1e-5


def test_angular_limit_stops_the_free_axis():
    o = L.joint_options(1, L.JOINT_HINGE, 0, 1, anchor_a=(0, 0, 0), anchor_b=(0, 0.5, 0))
    o["angular_limits"]["axes"] = 4
    o["angular_limits"]["lower"] = (0, 0, -0.2)
    o["angular_limits"]["upper"] = (0, 0, 0.2)
    w = rod_world(L.JOINT_HINGE, angle=0.0, gravity=(0, 0, 0), angular_limits=o["angular_limits"])
    kick = np.zeros(1, L.RIGID_INPUT)
    kick["body_index"] = 1
    kick["angular_impulse"] = (0.0, 0.0, 3.0)
    w.apply_inputs(kick)
    peak = 0.0
    for _ in range(120):
        w.step(1)
        q = w.states()[1]["orientation"]
        peak = max(peak, 2.0 * np.arctan2(abs(q[2]), q[3]))
    assert 0.19 < peak < 0.23, peak
    assert w.joints()[0]["avbd_count"] == 7          # 5 locks + upper and lower limit rows


def test_overloaded_joint_breaks_and_the_body_falls():
    w = rod_world(L.JOINT_POINT, angle=0.0, breaking_impulse_threshold=0.5)
    w.step(1)
    j = w.joints()[0]
    assert j["broken"] == 1 and j["enabled"] == 0
    assert j["applied_impulse"] > 0.5
    v0 = w.states()[1]["linear_velocity"][1]
    for _ in range(10):
        w.step(1)
    v = w.states()[1]["linear_velocity"][1]
    # free fall after the break: dv = -g dt per step, damping 0
    assert v - v0 == pytest.approx(-9.81 * 10 / 60.0, rel=1e-3)


def weld_world(disable=1, kind=L.JOINT_FIXED):
    w = O.OracleWorld(profile=0, gravity=(0, 0, 0))
    w.add_mesh(*S.box_mesh(0.5, 0.5, 0.5))
    p = S.params(2, mass=1.0, inertia=S.box_inertia(1.0, (0.5, 0.5, 0.5)), shape=0)
    w.add_bodies(p, S.states(2, np.array([[0, 0, 0], [0.9, 0, 0]], np.float32)))   # overlapping by 0.1 m
    w.add_joints(L.joint_options(1, kind, 0, 1, anchor_a=(0.45, 0, 0), anchor_b=(-0.45, 0, 0),
                                 disable_collisions=disable))
    return w


def test_weld_moves_as_one_body_without_contacts():
    w = weld_world()
    kick = np.zeros(1, L.RIGID_INPUT)
    kick["body_index"] = 0
    kick["linear_impulse"] = (0.0, 1.0, 0.0)
    kick["angular_impulse"] = (0.0, 0.0, 0.2)
    w.apply_inputs(kick)
    for _ in range(60):
        w.step(1)
        assert len(w.pairs()) == 0
    s = w.states()
    ra = rotate(s[0]["orientation"], np.array([0.45, 0, 0])) + s[0]["position"]
    rb = rotate(s[1]["orientation"], np.array([-0.45, 0, 0])) + s[1]["position"]
    assert np.linalg.norm(ra - rb) < 1e-4
    qa, qb = s[0]["orientation"], s[1]["orientation"]
    assert abs(abs(np.dot(qa, qb)) - 1.0) < 1e-6
    # momentum is shared: total linear momentum = the impulse
    assert s["linear_velocity"][:, 1].sum() == pytest.approx(1.0, rel=2e-2)


@pytest.mark.parametrize("disable", [0, 1])
def test_disable_collisions_excludes_the_pair(disable):
    w = weld_world(disable=disable, kind=L.JOINT_POINT)
    w.step(1)
    assert len(w.pairs()) == (0 if disable else 1)


def test_chain_colours_islands_and_budget():
    n = 5
    w = O.OracleWorld(profile=0, base_iterations=10, gravity=(0, 0, 0))
    w.add_mesh(*S.box_mesh(0.1, 0.1, 0.1))
    w.add_bodies(S.params(n, shape=0), S.states(n, np.array([[2.0 * k, 0, 0] for k in range(n)], np.float32)))
    o = L.joint_options(n - 1, L.JOINT_POINT, np.arange(n - 1), np.arange(1, n), anchor_a=(1.0, 0, 0),
                        anchor_b=(-1.0, 0, 0), iterations=20)
    w.add_joints(o)
    w.step(1)
    b = w.bodies()
    np.testing.assert_array_equal(b["color"], [0, 1, 0, 1, 0])
    np.testing.assert_array_equal(b["parent"], [0] * n)
    sched = w.schedule()
    assert sched["island_count"] == 1 and sched["budget"] == 20
    j = w.joints()
    # lists are built by pushing joints to the front: body 1's list holds
    # joint 1 then joint 0.
    assert b["joint_head"][1] == 1 and j["avbd_next_a"][1] == 0


@pytest.mark.parametrize("iterations,max_gap", [(10, 0.04), (40, 0.008)])
def test_capsule_chains_hold_together_and_rest_on_the_ground(iterations, max_gap):
    """Four 24-link capsule chains (ball and hinge joints in turn, every
    8th link 100x heavier) swing, hit each other and land on the ground.
    The joints stretch most next to the heavy links and less with more
    iterations (AVBD converges the hard rows over iterations and steps):
    after 150 steps the median anchor gap is under 1 mm and the largest
    under `max_gap`; no link sinks below the ground by more than the
    collision margin, hinge axes stay aligned and no joint breaks."""
    sc = S.chains(count=4, links=24)
    w = sc.build(O.OracleWorld(profile=0, base_iterations=iterations))
    for _ in range(150):
        w.step(1)
    s = w.states()
    o = sc.joints
    gaps, axis_err = [], []
    for k in range(len(o)):
        a, b = int(o["body_a"][k][0]), int(o["body_b"][k][0])
        pa = s["position"][a] + rotate(s["orientation"][a], o["local_anchor_a"][k].astype(np.float64))
        pb = s["position"][b] + rotate(s["orientation"][b], o["local_anchor_b"][k].astype(np.float64))
        gaps.append(np.linalg.norm(pa - pb))
        if o["type"][k] == L.JOINT_HINGE:
            za = rotate(s["orientation"][a], np.array([0.0, 0.0, 1.0]))
            zb = rotate(s["orientation"][b], np.array([0.0, 0.0, 1.0]))
            axis_err.append(1.0 - abs(float(np.dot(za, zb))))
    assert np.median(gaps) < 1e-3, np.median(gaps)
    assert max(gaps) < max_gap, max(gaps)
    assert max(axis_err) < 1e-3, max(axis_err)
    assert s["position"][2:, 1].min() > 0.05 - 0.006
    assert w.joints()["broken"].sum() == 0
    assert w.stats()["manifolds"] > 0


def test_the_seam_resolves_joint_bodies_through_ids():
    """The seam's joints name their bodies by id (rigid_avbd_find_dense,
    geometry_constraints.slang:318-325): the same joints under shifted ids
    solve exactly as under {dense index, 1}; a joint whose body id matches
    no body gets no rows."""
    import _oracle as O
    base = S.joint_seam_inputs(id_offset=0, id_generation=1)
    moved = S.joint_seam_inputs(id_offset=1000, id_generation=3)
    out = []
    for inp in (base, moved):
        st, man, bod, jo = inp["states"].copy(), inp["manifolds"].copy(), inp["bodies"].copy(), inp["joints"].copy()
        sched = O.prepare_solve(inp["params"], inp["previous"], st, man, inp["pairs"], bodies=bod, base=10,
                                profile=0, lanes=32, joints=jo, ids=inp["ids"])
        out.append((st, jo, bod, sched))
    (s0, j0, b0, c0), (s1, j1, b1, c1) = out
    assert s0.tobytes() == s1.tobytes()
    np.testing.assert_array_equal(j0["rows"], j1["rows"])
    np.testing.assert_array_equal(j0["avbd_count"], j1["avbd_count"])
    np.testing.assert_array_equal(b0["joint_head"], b1["joint_head"])
    assert (j0["avbd_count"][j0["alive"] != 0] > 0).any()
    assert c0.tobytes() == c1.tobytes()
    lost = moved["joints"].copy()
    lost["options"]["body_b"][0] = (999999, 3)          # names no body
    st = moved["states"].copy()
    O.prepare_solve(moved["params"], moved["previous"], st, moved["manifolds"].copy(), moved["pairs"],
                    bodies=moved["bodies"].copy(), base=10, profile=0, lanes=32, joints=lost, ids=moved["ids"])
    assert lost["avbd_count"][0] == 0 and lost["dense_b"][0] == 0xffffffff
