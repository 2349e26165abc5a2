# SPDX-License-Identifier: MIT
"""Pins of oracle functions that no reference-held vector covers, each by
what the mathematics fixes, and each shown to discriminate by a mistake
planted in the oracle (never set outside these tests):

- free_mean (geometry_constraints_solve.slang:617-675): an island touching
  no static body conserves linear momentum to float precision; without the
  mean correction (planted) it drifts by tenths of a newton-second;
- the crossing safeguard (:433-475): a box released 2 mm above the ground,
  inside the contact reach but outside the 1 mm rest offset, lands at the
  rest offset (+1e-5) at rest within one step; without the safeguard
  (planted) it ends above that height and still moving;
- cache matching (geometry.slang:2941-3009): a cached contact is reused
  only while the normals agree to cos 8.11 deg = 0.99;
- the swept face manifold (geometry.slang:1751-1853): a box translating
  through the ground in one step gets its contacts at the time of impact
  (gap - rest) / travel, at its four bottom corners on the mid-gap plane,
  with the depth the remaining travel adds;
- hinge contact frames (rigid_hinge_contact_frame, geometry.slang:404-491):
  a slider held by a static post comes to rest against a static wall at
  the guided rest offset (0.1 mm), not the 1 mm of an unguided contact; a
  hinged lid's frame holds the post's anchor and axis and its contacts do
  not persist; a piston's frame has the closed-form inverse moment about
  its axis; welded boxes are fixed members.  Without frames (planted) each
  of these fails.
"""
import numpy as np
import pytest

import _oracle as O
import scenarios as S


def _momentum_drift(planted, lanes):
    w = O.OracleWorld(gravity=(0.0, 0.0, 0.0), profile=0, base_iterations=10, lanes=lanes, planted=planted)
    box = w.add_mesh(*S.box_mesh(0.5, 0.5, 0.5))
    m = np.array([1.0, 3.0, 2.0, 1.5], np.float32)
    p = S.params(4, mass=m, inertia=(0.2, 0.2, 0.2), friction=0.5, margin=0.01, shape=box)
    st = S.states(4, np.array([[0, 0, 0], [1.02, 0.3, 0], [-1.05, -0.2, 0.1], [0.1, 1.03, 0.2]], np.float32))
    st["linear_velocity"][1] = (-3.0, 0.5, 0.0)
    st["linear_velocity"][2] = (2.0, 0.0, 0.4)
    st["linear_velocity"][3] = (0.0, -2.5, 0.0)
    st["angular_velocity"][1] = (0.0, 0.0, 2.0)
    w.add_bodies(p, st)
    p0 = (m[:, None].astype(np.float64) * st["linear_velocity"]).sum(0)
    worst, points = 0.0, 0
    for _ in range(10):
        w.step(1)
        s = w.states()
        worst = max(worst, float(np.abs((m[:, None].astype(np.float64) * s["linear_velocity"]).sum(0) - p0).max()))
        points = max(points, w.stats()["points"])
    return worst, points


@pytest.mark.parametrize("lanes", [0, 32])
def test_an_unanchored_island_conserves_momentum(lanes):
    drift, points = _momentum_drift(0, lanes)
    assert points > 10                       # the four boxes collide
    assert drift < 5e-6                      # |sum m v| ~ 10 N s: float precision
    planted, _ = _momentum_drift(1, lanes)   # no free-motion mean
    assert planted > 0.1


def _landing(gap, planted):
    w = O.OracleWorld(profile=0, base_iterations=10, planted=planted)
    ground = w.add_mesh(*S.box_mesh(5.0, 0.5, 5.0))
    box = w.add_mesh(*S.box_mesh(0.5, 0.5, 0.5))
    p = S.params(2, motion=np.array([0, 2], np.uint32), mass=1.0, inertia=(1 / 6, 1 / 6, 1 / 6), margin=0.01,
                 shape=np.array([ground, box]))
    w.add_bodies(p, S.states(2, np.array([[0, -0.5, 0], [0, 0.5 + gap, 0]], np.float32)))
    w.step(1)
    s = w.states()
    return float(s["position"][1][1]) - 0.5, float(s["linear_velocity"][1][1]), w.stats()["points"]


def test_the_crossing_safeguard_lands_a_box_at_the_rest_offset():
    height, speed, points = _landing(2e-3, 0)
    assert points == 4
    assert abs(height - (1e-3 + 1e-5)) < 2e-6   # rest offset min(margin, 1 mm) + 1e-5
    assert abs(speed) < 1e-6
    height, speed, _ = _landing(2e-3, 2)        # no crossing safeguard
    assert height > 2e-3 and abs(speed) > 1e-2


def _tilted_hits(degrees):
    """A box resting on the ground for one step, then both turned rigidly by
    `degrees` about z through the origin: the contact points keep their
    place in the box's frame, the normal turns by the angle."""
    w = O.OracleWorld(profile=0, base_iterations=10)
    ground = w.add_mesh(*S.box_mesh(5.0, 0.5, 5.0))
    box = w.add_mesh(*S.box_mesh(0.5, 0.5, 0.5))
    p = S.params(2, motion=np.array([0, 2], np.uint32), mass=1.0, inertia=(1 / 6, 1 / 6, 1 / 6), margin=0.01,
                 shape=np.array([ground, box]))
    w.add_bodies(p, S.states(2, np.array([[0, -0.5, 0], [0, 0.5 - 5e-4, 0]], np.float32)))
    w.step(1)
    s = w.states()
    th = np.deg2rad(degrees)
    rot = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    qx, qy, qz, qw = S.quat((0, 0, 1), th)
    new = s.copy()
    for i in range(2):
        new["position"][i] = (rot @ s["position"][i].astype(np.float64)).astype(np.float32)
        x, y, z, ww = s["orientation"][i].astype(np.float64)
        new["orientation"][i] = np.array([qw * x + qx * ww + qy * z - qz * y, qw * y - qx * z + qy * ww + qz * x,
                                          qw * z + qx * y - qy * x + qz * ww, qw * ww - qx * x - qy * y - qz * z],
                                         np.float32)
    w.set_rigid_states(0, new)
    w.step(1)
    return w.stats()["cache_hits"], w.stats()["points"]


def test_cached_contacts_need_normals_within_cos_0_99():
    assert np.cos(np.deg2rad(8.0)) > 0.99 > np.cos(np.deg2rad(8.2))
    assert _tilted_hits(0.0) == (4, 4)
    assert _tilted_hits(8.0) == (4, 4)
    assert _tilted_hits(8.2) == (0, 4)


def test_the_swept_face_manifold_finds_the_time_of_impact():
    w = O.OracleWorld(gravity=(0.0, 0.0, 0.0), profile=0, base_iterations=10)
    ground = w.add_mesh(*S.box_mesh(5.0, 0.5, 5.0))
    box = w.add_mesh(*S.box_mesh(0.5, 0.5, 0.5))
    p = S.params(2, motion=np.array([0, 2], np.uint32), mass=1.0, inertia=(1 / 6, 1 / 6, 1 / 6), margin=0.01,
                 shape=np.array([ground, box]))
    gap, speed = np.float32(0.3), np.float32(-60.0)
    st = S.states(2, np.array([[0, -0.5, 0], [0, 0.5 + gap, 0]], np.float32))
    st["linear_velocity"][1] = (0.0, speed, 0.0)
    w.add_bodies(p, st)
    w.step(1)
    m = w.manifolds()[0]
    c = m["contacts"][:m["count"]]
    y0 = np.float32(0.5) + gap
    travel = float(y0 - (y0 + speed * np.float32(1.0 / 60.0)))    # 1 m in the step
    rest = 1e-3                                                    # min(margin, 1 mm)
    entry = (float(gap) - rest) / travel
    assert m["count"] == 4
    np.testing.assert_allclose(c["impact_fraction"], entry, atol=1e-6)
    np.testing.assert_allclose(c["penetration"], travel * (1.0 - entry) + rest + 2e-5, atol=2e-6)
    np.testing.assert_allclose(c["normal"], np.tile([0.0, 1.0, 0.0], (4, 1)), atol=1e-7)
    np.testing.assert_allclose(np.sort(np.abs(c["point"][:, [0, 2]]).ravel()), 0.5, atol=1e-6)
    np.testing.assert_allclose(c["point"][:, 1], rest / 2, atol=1e-6)


def _hinge_world(planted, steps):
    scene = S.hinge_contacts()
    w = scene.build(O.OracleWorld(profile=0, base_iterations=10, planted=planted))
    frames = []
    for _ in range(steps):
        w.step(1)
        frames.append((w.pairs().copy(), w.manifolds().copy()))
    return scene, w, frames


def _contacts(scene, pairs, manifolds, a, b):
    n = scene.n
    for k, code in enumerate(pairs):
        if code // n == a and code % n == b:
            m = manifolds[k]
            return m["contacts"][:m["count"]]
    return None


def test_hinge_contact_frames():
    for planted in (0, 4):
        scene, w, frames = _hinge_world(planted, 40)
        s = w.states()
        # slider (body 2) against its wall (body 3, face at x = 0.9)
        gap = 0.9 - (float(s["position"][2][0]) + 0.25)
        lid = _contacts(scene, *frames[-1], 8, 0)
        weld = _contacts(scene, *frames[-1], 10, 9)
        piston = [c for pr, ms in frames for c in [_contacts(scene, pr, ms, 5, 6)] if c is not None and len(c)]
        if planted == 0:
            assert 0.0 < gap < 1.5e-4, gap
            h = lid["body_hinge"]
            assert (h["present"] == 1).all() and (h["fixed"] == 1).all() and (lid["persistent"] == 0).all()
            np.testing.assert_allclose(h["anchor"], np.tile([-2.25, 0.0, 0.0], (len(lid), 1)), atol=1e-6)
            np.testing.assert_allclose(h["axis"], np.tile([0.0, 0.0, 1.0], (len(lid), 1)), atol=1e-6)
            assert (weld["body_hinge"]["fixed_member"] == 1).all() and (weld["persistent"] == 0).all()
            pf = piston[0]["body_hinge"]
            assert (pf["axial"] == 1).all() and (pf["axial_rotation"] == 1).all()
            np.testing.assert_allclose(pf["inverse_mass"], 0.5)
            np.testing.assert_allclose(pf["inverse_moment"], 12.0, rtol=1e-3)   # 1 / I_xx of the 0.5 m cube, 2 kg
            assert (piston[0]["persistent"] == 1).all()                          # guided contacts persist
        else:
            assert gap > 5e-4, gap
            assert (lid["body_hinge"]["present"] == 0).all() and (lid["persistent"] == 1).all()
            assert (weld["body_hinge"]["fixed_member"] == 0).all()
            assert all((c["body_hinge"]["axial"] == 0).all() for c in piston)
