# SPDX-License-Identifier: MIT
"""One-step inputs of the predictor (pm_integrate_rigid_bodies,
geometry.slang:524-612): forces, torques and impulses on dynamic bodies
(apply_rigid_input, runtime_rigid.cu.inc:265-289) and kinematic targets
(set_kinematic_target, :197-230), pinned by the closed forms of one
semi-implicit Euler step."""
import numpy as np
import pytest

import _oracle as O
import scenarios as S
from pmamd import layouts as L

DT = np.float32(1.0) / np.float32(60.0)


def free_world(motion=L.MOTION_DYNAMIC, mass=2.0, inertia=(0.5, 0.25, 0.125), damping=0.0):
    w = O.OracleWorld(gravity=(0, 0, 0), profile=0)
    w.add_mesh(*S.box_mesh(0.5, 0.5, 0.5))
    p = S.params(1, motion=np.array([motion], np.uint32), mass=mass, inertia=inertia, damping=damping)
    w.add_bodies(p, S.states(1, np.array([[1.0, 2.0, 3.0]], np.float32)))
    return w


def inputs(**kw):
    a = np.zeros(1, L.RIGID_INPUT)
    for k, v in kw.items():
        a[k] = v
    return a


def test_linear_impulse_and_force():
    w = free_world()
    w.apply_inputs(inputs(linear_impulse=(3.0, 0.0, 0.0), force=(0.0, 60.0, 0.0)))
    w.step(1)
    s = w.states()[0]
    # v = J/m + F/m dt; x = x0 + v dt
    v = np.array([np.float32(3.0) * np.float32(0.5), np.float32(60.0) * np.float32(0.5) * DT, 0.0], np.float32)
    np.testing.assert_allclose(s["linear_velocity"], v, rtol=1e-6)
    np.testing.assert_allclose(s["position"], np.array([1, 2, 3], np.float32) + v * DT, rtol=1e-6)
    w.step(1)                                        # inputs last one step
    np.testing.assert_allclose(w.states()[0]["linear_velocity"], v, rtol=1e-6)


def test_accumulated_inputs_add():
    w = free_world()
    w.apply_inputs(inputs(linear_impulse=(1.0, 0.0, 0.0)))
    w.apply_inputs(inputs(linear_impulse=(1.0, 0.5, 0.0)))
    w.step(1)
    np.testing.assert_allclose(w.states()[0]["linear_velocity"], [1.0, 0.25, 0.0], rtol=1e-6)


def test_angular_impulse_and_torque():
    w = free_world()
    w.apply_inputs(inputs(angular_impulse=(0.5, 0.0, 0.0), torque=(0.0, 0.0, 6.0)))
    w.step(1)
    s = w.states()[0]
    # omega = I^-1 L + I^-1 tau dt (identity orientation, diagonal inertia)
    np.testing.assert_allclose(s["angular_velocity"], [0.5 / 0.5, 0.0, 6.0 / 0.125 * float(DT)], rtol=1e-5)
    q = s["orientation"]
    assert np.isclose(np.linalg.norm(q), 1.0, atol=1e-6) and q[0] > 0 and q[2] > 0


def test_damping_divides_velocity():
    w = free_world(damping=0.05)
    w.apply_inputs(inputs(linear_impulse=(2.0, 0.0, 0.0)))
    w.step(1)
    expect = np.float32(1.0) / (np.float32(1.0) + np.float32(0.05) * DT)
    assert w.states()[0]["linear_velocity"][0] == pytest.approx(float(expect), rel=1e-6)


def test_kinematic_target_is_reached_in_one_step():
    w = free_world(motion=L.MOTION_KINEMATIC)
    t = S.states(1, np.array([[1.5, 2.0, 3.0]], np.float32))
    t["orientation"][0] = S.quat((0, 1, 0), 0.3)
    w.set_kinematic_targets(0, t)
    w.step(1)
    s = w.states()[0]
    np.testing.assert_allclose(s["position"], [1.5, 2.0, 3.0], rtol=1e-7)
    np.testing.assert_allclose(s["linear_velocity"], [0.5 / float(DT), 0, 0], rtol=1e-5)
    np.testing.assert_allclose(s["angular_velocity"], [0, 0.3 / float(DT), 0], rtol=1e-4)
    np.testing.assert_allclose(np.abs(s["orientation"]), np.abs(S.quat((0, 1, 0), 0.3)), atol=1e-6)
    w.step(1)                                        # no target: the body stops
    s = w.states()[0]
    np.testing.assert_allclose(s["position"], [1.5, 2.0, 3.0], rtol=1e-7)
    np.testing.assert_array_equal(s["linear_velocity"], [0, 0, 0])


def test_invalid_inputs_are_rejected():
    w = free_world(motion=L.MOTION_KINEMATIC)
    with pytest.raises(ValueError):
        w.apply_inputs(inputs(force=(1.0, 0.0, 0.0)))          # not dynamic
    d = free_world()
    with pytest.raises(ValueError):
        d.set_kinematic_targets(0, S.states(1))                # not kinematic
    with pytest.raises(ValueError):
        d.apply_inputs(inputs(force=(np.inf, 0.0, 0.0)))
    k = free_world(motion=L.MOTION_KINEMATIC)
    zero = S.states(1)
    zero["orientation"][0] = 0.0
    with pytest.raises(ValueError):
        k.set_kinematic_targets(0, zero)
