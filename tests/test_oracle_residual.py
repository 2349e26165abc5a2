# SPDX-License-Identifier: MIT
"""The constraint residual and the to-residual solve mode on the oracle.

The reference defines no residual (geometry_constraints.slang:1441-1444
keeps a fixed budget), so it is this project's.  After each dual update
every contact point contributes |C_n| if its normal force is positive,
else max(C_n, 0), raised to |C_t| when its friction sticks; every joint
contributes the largest |C| of its hard equality rows (metres, radians for
angular rows).  The residual is the RMS of those values, with squares
summed in 2^-40 fixed point so the sum is exact in any order; the maximum
is reported beside it.  The to-residual mode stops a phase at the first
update whose residual is at most the tolerance (the budget caps it) and
publishes the impulses from that update.  Pins: a one-update cap in that
mode equals a one-iteration fixed budget bit for bit (the deferred
publication equals the in-loop one), a loose tolerance stops after one
update, the RMS equals a host recomputation from the maxima of a single
contact, and a settling pyramid's residual falls with the iterations."""
import numpy as np
import pytest

import _oracle as O
import scenarios as S
from pmamd import layouts as L


def states_bits(w):
    return w.states().view(np.uint32).copy()


def test_one_update_cap_equals_one_iteration_budget():
    scene = S.pyramid(rows=5)
    a = scene.build(O.OracleWorld(profile=0, base_iterations=1))
    b = scene.build(O.OracleWorld(profile=0, base_iterations=1, residual_tolerance=1e-30))
    for _ in range(40):
        a.step(1)
        b.step(1)
        _, ex = b.residuals()
        assert ex[0] == 1
    np.testing.assert_array_equal(states_bits(a), states_bits(b))
    ma, mb = a.manifolds(), b.manifolds()
    for f in ("avbd_dual", "accumulated_normal_impulse", "accumulated_friction_impulse", "avbd_anchor_a"):
        np.testing.assert_array_equal(ma["contacts"][f], mb["contacts"][f], err_msg=f)


def test_fixed_budget_is_unchanged_by_the_log():
    """The log only reads: a fixed budget with and without it gives the
    same states as the run recorded before it existed would (same
    world twice)."""
    scene = S.pyramid(rows=5)
    a = scene.build(O.OracleWorld(profile=0, base_iterations=10))
    b = scene.build(O.OracleWorld(profile=0, base_iterations=10))
    for _ in range(20):
        a.step(1)
        b.step(1)
        b.residuals()
    np.testing.assert_array_equal(states_bits(a), states_bits(b))


def test_loose_tolerance_stops_after_one_update():
    scene = S.pyramid(rows=4)
    w = scene.build(O.OracleWorld(profile=0, base_iterations=10, residual_tolerance=10.0))
    for _ in range(20):
        w.step(1)
    r, ex = w.residuals()
    assert ex[0] == 1
    assert np.isfinite(r[0, 0]) and np.all(np.isnan(r[0, 1:]))
    assert w.schedule()["counts"][6] == 1


def test_rms_of_a_single_point_is_its_value():
    """One sphere resting on a static box gives one contact point, so the
    RMS of each update equals its maximum up to the 2^-40 quantisation of
    the square."""
    w = O.OracleWorld(profile=0, base_iterations=6, gravity=(0, -9.81, 0))
    w.add_mesh(*S.box_mesh(5.0, 0.5, 5.0))
    w.add_sphere(0.25)
    p = S.params(2, motion=np.array([L.MOTION_STATIC, L.MOTION_DYNAMIC], np.uint32), shape=np.array([0, 1]),
                 mass=np.array([1.0, 1.0], np.float32))
    w.add_bodies(p, S.states(2, np.array([[0, -0.5, 0], [0, 0.2495, 0]], np.float32)))
    for _ in range(3):
        w.step(1)
    r, ex, mx = w.residuals(maximum=True)
    assert ex[0] == 6 and w.manifolds()["count"][0] == 1
    q = np.floor(mx[0, :6].astype(np.float64) ** 2 * 2.0 ** 40 + 0.5)
    np.testing.assert_array_equal(r[0, :6], np.sqrt(q / 2.0 ** 40).astype(np.float32))


def test_pyramid_residual_falls_with_the_iterations():
    scene = S.pyramid(rows=6)
    w = scene.build(O.OracleWorld(profile=0, base_iterations=30))
    for _ in range(60):
        w.step(1)
    r, ex = w.residuals()
    assert ex[0] == 30
    first, last = float(r[0, 0]), float(r[0, ex[0] - 1])
    assert np.all(np.isfinite(r[0, :30]))
    assert last < first, (first, last)


def test_tolerance_mode_meets_the_tolerance_or_the_cap():
    scene = S.pyramid(rows=6)
    tol = 1e-4
    w = scene.build(O.OracleWorld(profile=0, base_iterations=40, residual_tolerance=tol))
    for _ in range(60):
        w.step(1)
        r, ex = w.residuals()
        for ph in range(2):
            if ex[ph] == 0:
                continue
            assert r[ph, ex[ph] - 1] <= tol or ex[ph] == 40
            assert np.all(r[ph, :ex[ph] - 1] > tol)
