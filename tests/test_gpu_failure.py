# SPDX-License-Identifier: MIT
"""Failure handling of the step, forced from the host so that every kernel
still ends normally.

- A colouring wait that gives up (here: the wait limit set to 0, so the first
  wait on an uncoloured neighbour gives up) ends the colouring, the substep
  returns PM_ERROR_INVALID_STATE before its solve, and the states stay those
  before that step; with the default limit the world then steps on.
- A dataflow-sweep wait that runs out (the sweep's wait limit set to 0, so
  the first unmet wait on a lower-colour neighbour runs out) ends every wait
  of the step; the step keeps its states and returns
  PM_ERROR_INVALID_STATE, and the world then steps on.
- Worlds of fewer than four bodies are refused before any device work.
"""
import sys

import numpy as np
import pytest

import conformance as C
import pmamd
import scenarios as S

pytestmark = pytest.mark.gpu
INVALID_STATE = -2
FEATURE_NOT_PRESENT = -5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_a_colouring_that_gives_up_fails_the_step_and_keeps_the_states():
    say("pyramid of 55 boxes: 2 steps, one step with colour wait limit 0, 3 more steps")
    scene = S.pyramid()
    g = scene.build(pmamd.World(profile=0, base_iterations=10))
    g.step(2)
    before = g.states()
    # After the first step the world colours with per-lane tickets (few
    # neighbours per body): the lane of a box waits on the lane of the box
    # below it in the same wavefront, so a limit of 0 always gives up.
    g.set_color_wait_limit(0)
    with pytest.raises(pmamd.PmhError) as err:
        g.step(1)
    assert err.value.code == INVALID_STATE
    assert g.stats()["status"] & 16
    after = g.states()
    for f in ("position", "orientation", "linear_velocity", "angular_velocity"):
        np.testing.assert_array_equal(after[f], before[f], err_msg=f)
    g.set_color_wait_limit(1 << 22)
    g.step(3)
    assert g.stats()["status"] == 0
    s = g.states()
    assert np.isfinite(s["position"]).all()
    # the pyramid still stands: every box within 5 cm of where it started
    assert np.abs(s["position"][1:] - scene.states["position"][1:]).max() < 0.05
    g.close()


def test_a_sweep_wait_that_runs_out_fails_the_step_and_keeps_the_states():
    say("pyramid of 55 boxes: 2 steps, steps with sweep wait limit 0 until one fails, 3 more steps")
    scene = S.pyramid()
    g = scene.build(pmamd.World(profile=0, base_iterations=10))
    g.step(2)
    g.set_sweep_wait_limit(0)
    failed = None
    for attempt in range(4):
        before = g.states()
        try:
            g.step(1)
        except pmamd.PmhError as e:
            failed = e.code
            break
    # The boxes of the second colour rest on those of the first and are
    # ticketed right after them: one of them meets an unpublished neighbour.
    assert failed == INVALID_STATE
    say(f"failed at attempt {attempt}")
    after = g.states()
    for f in ("position", "orientation", "linear_velocity", "angular_velocity"):
        np.testing.assert_array_equal(after[f], before[f], err_msg=f)
    g.set_sweep_wait_limit(1 << 22)
    g.step(3)
    assert g.stats()["status"] == 0
    s = g.states()
    assert np.isfinite(s["position"]).all()
    assert np.abs(s["position"][1:] - scene.states["position"][1:]).max() < 0.05
    g.close()


def test_worlds_under_four_bodies_are_refused():
    say("three-body world: step refused before any device work")
    g = pmamd.World(profile=0, base_iterations=10)
    C.build("rigid-direct", g)
    with pytest.raises(pmamd.PmhError) as err:
        g.step(1)
    assert err.value.code == FEATURE_NOT_PRESENT
    g.close()
