# SPDX-License-Identifier: MIT
"""The oracle against the reference's CUDA conformance captures of its three
analytic rigid cases (conformance/v1/golden/cuda/{rigid-direct,
constraint-breaking,compound-weld-lifecycle}.json, kept as fixtures under
tests/golden/conformance) and of its nine integrated rigid cases, whose
scenes are GLB files (examples/assets/Constraint*.glb, PassiveActive.glb,
kept as fixtures under tests/golden/glb).

The captures pin, within the cases' own tolerance classes: free flight under
4 substeps (direct class, 1e-5), an applied impulse, a kinematic target
reached over a frame's substeps, a body replaced mid-run
(set_rigid_body_state), a fixed joint breaking in its first substep (applied
impulse 0.896216, broken and disabled), a three-body weld compound moving as
one, released (its members then touch: 8 contact events with their
positions, normals and impulses) and rebuilt (fresh joint state).
"""
import numpy as np
import pytest

import _oracle as O
import conformance as C
import glb_scenes as G


def oracle_world(**kw):
    args = dict(dt=1.0 / 60.0, profile=0, base_iterations=C.ITERATIONS, substeps=C.SUBSTEPS)
    args.update(kw)
    return O.OracleWorld(**args)


@pytest.mark.parametrize("name", C.CASES)
def test_oracle_matches_the_cuda_capture(name):
    result = C.run(name, oracle_world())
    assert C.compare(name, result, C.EXACT_FRAMES.get(name)) == []
    if name == "rigid-direct":
        assert C.impact_band(result) == []


@pytest.mark.parametrize("name", G.CASES)
def test_oracle_follows_the_cuda_capture_of_the_glb_cases(name):
    """The integrated cases, loaded from their GLB scenes as the reference's
    loader does: every record at frame 0 (centred shapes, joint frames), and
    the records of C.GLB_FOLLOWS at the later checkpoints: the piston case
    whole; the slider case whole up to frame 30 (its sphere thrown at the
    rail's end, the rebound and the contacts) and its block and joint to
    frame 60; the spring-driven generic joint's block and impulse; the motor
    car's chassis and four wheels; the gear train's first and third hinges
    and its return spring (its second hinge to frame 45: the gear teeth
    make the rest depend on the summation order); the thrown case's second
    monkey head to frame 24."""
    result = C.run(name, oracle_world())
    assert C.glb_failures(name, result) == []


@pytest.mark.parametrize("name, change", [("constraint-piston", dict(substeps=1)),
                                          ("constraint-slider", dict(substeps=2)),
                                          ("rigid-direct", dict(substeps=1)),
                                          ("constraint-breaking", dict(profile=1, substeps=1)),
                                          ("compound-weld-lifecycle", dict(substeps=2))])
def test_a_different_step_contract_fails_the_capture(name, change):
    """One or two substeps per frame, or the realtime profile (whose contract
    is one substep), each move the oracle off the capture: the comparison
    discriminates."""
    result = C.run(name, oracle_world(**change))
    if name in G.CASES:
        assert C.glb_failures(name, result) != []
    else:
        assert C.compare(name, result, C.EXACT_FRAMES.get(name)) != []


@pytest.mark.parametrize("dt, substeps, ok", [(1.0 / 60.0, 1, True), (np.nextafter(np.float32(1.0 / 60.0), 1.0), 1, True),
                                              (1.0 / 60.0, 4, False), (1.0 / 120.0, 1, False),
                                              (np.float32(1.0 / 60.0) * np.float32(1.0 + 3e-7), 1, False)])
def test_realtime_worlds_take_the_realtime_step_only(dt, substeps, ok):
    """Realtime v1 is a 1/60 s step (within one float ulp) of one substep
    (rigid_realtime.hpp:16-26, pmCmdStepSimulation at resources.cpp:1240-1241);
    anything else is refused at creation.  The legacy profile takes any."""
    if ok:
        O.OracleWorld(dt=float(dt), profile=1, substeps=substeps)
    else:
        with pytest.raises(ValueError):
            O.OracleWorld(dt=float(dt), profile=1, substeps=substeps)
    O.OracleWorld(dt=float(dt), profile=0, substeps=substeps)
