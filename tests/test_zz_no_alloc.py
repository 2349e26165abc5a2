# SPDX-License-Identifier: MIT
"""The seam without allocation (pmhPrepareSolveGeometryConstraintsNoAlloc)
against the seam (pmhPrepareSolveGeometryConstraints), both through the C
ABI on the same inputs: states, manifolds, AvbdBody records and schedule
bit for bit, on two calls in a row, the second with a smaller world, so the
static array's leftovers from the first call cannot reach the second."""
import numpy as np
import pytest

import pmamd
import scenarios as S

pytestmark = pytest.mark.gpu

SEAM = "pmhPrepareSolveGeometryConstraints"
NO_ALLOC = "pmhPrepareSolveGeometryConstraintsNoAlloc"


def say(msg):
    print(msg, flush=True)


def pyramid_inputs():
    """A 10-row pyramid's contacts after three steps of a world: its
    parameters, the states before and after the fourth step's predictor,
    the manifolds and the pair codes."""
    scene = S.pyramid(rows=10)
    w = scene.build(pmamd.World(profile=0, base_iterations=10))
    w.step(3)
    previous = w.states()
    w.step(1)
    return (scene.params.copy(), previous, w.states(), w.manifolds(), w.pairs(), w.bodies(),
            dict(profile=0, base=10))


def fixture_inputs():
    p, prev, cur, m, pairs = S.realtime_fixture()
    return p, prev, cur, m, pairs, None, dict(profile=1, base=10, gravity=(0.0, 0.0, 0.0))


def run(entry, inputs):
    p, prev, cur, m, pairs, bodies, kw = inputs
    cur, m = cur.copy(), m.copy()
    bodies = np.zeros(len(p), dtype=pmamd.AVBD_BODY) if bodies is None else bodies.copy()
    sched = pmamd.prepare_solve(p, prev, cur, m, pairs, bodies=bodies, entry=entry, **kw)
    return cur, m, bodies, sched


def test_the_seam_without_allocation_matches_the_seam():
    for name, make in (("pyramid, 55 bodies", pyramid_inputs), ("realtime fixture, 8 bodies", fixture_inputs)):
        inputs = make()
        say(f"no-alloc seam vs seam: {name}, {len(inputs[4])} pairs")
        want = run(SEAM, inputs)
        got = run(NO_ALLOC, inputs)
        for label, a, b in zip(("states", "manifolds", "bodies", "schedule"), want, got):
            a, b = np.asarray(a), np.asarray(b)
            assert a.tobytes() == b.tobytes(), f"{name}: {label} differ"
        assert int(want[3]["budget"]) > 0


def test_the_seam_without_allocation_refuses_oversized_worlds():
    p, prev, cur, m, pairs = S.realtime_fixture()
    big = np.resize(p, 4097)
    with pytest.raises(pmamd.PmhError):
        pmamd.prepare_solve(big, np.resize(prev, 4097), np.resize(cur, 4097), m.copy(), pairs, entry=NO_ALLOC)
