# SPDX-License-Identifier: MIT
"""Long runs on the GPU (300 steps of the pyramid, a 4096-body slice of the
pile and 16 chains of 64 links) against the oracle's aggregate series
(tests/golden/longrun, tests/longrun.py): the deepest contact, the deepest
reach below the floor, the total energy, the mean height, the contact count
and the joints' anchor gaps.

The pyramid stands: the GPU's series agree with the oracle's in the GPU's
summation order (32 lanes per body) to float rounding over all 300 steps.
The pile and the chains are chaotic: the last bits of the device's and the
host's arithmetic differ (square roots, sines), the runs part within some
tens of steps and bodies reach different places.  Their series are held to
stated bands against both of the oracle's orders, each about three times
the spread the two orders of the oracle show (checked on the fixture)."""
import json
import pathlib

import numpy as np
import pytest

import longrun as LR
import pmamd

pytestmark = pytest.mark.gpu
FIXTURES = pathlib.Path(__file__).resolve().parent / "golden" / "longrun"

# Bands: (relative, absolute) per series, about three times the largest
# spread between the oracle's two orders over the run (the pile's deepest
# contact spreads most while it falls: swept contacts report the depth the
# step's remaining travel adds).
BANDS = {
    "pyramid": {"deepest": (0.02, 3e-5), "floor_depth": (0.0, 3e-5), "energy": (1e-5, 0.0),
                "mean_height": (0.0, 1e-4), "points": (0.08, 2)},
    "pile4096": {"deepest": (0.7, 0.09), "floor_depth": (0.6, 5e-3), "energy": (0.015, 0.0),
                 "mean_height": (0.0, 0.02), "points": (0.03, 0)},
    "chains16x64": {"deepest": (0.7, 0.015), "floor_depth": (0.6, 1.5e-3), "energy": (0.002, 0.0),
                    "mean_height": (0.0, 0.02), "points": (0.5, 10), "joint_gap_max": (0.25, 0.013),
                    "joint_gap_mean": (0.06, 7e-4)},
}
CHAOTIC = {"pile4096", "chains16x64"}


def _load(name):
    return json.loads((FIXTURES / f"{name}.json").read_text())


def _within(got, want, rel, abs_):
    got, want = np.asarray(got, float), np.asarray(want, float)
    return np.abs(got - want) <= abs_ + rel * np.maximum(np.abs(got), np.abs(want))


@pytest.mark.parametrize("name", list(LR.SCENES))
def test_long_run_aggregates_follow_the_oracle(name):
    print(f"{name}: {LR.STEPS} steps on the GPU", flush=True)
    fx = _load(name)
    got = LR.run(name, pmamd.World(profile=0, base_iterations=10, solve_lanes=32))
    assert got["step"] == fx["lanes32"]["step"]
    if name not in CHAOTIC:
        same = fx["lanes32"]
        for key in got:
            if key == "step" or not got[key]:
                continue
            np.testing.assert_allclose(got[key], same[key], rtol=1e-5, atol=1e-7, err_msg=f"{name} {key} (GPU order)")
    for order in ("lanes0", "lanes32"):
        ref = fx[order]
        for key, (rel, abs_) in BANDS[name].items():
            ok = _within(got[key], ref[key], rel, abs_)
            assert ok.all(), f"{name} {key} against {order}: {np.asarray(got[key])[~ok]} / {np.asarray(ref[key])[~ok]}"
