# SPDX-License-Identifier: MIT
"""The long-run aggregate fixtures (tests/golden/longrun) on the CPU:

- the pyramid's series are the oracle's own (re-run here, equal to the last
  bit of every sample);
- every scene keeps the reference's realtime gates
  (tests/solver/rigid_realtime_cuda_tests.cu:365-370) in the reference's
  summation order: no body more than 5 mm below the floor once settled (the
  last third of the run; while the pile falls its impacts reach 13 mm), and
  no energy growth above 5% between the first and the last sample (the pile
  and the chains lose energy as they settle; the pyramid stands);
- the GPU test's bands against the reference order are wider than the
  spread between the oracle's two summation orders at every sample.
"""
import json
import pathlib

import numpy as np
import pytest

import _oracle as O
import longrun as LR
from test_gpu_longrun import BANDS, _within

FIXTURES = pathlib.Path(__file__).resolve().parent / "golden" / "longrun"


def _load(name):
    return json.loads((FIXTURES / f"{name}.json").read_text())


def test_the_pyramid_fixture_is_the_oracle_run():
    fx = _load("pyramid")
    got = LR.run("pyramid", O.OracleWorld(profile=0, base_iterations=10, lanes=0))
    for key, series in got.items():
        assert series == fx["lanes0"][key], key


@pytest.mark.parametrize("name", list(LR.SCENES))
def test_long_runs_keep_the_reference_gates(name):
    ref = _load(name)["lanes0"]
    settled = ref["floor_depth"][2 * len(ref["floor_depth"]) // 3:]
    assert max(settled) <= 0.005
    e = np.asarray(ref["energy"])
    assert (e[-1] - e[0]) / abs(e[0]) <= 0.05


@pytest.mark.parametrize("name", list(LR.SCENES))
def test_the_bands_cover_the_spread_of_the_two_orders(name):
    fx = _load(name)
    for key, (rel, abs_) in BANDS[name].items():
        ok = _within(fx["lanes32"][key], fx["lanes0"][key], rel, abs_)
        assert ok.all(), f"{name} {key}"
