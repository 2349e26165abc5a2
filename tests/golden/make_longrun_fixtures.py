# SPDX-License-Identifier: MIT
"""Writes tests/golden/longrun/<scene>.json: the long-run aggregate series
of tests/longrun.py for each of its scenes, from the oracle in the
reference's summation order (lanes 0) and in the GPU's (lanes 32), legacy
profile, 10 iterations, dt 1/60, 300 steps.

    python tests/golden/make_longrun_fixtures.py [scene ...]
"""
import json
import pathlib
import sys
import time

HERE = pathlib.Path(__file__).resolve().parent
sys.path[:0] = [str(HERE.parent), str(HERE.parent.parent / "parallel-mater_amd")]

import _oracle as O  # noqa: E402
import longrun as LR  # noqa: E402


def main(names):
    out_dir = HERE / "longrun"
    out_dir.mkdir(exist_ok=True)
    for name in names or list(LR.SCENES):
        fixture = {"scene": name, "steps": LR.STEPS, "every": LR.SCENES[name]["every"], "profile": 0,
                   "iterations": 10, "dt": 1.0 / 60.0}
        for lanes in (0, 32):
            t = time.time()
            fixture[f"lanes{lanes}"] = LR.run(name, O.OracleWorld(profile=0, base_iterations=10, lanes=lanes))
            print(f"{name} lanes {lanes}: {time.time() - t:.1f} s", flush=True)
        (out_dir / f"{name}.json").write_text(json.dumps(fixture, indent=1) + "\n")


if __name__ == "__main__":
    main(sys.argv[1:])
