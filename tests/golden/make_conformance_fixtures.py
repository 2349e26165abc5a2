# SPDX-License-Identifier: MIT
"""Writes tests/golden/conformance/<case>.json from the reference's
conformance inputs (conformance/v1/cases/<case>.json) and CUDA captures
(conformance/v1/golden/cuda/<case>.json) for the three analytic rigid
cases.  Each fixture holds the case's frames, checkpoints, commands,
resources and tolerances, and per checkpoint the rigid-body and constraint
records and the rigid contacts of the capture (data only).

Run in a checkout that holds the reference:
    python tests/golden/make_conformance_fixtures.py /root/reference
"""
import json
import pathlib
import sys

CASES = ("rigid-direct", "constraint-breaking", "compound-weld-lifecycle",
         "constraint-point", "constraint-fixed", "constraint-hinge", "constraint-slider", "constraint-piston",
         "constraint-generic", "constraint-generic-spring", "constraint-motor", "passive-active")
KEEP_CASE = ("id", "checkpoints", "commands", "resources", "tolerances", "world")
KEEP_RESOURCE = ("id", "type", "tolerance_class", "present", "position", "orientation",
                 "linear_velocity", "angular_velocity", "applied_impulse", "broken", "enabled")


def main(root):
    base = pathlib.Path(root) / "conformance" / "v1"
    out_dir = pathlib.Path(__file__).resolve().parent / "conformance"
    out_dir.mkdir(exist_ok=True)
    for name in CASES:
        case = json.loads((base / "cases" / f"{name}.json").read_text())
        capture = json.loads((base / "golden" / "cuda" / f"{name}.json").read_text())
        checkpoints = []
        for cp in capture["checkpoints"]:
            checkpoints.append({
                "frame": cp["frame"],
                "resources": [{k: r[k] for k in KEEP_RESOURCE if k in r} for r in cp["resources"]],
                "rigid_contact_count": cp["contacts"]["rigid_contact_count"],
                "rigid_contacts": cp["contacts"]["rigid_contacts"],
            })
        fixture = {
            "source": {"case": f"conformance/v1/cases/{name}.json",
                       "capture": f"conformance/v1/golden/cuda/{name}.json",
                       "backend": capture["backend"], "case_sha256": capture["case_sha256"]},
            "case": {k: case[k] for k in KEEP_CASE},
            "checkpoints": checkpoints,
        }
        (out_dir / f"{name}.json").write_text(json.dumps(fixture, indent=1) + "\n")
        print(f"wrote {name}: {len(checkpoints)} checkpoints")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
