# SPDX-License-Identifier: MIT
"""The C-ABI library loads, exports every entry point include/pm_hip_rigid.h
declares, its record layouts equal the numpy views (and so the reference's
layouts), and with no GPU it refuses to run instead of falling back."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import pmamd
from pmamd import layouts as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pm_hip_rigid.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pmh[A-Z]\w*)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = pmamd.lib()
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_header_record_sizes_match_layouts(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include "pm_hip_rigid.h"\n#include <stdio.h>\n#include <stddef.h>\nint main(void){'
                   'printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(PmhRigidBodyState),'
                   'sizeof(PmhBodyParameters), sizeof(PmhHingeContactFrame), sizeof(PmhContact),'
                   'sizeof(PmhContactManifold), sizeof(PmhAvbdBody), sizeof(PmhContactSchedule),'
                   'offsetof(PmhContact, avbd_dual)); return 0;}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()))
    assert out == [L.STATE.itemsize, L.PARAMS.itemsize, L.HINGE.itemsize, L.CONTACT.itemsize,
                   L.MANIFOLD.itemsize, L.AVBD_BODY.itemsize, L.SCHEDULE.itemsize,
                   L.CONTACT.fields["avbd_dual"][1]]
    # Reference sizes (sizeof probe over src/simulation/geometry_storage.hpp).
    assert out[:7] == [52, 64, 72, 284, 2304, 236, 264]


def _has_gpu():
    try:
        pmamd.device_info()
        return True
    except pmamd.PmhError:
        return False


def test_no_gpu_means_no_world():
    if _has_gpu():
        pytest.skip("a GPU is present")
    with pytest.raises(pmamd.PmhError) as e:
        pmamd.World()
    assert e.value.code == -5   # PM_ERROR_FEATURE_NOT_PRESENT, no CPU fallback


def test_invalid_arguments_are_rejected():
    lib = pmamd.lib()
    assert lib.pmhCreateWorld(None, None) == -1
    out = ctypes.c_void_p()
    assert lib.pmhPrepareSolveGeometryConstraints(None, None, None, None, 0, None, None, None, None,
                                                  None, 0.0, None, 0, 0, 1) == -1
    assert lib.pmhWorldAddMesh(None, None, 0, None, 0, ctypes.byref(out)) == -1


def test_realtime_worlds_take_the_realtime_step_only():
    """Realtime v1 is a 1/60 s step (one float ulp either side) of one
    substep (pmCmdStepSimulation, resources.cpp:1240-1241): anything else is
    an invalid argument, refused before any device is touched."""
    lib = pmamd.lib()
    out = ctypes.c_void_p()
    for dt, substeps in ((1.0 / 120.0, 1), (1.0 / 60.0, 4), (float(np.float32(1.0 / 60.0) * np.float32(1.0 + 3e-7)), 1)):
        desc = pmamd.WorldDesc((ctypes.c_float * 3)(0.0, -9.81, 0.0), dt, 1, 10, 0, 0, 0, 0, 0, 0, 0, 0, 0.0, substeps)
        assert lib.pmhCreateWorld(ctypes.byref(desc), ctypes.byref(out)) == -1
