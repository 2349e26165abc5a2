# SPDX-License-Identifier: MIT
"""Build checks on the device code (CPU only: hipcc cross-compiles gfx950).

Packed FP32 arithmetic (v_pk_add_f32, v_pk_mul_f32, v_pk_fma_f32) is kept
out of every kernel: the build passes the target feature -packed-fp32-ops
to the device compile, and build.py disassembles the linked library and
fails on any such instruction.  The host pass of hipcc warns that it does
not know the feature; the device pass applies it, which the probe below
shows: the same float2 kernel compiles to packed instructions without the
feature flag and to none with the library's flags.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "parallel-mater_amd"))
import build  # noqa: E402

PROBE = r"""
#include <hip/hip_runtime.h>
__global__ void axpy2(float2 *a, const float2 *b, float s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float2 x = a[i], y = b[i];
    x.x = x.x * s + y.x;
    x.y = x.y * s + y.y;
    a[i] = x;
}
"""

needs_hipcc = pytest.mark.skipif(not os.path.exists(build.HIPCC) or not os.path.exists(build.OBJDUMP),
                                 reason="hipcc or llvm-objdump missing")


@needs_hipcc
def test_the_library_has_no_packed_fp32_arithmetic():
    if not os.path.exists(build.LIB):
        pytest.skip("library not built")
    assert build.packed_fp32_ops(build.LIB) == {}


def _probe(tmp_path, flags, name):
    src = tmp_path / "probe.hip"
    src.write_text(PROBE)
    out = tmp_path / name
    subprocess.run([build.HIPCC, *flags, "-shared", str(src), "-o", str(out)], check=True, capture_output=True)
    return build.packed_fp32_ops(str(out))


@needs_hipcc
def test_the_feature_flag_keeps_packed_fp32_out(tmp_path):
    flags = list(build.FLAGS)
    assert _probe(tmp_path, flags, "with.so") == {}
    i = flags.index("-packed-fp32-ops")
    without = flags[:i - 3] + flags[i + 1:]      # drop -Xclang -target-feature -Xclang -packed-fp32-ops
    assert "-packed-fp32-ops" not in without and len(without) == len(flags) - 4
    found = _probe(tmp_path, without, "without.so")
    assert sum(found.values()) > 0, found
