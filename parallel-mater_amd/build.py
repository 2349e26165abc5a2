# SPDX-License-Identifier: MIT
"""Builds libpm_hip_rigid.so for gfx950 in-tree (parallel-mater_amd/lib/).

Every translation unit is compiled unfused (-ffp-contract=off) so the GPU
evaluates each float expression in the reference's operation order, and
without packed-FP32 instructions.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib")
LIB = os.path.join(OUT, "libpm_hip_rigid.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PMH_ARCH", "gfx950")

FLAGS = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
         "-Wno-unused-command-line-argument"]


def sources():
    return sorted(f for f in os.listdir(SRC) if f.endswith(".hip"))


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    objs = []

    def compile_one(name):
        src = os.path.join(SRC, name)
        obj = os.path.join(OUT, name.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0 or verbose:
            sys.stderr.write(r.stdout + r.stderr)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {name}")
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as pool:
        objs = list(pool.map(compile_one, sources()))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    for obj in objs:
        os.remove(obj)
    found = packed_fp32_ops(LIB)
    if found:
        raise RuntimeError(f"packed FP32 arithmetic in the device code: {found}")
    build_driver()
    return LIB


OBJDUMP = os.environ.get("PMH_OBJDUMP", "/opt/rocm/llvm/bin/llvm-objdump")


def packed_fp32_ops(lib):
    """Counts v_pk_add_f32 / v_pk_mul_f32 / v_pk_fma_f32 in the gfx950 code
    objects of a library built here (the compiler forms them from plain
    float code; they are kept out of every kernel).  The code objects are
    unbundled in a scratch directory."""
    import re
    import shutil
    import tempfile
    pattern = re.compile(r"\bv_pk_(?:add|mul|fma)_f32\b")
    counts = {}
    with tempfile.TemporaryDirectory() as tmp:
        copy = os.path.join(tmp, os.path.basename(lib))
        shutil.copyfile(lib, copy)
        subprocess.run([OBJDUMP, "--offloading", copy], cwd=tmp, check=True, capture_output=True)
        for name in sorted(os.listdir(tmp)):
            if not name.endswith(ARCH):
                continue
            text = subprocess.run([OBJDUMP, "-d", f"--mcpu={ARCH}", os.path.join(tmp, name)], check=True,
                                  capture_output=True, text=True).stdout
            for op in pattern.findall(text):
                counts[op] = counts.get(op, 0) + 1
    return counts


DRIVER_SRC = os.path.join(HERE, "driver", "pmh_driver.cpp")
DRIVER = os.path.join(OUT, "libpm_hip_driver.so")
MANIFEST = os.path.join(OUT, "pm_hip_driver.json")


def build_driver():
    """The installable driver: host C++ over libpm_hip_rigid.so, exporting
    only the four pm_driver* entry points, and its loader manifest
    (config/driver-manifest.json.in's format; a relative library_path is
    resolved against the manifest's directory)."""
    tmp = DRIVER + ".tmp"
    cmd = [HIPCC, "-std=c++17", "-O2", "-fPIC", "-shared", "-fvisibility=hidden", "-Wall",
           "-Wno-unused-command-line-argument", DRIVER_SRC, "-o", tmp, "-L" + OUT, "-lpm_hip_rigid",
           "-Wl,-rpath,$ORIGIN", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("driver build failed")
    os.replace(tmp, DRIVER)
    with open(MANIFEST, "w") as f:
        f.write('{\n  "file_format_version": "1.0.0",\n  "PMD": {\n'
                '    "library_path": "libpm_hip_driver.so",\n    "api_version": "0.3.0",\n'
                '    "library_arch": "64"\n  }\n}\n')
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
