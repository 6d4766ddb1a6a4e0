"""Builds libpa_device.so (gfx950) in-tree with hipcc.

    python -m pairwise_aligner_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpa_device.so")
SOURCES = ["pa_kernels.hip", "pa_encode.hip", "pa_learned.hip", "pa_capi.cpp", "pa_plan.cpp", "pa_fastx.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-shared",
    # no packed-FP32 instructions anywhere (the kernels are integer-only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unknown-warning-option", "-pthread",
]


def hipcc() -> str:
    for cand in [os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"]:
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           *[os.path.join(CSRC, s) for s in SOURCES], "-o", OUT + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libpa_device.so")
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


PROBE_SRC = os.path.join(CSRC, "valu_probe.hip")
PROBE_OUT = os.path.join(HERE, "valu_probe")


def build_probe() -> str:
    """Compiles the VALU issue-rate probe (csrc/valu_probe.hip), run by `bench.py --valu-probe`."""
    flags = [f for f in HIPCC_FLAGS if f not in ("-fPIC", "-shared")]
    cmd = [hipcc(), *flags, PROBE_SRC, "-o", PROBE_OUT + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building valu_probe")
    os.replace(PROBE_OUT + ".tmp", PROBE_OUT)
    return PROBE_OUT


def build_cpp_tests() -> str:
    """Compiles tests/cpp/test_host_api (the header-only C++ host API against libpa_device.so) and
    tests/cpp/test_planner (the host planner on its own, no device code)."""
    src = os.path.join(ROOT, "tests", "cpp", "test_host_api.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "test_host_api")
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), src,
           "-L", HERE, "-lpa_device", "-Wl,-rpath,$ORIGIN/../../pairwise_aligner_amd",
           "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building tests/cpp/test_host_api")
    planner = os.path.join(ROOT, "tests", "cpp", "test_planner")
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-pthread", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           os.path.join(ROOT, "tests", "cpp", "test_planner.cpp"), os.path.join(CSRC, "pa_plan.cpp"), "-o", planner]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building tests/cpp/test_planner")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_cpp_tests())
    print(build_probe())
