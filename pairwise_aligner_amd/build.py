"""Builds libpa_device.so (gfx950) in-tree with hipcc.

    python -m pairwise_aligner_amd.build

Every source is compiled to an object of its own, in parallel: the DP kernels once per
(lane width, rows per lane) from csrc/pa_kernels_inst.hip, the rest once each; then one
hipcc -shared link.  Objects go to pairwise_aligner_amd/_build/.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libpa_device.so")
SOURCES = ["pa_kernels.hip", "pa_fused.hip", "pa_encode.hip", "pa_learned.hip", "pa_noalloc.hip", "pa_capi.cpp",
           "pa_plan.cpp", "pa_fastx.cpp"]
# (lane width, rows per lane) instantiations of the DP kernels
INSTANCES = [(w, r) for w in (16, 32) for r in (4, 6, 8, 10, 12, 14, 16)]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    # no packed-FP32 instructions anywhere (the kernels are integer-only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unknown-warning-option", "-pthread",
]


def hipcc() -> str:
    for cand in [os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"]:
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _jobs():
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    jobs = []
    for s in SOURCES:
        if not os.path.exists(os.path.join(CSRC, s)):
            continue
        obj = os.path.join(OBJ, s.replace(".", "_") + ".o")
        jobs.append((obj, [hipcc(), *HIPCC_FLAGS, *inc, "-c", os.path.join(CSRC, s), "-o", obj]))
    for w, r in INSTANCES:
        obj = os.path.join(OBJ, f"pa_kernels_w{w}_r{r}.o")
        jobs.append((obj, [hipcc(), *HIPCC_FLAGS, *inc, f"-DPA_INST_W={w}", f"-DPA_INST_R={r}", "-c",
                           os.path.join(CSRC, "pa_kernels_inst.hip"), "-o", obj]))
    return jobs


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = _jobs()
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1)), 16))

    def compile_one(job):
        obj, cmd = job
        return obj, subprocess.run(cmd, capture_output=True, text=True)

    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        results = list(ex.map(compile_one, jobs))
    failed = [(obj, res) for obj, res in results if res.returncode != 0]
    for obj, res in failed:
        sys.stderr.write(f"--- {os.path.basename(obj)}\n{res.stdout}{res.stderr}")
    if failed:
        raise RuntimeError("hipcc failed building libpa_device.so")
    if verbose:
        for _, res in results:
            sys.stderr.write(res.stderr)
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", *[obj for obj, _ in results],
           "-o", OUT + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libpa_device.so")
    os.replace(OUT + ".tmp", OUT)
    return OUT


def build_cpp_tests() -> str:
    """Compiles tests/cpp/test_host_api (the header-only C++ host API against libpa_device.so) and
    tests/cpp/test_planner (the host planner on its own, no device code)."""
    src = os.path.join(ROOT, "tests", "cpp", "test_host_api.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "test_host_api")
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), src,
           "-L", HERE, "-lpa_device", "-Wl,-rpath,$ORIGIN/../../pairwise_aligner_amd",
           "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-o", out]
    planner = os.path.join(ROOT, "tests", "cpp", "test_planner")
    cmd2 = ["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-pthread", "-I", os.path.join(ROOT, "include"), "-I",
            CSRC, os.path.join(ROOT, "tests", "cpp", "test_planner.cpp"), os.path.join(CSRC, "pa_plan.cpp"), "-o",
            planner]
    with concurrent.futures.ThreadPoolExecutor(2) as ex:
        r1, r2 = ex.map(lambda c: subprocess.run(c, capture_output=True, text=True), [cmd, cmd2])
    if r1.returncode != 0:
        sys.stderr.write(r1.stdout + r1.stderr)
        raise RuntimeError("g++ failed building tests/cpp/test_host_api")
    if r2.returncode != 0:
        sys.stderr.write(r2.stdout + r2.stderr)
        raise RuntimeError("g++ failed building tests/cpp/test_planner")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_cpp_tests())
