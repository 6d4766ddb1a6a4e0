"""Runs the host planner test (tests/cpp/test_planner.cpp, built by __graft_entry__.build()): plans
of seeded random batches checked against the invariants the kernels rely on, on the CPU."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(__file__), "cpp", "test_planner")


def test_planner_invariants():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing; run __graft_entry__.build()")
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(res.stdout, res.stderr)
    assert res.returncode == 0 and "OK: 0 failure(s)" in res.stdout
