"""Runs the C++ host API test (tests/cpp/test_host_api.cpp, built by __graft_entry__.build()):
the reference's known answers written against the header-only C++20 API."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(__file__), "cpp", "test_host_api")


def _run(*args):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing; run __graft_entry__.build()")
    res = subprocess.run([BIN, *args], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    return res


def test_cpp_configuration_layer():
    res = _run("--config")
    assert res.returncode == 0 and "OK" in res.stdout


@pytest.mark.gpu
def test_cpp_known_answers_on_gpu():
    res = _run()
    assert res.returncode == 0, res.stdout
