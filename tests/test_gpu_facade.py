"""The reference's C++ APIs (pgmcore::Engine, pgmcuda::Engine method sets) through include/pgm_hip.hpp, from a
C++ program compiled against the C ABI, checked against the FP64 oracle inside that program."""

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_facade")


def test_cpp_facade():
    assert os.path.exists(BIN), "tests/cpp/test_facade is built by __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "PASSED" in r.stdout
