"""Runs the C++ facade test program (tests/cpp/test_facades.cpp, built by the
library Makefile into spectrum_amd/lib/test_facades): the reference's own
facade-level test cases against the drop-in C++ classes."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "spectrum_amd", "lib",
                   "test_facades")


def test_cpp_facades():
    if not os.path.exists(BIN):
        raise RuntimeError(f"{BIN} missing: build with __graft_entry__.build()")
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "ALL PASSED" in p.stdout
