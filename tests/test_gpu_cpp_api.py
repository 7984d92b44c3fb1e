"""Runs the C++ API test driver (tests/cpp/test_api.cpp): the Dense /
Workspace / registry / lsqr_solve surface exercised as the reference's own
C++ tests exercise it."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(__file__), "cpp", "test_api")


def test_cpp_api_driver():
    assert os.path.exists(BIN), "build first: __graft_entry__.build()"
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0, p.stdout[-3000:]
    assert "FAIL" not in p.stdout
