"""Runs the C++ drop-in test (tests/cpp/test_dft.cpp, built by
__graft_entry__.build()) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_dft")


def test_cpp_drop_in_interface():
    assert os.path.exists(BIN), "tests/cpp/test_dft is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "DFT check passed" in r.stdout
