"""Runs the C++ test of the ITK-style layer (tests/cpp/skm_itk_api_test.cpp,
built by __graft_entry__.build()) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build", "skm_itk_api_test")


def test_cpp_itk_layer():
    assert os.path.exists(BIN), "build the C++ tests with __graft_entry__.build()"
    res = subprocess.run([BIN], capture_output=True, text=True, timeout=170)
    print(res.stdout)
    print(res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
