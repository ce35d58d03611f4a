"""The header-only C++ facades (include/tetmag_gpu.hpp) driven by a C++
program the way TheLLG / DemagField call the reference classes, checked
against the C oracle inside the program (tests/facade/facade_test.cpp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "facade", "facade_test")


def test_cpp_facades(gpu_available):
    if not os.path.exists(BIN):
        pytest.fail("tests/facade/facade_test not built: run __graft_entry__.build()")
    r = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "PASSED" in r.stdout
