"""Runs the C++ drop-in API test (tests/cpp/test_gpu_offset.cpp, built by build()):
the reference's end-to-end tests written against include/emida/gpu_offset.hpp and
the reference's threaded loader / GPU / finalize pipeline."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_gpu_offset")


@pytest.mark.gpu
def test_cpp_drop_in_api():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/test_gpu_offset is not built (run __graft_entry__.build())")
    p = subprocess.run([BIN], cwd=ROOT, capture_output=True, text=True, timeout=110)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0, p.stderr


def test_cpp_api_test_is_built():
    assert os.path.exists(BIN), "tests/cpp/test_gpu_offset missing: build() compiles it"
