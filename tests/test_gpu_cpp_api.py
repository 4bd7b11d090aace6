"""Runs the C++ drop-in API tests (tests/cpp, built by build()):
  * test_gpu_offset: the reference's end-to-end tests written against
    include/emida/gpu_offset.hpp and the reference's threaded loader / GPU / finalize
    pipeline;
  * test_cross_corr_fft: the reference's cross_corr_fft test (test/cross_corr_fft_test.cpp)
    with assertions, get_cu_pic / get_cu_temp and the public streams;
  * test_threads: the request server under four submitting threads and a statistics
    reader (every served result equal to the same request served alone) and the file
    pipeline with four loaders, results in completion order and one unreadable file."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINS = [os.path.join(ROOT, "tests", "cpp", n) for n in ("test_gpu_offset", "test_cross_corr_fft", "test_threads")]


@pytest.mark.gpu
@pytest.mark.parametrize("binary", BINS, ids=os.path.basename)
def test_cpp_drop_in_api(binary):
    if not os.path.exists(binary):
        pytest.fail(f"{os.path.relpath(binary, ROOT)} is not built (run __graft_entry__.build())")
    p = subprocess.run([binary], cwd=ROOT, capture_output=True, text=True, timeout=110)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0, p.stderr


def test_cpp_api_tests_are_built():
    for b in BINS:
        assert os.path.exists(b), f"{os.path.relpath(b, ROOT)} missing: build() compiles it"
