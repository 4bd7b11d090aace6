"""The allocation-free FFT correlation entry point (emida_run_cross_corr_fft_noalloc).

It has no counterpart in the reference; it serves callers that cannot allocate device
memory.  Its results must equal emida_run_cross_corr_fft's (the same kernels, working
in a static device array instead of allocated buffers) bit for bit, on consecutive
calls that reuse the array with different sizes, and sizes beyond the array must be
refused before any work.
"""
import numpy as np
import pytest

import emida_amd
from emida_amd import kernels as K

pytestmark = pytest.mark.gpu

# (T, slice side, ROIs, patterns of the first call, ROIs and patterns of the second)
CASES = [
    (np.float32, 32, 3, 4, (2, 3)),
    (np.float64, 64, 2, 3, (1, 2)),
    (np.float32, 20, 4, 2, (3, 1)),
    (np.float32, 96, 5, 3, (4, 2)),     # 256-point path
    (np.float32, 128, 40, 3, (36, 2)),  # 256-point path, 10 MiB of reference spectra
    (np.float64, 128, 4, 3, (3, 1)),    # three-kernel path (double 256-point transforms)
    (np.float32, (64, 48), 5, 2, (2, 2)),  # three-kernel path (non-square)
    (np.float64, 256, 12, 2, (5, 1)),   # three-kernel path, 512-point transforms in several chunks
]


def _call(T, side, nb, npat, noalloc, seed):
    rng = np.random.default_rng(seed)
    size = (side, side) if np.isscalar(side) else side
    a = size[0] * size[1]
    p = (rng.standard_normal(a * nb * npat) * 100).astype(T)
    r = (rng.standard_normal(a * nb) * 100).astype(T)
    dp, dr = emida_amd.DeviceBuffer.from_array(p), emida_amd.DeviceBuffer.from_array(r)
    res = emida_amd.DeviceBuffer(4 * a * nb * npat * np.dtype(T).itemsize)
    K.run_cross_corr_fft_device(dp, dr, res, T, size, nb, npat, noalloc=noalloc)
    emida_amd.synchronize()
    return res.download(T)


@pytest.mark.parametrize("T,side,nb,npat,second", CASES)
def test_noalloc_equals_allocating_entry_point(T, side, nb, npat, second):
    for i, (n_roi, n_pat) in enumerate([(nb, npat), second]):
        exp = _call(T, side, n_roi, n_pat, False, seed=7 + i)
        got = _call(T, side, n_roi, n_pat, True, seed=7 + i)
        np.testing.assert_array_equal(got, exp, err_msg=f"call {i + 1}: {n_roi} ROIs x {n_pat} patterns")
        assert np.isfinite(got).all()


def test_noalloc_refuses_sizes_beyond_its_scratch():
    cap = K.fft_noalloc_scratch_bytes()
    assert cap == 32 << 20
    # reference spectra of 40 MiB (double, 128-point), 52 MiB (float, 256-point) and the
    # three-kernel path's spectra plus work buffer; the buffers have the sizes the call
    # reads and writes, so a call that were not refused would still stay in bounds
    for T, side, nb in [(np.float64, 64, 300), (np.float32, 128, 200), (np.float64, 256, 16)]:
        n = side * side * nb * np.dtype(T).itemsize
        pics, ref, res = emida_amd.DeviceBuffer(n), emida_amd.DeviceBuffer(n), emida_amd.DeviceBuffer(4 * n)
        with pytest.raises(emida_amd.EmidaError) as e:
            K.run_cross_corr_fft_device(pics, ref, res, T, (side, side), nb, 1, noalloc=True)
        assert e.value.code == emida_amd._capi.ERR_ARG and "static scratch" in str(e.value)
