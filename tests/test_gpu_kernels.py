"""Kernel-level parity on the GPU, through the C ABI seam (include/emida_amd.h).

Restates the reference's kernel tests (test/sum_test.cpp, test/prepare_test.cpp,
test/cross_corr_test.cpp, test/maxarg_test.cpp, test/subpixel_max_test.cpp) and
compares every kernel with the CPU oracle on seeded random inputs: bit-exact for
integer work, sums, prepare, brute correlation (same fma order), arg-max, gather and
the double-precision fit; within a stated tolerance for the FFT correlation.
"""
import os

import numpy as np
import pytest

from emida_amd import _capi as C
from emida_amd import kernels as K
from emida_amd.gpu_offset import data_index_dtype
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ------------------------------------------------------------------ sums
def test_sum_slice_size_9():  # test/sum_test.cpp:53-66
    v = np.array([1, 3, 5, 8, 8, 10, 4, 1, 3, 2, 5, 15], np.float64)
    assert list(K.run_sum(v, [(0, 0), (1, 1)], (4, 3), (2, 2))) == [22, 21]


@pytest.mark.parametrize("src,sl,begins,batch", [
    ((31, 43), (29, 40), [(0, 0), (1, 1)], 1),
    ((64, 64), (32, 32), [(0, 0), (0, 16), (16, 16), (32, 32)], 1),
    ((64, 64), (32, 32), [(0, 0), (0, 16), (16, 16), (32, 32)], 97),
    ((23, 43), (21, 41), [(0, 0), (1, 1)], 8),
])
def test_sum_slice_fixture(src, sl, begins, batch):  # test/sum_test.cpp:99-134
    data = np.arange(src[0] * src[1] * batch, dtype=np.float64)
    np.testing.assert_array_equal(K.run_sum(data, begins, src, sl, batch), O.sums(data, begins, src, sl, batch))


@pytest.mark.parametrize("dt", [np.uint16, np.float32, np.float64])
@pytest.mark.parametrize("sl", [(32, 32), (64, 64), (128, 128), (120, 100)])
def test_sum_bitexact_vs_oracle(dt, sl):
    rng = np.random.default_rng(11)
    src = (300, 260)
    if dt == np.uint16:
        pics = rng.integers(0, 65536, size=src[0] * src[1] * 3, dtype=np.uint16)
    else:
        pics = (rng.standard_normal(src[0] * src[1] * 3) * 1000).astype(dt)
    begins = [(0, 0), (7, 13), (src[0] - sl[0], src[1] - sl[1])]
    got = K.run_sum(pics, begins, src, sl, 3)
    np.testing.assert_array_equal(got, O.sums(pics, begins, src, sl, 3))


# ------------------------------------------------------------------ prepare
PIC30 = np.arange(1, 31, dtype=np.uint16)


def test_prepare_no_window():  # test/prepare_test.cpp:47-78
    s = O.sums(PIC30.astype(np.float64), [(1, 2), (0, 1)], (6, 5), (5, 3))
    got = K.run_prepare_pics(PIC30, np.ones(6), np.ones(5), s, [(1, 2), (0, 1)], (6, 5), (5, 3))
    row = [-8, -7, -6, -5, -4, -2, -1, 0, 1, 2, 4, 5, 6, 7, 8]
    assert list(got) == row + row


def test_prepare_no_window_batch():  # test/prepare_test.cpp:80-125
    pic = np.concatenate([PIC30, PIC30.reshape(5, 6)[::-1].ravel()])
    got = K.run_prepare_pics(pic, np.ones(6), np.ones(5), np.array([330, 225, 0, 0], float), [(1, 2), (0, 1)],
                             (6, 5), (5, 3), batch_size=2)
    exp = [-8, -7, -6, -5, -4, -2, -1, 0, 1, 2, 4, 5, 6, 7, 8] * 2 + \
          [14, 15, 16, 17, 18, 8, 9, 10, 11, 12, 2, 3, 4, 5, 6] + \
          [19, 20, 21, 22, 23, 13, 14, 15, 16, 17, 7, 8, 9, 10, 11]
    assert list(got) == exp


def test_prepare_window():  # test/prepare_test.cpp:128-161
    wx = np.ones(6); wx[2] = 0.5
    wy = np.ones(5); wy[1] = 0.25
    s = O.sums(PIC30.astype(np.float64), [(1, 2), (0, 1)], (6, 5), (5, 3))
    got = K.run_prepare_pics(PIC30, wx, wy, s, [(1, 2), (0, 1)], (6, 5), (5, 3))
    row = [-8, -7, -3, -5, -4, -0.5, -.25, 0, .25, .5, 4, 5, 3, 7, 8]
    assert list(got) == row + row


def test_prepare_fft_padded():  # test/prepare_test.cpp:163-210
    s = O.sums(PIC30.astype(np.float64), [(1, 2), (0, 1)], (6, 5), (5, 3))
    got = K.run_prepare_pics(PIC30, np.ones(6), np.ones(5), s, [(1, 2), (0, 1)], (6, 5), (5, 3),
                             out_size=(10, 6)).reshape(2, 6, 10)
    inner = np.array([-8, -7, -6, -5, -4, -2, -1, 0, 1, 2, 4, 5, 6, 7, 8], float).reshape(3, 5)
    exp = np.zeros((6, 10)); exp[:3, :5] = inner
    for k in range(2):
        np.testing.assert_array_equal(got[k], exp)


@pytest.mark.parametrize("IN,T", [(np.uint16, np.float32), (np.uint16, np.float64), (np.float64, np.float64),
                                  (np.float32, np.float32)])
def test_prepare_bitexact_vs_oracle(IN, T):
    rng = np.random.default_rng(12)
    src, sl = (200, 150), (64, 48)
    if IN == np.uint16:
        pics = rng.integers(0, 65536, size=src[0] * src[1] * 2, dtype=np.uint16)
    else:
        pics = (rng.standard_normal(src[0] * src[1] * 2) * 1000).astype(IN)
    begins = [(0, 0), (100, 50), (136, 102)]
    sums = O.sums(pics, begins, src, sl, 2)
    hx, hy = O.hann(sl[0], T), O.hann(sl[1], T)
    got = K.run_prepare_pics(pics, hx, hy, sums, begins, src, sl, batch_size=2)
    np.testing.assert_array_equal(got, O.prepare(pics, hx, hy, sums, begins, src, sl, None, 2))


# ------------------------------------------------------------------ brute correlation
A33 = np.arange(1, 10, dtype=np.int32)
A43 = np.arange(1, 13, dtype=np.int32)
B43 = np.array([8, 7, 6, 5, 16, 15, 14, 13, 4, 3, 2, 1], np.int32)
A43b = np.array([3, 2, 1, 4, 2, 6, 10, 8, 9, 8, 11, 5], np.int32)
B43b = np.array([1, 2, 3, 4, 16, 5, 14, 13, 8, 3, 2, 1], np.int32)


def test_cross_corr_3x3():  # test/cross_corr_test.cpp:17-43
    exp = [9, 26, 50, 38, 21, 42, 94, 154, 106, 54, 90, 186, 285, 186, 90, 54, 106, 154, 94, 42, 21, 38, 50, 26, 9]
    assert list(K.run_cross_corr(A33, A33, (3, 3), (5, 5))) == exp


def test_cross_corr_3x3_res3x3():  # test/cross_corr_test.cpp:45-69
    assert list(K.run_cross_corr(A33, A33, (3, 3), (3, 3))) == [94, 154, 106, 186, 285, 186, 106, 154, 94]


def test_cross_corr_4x3_and_batches():  # test/cross_corr_test.cpp:71-202
    a, b = np.concatenate([A43, A43b]), np.concatenate([B43, B43b])
    np.testing.assert_array_equal(K.run_cross_corr(A43, B43, (4, 3), (7, 5)), O.cross_corr(A43, B43, (4, 3), (7, 5)))
    np.testing.assert_array_equal(K.run_cross_corr(a, b, (4, 3), (7, 5), ref_slices=2),
                                  O.cross_corr(a, b, (4, 3), (7, 5), ref_slices=2))
    np.testing.assert_array_equal(K.run_cross_corr(a, b, (4, 3), (5, 5), ref_slices=2),
                                  O.cross_corr(a, b, (4, 3), (5, 5), ref_slices=2))
    assert list(K.run_cross_corr(A43, B43, (4, 3), (7, 5))[:7]) == [1, 4, 10, 20, 25, 24, 16]


def test_cross_corr_64_fixture():  # test/cross_corr_test.cpp:205-217
    a, b, res = load("cross_corr_64_A.npy"), load("cross_corr_64_B.npy"), load("cross_corr_64_res.npy")
    np.testing.assert_array_equal(K.run_cross_corr(a.ravel(), b.ravel(), (64, 64), (127, 127)), res.ravel())


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("size,res,ref_slices,batch", [
    ((32, 32), (63, 63), 3, 2),
    ((17, 23), (33, 45), 2, 1),      # the x8 kernel's last thread of a row owns one shift
    ((64, 64), (9, 9), 2, 2),        # double: thread-per-shift kernel over LDS (x8 rows do not fit)
    ((128, 128), (255, 255), 1, 1),  # thread-per-shift kernel from HBM (slices larger than LDS)
    ((64, 48), (127, 95), 2, 17),    # two pattern tiles of the register-blocked kernel, the second partial
    ((40, 24), (81, 49), 1, 3),      # the widest window the reference allows (2 size + 1)
    ((24, 16), (32, 20), 2, 5),      # even window sizes
    ((100, 30), (199, 59), 1, 2),    # 3 waves along x-shifts
    ((136, 8), (15, 15), 1, 2),      # wider than 128: not the register-blocked kernel
    ((128, 4), (257, 7), 1, 1),      # window wider than 256: not the register-blocked kernel
])
def test_cross_corr_bitexact_vs_oracle(dt, size, res, ref_slices, batch):
    rng = np.random.default_rng(13)
    n = size[0] * size[1]
    pics = rng.standard_normal(n * ref_slices * batch).astype(dt)
    ref = rng.standard_normal(n * ref_slices).astype(dt)
    got = K.run_cross_corr(pics, ref, size, res, ref_slices, batch)
    np.testing.assert_array_equal(got, O.cross_corr(pics, ref, size, res, ref_slices, batch))


# Launches of at least 256 workgroups take the register-blocked kernel (16 patterns x
# a band of shifts per workgroup); the shapes cover 1, 2, 3 and 4 waves along the
# x-shifts, partial pattern tiles, even windows and the widest window allowed.
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("size,res,ref_slices,batch", [
    ((32, 32), (63, 63), 70, 1),
    ((64, 48), (127, 95), 12, 17),
    ((128, 128), (255, 255), 4, 1),
    ((40, 24), (81, 49), 37, 1),
    ((100, 30), (199, 59), 18, 1),
    ((24, 16), (32, 20), 130, 2),
])
def test_cross_corr_tiled_bitexact_vs_oracle(dt, size, res, ref_slices, batch):
    rng = np.random.default_rng(17)
    n = size[0] * size[1]
    pics = rng.standard_normal(n * ref_slices * batch).astype(dt)
    ref = rng.standard_normal(n * ref_slices).astype(dt)
    got = K.run_cross_corr(pics, ref, size, res, ref_slices, batch)
    np.testing.assert_array_equal(got, O.cross_corr(pics, ref, size, res, ref_slices, batch))


def test_cross_corr_int32_wraps_like_oracle():
    # the reference instantiates the brute kernel for int (test/cross_corr_test.cpp); sums wrap mod 2^32
    rng = np.random.default_rng(2)
    size, res = (48, 40), (95, 79)
    for nb, batch in ((1, 3), (16, 17)):  # thread-per-shift and register-blocked kernels
        pics = rng.integers(-40000, 40000, size=size[0] * size[1] * nb * batch, dtype=np.int32)
        ref = rng.integers(-40000, 40000, size=size[0] * size[1] * nb, dtype=np.int32)
        np.testing.assert_array_equal(K.run_cross_corr(pics, ref, size, res, nb, batch),
                                      O.cross_corr(pics, ref, size, res, nb, batch))


# ------------------------------------------------------------------ FFT correlation
@pytest.mark.parametrize("s", [2, 4, 8, 16, 32, 64, 3, 5, 6, 20, 33, 48, 63])
@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_cross_corr_fft_vs_oracle(s, dt, tol):
    # sides that are not powers of two: padded transforms, reference (2s)^2 layout
    rng = np.random.default_rng(14 + s)
    ref_slices, batch = 3, 2
    pics = rng.standard_normal(s * s * ref_slices * batch).astype(dt)
    ref = rng.standard_normal(s * s * ref_slices).astype(dt)
    got = K.run_cross_corr_fft(pics, ref, (s, s), ref_slices, batch).reshape(-1, 4 * s * s)
    exp = O.fft_corr(pics, ref, (s, s), ref_slices, batch).reshape(-1, 4 * s * s).astype(np.float64)
    # error relative to the largest correlation magnitude of each slice
    scale = np.abs(exp).max(axis=1, keepdims=True)
    assert (np.abs(got - exp) / scale).max() < tol * max(1, np.log2(2 * s))


@pytest.mark.parametrize("s,ref_slices,batch", [(128, 2, 3), (128, 1, 1), (96, 2, 2), (65, 1, 2), (127, 1, 1)])
def test_cross_corr_fft_128_float_vs_oracle(s, ref_slices, batch):
    # slices of side 65..128 (256-point transforms) run as two kernels with a work buffer
    rng = np.random.default_rng(99 + batch)
    pics = rng.standard_normal(s * s * ref_slices * batch).astype(np.float32)
    ref = rng.standard_normal(s * s * ref_slices).astype(np.float32)
    got = K.run_cross_corr_fft(pics, ref, (s, s), ref_slices, batch).reshape(-1, 4 * s * s)
    exp = O.fft_corr(pics, ref, (s, s), ref_slices, batch).reshape(-1, 4 * s * s).astype(np.float64)
    scale = np.abs(exp).max(axis=1, keepdims=True)
    assert (np.abs(got - exp) / scale).max() < 2e-6 * np.log2(2 * s)


def _rocfft_corr(pics, ref, s, ref_slices, batch):
    """The reference's R2C (2s.x x 2s.y, zero-padded), p * conj(r) and unnormalised C2R
    (src/gpu_offset.hpp:303-316, src/cufft_helpers.hpp:105-155) with rocFFT through
    the system hipFFT library, in the HIP runtime the library under test uses.  The
    plan's dimensions are (rows, columns) = (2 s.y, 2 s.x)."""
    import ctypes
    from emida_amd.gpu_offset import DeviceBuffer
    fft = ctypes.CDLL("libhipfft.so")
    dt = pics.dtype
    double = dt == np.float64
    cdt = np.complex128 if double else np.complex64
    sx, sy = (s, s) if np.isscalar(s) else s
    n, mx, my = ref_slices * batch, 2 * sx, 2 * sy
    nc = my * (mx // 2 + 1)
    P = np.zeros((n, my, mx), dt)
    P[:, :sy, :sx] = pics.reshape(n, sy, sx)
    R = np.zeros((n, my, mx), dt)
    R[:, :sy, :sx] = np.tile(ref.reshape(ref_slices, sy, sx), (batch, 1, 1))  # slice i uses ref i % ref_slices

    def plan(kind):
        h = ctypes.c_void_p()
        dims = (ctypes.c_int * 2)(my, mx)
        assert fft.hipfftPlanMany(ctypes.byref(h), 2, dims, None, 1, 0, None, 1, 0, kind, n) == 0
        return h

    r2c, c2r = plan(0x6A if double else 0x2A), plan(0x6C if double else 0x2C)  # D2Z/R2C, Z2D/C2R
    fwd = fft.hipfftExecD2Z if double else fft.hipfftExecR2C
    inv = fft.hipfftExecZ2D if double else fft.hipfftExecC2R
    spectra = []
    for a in (P, R):
        d_in, d_out = DeviceBuffer.from_array(a), DeviceBuffer(n * nc * np.dtype(cdt).itemsize)
        assert fwd(r2c, d_in.ptr, d_out.ptr) == 0
        spectra.append(d_out.download(cdt))
    H = (spectra[0] * np.conj(spectra[1])).astype(cdt)
    d_h, d_c = DeviceBuffer.from_array(H), DeviceBuffer(n * mx * my * dt.itemsize)
    assert inv(c2r, d_h.ptr, d_c.ptr) == 0
    out = d_c.download(dt).reshape(n, my, mx)
    fft.hipfftDestroy(r2c)
    fft.hipfftDestroy(c2r)
    return out


@pytest.mark.parametrize("s", [64, 20, 33, 128, 96, 160, (64, 48), (20, 100), (130, 9)])
@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 3e-6)])
def test_cross_corr_fft_vs_rocfft(s, dt, tol):
    # cross-check of the FFT correlation against rocFFT (used in this test only),
    # every path: one kernel, the 256-point float kernels and the three-kernel path
    # (double sides above 64, sides above 128, non-square slices)
    sx, sy = (s, s) if np.isscalar(s) else s
    rng = np.random.default_rng(31 + sx + 7 * sy)
    ref_slices, batch = 3, 2
    n = ref_slices * batch
    pics = rng.standard_normal(sx * sy * n).astype(dt)
    ref = rng.standard_normal(sx * sy * ref_slices).astype(dt)
    got = K.run_cross_corr_fft(pics, ref, (sx, sy), ref_slices, batch).reshape(n, 2 * sy, 2 * sx)
    exp = _rocfft_corr(pics, ref, (sx, sy), ref_slices, batch).astype(np.float64)
    scale = np.abs(exp).reshape(n, -1).max(axis=1)[:, None, None]
    assert (np.abs(got.astype(np.float64) - exp) / scale).max() < tol * np.log2(2 * max(sx, sy))


def _numpy_fft_corr(pics, ref, size, ref_slices, batch):
    """(2 s.x) x (2 s.y) unnormalised circular correlation of zero-padded slices by
    numpy's FFT in double: irfft2(rfft2(p) conj(rfft2(r))) x 4 s.x s.y."""
    sx, sy = size
    p = pics.astype(np.float64).reshape(batch, ref_slices, sy, sx)
    r = ref.astype(np.float64).reshape(1, ref_slices, sy, sx)
    P = np.fft.rfft2(p, s=(2 * sy, 2 * sx))
    R = np.fft.rfft2(r, s=(2 * sy, 2 * sx))
    return (np.fft.irfft2(P * np.conj(R), s=(2 * sy, 2 * sx)) * (4 * sx * sy)).reshape(batch * ref_slices, -1)


@pytest.mark.parametrize("size,ref_slices,batch", [
    ((96, 96), 2, 2), ((128, 128), 2, 3), ((65, 65), 1, 2),       # double: 256-point transforms
    ((160, 160), 2, 2), ((256, 256), 1, 2), ((129, 129), 1, 1),  # 512-point transforms
    ((64, 48), 3, 2), ((48, 64), 2, 3), ((20, 100), 2, 2), ((200, 3), 1, 3), ((2, 2), 1, 1)])
@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 3e-6)])
def test_cross_corr_fft_three_kernel_path_vs_numpy(size, ref_slices, batch, dt, tol):
    # the shapes of the three-kernel path (and a few single-kernel ones beside them)
    # against numpy's double FFT of the same zero-padded slices
    rng = np.random.default_rng(size[0] * 3 + size[1] + batch)
    a = size[0] * size[1]
    pics = rng.standard_normal(a * ref_slices * batch).astype(dt)
    ref = rng.standard_normal(a * ref_slices).astype(dt)
    got = K.run_cross_corr_fft(pics, ref, size, ref_slices, batch).reshape(ref_slices * batch, -1)
    exp = _numpy_fft_corr(pics, ref, size, ref_slices, batch)
    scale = np.abs(exp).max(axis=1, keepdims=True)
    assert (np.abs(got.astype(np.float64) - exp) / scale).max() < tol * np.log2(2 * max(size))


@pytest.mark.parametrize("size", [(20, 12), (12, 20), (64, 48), (33, 5)])
@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-12), (np.float32, 3e-6)])
def test_cross_corr_fft_non_square_vs_oracle(size, dt, tol):
    rng = np.random.default_rng(size[0] + 11 * size[1])
    ref_slices, batch = 2, 2
    a = size[0] * size[1]
    pics = rng.standard_normal(a * ref_slices * batch).astype(dt)
    ref = rng.standard_normal(a * ref_slices).astype(dt)
    got = K.run_cross_corr_fft(pics, ref, size, ref_slices, batch).reshape(ref_slices * batch, -1)
    exp = O.fft_corr(pics, ref, size, ref_slices, batch).reshape(ref_slices * batch, -1).astype(np.float64)
    scale = np.abs(exp).max(axis=1, keepdims=True)
    assert (np.abs(got.astype(np.float64) - exp) / scale).max() < tol * np.log2(2 * max(size))


def test_cross_corr_fft_smoke_4x4():  # test/cross_corr_fft_test.cpp:8-42 (same inputs, with an assertion)
    pic = np.array([[1, 2, 3, 1], [4, 5, 6, 1], [7, 8, 9, 1], [7, 8, 9, 1]], float)
    ref = np.array([[6, 3, 2, 1], [12, 23, 3, 1], [1, 8, 9, 1], [7, 8, 9, 1]], float)
    got = K.run_cross_corr_fft(pic.ravel(), ref.ravel(), (4, 4))
    P = np.fft.rfft2(pic, s=(8, 8))
    R = np.fft.rfft2(ref, s=(8, 8))
    exp = np.fft.irfft2(P * np.conj(R), s=(8, 8)) * 64
    np.testing.assert_allclose(got.reshape(8, 8), exp, atol=1e-10)


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_hadamard_vs_numpy(dt):  # src/hadamard.cu:27-41
    rng = np.random.default_rng(15)
    one, ref_slices, batch = (5, 8), 3, 4
    n = one[0] * one[1] * ref_slices
    ref = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dt)
    pics = (rng.standard_normal(n * batch) + 1j * rng.standard_normal(n * batch)).astype(dt)
    got = K.run_hadamard(pics, ref, one, ref_slices, batch)
    exp = pics * np.conj(np.tile(ref, batch))
    np.testing.assert_allclose(got, exp, rtol=1e-6 if dt == np.complex64 else 1e-15)


# ------------------------------------------------------------------ arg-max
def _saw(n):
    out, mom = [], 0.0
    for _ in range(n):
        out.append(mom)
        mom += 0.007
        if mom > 1:
            mom = -2
    return np.array(out)


def _pairs(m):
    return [(float(x["data"]), int(x["index"])) for x in m]


def test_maxarg_known_answers():  # test/maxarg_test.cpp:63-230
    assert _pairs(K.run_maxarg_reduce(np.array([1, 3, 5, 8, 10, 4, 3, 2], float), (4, 2))) == [(10.0, 4)]
    v = np.array([1, 3, 0, 8, 10, 4, 0, 3, 0, 0, 0, 0, 1, 1, 0, 1], float)
    assert _pairs(K.run_maxarg_reduce(np.concatenate([v, v]), (4, 4), 2)) == [(10.0, 4), (10.0, 4)]
    v = np.array([1, 3, 5, 8, 10, 4, 3, 2, 1, 12, 15, 18, 10, 4, 3, 2], float)
    assert _pairs(K.run_maxarg_reduce(v, (4, 2), 2)) == [(10.0, 4), (18.0, 3)]
    d = _saw(1333); d[1001] = 3
    assert _pairs(K.run_maxarg_reduce(d, (31, 43))) == [(3.0, 1001)]
    d = _saw(3999); d[1001], d[1024], d[1333], d[3000] = 3, 4, 4, 4
    assert _pairs(K.run_maxarg_reduce(d, (31, 43), 3)) == [(4.0, 1024), (4.0, 0), (4.0, 334)]
    d = _saw(4096); d[1024] = 3
    assert _pairs(K.run_maxarg_reduce(d, (31, 43))) == [(3.0, 1024)]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_maxarg_vs_oracle_with_ties(dt):
    rng = np.random.default_rng(16)
    d = rng.integers(0, 40, size=9 * 65025).astype(dt)  # many ties
    got = K.run_maxarg_reduce(d, (255, 255), 9)
    exp = O.maxarg(d, 65025, 9)
    assert _pairs(got) == _pairs(exp)


# ------------------------------------------------------------------ gather
def test_extract_known_answers():  # test/maxarg_test.cpp:8-61
    v = np.arange(50, dtype=float)
    mi = np.array([(18.0, 18), (6.0, 6)], data_index_dtype(np.float64))
    got = K.run_extract_neighbors(v, mi, 3, (5, 5), 2)
    assert list(got) == [12, 13, 14, 17, 18, 19, 22, 23, 24, 25, 26, 27, 30, 31, 32, 35, 36, 37]


@pytest.mark.parametrize("policy", [C.CROSS_POLICY_BRUTE, C.CROSS_POLICY_FFT])
@pytest.mark.parametrize("s", [3, 5, 7, 9])
def test_extract_vs_oracle_including_edges(policy, s):
    rng = np.random.default_rng(17 + s)
    size = (13, 13)
    area = 14 * 14 if policy == C.CROSS_POLICY_FFT else 13 * 13
    n = 40
    data = rng.standard_normal(area * n)
    idx = rng.integers(0, area, n).astype(np.uint32)
    idx[:6] = [0, area - 1, 12, 13 * 12, 13, area // 2]  # corners and edges
    mi = np.zeros(n, data_index_dtype(np.float64))
    mi["index"] = idx
    got = K.run_extract_neighbors(data, mi, s, size, n, policy)
    np.testing.assert_array_equal(got, O.extract(data, mi, s, size, n, policy))


# ------------------------------------------------------------------ fit
def test_subpixel_known_answers():  # test/subpixel_max_test.cpp:21-46
    o, c = K.run_subpixel_max(np.arange(1, 10, dtype=float), 3)
    np.testing.assert_allclose(c[0], [1, 1, 3, 0, 0, 0], atol=1e-15)
    o, c = K.run_subpixel_max(np.array([1, 2, 3, 7, 9, 8, 4, 5, 6], float), 3)
    eo, ec = O.subpixel(np.array([1, 2, 3, 7, 9, 8, 4, 5, 6], float), 3)
    np.testing.assert_array_equal(c, ec)
    np.testing.assert_array_equal(o, eo)


@pytest.mark.parametrize("s", [3, 5, 7, 9])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_subpixel_bitexact_vs_oracle(s, dt):
    rng = np.random.default_rng(18 + s)
    nb = rng.standard_normal(200 * s * s).astype(dt)
    o, c = K.run_subpixel_max(nb, s)
    eo, ec = O.subpixel(nb, s)
    np.testing.assert_array_equal(c, ec)
    np.testing.assert_array_equal(o, eo)
