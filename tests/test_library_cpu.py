"""CPU-only checks of the HIP library: it loads, exports every declared entry point,
and its host-side pieces (tables, windows, argument validation) agree with the
oracle.  No GPU call is made here."""
import ctypes

import numpy as np
import pytest

import emida_amd
from emida_amd import _capi as C
from emida_amd import kernels as K
from emida_amd import synth
from oracle import oracle as O


def test_library_exports_every_declared_symbol():
    L = C.lib()
    names = C.exported_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("s", [3, 5, 7, 9])
def test_lstsq_table_matches_oracle_bitwise(s):
    np.testing.assert_array_equal(K.lstsq_matrix(s), O.lstsq_matrix(s))


@pytest.mark.parametrize("n", [2, 5, 32, 64, 128, 101])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_hanning_matches_oracle_bitwise(n, dt):
    np.testing.assert_array_equal(K.generate_hanning(n, dt), O.hann(n, dt))


def _create(**kw):
    args = dict(T=np.float64, IN=np.uint16, src_size=(128, 128), begins=[(0, 0)], slice_size=(32, 32),
                cross_size=(63, 63), batch_size=1, policy=emida_amd.CROSS_POLICY_BRUTE, s=3)
    args.update(kw)
    return emida_amd.GpuOffset(**args)


def test_create_valid_configurations_without_gpu():
    g = _create()
    assert g.total_slices == 1 and g.cross_size == (63, 63) and g.neigh_size == 9
    g = _create(policy=emida_amd.CROSS_POLICY_FFT, cross_size=(5, 5), s=5, batch_size=3, begins=[(0, 0), (10, 20)])
    # the FFT policy forces cross = 2s - 1 (src/gpu_offset.hpp:114)
    assert g.cross_size == (63, 63) and g.corr_buffer == (64, 64) and g.fft_size == 64
    assert g.total_slices == 6 and g.neigh_size == 150
    assert g.data_index.itemsize == 16 and _create(T=np.float32).data_index.itemsize == 8


@pytest.mark.parametrize("kw,code", [
    (dict(s=4), C.ERR_ARG),
    (dict(begins=[(100, 0)]), C.ERR_ARG),               # ROI does not fit
    (dict(cross_size=(64, 63)), C.ERR_ARG),             # > 2s-1
    (dict(batch_size=0), C.ERR_ARG),
    (dict(begins=[]), C.ERR_ARG),
    (dict(IN=np.int32), C.ERR_ARG),
    (dict(policy=emida_amd.CROSS_POLICY_FFT, slice_size=(257, 16), src_size=(300, 300)), C.ERR_UNSUPPORTED),
    (dict(policy=emida_amd.CROSS_POLICY_FFT, T=np.float32, slice_size=(16, 257), src_size=(300, 300)),
     C.ERR_UNSUPPORTED),
    (dict(policy=emida_amd.CROSS_POLICY_BRUTE_MFMA, mfma_refine=False), C.ERR_ARG),  # float screen, T double
    (dict(policy=emida_amd.CROSS_POLICY_BRUTE_MFMA, T=np.float32, slice_size=(36, 32)), C.ERR_UNSUPPORTED),
    (dict(policy=emida_amd.CROSS_POLICY_BRUTE_MFMA, T=np.float32, mfma_mode=7), C.ERR_ARG),
    (dict(policy=7), C.ERR_ARG),
])
def test_create_rejects_invalid_arguments(kw, code):
    with pytest.raises(emida_amd.EmidaError) as e:
        _create(**kw)
    assert e.value.code == code


def test_fft_limits():
    assert emida_amd.fft_max_slice(np.float32) == 256 and emida_amd.fft_max_slice(np.float64) == 256
    g = _create(policy=emida_amd.CROSS_POLICY_FFT, T=np.float32, slice_size=(128, 128), cross_size=(1, 1))
    assert g.cross_size == (255, 255) and g.fft_size == 256
    # any side: transforms padded to the next power of two >= 2s, reference buffer 2s
    for T, side, m in [(np.float64, 48, 128), (np.float64, 3, 8), (np.float32, 96, 256), (np.float32, 65, 256),
                       (np.float64, 64, 128), (np.float32, 33, 128), (np.float64, 128, 256), (np.float64, 65, 256),
                       (np.float64, 256, 512), (np.float32, 129, 512), (np.float32, 256, 512)]:
        g = _create(policy=emida_amd.CROSS_POLICY_FFT, T=T, slice_size=(side, side), src_size=(300, 300))
        assert g.fft_size == m and g.corr_buffer == (2 * side, 2 * side) and g.cross_size == (2 * side - 1,) * 2
    # non-square slices: (2 s.x) x (2 s.y) buffers, fft_size is the x axis' transform
    for T, sl, m in [(np.float64, (64, 48), 128), (np.float32, (20, 100), 64), (np.float32, (200, 3), 512)]:
        g = _create(policy=emida_amd.CROSS_POLICY_FFT, T=T, slice_size=sl, src_size=(300, 300))
        assert g.fft_size == m and g.corr_buffer == (2 * sl[0], 2 * sl[1])
        assert g.cross_size == (2 * sl[0] - 1, 2 * sl[1] - 1)


def test_create_matrix_core_policy_without_gpu():
    for mode in (emida_amd.MFMA_BF16, emida_amd.MFMA_BF16_SPLIT, emida_amd.MFMA_F16, emida_amd.MFMA_F16_SPLIT):
        g = _create(T=np.float32, policy=emida_amd.CROSS_POLICY_BRUTE_MFMA, mfma_mode=mode, cross_size=(21, 41))
        assert g.cross_size == (21, 41) and g.neigh_size == 9 and g.mfma_mode == mode
        g = _create(T=np.float64, policy=emida_amd.CROSS_POLICY_BRUTE_MFMA, mfma_mode=mode)  # refined: double
        assert g.data_index.itemsize == 16


def test_fp16_matrix_core_modes_need_uint16_patterns():
    # a float or double pattern can hold prepared values beyond fp16's range (131008
    # after the halving), which would put inf into the screen: those patterns default
    # to a bf16 mode and an explicit fp16 mode is refused before any device work
    for IN in (np.float32, np.float64):
        g = _create(T=np.float64, IN=IN, policy=emida_amd.CROSS_POLICY_BRUTE_MFMA)
        assert g.mfma_mode == emida_amd.MFMA_BF16
        for mode in (emida_amd.MFMA_BF16, emida_amd.MFMA_BF16_SPLIT):
            assert _create(T=np.float64, IN=IN, policy=emida_amd.CROSS_POLICY_BRUTE_MFMA, mfma_mode=mode).mfma_mode == mode
        for mode in (emida_amd.MFMA_F16, emida_amd.MFMA_F16_SPLIT):
            with pytest.raises(emida_amd.EmidaError) as e:
                _create(T=np.float64, IN=IN, policy=emida_amd.CROSS_POLICY_BRUTE_MFMA, mfma_mode=mode)
            assert e.value.code == C.ERR_UNSUPPORTED and "bf16" in str(e.value)
    assert _create(T=np.float64, policy=emida_amd.CROSS_POLICY_BRUTE_MFMA).mfma_mode == emida_amd.MFMA_F16


def test_matrix_core_default_mode_follows_refinement():
    # the default (MFMA_AUTO): the plain half type when the peak is refined (exact
    # results; a third of the products and of the fp32 accumulations), the split one
    # when the screen values are the results
    mfma = emida_amd.CROSS_POLICY_BRUTE_MFMA
    for IN, plain, split in ((np.uint16, emida_amd.MFMA_F16, emida_amd.MFMA_F16_SPLIT),
                             (np.float32, emida_amd.MFMA_BF16, emida_amd.MFMA_BF16_SPLIT)):
        assert _create(T=np.float32, IN=IN, policy=mfma).mfma_mode == plain
        assert _create(T=np.float32, IN=IN, policy=mfma, mfma_refine=False).mfma_mode == split
        assert _create(T=np.float32, IN=IN, policy=mfma, mfma_mode=emida_amd.MFMA_AUTO).mfma_mode == plain
        assert _create(T=np.float32, IN=IN, policy=mfma, mfma_mode=split).mfma_mode == split


def test_synthetic_grid_matches_benchmark_definition():
    # 5 x 5 centres on 256 px with radius 16 are exactly 16, 72, ..., 240
    assert list(synth.grid_centres(256, 5, 16)) == [16, 72, 128, 184, 240]
    assert list(synth.grid_centres(1024, 31, 32)) == list(range(32, 993, 32))
    b = synth.grid_begins(873, 20, 32)
    assert len(b) == 400 and all(0 <= x and x + 64 <= 873 and 0 <= y and y + 64 <= 873 for x, y in b)


def test_synthetic_patterns_are_deterministic():
    a = synth.PatternSet("det", 64, 2)
    b = synth.PatternSet("det", 64, 2)
    np.testing.assert_array_equal(a.pics, b.pics)
    np.testing.assert_array_equal(a.ref, b.ref)
    assert a.pics.dtype == np.uint16 and a.pics.shape == (2, 64, 64)


def test_entry_points_validate_arguments_before_any_device_call():
    # argument errors are reported without touching the GPU (this runs without one)
    L = C.lib()
    d = ctypes.c_void_p(1)  # never dereferenced: validation fails first
    calib = C.Calib(436.0, 436.0, 436.0)

    def err(rc):
        return rc, C.lib().emida_last_error().decode()

    fit = L.emida_fit_deformation
    assert err(fit(None, 0, 10, None, None, 1, calib, 1e-6, 10, *[None] * 7)) == (C.ERR_ARG, "null argument")
    for n in (0, 8193):
        assert err(fit(d, 0, n, d, d, 1, calib, 1e-6, 10, *[None] * 7))[1] == "n_rois must be in 1..8192"
    assert "more than 4 ROIs" in err(fit(d, 0, 4, d, None, 1, calib, 1e-6, 10, *[None] * 7))[1]
    assert fit(d, 0, 10, d, d, 0, calib, 1e-6, 10, *[None] * 7) == C.OK  # no patterns: nothing to do

    h = ctypes.c_void_p()
    for cfg, msg in ((C.ServerConfig(0, 1, 1, 1, 1, 0), "must be positive"),
                     (C.ServerConfig(1 << 20, 1, 17, 1, 1, 0), "at most 16 streams")):
        rc, m = err(L.emida_server_create(ctypes.byref(h), ctypes.byref(cfg)))
        assert rc == C.ERR_ARG and msg in m and not h.value

    rc, m = err(L.emida_run_cross_corr_mfma(0, d, d, d, C.Size2(12, 12), C.Size2(23, 23), 1, 1, None))
    assert rc == C.ERR_UNSUPPORTED and "multiple of 8" in m
    rc, m = err(L.emida_run_cross_corr_mfma(0, d, d, d, C.Size2(16, 16), C.Size2(33, 31), 1, 1, None))
    assert rc == C.ERR_ARG and "cross size" in m
    rc, m = err(L.emida_run_cross_corr(1, d, d, d, C.Size2(4, 4), C.Size2(10, 9), 1, 1, None))
    assert rc == C.ERR_ARG and "at most 2*size+1" in m


def test_cross_corr_fft_device_checks_buffer_sizes_first():
    # buffers of known size smaller than the call reads or writes are refused on the
    # host, before the library is called (this runs without a GPU)
    class Buf:
        def __init__(self, nbytes):
            self.nbytes, self.ptr = nbytes, ctypes.c_void_p(1)

    n = 128 * 128 * 3 * 2 * 4  # float, 3 ROIs x 2 patterns of 128 x 128
    for pics, ref, res, what in ((n - 4, n, 4 * n, "pattern slices"), (n, n // 2 - 4, 4 * n, "reference slices"),
                                 (n, n, 4 * n - 4, "result buffers")):
        with pytest.raises(ValueError, match=what):
            K.run_cross_corr_fft_device(Buf(pics), Buf(ref), Buf(res), np.float32, (128, 128), 3, 2)
