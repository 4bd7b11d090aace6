"""End-to-end parity of the drop-in gpu_offset object on the GPU (C ABI).

  * the reference's own end-to-end goldens (test/get_offset_test.cpp), double;
  * the oracle on synthetic uint16 band patterns: brute double and float results are
    bit-identical (same operation order); FFT double offsets within 1e-9 px and FFT
    float within 2e-3 px, with equal integer peaks, on the well-conditioned ROIs
    (oracle.well_conditioned: the oracle's peak is separated from the runner-up by
    more than the rounding bound and its fitted maximum lies inside the neighbourhood);
  * the asynchronous pipeline API (transfer / get_offset_core / finalize), the
    device-resident entry point and argument errors.
"""
import os

import numpy as np
import pytest

import emida_amd
from emida_amd import synth
from oracle import oracle as O
import get_offset_goldens as G

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
BRUTE, FFT = emida_amd.CROSS_POLICY_BRUTE, emida_amd.CROSS_POLICY_FFT


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ------------------------------------------------------------------ reference goldens
@pytest.mark.parametrize("policy", [BRUTE, FFT])
@pytest.mark.parametrize("batch", [1, 5])
def test_batched_goldens(policy, batch):  # test/get_offset_test.cpp:108-199
    pic, temp = load("data_pic.npy"), load("data_temp.npy")
    pics = np.concatenate([pic.ravel()] * batch)
    begins = G.slice_begins()
    with emida_amd.GpuOffset(np.float64, np.float64, (256, 256), begins, (64, 64), (127, 127), batch, policy) as g:
        g.allocate_memory(temp)
        offs, _ = g.get_offset(pics)
    exp = np.tile(np.array(G.BATCHED), (batch, 1))
    assert np.abs(offs - exp).max() <= G.BATCHED_TOL


def test_bigger_golden():  # test/get_offset_test.cpp:10-27 (free function get_offset<double>)
    pic, temp = load("data_pic.npy"), load("data_temp.npy")
    a, b = pic[:64, :64].copy(), temp[:64, :64].copy()
    off = emida_amd.get_offset(a, b, (64, 64), (127, 127))
    assert np.abs(off - np.array(G.BIGGER)).max() <= G.BIGGER_TOL


@pytest.mark.parametrize("policy", [BRUTE, FFT])
def test_goldens_bitexact_to_oracle_brute_close_fft(policy):
    pic, temp = load("data_pic.npy"), load("data_temp.npy")
    begins = G.slice_begins()
    with emida_amd.GpuOffset(np.float64, np.float64, (256, 256), begins, (64, 64), (127, 127), 1, policy) as g:
        g.allocate_memory(temp)
        offs, coefs = g.get_offset(pic)
    eo, ec = O.get_offset(pic, temp, (256, 256), begins, (64, 64), (127, 127), 1, 3, policy)
    if policy == BRUTE:
        np.testing.assert_array_equal(offs, eo)
        np.testing.assert_array_equal(coefs, ec)
    else:
        assert np.abs(offs - eo).max() < 1e-9


# ------------------------------------------------------------------ synthetic uint16 patterns
def _patterns(size, n, name="parity", t_range=2.0):
    return synth.PatternSet(name, size, n, t_range=t_range)


def _run(T, policy, ps, begins, slice_size, cross, fit, batch):
    with emida_amd.GpuOffset(T, np.uint16, ps.size, begins, slice_size, cross, batch, policy, fit) as g:
        g.allocate_memory(ps.ref)
        maxes, neigh = g.new_result_buffers(pinned=False)
        idx = g.transfer_pic_to_device_async(ps.pics[:batch])
        g.wait_uploads()
        g.get_offset_core(maxes, neigh, idx)
        offs, coefs = g.finalize(maxes, neigh)
        return offs, coefs, maxes.copy(), neigh.copy()


@pytest.mark.parametrize("fit", [3, 5, 7, 9])
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_brute_bitexact_vs_oracle(T, fit):
    ps = _patterns(256, 3)
    begins = synth.grid_begins(256, 5, 16)
    offs, coefs, maxes, neigh = _run(T, BRUTE, ps, begins, (32, 32), (63, 63), fit, 3)
    eo, ec, em, en = O.get_offset(ps.pics.ravel(), ps.ref, ps.size, begins, (32, 32), (63, 63), 3, fit, BRUTE, T,
                                  return_intermediates=True)
    np.testing.assert_array_equal(maxes["index"], em["index"])
    np.testing.assert_array_equal(maxes["data"], em["data"])
    np.testing.assert_array_equal(neigh, en)
    np.testing.assert_array_equal(coefs, ec)
    np.testing.assert_array_equal(offs, eo)


@pytest.mark.parametrize("side", [128, 136])
@pytest.mark.parametrize("T", [np.float64, np.float32])
def test_brute_both_preparation_paths_vs_oracle(T, side):
    # uint16 slices of up to 128 x 128 (32 KB of pixels) take the one-kernel sums and
    # preparation, larger ones the separate sum and prepare kernels
    ps = _patterns(256, 2)
    begins = synth.grid_begins(256, 2, side // 2)
    offs, coefs, maxes, neigh = _run(T, BRUTE, ps, begins, (side, side), (9, 9), 3, 2)
    eo, ec, em, en = O.get_offset(ps.pics.ravel(), ps.ref, ps.size, begins, (side, side), (9, 9), 2, 3, BRUTE, T,
                                  return_intermediates=True)
    np.testing.assert_array_equal(maxes["index"], em["index"])
    np.testing.assert_array_equal(maxes["data"], em["data"])
    np.testing.assert_array_equal(neigh, en)
    np.testing.assert_array_equal(offs, eo)


@pytest.mark.parametrize("cross", [(9, 9), (5, 7), (3, 3)])
def test_brute_cropped_cross_with_edge_peaks_vs_oracle(cross):
    # displacements of +-3 px with a small correlation window put maxima on the
    # window edge, exercising the one-past-the-edge clamp (src/extract_neighbors.cu:31-35)
    ps = _patterns(256, 2, "edges", t_range=3.0)
    begins = synth.grid_begins(256, 4, 16)
    offs, coefs, maxes, neigh = _run(np.float64, BRUTE, ps, begins, (32, 32), cross, 3, 2)
    eo, ec, em, en = O.get_offset(ps.pics.ravel(), ps.ref, ps.size, begins, (32, 32), cross, 2, 3, BRUTE,
                                  np.float64, return_intermediates=True)
    np.testing.assert_array_equal(maxes["index"], em["index"])
    np.testing.assert_array_equal(neigh, en)
    np.testing.assert_array_equal(offs, eo)


@pytest.mark.parametrize("s,fit", [(64, 3), (64, 5), (64, 9), (48, 5), (33, 3), (20, 7)])
def test_fft_double_vs_oracle(s, fit):
    # sides that are not powers of two run zero-padded to the next power of two
    ps = _patterns(256, 2)
    begins = synth.grid_begins(256, 4, (s + 1) // 2)
    c = (2 * s - 1, 2 * s - 1)
    offs, coefs, maxes, neigh = _run(np.float64, FFT, ps, begins, (s, s), c, fit, 2)
    eo, ec, em, en, st = O.get_offset(ps.pics.ravel(), ps.ref, ps.size, begins, (s, s), c, 2, fit, FFT,
                                      np.float64, return_intermediates=True, return_stats=True)
    ok = O.well_conditioned(ec, fit, st, rel_margin=1e-11)
    assert ok.mean() > 0.6, ok.mean()
    np.testing.assert_array_equal(maxes["index"][ok], em["index"][ok])
    assert np.abs(offs - eo)[ok].max() < 1e-9
    # values carry the reference's (2s)^2 scale also when the transform is padded
    np.testing.assert_allclose(maxes["data"][ok], em["data"][ok], rtol=1e-9)
    assert (np.abs(coefs - ec)[ok].max(axis=1) <= 1e-9 * np.abs(ec)[ok].max(axis=1)).all()


@pytest.mark.parametrize("s,fit", [(32, 5), (16, 3), (8, 3), (64, 7), (128, 7), (128, 9), (48, 5), (21, 3),
                                   (96, 7), (100, 5), (127, 9), (65, 3)])
def test_fft_float_vs_oracle(s, fit):
    ps = _patterns(256, 2)
    begins = synth.grid_begins(256, 5, (s + 1) // 2)
    offs, coefs, maxes, neigh = _run(np.float32, FFT, ps, begins, (s, s), (2 * s - 1, 2 * s - 1), fit, 2)
    eo, ec, em, en, st = O.get_offset(ps.pics.ravel(), ps.ref, ps.size, begins, (s, s), (2 * s - 1, 2 * s - 1), 2,
                                      fit, FFT, np.float64, return_intermediates=True, return_stats=True)
    ok = O.well_conditioned(ec, fit, st)
    assert ok.mean() > 0.6, ok.mean()
    np.testing.assert_array_equal(maxes["index"][ok], em["index"][ok])
    assert np.abs(offs - eo)[ok].max() < 2e-3


@pytest.mark.parametrize("T,size,fit", [
    (np.float64, (96, 96), 5), (np.float64, (128, 128), 7),   # double 256-point transforms
    (np.float64, (160, 160), 5), (np.float32, (160, 160), 5),  # 512-point transforms
    (np.float32, (200, 200), 3),
    (np.float64, (64, 48), 5), (np.float32, (64, 48), 5), (np.float32, (40, 100), 3), (np.float64, (100, 40), 3)])
def test_fft_three_kernel_path_vs_oracle(T, size, fit):
    # shapes the single-kernel pipelines do not take: non-square slices, double sides
    # above 64, sides above 128
    ps = _patterns(512, 2, "fftg", t_range=3.0)
    sx, sy = size
    xs = synth.grid_centres(512, 2, sx // 2 + 1) - sx // 2
    ys = synth.grid_centres(512, 2, sy // 2 + 1) - sy // 2
    begins = [(int(x), int(y)) for y in ys for x in xs]
    c = (2 * sx - 1, 2 * sy - 1)
    offs, coefs, maxes, neigh = _run(T, FFT, ps, begins, size, c, fit, 2)
    eo, ec, em, en, st = O.get_offset(ps.pics.ravel(), ps.ref, ps.size, begins, size, c, 2, fit, FFT, np.float64,
                                      return_intermediates=True, return_stats=True)
    double = T == np.float64
    ok = O.well_conditioned(ec, fit, st, rel_margin=1e-11 if double else O.FLOAT_MARGIN)
    assert ok.sum() >= 4, ok
    np.testing.assert_array_equal(maxes["index"][ok], em["index"][ok])
    assert np.abs(offs - eo)[ok].max() < (1e-9 if double else 2e-3)
    if double:  # values carry the reference's (2s.x)(2s.y) scale
        np.testing.assert_allclose(maxes["data"][ok], em["data"][ok], rtol=1e-9)


def test_fft_float_873_config_subset_vs_fp64_oracle():
    # the benchmark workload's shape: 873 x 873 patterns, 20 x 20 ROIs of 64 x 64
    ps = _patterns(873, 1, "fft873", t_range=3.0)
    begins = synth.grid_begins(873, 20, 32)[::4]  # 100 of the 400 ROIs, all rows/columns
    offs, _, maxes, _ = _run(np.float32, FFT, ps, begins, (64, 64), (127, 127), 5, 1)
    eo, ec, em, _, st = O.get_offset(ps.pics.ravel(), ps.ref, ps.size, begins, (64, 64), (127, 127), 1, 5, FFT,
                                     np.float64, return_intermediates=True, return_stats=True)
    ok = O.well_conditioned(ec, 5, st)
    np.testing.assert_array_equal(maxes["index"][ok], em["index"][ok])
    worst = np.abs(offs - eo)[ok].max()
    print(f"873 x 873 subset, float FFT: worst well-conditioned ROI {worst:.2e} px, all ROIs {np.abs(offs - eo).max():.2e} px")
    assert worst < 2e-3, worst
    assert np.abs(offs - eo).max() < 1e-2


# ------------------------------------------------------------------ host API forms
def test_fft_float_128_chunked_launch_vs_oracle():
    # more slices than one launch chunk of the 256-point path (4096 slices): 2 patterns
    # x 2209 ROIs of 128 x 128 on a 873 x 873 pattern, checked on every 97th ROI
    ps = _patterns(873, 2, "fft128", t_range=3.0)
    begins = synth.grid_begins(873, 47, 64)
    offs, _, maxes, _ = _run(np.float32, FFT, ps, begins, (128, 128), (255, 255), 5, 2)
    sub = begins[::97]
    idx = np.concatenate([np.arange(0, len(begins), 97) + p * len(begins) for p in range(2)])
    eo, _, em, _ = O.get_offset(ps.pics.ravel(), ps.ref, ps.size, sub, (128, 128), (255, 255), 2, 5, FFT,
                                np.float64, return_intermediates=True)
    assert np.abs(offs[idx] - eo).max() < 2e-3


def test_pipeline_api_equals_get_offset():
    ps = _patterns(256, 4)
    begins = synth.grid_begins(256, 5, 16)
    with emida_amd.GpuOffset(np.float32, np.uint16, ps.size, begins, (32, 32), (63, 63), 2, FFT, 5, 2) as g:
        g.allocate_memory(ps.ref)
        sync0, _ = g.get_offset(ps.pics[:2])
        sync1, _ = g.get_offset(ps.pics[2:4])
        bufs = [g.new_result_buffers(pinned=True), g.new_result_buffers(pinned=True)]
        i0 = g.transfer_pic_to_device_async(ps.pics[:2].copy())
        i1 = g.transfer_pic_to_device_async(ps.pics[2:4].copy())
        g.wait_uploads()
        g.get_offset_core(*bufs[0], i0)
        g.get_offset_core(*bufs[1], i1)
        a0, _ = g.finalize(*bufs[0])
        a1, _ = g.finalize(*bufs[1])
        np.testing.assert_array_equal(a0, sync0)
        np.testing.assert_array_equal(a1, sync1)
        # foreign buffers: finalize fits what it is given (GPU fit)
        m = bufs[0][0].copy()
        n = bufs[0][1].copy() * np.float32(2)
        b0, bc = g.finalize(m, n)
        eo, ec = O.subpixel(n, 5)
        np.testing.assert_array_equal(bc, ec)


def test_process_device_equals_get_offset():
    ps = _patterns(256, 3)
    begins = synth.grid_begins(256, 5, 16)
    with emida_amd.GpuOffset(np.float32, np.uint16, ps.size, begins, (32, 32), (63, 63), 3, FFT, 3) as g:
        g.allocate_memory(ps.ref)
        ref_offs, ref_coefs = g.get_offset(ps.pics)
        d_pics = emida_amd.DeviceBuffer.from_array(ps.pics)
        d_offs = emida_amd.DeviceBuffer(g.total_slices * 16)
        d_coefs = emida_amd.DeviceBuffer(g.total_slices * 48)
        g.process_device(d_pics, 3, d_offsets=d_offs, d_coefs=d_coefs)
        emida_amd.synchronize()
        np.testing.assert_array_equal(d_offs.download(np.float64).reshape(-1, 2), ref_offs)
        np.testing.assert_array_equal(d_coefs.download(np.float64).reshape(-1, 6), ref_coefs)
        # a partial batch gives the leading slices' results
        g.process_device(d_pics, 1, d_offsets=d_offs)
        emida_amd.synchronize()
        part = d_offs.download(np.float64, 2 * len(begins)).reshape(-1, 2)
        np.testing.assert_array_equal(part, ref_offs[:len(begins)])


def test_calls_in_wrong_order_raise():
    with emida_amd.GpuOffset(np.float64, np.uint16, (64, 64), [(0, 0)], (32, 32), (63, 63), 1) as g:
        with pytest.raises(emida_amd.EmidaError):
            g.get_offset(np.zeros((64, 64), np.uint16))
        g.allocate_memory(np.ones((64, 64), np.uint16))
        with pytest.raises(emida_amd.EmidaError):
            g.allocate_memory(np.ones((64, 64), np.uint16))
        with pytest.raises(ValueError):
            g.get_offset(np.zeros((63, 64), np.uint16))


def test_constant_patterns_give_finite_or_nan_offsets_without_fault():
    # all-zero correlation: arg-max index 0 (lowest), degenerate fit -> NaN like the reference
    with emida_amd.GpuOffset(np.float64, np.uint16, (64, 64), [(0, 0), (32, 32)], (32, 32), (63, 63), 1) as g:
        g.allocate_memory(np.full((64, 64), 7, np.uint16))
        offs, coefs = g.get_offset(np.full((64, 64), 9, np.uint16))
    eo, ec = O.get_offset(np.full(64 * 64, 9, np.uint16), np.full(64 * 64, 7, np.uint16), (64, 64),
                          [(0, 0), (32, 32)], (32, 32), (63, 63), 1)
    np.testing.assert_array_equal(np.isnan(offs), np.isnan(eo))
    np.testing.assert_array_equal(coefs, ec)
