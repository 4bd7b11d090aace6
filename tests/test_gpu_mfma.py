"""Matrix-core brute correlation (the precision form) on the GPU, through the C ABI.

The reference has one brute-force correlation, src/cross_corr.cu:11-66, in T = float
or double.  The matrix-core form computes the same sums from bf16/fp16 halves of the
prepared float values with fp32 accumulation, optionally as hi*hi + hi*lo + lo*hi.
Checks:
  * integer-valued inputs are exact in every half type and every partial sum stays
    below 2^24, so the result must equal the FP64 oracle bit for bit - this pins the
    operand layout, the Toeplitz indexing and the output order;
  * random inputs: every element within the worst-case rounding bound of its mode,
    and the split modes far closer than the plain ones;
  * whole pipeline: offsets against the FP64 oracle (brute, double), with the
    maximum error stated per mode.
"""
import numpy as np
import pytest

import emida_amd
from emida_amd import kernels as K
from emida_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MODES = {  # mode: (unit roundoff of the half type, split, subnormal half-spacing)
    emida_amd.MFMA_BF16: (2.0 ** -8, False, 0.0),
    emida_amd.MFMA_BF16_SPLIT: (2.0 ** -8, True, 0.0),
    # fp16 rounds the halved value: its subnormal half-spacing 2^-25 is 2^-24 unscaled
    emida_amd.MFMA_F16: (2.0 ** -11, False, 2.0 ** -24),
    emida_amd.MFMA_F16_SPLIT: (2.0 ** -11, True, 2.0 ** -24),
}
SHAPES = [  # slice (x, y), cross (x, y), ROIs, patterns
    ((128, 128), (255, 255), 2, 3),
    ((64, 64), (127, 127), 3, 33),
    ((40, 24), (33, 17), 2, 5),
    ((8, 8), (15, 15), 1, 2),
    ((16, 40), (31, 79), 2, 3),
    ((24, 16), (32, 20), 2, 4),
]


def _oracle(p, r, size, cross, nb, npat):
    return O.cross_corr(p.astype(np.float64), r.astype(np.float64), size, cross, nb, npat)


class RoundingBound:
    """Worst case |mfma - exact| per element.  Operands: |v - hi| <= u|v| (+ the
    subnormal half-spacing h for fp16); split: |v - hi - lo| <= u^2|v| + h, and the
    omitted lo*lo term is <= u^2(1+u)^2|v||r|.  Products of halves are exact in fp32;
    the fp32 sum of n terms adds gamma_n = n 2^-24 / (1 - n 2^-24) of sum|terms|."""

    def __init__(self, p, r, size, cross, nb, npat):
        ap, ar = np.abs(p).astype(np.float64), np.abs(r).astype(np.float64)
        self.n = size[0] * size[1]
        self.pr = _oracle(ap, ar, size, cross, nb, npat)
        self.p1 = _oracle(ap, np.ones_like(ar), size, cross, nb, npat)
        self.r1 = _oracle(np.ones_like(ap), ar, size, cross, nb, npat)

    def __call__(self, mode):
        u, split, h = MODES[mode]
        n = self.n * (3 if split else 1)
        gamma = n * 2.0 ** -24 / (1 - n * 2.0 ** -24)
        rel = (3.1 * u * u if split else 2 * u + u * u) * (1 + gamma) + gamma * (1 + u) ** 2
        return rel * self.pr + 1.01 * h * (self.p1 + self.r1) + 1e-30


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("size,cross,nb,npat", SHAPES)
def test_integer_inputs_bitexact(size, cross, nb, npat, mode):
    rng = np.random.default_rng(size[0] * 7 + size[1] + npat)
    p = rng.integers(-4, 5, size=size[0] * size[1] * nb * npat).astype(np.float32)
    r = rng.integers(-4, 5, size=size[0] * size[1] * nb).astype(np.float32)
    got = K.run_cross_corr_mfma(p, r, size, cross, nb, npat, mode)
    np.testing.assert_array_equal(got.astype(np.float64), _oracle(p, r, size, cross, nb, npat))


@pytest.mark.parametrize("size,cross,nb,npat", SHAPES[:3])
def test_random_inputs_within_rounding_bound(size, cross, nb, npat):
    rng = np.random.default_rng(5)
    p = (rng.standard_normal(size[0] * size[1] * nb * npat) * 40).astype(np.float32)
    r = (rng.standard_normal(size[0] * size[1] * nb) * 40).astype(np.float32)
    exact = _oracle(p, r, size, cross, nb, npat)
    rb = RoundingBound(p, r, size, cross, nb, npat)
    worst = {}
    for mode in MODES:
        got = K.run_cross_corr_mfma(p, r, size, cross, nb, npat, mode).astype(np.float64)
        err = np.abs(got - exact)
        bound = rb(mode)
        assert (err <= bound).all(), (mode, float((err / bound).max()))
        worst[mode] = err.max()
    # hi + lo products recover most of what the single half type loses
    assert worst[emida_amd.MFMA_BF16_SPLIT] * 16 < worst[emida_amd.MFMA_BF16]
    assert worst[emida_amd.MFMA_F16_SPLIT] * 16 < worst[emida_amd.MFMA_F16]
    print("max abs error per mode", {k: float(v) for k, v in worst.items()})


def test_rejects_unsupported_shapes():
    p = np.zeros(60 * 60, np.float32)
    with pytest.raises(emida_amd.EmidaError) as e:
        K.run_cross_corr_mfma(p, p, (60, 60), (119, 119))  # width not a multiple of 8
    assert e.value.code == emida_amd._capi.ERR_UNSUPPORTED
    p = np.zeros(136 * 8, np.float32)
    with pytest.raises(emida_amd.EmidaError):
        K.run_cross_corr_mfma(p, p, (136, 8), (15, 15))  # wider than 128
    p = np.zeros(16 * 16, np.float32)
    with pytest.raises(emida_amd.EmidaError) as e:
        K.run_cross_corr_mfma(p, p, (16, 16), (33, 31))  # cross > 2s - 1
    assert e.value.code == emida_amd._capi.ERR_ARG


# ------------------------------------------------------------------ whole pipeline
def well_posed(coefs, fit):
    """ROIs whose FP64 quadratic fit has its maximum inside the F x F neighbourhood;
    elsewhere the fitted position is an extrapolation, ill-conditioned in any precision."""
    b, c, d, e, f = (coefs[:, i] for i in range(1, 6))
    det = 4 * f * d - e * e
    with np.errstate(divide="ignore", invalid="ignore"):
        py = (b * e - 2 * c * d) / det
        px = (-e * py - b) / (2 * d)
    return (d < 0) & (det > 0) & (px >= 0) & (px <= fit - 1) & (py >= 0) & (py <= fit - 1)


def _run(ps, begins, slice_size, cross, fit, batch, T=np.float32, policy=emida_amd.CROSS_POLICY_BRUTE_MFMA,
         mode=emida_amd.MFMA_F16_SPLIT, refine=True):
    with emida_amd.GpuOffset(T, np.uint16, ps.size, begins, slice_size, cross, batch, policy, fit, mfma_mode=mode,
                             mfma_refine=refine) as g:
        g.allocate_memory(ps.ref)
        maxes, neigh = g.new_result_buffers(pinned=False)
        idx = g.transfer_pic_to_device_async(ps.pics[:batch])
        g.wait_uploads()
        g.get_offset_core(maxes, neigh, idx)
        offs, coefs = g.finalize(maxes, neigh)
        p_offs, _ = g.get_offset(ps.pics[:batch])
        np.testing.assert_array_equal(offs, p_offs)  # both host forms, same results
        return offs, coefs, maxes.copy(), neigh.copy()


# max offset error vs FP64 on well-posed ROIs of the unrefined float screen (measured
# on MI355X: bf16 2.4e-3, bf16 split 3.5e-5, fp16 2.8e-4, fp16 split 2.9e-5 px)
PIPE_TOL = {emida_amd.MFMA_BF16: 0.01, emida_amd.MFMA_BF16_SPLIT: 5e-4, emida_amd.MFMA_F16: 2e-3,
            emida_amd.MFMA_F16_SPLIT: 5e-4}


@pytest.mark.parametrize("mode", list(MODES))
def test_screen_only_pipeline_vs_fp64_oracle(mode):
    ps = synth.PatternSet("mfma", 256, 3, t_range=2.0)
    begins = synth.grid_begins(256, 4, 32)
    eo, ec = O.get_offset(ps.pics.ravel(), ps.ref, ps.size, begins, (64, 64), (127, 127), 3, 7, O.POLICY_BRUTE,
                          np.float64)
    offs = _run(ps, begins, (64, 64), (127, 127), 7, 3, mode=mode, refine=False)[0]
    ok = well_posed(ec, 7)
    assert ok.mean() > 0.5
    err = np.abs(offs - eo)[ok].max()
    print(f"mode {mode}: max error {err:.2e} px over {ok.sum()} well-posed ROIs")
    assert err < PIPE_TOL[mode]


# Refined: every output equals the brute policy's in the same precision, bit for bit,
# whatever the screen's precision (even a single bf16 half).
@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("mode", list(MODES))
def test_refined_equals_brute_policy(T, mode):
    ps = synth.PatternSet("mfma", 256, 3, t_range=2.0)
    begins = synth.grid_begins(256, 4, 32)
    got = _run(ps, begins, (64, 64), (127, 127), 7, 3, T=T, mode=mode)
    exp = _run(ps, begins, (64, 64), (127, 127), 7, 3, T=T, policy=emida_amd.CROSS_POLICY_BRUTE)
    for g, e, what in zip(got, exp, ("offsets", "coefs", "maxes", "neighbourhoods")):
        np.testing.assert_array_equal(g, e, err_msg=what)


@pytest.mark.parametrize("cross", [(9, 9), (15, 7)])
def test_refined_equals_brute_policy_edge_peaks(cross):
    # displacements up to 3 px in a small window put maxima on the window edge: the
    # neighbourhood clamp reads one past the edge, into the next row or slice
    ps = synth.PatternSet("mfma-edges", 256, 2, t_range=3.0)
    begins = synth.grid_begins(256, 4, 16)
    got = _run(ps, begins, (32, 32), cross, 3, 2, T=np.float64, mode=emida_amd.MFMA_BF16)
    exp = _run(ps, begins, (32, 32), cross, 3, 2, T=np.float64, policy=emida_amd.CROSS_POLICY_BRUTE)
    for g, e in zip(got, exp):
        np.testing.assert_array_equal(g, e)


@pytest.mark.parametrize("mode", [emida_amd.MFMA_F16_SPLIT, emida_amd.MFMA_BF16])
def test_refined_equals_brute_policy_tall_window(mode):
    # 64 x 144 slices with the full 127 x 287 window: more screen rows than the
    # refinement's row list holds (only rows that can hold a candidate are listed; a
    # slice with more of them than the list holds is swept)
    ps = synth.PatternSet("mfma-tall", 256, 2, t_range=2.0)
    begins = [(16, 8), (150, 100), (96, 40)]
    got = _run(ps, begins, (64, 144), (127, 287), 5, 2, T=np.float64, mode=mode)
    exp = _run(ps, begins, (64, 144), (127, 287), 5, 2, T=np.float64, policy=emida_amd.CROSS_POLICY_BRUTE)
    for g, e, what in zip(got, exp, ("offsets", "coefs", "maxes", "neighbourhoods")):
        np.testing.assert_array_equal(g, e, err_msg=what)


def _spread_patterns(batch):
    """384 x 128 uint16 patterns of three 128 x 128 ROIs: vertical stripes (a ridge of
    candidates along the y-shifts), independent noise in pattern and reference (a weak
    peak: candidates scattered over a box of about 30 x 30 to 70 x 70 shifts depending
    on the mode's rounding bound, more than one box evaluation holds) and diagonal
    stripes with noise (a compact peak)."""
    rng = np.random.default_rng(5)
    yy, xx = np.mgrid[0:128, 0:384].astype(float)

    def one():
        a = 30000 + 20000 * np.sin(2 * np.pi * xx / 16)
        a[:, 128:256] = 30000 + rng.normal(0, 3000, (128, 128))
        a[:, 256:] = (30000 + 15000 * np.sin(2 * np.pi * (xx[:, 256:] + yy[:, 256:]) / 24)
                      + rng.normal(0, 2000, (128, 128)))
        return np.clip(a, 0, 65535).astype(np.uint16)

    ref = one()
    pics = np.stack([one() for _ in range(batch)])

    class P:
        pass

    ps = P()
    ps.ref, ps.pics, ps.size = ref, pics, (384, 128)
    return ps, [(0, 0), (128, 0), (256, 0)]


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("mode", list(MODES))
def test_refined_equals_brute_policy_spread_candidates(T, mode):
    # candidate sets wider than one box are evaluated one chain per candidate, from the
    # prepared slices in HBM; results still the brute policy's
    ps, begins = _spread_patterns(2)
    got = _run(ps, begins, (128, 128), (255, 255), 7, 2, T=T, mode=mode)
    exp = _run(ps, begins, (128, 128), (255, 255), 7, 2, T=T, policy=emida_amd.CROSS_POLICY_BRUTE)
    for g, e, what in zip(got, exp, ("offsets", "coefs", "maxes", "neighbourhoods")):
        np.testing.assert_array_equal(g, e, err_msg=what)


@pytest.mark.slow
def test_precision_form_workload_subset_vs_fp64_oracle():
    # the precision-form workload's shape: 873 x 873 patterns, ROIs of 128 x 128 on a
    # 10 x 10 grid, full 255 x 255 correlation, fitsize 7 (25 of the 100 ROIs): refined
    # double results are the FP64 oracle's, screen-only split modes within 0.01 px
    ps = synth.PatternSet("mfma873", 873, 1, t_range=3.0)
    begins = synth.grid_begins(873, 10, 64)[::4]
    eo, ec = O.get_offset(ps.pics.ravel(), ps.ref, ps.size, begins, (128, 128), (255, 255), 1, 7, O.POLICY_BRUTE,
                          np.float64)
    offs = _run(ps, begins, (128, 128), (255, 255), 7, 1, T=np.float64)[0]
    np.testing.assert_array_equal(offs, eo)
    ok = well_posed(ec, 7)
    for mode in (emida_amd.MFMA_BF16_SPLIT, emida_amd.MFMA_F16_SPLIT):
        offs = _run(ps, begins, (128, 128), (255, 255), 7, 1, mode=mode, refine=False)[0]
        assert np.abs(offs - eo)[ok].max() < 0.01


# ------------------------------------------------------------------ flat screens
def _tau(p, r, mode, size, t_double):
    """The refinement's candidate bound tau = tol_rel ||p|| ||r|| + tol_abs (||p|| + ||r||)
    for mode (the arithmetic of mfma_refine_bounds, restated)."""
    u = 2.0 ** -11 if mode in (emida_amd.MFMA_F16, emida_amd.MFMA_F16_SPLIT) else 2.0 ** -8
    split = mode in (emida_amd.MFMA_BF16_SPLIT, emida_amd.MFMA_F16_SPLIT)
    n = size[0] * size[1]
    e_op = (3.1 * u * u if split else 2 * u + u * u) + (2.01 * 2.0 ** -24 if t_double else 0.0)
    nt, u32 = n * (3 if split else 1), 2.0 ** -24
    g_acc = nt * u32 / (1 - nt * u32)
    ut = 2.0 ** -53 if t_double else u32
    g_t = n * ut / (1 - n * ut)
    h = 2.1 * 2.0 ** -24 if u == 2.0 ** -11 else 0.0
    tol_rel = 2.02 * (e_op * (1 + g_acc) + g_acc * (1 + u) ** 2 + g_t)
    tol_abs = 2.02 * h * np.sqrt(n) * (1 + g_acc) * (2 if split else 1)
    npn, nrn = np.sqrt((p.astype(np.float64) ** 2).sum()), np.sqrt((r.astype(np.float64) ** 2).sum())
    return tol_rel * npn * nrn + tol_abs * (npn + nrn)


def _plateau_patterns(batch):
    """256 x 128 uint16 patterns of two 128 x 128 ROIs.  The deformed patterns hold one
    bright pixel on a dark ROI, the reference a smooth blob: the correlation is the
    windowed blob, shifted, whose broad top leaves thousands of shifts within a plain
    half-type screen's rounding bound of the maximum."""
    yy, xx = np.mgrid[0:128, 0:256].astype(float)
    ref = 500 + 40000 * np.exp(-((xx - 60) ** 2 + (yy - 70) ** 2) / (2 * 22.0 ** 2))
    ref += 30000 * np.exp(-((xx - 200) ** 2 + (yy - 50) ** 2) / (2 * 28.0 ** 2))
    ref = np.clip(ref, 0, 65535).astype(np.uint16)
    pics = np.zeros((batch, 128, 256), np.uint16)
    for b in range(batch):
        pics[b, 50 + 7 * b, 70 - 5 * b] = 60000
        pics[b, 40 + 3 * b, 128 + 80 + 4 * b] = 50000

    class P:
        pass

    ps = P()
    ps.ref, ps.pics, ps.size = ref, pics, (256, 128)
    return ps, [(0, 0), (128, 0)]


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("mode", [emida_amd.MFMA_BF16])
def test_refined_equals_brute_policy_beyond_candidate_cap(T, mode):
    # more than 512 candidates: the refinement sweeps the whole window exactly, so the
    # arg-max is still the brute policy's (it used to keep the screen's arg-max)
    ps, begins = _plateau_patterns(2)
    size, cross = (128, 128), (255, 255)
    # the candidates the screen leaves, counted as the refinement counts them
    hx = O.hann(128, np.float32)
    sp = O.sums(ps.pics.ravel(), begins, ps.size, size, 2)
    sr = O.sums(ps.ref.ravel(), begins, ps.size, size, 1)
    pf = O.prepare(ps.pics.ravel(), hx, hx, sp, begins, ps.size, size, None, 2).astype(np.float32)
    rf = O.prepare(ps.ref.ravel(), hx, hx, sr, begins, ps.size, size, None, 1).astype(np.float32)
    screen = K.run_cross_corr_mfma(pf, rf, size, cross, 2, 2, mode).reshape(4, -1).astype(np.float64)
    hd = O.hann(128)
    pd = O.prepare(ps.pics.ravel(), hd, hd, sp, begins, ps.size, size, None, 2).reshape(4, -1)
    rd = O.prepare(ps.ref.ravel(), hd, hd, sr, begins, ps.size, size, None, 1).reshape(2, -1)
    for sn in range(4):
        tau = _tau(pd[sn], rd[sn % 2], mode, size, T == np.float64)
        n_cand = int((screen[sn] >= screen[sn].max() - tau).sum())
        assert n_cand > 512, (sn, n_cand)
    got = _run(ps, begins, size, cross, 5, 2, T=T, mode=mode)
    exp = _run(ps, begins, size, cross, 5, 2, T=T, policy=emida_amd.CROSS_POLICY_BRUTE)
    for g, e, what in zip(got, exp, ("offsets", "coefs", "maxes", "neighbourhoods")):
        np.testing.assert_array_equal(g, e, err_msg=what)


@pytest.mark.parametrize("mode", [emida_amd.MFMA_BF16, emida_amd.MFMA_BF16_SPLIT])
@pytest.mark.parametrize("data", ["plateau", "stripes"])
def test_screen_only_row_maxima_equal_full_scan(mode, data):
    # the screen-only pipeline finds the arg-max from the matrix-core kernel's per-row
    # maxima; the kernel seam's full scan of the same screen (maxarg_reduce, lowest
    # index on ties) and gather must give the same records and neighbourhoods bit for
    # bit, on inputs with many ties
    if data == "plateau":
        ps, begins = _plateau_patterns(2)
        size, cross, batch = (128, 128), (255, 255), 2
    else:
        ps, begins = _spread_patterns(2)
        size, cross, batch = (128, 128), (255, 255), 2
    nb = len(begins)
    hx = O.hann(size[0], np.float32)
    sp = O.sums(ps.pics[:batch].ravel(), begins, ps.size, size, batch)
    sr = O.sums(ps.ref.ravel(), begins, ps.size, size, 1)
    pf = O.prepare(ps.pics[:batch].ravel(), hx, hx, sp, begins, ps.size, size, None, batch).astype(np.float32)
    rf = O.prepare(ps.ref.ravel(), hx, hx, sr, begins, ps.size, size, None, 1).astype(np.float32)
    screen = K.run_cross_corr_mfma(pf, rf, size, cross, nb, batch, mode)
    full = K.run_maxarg_reduce(screen, cross, nb * batch)
    neigh = K.run_extract_neighbors(screen, full, 7, cross, nb * batch)
    _, _, maxes, got_neigh = _run(ps, begins, size, cross, 7, batch, mode=mode, refine=False)
    np.testing.assert_array_equal(maxes["index"], full["index"])
    np.testing.assert_array_equal(maxes["data"], full["data"])
    np.testing.assert_array_equal(got_neigh, neigh)


def test_matrix_core_objects_release_their_memory():
    # create / allocate / destroy in a loop: every device buffer of the matrix-core
    # policy is given back.  8 x 128 slices with a 15 x 255 window and 2048 slices per
    # batch give per-row screen maxima of 2 MiB, a buffer the destructor once missed
    ps = synth.PatternSet("mfma-leak", 256, 1, t_range=2.0)
    begins = [(16 * i, 64) for i in range(16)]

    def once():
        with emida_amd.GpuOffset(np.float64, np.uint16, ps.size, begins, (8, 128), (15, 255), 128,
                                 emida_amd.CROSS_POLICY_BRUTE_MFMA, 3) as g:
            g.allocate_memory(ps.ref)
            return g.device_bytes

    once()
    emida_amd.synchronize()
    free0, _ = emida_amd.device_memory()
    held = 0
    for _ in range(30):
        held = once()
    emida_amd.synchronize()
    free1, _ = emida_amd.device_memory()
    assert held > 32 << 20
    # 30 objects; a 2 MiB buffer left behind by each would take 60 MiB
    assert free0 - free1 < 16 << 20, (free0 - free1, held)
