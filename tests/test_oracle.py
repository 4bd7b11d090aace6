"""Pins the CPU oracle against the reference's own tests, fixtures and mathematics.

Each test names the reference test it restates (file:line). Independent checks
(numpy FFT, numpy lstsq, brute force) are chosen so that a plausible mistake in
the oracle (transposed operand, wrong sign, dropped term, off-by-one shift) fails.
"""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
import get_offset_goldens as G

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF_TREE = "/root/reference"


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ------------------------------------------------------------------ Hann window
def ulp_equal(a, b, ulps=4):
    """gtest EXPECT_DOUBLE_EQ: within 4 units in the last place."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    ia, ib = a.view(np.int64), b.view(np.int64)
    ia = np.where(ia < 0, np.int64(-2**63) - ia, ia)
    ib = np.where(ib < 0, np.int64(-2**63) - ib, ib)
    return bool(np.all(np.abs(ia - ib) <= ulps))


def test_hann_closed_form():  # test/prepare_test.cpp:18-23 (EXPECT_DOUBLE_VECTORS_EQ, 4 ulp)
    assert ulp_equal(O.hann(5), [0, 0.5, 1, 0.5, 0])


@pytest.mark.parametrize("n", [5, 128])
def test_hann_fixture(n):  # fixtures test/res/hann/{5,128}_{A,res}.txt (python/hann_gen.py)
    a, res = load(f"hann_{n}_A.npy"), load(f"hann_{n}_res.npy")
    w = O.hann(n)
    got = (a * w[:, None]) * w[None, :]
    np.testing.assert_allclose(got, res, rtol=4 * np.finfo(float).eps, atol=1e-300)


def test_hann_float_is_rounded_double():
    np.testing.assert_array_equal(O.hann(77, np.float32), O.hann(77).astype(np.float32))


# ------------------------------------------------------------------ sums
def test_sum_slice_size_9():  # test/sum_test.cpp:53-66
    v = np.array([1, 3, 5, 8, 8, 10, 4, 1, 3, 2, 5, 15], np.float64)
    s = O.sums(v, [(0, 0), (1, 1)], (4, 3), (2, 2))
    assert list(s) == [22, 21]


@pytest.mark.parametrize("src,sl,begins,batch", [
    ((31, 43), (29, 40), [(0, 0), (1, 1)], 1),
    ((64, 64), (32, 32), [(0, 0), (0, 16), (16, 16), (32, 32)], 1),
    ((64, 64), (32, 32), [(0, 0), (0, 16), (16, 16), (32, 32)], 97),
    ((23, 43), (21, 41), [(0, 0), (1, 1)], 8),
])
def test_sum_slice_arithmetic_series(src, sl, begins, batch):  # test/sum_test.cpp:99-134
    data = np.arange(src[0] * src[1] * batch, dtype=np.float64)
    got = O.sums(data, begins, src, sl, batch)
    img = data.reshape(batch, src[1], src[0])
    exp = [img[b, y:y + sl[1], x:x + sl[0]].sum() for b in range(batch) for (x, y) in begins]
    np.testing.assert_array_equal(got, exp)  # integers < 2^53: exact in any order


def test_sum_u16_is_exact_uint32():
    rng = np.random.default_rng(1)
    pic = rng.integers(0, 65536, size=300 * 200, dtype=np.uint16)
    begins = [(0, 0), (10, 7), (150, 90)]
    got = O.sums(pic, begins, (300, 200), (128, 100), 1)
    img = pic.reshape(200, 300).astype(np.uint64)
    exp = [img[y:y + 100, x:x + 128].sum() for (x, y) in begins]
    assert got.dtype == np.uint32
    np.testing.assert_array_equal(got, np.array(exp) % 2**32)


def test_sum_float_tree_order_within_bound():
    rng = np.random.default_rng(2)
    pic = rng.standard_normal(90 * 90)
    got = O.sums(pic, [(0, 0)], (90, 90), (90, 90))[0]
    exact = __import__("math").fsum(pic)
    assert abs(got - exact) <= 90 * 90 * np.finfo(float).eps * np.abs(pic).sum()


# ------------------------------------------------------------------ prepare
PIC30 = np.arange(1, 31, dtype=np.uint16)


def _prep(pic, begins, wx, wy, sumv, src, sl, out=None, batch=1):
    return O.prepare(pic, np.asarray(wx, np.float64), np.asarray(wy, np.float64), np.asarray(sumv, np.float64),
                     begins, src, sl, out, batch)


def test_prepare_no_window():  # test/prepare_test.cpp:47-78
    s = O.sums(PIC30.astype(np.float64), [(1, 2), (0, 1)], (6, 5), (5, 3))
    got = _prep(PIC30, [(1, 2), (0, 1)], np.ones(6), np.ones(5), s, (6, 5), (5, 3))
    row = [-8, -7, -6, -5, -4, -2, -1, 0, 1, 2, 4, 5, 6, 7, 8]
    assert list(got) == row + row


def test_prepare_no_window_batch():  # test/prepare_test.cpp:80-125
    pic = np.concatenate([PIC30, PIC30.reshape(5, 6)[::-1].ravel()])
    got = _prep(pic, [(1, 2), (0, 1)], np.ones(6), np.ones(5), [330, 225, 0, 0], (6, 5), (5, 3), batch=2)
    exp = [-8, -7, -6, -5, -4, -2, -1, 0, 1, 2, 4, 5, 6, 7, 8] * 2 + \
          [14, 15, 16, 17, 18, 8, 9, 10, 11, 12, 2, 3, 4, 5, 6] + \
          [19, 20, 21, 22, 23, 13, 14, 15, 16, 17, 7, 8, 9, 10, 11]
    assert list(got) == exp


def test_prepare_window():  # test/prepare_test.cpp:128-161
    wx = np.ones(6); wx[2] = 0.5
    wy = np.ones(5); wy[1] = 0.25
    s = O.sums(PIC30.astype(np.float64), [(1, 2), (0, 1)], (6, 5), (5, 3))
    got = _prep(PIC30, [(1, 2), (0, 1)], wx, wy, s, (6, 5), (5, 3))
    row = [-8, -7, -3, -5, -4, -0.5, -.25, 0, .25, .5, 4, 5, 3, 7, 8]
    assert list(got) == row + row


def test_prepare_fft_padded():  # test/prepare_test.cpp:163-210
    s = O.sums(PIC30.astype(np.float64), [(1, 2), (0, 1)], (6, 5), (5, 3))
    got = _prep(PIC30, [(1, 2), (0, 1)], np.ones(6), np.ones(5), s, (6, 5), (5, 3), out=(10, 6)).reshape(2, 6, 10)
    inner = np.array([-8, -7, -6, -5, -4, -2, -1, 0, 1, 2, 4, 5, 6, 7, 8], float).reshape(3, 5)
    for k in range(2):
        exp = np.zeros((6, 10)); exp[:3, :5] = inner
        np.testing.assert_array_equal(got[k], exp)


# ------------------------------------------------------------------ brute correlation
A33 = np.arange(1, 10, dtype=np.int32)


def test_cross_corr_3x3():  # test/cross_corr_test.cpp:17-43
    exp = [9, 26, 50, 38, 21, 42, 94, 154, 106, 54, 90, 186, 285, 186, 90,
           54, 106, 154, 94, 42, 21, 38, 50, 26, 9]
    assert list(O.cross_corr(A33, A33, (3, 3), (5, 5))) == exp


def test_cross_corr_3x3_res3x3():  # test/cross_corr_test.cpp:45-69
    assert list(O.cross_corr(A33, A33, (3, 3), (3, 3))) == [94, 154, 106, 186, 285, 186, 106, 154, 94]


A43 = np.arange(1, 13, dtype=np.int32)
B43 = np.array([8, 7, 6, 5, 16, 15, 14, 13, 4, 3, 2, 1], np.int32)
RES43 = [1, 4, 10, 20, 25, 24, 16, 18, 56, 116, 200, 194, 160, 96, 79, 192, 342, 532, 471, 364, 208,
         142, 316, 524, 768, 638, 468, 256, 45, 104, 178, 268, 229, 172, 96]
A43b = np.array([3, 2, 1, 4, 2, 6, 10, 8, 9, 8, 11, 5], np.int32)
B43b = np.array([1, 2, 3, 4, 16, 5, 14, 13, 8, 3, 2, 1], np.int32)
RES43b = [3, 8, 14, 36, 27, 20, 32, 41, 78, 84, 186, 187, 140, 128, 47, 149, 294, 455, 381, 312, 172,
          125, 260, 362, 479, 303, 227, 88, 36, 59, 86, 78, 45, 21, 5]


def test_cross_corr_4x3():  # test/cross_corr_test.cpp:71-104
    assert list(O.cross_corr(A43, B43, (4, 3), (7, 5))) == RES43


def test_cross_corr_4x3x2():  # test/cross_corr_test.cpp:106-153 (two slices, two refs)
    got = O.cross_corr(np.concatenate([A43, A43b]), np.concatenate([B43, B43b]), (4, 3), (7, 5), ref_slices=2)
    assert list(got) == RES43 + RES43b


def test_cross_corr_4x3x2_res5x5():  # test/cross_corr_test.cpp:155-202
    got = O.cross_corr(np.concatenate([A43, A43b]), np.concatenate([B43, B43b]), (4, 3), (5, 5), ref_slices=2)
    crop = lambda r: list(np.array(r).reshape(5, 7)[:, 1:6].ravel())
    assert list(got) == crop(RES43) + crop(RES43b)


def test_cross_corr_64_fixture():  # test/cross_corr_test.cpp:205-217
    a, b, res = load("cross_corr_64_A.npy"), load("cross_corr_64_B.npy"), load("cross_corr_64_res.npy")
    got = O.cross_corr(a.ravel(), b.ravel(), (64, 64), (127, 127))
    np.testing.assert_array_equal(got, res.ravel())


def test_cross_corr_matches_numpy_bruteforce_double():
    rng = np.random.default_rng(3)
    p, r = rng.standard_normal((5, 7)), rng.standard_normal((5, 7))
    got = O.cross_corr(p.ravel(), r.ravel(), (7, 5), (13, 9)).reshape(9, 13)
    for sy in range(-4, 5):
        for sx in range(-6, 7):
            ys, xs = slice(max(-sy, 0), min(5 - sy, 5)), slice(max(-sx, 0), min(7 - sx, 7))
            ref = (p[ys.start + sy:ys.stop + sy, xs.start + sx:xs.stop + sx] * r[ys, xs]).sum()
            assert abs(got[sy + 4, sx + 6] - ref) < 1e-12


def test_cross_corr_double_uses_fma_order():
    # acc + a*b rounds (1+2^-30)(1-2^-30) to 1 and gives 0; fma keeps -2^-60.
    a = np.array([1.0, 1 + 2**-30], np.float64)
    b = np.array([-1.0, 1 - 2**-30], np.float64)
    got = O.cross_corr(a, b, (2, 1), (1, 1))[0]
    assert got == -2.0**-60


# ------------------------------------------------------------------ FFT-policy correlation
@pytest.mark.parametrize("s", [(4, 4), (5, 3), (8, 8)])
def test_fft_corr_matches_numpy_fft_identity(s):
    rng = np.random.default_rng(4)
    sx, sy = s
    p, r = rng.standard_normal((sy, sx)), rng.standard_normal((sy, sx))
    got = O.fft_corr(p.ravel(), r.ravel(), (sx, sy)).reshape(2 * sy, 2 * sx)
    P = np.fft.rfft2(p, s=(2 * sy, 2 * sx))
    R = np.fft.rfft2(r, s=(2 * sy, 2 * sx))
    ref = np.fft.irfft2(P * np.conj(R), s=(2 * sy, 2 * sx)) * (4 * sx * sy)  # unnormalised C2R
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_fft_corr_relates_to_brute():
    rng = np.random.default_rng(5)
    p, r = rng.standard_normal(36), rng.standard_normal(36)
    fft = O.fft_corr(p, r, (6, 6)).reshape(12, 12)
    brute = O.cross_corr(p, r, (6, 6), (11, 11)).reshape(11, 11)
    # circular index q <-> shift q (q<=s) or q-2s; brute index = shift + 5
    for qy in range(12):
        for qx in range(12):
            sy = qy if qy <= 6 else qy - 12
            sx = qx if qx <= 6 else qx - 12
            if abs(sx) == 6 or abs(sy) == 6:
                assert fft[qy, qx] == 0
            else:
                assert abs(fft[qy, qx] - 144 * brute[sy + 5, sx + 5]) < 1e-9


# ------------------------------------------------------------------ arg-max
def test_maxarg_size_8():  # test/maxarg_test.cpp:63-77
    m = O.maxarg(np.array([1, 3, 5, 8, 10, 4, 3, 2], float), 8)
    assert (m[0]["data"], m[0]["index"]) == (10.0, 4)


def test_maxarg_fftpos():  # test/maxarg_test.cpp:79-130
    v = np.array([1, 3, 0, 8, 10, 4, 0, 3, 0, 0, 0, 0, 1, 1, 0, 1], float)
    m = O.maxarg(np.concatenate([v, v]), 16, 2)
    assert [(x["data"], x["index"]) for x in m] == [(10.0, 4), (10.0, 4)]


def test_maxarg_size_8x2():  # test/maxarg_test.cpp:132-150
    v = np.array([1, 3, 5, 8, 10, 4, 3, 2, 1, 12, 15, 18, 10, 4, 3, 2], float)
    m = O.maxarg(v, 8, 2)
    assert [(x["data"], x["index"]) for x in m] == [(10.0, 4), (18.0, 3)]


def _saw(n):
    out, mom = [], 0.0
    for _ in range(n):
        out.append(mom)
        mom += 0.007
        if mom > 1:
            mom = -2
    return np.array(out)


def test_maxarg_size_1333x3():  # test/maxarg_test.cpp:152-204
    d = _saw(1333)
    d[1001] = 3
    m = O.maxarg(d, 1333)
    assert (m[0]["data"], m[0]["index"]) == (3.0, 1001)
    d = _saw(3999)
    d[1001], d[1024], d[1333], d[3000] = 3, 4, 4, 4
    m = O.maxarg(d, 1333, 3)
    assert [(x["data"], x["index"]) for x in m] == [(4.0, 1024), (4.0, 0), (4.0, 334)]


def test_maxarg_size_4096():  # test/maxarg_test.cpp:206-231
    d = _saw(4096)
    d[1024] = 3
    m = O.maxarg(d, 31 * 43)
    assert (m[0]["data"], m[0]["index"]) == (3.0, 1024)


def test_maxarg_matches_numpy_argmax():
    rng = np.random.default_rng(6)
    d = rng.integers(0, 50, size=7 * 333).astype(np.float32)  # many ties
    m = O.maxarg(d, 333, 7)
    for b in range(7):
        seg = d[b * 333:(b + 1) * 333]
        assert m[b]["index"] == np.argmax(seg) and m[b]["data"] == seg.max()


# ------------------------------------------------------------------ neighbourhood gather
def test_extract_5x5():  # test/maxarg_test.cpp:8-29
    v = np.arange(25, dtype=float)
    mi = np.array([(18.0, 18)], O.data_index_dtype(np.float64))
    assert list(O.extract(v, mi, 3, (5, 5))) == [12, 13, 14, 17, 18, 19, 22, 23, 24]


def test_extract_5x5x2():  # test/maxarg_test.cpp:31-61
    v = np.arange(50, dtype=float)
    mi = np.array([(18.0, 18), (6.0, 6)], O.data_index_dtype(np.float64))
    assert list(O.extract(v, mi, 3, (5, 5), 2)) == [12, 13, 14, 17, 18, 19, 22, 23, 24,
                                                    25, 26, 27, 30, 31, 32, 35, 36, 37]


def test_extract_clamps_one_past_edge():
    # src/extract_neighbors.cu:31-35 clamps to [0, size]: x = size reads the next row's first value
    v = np.arange(50, dtype=float)
    mi = np.array([(0.0, 24), (0.0, 0)], O.data_index_dtype(np.float64))  # (4,4) and (0,0)
    got = O.extract(v, mi, 3, (5, 5), 2).reshape(2, 3, 3)
    # slice 0 at (4,4): x in {3,4,5->5}, y in {3,4,5->5}; (5,y) -> index y*5+5 = next row's col 0
    assert list(got[0].ravel()) == [18, 19, 20, 23, 24, 25, 28, 29, 30]
    # slice 1 at (0,0): x,y in {-1->0, 0, 1}
    assert list(got[1].ravel()) == [25, 25, 26, 25, 25, 26, 30, 30, 31]


def test_extract_fft_policy_maps_through_shift():
    # fft layout: buffer (2s)^2, position p <-> buffer index (p + s + 1) mod 2s
    s = 4
    buf = np.arange(64, dtype=float)
    idx = 2 * 8 + 3  # buffer (3,2) -> position ((3+3)%8, (2+3)%8) = (6,5)
    mi = np.array([(0.0, idx)], O.data_index_dtype(np.float64))
    got = O.extract(buf, mi, 3, (7, 7), 1, O.POLICY_FFT).reshape(3, 3)
    for j in range(3):
        for i in range(3):
            px, py = min(max(6 - 1 + i, 0), 7), min(max(5 - 1 + j, 0), 7)
            qx, qy = (px + s + 1) % 8, (py + s + 1) % 8
            assert got[j, i] == buf[qy * 8 + qx]


# ------------------------------------------------------------------ least squares table and fit
@pytest.mark.parametrize("s", [3, 5, 7, 9])
def test_lstsq_matrix_is_pseudo_inverse(s):
    yi, xi = np.indices((s, s))
    A = np.stack([np.ones(s * s), xi.ravel(), yi.ravel(), xi.ravel() ** 2, (xi * yi).ravel(), yi.ravel() ** 2], 1)
    np.testing.assert_allclose(O.lstsq_matrix(s), np.linalg.pinv(A), atol=1e-13)


@pytest.mark.skipif(not os.path.isdir(REF_TREE), reason="reference tree not mounted")
def test_lstsq_matrix_equals_reference_table_bitwise():
    # src/subpixel_max.cpp:22-58 writes each entry as n.0/d; compare every double bitwise.
    text = open(os.path.join(REF_TREE, "src/subpixel_max.cpp")).read()
    blocks = re.findall(r"lstsq_matrix<double,\s*(\d)>::mat\{(.*?)\};", text, re.S)
    assert len(blocks) == 4
    for s, body in blocks:
        s = int(s)
        vals = []
        for tok in body.replace("\n", " ").split(","):
            tok = tok.strip()
            if not tok:
                continue
            if "/" in tok:
                n, d = tok.split("/")
                vals.append(float(n) / float(d))
            else:
                vals.append(float(tok))
        np.testing.assert_array_equal(O.lstsq_matrix(s).ravel(), np.array(vals))


def test_subpixel_coefs_plane():  # test/subpixel_max_test.cpp:21-28
    _, c = O.subpixel(np.arange(1, 10, dtype=float), 3)
    np.testing.assert_allclose(c[0], [1, 1, 3, 0, 0, 0], atol=1e-15)


def test_subpixel_coefs2_and_peak():  # test/subpixel_max_test.cpp:30-46 (EXPECT_DOUBLE_EQ, 4 ulp)
    o, c = O.subpixel(np.array([1, 2, 3, 7, 9, 8, 4, 5, 6], float), 3)
    assert ulp_equal(c[0], [1.0, 1.833333333333333, 10.5, -0.5, 0, -4.5])
    assert ulp_equal(o[0], (1.8333333333333333, 1.1666666666666666))


@pytest.mark.parametrize("s", [3, 5, 7, 9])
def test_subpixel_recovers_sampled_quadratic(s):
    rng = np.random.default_rng(s)
    for _ in range(20):
        x0, y0 = rng.uniform(0, s - 1, 2)
        d, f = -rng.uniform(0.5, 2, 2)
        e = rng.uniform(-0.3, 0.3)
        yi, xi = np.indices((s, s)).astype(float)
        # f(x,y) = d(x-x0)^2 + e(x-x0)(y-y0) + f(y-y0)^2 + 7
        z = d * (xi - x0) ** 2 + e * (xi - x0) * (yi - y0) + f * (yi - y0) ** 2 + 7
        o, c = O.subpixel(z.ravel(), s)
        assert abs(o[0, 0] - x0) < 1e-9 and abs(o[0, 1] - y0) < 1e-9


def test_subpixel_matches_numpy_lstsq_on_noise():
    rng = np.random.default_rng(9)
    nb = rng.standard_normal((10, 25))
    o, c = O.subpixel(nb.ravel(), 5)
    yi, xi = np.indices((5, 5))
    A = np.stack([np.ones(25), xi.ravel(), yi.ravel(), xi.ravel() ** 2, (xi * yi).ravel(), yi.ravel() ** 2], 1)
    for k in range(10):
        q = np.linalg.lstsq(A, nb[k], rcond=None)[0]
        np.testing.assert_allclose(c[k], q, atol=1e-12)
        xs, ys = np.linalg.solve([[2 * q[3], q[4]], [q[4], 2 * q[5]]], [-q[1], -q[2]])
        np.testing.assert_allclose(o[k], [xs, ys], rtol=1e-9, atol=1e-9)


# ------------------------------------------------------------------ whole path
def _fixture_pics(batch):
    pic = load("data_pic.npy")
    return np.concatenate([pic.ravel()] * batch), load("data_temp.npy")


@pytest.mark.parametrize("policy", [O.POLICY_BRUTE, O.POLICY_FFT])
@pytest.mark.parametrize("batch", [1, 5])
def test_get_offset_batched_goldens(policy, batch):  # test/get_offset_test.cpp:108-199
    pics, temp = _fixture_pics(batch)
    offs, _ = O.get_offset(pics, temp, (256, 256), G.slice_begins(), (64, 64), (127, 127), batch, 3, policy)
    exp = np.tile(np.array(G.BATCHED), (batch, 1))
    assert np.abs(offs - exp).max() <= G.BATCHED_TOL


def test_get_offset_bigger_golden():  # test/get_offset_test.cpp:10-27
    pic, temp = load("data_pic.npy"), load("data_temp.npy")
    a, b = pic[:64, :64].copy(), temp[:64, :64].copy()
    offs, _ = O.get_offset(a, b, (64, 64), [(0, 0)], (64, 64), (127, 127), 1, 3, O.POLICY_BRUTE)
    assert np.abs(offs[0] - np.array(G.BIGGER)).max() <= G.BIGGER_TOL


def test_get_offset_integer_translation_sign():
    # content moved by (+3, -2) in the deformed pattern -> offset = -(displacement)
    rng = np.random.default_rng(10)
    big = rng.standard_normal((80, 80))
    ref = big[10:58, 10:58]
    dfm = big[10 + 2:58 + 2, 10 - 3:58 - 3]  # dfm[y][x] = ref[y+2][x-3]
    offs, _ = O.get_offset(dfm.ravel().copy(), ref.ravel().copy(), (48, 48), [(8, 8)], (32, 32), (63, 63), 1, 3)
    assert np.allclose(offs[0], [-3, 2], atol=0.05)


# ------------------------------------------------------------------ conditioning of a result
def test_peak_stats_equal_numpy_of_the_correlation():
    # get_offset's per-slice statistics: largest value, second largest (ties count
    # twice) and largest magnitude of the correlation it took the arg-max over
    rng = np.random.default_rng(21)
    src, sl, cross, nb, batch = (40, 24), (8, 8), (9, 11), 3, 2
    begins = [(0, 0), (16, 8), (30, 14)]
    pics = rng.integers(0, 4000, size=src[0] * src[1] * batch).astype(np.uint16)
    temp = rng.integers(0, 4000, size=src[0] * src[1]).astype(np.uint16)
    for policy, c in ((O.POLICY_BRUTE, cross), (O.POLICY_FFT, (15, 15))):
        *_, st = O.get_offset(pics, temp, src, begins, sl, cross, batch, 3, policy, np.float64, return_stats=True)
        hx = O.hann(8)
        p = O.prepare(pics, hx, hx, O.sums(pics, begins, src, sl, batch), begins, src, sl, None, batch)
        r = O.prepare(temp, hx, hx, O.sums(temp, begins, src, sl, 1), begins, src, sl, None, 1)
        corr = (O.cross_corr(p, r, sl, c, nb, batch) if policy == O.POLICY_BRUTE else
                O.fft_corr(p, r, sl, nb, batch)).reshape(nb * batch, -1)
        top = np.sort(corr, axis=1)
        np.testing.assert_array_equal(st[:, 0], top[:, -1])
        np.testing.assert_array_equal(st[:, 1], top[:, -2])
        np.testing.assert_array_equal(st[:, 2], np.abs(corr).max(axis=1))


def test_well_conditioned_definition():
    # f = a + bx + cy + dx^2 + exy + fy^2 with its maximum at (1.3, 0.8) inside a 3 x 3
    # neighbourhood, then a saddle, a maximum outside, and a peak without margin
    def coefs(px, py, d=-1.0, f=-2.0, e=0.3):
        b = -(2 * d * px + e * py)
        c = -(2 * f * py + e * px)
        return [0.0, b, c, d, e, f]

    cf = np.array([coefs(1.3, 0.8), coefs(1.3, 0.8, d=1.0), coefs(2.5, 1.0), coefs(1.0, 1.0)])
    st = np.array([[10.0, 5.0, 10.0], [10.0, 5.0, 10.0], [10.0, 5.0, 10.0], [10.0, 10.0 - 1e-6, 10.0]])
    np.testing.assert_array_equal(O.well_conditioned(cf, 3, st), [True, False, False, False])
    # the margin is relative to the largest magnitude
    st2 = st.copy()
    st2[3] = [10.0, 10.0 - 1e-2, 10.0]
    assert O.well_conditioned(cf, 3, st2)[3]
