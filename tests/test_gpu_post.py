"""The deformation post-step on the GPU (emida_fit_deformation) against the numpy
restatement of the reference's post-processing (oracle/post.py):

  * ROI selection, mask and median quality bit-identical (same arithmetic, no
    contraction; NaN determinants give NaN quality);
  * H (map / [2][2]), F (determinant 1), strain and rotation within 1e-9 of numpy's
    SVD-based results, for shared and per-pattern ROI middles, with and without
    coefficients, at the selection boundary (10 vs 11 ROIs);
  * composition: the brute double pipeline followed by the post-step equals the
    oracle's offsets followed by the numpy post-processing.
"""
import numpy as np
import pytest

import emida_amd
from emida_amd import post as G
from emida_amd import synth
from oracle import post as P

pytestmark = pytest.mark.gpu
CALIB = np.array([436.0, 436.0, 436.0])


def grid(k=20, n=873, r=32):
    c = np.linspace(r, n - r, k)
    x, y = np.meshgrid(c, c)
    return np.stack([x.ravel(), y.ravel()], 1)


def _data(n_pat, k, seed, noise=0.02, bad_frac=0.2, nan_rois=0):
    rng = np.random.default_rng(seed)
    mids = grid(k)
    offs, coefs = [], []
    for p in range(n_pat):
        Ft = np.eye(3) + rng.normal(0, 2e-3, (3, 3))
        new = P.apply_transform(Ft, mids, CALIB) + rng.normal(0, noise, mids.shape)
        q = rng.normal(0, 1, (len(mids), 6))
        q[:, 3] = -rng.uniform(0.5, 2.0, len(mids))
        q[:, 5] = -rng.uniform(0.5, 2.0, len(mids))
        q[:, 4] = rng.normal(0, 0.3, len(mids))
        bad = rng.random(len(mids)) < bad_frac
        q[bad, 5] = rng.uniform(0.0, 0.5, bad.sum())  # c < 0: not a maximum
        if nan_rois:
            q[rng.choice(len(mids), nan_rois, replace=False), 3] = np.nan
        offs.append(new - mids)
        coefs.append(q)
    return mids, np.array(offs), np.array(coefs)


def _compare(res, mids, offs, coefs, min_rois=10):
    data = [np.column_stack([mids if mids.ndim == 2 else mids[i], offs[i], coefs[i]]) for i in range(len(offs))]
    mask, quality, H, F, eps, omega = P.post_process(data, CALIB, min_rois=min_rois)
    np.testing.assert_array_equal(res["mask"], mask)
    np.testing.assert_array_equal(res["quality"], quality)  # bit for bit, NaN included
    for name, exp in (("H", H), ("F", F), ("epsilon", eps), ("omega", omega)):
        got = res[name]
        assert np.array_equal(np.isnan(got[~mask]), np.isnan(exp[~mask]))
        assert np.abs(got[mask] - exp[mask]).max() < 1e-9, name


@pytest.mark.parametrize("n_pat,k,seed", [(1, 20, 0), (37, 20, 1), (8, 31, 2), (5, 4, 3)])
def test_fit_deformation_vs_oracle(n_pat, k, seed):
    mids, offs, coefs = _data(n_pat, k, seed)
    res = G.fit_deformation(mids, offs, coefs, CALIB)
    _compare(res, mids, offs, coefs)


def test_per_pattern_mids_and_nan_quality():
    mids, offs, coefs = _data(6, 12, 4, nan_rois=3)
    rng = np.random.default_rng(5)
    mids_p = mids[np.newaxis] + rng.uniform(-3, 3, (6,) + mids.shape)
    res = G.fit_deformation(mids_p, offs, coefs, CALIB)
    assert np.isnan(res["quality"]).all()
    _compare(res, mids_p, offs, coefs)


def test_selection_boundary():
    mids, offs, coefs = _data(2, 4, 6, bad_frac=0.0)  # 16 ROIs each
    coefs[0, :5, 5] = 1.0   # 11 ROIs left: fitted
    coefs[1, :6, 5] = 1.0   # 10 left: not fitted
    res = G.fit_deformation(mids, offs, coefs, CALIB)
    assert list(res["mask"]) == [True, False]
    _compare(res, mids, offs, coefs)


def test_algebraic_fit_without_coefficients():
    rng = np.random.default_rng(7)
    old = grid(7)
    Ft = np.eye(3) + rng.normal(0, 1e-3, (3, 3))
    new = P.apply_transform(Ft, old, CALIB) + rng.normal(0, 0.01, old.shape)
    f = G.algebraic_fit(old, new, CALIB)
    e = P.algebraic_fit(old, new, CALIB)
    e = e * np.sign(e[2, 2])
    assert np.abs(f - e).max() < 1e-10
    with pytest.raises(ValueError):
        G.algebraic_fit(old[:4], new[:4], CALIB)


def test_post_process_api_and_output_parser():
    mids, offs, coefs = _data(4, 6, 8)
    data = np.concatenate([np.repeat(mids[np.newaxis], 4, 0), offs, coefs], axis=2)
    mask, F, eps, omega, quality = G.post_process(data, CALIB)
    em, eq, _, eF, ee, eo = P.post_process(list(data), CALIB)
    np.testing.assert_array_equal(mask, em)
    assert F.shape == (mask.sum(), 3, 3) and np.abs(F - eF[em]).max() < 1e-9
    text = "".join(f"{60.0 * i:f} 0.000000 {len(mids)} p{i}.tif\n" +
                   "".join(f"{int(x)} {int(y)} " + " ".join(f"{v:f}" for v in row) + "\n"
                           for (x, y), row in zip(mids, np.column_stack([offs[i], coefs[i]])))
                   for i in range(4))
    parsed = G.read_output(text)
    assert [p[2] for p in parsed] == [f"p{i}.tif" for i in range(4)]
    assert parsed[1][3].shape == (len(mids), 10)


def test_pipeline_then_post_step_equals_oracle_composition():
    # uint16 band patterns through the brute double pipeline (bit-identical to the
    # oracle's offsets and coefficients), then the post-step on the GPU, against the
    # oracle's offsets fed to the numpy post-processing
    from oracle import oracle as O

    ps = synth.PatternSet("post", 256, 4, t_range=2.0)
    begins = synth.grid_begins(256, 6, 16)
    with emida_amd.GpuOffset(np.float64, np.uint16, ps.size, begins, (32, 32), (63, 63), 4,
                             emida_amd.CROSS_POLICY_BRUTE, 3) as g:
        g.allocate_memory(ps.ref)
        offs, coefs = g.get_offset(ps.pics)
    eo, ec = O.get_offset(ps.pics.ravel(), ps.ref, ps.size, begins, (32, 32), (63, 63), 4, 3, O.POLICY_BRUTE,
                          np.float64)
    np.testing.assert_array_equal(offs, eo)
    mids = np.array([(x + 16, y + 16) for x, y in begins], float)
    calib = np.array([128.0, 128.0, 128.0])
    res = G.fit_deformation(mids, offs.reshape(4, -1, 2), coefs.reshape(4, -1, 6), calib)
    data = [np.column_stack([mids, eo.reshape(4, -1, 2)[i], ec.reshape(4, -1, 6)[i]]) for i in range(4)]
    mask, quality, H, F, eps, omega = P.post_process(data, calib)
    np.testing.assert_array_equal(res["mask"], mask)
    np.testing.assert_array_equal(res["quality"], quality)
    assert mask.all()
    for got, exp in ((res["H"], H), (res["F"], F), (res["epsilon"], eps), (res["omega"], omega)):
        assert np.abs(got - exp).max() < 1e-9
