"""Pins the numpy restatement of the reference's deformation post-processing
(oracle/post.py) with closed forms and independent computations."""
import numpy as np
import pytest

from oracle import post as P

CALIB = np.array([436.0, 436.0, 436.0])


def grid(k=20, n=873, r=32):
    c = np.linspace(r, n - r, k)
    x, y = np.meshgrid(c, c)
    return np.stack([x.ravel(), y.ravel()], 1)


def _norm(f):
    return f / f[2, 2]


def test_exact_projective_map_is_recovered():
    rng = np.random.default_rng(1)
    for _ in range(5):
        Ft = np.eye(3) + rng.normal(0, 2e-3, (3, 3))
        old = grid()
        new = P.apply_transform(Ft, old, CALIB)
        f = P.algebraic_fit(old, new, CALIB)
        assert np.abs(np.linalg.norm(f) - 1) < 1e-12
        np.testing.assert_allclose(_norm(f), _norm(Ft), atol=1e-10)


def test_null_vector_matches_eigh_and_is_minimal():
    rng = np.random.default_rng(2)
    old = grid(10)
    new = P.apply_transform(np.eye(3) + rng.normal(0, 1e-3, (3, 3)), old, CALIB) + rng.normal(0, 0.05, old.shape)
    f = P.algebraic_fit(old, new, CALIB).ravel()
    A = P._rows(old, new, CALIB)
    w, Q = np.linalg.eigh(A.T @ A)
    v = Q[:, 0] * np.sign(Q[:, 0] @ f)
    assert np.abs(v - f).max() < 1e-12
    x = rng.normal(size=(2000, 9))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    assert np.linalg.norm(A @ f) <= np.linalg.norm(x @ A.T, axis=1).min()


def test_rotation_about_the_optical_axis():
    t = 0.01
    R = np.array([[np.cos(t), -np.sin(t), 0], [np.sin(t), np.cos(t), 0], [0, 0, 1]])
    old = grid()
    f = P.algebraic_fit(old, P.apply_transform(R, old, CALIB), CALIB)
    F, eps, omega = P.strain(f)
    np.testing.assert_allclose(F, R, atol=1e-12)
    assert np.abs(eps).max() < 1e-12
    np.testing.assert_allclose(omega, [[0, -t, 0], [t, 0, 0], [0, 0, 0]], atol=1e-12)


def test_stretch_gives_deviatoric_log_strain():
    a, b, c = 2e-3, -1e-3, 5e-4
    S = np.diag(np.exp([a, b, c]))
    old = grid()
    f = P.algebraic_fit(old, P.apply_transform(S, old, CALIB), CALIB)
    with np.errstate(invalid="ignore", divide="ignore"):
        F, eps, omega = P.strain(f)
    m = (a + b + c) / 3
    np.testing.assert_allclose(eps, np.diag([a - m, b - m, c - m]), atol=1e-12)
    assert np.isclose(np.linalg.det(F), 1.0, atol=1e-14)
    assert np.all(np.isnan(omega) | (np.abs(omega) < 1e-10))


def _coefs_for(dets_abc):
    """Coefficients whose quality term a c - b^2 is given by (a, b, c) triples."""
    q = np.zeros((len(dets_abc), 6))
    for i, (a, b, c) in enumerate(dets_abc):
        q[i, 3], q[i, 4], q[i, 5] = -a, -2 * b, -c
    return q


def test_selection_and_median():
    abc = [(1.0, 0.0, 2.0), (0.5, 0.5, 0.5), (-1.0, 0.0, 1.0), (1.0, 0.0, 1e-7), (3.0, 1.0, 1.0)]
    d = P.quality_terms(_coefs_for(abc))
    np.testing.assert_array_equal(d, [2.0, 0.0, -1.0, 1e-7, 2.0])
    assert np.median(d) == 1e-7  # odd count: the middle element
    assert np.median(d[:4]) == (1e-7 + 0.0) / 2  # even: mean of the two middle ones
    assert np.isnan(np.median(np.append(d, np.nan)))
    assert list(d > 1e-6) == [True, False, False, False, True]


def test_post_process_masks_patterns_with_few_good_rois():
    rng = np.random.default_rng(3)
    old = grid(4)  # 16 ROIs
    new = P.apply_transform(np.eye(3) + rng.normal(0, 1e-3, (3, 3)), old, CALIB)
    good = _coefs_for([(1.0, 0.1, 1.0)] * 16)
    bad = good.copy()
    bad[:6, 5] = 0.0  # c = 0: a c - b^2 < 0 on 6 ROIs, 10 left: not more than 10
    data = [np.column_stack([old, new - old, q]) for q in (good, bad)]
    mask, quality, H, F, eps, omega = P.post_process(data, CALIB)
    assert list(mask) == [True, False]
    assert quality[0] == 1.0 - 0.01 and np.isnan(H[1]).all()
    np.testing.assert_allclose(np.linalg.det(F[0]), 1.0, atol=1e-13)
