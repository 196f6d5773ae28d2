"""Pins the FP64 OLS oracle (oracle/ols.py) with computations that share no
code with it: least squares through QR / lstsq, scipy's linregress for the
one-predictor case, and invariances the t statistic must satisfy."""
import numpy as np
import pytest
from scipy import stats

from oracle import ols


def _data(n=60, k=4, V=50, seed=0):
    rng = np.random.default_rng(seed)
    X = np.column_stack([np.ones(n)] + [rng.standard_normal(n) for _ in range(k - 2)]
                        + [(rng.random(n) < 0.5).astype(float)])
    Y = rng.standard_normal((n, V)) + 0.3 * X[:, -1:] * rng.standard_normal(V)
    return X, Y


def _qr_t(X, Y):
    Q, R = np.linalg.qr(X)
    a = np.linalg.solve(R, Q.T @ Y)
    res = Y - X @ a
    s2 = (res ** 2).sum(0) / (X.shape[0] - X.shape[1])
    Rinv = np.linalg.inv(R)
    d = (Rinv ** 2).sum(1)
    return a / np.sqrt(s2[None, :] * d[:, None])


def test_tval_int_matches_qr():
    X, Y = _data()
    t = ols.tval_int(X, np.linalg.inv(X.T @ X), Y, X.shape[0], X.shape[1], Y.shape[1])
    np.testing.assert_allclose(t, _qr_t(X, Y), rtol=2e-7)  # se is float32 in the reference


def test_single_predictor_matches_linregress():
    rng = np.random.default_rng(3)
    n = 40
    x = rng.standard_normal(n)
    y = 0.4 * x + rng.standard_normal(n)
    X = np.column_stack([np.ones(n), x])
    t = ols.tval_int(X, np.linalg.inv(X.T @ X), y[:, None], n, 2, 1)
    lr = stats.linregress(x, y)
    assert t[1, 0] == pytest.approx(lr.slope / lr.stderr, rel=1e-6)
    beta, se = ols.calc_beta_se(x, y[:, None], n, 1)
    assert beta[0] == pytest.approx(lr.slope, rel=1e-10)
    assert se[1, 0] == pytest.approx(lr.stderr, rel=1e-6)


def test_invariances():
    X, Y = _data(seed=5)
    n, k = X.shape
    t0 = ols.tval_int(X, np.linalg.inv(X.T @ X), Y, n, k, Y.shape[1])
    X2 = X.copy(); X2[:, 1] *= -3.0
    t1 = ols.tval_int(X2, np.linalg.inv(X2.T @ X2), Y, n, k, Y.shape[1])
    np.testing.assert_allclose(t1[1], -t0[1], rtol=1e-6)
    np.testing.assert_allclose(t1[2:], t0[2:], rtol=1e-6)
    Y2 = Y + 2.5 * X[:, 2:3]  # add an in-model column: only that coefficient moves
    t2 = ols.tval_int(X, np.linalg.inv(X.T @ X), Y2, n, k, Y.shape[1])
    np.testing.assert_allclose(t2[[1, 3]], t0[[1, 3]], rtol=1e-6)
    perm = np.random.default_rng(1).permutation(n)
    tp = ols.perm_t(X, Y, [perm], [1, 2, 3])[:, 0]
    nx = X[perm]
    np.testing.assert_allclose(tp, _qr_t(nx, Y)[1:], rtol=2e-7)


def test_resid_covars_orthogonal():
    X, Y = _data(seed=2)
    r = ols.resid_covars(X, Y.T)
    assert np.abs(X.T @ r).max() < 1e-10
    np.testing.assert_allclose(r, Y - X @ np.linalg.lstsq(X, Y, rcond=None)[0], atol=1e-10)


def test_sobel_forms():
    # equal path t-values: closed forms of the three variants
    t = np.array([2.0, 3.0, 5.0])
    np.testing.assert_allclose(ols.sobel_from_t(t, t, "sobel"), t / np.sqrt(2))
    np.testing.assert_allclose(ols.sobel_from_t(t, t, "aroian"), 1 / np.sqrt(2 / t ** 2 + 1 / t ** 4))
    np.testing.assert_allclose(ols.sobel_from_t(t, t, "goodman"), 1 / np.sqrt(2 / t ** 2 - 1 / t ** 4))
    # type M: path b is the partial t of the mediator given x (symmetric in M and y)
    rng = np.random.default_rng(9)
    n = 80
    x = rng.standard_normal(n)
    m = 0.5 * x + rng.standard_normal(n)
    y = 0.4 * m + rng.standard_normal(n)
    z = ols.calc_sobelz("M", x, y, m[:, None], n, 1)
    ta = stats.linregress(x, m).slope / stats.linregress(x, m).stderr
    tb = _qr_t(np.column_stack([np.ones(n), y, x]), m[:, None])[1, 0]
    tb2 = _qr_t(np.column_stack([np.ones(n), m, x]), y[:, None])[1, 0]
    assert tb == pytest.approx(tb2, rel=1e-9)
    assert z[0] == pytest.approx(1 / np.sqrt(1 / tb ** 2 + 1 / ta ** 2 + 1 / (ta ** 2 * tb ** 2)), rel=1e-6)


def test_calcF_two_columns_is_t_squared():
    """k = 2: the F of [1, x] against the intercept-only model is the square
    of the slope's t (scipy's linregress, no shared code)."""
    rng = np.random.default_rng(11)
    n = 57
    x = rng.standard_normal(n)
    y = 0.3 * x + rng.standard_normal(n)
    X = np.column_stack([np.ones(n), x])
    F = ols.calcF(X, y[:, None], n, 2)
    lr = stats.linregress(x, y)
    assert F.shape == (1,)
    assert F[0] == pytest.approx((lr.slope / lr.stderr) ** 2, rel=1e-10)


def test_calcF_matches_r_squared_of_lstsq():
    """F = (R^2 / (k-1)) / ((1 - R^2) / (n-k)) with R^2 from np.linalg.lstsq,
    for every column of a 2-D y whose columns share one mean (calcF takes
    the TSS about the grand mean of all of y, which is then each column's)."""
    rng = np.random.default_rng(12)
    n, k, V = 80, 4, 30
    X = np.column_stack([np.ones(n), rng.standard_normal((n, k - 1))])
    Y = rng.standard_normal((n, V)) + 0.2 * X[:, 1:2] * rng.standard_normal(V)
    Y = Y - Y.mean(0) + 1.5  # every column's mean is 1.5
    F = ols.calcF(X, Y, n, k)
    coef, rss, _, _ = np.linalg.lstsq(X, Y, rcond=None)
    tss = ((Y - Y.mean(0)) ** 2).sum(0)
    r2 = 1 - rss / tss
    np.testing.assert_allclose(F, (r2 / (k - 1)) / ((1 - r2) / (n - k)), rtol=1e-9)
    # a column with its own mean: the grand-mean TSS adds n * (mean - grand)^2
    Y2 = Y.copy()
    Y2[:, 0] += 2.0
    F2 = ols.calcF(X, Y2, n, k)
    g = Y2.mean()
    tss0 = ((Y2[:, 0] - g) ** 2).sum()
    rss0 = rss[0]
    assert F2[0] == pytest.approx(((tss0 - rss0) / (k - 1)) / (rss0 / (n - k)), rel=1e-9)
    assert tss0 == pytest.approx(tss[0] + n * (Y2[:, 0].mean() - g) ** 2, rel=1e-12)


def _dummies(labels, g):
    """Treatment coding: g - 1 indicator columns (level 0 is the reference)."""
    return np.column_stack([(labels == j).astype(float) for j in range(1, g)])


def test_glm_typeI_one_factor_is_one_way_anova():
    """Intercept + one dummy-coded factor of g levels: the factor's F (and the
    model F) is scipy's one-way ANOVA F of the groups, element by element."""
    rng = np.random.default_rng(21)
    n, g, V = 45, 4, 12
    labels = np.arange(n) % g
    rng.shuffle(labels)
    Y = rng.standard_normal((n, V)) + 0.5 * labels[:, None] * rng.standard_normal(V)
    fmodel, fvar, _ = ols.glm_typeI(Y, [_dummies(labels, g)])
    ref = np.array([stats.f_oneway(*[Y[labels == j, v] for j in range(g)]).statistic for v in range(V)])
    np.testing.assert_allclose(fvar[0], ref, rtol=1e-9)
    np.testing.assert_allclose(fmodel, ref, rtol=1e-9)


def test_glm_typeI_balanced_two_factor_closed_form():
    """Balanced additive two-factor design (a x b cells, r per cell): the F of
    each factor is SS_factor / (levels - 1) over the additive model's mean
    square error, with SS_A = b r sum_i (mean_i - grand)^2 (no shared code)."""
    rng = np.random.default_rng(22)
    A, B, r, V = 3, 4, 5, 7
    la = np.repeat(np.arange(A), B * r)
    lb = np.tile(np.repeat(np.arange(B), r), A)
    n = A * B * r
    Y = rng.standard_normal((n, V)) + 0.4 * la[:, None] - 0.3 * (lb == 2)[:, None]
    _, fvar, _ = ols.glm_typeI(Y, [_dummies(la, A), _dummies(lb, B)])
    g = Y.mean(0)
    ssa = B * r * sum((Y[la == i].mean(0) - g) ** 2 for i in range(A))
    ssb = A * r * sum((Y[lb == j].mean(0) - g) ** 2 for j in range(B))
    sst = ((Y - g) ** 2).sum(0)
    mse = (sst - ssa - ssb) / (n - (1 + (A - 1) + (B - 1)))
    np.testing.assert_allclose(fvar[0], (ssa / (A - 1)) / mse, rtol=1e-9)
    np.testing.assert_allclose(fvar[1], (ssb / (B - 1)) / mse, rtol=1e-9)


def test_glm_typeI_single_columns_are_t_squared_and_reparametrisation_free():
    """A one-column block's F is the square of that column's t (tval_int,
    pinned by QR above; se is float32 there); a block's F does not change when
    its columns are replaced by an invertible mix of them, nor do the other
    blocks' F; the t-values are tval_int's on X[rand_array]."""
    rng = np.random.default_rng(23)
    n, V = 70, 15
    x1 = rng.standard_normal(n)
    blk = rng.standard_normal((n, 3))
    cov = rng.standard_normal((n, 2))
    Y = rng.standard_normal((n, V)) + 0.4 * x1[:, None] + 0.3 * blk[:, :1]
    perm = rng.permutation(n)
    fmodel, fvar, t = ols.glm_typeI(Y, [x1, blk], dmy_covariates=cov, rand_array=perm)
    X = np.column_stack([np.ones(n), x1, blk, cov])[perm]
    tq = _qr_t(X, Y)
    np.testing.assert_allclose(fvar[0], tq[1] ** 2, rtol=2e-6)
    np.testing.assert_allclose(t, tq, rtol=1e-6)
    T = rng.standard_normal((3, 3)) + 3 * np.eye(3)
    _, fvar2, _ = ols.glm_typeI(Y, [x1, blk @ T], dmy_covariates=cov, rand_array=perm)
    np.testing.assert_allclose(fvar2, fvar, rtol=1e-9)
    # the model F against R^2 of lstsq
    _, rss, _, _ = np.linalg.lstsq(X, Y, rcond=None)
    r2 = 1 - rss / ((Y - Y.mean(0)) ** 2).sum(0)
    k = X.shape[1]
    np.testing.assert_allclose(fmodel, (r2 / (k - 1)) / ((1 - r2) / (n - k)), rtol=1e-9)
    # the batched form
    blocks = [(1, 1), (2, 3)]
    fp = ols.perm_partial_f(np.column_stack([np.ones(n), x1, blk, cov]), Y, [perm], blocks)
    np.testing.assert_allclose(fp[:, 0], fvar, rtol=1e-12)
