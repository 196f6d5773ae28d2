"""The oracle of the step-1 element-wise regressor and VIF image
(oracle/imgreg.py, pyfunc.py:1024-1114), pinned by what the algebra fixes,
and the C-ABI's argument checks (host only, no GPU)."""
import ctypes

import numpy as np
import pytest
from scipy import stats

from oracle import imgreg as oimg
from oracle import ols as ools
from tfce_mediation_amd import _lib


def data(seed, V=60, n=40, p=2, c=1):
    rng = np.random.default_rng(seed)
    pred = rng.standard_normal((n, p))
    cov = rng.standard_normal((n, c))
    img = rng.standard_normal((V, n)).astype(np.float32)
    y = (0.4 * img + pred[:, 0] * 0.3 + rng.standard_normal((V, n))).astype(np.float32)
    return y, img, pred, cov


def test_same_regressor_everywhere_is_one_shared_design():
    """With the same regressor at every element, the model is one OLS with
    the design [1, pred, cov, z]: every t equals tval_int's (se stored as
    float32 there, double here)."""
    y, img, pred, cov = data(1)
    z = img[0]
    img_same = np.tile(z, (y.shape[0], 1))
    Tval, Timg = oimg.image_regression(y, img_same, pred, cov)
    X = np.column_stack([np.ones(len(z)), pred, cov, z.astype(np.float64)])
    n, k = X.shape
    t = ools.tval_int(X, np.linalg.inv(X.T @ X), y.T.astype(np.float64), n, k, y.shape[0])
    np.testing.assert_allclose(Tval, t[: pred.shape[1] + 1].T, rtol=1e-6)
    np.testing.assert_allclose(Timg, t[-1], rtol=1e-6)


def test_regressor_t_by_frisch_waugh_lovell():
    """The t of the element's regressor from the partialled-out fit
    (Frisch-Waugh-Lovell: residualise y and z on the shared columns, slope
    of one on the other, sigma2 of the full model), with least squares by
    QR (np.linalg.lstsq)."""
    y, img, pred, cov = data(2)
    _, Timg = oimg.image_regression(y, img, pred, cov)
    n = y.shape[1]
    R = np.column_stack([np.ones(n), pred, cov])
    k = R.shape[1] + 1
    for i in (0, 7, 33):
        yi = y[i].astype(np.float64)
        zi = img[i].astype(np.float64)
        ry = yi - R @ np.linalg.lstsq(R, yi, rcond=None)[0]
        rz = zi - R @ np.linalg.lstsq(R, zi, rcond=None)[0]
        b = (rz @ ry) / (rz @ rz)
        sse = np.sum((ry - b * rz) ** 2)
        t = b / np.sqrt(sse / (n - k) / (rz @ rz))
        assert abs(Timg[i] - t) <= 1e-6 * abs(t)


def test_tval_columns_follow_len_of_pred_x_T():
    """Tval keeps the first len(pred_x.T) + 1 columns: intercept and pred_x
    for a two-dimensional pred_x, every column for a one-dimensional one
    (len(pred_x.T) is then n)."""
    y, img, pred, cov = data(3)
    Tval, _ = oimg.image_regression(y, img, pred[:, 0], cov)
    assert Tval.shape == (y.shape[0], 1 + 1 + cov.shape[1] + 1)
    Tval2, _ = oimg.image_regression(y, img, pred, cov)
    assert Tval2.shape == (y.shape[0], 1 + pred.shape[1])


def test_flat_or_nan_regressor_gives_zeros():
    y, img, pred, cov = data(4)
    img = img.copy()
    img[3] = 5.0
    img[4, 0] = np.nan
    img[5] = img[5] * 0.004  # std below 0.01
    Tval, Timg = oimg.image_regression(y, img, pred, cov)
    for i in (3, 4, 5):
        assert Timg[i] == 0 and np.all(Tval[i] == 0)
    assert np.all(Timg[6:] != 0)


def test_vif_single_regressor_is_one_over_one_minus_r_squared():
    """One regressor: R2 is the squared correlation (scipy.stats.linregress);
    float64 y takes the sum of squares in float64."""
    rng = np.random.default_rng(5)
    n, V = 50, 30
    x = rng.standard_normal(n)
    y = 0.7 * x[:, None] + rng.standard_normal((n, V))
    vif = oimg.image_reg_VIF(y, x)
    for v in range(V):
        r = stats.linregress(x, y[:, v]).rvalue
        assert abs(vif[v] - 1.0 / (1.0 - r * r)) <= 1e-10 * vif[v]


def test_vif_float32_sum_of_squares_and_degenerate_columns():
    """float32 y: the sum of squares about the mean is numpy's float32
    reduction over the first axis (rounding of order 1e-7); a constant column
    gives RSS = TSS = 0, NaN, set to 0."""
    rng = np.random.default_rng(6)
    n, V = 50, 20
    reg = rng.standard_normal((n, 2))
    y = (reg @ rng.standard_normal((2, V)) + rng.standard_normal((n, V))).astype(np.float32)
    y[:, 3] = 2.5
    vif32 = oimg.image_reg_VIF(y, reg)
    vif64 = oimg.image_reg_VIF(y.astype(np.float64), reg)
    assert vif32[3] == 0
    ok = np.arange(V) != 3
    np.testing.assert_allclose(vif32[ok], vif64[ok], rtol=1e-5)
    assert not np.array_equal(vif32[ok], vif64[ok])


def test_zscaler_columns():
    rng = np.random.default_rng(7)
    X = rng.standard_normal((40, 6)) * 3 + 2
    Z = oimg.zscaler(X, axis=0)
    np.testing.assert_allclose(Z.mean(0), 0, atol=1e-12)
    np.testing.assert_allclose(Z.std(0), 1, rtol=1e-12)


def test_abi_refuses_bad_arguments():
    """Argument checks happen on the host before any GPU work."""
    lib = _lib.load_library()
    EINVAL = -1000
    R = np.ones((10, 2))
    buf = ctypes.c_void_p(8)  # never dereferenced: the checks fail first
    assert lib.tfce_image_regression(buf, 5, buf, 5, R.ctypes.data, 10, 0, 5, 0.01, buf, 5, None) == EINVAL
    assert lib.tfce_image_regression(buf, 5, buf, 5, R.ctypes.data, 2, 2, 5, 0.01, buf, 5, None) == EINVAL
    assert lib.tfce_image_regression(None, 5, buf, 5, R.ctypes.data, 10, 2, 5, 0.01, buf, 5, None) == EINVAL
    assert lib.tfce_image_vif(buf, 0, 5, R.ctypes.data, 10, 9, 5, buf, None) == EINVAL
    assert lib.tfce_image_vif(buf, 0, 4, R.ctypes.data, 10, 2, 5, buf, None) == EINVAL
    Rs = np.ones((10, 2))  # two equal columns: R'R singular
    assert lib.tfce_image_regression(buf, 5, buf, 5, Rs.ctypes.data, 10, 2, 5, 0.01, buf, 5, None) == EINVAL
    assert b"singular" in lib.tfce_last_error()
