"""numpy restatement of the step-1 element-wise regressor and VIF image.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``zscaler``           pyfunc.py:264-271.
* ``image_regression``  pyfunc.py:1024-1086: per element i, the design
                        [1, pred_x, covars, image_x[i]] (float64), inv(X'X),
                        a = cy_lin_lstsqr_mat(X, y[i]), sigma2 = sum of squared
                        residuals / (n - k), se = sqrt(diag(sigma2 * invXX)) in
                        float64, t = a / se; zeros where image_x[i] holds a NaN
                        or has std() < 0.01.  Returns float32 (Tval [V, p + 1]
                        with p = len(pred_x.T), Timg [V]).
* ``image_reg_VIF``     pyfunc.py:1088-1114: R2 = 1 - RSS / TSS of y on
                        [1, regressors], VIF = 1 / (1 - R2), NaN set to 0; TSS
                        is numpy's ``np.sum((y - np.mean(y, axis=0))**2, axis=0)``
                        in y's own dtype.
"""
import numpy as np

from .ols import cy_lin_lstsqr_mat


def zscaler(X, axis=0, w_mean=True, w_std=True):
    data = np.zeros_like(X)
    data[:] = np.copy(X)
    if w_mean:
        data -= np.mean(data, axis)
    if w_std:
        data /= np.std(data, axis)
    return data


def image_regression(y, image_x, pred_x, covars=None, normalize=False):
    nv, n = y.shape
    regressors = pred_x if covars is None else np.column_stack((pred_x, covars))
    regressors = np.column_stack([np.ones(len(regressors)), regressors])
    arr = np.zeros((nv, regressors.shape[1] + 1))
    if normalize:
        image_x = zscaler(image_x, axis=0, w_mean=True, w_std=True)
    for i in range(nv):
        zi = image_x[i, :]
        if zi.std() < 0.01 or np.any(np.isnan(zi)):
            continue
        yi = y[i, :]
        X = np.column_stack((regressors, zi))
        k = X.shape[1]
        invXX = np.linalg.inv(np.dot(X.T, X))
        a = cy_lin_lstsqr_mat(X, yi)
        sigma2 = np.sum((yi - np.dot(X, a)) ** 2, axis=0) / (n - k)
        se = np.sqrt(np.diag(sigma2 * invXX))
        arr[i] = a / se
    Tval = np.array(arr[:, :len(np.asarray(pred_x).T) + 1], dtype=np.float32)
    Timg = np.array(np.squeeze(arr[:, -1:]), dtype=np.float32)
    return Tval, Timg


def image_reg_VIF(y, regressors):
    X = np.column_stack([np.ones(len(regressors)), regressors])
    a = cy_lin_lstsqr_mat(X, y)
    resids = y - np.dot(X, a)
    RSS = np.sum(resids ** 2, axis=0)
    TSS = np.sum((y - np.mean(y, axis=0)) ** 2, axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):
        R2 = 1 - (RSS / TSS)
        VIF = 1 / (1 - R2)
    VIF[np.isnan(VIF)] = 0
    return VIF
