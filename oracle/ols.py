"""FP64 numpy restatement of the reference's OLS statistics.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``tval_int``      cynumstats.pyx:59-64 with cy_lin_lstsqr_mat (:28-29) and
                    se_of_slope (:47-52): explicit inverse of X'X, se stored as
                    float32, t = a / se.
* ``calc_beta_se``  cynumstats.pyx:66-74 (X = [1, x]).
* ``resid_covars``  cynumstats.pyx:54-57.
* ``calcF``         cynumstats.pyx:31-36 (the builtin ``sum`` over the rows of a
                    2-D y gives one RSS and one TSS per column; the TSS is taken
                    about the grand mean of all of y, ``np.mean(y)``).
* ``calc_sobelz``   pyfunc.py:130-162 for mediation types I, M (Y needs
                    scipy's linregress on two vectors).
* ``perm_t``        tval_int on X[perm] for a batch of permutations, the
                    permutation driver of vertex_tfce_multiple_regression_
                    randomise.py:105-107.
* ``glm_typeI``     pyfunc.py:2282-2360: design [1, exog blocks, covariates],
                    rows permuted by rand_array; the model F, one F per exog
                    block from the fit without that block (SS_Total minus the
                    reduced model's residual sum of squares, cynumstats.pyx:
                    109-112), and t = a / se of every column (se float32; FP64
                    for a one-dimensional endog, :2365-2366).
* ``perm_partial_f`` glm_typeI's per-block F for a batch of permutations
                    (tm_models_randomise.py:190-196).
"""
import numpy as np


def cy_lin_lstsqr_mat(X, y):
    return (np.linalg.inv(X.T.dot(X)).dot(X.T)).dot(y)


def se_of_slope(num_voxel, invXX, sigma2, k):
    # vectorised form of the per-voxel loop; float32 storage as in the reference
    d = np.diag(invXX)
    return np.sqrt(sigma2[None, :] * d[:, None]).astype(np.float32)


def tval_int(X, invXX, y, n, k, numvoxel):
    y = np.asarray(y)
    a = cy_lin_lstsqr_mat(X, y)
    sigma2 = np.sum((y - np.dot(X, a)) ** 2, axis=0) / (n - k)
    se = se_of_slope(numvoxel, invXX, sigma2, k)
    return a / se


def calc_beta_se(x, y, n, num_voxel):
    X = np.column_stack([np.ones(n), x])
    invXX = np.linalg.inv(np.dot(X.T, X))
    k = len(X.T)
    a = cy_lin_lstsqr_mat(X, y)
    beta = a[1]
    sigma2 = np.sum((y - np.dot(X, a)) ** 2, axis=0) / (n - k)
    se = se_of_slope(num_voxel, invXX, sigma2, k)
    return beta, se


def resid_covars(x_covars, data):
    a_c = cy_lin_lstsqr_mat(x_covars, data.T)
    return data.T - np.dot(x_covars, a_c)


def calcF(X, y, n, k):
    a = cy_lin_lstsqr_mat(X, y)
    resids = y - np.dot(X, a)
    RSS = sum(resids ** 2)
    TSS = sum((y - np.mean(y)) ** 2)
    return ((TSS - RSS) / (k - 1)) / (RSS / (n - k))


def sobel_from_t(ta, tb, alg="aroian"):
    if alg == "aroian":
        return 1 / np.sqrt((1 / (tb ** 2)) + (1 / (ta ** 2)) + (1 / (ta ** 2 * tb ** 2)))
    if alg == "sobel":
        return 1 / np.sqrt((1 / (tb ** 2)) + (1 / (ta ** 2)))
    if alg == "goodman":
        return 1 / np.sqrt((1 / (tb ** 2)) + (1 / (ta ** 2)) - (1 / (ta ** 2 * tb ** 2)))
    raise ValueError(alg)


def calc_sobelz(medtype, pred_x, depend_y, merge_y, n, num_vertex, alg="aroian"):
    if medtype == "I":
        a_b, a_se = calc_beta_se(pred_x, merge_y, n, num_vertex)
        b_b, b_se = calc_beta_se(np.column_stack([pred_x, depend_y]), merge_y, n, num_vertex)
    elif medtype == "M":
        a_b, a_se = calc_beta_se(pred_x, merge_y, n, num_vertex)
        b_b, b_se = calc_beta_se(np.column_stack([depend_y, pred_x]), merge_y, n, num_vertex)
    else:
        raise ValueError("type Y needs scipy.stats.linregress; see calc_sobelz_y")
    return sobel_from_t(a_b / a_se[1], b_b / b_se[1], alg)


def perm_t(X, Y, perms, contrasts, perm_columns=None):
    """t [len(contrasts), P, V] in FP64 for X[perm] (all columns or perm_columns)."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    n, k = X.shape
    out = np.zeros((len(contrasts), len(perms), Y.shape[1]))
    for p, perm in enumerate(perms):
        if perm_columns is None:
            nx = X[np.asarray(perm)]
        else:
            nx = X.copy()
            cols = list(perm_columns)
            nx[:, cols] = X[np.asarray(perm)][:, cols]
        invXX = np.linalg.inv(nx.T.dot(nx))
        t = tval_int(nx, invXX, Y, n, k, Y.shape[1])
        out[:, p, :] = t[list(contrasts)]
    return out


def cy_lin_lstsqr_mat_residual(X, y):
    a = cy_lin_lstsqr_mat(X, y)
    resids = y - np.dot(X, a)
    return a, np.sum(resids ** 2, axis=0)


def glm_typeI(endog, exog, dmy_covariates=None, rand_array=None):
    """(Fmodel, Fvar [n_blocks, V], Tvalues [k, V]) of pyfunc.py:2282-2360 with
    output_fvalues and output_tvalues (the p-value and residual outputs are
    left out).  ``exog`` is a list of [n] or [n, w] arrays, one per variable."""
    endog = np.asarray(endog, dtype=np.float64)
    n = endog.shape[0]
    kvars = []
    exog_vars = np.ones(n)
    for var in exog:
        var = np.asarray(var, dtype=np.float64)
        kvars.append(1 if var.ndim == 1 else var.shape[1])
        exog_vars = np.column_stack((exog_vars, var))
    if dmy_covariates is not None:
        exog_vars = np.column_stack((exog_vars, dmy_covariates))
    if rand_array is not None:
        exog_vars = exog_vars[np.asarray(rand_array)]
    k = exog_vars.shape[1]
    df_between, df_within = k - 1, n - k
    ss_total = np.sum((endog - np.mean(endog, 0)) ** 2, 0)
    a, ss_res = cy_lin_lstsqr_mat_residual(exog_vars, endog)
    ss_between = ss_total - ss_res
    ms_res = ss_res / df_within
    fmodel = (ss_between / df_between) / ms_res
    fvar = []
    start = 1
    for w in kvars:
        stop = start + w
        ss_model = ss_total - cy_lin_lstsqr_mat_residual(np.delete(exog_vars, np.s_[start:stop], 1), endog)[1]
        fvar.append((ss_between - ss_model) / (ms_res * w))
        start += w
    sigma2 = np.sum((endog - np.dot(exog_vars, a)) ** 2, axis=0) / (n - k)
    invXX = np.linalg.inv(np.dot(exog_vars.T, exog_vars))
    if endog.ndim == 1:
        se = np.sqrt(np.diag(sigma2 * invXX))  # pyfunc.py:2365-2366: one column, FP64 se
    else:
        se = se_of_slope(endog.shape[1], invXX, sigma2, k)
    return fmodel, np.array(fvar), a / se


def perm_partial_f(X, Y, perms, blocks, perm_columns=None):
    """F [len(blocks), P, V] in FP64: for each permutation, the F of each
    column block (start, width) of X[perm] (all columns, or only
    perm_columns permuted) against the fit without it."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    n, k = X.shape
    out = np.zeros((len(blocks), len(perms), Y.shape[1]))
    ss_total = np.sum((Y - Y.mean(0)) ** 2, 0)
    for p, perm in enumerate(perms):
        if perm_columns is None:
            nx = X[np.asarray(perm)]
        else:
            nx = X.copy()
            cols = list(perm_columns)
            nx[:, cols] = X[np.asarray(perm)][:, cols]
        _, ss_res = cy_lin_lstsqr_mat_residual(nx, Y)
        ms_res = ss_res / (n - k)
        for b, (s0, w) in enumerate(blocks):
            ss_model = ss_total - cy_lin_lstsqr_mat_residual(np.delete(nx, np.s_[s0:s0 + w], 1), Y)[1]
            out[b, p] = ((ss_total - ss_res) - ss_model) / (ms_res * w)
    return out
