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
