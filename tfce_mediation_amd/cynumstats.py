"""Drop-in for ``tfce_mediation.cynumstats`` (cynumstats.pyx:28-112).

Same names, arguments and return types as the reference; the per-element work
runs on the GPU in FP64 (``tfce_ols_direct``), the way numpy runs it on the
host in the reference.  ``y`` / ``data`` arrays are read as float32, which is
how every reference call site stores them (merge_y is saved as float32,
STEP_1_vertex_tfce_multiple_regression.py:263).  ``calc_gd_fwhm`` (geodesic
smoothing, :83-107) is not part of the permutation path and is not provided.

Batched permutations do not come through here: see ``ols.PermutedOLS``.
"""
import ctypes

import numpy as np

from . import _lib


def _as_device_y(y):
    import torch
    if isinstance(y, torch.Tensor):
        t = y.to(device="cuda", dtype=torch.float32)
    else:
        t = torch.as_tensor(np.ascontiguousarray(np.asarray(y), dtype=np.float32)).cuda()
    if t.dim() == 1:
        t = t[:, None]
    return t.contiguous()


def _direct(X, y, invXX_se=None, coef=False, se=False, t=False, resid=False, sse=False, sums=False):
    import torch
    lib = _lib.lib()
    X = np.ascontiguousarray(np.asarray(X, dtype=np.float64))
    if X.ndim == 1:
        X = X[:, None]
    n, k = X.shape
    Y = _as_device_y(y)
    if Y.shape[0] != n:
        raise ValueError("shapes of X and y do not align")
    V = int(Y.shape[1])
    dev = Y.device
    out = {}
    def buf(name, shape, dt):
        out[name] = torch.empty(shape, dtype=dt, device=dev)
        return out[name]
    d_coef = buf("coef", (k, V), torch.float64) if coef else None
    d_se = buf("se", (k, V), torch.float32) if se else None
    d_t = buf("t", (k, V), torch.float64) if t else None
    d_res = buf("resid", (n, V), torch.float64) if resid else None
    d_sse = buf("sse", (V,), torch.float64) if sse else None
    d_sy = buf("sumy", (V,), torch.float64) if sums else None
    d_syy = buf("sumyy", (V,), torch.float64) if sums else None
    inv = None
    if invXX_se is not None:
        inv = np.ascontiguousarray(np.asarray(invXX_se, dtype=np.float64))
    _lib.check(lib.tfce_ols_direct(_lib.ptr(Y), n, V, Y.stride(0), X.ctypes.data, k,
                                   inv.ctypes.data if inv is not None else None,
                                   _lib.ptr(d_coef), _lib.ptr(d_se), _lib.ptr(d_t), _lib.ptr(d_res),
                                   _lib.ptr(d_sse), _lib.ptr(d_sy), _lib.ptr(d_syy), _lib.stream_ptr()))
    return {k_: v.cpu().numpy() for k_, v in out.items()}


def cy_lin_lstsqr_mat(X, y):
    """inv(X'X) X' y  (cynumstats.pyx:28-29)."""
    return _direct(X, y, coef=True)["coef"]


def cy_lin_lstsqr_mat_residual(exog_vars, endog_arr):
    """(coefficients, residual sum of squares)  (cynumstats.pyx:109-112)."""
    r = _direct(exog_vars, endog_arr, coef=True, sse=True)
    return r["coef"], r["sse"]


def calcF(X, y, n, k):
    """F of the model against the grand-mean total sum of squares
    (cynumstats.pyx:31-36: TSS uses np.mean over every element of y)."""
    r = _direct(X, y, sse=True, sums=True)
    N = r["sumy"].size
    m = r["sumy"].sum() / (float(n) * N)
    tss = r["sumyy"] - 2.0 * m * r["sumy"] + float(n) * m * m
    rss = r["sse"]
    return ((tss - rss) / (k - 1)) / (rss / (n - k))


def cython_lstsqr(x, y):
    """(intercept, slope) of y on x  (cynumstats.pyx:38-45)."""
    x = np.asarray(x, dtype=np.float64)
    a = _direct(np.column_stack([np.ones(len(x)), x]), np.asarray(y, dtype=np.float32), coef=True)["coef"]
    return float(a[0, 0]), float(a[1, 0])


def se_of_slope(num_voxel, invXX, sigma2, k):
    """float32 [k, V] of sqrt(diag(sigma2[j] * invXX))  (cynumstats.pyx:47-52)."""
    import torch
    s2 = torch.as_tensor(np.ascontiguousarray(np.asarray(sigma2, dtype=np.float64)[:num_voxel])).cuda()
    inv = np.ascontiguousarray(np.asarray(invXX, dtype=np.float64))
    out = torch.empty((k, int(num_voxel)), dtype=torch.float32, device="cuda")
    _lib.check(_lib.lib().tfce_se_of_slope(_lib.ptr(s2), inv.ctypes.data, int(k), int(num_voxel),
                                           _lib.ptr(out), _lib.stream_ptr()))
    return out.cpu().numpy()


def resid_covars(x_covars, data):
    """Residuals [n, V] of data.T on x_covars  (cynumstats.pyx:54-57)."""
    return _direct(x_covars, np.asarray(data).T, resid=True)["resid"]


def tval_int(X, invXX, y, n, k, numvoxel):
    """t-values [k, V] float64; se from the caller's invXX, stored float32
    (cynumstats.pyx:59-64)."""
    return _direct(X, y, invXX_se=invXX, t=True)["t"]


def calc_beta_se(x, y, n, num_voxel):
    """(beta of x, se [k, V] float32) with X = [1, x]  (cynumstats.pyx:66-74)."""
    X = np.column_stack([np.ones(n), x])
    r = _direct(X, y, coef=True, se=True)
    return r["coef"][1], r["se"]
