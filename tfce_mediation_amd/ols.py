"""Batched permuted OLS t-values on the GPU (tfce_ols_* of libtfcehip).

``PermutedOLS(Y, X, ...)`` prepares Y once (centring, FP64 y'y, split planes)
and binds the design; ``t(perms)`` returns the t-values of the requested
contrasts for every permutation of a batch, ``[n_contrasts, P, V]`` float32 on
the device.  Row i of the permuted design is ``X[perm[i]]`` as in
``vertex_tfce_multiple_regression_randomise.py:105`` (all columns) or, with
``perm_columns``, only the listed columns move (``-v``, :93-97).
``fstat(perms, tss)`` gives sqrt(max(F, 0)) of the whole model instead (the
ancova branch, voxel_tfce_multiple_regression_randomise.py:79-89), with
``total_ss(Y)`` the total sum of squares calcF uses (cynumstats.pyx:31-36).
"""
import ctypes

import numpy as np

from . import _lib


class PermutedOLS:
    def __init__(self, Y, X, contrasts=None, perm_columns=None, precision="fp16x3", center=None,
                 stream=None, data_from=None):
        """data_from: another PermutedOLS whose prepared data (same Y, same
        precision and centring) this model binds instead of preparing Y again;
        Y is then ignored."""
        import torch
        lib = _lib.lib()
        X = np.ascontiguousarray(np.asarray(X, dtype=np.float64))
        if X.ndim == 1:
            X = X[:, None]
        n, k = X.shape
        const = [bool(np.all(X[:, j] == X[0, j]) and X[0, j] != 0) for j in range(k)]
        if data_from is not None:
            assert data_from.n == n, "the shared data has another number of subjects"
            self.n, self.k, self.V = n, k, data_from.V
            self._owner = data_from  # keeps the data handle alive
            self.precision = data_from.precision
            self._data = None
            self._bind(X, k, const, contrasts, perm_columns, data_from._data, stream)
            return
        if not (isinstance(Y, torch.Tensor) and Y.is_cuda):
            Y = torch.as_tensor(np.ascontiguousarray(Y, dtype=np.float32)).cuda()
        Y = Y.to(torch.float32)
        if Y.dim() == 1:
            Y = Y[:, None]
        if Y.stride(1) != 1:
            Y = Y.contiguous()
        assert Y.shape[0] == n, "Y must be [n_subjects, n_elements]"
        self.n, self.k, self.V = n, k, int(Y.shape[1])
        if center is None:
            center = any(const)
        self.precision = precision
        self._Y = Y  # kept alive until the prep kernel has run
        data = ctypes.c_void_p()
        _lib.check(lib.tfce_ols_data_create(_lib.ptr(Y), n, self.V, Y.stride(0), _lib.PREC[precision],
                                            1 if center else 0, ctypes.byref(data),
                                            _lib.stream_ptr(stream)))
        self._data = data
        self._bind(X, k, const, contrasts, perm_columns, self._data, stream)
        self._Y = None

    def _bind(self, X, k, const, contrasts, perm_columns, data, stream):
        lib = _lib.lib()
        if contrasts is None:
            contrasts = [j for j in range(k) if not const[j]]
        self.contrasts = [int(c) for c in contrasts]
        if perm_columns is None:
            perm_columns = [j for j in range(k) if not const[j]]
        mask = 0
        for j in perm_columns:
            mask |= 1 << int(j)
        con = np.ascontiguousarray(self.contrasts, dtype=np.int32)
        model = ctypes.c_void_p()
        _lib.check(lib.tfce_ols_model_create(data, X.ctypes.data, k, mask, con.ctypes.data,
                                             len(self.contrasts), ctypes.byref(model),
                                             _lib.stream_ptr(stream)))
        self._model = model

    def __del__(self):
        if _lib._lib is None:
            return
        # the model first: a model bound to another's data keeps that owner alive
        m = getattr(self, "_model", None)
        if m is not None and m.value:
            _lib._lib.tfce_ols_model_destroy(m)
        d = getattr(self, "_data", None)
        if d is not None and d.value:
            _lib._lib.tfce_ols_data_destroy(d)

    def t(self, perms=None, out=None, stream=None):
        """t-values [n_contrasts, P, V] for permutation indices perms [P, n] (int32, cuda)."""
        import torch
        if perms is None:
            P = 1
        else:
            assert perms.is_cuda and perms.dtype == torch.int32 and perms.is_contiguous()
            assert perms.dim() == 2 and perms.shape[1] == self.n
            P = int(perms.shape[0])
        if out is None:
            out = torch.empty((len(self.contrasts), P, self.V), dtype=torch.float32, device="cuda")
        assert out.is_contiguous() and out.shape == (len(self.contrasts), P, self.V)
        _lib.check(_lib.lib().tfce_ols_perm_t(self._model, _lib.ptr(perms), P, _lib.ptr(out), self.V,
                                             _lib.stream_ptr(stream)))
        return out

    def fstat(self, perms, tss, out=None, stream=None):
        """sqrt(max(F, 0)) [P, V] float32 of the model of X[perm] for every
        permutation of perms [P, n]; tss [V] float64 on the device."""
        import torch
        assert perms.is_cuda and perms.dtype == torch.int32 and perms.is_contiguous()
        assert perms.dim() == 2 and perms.shape[1] == self.n
        assert tss.is_cuda and tss.dtype == torch.float64 and tss.numel() == self.V and tss.is_contiguous()
        P = int(perms.shape[0])
        if out is None:
            out = torch.empty((P, self.V), dtype=torch.float32, device="cuda")
        assert out.is_contiguous() and out.shape == (P, self.V)
        _lib.check(_lib.lib().tfce_ols_perm_f(self._model, _lib.ptr(perms), P, _lib.ptr(tss), _lib.ptr(out),
                                             self.V, _lib.stream_ptr(stream)))
        return out

    def partial_f(self, perms, blocks, out=None, stream=None):
        """F [len(blocks), P, V] float32 of each column block (start, width) of
        X[perm] against the fit without it, for every permutation of perms
        [P, n] (glm_typeI's per-variable F, pyfunc.py:2340-2352)."""
        import numpy as np
        import torch
        assert perms.is_cuda and perms.dtype == torch.int32 and perms.is_contiguous()
        assert perms.dim() == 2 and perms.shape[1] == self.n
        P = int(perms.shape[0])
        starts = np.ascontiguousarray([int(b[0]) for b in blocks], dtype=np.int32)
        widths = np.ascontiguousarray([int(b[1]) for b in blocks], dtype=np.int32)
        if out is None:
            out = torch.empty((len(blocks), P, self.V), dtype=torch.float32, device="cuda")
        assert out.is_contiguous() and out.shape == (len(blocks), P, self.V)
        _lib.check(_lib.lib().tfce_ols_perm_partial_f(self._model, _lib.ptr(perms), P, starts.ctypes.data,
                                                     widths.ctypes.data, len(blocks), _lib.ptr(out), self.V,
                                                     _lib.stream_ptr(stream)))
        return out


def total_ss(Y):
    """sum_i (y_iv - mean(Y))^2 per element [V] (float64, device): the TSS of
    calcF (cynumstats.pyx:34), about the grand mean of every element of Y,
    computed in FP64 by ``tfce_total_ss``."""
    import torch
    Yd = Y if isinstance(Y, torch.Tensor) else torch.as_tensor(np.ascontiguousarray(Y, dtype=np.float32))
    Yd = Yd.to(device="cuda", dtype=torch.float32)
    if Yd.dim() == 1:
        Yd = Yd[:, None]
    if Yd.stride(1) != 1:
        Yd = Yd.contiguous()
    n, V = int(Yd.shape[0]), int(Yd.shape[1])
    out = torch.empty(V, dtype=torch.float64, device=Yd.device)
    _lib.check(_lib.lib().tfce_total_ss(_lib.ptr(Yd), n, V, Yd.stride(0), _lib.ptr(out), _lib.stream_ptr()))
    return out
