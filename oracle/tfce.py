"""ctypes front end of the C TFCE oracle and of the reference build.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``tfce(...)`` follows ``lib/fast_tfce.hpp:11-95`` as bound by
``tfce.pyx:24-45`` (``CreateAdjSet(H, E, adj).run(image, enhn)`` passes
minT = 0 and deltaT = 0).  ``ref_tfce(...)`` runs the reference's own header,
compiled by ``oracle/Makefile`` into ``oracle/_ref/libreftfce.so``; it is None
when that library was not built (e.g. on a machine without the reference tree).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE_SO = os.path.join(_HERE, "_build", "liboracle.so")
_REF_SO = os.path.join(_HERE, "_ref", "libreftfce.so")

DEGENERATE = -2

_lib = None
_ref = None


def _load():
    global _lib, _ref
    if _lib is None:
        if not os.path.exists(_ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        lib = ctypes.CDLL(_ORACLE_SO)
        lib.oracle_tfce.restype = ctypes.c_int64
        lib.oracle_tfce.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                    ctypes.c_int]
        lib.oracle_tfce_levels.restype = ctypes.c_int64
        lib.oracle_tfce_levels.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_void_p, ctypes.c_int64]
        _lib = lib
        if os.path.exists(_REF_SO):
            ref = ctypes.CDLL(_REF_SO)
            ref.ref_tfce.restype = ctypes.c_int
            ref.ref_tfce.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]
            _ref = ref
    return _lib


def have_reference():
    _load()
    return _ref is not None


def _csr(row_ptr, col):
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    return row_ptr, col


def levels(max_t, min_t=0.0, delta_t=0.0):
    """Thresholds the reference visits (``fast_tfce.hpp:34-39``) as float32."""
    lib = _load()
    n = lib.oracle_tfce_levels(np.float32(max_t), np.float32(min_t), np.float32(delta_t), None, 0)
    if n < 0:
        return None
    out = np.zeros(max(n, 1), dtype=np.float32)
    lib.oracle_tfce_levels(np.float32(max_t), np.float32(min_t), np.float32(delta_t),
                           out.ctypes.data, n)
    return out[:n]


def tfce(row_ptr, col, image, H, E, enhn=None, min_t=0.0, delta_t=0.0, labels_level=-1,
         exact_pow=False):
    """enhn += TFCE(image); returns (enhn, n_levels_or_status, labels or None).

    exact_pow=False reproduces the reference bit for bit (libm powf for T^H);
    exact_pow=True uses the correctly rounded T^H, the GPU's definition.
    """
    lib = _load()
    row_ptr, col = _csr(row_ptr, col)
    image = np.ascontiguousarray(image, dtype=np.float32)
    V = image.shape[0]
    if enhn is None:
        enhn = np.zeros(V, dtype=np.float32)
    assert enhn.dtype == np.float32 and enhn.flags.c_contiguous and enhn.shape == (V,)
    labels = None
    if labels_level >= 0:
        labels = np.full(V, -1, dtype=np.int32)
    st = lib.oracle_tfce(np.float32(H), np.float32(E), np.float32(min_t), np.float32(delta_t),
                         V, row_ptr.ctypes.data, col.ctypes.data, image.ctypes.data,
                         enhn.ctypes.data, labels_level,
                         labels.ctypes.data if labels is not None else None, 1 if exact_pow else 0)
    return enhn, st, labels


def ref_tfce(row_ptr, col, image, H, E, enhn=None, min_t=0.0, delta_t=0.0):
    """The reference's own fast_tfce.hpp; None when oracle/_ref was not built."""
    _load()
    if _ref is None:
        return None
    row_ptr, col = _csr(row_ptr, col)
    image = np.ascontiguousarray(image, dtype=np.float32)
    V = image.shape[0]
    if enhn is None:
        enhn = np.zeros(V, dtype=np.float32)
    st = _ref.ref_tfce(np.float32(H), np.float32(E), np.float32(min_t), np.float32(delta_t),
                       V, row_ptr.ctypes.data, col.ctypes.data, image.ctypes.data, enhn.ctypes.data)
    if st != 0:
        raise ValueError("reference TFCE refuses a degenerate image (it would not terminate)")
    return enhn
