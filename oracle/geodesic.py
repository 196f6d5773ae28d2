"""ctypes front end of the reference's exact geodesic library.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

``ref_pairs(v, f, max_distance)`` runs adjacency.pyx:44-106's loops over the
reference's own lib/geodesic (``GeodesicAlgorithmExact``: propagate from
vertex i up to max_distance + 1e-10, then ``best_source`` for every j > i),
compiled by ``oracle/Makefile`` into ``oracle/_ref/libgeoref.so``; None when
that library was not built (no reference tree on this machine).
"""
import contextlib
import ctypes
import os

import numpy as np

_SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libgeoref.so")
_lib = None


def have_reference():
    return os.path.exists(_SO)


@contextlib.contextmanager
def _quiet():
    # the library prints a mesh summary on stdout for every mesh it loads
    fd = os.dup(1)
    null = os.open(os.devnull, os.O_WRONLY)
    try:
        os.dup2(null, 1)
        yield
    finally:
        os.dup2(fd, 1)
        os.close(fd)
        os.close(null)


def ref_pairs(vertices, faces, max_distance):
    global _lib
    if not have_reference():
        return None
    if _lib is None:
        _lib = ctypes.CDLL(_SO)
        _lib.ref_geodesic_pairs.restype = ctypes.c_int64
        _lib.ref_geodesic_pairs.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                            ctypes.c_double, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p]
    v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float32).reshape(-1, 3))
    f = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).reshape(-1, 3))
    with _quiet():
        n = _lib.ref_geodesic_pairs(v.ctypes.data, len(v), f.ctypes.data, len(f), float(max_distance), 0,
                                    None, None, None)
        i = np.empty(n, np.int32)
        j = np.empty(n, np.int32)
        d = np.empty(n, np.float64)
        _lib.ref_geodesic_pairs(v.ctypes.data, len(v), f.ctypes.data, len(f), float(max_distance), n,
                                i.ctypes.data, j.ctypes.data, d.ctypes.data)
    return i, j, d
