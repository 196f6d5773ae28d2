"""Drop-in for ``tfce_mediation.tfce`` (the Cython operator of ``tfce.pyx:24-45``).

``CreateAdjSet(H, E, adjacency).run(image, enhn)`` keeps the reference's
contract: ``image`` and ``enhn`` are 1-D float32 C-contiguous arrays of the
adjacency's length, ``enhn`` is accumulated in place (callers zero it), H and E
are stored as float32 (``tfce.pyx:27-28``) and the step is max/100
(``tfce.pyx:45`` passes minT = deltaT = 0).  The sweep runs on the GPU through
``tfce_run`` of libtfcehip; numpy inputs are staged to the device and back,
torch CUDA tensors are used in place.

Defined where the reference is not: an image whose maximum is exactly 0 (the
reference's loop never ends, ``fast_tfce.hpp:34-39``) leaves ``enhn`` unchanged;
NaN elements never join a cluster.
"""
import ctypes

import numpy as np

from . import _lib
from .adjacency import to_csr

TFCE_ADJ_IMPLICIT_ZERO = 2
TFCE_ADJ_ZERO_IN_SCALE = 4


class Adjacency:
    """Device-resident CSR adjacency with its TFCE settings (a tfce_adj_t)."""

    def __init__(self, adjacency, H, E, flags=0):
        lib = _lib.lib()
        row_ptr, col = to_csr(adjacency)
        self.V = int(row_ptr.shape[0] - 1)
        self.nnz = int(col.shape[0])
        self.H = float(np.float32(H))
        self.E = float(np.float32(E))
        self.flags = int(flags)
        handle = ctypes.c_void_p()
        _lib.check(lib.tfce_adj_create_ex(row_ptr.ctypes.data, col.ctypes.data if self.nnz else None,
                                          self.V, self.nnz, self.H, self.E, self.flags,
                                          ctypes.byref(handle)))
        self.handle = handle
        sym = ctypes.c_int32()
        _lib.check(lib.tfce_adj_info(self.handle, None, None, ctypes.byref(sym)))
        self.symmetric = bool(sym.value)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.tfce_adj_destroy(h)
            self.handle = None

    # -- single image, full map (the CreateAdjSet.run contract) --------------
    def run_device(self, image, enhn, dh=0.0, accumulate=True, min_t=0.0,
                   labels_level=-1, labels=None, status=None, stream=None):
        import torch
        assert image.is_cuda and enhn.is_cuda
        assert image.dtype == torch.float32 and enhn.dtype == torch.float32
        assert image.numel() == self.V and enhn.numel() == self.V
        assert image.is_contiguous() and enhn.is_contiguous()
        lib = _lib.lib()
        _lib.check(lib.tfce_run_ex(self.handle, _lib.ptr(image), _lib.ptr(enhn), float(min_t),
                                   float(dh), 1 if accumulate else 0, int(labels_level),
                                   _lib.ptr(labels), _lib.ptr(status), _lib.stream_ptr(stream)))

    # -- batched maxima ------------------------------------------------------
    def batch_max(self, maps, sign_mask=3, vdensity=None, out=None, stream=None):
        """[P, 2] rescaled maxima of +map / -map (pyfunc.py:107-126 semantics)."""
        import torch
        assert maps.is_cuda and maps.dtype == torch.float32 and maps.is_contiguous()
        P = maps.shape[0]
        assert maps.numel() == P * self.V
        if out is None:
            out = torch.empty((P, 2), dtype=torch.float32, device=maps.device)
        if vdensity is not None:
            assert vdensity.is_cuda and vdensity.dtype == torch.float32 and vdensity.numel() == self.V
        _lib.check(_lib.lib().tfce_batch_max(self.handle, _lib.ptr(maps), P, int(sign_mask),
                                            _lib.ptr(vdensity), _lib.ptr(out), _lib.stream_ptr(stream)))
        return out


class CreateAdjSet:
    """``CreateAdjSet(H, E, pyAdjacency)``; ``run(image, enhn)`` accumulates TFCE."""

    def __init__(self, H, E, pyAdjacency):
        self._adj = Adjacency(pyAdjacency, H, E)
        self.H = self._adj.H
        self.E = self._adj.E

    @property
    def adjacency(self):
        return self._adj

    def run(self, image, enhn):
        import torch
        if isinstance(image, torch.Tensor):
            self._adj.run_device(image, enhn, accumulate=True)
            return
        if not (isinstance(image, np.ndarray) and image.dtype == np.float32 and image.ndim == 1
                and image.flags.c_contiguous):
            raise ValueError("Buffer dtype mismatch: image must be a 1-D C-contiguous float32 array")
        if not (isinstance(enhn, np.ndarray) and enhn.dtype == np.float32 and enhn.ndim == 1
                and enhn.flags.c_contiguous):
            raise ValueError("Buffer dtype mismatch: enhn must be a 1-D C-contiguous float32 array")
        if image.shape[0] != self._adj.V or enhn.shape[0] != self._adj.V:
            raise IndexError("image/enhn length differs from the adjacency")
        dev = torch.device("cuda", torch.cuda.current_device())
        d_img = torch.from_numpy(image).to(dev)
        d_enh = torch.from_numpy(enhn).to(dev)
        self._adj.run_device(d_img, d_enh, accumulate=True)
        enhn[:] = d_enh.cpu().numpy()
