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


def default_batch(items_per_perm, device=None):
    """Permutations per batch for a TFCE launch with ``items_per_perm``
    (surface, sign) items per permutation: the smallest multiple of 64, at
    least 256 and at most 1024, that gives the launch four items per compute
    unit, the number of sweep workgroups a compute unit holds at once.  With
    fewer items (one surface with +/-t, or two with +z only, at 256
    permutations) half of every compute unit idles: measured on MI355X, 512
    permutations instead of 256 give 8% more permutations/s on a 232k-voxel
    volume (one surface, +/-t) and 25% more for mediation on a 327,684-vertex
    mesh (two surfaces, +z)."""
    import torch
    dev = device if device is not None else torch.cuda.current_device()
    return batch_for(torch.cuda.get_device_properties(dev).multi_processor_count, items_per_perm)


def batch_for(compute_units, items_per_perm):
    """``default_batch`` for a device of ``compute_units`` compute units."""
    want = -(-4 * int(compute_units) // max(1, int(items_per_perm)))
    return int(min(1024, max(256, -(-want // 64) * 64)))


class SurfaceSet:
    """Several adjacencies laid out along one map row (tfce_surfset_t).

    ``batch(maps)`` runs every (surface, map, sign) item in one persistent
    launch and returns the rescaled maxima as ``[n_surf, P, n_signs]`` in the
    caller's surface order (optionally also the raw TFCE maxima, the maxima of
    the maps themselves and the per-item status).
    """

    def __init__(self, adjacencies, offsets):
        lib = _lib.lib()
        self.adjs = list(adjacencies)
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        n = len(self.adjs)
        arr = (ctypes.c_void_p * n)(*[a.handle.value for a in self.adjs])
        h = ctypes.c_void_p()
        _lib.check(lib.tfce_surfset_create(arr, n, self.offsets.ctypes.data, ctypes.byref(h)))
        self.handle = h
        order = np.zeros(n, dtype=np.int32)
        _lib.check(lib.tfce_surfset_order(self.handle, order.ctypes.data))
        self.launch_order = order  # launch position -> caller's index
        self._inv = np.argsort(order)
        self._inv_dev = {}  # device -> index tensor (made once, outside any graph capture)
        self.n = n
        self.width = int(max(o + a.V for o, a in zip(self.offsets, self.adjs)))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.tfce_surfset_destroy(h)
            self.handle = None

    def batch(self, maps, sign_mask=3, T0=None, vdensity=None, want_raw=False, want_own=False,
              want_status=False, stream=None, shared_thresholds=False, zero_in_image=False):
        import torch
        assert maps.is_cuda and maps.dtype == torch.float32 and maps.stride(-1) == 1
        P = maps.shape[0]
        stride = maps.stride(0)
        assert stride >= self.width
        ns = 2 if sign_mask == 3 else 1
        n_items = self.n * P * ns
        dev = maps.device
        scaled = torch.empty(n_items, dtype=torch.float32, device=dev)
        raw = torch.empty(n_items, dtype=torch.float32, device=dev) if want_raw else None
        own = torch.empty(n_items, dtype=torch.float32, device=dev) if want_own else None
        st = torch.empty(n_items, dtype=torch.int32, device=dev) if want_status else None
        if T0 is not None:
            assert T0.is_cuda and T0.numel() == n_items
        if vdensity is not None:
            assert vdensity.is_cuda and vdensity.dtype == torch.float32 and vdensity.numel() >= self.width
        if shared_thresholds:
            assert T0 is None
            flags = 1 | (2 if zero_in_image else 0)
            _lib.check(_lib.lib().tfce_batch_ex(self.handle, _lib.ptr(maps), stride, P, int(sign_mask), flags,
                                               _lib.ptr(vdensity), _lib.ptr(scaled), _lib.ptr(raw), _lib.ptr(own),
                                               _lib.ptr(st), _lib.stream_ptr(stream)))
        else:
            _lib.check(_lib.lib().tfce_batch(self.handle, _lib.ptr(maps), stride, P, int(sign_mask),
                                            _lib.ptr(T0), _lib.ptr(vdensity), _lib.ptr(scaled), _lib.ptr(raw),
                                            _lib.ptr(own), _lib.ptr(st), _lib.stream_ptr(stream)))
        inv = self._inv_dev.get(str(dev))
        if inv is None:
            inv = self._inv_dev[str(dev)] = torch.as_tensor(self._inv, device=dev, dtype=torch.long)

        def user(x):
            return None if x is None else x.view(self.n, P, ns).index_select(0, inv)
        out = user(scaled)
        if not (want_raw or want_own or want_status):
            return out
        return out, user(raw), user(own), user(st)
