"""ctypes binding of libtfcehip.so (C ABI declared in include/tfcehip.h).

The library is the only compute path: if it is missing, or no GPU is visible,
every entry point raises.  There is no CPU fallback.
"""
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libtfcehip.so")

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u32 = ctypes.c_uint32
c_f32 = ctypes.c_float
c_p = ctypes.c_void_p
c_str = ctypes.c_char_p

# name -> (restype, argtypes); every function declared in include/tfcehip.h
SIGNATURES = {
    "tfce_last_error": (c_str, []),
    "tfce_version": (c_str, []),
    "tfce_adj_create": (c_int, [c_p, c_p, c_i32, c_i64, c_f32, c_f32, c_p]),
    "tfce_adj_create_ex": (c_int, [c_p, c_p, c_i32, c_i64, c_f32, c_f32, c_u32, c_p]),
    "tfce_adj_destroy": (c_int, [c_p]),
    "tfce_locality_order": (c_int, [c_p, c_p, c_i32, c_p]),
    "tfce_adj_info": (c_int, [c_p, c_p, c_p, c_p]),
    "tfce_run": (c_int, [c_p, c_p, c_p, c_f32, c_int, c_p]),
    "tfce_run_ex": (c_int, [c_p, c_p, c_p, c_f32, c_f32, c_int, c_i32, c_p, c_p, c_p]),
    "tfce_batch_max": (c_int, [c_p, c_p, c_int, c_int, c_p, c_p, c_p]),
    "tfce_batch_max_noalloc": (c_int, [c_p, c_p, c_int, c_int, c_p, c_p, c_p]),
    "tfce_surfset_create": (c_int, [c_p, c_i32, c_p, c_p]),
    "tfce_surfset_destroy": (c_int, [c_p]),
    "tfce_surfset_order": (c_int, [c_p, c_p]),
    "tfce_batch": (c_int, [c_p, c_p, c_i64, c_int, c_int, c_p, c_p, c_p, c_p, c_p, c_p, c_p]),
    "tfce_set_occupancy": (c_int, [c_i32]),
    "tfce_set_single_image_mode": (c_int, [c_i32, c_i32]),
    "tfce_adj_voxel_csr": (c_int, [c_p, c_i32, c_i32, c_i32, c_i32, c_p, c_p, c_p, c_p, c_p]),
    "tfce_adj_mesh_csr": (c_int, [c_p, c_i64, c_i32, c_p, c_p, c_p, c_p]),
    "tfce_geodesic_pairs": (c_int, [c_p, c_i32, c_p, c_i32, ctypes.c_double, c_i32, c_i32, c_i32, c_p]),
    "tfce_geodesic_count": (c_int, [c_p, c_p]),
    "tfce_geodesic_copy": (c_int, [c_p, c_p, c_p, c_p]),
    "tfce_geodesic_destroy": (c_int, [c_p]),
    "tfce_set_lds_sweep": (c_int, [c_i32]),
    "tfce_set_pair_sweep": (c_int, [c_i32]),
    "tfce_batch_ex": (c_int, [c_p, c_p, c_i64, c_int, c_int, c_u32, c_p, c_p, c_p, c_p, c_p, c_p]),
    "tfce_ols_data_create": (c_int, [c_p, c_i32, c_i64, c_i64, c_i32, c_i32, c_p, c_p]),
    "tfce_ols_data_destroy": (c_int, [c_p]),
    "tfce_ols_model_create": (c_int, [c_p, c_p, c_i32, c_u32, c_p, c_i32, c_p, c_p]),
    "tfce_ols_model_destroy": (c_int, [c_p]),
    "tfce_ols_info": (c_int, [c_p, c_p, c_p, c_p]),
    "tfce_ols_perm_t": (c_int, [c_p, c_p, c_i32, c_p, c_i64, c_p]),
    "tfce_ols_perm_f": (c_int, [c_p, c_p, c_i32, c_p, c_p, c_i64, c_p]),
    "tfce_ols_perm_sobel": (c_int, [c_p, c_p, c_p, c_i32, c_i32, c_p, c_i64, c_p]),
    "tfce_ols_perm_partial_f": (c_int, [c_p, c_p, c_i32, c_p, c_p, c_i32, c_p, c_i64, c_p]),
    "tfce_ols_set_tile": (c_int, [c_i32]),
    "tfce_ols_tile": (c_int, []),
    "tfce_fwe_p": (c_int, [c_p, c_i32, c_p, c_i64, c_i32, c_p, c_p]),
    "tfce_fwe_p64": (c_int, [c_p, c_i32, c_p, c_i64, c_i32, c_p, c_p]),
    "tfce_ols_direct": (c_int, [c_p, c_i32, c_i64, c_i64, c_p, c_i32, c_p, c_p, c_p, c_p, c_p, c_p, c_p,
                                c_p, c_p]),
    "tfce_se_of_slope": (c_int, [c_p, c_p, c_i32, c_i64, c_p, c_p]),
    "tfce_image_regression": (c_int, [c_p, c_i64, c_p, c_i64, c_p, c_i32, c_i32, c_i64, ctypes.c_double, c_p, c_i64,
                                      c_p]),
    "tfce_image_vif": (c_int, [c_p, c_i32, c_i64, c_p, c_i32, c_i32, c_i64, c_p, c_p]),
    "tfce_total_ss": (c_int, [c_p, c_i32, c_i64, c_i64, c_p, c_p]),
    "tfce_mlp_param_count": (c_int, []),
    "tfce_mlp_forward": (c_int, [c_p, c_p, c_i32, c_p, c_p]),
    "tfce_mlp_train": (c_int, [c_p, c_p, c_p, c_p, c_p, c_i32, c_p, c_i32, c_i32, c_i32, c_f32, c_f32, c_p, c_p, c_p]),
    "ols_perm_t": (c_int, [c_p, c_p, c_p, c_i32, c_i32, c_i64, c_i32, c_i32, c_i32, c_p, c_p]),
    "sobel_perm_z": (c_int, [c_p, c_i32, c_i64, c_p, c_p, c_i32, c_i32, c_p, c_i32, c_i32, c_p, c_p]),
    "tfce_sobel_z": (c_int, [c_p, c_p, c_i64, c_i32, c_f32, c_p, c_p]),
}

PREC = {"fp32": 0, "bf16x3": 1, "bf16x6": 2, "fp16x3": 3}

_lock = threading.Lock()
_lib = None


class TfceHipError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Load the shared library without touching the GPU (symbol checks)."""
    if not os.path.exists(path):
        raise ImportError(f"libtfcehip.so not built at {path}: run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    """The library, after checking that a GPU is present (raises otherwise)."""
    global _lib
    with _lock:
        if _lib is None:
            import torch
            if not torch.cuda.is_available():
                raise TfceHipError("tfce_mediation_amd needs a ROCm GPU; none is visible")
            _lib = load_library()
    return _lib


_host = None


def host_lib():
    """The library for its host-only entry points (the geodesic adjacency),
    which need no GPU."""
    global _host
    with _lock:
        if _host is None:
            _host = _lib if _lib is not None else load_library()
    return _host


def check(rc):
    if rc != 0:
        src = _lib if _lib is not None else _host
        msg = src.tfce_last_error().decode() if src is not None else ""
        raise TfceHipError(f"libtfcehip error {rc}: {msg}")
    return rc


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())
