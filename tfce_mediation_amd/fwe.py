"""Family-wise error outputs from max-TFCE nulls.

* ``fwe_p(values, null, mode)``      1 - p(FWE) per element on the GPU (FP64
  lookups in the sorted null, ``tfce_fwe_p``):
    mode "left"    = calculate_fweP_vertex.py:37-69 / calculate_fweP_voxel.py:23-57
    mode "nearest" = find_nearest of tm_func.py:566-573 (apply_mfwer).
* ``calculate_fwep(tfce_values, null)`` the array part of the vertex / voxel
  scripts (elements <= 0 stay 0) plus the accuracy they print.
* ``critical_value(null, alpha)``    sorted(null, reverse=True)[int(N*alpha)]
  (tools/plot_permutations.py:71-81).
* ``apply_mfwer(...)``               tm_func.py:403-491 on in-memory nulls.
* ``lowest_length`` / null CSV I/O    tm_func.py:544-552 and the writers'
  formats: vertex "%.4f" (pyfunc.py:119), voxel "%1.4f" (:126), mmr "%f"
  (tm_func.py:94-98).
"""
import os

import numpy as np

from . import _lib


def fwe_p(values, null, mode="left"):
    """1 - p(FWE) of every element of ``values`` against ``null`` (float64 numpy)."""
    import torch
    lib = _lib.lib()
    vals = values if isinstance(values, torch.Tensor) else torch.as_tensor(
        np.ascontiguousarray(values, dtype=np.float32))
    vals = vals.to(device="cuda", dtype=torch.float32).contiguous().view(-1)
    srt = torch.as_tensor(np.sort(np.asarray(null, dtype=np.float64))).cuda()
    out = torch.empty(vals.numel(), dtype=torch.float64, device="cuda")
    _lib.check(lib.tfce_fwe_p(_lib.ptr(srt), int(srt.numel()), _lib.ptr(vals), int(vals.numel()),
                              {"left": 0, "nearest": 1}[mode], _lib.ptr(out), _lib.stream_ptr()))
    return out.cpu().numpy()


def accuracy(num_perm):
    """Half-width the scripts print: p = 0.05 +/- 2 sqrt(0.05*0.95/N)."""
    return 2 * (np.sqrt(0.05 * 0.95 / num_perm))


def calculate_fwep(tfce_values, perm_tfce_max):
    """1-p(FWE) map for the elements of a TFCE image (values <= 0 give 0)."""
    return fwe_p(tfce_values, perm_tfce_max, "left")


def critical_value(null, alpha=0.05):
    s = sorted(np.asarray(null, dtype=np.float64), reverse=True)
    return s[int(len(s) * alpha)]


def lowest_length(nulls):
    return int(min(len(np.atleast_1d(v)) for v in nulls))


def write_null(path, values, kind="vertex"):
    """Append null values in the reference's text format (one per line)."""
    fmt = {"vertex": "%.4f", "voxel": "%1.4f", "mmr": "%f"}[kind]
    with open(path, "a") as f:
        for v in np.asarray(values, dtype=np.float64).ravel():
            f.write((fmt % v) + "\n")


def read_null(path):
    return np.atleast_1d(np.genfromtxt(path, delimiter=","))


def apply_mfwer(image_array, num_contrasts, surface_range, num_perm, num_surf, nulls, position_array,
                pos_range, neg_range=None, weight=None, mediation=False):
    """Joint FWER over modalities (tm_func.py:403-491).

    ``nulls(surface, contrast)`` returns that surface's max-TFCE null (the
    reference reads output_<tmi>/perm_maxTFCE_surf<s>_tcon<c>.csv).  The
    standardisation stays in numpy FP64 exactly as the reference computes it;
    the per-element nearest lookups run on the GPU.
    """
    maxvalue_array = np.zeros((num_perm, num_contrasts))
    temp_max = np.zeros((num_perm, num_surf))
    positive_data = np.zeros((image_array[0].shape[0], num_contrasts))
    negative_data = None if mediation else np.zeros((image_array[0].shape[0], num_contrasts))
    if weight == "logmasksize":
        x = [position_array[i + 1] - position_array[i] for i in range(len(position_array) - 1)]
        weights = (np.log(x) / np.log(x).sum()) / np.mean(np.log(x) / np.log(x).sum())
        w_temp_max = np.zeros((num_perm, num_surf))
    with np.errstate(divide="ignore", invalid="ignore"):
        for contrast in range(num_contrasts):
            for surface in surface_range:
                log_perm = np.log(np.asarray(nulls(surface, contrast), dtype=np.float64)[:num_perm])
                log_perm[np.isinf(log_perm)] = 0
                start, end = position_array[surface], position_array[surface + 1]
                signs = [(positive_data, pos_range)] + ([] if mediation else [(negative_data, neg_range)])
                for data, rng in signs:
                    col = image_array[0][start:end, rng[contrast]]
                    m = np.log(col) > 0
                    lt = np.log(col[m])
                    lt -= log_perm.mean()
                    lt /= log_perm.std()
                    lt += 10
                    data[start:end, contrast][m] = lt
                log_perm -= log_perm.mean()
                log_perm /= log_perm.std()
                if weight == "logmasksize":
                    w = log_perm * weights[surface]
                    w += 10
                    w_temp_max[:, surface] = w
                log_perm += 10
                temp_max[:, surface] = log_perm
            if weight is not None:
                w_temp_max[np.isnan(w_temp_max)] = 0
                idx = np.argmax(w_temp_max, axis=1)
                mv = temp_max[np.arange(len(temp_max)), idx].astype(np.float32)
                maxvalue_array[:, contrast] = np.sort(mv)
            else:
                maxvalue_array[:, contrast] = np.sort(temp_max.max(axis=1))
    for contrast in range(num_contrasts):
        null = maxvalue_array[:, contrast]
        positive_data[:, contrast] = _nearest(positive_data[:, contrast], null)
        if not mediation:
            negative_data[:, contrast] = _nearest(negative_data[:, contrast], null)
    return positive_data if mediation else (positive_data, negative_data)


def _nearest(values, sorted_null):
    return fwe_p64(values, sorted_null)


def fwe_p64(values, sorted_null):
    """Nearest lookup of float64 element values (tfce_fwe_p64)."""
    import torch
    lib = _lib.lib()
    v = values if isinstance(values, torch.Tensor) else torch.as_tensor(np.asarray(values, np.float64)).cuda()
    srt = torch.as_tensor(np.ascontiguousarray(sorted_null, dtype=np.float64)).cuda()
    out = torch.empty(v.numel(), dtype=torch.float64, device="cuda")
    _lib.check(lib.tfce_fwe_p64(_lib.ptr(srt), int(srt.numel()), _lib.ptr(v.contiguous()), int(v.numel()), 1,
                                _lib.ptr(out), _lib.stream_ptr()))
    return out.cpu().numpy()
