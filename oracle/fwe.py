"""Loop restatements of the reference's FWER functions.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``find_nearest_left``     calculate_fweP_vertex.py:37-42 (and the voxel twin).
* ``fwep_map``              calculate_fweP_vertex.py:58-69 array part (elements
                            with value > 0 get find_nearest_left, others 0).
* ``find_nearest``          tm_func.py:566-573 / pyfunc.py:254-261.
* ``apply_mfwer``           tm_func.py:403-491 with the per-element Python loops,
                            nulls passed in memory.
* ``perm_max_voxel``        write_perm_maxTFCE_voxel scaling (pyfunc.py:121-126)
                            given a TFCE map.
"""
import math

import numpy as np


def find_nearest_left(array, value, p_array):
    idx = np.searchsorted(array, value, side="left")
    if idx == 0:
        return p_array[idx]
    return p_array[idx - 1]


def p_array_of(n):
    p = np.zeros(n)
    for j in range(n):
        p[j] = np.true_divide(j, n)
    return p


def fwep_map(values, null):
    values = np.asarray(values)
    s = np.sort(np.asarray(null, dtype=np.float64))
    p = p_array_of(len(s))
    out = np.zeros(values.shape[0])
    for cV, k in enumerate(values):
        if k > 0:
            out[cV] = find_nearest_left(s, k, p)
    return out


def find_nearest(array, value, p_array):
    idx = np.searchsorted(array, value, side="left")
    if idx == len(p_array):
        return p_array[idx - 1]
    elif math.fabs(value - array[idx - 1]) < math.fabs(value - array[idx]):
        return p_array[idx - 1]
    else:
        return p_array[idx]


def apply_mfwer(image_array, num_contrasts, surface_range, num_perm, num_surf, nulls, position_array,
                pos_range, neg_range=None, weight=None, mediation=False):
    maxvalue_array = np.zeros((num_perm, num_contrasts))
    temp_max = np.zeros((num_perm, num_surf))
    positive_data = np.zeros((image_array[0].shape[0], num_contrasts))
    if not mediation:
        negative_data = np.zeros((image_array[0].shape[0], num_contrasts))
    if weight == "logmasksize":
        x = []
        for i in range(len(position_array) - 1):
            x.append((position_array[i + 1] - position_array[i]))
        weights = (np.log(x) / np.log(x).sum()) / np.mean(np.log(x) / np.log(x).sum())
        w_temp_max = np.zeros((num_perm, num_surf))
    with np.errstate(divide="ignore", invalid="ignore"):
        for contrast in range(num_contrasts):
            for surface in surface_range:
                log_perm_results = np.log(np.asarray(nulls(surface, contrast), dtype=np.float64)[:num_perm])
                log_perm_results[np.isinf(log_perm_results)] = 0
                start = position_array[surface]
                end = position_array[surface + 1]
                posvmask = np.log(image_array[0][start:end, pos_range[contrast]]) > 0
                temp_lt = np.log(image_array[0][start:end, pos_range[contrast]][posvmask])
                temp_lt -= log_perm_results.mean()
                temp_lt /= log_perm_results.std()
                temp_lt += 10
                positive_data[start:end, contrast][posvmask] = temp_lt
                if not mediation:
                    posvmask = np.log(image_array[0][start:end, neg_range[contrast]]) > 0
                    temp_lt = np.log(image_array[0][start:end, neg_range[contrast]][posvmask])
                    temp_lt -= log_perm_results.mean()
                    temp_lt /= log_perm_results.std()
                    temp_lt += 10
                    negative_data[start:end, contrast][posvmask] = temp_lt
                log_perm_results -= log_perm_results.mean()
                log_perm_results /= log_perm_results.std()
                if weight == "logmasksize":
                    w_log_perm_results = log_perm_results * weights[surface]
                    w_log_perm_results += 10
                    w_temp_max[:, surface] = w_log_perm_results
                log_perm_results += 10
                temp_max[:, surface] = log_perm_results
            if weight is not None:
                w_temp_max[np.isnan(w_temp_max)] = 0
                max_index = np.argmax(w_temp_max, axis=1)
                max_value_list = np.zeros((len(max_index)), dtype=np.float32)
                for i in range(len(temp_max)):
                    max_value_list[i] = temp_max[i, max_index[i]]
                maxvalue_array[:, contrast] = np.sort(max_value_list)
            else:
                maxvalue_array[:, contrast] = np.sort(temp_max.max(axis=1))
    for contrast in range(num_contrasts):
        sorted_perm_tfce_max = maxvalue_array[:, contrast]
        p_array = np.zeros_like(sorted_perm_tfce_max)
        for j in range(num_perm):
            p_array[j] = np.true_divide(j, num_perm)
        corrp = np.zeros((positive_data.shape[0]))
        for cV, k in enumerate(positive_data[:, contrast]):
            corrp[cV] = find_nearest(sorted_perm_tfce_max, k, p_array)
        positive_data[:, contrast] = np.copy(corrp)
        if not mediation:
            corrp = np.zeros((negative_data.shape[0]))
            for cV, k in enumerate(negative_data[:, contrast]):
                corrp[cV] = find_nearest(sorted_perm_tfce_max, k, p_array)
            negative_data[:, contrast] = np.copy(corrp)
    if mediation:
        return positive_data
    return positive_data, negative_data


def perm_max_voxel(tfce, t):
    return tfce.max() * (t.max() / np.float32(100))
