"""Drop-in for the permutation-path subset of ``tfce_mediation.tm_func``
(tm_func.py:54-573): the mmr / mmr-lr per-permutation functions, the joint
FWER and the mask / adjacency helpers.

The statistics and TFCE run on the GPU (``cynumstats`` FP64 kernels and the
``CreateAdjSet`` operator); null values are appended to the reference's file
names in its ``%f`` format.  For many permutations at once use
``engine.PermutationEngine`` / ``randomise`` instead of calling these in a
loop.  Time-seeded permutations (``perm_number + int(float(str(time())[-6:])*100)``,
tm_func.py:57) are kept for drop-in behaviour; pass ``perm_seed`` to the
low-RAM variants for reproducible draws (tm_func.py:147-148).
"""
import os
from time import time

import numpy as np

from . import fwe as _fwe
from .adjacency import merge_adjacency_array  # noqa: F401
from .cynumstats import tval_int
from .pyfunc import calc_sobelz, find_nearest  # noqa: F401


def _seed(perm_number, perm_seed=None):
    if perm_seed is not None:
        np.random.seed(perm_number + perm_seed)
    else:
        np.random.seed(perm_number + int(float(str(time())[-6:]) * 100))


def _scaled(tfce, t, vdensity):
    return tfce * (t.max() / np.float32(100)) * vdensity


def _run_pm(calcTFCE, image):
    pos = np.zeros_like(image)
    neg = np.zeros_like(image)
    calcTFCE.run(image, pos)
    calcTFCE.run(np.ascontiguousarray(image * -1), neg)
    return pos, neg


def _echo(path, value):
    with open(path, "a") as f:
        f.write("%f\n" % value)


def _tfce_maps(merge_y, masking_array, X, calcTFCE, vdensity, position_array, fullmask, no_intercept):
    """t-values and rescaled +/- TFCE of every contrast (no I/O)."""
    n = merge_y.shape[0]
    k = len(X.T)
    invXX = np.linalg.inv(np.dot(X.T, X))
    tvals = tval_int(X, invXX, merge_y, n, k, merge_y.shape[1])
    if no_intercept:
        tvals = tvals[1:, :]
    tvals = tvals.astype(np.float32, order="C")
    tfce_tvals = np.zeros_like(tvals)
    neg_tfce_tvals = np.zeros_like(tvals)
    for c in range(tvals.shape[0]):
        img = np.zeros_like(fullmask).astype(np.float32, order="C")
        img[fullmask == 1] = tvals[0] if tvals.shape[0] == 1 else tvals[c]
        pos, neg = _run_pm(calcTFCE, img)
        t_in = img[fullmask == 1]
        pos, neg = pos[fullmask == 1], neg[fullmask == 1]
        for s in range(len(masking_array)):
            a, b = position_array[s], position_array[s + 1]
            vd = vdensity if isinstance(vdensity, int) else vdensity[a:b]
            tfce_tvals[c, a:b] = _scaled(pos[a:b], t_in[a:b], vd)
            neg_tfce_tvals[c, a:b] = _scaled(neg[a:b], (t_in * -1)[a:b], vd)
    return tvals, tfce_tvals, neg_tfce_tvals


def calculate_tfce(merge_y, masking_array, pred_x, calcTFCE, vdensity, position_array, fullmask,
                   perm_number=None, randomise=False, verbose=False, no_intercept=True, set_surf_count=None,
                   print_interation=False):
    n = merge_y.shape[0]
    X = np.column_stack([np.ones(n), pred_x])
    if randomise:
        _seed(perm_number)
        X = X[np.random.permutation(list(range(n)))]
    tvals, tfce_tvals, neg_tfce_tvals = _tfce_maps(merge_y, masking_array, X, calcTFCE, vdensity,
                                                   position_array, fullmask, no_intercept)
    for c in range(tvals.shape[0]):
        for s in range(len(masking_array)):
            a, b = position_array[s], position_array[s + 1]
            sid = int(set_surf_count[s]) if set_surf_count is not None else s
            if randomise:
                path = "perm_maxTFCE_surf%d_tcon%d.csv" % (sid, c + 1)
                _echo(path, np.nanmax(tfce_tvals[c, a:b]))
                _echo(path, np.nanmax(neg_tfce_tvals[c, a:b]))
            else:
                print("Maximum (untransformed) postive tfce value for surface %s, tcon %d: %f"
                      % (sid, c + 1, np.nanmax(tfce_tvals[c, a:b])))
                print("Maximum (untransformed) negative tfce value for surface %s, tcon %d: %f"
                      % (sid, c + 1, np.nanmax(neg_tfce_tvals[c, a:b])))
        if verbose:
            print("T-contrast: %d" % c)
            print("Max tfce from all surfaces = %f" % tfce_tvals[c].max())
            print("Max negative tfce from all surfaces = %f" % neg_tfce_tvals[c].max())
    if randomise:
        if print_interation:
            print("Interation number: %d" % perm_number)
        return None
    return (tvals.astype(np.float32, order="C"), tfce_tvals.astype(np.float32, order="C"),
            neg_tfce_tvals.astype(np.float32, order="C"))


def low_ram_calculate_tfce(data, mask, pred_x, calcTFCE, vdensity, set_surf_count=0, perm_number=None,
                           randomise=False, no_intercept=True, output_dir=None, perm_seed=None):
    n = data.shape[0]
    X = np.column_stack([np.ones(n), pred_x])
    if randomise:
        _seed(perm_number, perm_seed)
        X = X[np.random.permutation(list(range(n)))]
    k = len(X.T)
    invXX = np.linalg.inv(np.dot(X.T, X))
    tvals = tval_int(X, invXX, data, n, k, data.shape[1])
    if no_intercept:
        tvals = tvals[1:, :]
    tvals = tvals.astype(np.float32, order="C")
    tfce_tvals = np.zeros_like(tvals)
    neg_tfce_tvals = np.zeros_like(tvals)
    for c in range(tvals.shape[0]):
        img = np.zeros_like(mask).astype(np.float32, order="C")
        img[mask == 1] = tvals[0] if tvals.shape[0] == 1 else tvals[c]
        pos, neg = _run_pm(calcTFCE, img)
        tfce_tvals[c, :] = pos[mask == 1] * (img.max() / np.float32(100)) * vdensity
        neg_tfce_tvals[c, :] = neg[mask == 1] * ((img * -1).max() / np.float32(100)) * vdensity
        if randomise:
            name = "perm_maxTFCE_surf%d_tcon%d.csv" % (int(set_surf_count), c + 1)
            path = os.path.join(output_dir, name) if output_dir is not None else name
            _echo(path, np.nanmax(tfce_tvals[c, :]))
            _echo(path, np.nanmax(neg_tfce_tvals[c, :]))
    if not randomise:
        return (tvals.astype(np.float32, order="C"), tfce_tvals.astype(np.float32, order="C"),
                neg_tfce_tvals.astype(np.float32, order="C"))


def _mediation_inputs(medtype, n, pred_x, depend_y, perm_number, perm_seed, randomise):
    if randomise:
        _seed(perm_number, perm_seed)
        idx = np.random.permutation(list(range(n)))
        pred_x = pred_x[idx]
        if medtype not in ("M", "I"):
            depend_y = depend_y[idx]
    return pred_x, depend_y


def calculate_mediation_tfce(medtype, merge_y, masking_array, pred_x, depend_y, calcTFCE, vdensity,
                             position_array, fullmask, perm_number=None, randomise=False, verbose=False,
                             no_intercept=True, print_interation=False):
    n = merge_y.shape[0]
    pred_x, depend_y = _mediation_inputs(medtype, n, pred_x, depend_y, perm_number, None, randomise)
    z = calc_sobelz(medtype, pred_x, depend_y, merge_y, n, merge_y.shape[1]).astype(np.float32, order="C")
    tfce_z = np.zeros_like(z)
    img = np.zeros_like(fullmask).astype(np.float32, order="C")
    img[fullmask == 1] = z
    enh = np.zeros_like(img)
    calcTFCE.run(img, enh)
    z_in, enh = img[fullmask == 1], enh[fullmask == 1]
    for s in range(len(masking_array)):
        a, b = position_array[s], position_array[s + 1]
        vd = vdensity if isinstance(vdensity, int) else vdensity[a:b]
        tfce_z[a:b] = _scaled(enh[a:b], z_in[a:b], vd)
        if randomise:
            _echo("perm_maxTFCE_surf%d_%s_zstat.csv" % (s, medtype), np.nanmax(tfce_z[a:b]))
        else:
            print("Max Sobel Z tfce value for surface %s:\t %1.5f" % (s, np.nanmax(tfce_z[a:b])))
    if verbose:
        print("Max Zstat tfce from all surfaces = %f" % tfce_z.max())
    if randomise:
        print("Interation number: %d" % perm_number)
        return None
    return z.astype(np.float32, order="C"), tfce_z.astype(np.float32, order="C")


def low_ram_calculate_mediation_tfce(medtype, data, mask, pred_x, depend_y, calcTFCE, vdensity,
                                     set_surf_count=0, perm_number=None, randomise=False, no_intercept=True,
                                     output_dir=None, perm_seed=None):
    n = data.shape[0]
    pred_x, depend_y = _mediation_inputs(medtype, n, pred_x, depend_y, perm_number, perm_seed, randomise)
    z = calc_sobelz(medtype, pred_x, depend_y, data, n, data.shape[1]).astype(np.float32, order="C")
    img = np.zeros_like(mask).astype(np.float32, order="C")
    img[mask == 1] = z
    enh = np.zeros_like(img)
    calcTFCE.run(img, enh)
    zin = img[mask == 1]
    tz = enh[mask == 1] * (zin.max() / np.float32(100)) * vdensity
    if randomise:
        name = "perm_maxTFCE_surf%d_%s_zstat.csv" % (set_surf_count, medtype)
        _echo(os.path.join(output_dir, name) if output_dir is not None else name, np.nanmax(tz))
        return None
    return zin.astype(np.float32, order="C"), tz.astype(np.float32, order="C")


def calc_mixed_tfce(assigntfcesettings, merge_y, masking_array, position_array, vdensity, pred_x, calcTFCE,
                    perm_number=None, randomise=False, medtype=None, depend_y=None):
    """One TFCE pass per TFCE-settings group (tm_func.py:327-378).  Unlike the
    reference, every group of one permutation sees the same permutation: the
    draw is made once here (the reference re-seeds from the clock per group,
    so its groups see different permutations)."""
    assigntfcesettings = np.asarray(assigntfcesettings)
    n = merge_y.shape[0]
    X = np.column_stack([np.ones(n), pred_x])
    if randomise:
        _seed(perm_number)
        X = X[np.random.permutation(list(range(n)))]
    out = None
    for i in np.unique(assigntfcesettings):
        data_mask = np.zeros(merge_y.shape[1], dtype=bool)
        extract_range = np.argwhere(assigntfcesettings == i).ravel()
        for surface in extract_range:
            data_mask[position_array[int(surface)]:position_array[int(surface) + 1]] = True
        try:
            temp_vd = vdensity[data_mask]
        except (TypeError, IndexError):
            temp_vd = 1
        masks = [masking_array[int(j)] for j in extract_range]
        pos_arr = create_position_array(masks)
        tv, tp, tn = _tfce_maps(merge_y[:, data_mask], masks, X, calcTFCE[i], temp_vd, pos_arr,
                                create_full_mask(masks), True)
        if randomise:
            for c in range(tp.shape[0]):
                for s_local, sid in enumerate(extract_range):
                    a, b = pos_arr[s_local], pos_arr[s_local + 1]
                    path = "perm_maxTFCE_surf%d_tcon%d.csv" % (int(sid), c + 1)
                    _echo(path, np.nanmax(tp[c, a:b]))
                    _echo(path, np.nanmax(tn[c, a:b]))
            continue
        if out is None:
            out = [np.zeros((tv.shape[0], merge_y.shape[1])) for _ in range(3)]
        out[0][:, data_mask] = tv
        out[1][:, data_mask] = tp
        out[2][:, data_mask] = tn
    if not randomise:
        return tuple(out)


def apply_mfwer(image_array, num_contrasts, surface_range, num_perm, num_surf, tminame, position_array,
                pos_range, neg_range=None, method="scale", weight=None, mediation=False, medtype=None):
    """tm_func.py:403-491, reading the per-surface nulls from output_<tminame>/."""
    def nulls(surface, contrast):
        if mediation:
            return _fwe.read_null("output_%s/perm_maxTFCE_surf%d_%s_zstat.csv" % (tminame, surface, str(medtype)))
        return _fwe.read_null("output_%s/perm_maxTFCE_surf%d_tcon%d.csv" % (tminame, surface, contrast + 1))
    return _fwe.apply_mfwer(image_array, num_contrasts, surface_range, num_perm, num_surf, nulls, position_array,
                            pos_range, neg_range, weight=weight, mediation=mediation)


def create_full_mask(masking_array):
    fullmask = []
    for m in masking_array:
        if m.shape[2] == 1:
            fullmask = np.hstack((fullmask, m[:, 0, 0]))
        else:
            fullmask = np.hstack((fullmask, m[m == True]))  # noqa: E712
    return fullmask


def lowest_length(num_contrasts, surface_range, tmifilename, medtype=None):
    lengths = []
    for contrast in range(num_contrasts):
        for surface in surface_range:
            if medtype is not None:
                p = "output_%s/perm_maxTFCE_surf%d_%s_zstat.csv" % (tmifilename, surface, medtype)
            else:
                p = "output_%s/perm_maxTFCE_surf%d_tcon%d.csv" % (tmifilename, surface, contrast + 1)
            lengths.append(_fwe.read_null(p).shape[0])
    return np.array(lengths).min()


def create_position_array(masking_array):
    pointer = 0
    position_array = [0]
    for m in masking_array:
        pointer += len(m[m == True])  # noqa: E712
        position_array.append(pointer)
    return position_array
