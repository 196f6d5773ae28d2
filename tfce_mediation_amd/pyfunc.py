"""Drop-in for the permutation-path subset of ``tfce_mediation.pyfunc``.

* ``create_adjac_vertex`` / ``create_adjac_voxel``   pyfunc.py:37-76
* ``write_vertStat_img`` / ``write_voxelStat_img``     pyfunc.py:80-103
  (observed TFCE images; nibabel is not a dependency, so images go through a
  ``save(fname, data, affine)`` hook, by default ``numpy.save`` to the same
  file name with ``.npy`` in place of ``.mgh`` / ``.nii.gz``)
* ``write_perm_maxTFCE_vertex`` / ``_voxel``          pyfunc.py:107-126
* ``calc_sobelz`` / ``calc_indirect``                 pyfunc.py:130-162, 2678-2709
* ``find_nearest``                                    pyfunc.py:254-261
* ``zscaler``                                         pyfunc.py:264-271
* ``image_regression`` / ``image_reg_VIF``            pyfunc.py:1024-1114
* ``glm_typeI``                                       pyfunc.py:2282-2404
* ``check_blocks`` / ``rand_blocks``                  pyfunc.py:2711-2757
The TFCE and statistics run on the GPU; the null writers append to the same
files in the same text format as the reference, with Python file I/O instead
of one ``os.system('echo ...')`` per value.
"""
import math
import os

import numpy as np

from . import _lib
from .adjacency import create_adjac_vertex, create_adjac_voxel  # noqa: F401
from .cynumstats import calc_beta_se


def _tfce_max_scaled(calc, stat_full, vdensity):
    """max over elements of TFCE(stat) * (max(stat)/100) * vdensity (one map)."""
    import torch
    adj = calc.adjacency
    d = torch.as_tensor(np.ascontiguousarray(stat_full, dtype=np.float32)).cuda()[None, :]
    vd = None
    if not np.isscalar(vdensity) and np.asarray(vdensity).ndim > 0:
        vd = torch.as_tensor(np.ascontiguousarray(vdensity, dtype=np.float32)).cuda()
    out = adj.batch_max(d, 1, vdensity=vd).cpu().numpy()[0, 0]
    if vd is None and float(vdensity) != 1.0:
        out = np.float32(out) * vdensity
    return out


def save_npy(fname, data, affine=None):
    """Default image writer: ``numpy.save`` under the image name with its
    extension (.mgh, .nii, .nii.gz) replaced by .npy; the affine is dropped."""
    for ext in (".nii.gz", ".mgh", ".nii"):
        if fname.endswith(ext):
            fname = fname[: -len(ext)]
            break
    np.save(fname + ".npy", np.asarray(data))


def _append_max(statline):
    with open("max_TFCE_contrast_values.csv", "a") as f:
        f.write(statline + "\n")


def _tfce_full(calc, image):
    enh = np.zeros_like(image)
    calc.run(image, enh)
    return enh


def write_vertStat_img(statname, vertStat, outdata_mask, affine_mask, surf, hemi, bin_mask, TFCEfunc,
                       all_vertex, density_corr=1, TFCE=True, save=save_npy):
    """Observed statistic image of one hemisphere (pyfunc.py:80-92).

    The masked statistic is scattered into a zero image of ``all_vertex``
    elements; with TFCE the image ``TFCE(stat) * (max(finite masked stat)/100)
    * density_corr`` is saved as ``<statname>_<surf>_<hemi>_TFCE.mgh`` and its
    maximum appended to max_TFCE_contrast_values.csv ("%s_%s_%s,%f"); then the
    statistic image is saved as ``<statname>_<surf>_<hemi>.mgh``.  The product
    is the reference's numpy expression on the caller's dtypes (a float64
    statistic gives a float64 scale and product), rounded where it is stored in
    ``outdata_mask`` ([all_vertex, 1, 1] float32, made when None is passed).
    Returns (statistic image, scaled TFCE image or None).
    """
    vertStat = np.asarray(vertStat)
    if outdata_mask is None:
        outdata_mask = np.zeros((all_vertex, 1, 1), dtype=np.float32)
    vertStat_out = np.zeros(all_vertex).astype(np.float32, order="C")
    vertStat_out[bin_mask] = vertStat
    tfce_img = None
    if TFCE:
        enh = _tfce_full(TFCEfunc, vertStat_out)
        outdata_mask[:, 0, 0] = enh * (vertStat[np.isfinite(vertStat)].max() / 100) * density_corr
        tfce_img = outdata_mask[:, 0, 0].copy()
        _append_max("%s_%s_%s,%f" % (statname, surf, hemi,
                                      outdata_mask[np.isfinite(outdata_mask[:, 0, 0])].max()))
        save("%s_%s_%s_TFCE.mgh" % (statname, surf, hemi), outdata_mask, affine_mask)
    outdata_mask[:, 0, 0] = vertStat_out
    save("%s_%s_%s.mgh" % (statname, surf, hemi), outdata_mask, affine_mask)
    return vertStat_out, tfce_img


def write_voxelStat_img(statname, voxelStat, out_path, data_index, affine, TFCEfunc, imgext=".nii.gz",
                        TFCE=True, save=save_npy):
    """Observed statistic volume (pyfunc.py:94-103): with TFCE,
    ``TFCE(stat) * (max(stat)/100)`` (float32) goes into ``out_path[data_index]``,
    is saved as ``<statname>_TFCE<imgext>`` and its volume maximum appended to
    max_TFCE_contrast_values.csv ("%s,%f"); then the statistic itself is saved
    as ``<statname><imgext>``.  Returns (statistic, scaled TFCE or None) over
    the in-mask voxels."""
    tfce_vals = None
    if TFCE:
        voxelStat_out = np.asarray(voxelStat).astype(np.float32, order="C")
        enh = _tfce_full(TFCEfunc, voxelStat_out)
        out_path[data_index] = enh * (voxelStat_out.max() / 100)
        tfce_vals = np.asarray(out_path[data_index]).copy()
        save("%s_TFCE%s" % (statname, imgext), out_path, affine)
        _append_max("%s,%f" % (statname, out_path.max()))
    out_path[data_index] = voxelStat
    save("%s%s" % (statname, imgext), out_path, affine)
    return np.asarray(out_path[data_index]).copy(), tfce_vals


def write_perm_maxTFCE_vertex(statname, vertStat, num_vertex, bin_mask_lh, bin_mask_rh, calcTFCE_lh,
                              calcTFCE_rh, density_corr_lh=1, density_corr_rh=1):
    vertStat = np.asarray(vertStat)
    lh = np.zeros(bin_mask_lh.shape[0], dtype=np.float32)
    rh = np.zeros(bin_mask_rh.shape[0], dtype=np.float32)
    lh[bin_mask_lh] = vertStat[:num_vertex]
    rh[bin_mask_rh] = vertStat[num_vertex:]
    m_lh = _tfce_max_scaled(calcTFCE_lh, lh, density_corr_lh)
    m_rh = _tfce_max_scaled(calcTFCE_rh, rh, density_corr_rh)
    maxTFCE = np.array([m_lh, m_rh]).max()
    with open("perm_%s_TFCE_maxVertex.csv" % statname, "a") as f:
        f.write("%.4f\n" % maxTFCE)
    return maxTFCE


def write_perm_maxTFCE_voxel(statname, voxelStat, TFCEfunc):
    v = np.asarray(voxelStat).astype(np.float32, order="C")
    maxval = _tfce_max_scaled(TFCEfunc, v, 1)
    with open("perm_%s_TFCE_maxVoxel.csv" % statname, "a") as f:
        f.write("%1.4f\n" % maxval)
    return maxval


def zscaler(X, axis=0, w_mean=True, w_std=True):
    """Centre and/or scale X along ``axis`` by its mean and (population)
    standard deviation, in X's dtype (pyfunc.py:264-271)."""
    out = np.array(X, copy=True)
    if w_mean:
        out -= np.mean(out, axis)
    if w_std:
        out /= np.std(out, axis)
    return out


def image_regression(y, image_x, pred_x, covars=None, normalize=False, verbose=True, stream=None):
    """Regression with an element-wise regressor (pyfunc.py:1024-1086).

    ``y`` and ``image_x`` are [elements, subjects]; per element the design is
    [1, pred_x, covars, image_x[element]] (covars optional).  Returns float32
    ``(Tval, Timg)``: the t of the intercept and pred_x columns
    (``arr[:, :len(pred_x.T) + 1]`` of the reference's per-element t array, so a
    one-dimensional pred_x keeps every column) and the t of the element's
    regressor; elements whose regressor holds a NaN or has std() < 0.01 are 0.
    FP64 per element on the GPU (``tfce_image_regression``); with
    ``normalize`` image_x is first z-scaled along axis 0 (``zscaler``).
    """
    import time
    import torch
    t0 = time.time()
    y = np.asarray(y)
    image_x = np.asarray(image_x)
    pred_x = np.asarray(pred_x)
    nv, n = y.shape
    regressors = pred_x if covars is None else np.column_stack((pred_x, covars))
    R = np.ascontiguousarray(np.column_stack([np.ones(n), regressors]), dtype=np.float64)
    if normalize:
        image_x = zscaler(image_x, axis=0, w_mean=True, w_std=True)
    r = R.shape[1]
    dy = torch.as_tensor(np.ascontiguousarray(y.T, dtype=np.float32)).cuda()
    dz = torch.as_tensor(np.ascontiguousarray(image_x.T, dtype=np.float32)).cuda()
    dt = torch.empty((r + 1, nv), dtype=torch.float32, device=dy.device)
    lib = _lib.lib()
    _lib.check(lib.tfce_image_regression(_lib.ptr(dy), nv, _lib.ptr(dz), nv, R.ctypes.data, n, r, nv, 0.01,
                                         _lib.ptr(dt), nv, _lib.stream_ptr(stream)))
    arr = dt.cpu().numpy().T
    Tval = np.array(arr[:, :len(pred_x.T) + 1], dtype=np.float32)
    Timg = np.array(np.squeeze(arr[:, -1:]), dtype=np.float32)
    if verbose:
        print("Finished. Image-wise independent variable regression took %.1f seconds" % (time.time() - t0))
    return Tval, Timg


def image_reg_VIF(y, regressors, stream=None):
    """Variance inflation image (pyfunc.py:1088-1114): per element (column
    of the [subjects, elements] array y) 1 / (1 - R2) of y on [1, regressors],
    NaN set to 0, as float64.  The sum of squares about the element's mean is
    formed in y's own type (float32 or float64) as numpy forms it; the fit in
    FP64 on the GPU (``tfce_image_vif``)."""
    import torch
    y = np.asarray(y)
    f64 = y.dtype == np.float64
    X = np.ascontiguousarray(np.column_stack([np.ones(len(regressors)), regressors]), dtype=np.float64)
    n, V = y.shape
    dy = torch.as_tensor(np.ascontiguousarray(y, dtype=np.float64 if f64 else np.float32)).cuda()
    dv = torch.empty(V, dtype=torch.float64, device=dy.device)
    lib = _lib.lib()
    _lib.check(lib.tfce_image_vif(_lib.ptr(dy), 1 if f64 else 0, V, X.ctypes.data, n, X.shape[1], V,
                                  _lib.ptr(dv), _lib.stream_ptr(stream)))
    return dv.cpu().numpy()


def calc_indirect(ta, tb, alg="aroian"):
    """Sobel-type z from path t-values, on the GPU (float32 result)."""
    import torch
    code = {"aroian": 0, "sobel": 1, "goodman": 2}.get(alg)
    if code is None:
        raise ValueError("Unknown indirect test algorithm")
    tb_t = torch.as_tensor(np.ascontiguousarray(np.atleast_1d(tb), dtype=np.float32)).cuda()
    ta_arr = np.atleast_1d(ta)
    ta_t = None
    ta_c = 0.0
    if ta_arr.size == 1 and tb_t.numel() != 1:
        ta_c = float(ta_arr[0])
    else:
        ta_t = torch.as_tensor(np.ascontiguousarray(ta_arr, dtype=np.float32)).cuda()
    z = torch.empty_like(tb_t)
    _lib.check(_lib.lib().tfce_sobel_z(_lib.ptr(ta_t), _lib.ptr(tb_t), tb_t.numel(), code, ta_c, _lib.ptr(z),
                                       _lib.stream_ptr()))
    return z.cpu().numpy()


def calc_sobelz(medtype, pred_x, depend_y, merge_y, n, num_vertex, alg="aroian"):
    if medtype == "I":
        a_b, a_se = calc_beta_se(pred_x, merge_y, n, num_vertex)
        b_b, b_se = calc_beta_se(np.column_stack([pred_x, depend_y]), merge_y, n, num_vertex)
        ta, tb = a_b / a_se[1], b_b / b_se[1]
    elif medtype == "M":
        a_b, a_se = calc_beta_se(pred_x, merge_y, n, num_vertex)
        b_b, b_se = calc_beta_se(np.column_stack([depend_y, pred_x]), merge_y, n, num_vertex)
        ta, tb = a_b / a_se[1], b_b / b_se[1]
    elif medtype == "Y":
        from scipy.stats import linregress
        lr = linregress(pred_x, depend_y)
        ta = lr.slope / lr.stderr
        b_b, b_se = calc_beta_se(np.column_stack([depend_y, pred_x]), merge_y, n, num_vertex)
        tb = b_b / b_se[1]
    else:
        raise ValueError("Invalid mediation type")
    return calc_indirect(ta, tb, alg).astype(np.float64)


def glm_typeI(endog, exog, dmy_covariates=None, output_fvalues=True, output_tvalues=False, output_pvalues=False,
              verbose=True, rand_array=None, use_reduced_residuals=False, output_reduced_residuals=False,
              exog_names=None):
    """GLM with one F per exog variable (pyfunc.py:2282-2404): design [1, exog
    variables, covariates], rows permuted by rand_array; the model F, each
    variable's F from the fit without its columns, and t = a / se of every
    column (se float32).  The least-squares fits run on the GPU in FP64
    (``tfce_ols_direct``) on float32 data; with use_reduced_residuals the
    residuals that replace endog are rounded to float32 as well (the
    reference keeps them in float64).  Same return tuples as the reference."""
    from scipy.stats import f as fdist
    from scipy.stats import t as tdist
    from .cynumstats import _direct
    from .randomise import glm_design
    endog = np.asarray(endog)
    n = endog.shape[0]
    X, blocks = glm_design(n, exog, dmy_covariates)
    if rand_array is not None:
        if use_reduced_residuals:
            endog = _direct(X, endog, resid=True)["resid"]
        X = X[np.asarray(rand_array)]
    reduced_data = None
    if output_reduced_residuals:
        reduced_data = _direct(X, endog, resid=True)["resid"]
    k = X.shape[1]
    df_between, df_within, df_total = k - 1, n - k, n - 1
    full = _direct(X, endog, coef=True, se=output_tvalues, sse=True)
    ss_total = _direct(np.ones((n, 1)), endog, sse=True)["sse"]
    ss_res = full["sse"]
    if output_fvalues:
        ss_between = ss_total - ss_res
        ms_res = ss_res / df_within
        fvalues = (ss_between / df_between) / ms_res
        if verbose:
            print("Source\t\tDF\tF(Max)")
            print("Model\t\t(%d,%d)\t%.2f" % (df_between, df_within, fvalues.max()))
        fvar, pvar = [], []
        for i, (s0, w) in enumerate(blocks):
            ss_model = ss_total - _direct(np.delete(X, np.s_[s0:s0 + w], 1), endog, sse=True)["sse"]
            ftemp = (ss_between - ss_model) / (ms_res * w)
            fvar.append(ftemp)
            if verbose:
                label = exog_names[i] if exog_names is not None else "Exog%d" % (i + 1)
                print("%s\t\t(%d,%d)\t%.2f" % (label, w, df_within, ftemp.max()))
            if output_pvalues:
                pvar.append(fdist.sf(ftemp, w, df_within))
    if output_tvalues:
        tvalues = full["coef"] / full["se"]
        if endog.ndim == 1:
            tvalues = tvalues[:, 0]
    if output_fvalues and endog.ndim == 1:
        fvalues = fvalues[0]
        fvar = [v[0] for v in fvar]
        pvar = [v[0] for v in pvar]
    if output_tvalues and output_fvalues:
        if output_pvalues:
            return (fvalues, np.array(fvar), tvalues, fdist.sf(fvalues, df_between, df_within), np.array(pvar),
                    tdist.sf(np.abs(tvalues), df_total) * 2)
        if output_reduced_residuals:
            return fvalues, np.array(fvar), tvalues, reduced_data
        return fvalues, np.array(fvar), tvalues
    if output_tvalues:
        if output_pvalues:
            return tvalues, tdist.sf(np.abs(tvalues), df_total) * 2
        if output_reduced_residuals:
            return tvalues, reduced_data
        return tvalues
    if output_fvalues:
        if output_pvalues:
            return fvalues, np.array(fvar), fdist.sf(fvalues, df_between, df_within), np.array(pvar)
        if output_reduced_residuals:
            return fvalues, np.array(fvar), reduced_data
        return fvalues, np.array(fvar)
    print("No output has been selected")


def find_nearest(array, value, p_array):
    idx = np.searchsorted(array, value, side="left")
    if idx == len(p_array):
        return p_array[idx - 1]
    elif math.fabs(value - array[idx - 1]) < math.fabs(value - array[idx]):
        return p_array[idx - 1]
    else:
        return p_array[idx]


def check_blocks(block_list):
    unique_blocks = np.unique(block_list)
    sizes = [len(block_list[block_list == b]) for b in unique_blocks]
    equal = all(x == sizes[0] for x in sizes)
    if not equal:
        print("Warning: blocks are not equal. Swaping with only occur within blocks, but not among blocks.")
    return equal


def rand_blocks(block_list, is_equal_sizes):
    """Permutation index array within (and, for equal sizes, among) blocks,
    drawn from numpy's global RNG exactly as pyfunc.py:2738-2757."""
    indexer = np.array(range(len(block_list)))
    randindex = []
    if is_equal_sizes is True:
        for block in np.random.permutation(list(np.unique(block_list))):
            randindex.append(np.random.permutation(indexer[block_list == block]))
    else:
        for block in np.unique(block_list):
            randindex.append(np.random.permutation(indexer[block_list == block]))
    return np.concatenate(randindex)
