"""Observed-statistic images: the output half of the step-1 scripts.

* ``vertex_regression_images``
  STEP_1_vertex_tfce_multiple_regression.py:273-398: t of every coefficient by
  ``tval_int`` with inv(X'X), or with a vertex-wise regressor (:276-345) by
  ``image_regression`` with the images ``tstat_imgcovar``,
  ``negtstat_imgcovar`` and ``VIF_imgcovar`` (no TFCE) first; then per contrast
  j = 1..k-1 the images ``tstat_con<j>`` and ``negtstat_con<j>`` of both
  hemispheres through ``write_vertStat_img`` (TFCE of the statistic scattered
  into the whole hemisphere, rescaled by the masked maximum / 100 and by
  vdensity; its maximum appended to max_TFCE_contrast_values.csv).
* ``voxel_regression_images``
  STEP_1_voxel_tfce_multiple_regression.py:160-230: the same for one volume
  (with a voxel-wise regressor :162-208; NaN t set to 0 first, :212), or with
  ``ancova`` the ``fstat`` image sqrt(max(calcF, 0)) (:218-227).
* ``vertex_mediation_images`` / ``voxel_mediation_images``
  STEP_1_vertex_tfce_mediation.py:279-288 / STEP_1_voxel_tfce_mediation.py:
  97-104: the Sobel-family z of ``calc_sobelz`` written as ``SobelZ_<medtype>``
  (TFCE image and statistic) of both hemispheres or of the volume.
* ``save_vertex_handoff`` / ``save_vertex_mediation_handoff`` /
  ``save_voxel_handoff`` / ``save_voxel_mediation_handoff``
  the step-1 -> step-2 directories as the step-1 scripts write them
  (STEP_1_vertex_tfce_multiple_regression.py:245-266,
  STEP_1_vertex_tfce_mediation.py:244-277,
  STEP_1_voxel_tfce_multiple_regression.py:85-116 with STEP_0_load_volumes.py:
  89-96, STEP_1_voxel_tfce_mediation.py:90-95): one ``np.save`` per variable
  under the reference's names, ragged adjacency as an object array (which
  ``step2`` reads back only with ``allow_pickle=True``).
* ``fwep_image``  calculate_fweP_vertex.py:45-79 / calculate_fweP_voxel.py:
  1 - p(FWE) of every element > 0 of an observed TFCE image against the max-TFCE
  null (left-neighbour rule), 0 elsewhere; optionally -log10(1 - that).

The t-values come from the FP64 GPU path (``cynumstats.tval_int``) and the TFCE
images from the GPU sweep in FULL mode (``CreateAdjSet.run``).  Files go through
the writers' ``save`` hook (numpy by default; nibabel is not a dependency).
"""
import contextlib
import os

import numpy as np

from .cynumstats import calcF, tval_int
from .fwe import calculate_fwep
from .pyfunc import (calc_sobelz, image_reg_VIF, image_regression, save_npy, write_vertStat_img,
                     write_voxelStat_img)
from .tfce import CreateAdjSet


@contextlib.contextmanager
def _in_dir(path):
    if path is None:
        yield
        return
    os.makedirs(path, exist_ok=True)
    old = os.getcwd()
    os.chdir(path)
    try:
        yield
    finally:
        os.chdir(old)


def _regressors(pred_x, covars):
    return pred_x if covars is None else np.column_stack((pred_x, covars))


def vertex_regression_images(merge_y, pred_x, adjac_lh, adjac_rh, bin_mask_lh, bin_mask_rh, optstfce,
                             vdensity_lh=1, vdensity_rh=1, surface="area", out_dir=None, save=save_npy,
                             img_x=None, covars=None):
    """{(statname, hemi): (statistic image, scaled TFCE image)} for every
    contrast and sign; merge_y is [n, V_lh + V_rh] (masked elements).  With
    ``img_x`` ([n, V_lh + V_rh], the vertex-wise regressor) the t-values come
    from ``image_regression`` on [1, pred_x, covars, img_x] and the regressor's
    own images and the VIF image (of merge_y on pred_x and covars) are written
    first."""
    merge_y = np.asarray(merge_y)
    n = merge_y.shape[0]
    bin_mask_lh = np.asarray(bin_mask_lh, dtype=bool)
    bin_mask_rh = np.asarray(bin_mask_rh, dtype=bool)
    nv_lh = int(bin_mask_lh.sum())
    num_vertex = merge_y.shape[1]
    X = np.column_stack([np.ones(n), pred_x])
    k = X.shape[1]
    H, E = float(optstfce[0]), float(optstfce[1])
    calc = {"lh": CreateAdjSet(H, E, adjac_lh), "rh": CreateAdjSet(H, E, adjac_rh)}
    masks = {"lh": bin_mask_lh, "rh": bin_mask_rh}
    vds = {"lh": vdensity_lh, "rh": vdensity_rh}
    parts = {"lh": slice(0, nv_lh), "rh": slice(nv_lh, num_vertex)}
    out = {}
    with _in_dir(out_dir):
        if img_x is not None:
            tvals, timage = image_regression(merge_y.T.astype(np.float32), np.asarray(img_x).T.astype(np.float32),
                                             pred_x, covars, verbose=False)
            VIF = image_reg_VIF(merge_y, _regressors(pred_x, covars))
            tvals = tvals.T
            timage = timage.T
            for name, stat, tfce in (("tstat_imgcovar", timage, True), ("negtstat_imgcovar", -timage, True),
                                     ("VIF_imgcovar", VIF, False)):
                for hemi in ("lh", "rh"):
                    out[(name, hemi)] = write_vertStat_img(
                        name, stat[parts[hemi]], None, None, surface, hemi, masks[hemi], calc[hemi],
                        masks[hemi].shape[0], vds[hemi], TFCE=tfce, save=save)
        else:
            invXX = np.linalg.inv(np.dot(X.T, X))
            tvals = tval_int(X, invXX, merge_y, n, k, num_vertex)
        for j in range(k - 1):
            tnum = j + 1
            for name, sgn in (("tstat_con%d" % tnum, 1), ("negtstat_con%d" % tnum, -1)):
                for hemi in ("lh", "rh"):
                    stat = tvals[tnum, parts[hemi]]
                    if sgn < 0:
                        stat = stat * -1
                    out[(name, hemi)] = write_vertStat_img(
                        name, stat, None, None, surface, hemi, masks[hemi], calc[hemi],
                        masks[hemi].shape[0], vds[hemi], save=save)
    return out


def voxel_regression_images(y, pred_x, adjac, mask, optstfce, affine=None, imgext=".nii.gz", out_dir=None,
                            ancova=False, save=save_npy, image_x=None, covars=None):
    """{statname: (statistic, scaled TFCE)} over the in-mask voxels; y is
    [n, num_voxel], ``mask`` the boolean volume (data_index) whose True
    elements, in C order, are the voxels of y and of the adjacency.  With
    ``image_x`` ([n, num_voxel], the voxel-wise regressor) the t-values come
    from ``image_regression`` and the regressor's images and the VIF image (of
    image_x on pred_x and covars) are written first."""
    y = np.asarray(y)
    n, num_voxel = y.shape
    mask = np.asarray(mask, dtype=bool)
    X = np.column_stack([np.ones(n), pred_x])
    k = X.shape[1]
    calc = CreateAdjSet(float(optstfce[0]), float(optstfce[1]), adjac)
    vol = np.zeros(mask.shape, dtype=np.float32)
    out = {}
    with _in_dir(out_dir):
        if ancova:
            fvals = calcF(X, y, n, k)
            fvals[fvals < 0] = 0
            out["fstat"] = write_voxelStat_img("fstat", np.sqrt(fvals), vol, mask, affine, calc, imgext,
                                               save=save)
            return out
        if image_x is not None:
            image_x = np.asarray(image_x, dtype=np.float32)
            tvalues, timage = image_regression(y.T.astype(np.float32), image_x.T, pred_x, covars, verbose=False)
            out["tstat_imgcovar"] = write_voxelStat_img("tstat_imgcovar", timage, vol, mask, affine, calc, imgext,
                                                        save=save)
            out["negtstat_imgcovar"] = write_voxelStat_img("negtstat_imgcovar", -timage, vol, mask, affine, calc,
                                                           imgext, save=save)
            tvalues = tvalues.T
            VIF = image_reg_VIF(image_x, _regressors(pred_x, covars))
            out["VIF_imgcovar"] = write_voxelStat_img("VIF_imgcovar", VIF, vol, mask, affine, calc, imgext,
                                                      TFCE=False, save=save)
        else:
            invXX = np.linalg.inv(np.dot(X.T, X))
            tvalues = tval_int(X, invXX, y, n, k, num_voxel)
        tvalues[np.isnan(tvalues)] = 0
        for j in range(k - 1):
            tnum = j + 1
            out["tstat_con%d" % tnum] = write_voxelStat_img("tstat_con%d" % tnum, tvalues[tnum], vol, mask,
                                                            affine, calc, imgext, save=save)
            out["negtstat_con%d" % tnum] = write_voxelStat_img("negtstat_con%d" % tnum, tvalues[tnum] * -1, vol,
                                                               mask, affine, calc, imgext, save=save)
    return out


def vertex_mediation_images(merge_y, pred_x, depend_y, medtype, adjac_lh, adjac_rh, bin_mask_lh, bin_mask_rh,
                            optstfce, vdensity_lh=1, vdensity_rh=1, surface="area", out_dir=None, save=save_npy,
                            alg="aroian"):
    """{("SobelZ_<medtype>", hemi): (statistic image, scaled TFCE image)}:
    the observed z of mediation type I, M or Y over merge_y ([n, V_lh + V_rh],
    masked elements, covariates already removed) and its TFCE images."""
    merge_y = np.asarray(merge_y)
    n, num_vertex = merge_y.shape
    bin_mask_lh = np.asarray(bin_mask_lh, dtype=bool)
    bin_mask_rh = np.asarray(bin_mask_rh, dtype=bool)
    nv_lh = int(bin_mask_lh.sum())
    H, E = float(optstfce[0]), float(optstfce[1])
    sobel = calc_sobelz(medtype, pred_x, depend_y, merge_y, n, num_vertex, alg)
    name = "SobelZ_%s" % medtype
    out = {}
    with _in_dir(out_dir):
        for hemi, part, mask, adj, vd in (("lh", slice(0, nv_lh), bin_mask_lh, adjac_lh, vdensity_lh),
                                          ("rh", slice(nv_lh, num_vertex), bin_mask_rh, adjac_rh, vdensity_rh)):
            out[(name, hemi)] = write_vertStat_img(name, sobel[part], None, None, surface, hemi, mask,
                                                   CreateAdjSet(H, E, adj), mask.shape[0], vd, save=save)
    return out


def voxel_mediation_images(y, pred_x, depend_y, medtype, adjac, mask, optstfce, affine=None, imgext=".nii.gz",
                           out_dir=None, save=save_npy, alg="aroian"):
    """{"SobelZ_<medtype>": (statistic, scaled TFCE)} over the in-mask voxels
    of y ([n, num_voxel], covariates already removed)."""
    y = np.asarray(y)
    n, num_voxel = y.shape
    mask = np.asarray(mask, dtype=bool)
    calc = CreateAdjSet(float(optstfce[0]), float(optstfce[1]), adjac)
    sobel = calc_sobelz(medtype, pred_x, depend_y, y, n, num_voxel, alg)
    name = "SobelZ_%s" % medtype
    with _in_dir(out_dir):
        return {name: write_voxelStat_img(name, sobel, np.zeros(mask.shape, dtype=np.float32), mask, affine, calc,
                                          imgext, save=save)}


def _adjacency_object_array(adjac):
    """Ragged neighbour lists as the object array np.save writes for them;
    a CSR pair (row_ptr, col) is split into its rows first."""
    if isinstance(adjac, tuple) and len(adjac) == 2:
        row_ptr, col = (np.asarray(a) for a in adjac)
        rows = [col[row_ptr[i]:row_ptr[i + 1]].tolist() for i in range(len(row_ptr) - 1)]
    else:
        rows = [list(r) for r in adjac]
    arr = np.empty(len(rows), dtype=object)
    for i, r in enumerate(rows):
        arr[i] = r
    return arr


def _save_all(directory, items):
    os.makedirs(directory, exist_ok=True)
    for name, value in items:
        np.save(os.path.join(directory, name), value)
    return directory


def save_vertex_handoff(surface, pred_x, merge_y, adjac_lh, adjac_rh, bin_mask_lh, bin_mask_rh, optstfce,
                        vdensity_lh=1, vdensity_rh=1, affine_mask_lh=None, affine_mask_rh=None, root="."):
    """``python_temp_<surface>/`` of the vertex GLM; merge_y is [n, V_lh + V_rh]
    (masked elements), the masks are the full hemispheres'."""
    merge_y = np.asarray(merge_y)
    bin_mask_lh = np.asarray(bin_mask_lh, dtype=bool)
    bin_mask_rh = np.asarray(bin_mask_rh, dtype=bool)
    nl, nr = int(bin_mask_lh.sum()), int(bin_mask_rh.sum())
    eye = np.eye(4)
    return _save_all(os.path.join(root, "python_temp_%s" % surface), [
        ("pred_x", pred_x), ("num_subjects", merge_y.shape[0]), ("all_vertex", bin_mask_lh.shape[0]),
        ("num_vertex", nl + nr), ("num_vertex_lh", nl), ("num_vertex_rh", nr), ("bin_mask_lh", bin_mask_lh),
        ("bin_mask_rh", bin_mask_rh), ("affine_mask_lh", eye if affine_mask_lh is None else affine_mask_lh),
        ("affine_mask_rh", eye if affine_mask_rh is None else affine_mask_rh),
        ("adjac_lh", _adjacency_object_array(adjac_lh)), ("adjac_rh", _adjacency_object_array(adjac_rh)),
        ("merge_y", merge_y.astype(np.float32, order="C")), ("optstfce", np.asarray(optstfce)),
        ("vdensity_lh", vdensity_lh), ("vdensity_rh", vdensity_rh)])


def save_vertex_mediation_handoff(surface, pred_x, depend_y, merge_y, adjac_lh, adjac_rh, bin_mask_lh, bin_mask_rh,
                                  optstfce, vdensity_lh=1, vdensity_rh=1, root="."):
    """``python_temp_med_<surface>/`` of the vertex mediation."""
    merge_y = np.asarray(merge_y)
    bin_mask_lh = np.asarray(bin_mask_lh, dtype=bool)
    bin_mask_rh = np.asarray(bin_mask_rh, dtype=bool)
    nl, nr = int(bin_mask_lh.sum()), int(bin_mask_rh.sum())
    return _save_all(os.path.join(root, "python_temp_med_%s" % surface), [
        ("pred_x", pred_x), ("depend_y", depend_y), ("num_subjects", merge_y.shape[0]), ("num_vertex", nl + nr),
        ("num_vertex_lh", nl), ("num_vertex_rh", nr), ("all_vertex", bin_mask_lh.shape[0]),
        ("bin_mask_lh", bin_mask_lh), ("bin_mask_rh", bin_mask_rh),
        ("adjac_lh", _adjacency_object_array(adjac_lh)), ("adjac_rh", _adjacency_object_array(adjac_rh)),
        ("optstfce", np.asarray(optstfce)), ("vdensity_lh", vdensity_lh), ("vdensity_rh", vdensity_rh),
        ("merge_y", merge_y.astype(np.float32, order="C"))])


def _voxel_common(y, mask, affine_mask, raw=None):
    y = np.asarray(y)
    mask = np.asarray(mask)
    return [("raw_nonzero", (y.T if raw is None else np.asarray(raw)).astype(np.float32, order="C")),
            ("affine_mask", np.eye(4) if affine_mask is None else affine_mask),
            ("data_mask", mask.astype(np.float32)), ("num_voxel", int((mask > 0.99).sum())),
            ("num_subjects", y.shape[0])]


def save_voxel_handoff(y, pred_x, adjac, mask, optstfce, ancova=0, covars=None, raw_nonzero=None,
                       affine_mask=None, root="."):
    """``python_temp/`` of the voxel GLM: STEP_0's volume files (raw_nonzero
    = ``raw_nonzero`` or y transposed, affine, data_mask, num_voxel,
    num_subjects) and step 1's (covars when given, adjac, pred_x, ancova,
    optstfce, raw_nonzero_corr = y transposed); y is [n, num_voxel]."""
    items = _voxel_common(y, mask, affine_mask, raw_nonzero)
    if covars is not None:
        items.append(("covars", covars))
    items += [("adjac", _adjacency_object_array(adjac)), ("pred_x", pred_x), ("ancova", int(ancova)),
              ("optstfce", np.asarray(optstfce)), ("raw_nonzero_corr", np.asarray(y).T.astype(np.float32, order="C"))]
    return _save_all(os.path.join(root, "python_temp"), items)


def save_voxel_mediation_handoff(y, pred_x, depend_y, medtype, adjac, mask, optstfce, raw_nonzero=None,
                                 affine_mask=None, root="."):
    """``python_temp/`` of the voxel mediation (STEP_0's volume files and
    step 1's pred_x, depend_y, adjac, medtype, optstfce, raw_nonzero_corr)."""
    items = _voxel_common(y, mask, affine_mask, raw_nonzero)
    items += [("pred_x", pred_x), ("depend_y", depend_y), ("adjac", _adjacency_object_array(adjac)),
              ("medtype", medtype), ("optstfce", np.asarray(optstfce)),
              ("raw_nonzero_corr", np.asarray(y).T.astype(np.float32, order="C"))]
    return _save_all(os.path.join(root, "python_temp"), items)


def fwep_image(tfce_image, perm_tfce_max, neglog10=False):
    """1 - p(FWE) image (float64, the input's shape): elements > 0 get the
    left-neighbour p of the sorted null, the rest 0.  With ``neglog10`` also
    returns -log10(1 - p) on the same elements."""
    data = np.asarray(tfce_image)
    flat = data.reshape(-1)
    sel = flat > 0
    corrp = np.zeros(flat.shape, dtype=np.float64)
    if sel.any():
        corrp[sel] = calculate_fwep(flat[sel], np.asarray(perm_tfce_max, dtype=np.float64))
    corrp = corrp.reshape(data.shape)
    if not neglog10:
        return corrp
    nl = np.zeros_like(corrp)
    with np.errstate(divide="ignore"):
        nl.reshape(-1)[sel] = -np.log10(1 - corrp.reshape(-1)[sel])
    return corrp, nl
