"""Observed-statistic images: the output half of the step-1 scripts.

* ``vertex_regression_images``
  STEP_1_vertex_tfce_multiple_regression.py:273-398 (no image regressor):
  t of every coefficient by ``tval_int`` with inv(X'X), then per contrast
  j = 1..k-1 the images ``tstat_con<j>`` and ``negtstat_con<j>`` of both
  hemispheres through ``write_vertStat_img`` (TFCE of the statistic scattered
  into the whole hemisphere, rescaled by the masked maximum / 100 and by
  vdensity; its maximum appended to max_TFCE_contrast_values.csv).
* ``voxel_regression_images``
  STEP_1_voxel_tfce_multiple_regression.py:174-230: the same for one volume
  (NaN t set to 0 first, :212), or with ``ancova`` the ``fstat`` image
  sqrt(max(calcF, 0)) (:218-227).
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
from .pyfunc import save_npy, write_vertStat_img, write_voxelStat_img
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


def vertex_regression_images(merge_y, pred_x, adjac_lh, adjac_rh, bin_mask_lh, bin_mask_rh, optstfce,
                             vdensity_lh=1, vdensity_rh=1, surface="area", out_dir=None, save=save_npy):
    """{(statname, hemi): (statistic image, scaled TFCE image)} for every
    contrast and sign; merge_y is [n, V_lh + V_rh] (masked elements)."""
    merge_y = np.asarray(merge_y)
    n = merge_y.shape[0]
    bin_mask_lh = np.asarray(bin_mask_lh, dtype=bool)
    bin_mask_rh = np.asarray(bin_mask_rh, dtype=bool)
    nv_lh = int(bin_mask_lh.sum())
    num_vertex = merge_y.shape[1]
    X = np.column_stack([np.ones(n), pred_x])
    k = X.shape[1]
    invXX = np.linalg.inv(np.dot(X.T, X))
    tvals = tval_int(X, invXX, merge_y, n, k, num_vertex)
    H, E = float(optstfce[0]), float(optstfce[1])
    calc = {"lh": CreateAdjSet(H, E, adjac_lh), "rh": CreateAdjSet(H, E, adjac_rh)}
    masks = {"lh": bin_mask_lh, "rh": bin_mask_rh}
    vds = {"lh": vdensity_lh, "rh": vdensity_rh}
    parts = {"lh": slice(0, nv_lh), "rh": slice(nv_lh, num_vertex)}
    out = {}
    with _in_dir(out_dir):
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
                            ancova=False, save=save_npy):
    """{statname: (statistic, scaled TFCE)} over the in-mask voxels; y is
    [n, num_voxel], ``mask`` the boolean volume (data_index) whose True
    elements, in C order, are the voxels of y and of the adjacency."""
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
