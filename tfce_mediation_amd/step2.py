"""Step 2 from the step-1 hand-off directories: the ``run(opts)`` bodies of
the reference's randomise scripts, reading what step 1 saved and writing the
null CSVs where the reference writes them.

* ``vertex_regression``  vertex_tfce_multiple_regression_randomise.py:56-118
  reads ``python_temp_<surface>/`` (written by
  STEP_1_vertex_tfce_multiple_regression.py:248-266) and appends to
  ``output_<surface>/perm_Tstat_<surface>/perm_tstat_con<j>_TFCE_maxVertex.csv``;
* ``voxel_regression``   voxel_tfce_multiple_regression_randomise.py:51-120
  reads ``python_temp/`` (STEP_1_voxel_tfce_multiple_regression.py:112-116)
  and appends to ``output/perm_Tstat/``; with ``ancova`` saved as 1 it runs
  the F branch over twice the permutation numbers (:79-80);
* ``vertex_mediation``   vertex_tfce_mediation_randomise.py:48-92 reads
  ``python_temp_med_<surface>/`` (STEP_1_vertex_tfce_mediation.py:244-277)
  and appends to ``output_med_<surface>/perm_SobelZ_<medtype>/``;
* ``voxel_mediation``    voxel_tfce_mediation_randomise.py:42-76 reads
  ``python_temp/`` (STEP_1_voxel_tfce_mediation.py:90-95) and appends to
  ``output_med_<medtype>/perm_SobelZ/``;
* ``mmr_lr``             the permutation branch of tm_mmr_rand_low_ram_parallel.py:
  174-212 (mmr-lr, one surface of a multimodal set per call) reads
  ``tmi_temp/`` (saved by tm_mmr_rand_low_ram.py:150-201) and appends
  ``perm_maxTFCE_surf<s>_tcon<c>.csv`` (t, +t then -t) or
  ``perm_maxTFCE_surf<s>_<medtype>_zstat.csv`` (Sobel z) to the output path;
* ``tm_glm``             the GLM branch of tm_models_randomise.py:68-243 reads
  ``tmtemp_GLM_<surface>/`` or ``tmtemp_GLM_volume/`` (saved by
  STEP_1_tm_models.py:156-175, 664-697) and appends to
  ``output_GLM_<surface>/perm_GLM/`` or ``output_GLM_volume/perm_GLM/``;
* ``tm_mediation``       the mediation branch of tm_models_randomise.py:
  92-102, 430-510 reads ``tmtemp_mediation_<surface>/`` or
  ``tmtemp_mediation_volume/`` (saved by STEP_1_tm_models.py:952-1012) and
  appends ``perm_Zstat_<medtype>_TFCE_max*.csv`` to
  ``output_mediation_<surface|volume>/perm_mediation/``.

The permutations of permutation numbers ``start..stop`` come from
``randomise.permutation_indices`` (numpy's legacy RNG seeded with number +
seed): the reference seeds from the clock (e.g.
vertex_tfce_multiple_regression_randomise.py:91), which cannot be reproduced.

Step 1 saves ragged adjacency lists with ``np.save`` as object arrays, which
numpy can only read by unpickling, and unpickling runs code from the file.
``allow_pickle`` must therefore be set explicitly, and only for directories
that the caller's own step-1 run wrote.
"""
import os

import numpy as np

from . import randomise


def load_temp(directory, names, allow_pickle=False):
    """{name: array} of ``directory/<name>.npy``."""
    out = {}
    for name in names:
        path = os.path.join(directory, name + ".npy")
        try:
            out[name] = np.load(path, allow_pickle=allow_pickle)
        except ValueError as e:
            raise ValueError("%s holds Python objects (ragged adjacency lists); pass allow_pickle=True only "
                             "for a directory your own step-1 run wrote" % path) from e
    return out


def _out_dir(root, *parts):
    path = os.path.join(root, *parts)
    os.makedirs(path, exist_ok=True)
    return path


def _perms(n, numbers, seed, blocks=None):
    return randomise.permutation_indices(int(n), list(numbers), seed=seed, blocks=blocks)


def vertex_regression(surface, start, stop, root=".", specifyvars=None, exchangeblock=None, seed=0,
                      allow_pickle=False, batch=None, precision="fp16x3"):
    """Permutation numbers start..stop (inclusive) of the vertex GLM of
    ``surface``; returns the driver's nulls {contrast: [n_perm, 2]}."""
    tmp = os.path.join(root, "python_temp_%s" % surface)
    v = load_temp(tmp, ["merge_y", "num_subjects", "pred_x", "adjac_lh", "adjac_rh", "bin_mask_lh",
                        "bin_mask_rh", "optstfce", "vdensity_lh", "vdensity_rh"], allow_pickle)
    blocks = None
    if exchangeblock is not None:
        blocks = np.genfromtxt(exchangeblock, dtype=str)
    perms = _perms(v["num_subjects"], range(int(start), int(stop) + 1), seed, blocks)
    out = _out_dir(root, "output_%s" % surface, "perm_Tstat_%s" % surface)
    return randomise.vertex_regress(v["merge_y"], v["pred_x"], v["adjac_lh"], v["adjac_rh"], v["bin_mask_lh"],
                                    v["bin_mask_rh"], v["optstfce"], perms, vdensity_lh=v["vdensity_lh"],
                                    vdensity_rh=v["vdensity_rh"], specifyvars=specifyvars, out_dir=out,
                                    batch=batch, precision=precision)


def voxel_regression(start, stop, root=".", specifyvars=None, exchangeblock=None, seed=0, allow_pickle=False,
                     batch=None, precision="fp16x3"):
    """Permutation numbers start..stop of the voxel GLM saved in
    ``python_temp/`` (start..2*stop in the F branch)."""
    tmp = os.path.join(root, "python_temp")
    v = load_temp(tmp, ["num_subjects", "raw_nonzero_corr", "pred_x", "adjac", "ancova", "optstfce"],
                  allow_pickle)
    ancova = int(v["ancova"]) == 1
    numbers = range(int(start), 2 * int(stop) + 1) if ancova else range(int(start), int(stop) + 1)
    blocks = None
    if exchangeblock is not None and not ancova:
        blocks = np.genfromtxt(exchangeblock, dtype=str)
    perms = _perms(v["num_subjects"], numbers, seed, blocks)
    out = _out_dir(root, "output", "perm_Tstat")
    return randomise.voxel_regress(v["raw_nonzero_corr"].T, v["pred_x"], v["adjac"], v["optstfce"], perms,
                                   specifyvars=specifyvars, out_dir=out, batch=batch, precision=precision,
                                   ancova=ancova)


def vertex_mediation(surface, medtype, start, stop, root=".", seed=0, allow_pickle=False, batch=None,
                     precision="fp16x3"):
    """Permutation numbers start..stop of the vertex mediation of ``surface``."""
    tmp = os.path.join(root, "python_temp_med_%s" % surface)
    v = load_temp(tmp, ["merge_y", "num_subjects", "pred_x", "depend_y", "adjac_lh", "adjac_rh", "bin_mask_lh",
                        "bin_mask_rh", "optstfce", "vdensity_lh", "vdensity_rh"], allow_pickle)
    perms = _perms(v["num_subjects"], range(int(start), int(stop) + 1), seed)
    out = _out_dir(root, "output_med_%s" % surface, "perm_SobelZ_%s" % medtype)
    return randomise.vertex_mediation(v["merge_y"], v["pred_x"], v["depend_y"], medtype, v["adjac_lh"],
                                      v["adjac_rh"], v["bin_mask_lh"], v["bin_mask_rh"], v["optstfce"], perms,
                                      vdensity_lh=v["vdensity_lh"], vdensity_rh=v["vdensity_rh"], out_dir=out,
                                      batch=batch, precision=precision)


def voxel_mediation(medtype, start, stop, root=".", seed=0, allow_pickle=False, batch=None, precision="fp16x3"):
    """Permutation numbers start..stop of the voxel mediation saved in ``python_temp/``."""
    tmp = os.path.join(root, "python_temp")
    v = load_temp(tmp, ["num_subjects", "raw_nonzero_corr", "pred_x", "depend_y", "adjac", "optstfce"],
                  allow_pickle)
    perms = _perms(v["num_subjects"], range(int(start), int(stop) + 1), seed)
    out = _out_dir(root, "output_med_%s" % medtype, "perm_SobelZ")
    return randomise.voxel_mediation(v["raw_nonzero_corr"].T, v["pred_x"], v["depend_y"], medtype, v["adjac"],
                                     v["optstfce"], perms, out_dir=out, batch=batch, precision=precision)


def tm_glm(start, stop, surface=None, root=".", exchangeblock=None, seed=0, allow_pickle=False, batch=None,
           precision="fp16x3"):
    """Permutation numbers start..stop (inclusive) of the tm-models GLM saved
    for ``surface`` (None: the volume); returns {statname: null}."""
    name = surface if surface is not None else "volume"
    tmp = os.path.join(root, "tmtemp_GLM_%s" % name)
    common = ["exog_flat", "exog_shape", "varnames", "gstat", "data", "optstfce", "dmy_covariates"]
    if surface is not None:
        v = load_temp(tmp, common + ["mask_lh", "mask_rh", "adjac_lh", "adjac_rh", "vdensity_lh", "vdensity_rh"],
                      allow_pickle)
    else:
        v = load_temp(tmp, common + ["adjac"], allow_pickle)
    exog, count = [], 0
    for nc in np.atleast_1d(v["exog_shape"]):
        exog.append(v["exog_flat"][:, count:count + int(nc)])
        count += int(nc)
    cov = v["dmy_covariates"]
    if cov.dtype == object and cov.ndim == 0 and cov.item() is None:
        cov = None  # np.save of None (tm_models_randomise.py:177-178)
    data = v["data"]
    blocks = None
    if exchangeblock is not None:
        blocks = np.genfromtxt(exchangeblock, dtype=str)
    perms = _perms(data.shape[0], range(int(start), int(stop) + 1), seed, blocks)
    out = _out_dir(root, "output_GLM_%s" % name, "perm_GLM")
    varnames = [str(x) for x in np.atleast_1d(v["varnames"])]
    kw = dict(varnames=varnames, out_dir=out, batch=batch, precision=precision)
    if surface is not None:
        return randomise.tm_glm(data, exog, cov, str(v["gstat"]), v["optstfce"], perms, adjac_lh=v["adjac_lh"],
                                adjac_rh=v["adjac_rh"], mask_lh=v["mask_lh"], mask_rh=v["mask_rh"],
                                vdensity_lh=v["vdensity_lh"], vdensity_rh=v["vdensity_rh"], **kw)
    return randomise.tm_glm(data, exog, cov, str(v["gstat"]), v["optstfce"], perms, adjac=v["adjac"], **kw)


def tm_mediation(start, stop, surface=None, root=".", exchangeblock=None, seed=0, allow_pickle=False, batch=None,
                 precision="fp16x3"):
    """Permutation numbers start..stop (inclusive) of the tm-models mediation
    saved for ``surface`` (None: the volume); returns the null [n_perm]."""
    name = surface if surface is not None else "volume"
    tmp = os.path.join(root, "tmtemp_mediation_%s" % name)
    common = ["dmy_leftvar", "dmy_rightvar", "medtype", "data", "optstfce", "dmy_covariates"]
    if surface is not None:
        v = load_temp(tmp, common + ["mask_lh", "mask_rh", "adjac_lh", "adjac_rh", "vdensity_lh", "vdensity_rh"],
                      allow_pickle)
    else:
        v = load_temp(tmp, common + ["adjac"], allow_pickle)
    cov = v["dmy_covariates"]
    if cov.dtype == object and cov.ndim == 0 and cov.item() is None:
        cov = None  # np.save of None (tm_models_randomise.py:177-178)
    data = v["data"]
    blocks = None
    if exchangeblock is not None:
        blocks = np.genfromtxt(exchangeblock, dtype=str)
    perms = _perms(data.shape[0], range(int(start), int(stop) + 1), seed, blocks)
    out = _out_dir(root, "output_mediation_%s" % name, "perm_mediation")
    kw = dict(out_dir=out, batch=batch, precision=precision)
    args = (data, v["dmy_leftvar"], v["dmy_rightvar"], str(v["medtype"]), cov, v["optstfce"], perms)
    if surface is not None:
        return randomise.tm_mediation(*args, adjac_lh=v["adjac_lh"], adjac_rh=v["adjac_rh"], mask_lh=v["mask_lh"],
                                      mask_rh=v["mask_rh"], vdensity_lh=v["vdensity_lh"],
                                      vdensity_rh=v["vdensity_rh"], **kw)
    return randomise.tm_mediation(*args, adjac=v["adjac"], **kw)


def mmr_lr(surface, start, stop, out_path, seed, root=".", allow_pickle=False, batch=None, precision="fp16x3"):
    """Permutation numbers start..stop of surface ``surface`` of an mmr-lr
    analysis saved in ``root/tmi_temp`` (low_ram_calculate_tfce /
    low_ram_calculate_mediation_tfce per permutation, tm_func.py:144-185 and
    269-305, batched): permutation p is np.random.seed(p + seed) and a
    permutation of the subjects.  The saved options are a pickled argparse
    namespace and the adjacency a pickled list, so ``allow_pickle`` must be
    given for a folder the caller's own mmr-lr run wrote.  Returns the null
    {contrast: [n_perm, 2]} (t) or [n_perm] (mediation)."""
    import torch
    from .ols import PermutedOLS
    from .tfce import Adjacency, SurfaceSet, TFCE_ADJ_IMPLICIT_ZERO
    tmp = os.path.join(root, "tmi_temp")
    s = int(surface)
    v = load_temp(tmp, ["opts", "%d_adjacency_temp" % s, "%d_mask_temp" % s, "%d_data_temp" % s,
                        "%d_vdensity_temp" % s], allow_pickle)
    opts = v["opts"].tolist() if isinstance(v["opts"], np.ndarray) else v["opts"]
    if getattr(opts, "assigntfcesettings", None):
        ptr = int(opts.assigntfcesettings[s] * 2)
        H, E = float(opts.tfce[ptr]), float(opts.tfce[ptr + 1])
    else:
        H, E = float(opts.tfce[0]), float(opts.tfce[1])
    mask = np.asarray(v["%d_mask_temp" % s]).astype(bool)
    data = np.asarray(v["%d_data_temp" % s], dtype=np.float32)
    vd = np.broadcast_to(np.asarray(v["%d_vdensity_temp" % s], dtype=np.float32), (int(mask.sum()),))
    vd = torch.as_tensor(np.array(vd, dtype=np.float32)).cuda()  # a writable copy of the broadcast view
    n = data.shape[0]
    perms = _perms(n, range(int(start), int(stop) + 1), seed)
    csr, flags = randomise._hemi_csr(v["%d_adjacency_temp" % s], mask)

    def csv(path):
        return np.genfromtxt(os.path.join(root, path), delimiter=",")
    os.makedirs(out_path, exist_ok=True)
    if getattr(opts, "inputmediation", None):
        medtype = str(opts.inputmediation[0])
        # the z rescale takes the largest in-mask z only (tm_func.py:295-297)
        surfs = SurfaceSet([Adjacency(csr, H, E, flags & TFCE_ADJ_IMPLICIT_ZERO)], [0])
        med = randomise.MediationModel(data, csv(opts.inputmediation[1]), csv(opts.inputmediation[2]), medtype,
                                       precision=precision)
        null = []
        for _, pb in randomise._batches(perms, batch, 1):
            null.append(surfs.batch(med.z(pb), 1, vdensity=vd)[0, :, 0].cpu().numpy())
        null = np.concatenate(null)
        randomise._write(os.path.join(out_path, "perm_maxTFCE_surf%d_%s_zstat.csv" % (s, medtype)), null, "%f")
        return null
    cols = [csv(p) for p in opts.input]
    pred_x = np.column_stack(cols) if len(cols) > 1 else cols[0]
    X = np.column_stack([np.ones(n), pred_x])
    contrasts = list(range(1, X.shape[1]))
    surfs = SurfaceSet([Adjacency(csr, H, E, flags)], [0])
    model = PermutedOLS(data, X, contrasts=contrasts, precision=precision)
    nulls = {c: [] for c in contrasts}
    for _, pb in randomise._batches(perms, batch, 2):
        t = model.t(pb)
        for ci, c in enumerate(contrasts):
            nulls[c].append(surfs.batch(t[ci], 3, vdensity=vd)[0].cpu().numpy())
    nulls = {c: np.concatenate(x, 0) for c, x in nulls.items()}
    for c, x in nulls.items():
        randomise._write(os.path.join(out_path, "perm_maxTFCE_surf%d_tcon%d.csv" % (s, c)), x, "%f")
    return nulls
