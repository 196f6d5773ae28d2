"""Permutation drivers: the step-2 randomise loops of the reference, batched.

Each driver produces the same max-TFCE null the reference's per-permutation
loop appends to its CSV files, for a whole range of permutations at once on
the GPU (permuted t on MFMA, one persistent TFCE launch per batch):

* ``vertex_regress``   vertex_tfce_multiple_regression_randomise.py:56-118
  (two hemispheres, bin masks, vdensity, every contrast j = 1..k-1, +t then -t,
  ``perm_tstat_con<j>_TFCE_maxVertex.csv`` in "%.4f");
* ``voxel_regress``    voxel_tfce_multiple_regression_randomise.py:51-120
  (``perm_tstat_con<j>_TFCE_maxVoxel.csv`` in "%1.4f");
* ``vertex_mediation`` vertex_tfce_mediation_randomise.py:48-92 (Sobel z,
  positive pass only, ``perm_Zstat_<type>_TFCE_maxVertex.csv``);
* ``voxel_mediation``  voxel_tfce_mediation_randomise.py:42-76 (Sobel z of a
  voxel mask, positive pass only, ``perm_Zstat_<type>_TFCE_maxVoxel.csv`` in
  "%1.4f");
* ``tm_glm``           the GLM branch of tm_models_randomise.py:182-243
  (glm_typeI statistics of the whole permuted design: t per exog column and
  F per exog variable, ``perm_Tstat_con<j>`` / ``perm_Fstat_<var>`` files);
* ``mmr``              tm_func.calculate_tfce in a loop (the mmr randomise of
  tm_multimodality_multisurface_regression.py:523-572): per-surface nulls
  ``perm_maxTFCE_surf<s>_tcon<c>.csv`` in "%f"; ``merged=True`` shares the
  threshold sequence over the surfaces of a settings group as the merged
  adjacency does, ``merged=False`` gives mmr-lr's per-surface thresholds.

Permutation indices: ``permutation_indices`` draws them reproducibly from
(seed, permutation number) with numpy's legacy RNG, the scheme of
tm_func.py:147-149 (the reference's step-2 scripts seed from the clock,
vertex_tfce_multiple_regression_randomise.py:91, which cannot be reproduced).
Exchange blocks (``-e``) use pyfunc.rand_blocks' rule.
"""
import os

import numpy as np

from . import synth
from .adjacency import induced_subgraph, locality_layout, to_csr
from .ols import PermutedOLS, total_ss
from .tfce import Adjacency, SurfaceSet, TFCE_ADJ_IMPLICIT_ZERO, TFCE_ADJ_ZERO_IN_SCALE, default_batch


def permutation_indices(n, perm_numbers, seed=0, blocks=None):
    """int32 [len(perm_numbers), n]: np.random.seed(p + seed) then a
    permutation (or the exchange-block permutation of pyfunc.py:2738-2757)."""
    from .pyfunc import check_blocks, rand_blocks
    out = np.zeros((len(perm_numbers), n), dtype=np.int32)
    equal = check_blocks(np.asarray(blocks)) if blocks is not None else None
    for i, p in enumerate(perm_numbers):
        np.random.seed(int(p) + int(seed))
        if blocks is None:
            out[i] = np.random.permutation(list(range(n)))
        else:
            out[i] = rand_blocks(np.asarray(blocks), equal)
    return out


def _hemi_csr(adjacency, bin_mask):
    """(CSR induced on the mask, adjacency flags) of one hemisphere."""
    rp, col = to_csr(adjacency)
    bin_mask = np.asarray(bin_mask, dtype=bool)
    flags = (TFCE_ADJ_IMPLICIT_ZERO | TFCE_ADJ_ZERO_IN_SCALE) if not bin_mask.all() else 0
    return induced_subgraph(rp, col, bin_mask), flags


def _hemi_surface(adjacency, bin_mask, H, E):
    csr, flags = _hemi_csr(adjacency, bin_mask)
    return Adjacency(csr, H, E, flags)


def _in_layout(csrs, offsets, Y, vd=None):
    """The surfaces, data columns and vdensity in the locality layout
    (adjacency.locality_layout): only the null maxima leave these drivers, and
    they do not depend on the element order."""
    import torch
    csrs2, cols = locality_layout(csrs, offsets, int(Y.shape[1]))
    if np.array_equal(cols, np.arange(len(cols))):
        return csrs, Y, vd
    if isinstance(Y, torch.Tensor):
        Y = Y.index_select(1, torch.as_tensor(cols, device=Y.device))
    else:
        Y = np.asarray(Y)[:, cols]
    if vd is not None:
        vd = vd[torch.as_tensor(cols, device=vd.device)]
    return csrs2, Y, vd


def _batches(perms, batch, items_per_perm):
    """(first, device batch) over perms; ``batch`` None: tfce.default_batch
    for ``items_per_perm`` (surface, sign) items of each TFCE launch."""
    import torch
    if batch is None:
        batch = default_batch(items_per_perm)
    for a in range(0, len(perms), batch):
        yield a, torch.as_tensor(np.ascontiguousarray(perms[a:a + batch], dtype=np.int32)).cuda()


def _write(path, values, fmt):
    with open(path, "a") as f:
        for v in np.asarray(values, dtype=np.float64).ravel():
            f.write((fmt % v) + "\n")


def vertex_regress(ny, pred_x, adjac_lh, adjac_rh, bin_mask_lh, bin_mask_rh, optstfce, perms,
                   vdensity_lh=1, vdensity_rh=1, specifyvars=None, out_dir=None, batch=None,
                   precision="fp16x3"):
    """Null rows [n_perm, 2] (+t, -t) per contrast, dict {tnum: array}; written
    to perm_tstat_con<tnum>_TFCE_maxVertex.csv when out_dir is given."""
    import torch
    n = ny.shape[0]
    X = np.column_stack([np.ones(n), pred_x])
    k = X.shape[1]
    if specifyvars is not None:
        start, stop = specifyvars[0], specifyvars[1] + 1
        perm_cols = list(range(start, stop))
        contrasts = [j + 1 for j in range(stop - start)]  # tnum = j + 1 (randomise.py:110-113)
    else:
        perm_cols = None
        contrasts = list(range(1, k))
    H, E = float(optstfce[0]), float(optstfce[1])
    (lh, fl), (rh, fr) = _hemi_csr(adjac_lh, bin_mask_lh), _hemi_csr(adjac_rh, bin_mask_rh)
    nv_lh = int(np.asarray(bin_mask_lh, bool).sum())
    vd = None
    if not np.isscalar(vdensity_lh) or not np.isscalar(vdensity_rh):
        vl = np.broadcast_to(np.asarray(vdensity_lh, np.float32), (len(bin_mask_lh),))[np.asarray(bin_mask_lh, bool)]
        vr = np.broadcast_to(np.asarray(vdensity_rh, np.float32), (len(bin_mask_rh),))[np.asarray(bin_mask_rh, bool)]
        vd = torch.as_tensor(np.ascontiguousarray(np.concatenate([vl, vr]))).cuda()
    (lh, rh), ny, vd = _in_layout([lh, rh], [0, nv_lh], ny, vd)
    surfs = SurfaceSet([Adjacency(lh, H, E, fl), Adjacency(rh, H, E, fr)], [0, nv_lh])
    model = PermutedOLS(ny, X, contrasts=contrasts, perm_columns=perm_cols, precision=precision)
    nulls = {c: [] for c in contrasts}
    for _, pb in _batches(perms, batch, 2 * 2):
        t = model.t(pb)
        for ci, c in enumerate(contrasts):
            nulls[c].append(surfs.batch(t[ci], 3, vdensity=vd).max(0).values.cpu().numpy())
    nulls = {c: np.concatenate(v, 0) for c, v in nulls.items()}
    if out_dir is not None:
        for c, v in nulls.items():
            _write(os.path.join(out_dir, "perm_tstat_con%d_TFCE_maxVertex.csv" % c), v, "%.4f")
    return nulls


def voxel_regress(ny, pred_x, adjac, optstfce, perms, specifyvars=None, out_dir=None, batch=None,
                  precision="fp16x3", ancova=False):
    """Voxel randomise: null rows [n_perm, 2] per contrast.  With ``ancova``
    (the -f branch, voxel_tfce_multiple_regression_randomise.py:79-89) the
    image is sqrt(max(F, 0)) of the whole model of X[perm] and the null is
    {"fstat": [n_perm]} (positive TFCE only), ``perm_fstat_TFCE_maxVoxel.csv``;
    the reference runs twice as many permutations in that branch (:80), which
    is the caller's choice of ``perms`` here."""
    n = ny.shape[0]
    X = np.column_stack([np.ones(n), pred_x])
    k = X.shape[1]
    if specifyvars is not None and not ancova:
        start, stop = specifyvars[0], specifyvars[1] + 1
        perm_cols, contrasts = list(range(start, stop)), [j + 1 for j in range(stop - start)]
    else:
        perm_cols, contrasts = None, list(range(1, k))
    (csr,), ny, _ = _in_layout([to_csr(adjac)], [0], ny)
    surfs = SurfaceSet([Adjacency(csr, float(optstfce[0]), float(optstfce[1]))], [0])
    model = PermutedOLS(ny, X, contrasts=contrasts, perm_columns=perm_cols, precision=precision)
    if ancova:
        tss = total_ss(ny)
        null = []
        for _, pb in _batches(perms, batch, 1):
            null.append(surfs.batch(model.fstat(pb, tss), 1)[0, :, 0].cpu().numpy())
        nulls = {"fstat": np.concatenate(null, 0)}
        if out_dir is not None:
            _write(os.path.join(out_dir, "perm_fstat_TFCE_maxVoxel.csv"), nulls["fstat"], "%1.4f")
        return nulls
    nulls = {c: [] for c in contrasts}
    for _, pb in _batches(perms, batch, 2):
        t = model.t(pb)
        t = t.nan_to_num(nan=0.0)  # perm_tvalues[np.isnan(perm_tvalues)] = 0 (randomise.py:109)
        for ci, c in enumerate(contrasts):
            nulls[c].append(surfs.batch(t[ci], 3)[0].cpu().numpy())
    nulls = {c: np.concatenate(v, 0) for c, v in nulls.items()}
    if out_dir is not None:
        for c, v in nulls.items():
            _write(os.path.join(out_dir, "perm_tstat_con%d_TFCE_maxVoxel.csv" % c), v, "%1.4f")
    return nulls


class MediationModel:
    """Batched permuted Sobel z (calc_sobelz, pyfunc.py:130-162) for one image
    set: path A is y_img ~ [1, x]; path B is y_img ~ [1, y, x] (type M, the
    mediator t of y), [1, x, y] (type I, t of x) or type Y with path A the
    scalar x -> y regression.  Types M and I permute x only, type Y permutes
    x and y together (vertex_tfce_mediation_randomise.py:82-90)."""

    def __init__(self, M, pred_x, depend_y, medtype="M", alg="aroian", precision="fp16x3", fused=True):
        """fused (types M and I): both paths bind one prepared copy of M and
        share one contraction per batch (``tfce_ols_perm_sobel``); otherwise
        each path contracts on its own and ``tfce_sobel_z`` combines the
        t-values (same numbers, bit for bit)."""
        import torch
        x = np.asarray(pred_x, dtype=np.float64)
        y = np.asarray(depend_y, dtype=np.float64)
        n = x.shape[0]
        self.medtype, self.alg = medtype, {"aroian": 0, "sobel": 1, "goodman": 2}[alg]
        one = np.ones(n)
        self.ta_const = None
        self.fused = False
        if medtype in ("M", "I"):
            self.A = PermutedOLS(M, np.column_stack([one, x]), contrasts=[1], perm_columns=[1], precision=precision)
            if medtype == "M":
                XB, cb, pc = np.column_stack([one, y, x]), 1, [2]
            else:
                XB, cb, pc = np.column_stack([one, x, y]), 1, [1]
            if fused:
                self.B = PermutedOLS(None, XB, contrasts=[cb], perm_columns=pc, data_from=self.A)
                self.fused = True
            else:
                self.B = PermutedOLS(M, XB, contrasts=[cb], perm_columns=pc, precision=precision)
        elif medtype == "Y":
            from .cynumstats import tval_int
            Xa = np.column_stack([one, x])
            self.ta_const = float(tval_int(Xa, np.linalg.inv(Xa.T @ Xa), y[:, None].astype(np.float32), n, 2, 1)[1, 0])
            self.A = None
            self.B = PermutedOLS(M, np.column_stack([one, y, x]), contrasts=[1], perm_columns=[1, 2],
                                 precision=precision)
        else:
            raise ValueError("Invalid mediation type")
        self.V = self.B.V

    def z(self, perms=None, out=None, stream=None):
        """[P, V] z on ``stream`` (default: the current stream), into ``out`` if given."""
        import torch
        from . import _lib
        s = stream if stream is not None else torch.cuda.current_stream()
        if self.fused:
            if perms is None:
                P = 1
            else:
                assert perms.is_cuda and perms.dtype == torch.int32 and perms.is_contiguous()
                P = int(perms.shape[0])
            with torch.cuda.stream(s):
                z = out if out is not None else torch.empty((P, self.V), dtype=torch.float32, device="cuda")
                assert z.shape == (P, self.V) and z.is_contiguous()
                _lib.check(_lib.lib().tfce_ols_perm_sobel(self.A._model, self.B._model, _lib.ptr(perms), P,
                                                          self.alg, _lib.ptr(z), self.V, _lib.stream_ptr(s)))
            return z
        with torch.cuda.stream(s):
            tb = self.B.t(perms, stream=s)[0]
            ta = self.A.t(perms, stream=s)[0] if self.A is not None else None
            z = out if out is not None else torch.empty_like(tb)
            assert z.shape == tb.shape and z.is_contiguous()
            _lib.check(_lib.lib().tfce_sobel_z(_lib.ptr(ta), _lib.ptr(tb), tb.numel(), self.alg,
                                               float(self.ta_const or 0.0), _lib.ptr(z), _lib.stream_ptr(s)))
        return z


def vertex_mediation(y, pred_x, depend_y, medtype, adjac_lh, adjac_rh, bin_mask_lh, bin_mask_rh, optstfce, perms,
                     vdensity_lh=1, vdensity_rh=1, out_dir=None, batch=None, precision="fp16x3"):
    """Mediation randomise: null [n_perm] of the positive Sobel-z TFCE maxima."""
    import torch
    H, E = float(optstfce[0]), float(optstfce[1])
    (lh, fl), (rh, fr) = _hemi_csr(adjac_lh, bin_mask_lh), _hemi_csr(adjac_rh, bin_mask_rh)
    nv_lh = int(np.asarray(bin_mask_lh, bool).sum())
    vd = None
    if not np.isscalar(vdensity_lh) or not np.isscalar(vdensity_rh):
        vl = np.broadcast_to(np.asarray(vdensity_lh, np.float32), (len(bin_mask_lh),))[np.asarray(bin_mask_lh, bool)]
        vr = np.broadcast_to(np.asarray(vdensity_rh, np.float32), (len(bin_mask_rh),))[np.asarray(bin_mask_rh, bool)]
        vd = torch.as_tensor(np.ascontiguousarray(np.concatenate([vl, vr]))).cuda()
    (lh, rh), y, vd = _in_layout([lh, rh], [0, nv_lh], y, vd)
    surfs = SurfaceSet([Adjacency(lh, H, E, fl), Adjacency(rh, H, E, fr)], [0, nv_lh])
    med = MediationModel(y, pred_x, depend_y, medtype, precision=precision)
    out = []
    for _, pb in _batches(perms, batch, 2):
        z = med.z(pb)
        out.append(surfs.batch(z, 1, vdensity=vd).max(0).values[:, 0].cpu().numpy())
    null = np.concatenate(out)
    if out_dir is not None:
        _write(os.path.join(out_dir, "perm_Zstat_%s_TFCE_maxVertex.csv" % medtype), null, "%.4f")
    return null


def voxel_mediation(ny, pred_x, depend_y, medtype, adjac, optstfce, perms, out_dir=None, batch=None,
                    precision="fp16x3"):
    """Voxel mediation randomise: null [n_perm] of max(TFCE(+z)) * max(z)/100
    (write_perm_maxTFCE_voxel, pyfunc.py:121-126) for the Sobel z of every
    permutation (x permuted for types M and I, x and y together for type Y,
    voxel_tfce_mediation_randomise.py:63-75); ny is [n, V]."""
    (csr,), ny, _ = _in_layout([to_csr(adjac)], [0], ny)
    surfs = SurfaceSet([Adjacency(csr, float(optstfce[0]), float(optstfce[1]))], [0])
    med = MediationModel(ny, pred_x, depend_y, medtype, precision=precision)
    out = []
    for _, pb in _batches(perms, batch, 1):
        out.append(surfs.batch(med.z(pb), 1)[0, :, 0].cpu().numpy())
    null = np.concatenate(out)
    if out_dir is not None:
        _write(os.path.join(out_dir, "perm_Zstat_%s_TFCE_maxVoxel.csv" % medtype), null, "%1.4f")
    return null


def glm_design(n, exog, dmy_covariates=None):
    """(X, column blocks) of glm_typeI (pyfunc.py:2298-2311): [1, exog
    variables (one or more columns each), covariates]; block i = (first
    column, width) of exog variable i."""
    cols, blocks, start = [np.ones(n)], [], 1
    for var in exog:
        var = np.asarray(var, dtype=np.float64)
        w = 1 if var.ndim == 1 else var.shape[1]
        cols.append(var.reshape(n, w))
        blocks.append((start, w))
        start += w
    if dmy_covariates is not None:
        cols.append(np.asarray(dmy_covariates, dtype=np.float64).reshape(n, -1))
    return np.column_stack(cols), blocks


def tm_glm(data, exog, dmy_covariates, gstat, optstfce, perms, adjac=None, adjac_lh=None, adjac_rh=None,
           mask_lh=None, mask_rh=None, vdensity_lh=1, vdensity_rh=1, varnames=None, out_dir=None, batch=None,
           precision="fp16x3"):
    """The GLM branch of the tm-models randomise loop
    (tm_models_randomise.py:182-243): per permutation the rows of the whole
    design [1, exog, covariates] are permuted (rand_array) and glm_typeI's
    statistics of ``data`` [n, V] go through TFCE.  gstat "t": the t of every
    exog column j = 1..numcon, +t then -t, ``perm_Tstat_con<j>_TFCE_max*``;
    "f": the F of every exog variable (its block of columns against the fit
    without it), positive only, ``perm_Fstat_<varname>_TFCE_max*``; anything
    else: both.  Vertex mode (adjac_lh given): both hemispheres as
    write_perm_maxTFCE_vertex, files ``..._TFCE_maxVertex.csv`` in "%.4f";
    voxel mode (adjac): ``..._TFCE_maxVoxel.csv`` in "%1.4f", with NaN and
    infinite statistics set to 0 (the reference's max would be nan there).
    Returns {statname: null}, [n_perm, 2] for t and [n_perm] for F."""
    import torch
    data = np.asarray(data) if not isinstance(data, torch.Tensor) else data
    n = int(data.shape[0])
    X, blocks = glm_design(n, exog, dmy_covariates)
    numcon = sum(w for _, w in blocks)
    if varnames is None:
        varnames = ["Exog%d" % (i + 1) for i in range(len(blocks))]
    gstat = str(gstat)
    want_t, want_f = gstat != "f", gstat != "t"
    H, E = float(optstfce[0]), float(optstfce[1])
    vd = None
    if adjac_lh is not None:
        (lh, fl), (rh, fr) = _hemi_csr(adjac_lh, mask_lh), _hemi_csr(adjac_rh, mask_rh)
        nv_lh = int(np.asarray(mask_lh, bool).sum())
        if not np.isscalar(vdensity_lh) or not np.isscalar(vdensity_rh):
            vl = np.broadcast_to(np.asarray(vdensity_lh, np.float32), (len(mask_lh),))[np.asarray(mask_lh, bool)]
            vr = np.broadcast_to(np.asarray(vdensity_rh, np.float32), (len(mask_rh),))[np.asarray(mask_rh, bool)]
            vd = torch.as_tensor(np.ascontiguousarray(np.concatenate([vl, vr]))).cuda()
        (lh, rh), data, vd = _in_layout([lh, rh], [0, nv_lh], data, vd)
        surfs = SurfaceSet([Adjacency(lh, H, E, fl), Adjacency(rh, H, E, fr)], [0, nv_lh])
        suffix, fmt, vertex = "maxVertex", "%.4f", True
    else:
        (csr,), data, _ = _in_layout([to_csr(adjac)], [0], data)
        surfs = SurfaceSet([Adjacency(csr, H, E)], [0])
        suffix, fmt, vertex = "maxVoxel", "%1.4f", False
    model = PermutedOLS(data, X, contrasts=list(range(1, numcon + 1)), precision=precision)
    names_t = ["Tstat_con%d" % (j + 1) for j in range(numcon)] if want_t else []
    names_f = ["Fstat_%s" % varnames[b] for b in range(len(blocks))] if want_f else []
    nulls = {name: [] for name in names_t + names_f}

    def maxima(stat, signs):
        if not vertex:
            stat = torch.where(torch.isfinite(stat), stat, torch.zeros_like(stat))
        m = surfs.batch(stat, signs, vdensity=vd)
        return (m.max(0).values if vertex else m[0]).cpu().numpy()

    items = 2 * (2 if vertex else 1)
    for _, pb in _batches(perms, batch, items):
        if want_t:
            t = model.t(pb)
            for j, name in enumerate(names_t):
                nulls[name].append(maxima(t[j], 3))
        if want_f:
            f = model.partial_f(pb, blocks)
            for b, name in enumerate(names_f):
                nulls[name].append(maxima(f[b], 1)[:, 0])
    nulls = {name: np.concatenate(v, 0) for name, v in nulls.items()}
    if out_dir is not None:
        for name, v in nulls.items():
            _write(os.path.join(out_dir, "perm_%s_TFCE_%s.csv" % (name, suffix)), v, fmt)
    return nulls


def mmr(merge_y, pred_x, surfaces, offsets, settings, perms, groups=None, merged=True, zero_in_image=None,
        vdensity=None, out_dir=None, batch=None, precision="fp16x3"):
    """Multi-surface randomise.  surfaces: CSR per surface (in-mask elements),
    offsets: their first column in merge_y; settings: (H, E) per surface;
    groups: settings-group id per surface (default: one per distinct setting).
    Returns {(surface, tcon): [n_perm, 2]} in the order the reference appends
    (+ then -) and writes perm_maxTFCE_surf<s>_tcon<c>.csv when out_dir is set."""
    import torch
    n = merge_y.shape[0]
    X = np.column_stack([np.ones(n), pred_x])
    k = X.shape[1]
    contrasts = list(range(1, k))
    vd = None
    if vdensity is not None and not np.isscalar(vdensity):
        vd = torch.as_tensor(np.ascontiguousarray(vdensity, dtype=np.float32)).cuda()
    surfaces, merge_y, vd = _in_layout([to_csr(c) for c in surfaces], offsets, merge_y, vd)
    model = PermutedOLS(merge_y, X, contrasts=contrasts, precision=precision)
    if groups is None:
        keys = sorted(set(map(tuple, settings)))
        groups = [keys.index(tuple(s)) for s in settings]
    sets = {}
    for g in sorted(set(groups)):
        idx = [i for i, gg in enumerate(groups) if gg == g]
        adjs = [Adjacency(surfaces[i], settings[i][0], settings[i][1]) for i in idx]
        sets[g] = (idx, SurfaceSet(adjs, [offsets[i] for i in idx]))
    zi = zero_in_image or {}
    res = {}
    for _, pb in _batches(perms, batch, 2 * min(len(idx) for idx, _ in sets.values())):
        t = model.t(pb)
        for ci, c in enumerate(contrasts):
            for g, (idx, ss) in sets.items():
                per = ss.batch(t[ci], 3, vdensity=vd, shared_thresholds=merged,
                               zero_in_image=bool(zi.get(g, False)) if merged else False)
                per = per.cpu().numpy()
                for j, s in enumerate(idx):
                    res.setdefault((s, c), []).append(per[j])
    res = {key: np.concatenate(v, 0) for key, v in res.items()}
    if out_dir is not None:
        for (s, c), v in sorted(res.items()):
            _write(os.path.join(out_dir, "perm_maxTFCE_surf%d_tcon%d.csv" % (s, c)), v, "%f")
    return res
