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
* ``tm_mediation``     the mediation branch of tm_models_randomise.py:430-510
  (Sobel z of glm_typeI t-values with covariates, types I, M and Y,
  ``perm_Zstat_<type>`` files);
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
    x and y together (vertex_tfce_mediation_randomise.py:82-90).

    With ``covariates`` (columns of no interest, never permuted) every design
    gains them after its variables, as glm_typeI builds [1, exog...,
    covariates] (pyfunc.py:2301-2314), and each path's t is row 1 of that fit
    (tm_models_randomise.py:430-510); x and y may then hold several columns
    (dummy-coded factors), of which row 1 is the first.  Type Y's path A,
    y ~ [1, x, covariates] with x and y permuted together and the covariates
    fixed, then changes with the permutation: it is solved in FP64 on the
    host, one small fit per permutation."""

    def __init__(self, M, pred_x, depend_y, medtype="M", alg="aroian", precision="fp16x3", fused=True,
                 covariates=None):
        """fused (types M and I with a one-column x): both paths bind one
        prepared copy of M and share one contraction per batch
        (``tfce_ols_perm_sobel``); otherwise each path contracts on its own
        and ``tfce_sobel_z`` combines the t-values (same numbers, bit for
        bit)."""
        x = np.asarray(pred_x, dtype=np.float64)
        y = np.asarray(depend_y, dtype=np.float64)
        n = x.shape[0]
        x2, y2 = x.reshape(n, -1), y.reshape(n, -1)
        wx, wy = x2.shape[1], y2.shape[1]
        C = np.zeros((n, 0)) if covariates is None else np.asarray(covariates, dtype=np.float64).reshape(n, -1)
        self.medtype, self.alg = medtype, {"aroian": 0, "sobel": 1, "goodman": 2}[alg]
        one = np.ones((n, 1))
        self.ta_const = None
        self.ta_fit = None
        self.fused = False
        xcols = list(range(1, 1 + wx))
        if medtype in ("M", "I"):
            self.A = PermutedOLS(M, np.column_stack([one, x2, C]), contrasts=[1], perm_columns=xcols,
                                 precision=precision)
            if medtype == "M":
                XB, pc = np.column_stack([one, y2, x2, C]), [1 + wy + j for j in range(wx)]
            else:
                XB, pc = np.column_stack([one, x2, y2, C]), xcols
            if fused and wx == 1:
                self.B = PermutedOLS(None, XB, contrasts=[1], perm_columns=pc, data_from=self.A)
                self.fused = True
            else:
                self.B = PermutedOLS(M, XB, contrasts=[1], perm_columns=pc, precision=precision)
        elif medtype == "Y":
            if wy != 1:
                raise ValueError("mediation type Y needs a one-column dependent variable")
            self.n = n
            self.ta_fit = (np.column_stack([one, x2, C]), y2[:, 0], wx)
            if C.shape[1] == 0:
                # x and y move together: the x -> y fit (linregress's slope over
                # its standard error) is that of the saved rows
                self.ta_const = float(self.path_a_t(np.arange(n)[None])[0])
                self.ta_fit = None
            self.A = None
            self.B = PermutedOLS(M, np.column_stack([one, y2, x2, C]), contrasts=[1],
                                 perm_columns=list(range(1, 1 + wy + wx)), precision=precision)
        else:
            raise ValueError("Invalid mediation type")
        self.n = n
        self.V = self.B.V

    def path_a_t(self, perms):
        """FP64 [P] t of x in y[perm] ~ [1, x[perm], covariates] (type Y;
        perms [P, n] host indices): a / sqrt(sigma2 * inv(X'X)_11), the
        one-column se of glm_typeI (pyfunc.py:2363-2366)."""
        Xa, y, wx = self.ta_fit
        perms = np.asarray(perms, dtype=np.int64).reshape(-1, self.n)
        X = np.repeat(Xa[None], len(perms), 0)
        X[:, :, 1:1 + wx] = Xa[perms][:, :, 1:1 + wx]
        yp = y[perms]
        n, k = Xa.shape
        inv = np.linalg.inv(np.einsum("pik,pil->pkl", X, X))
        a = np.einsum("pkl,pil,pi->pk", inv, X, yp)
        sigma2 = np.sum((yp - np.einsum("pik,pk->pi", X, a)) ** 2, 1) / (n - k)
        return a[:, 1] / np.sqrt(sigma2 * inv[:, 1, 1])

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
            z = out if out is not None else torch.empty_like(tb)
            assert z.shape == tb.shape and z.is_contiguous()
            if self.ta_fit is not None:
                host = np.arange(self.n)[None] if perms is None else perms.cpu().numpy()
                for p, ta in enumerate(self.path_a_t(host)):
                    _lib.check(_lib.lib().tfce_sobel_z(None, _lib.ptr(tb[p]), self.V, self.alg, float(ta),
                                                       _lib.ptr(z[p]), _lib.stream_ptr(s)))
                return z
            ta = self.A.t(perms, stream=s)[0] if self.A is not None else None
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


class _TmSurfaces:
    """The TFCE side of the tm-models randomise loop: both hemispheres with
    masks and vdensity as write_perm_maxTFCE_vertex (pyfunc.py:107-119,
    files ``..._TFCE_maxVertex.csv`` in "%.4f"), or one voxel set as
    write_perm_maxTFCE_voxel (pyfunc.py:121-126, ``..._TFCE_maxVoxel.csv`` in
    "%1.4f", NaN and infinite statistics set to 0 where the reference's max
    would be nan).  ``data`` is moved into the sweep's element order."""

    def __init__(self, data, optstfce, adjac=None, adjac_lh=None, adjac_rh=None, mask_lh=None, mask_rh=None,
                 vdensity_lh=1, vdensity_rh=1):
        import torch
        H, E = float(optstfce[0]), float(optstfce[1])
        self.vd = None
        if adjac_lh is not None:
            (lh, fl), (rh, fr) = _hemi_csr(adjac_lh, mask_lh), _hemi_csr(adjac_rh, mask_rh)
            nv_lh = int(np.asarray(mask_lh, bool).sum())
            if not np.isscalar(vdensity_lh) or not np.isscalar(vdensity_rh):
                ml, mr = np.asarray(mask_lh, bool), np.asarray(mask_rh, bool)
                vl = np.broadcast_to(np.asarray(vdensity_lh, np.float32), (len(ml),))[ml]
                vr = np.broadcast_to(np.asarray(vdensity_rh, np.float32), (len(mr),))[mr]
                self.vd = torch.as_tensor(np.ascontiguousarray(np.concatenate([vl, vr]))).cuda()
            (lh, rh), self.data, self.vd = _in_layout([lh, rh], [0, nv_lh], data, self.vd)
            self.surfs = SurfaceSet([Adjacency(lh, H, E, fl), Adjacency(rh, H, E, fr)], [0, nv_lh])
            self.suffix, self.fmt, self.vertex = "maxVertex", "%.4f", True
        else:
            (csr,), self.data, _ = _in_layout([to_csr(adjac)], [0], data)
            self.surfs = SurfaceSet([Adjacency(csr, H, E)], [0])
            self.suffix, self.fmt, self.vertex = "maxVoxel", "%1.4f", False

    def maxima(self, stat, signs):
        """[P, n_signs] scaled maxima of the maps stat [P, V] (host)."""
        import torch
        if not self.vertex:
            stat = torch.where(torch.isfinite(stat), stat, torch.zeros_like(stat))
        m = self.surfs.batch(stat, signs, vdensity=self.vd)
        return (m.max(0).values if self.vertex else m[0]).cpu().numpy()


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
    tm = _TmSurfaces(data, optstfce, adjac, adjac_lh, adjac_rh, mask_lh, mask_rh, vdensity_lh, vdensity_rh)
    model = PermutedOLS(tm.data, X, contrasts=list(range(1, numcon + 1)), precision=precision)
    names_t = ["Tstat_con%d" % (j + 1) for j in range(numcon)] if want_t else []
    names_f = ["Fstat_%s" % varnames[b] for b in range(len(blocks))] if want_f else []
    nulls = {name: [] for name in names_t + names_f}
    maxima, suffix, fmt = tm.maxima, tm.suffix, tm.fmt
    items = 2 * (2 if tm.vertex else 1)
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


def tm_mediation(data, leftvar, rightvar, medtype, dmy_covariates, optstfce, perms, adjac=None, adjac_lh=None,
                 adjac_rh=None, mask_lh=None, mask_rh=None, vdensity_lh=1, vdensity_rh=1, out_dir=None,
                 batch=None, precision="fp16x3"):
    """The mediation branch of the tm-models randomise loop
    (tm_models_randomise.py:430-510): per permutation, path A is row 1 of
    glm_typeI's t for data ~ [1, left, covariates] and path B row 1 for
    [1, left, right, covariates] (type I) or [1, right, left, covariates]
    (type M), left permuted; type Y permutes left and right together, with
    path A the fit right ~ [1, left, covariates].  The Aroian z
    (calc_indirect, pyfunc.py:2678-2709) goes through TFCE, positive only,
    into ``perm_Zstat_<medtype>_TFCE_maxVertex.csv`` (both hemispheres) or
    ``..._maxVoxel.csv`` (see ``_TmSurfaces``).  The covariates never move.

    The reference permutes its already permuted variables again on every
    iteration (``dmy_leftvar = dmy_leftvar[rand_array]``), so iteration i
    sees the composition of the first i permutations; a composition of
    uniform permutations is uniform, so the null is the same, and here
    permutation p is applied to the saved variables, which keeps any range
    of permutation numbers reproducible on its own.  Returns the null
    [n_perm]."""
    import torch
    data = np.asarray(data) if not isinstance(data, torch.Tensor) else data
    medtype = str(medtype)
    tm = _TmSurfaces(data, optstfce, adjac, adjac_lh, adjac_rh, mask_lh, mask_rh, vdensity_lh, vdensity_rh)
    med = MediationModel(tm.data, leftvar, rightvar, medtype, precision=precision, covariates=dmy_covariates)
    out = []
    for _, pb in _batches(perms, batch, 2 if tm.vertex else 1):
        out.append(tm.maxima(med.z(pb), 1)[:, 0])
    null = np.concatenate(out)
    if out_dir is not None:
        _write(os.path.join(out_dir, "perm_Zstat_%s_TFCE_%s.csv" % (medtype, tm.suffix)), null, tm.fmt)
    return null


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
