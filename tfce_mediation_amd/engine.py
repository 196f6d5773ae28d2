"""Batched permutation engine: the randomise loop of the reference on one GPU.

One batch of P permutations is
  1. t-values of the contrast for every permutation  (PermutedOLS, MFMA),
  2. TFCE of +t and -t on every surface, rescaled as the reference does, and
     the per-surface maxima                           (SurfaceSet, one launch),
  3. the maximum over surfaces per sign               (tiny reduction).
This replaces the per-permutation body of
vertex_tfce_multiple_regression_randomise.py:90-117 (tval_int, then
write_perm_maxTFCE_vertex for +t and -t, pyfunc.py:107-119) and of
voxel_tfce_multiple_regression_randomise.py:95-118; the result rows are the
two null values the reference appends per permutation (+ first, then -).
"""
import numpy as np

from . import synth
from .ols import PermutedOLS
from .tfce import Adjacency, SurfaceSet, TFCE_ADJ_IMPLICIT_ZERO, TFCE_ADJ_ZERO_IN_SCALE


class PermutationEngine:
    def __init__(self, Y, X, surfaces, offsets, H=2.0, E=1.0, contrast=None, precision="bf16x3",
                 perm_columns=None, vdensity=None, implicit_zero=None, settings=None):
        """surfaces: list of CSR (row_ptr, col); offsets: first element of each
        surface in a map row; settings: optional per-surface (H, E)."""
        import torch
        X = np.asarray(X, dtype=np.float64)
        self.n, self.k = X.shape
        if contrast is None:
            contrast = self.k - 1
        self.contrast = int(contrast)
        self.ols = PermutedOLS(Y, X, contrasts=[self.contrast], perm_columns=perm_columns,
                               precision=precision)
        self.V = self.ols.V
        if implicit_zero is None:
            implicit_zero = [False] * len(surfaces)
        adjs = []
        for i, csr in enumerate(surfaces):
            h, e = settings[i] if settings is not None else (H, E)
            flags = (TFCE_ADJ_IMPLICIT_ZERO | TFCE_ADJ_ZERO_IN_SCALE) if implicit_zero[i] else 0
            adjs.append(Adjacency(csr, h, e, flags))
        self.surfaces = SurfaceSet(adjs, offsets)
        self.vdensity = None
        if vdensity is not None:
            self.vdensity = torch.as_tensor(np.asarray(vdensity, dtype=np.float32)).cuda()

    def tvalues(self, perms):
        return self.ols.t(perms)[0]

    def null_batch(self, perms, per_surface=False):
        """[P, 2] maxima (max over surfaces) of +t and -t, or [n_surf, P, 2]."""
        t = self.tvalues(perms)
        per = self.surfaces.batch(t, 3, vdensity=self.vdensity)
        return per if per_surface else per.max(0).values

    def run(self, n_perm, batch=256, seed=0, perms=None):
        """Null of n_perm permutations: [n_perm, 2] (host numpy)."""
        import torch
        out = []
        done = 0
        while done < n_perm:
            P = min(batch, n_perm - done)
            pb = perms[done:done + P] if perms is not None else synth.permutations(self.n, P, seed + done, "cuda")
            if not isinstance(pb, torch.Tensor):
                pb = torch.as_tensor(np.ascontiguousarray(pb, dtype=np.int32)).cuda()
            out.append(self.null_batch(pb).cpu().numpy())
            done += P
        return np.concatenate(out, 0)
