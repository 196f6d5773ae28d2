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
from .adjacency import locality_layout
from .tfce import Adjacency, SurfaceSet, TFCE_ADJ_IMPLICIT_ZERO, TFCE_ADJ_ZERO_IN_SCALE, default_batch


class PermutationEngine:
    def __init__(self, Y, X, surfaces, offsets, H=2.0, E=1.0, contrast=None, precision="fp16x3",
                 perm_columns=None, vdensity=None, implicit_zero=None, settings=None, locality=True):
        """surfaces: list of CSR (row_ptr, col); offsets: first element of each
        surface in a map row; settings: optional per-surface (H, E).
        locality: hold the elements of every symmetric surface in the
        library's locality order (Y's columns are permuted once here); the
        null maxima do not depend on the element order."""
        import torch
        X = np.asarray(X, dtype=np.float64)
        self.cols = None
        if locality:
            width = int(Y.shape[1])
            surfaces, cols = locality_layout(surfaces, offsets, width)
            if not np.array_equal(cols, np.arange(width)):
                self.cols = cols
                if isinstance(Y, torch.Tensor):
                    Y = Y.index_select(1, torch.as_tensor(cols, device=Y.device))
                else:
                    Y = np.asarray(Y)[:, cols]
                if vdensity is not None:
                    vdensity = np.asarray(vdensity, dtype=np.float32)[cols]
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
        self.csrs = surfaces  # in the engine's element layout
        self.surfaces = SurfaceSet(adjs, offsets)
        self.vdensity = None
        if vdensity is not None:
            self.vdensity = torch.as_tensor(np.asarray(vdensity, dtype=np.float32)).cuda()

    def tvalues(self, perms):
        """[P, width] t-values in the engine's element layout (column j is
        map column cols[j] when a locality layout is in use)."""
        return self.ols.t(perms)[0]

    def rebind(self, Y, X, contrast=None, perm_columns=None):
        """New data Y [n, width] (map column order) and design X on the same
        surfaces: the adjacencies, their tables and the locality layout are
        kept, only the prepared data and the design model are rebuilt."""
        import torch
        X = np.asarray(X, dtype=np.float64)
        if self.cols is not None:
            if isinstance(Y, torch.Tensor):
                Y = Y.index_select(1, torch.as_tensor(self.cols, device=Y.device))
            else:
                Y = np.asarray(Y)[:, self.cols]
        self.n, self.k = X.shape
        self.contrast = int(self.k - 1 if contrast is None else contrast)
        precision = self.ols.precision
        self.ols = None
        self.ols = PermutedOLS(Y, X, contrasts=[self.contrast], perm_columns=perm_columns, precision=precision)
        assert self.ols.V == self.V

    def null_batch(self, perms, per_surface=False):
        """[P, 2] maxima (max over surfaces) of +t and -t, or [n_surf, P, 2]."""
        t = self.tvalues(perms)
        per = self.surfaces.batch(t, 3, vdensity=self.vdensity)
        return per if per_surface else per.max(0).values

    sign_mask = 3  # +t and -t

    def _contract(self, perms, out, stream):
        self.ols.t(perms, out=out[None], stream=stream)

    def pipeline(self, perm_batches, timing=None):
        """Null maxima of several batches, [P, n_signs] device tensors in batch
        order, with the contraction of batch b+1 (on a second stream) running
        while the sweep of batch b runs on the current stream.  Two map
        buffers alternate; the contraction into a buffer waits for the sweep
        that last read it.  Same values as null_batch, batch by batch.  With
        ``timing`` (a dict), HIP events around every contraction (on its
        stream) and every sweep are appended to timing["contraction"] and
        timing["sweep"] as (start, end) pairs."""
        import torch
        main = torch.cuda.current_stream()
        side = getattr(self, "_side", None)
        if side is None:
            side = self._side = torch.cuda.Stream()
        P = max(int(p.shape[0]) for p in perm_batches)
        bufs = [torch.empty((P, self.V), dtype=torch.float32, device="cuda") for _ in range(2)]
        swept = [None, None]
        side.wait_stream(main)  # the permutation tensors and buffers exist

        def event(stream):
            e = torch.cuda.Event(enable_timing=timing is not None)
            e.record(stream)
            return e

        def contract(b):
            pb = perm_batches[b]
            slot = b % 2
            out = bufs[slot][:pb.shape[0]]
            if swept[slot] is not None:
                side.wait_event(swept[slot])
            e0 = event(side) if timing is not None else None
            self._contract(pb, out, side)
            ready = event(side)
            if timing is not None:
                timing.setdefault("contraction", []).append((e0, ready))
            return out, ready

        outs = []
        nxt = contract(0) if perm_batches else None
        for b in range(len(perm_batches)):
            t, ready = nxt
            main.wait_event(ready)
            if b + 1 < len(perm_batches):
                nxt = contract(b + 1)
            s0 = event(main) if timing is not None else None
            per = self.surfaces.batch(t, self.sign_mask, vdensity=self.vdensity, stream=main)
            done = event(main)
            swept[b % 2] = done
            if timing is not None:
                timing.setdefault("sweep", []).append((s0, done))
            outs.append(per.max(0).values)
        main.wait_stream(side)
        return outs

    def graph(self, P):
        """One batch of P permutations captured as a HIP graph.  Returns
        run(perms) -> [P, 2] device tensor (the graph's own output buffer,
        overwritten by the next replay).  The library's scratch is grown on
        the capture stream by two eager warm-up batches first, so nothing is
        allocated while capturing; every launch of the batch is a kernel (no
        memset nodes).  The graph holds the capture stream's scratch
        (``run.stream``): later calls that grow it leave the captured buffers
        allocated, and replays must not overlap other work on that stream's
        scratch (replay on the current stream, in order)."""
        import torch
        s = torch.cuda.Stream()
        static_perm = torch.arange(self.n, dtype=torch.int32, device="cuda").repeat(P, 1).contiguous()
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            for _ in range(2):
                self.null_batch(static_perm)
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            static_out = self.null_batch(static_perm)
        torch.cuda.synchronize()

        def run(perms):
            static_perm.copy_(perms)
            g.replay()
            return static_out
        run.graph = g
        run.stream = s
        return run

    def default_batch(self, device=None):
        """Permutations per batch (``tfce.default_batch`` for this engine's
        surfaces and signs)."""
        return default_batch(len(self.csrs) * (2 if self.sign_mask == 3 else 1), device)

    def run(self, n_perm, batch=None, seed=0, perms=None, pipelined=False):
        """Null of n_perm permutations: [n_perm, n_signs] (host numpy), batch
        after batch (``batch`` permutations each, default ``default_batch()``)
        on the current stream, or through ``pipeline``.  Measured
        on MI355X at 327,684 vertices the one-stream order is the faster one:
        a contraction launched beside the sweep takes compute units the
        sweep's workgroups would occupy and delays them by more than the
        contraction takes alone."""
        import torch
        if batch is None:
            batch = self.default_batch()
        batches = []
        done = 0
        while done < n_perm:
            P = min(batch, n_perm - done)
            pb = perms[done:done + P] if perms is not None else synth.permutations(self.n, P, seed + done, "cuda")
            if not isinstance(pb, torch.Tensor):
                pb = torch.as_tensor(np.ascontiguousarray(pb, dtype=np.int32)).cuda()
            batches.append(pb.contiguous())
            done += P
        if not batches:
            return np.zeros((0, 2 if self.sign_mask == 3 else 1), dtype=np.float32)
        if pipelined:
            return np.concatenate([o.cpu().numpy() for o in self.pipeline(batches)], 0)
        return np.concatenate([self.null_batch(pb).cpu().numpy() for pb in batches], 0)


class MediationEngine:
    """Batched mediation randomise (vertex_tfce_mediation_randomise.py:82-91):
    per permutation of x, the Sobel-family z of the image mediator model
    (randomise.MediationModel), then TFCE of +z only on every surface and the
    maximum over surfaces.  Same surface layout and graph capture as
    PermutationEngine."""

    def __init__(self, M, x, y, surfaces, offsets, H=2.0, E=1.0, medtype="M", alg="aroian",
                 precision="fp16x3", settings=None, locality=True):
        import torch
        from .randomise import MediationModel
        self.n = len(x)
        self.cols = None
        if locality:
            width = int(M.shape[1])
            surfaces, cols = locality_layout(surfaces, offsets, width)
            if not np.array_equal(cols, np.arange(width)):
                self.cols = cols
                if isinstance(M, torch.Tensor):
                    M = M.index_select(1, torch.as_tensor(cols, device=M.device))
                else:
                    M = np.asarray(M)[:, cols]
        self.model = MediationModel(M, x, y, medtype, alg=alg, precision=precision)
        self.ols = self.model.B
        self.k = 3
        self.V = self.model.V
        self.vdensity = None
        self.csrs = surfaces
        adjs = [Adjacency(csr, *(settings[i] if settings is not None else (H, E))) for i, csr in enumerate(surfaces)]
        self.surfaces = SurfaceSet(adjs, offsets)

    sign_mask = 1  # +z only

    def tvalues(self, perms):
        """[P, width] Sobel z in the engine's element layout."""
        return self.model.z(perms)

    def _contract(self, perms, out, stream):
        self.model.z(perms, out=out, stream=stream)

    def null_batch(self, perms, per_surface=False):
        """[P, 1] maxima of +z (max over surfaces), or [n_surf, P, 1]."""
        per = self.surfaces.batch(self.tvalues(perms), 1)
        return per if per_surface else per.max(0).values

    graph = PermutationEngine.graph
    default_batch = PermutationEngine.default_batch
    run = PermutationEngine.run
    pipeline = PermutationEngine.pipeline
