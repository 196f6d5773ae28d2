"""Learned early stop of the permutation loop.

The reference always runs the requested number of permutations and reads the
FWER critical value off the sorted null (tools/plot_permutations.py:71-85:
``sorted(null, reverse=True)[int(N * alpha)]``).  Here a small tail model
predicts the 95% and 99% critical values from the first ``m0`` null values, and
the loop may stop early only when that prediction agrees with the running
empirical critical values within ``delta`` (relative) for both levels; otherwise
it runs to ``n_max`` values and reports the plain empirical result.  The plain
run is the fallback and the measure of the model's error.

* features: 32 quantiles of log(null[:m0]) at fixed probabilities, centred and
  scaled by that sample's mean / std of log values;
* targets : the critical values of the long run, log-transformed with the same
  centring and scaling;
* model   : 32 -> 64 -> 64 -> 2 tanh MLP, FP32, trained on the GPU by the
  library's own kernels (``tfce_mlp_train``, deterministic Adam steps) and
  evaluated by ``tfce_mlp_forward``.
Training nulls come from the permutation engine itself (``generate_nulls``).
"""

import numpy as np

from . import _lib

N_FEAT = 32
PROBS = 1.0 - np.geomspace(0.97, 0.004, N_FEAT)  # denser in the upper tail
ALPHAS = (0.05, 0.01)


def critical_value(null, alpha):
    s = np.sort(np.asarray(null, dtype=np.float64))[::-1]
    return s[int(len(s) * alpha)]


def features(first_values):
    z = np.log(np.maximum(np.asarray(first_values, dtype=np.float64), 1e-12))
    mu, sd = z.mean(), max(z.std(), 1e-12)
    q = np.quantile(z, PROBS)
    return ((q - mu) / sd).astype(np.float32), mu, sd


def targets(full_null, mu, sd):
    return np.array([(np.log(critical_value(full_null, a)) - mu) / sd for a in ALPHAS], dtype=np.float32)


def make_examples(nulls, m0=500, per_null=64, seed=0, pairs=True):
    """Training examples from long nulls: random subsets of m0 values (whole
    +/- pairs when ``pairs``) as inputs, the long null's critical values as targets."""
    rng = np.random.default_rng(seed)
    X, Y = [], []
    for null in nulls:
        null = np.asarray(null, dtype=np.float64).ravel()
        for r in range(per_null):
            if r == 0:
                sub = null[:m0]
            elif pairs:
                rows = rng.choice(len(null) // 2, m0 // 2, replace=False)
                sub = np.stack([null[2 * rows], null[2 * rows + 1]], 1).ravel()
            else:
                sub = null[rng.choice(len(null), m0, replace=False)]
            f, mu, sd = features(sub)
            X.append(f)
            Y.append(targets(null, mu, sd))
    return np.stack(X), np.stack(Y)


class TailMLP:
    def __init__(self, seed=0, params=None):
        import torch
        lib = _lib.lib()
        self.n_params = lib.tfce_mlp_param_count()
        if params is None:
            rng = np.random.default_rng(seed)
            p = []
            for fan_in, shape in ((32, (64, 32)), (None, (64,)), (64, (64, 64)), (None, (64,)), (64, (2, 64)),
                                  (None, (2,))):
                if fan_in is None:
                    p.append(np.zeros(shape))
                else:
                    p.append(rng.standard_normal(shape) * np.sqrt(1.0 / fan_in))
            params = np.concatenate([a.ravel() for a in p])
        params = np.asarray(params, dtype=np.float32)
        assert params.size == self.n_params
        self.params = torch.as_tensor(params.copy()).cuda()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.t = 0

    def numpy_params(self):
        return self.params.cpu().numpy()

    def predict(self, X):
        import torch
        Xd = torch.as_tensor(np.ascontiguousarray(X, dtype=np.float32)).cuda()
        Y = torch.empty((Xd.shape[0], 2), dtype=torch.float32, device="cuda")
        _lib.check(_lib.lib().tfce_mlp_forward(_lib.ptr(self.params), _lib.ptr(Xd), Xd.shape[0], _lib.ptr(Y),
                                               _lib.stream_ptr()))
        return Y.cpu().numpy()

    def train(self, X, Y, steps=3000, batch=128, lr=2e-3, weight_decay=1e-5, seed=0, want_grad0=False,
              chunk=500, log=None):
        """``steps`` Adam steps on minibatches drawn from (X, Y) with ``seed``,
        in launches of ``chunk`` steps (the kernel continues the Adam step
        count, so the parameters equal those of one launch bit for bit);
        ``log`` receives a progress line per launch.  Returns the per-step
        losses (and the first step's gradient and the minibatch indices with
        ``want_grad0``)."""
        import torch
        rng = np.random.default_rng(seed)
        Xd = torch.as_tensor(np.ascontiguousarray(X, dtype=np.float32)).cuda()
        Yd = torch.as_tensor(np.ascontiguousarray(Y, dtype=np.float32)).cuda()
        n = Xd.shape[0]
        b = min(batch, n)
        idx = np.stack([rng.choice(n, b, replace=False) for _ in range(steps)]).astype(np.int32)
        idx_d = torch.as_tensor(idx).cuda()
        loss = torch.empty(steps, dtype=torch.float32, device="cuda")
        g0 = torch.empty(self.n_params, dtype=torch.float32, device="cuda") if want_grad0 else None
        lib = _lib.lib()
        chunk = max(1, int(chunk))
        for a in range(0, steps, chunk):
            c = min(chunk, steps - a)
            _lib.check(lib.tfce_mlp_train(_lib.ptr(self.params), _lib.ptr(self.m), _lib.ptr(self.v), _lib.ptr(Xd),
                                          _lib.ptr(Yd), n, _lib.ptr(idx_d[a:]), c, b, self.t, float(lr),
                                          float(weight_decay), _lib.ptr(loss[a:]),
                                          _lib.ptr(g0) if a == 0 else None, _lib.stream_ptr()))
            self.t += c
            if log is not None:
                torch.cuda.synchronize()
                log(f"[train] {a + c}/{steps} Adam steps, minibatch loss {float(loss[a + c - 1]):.5f}")
        out = loss.cpu().numpy()
        return (out, g0.cpu().numpy(), idx) if want_grad0 else out

    def predict_critical(self, first_values):
        f, mu, sd = features(first_values)
        y = self.predict(f[None, :])[0]
        return np.exp(mu + sd * y.astype(np.float64))


class LearnedStop:
    """Stopping rule over a stream of null values (pooled +/- maxima in
    permutation order).

    The model predicts the critical values once, from the first m0 values.
    From m_min values on, every check_every values, the prediction is compared
    with the empirical critical values of the values that came after the first
    m0 (data the prediction has not seen); the loop stops when both levels
    agree within delta (relative) at `consecutive` checks in a row, and then
    reports the prediction.  Otherwise the plain empirical result of all n_max
    values is reported."""

    def __init__(self, model, m0=500, n_max=10000, delta=0.0075, m_min=1000, consecutive=4, check_every=512):
        self.model, self.m0, self.n_max, self.delta = model, m0, n_max, delta
        self.m_min, self.consecutive, self.check_every = m_min, consecutive, check_every

    def agrees(self, pred, values):
        values = np.asarray(values, dtype=np.float64).ravel()
        emp = np.array([critical_value(values[self.m0:], a) for a in ALPHAS])
        return bool(np.all(np.abs(pred - emp) <= self.delta * emp))

    def plain(self, values):
        return np.array([critical_value(np.asarray(values, dtype=np.float64)[:self.n_max], a) for a in ALPHAS])

    def simulate(self, stream):
        """Replay a recorded null stream: (values used, thresholds, source)."""
        stream = np.asarray(stream, dtype=np.float64).ravel()
        n_max = min(self.n_max, len(stream))
        pred = self.model.predict_critical(stream[:self.m0])
        ok, m = 0, max(self.m_min, self.m0 + 2)
        while m < n_max:
            ok = ok + 1 if self.agrees(pred, stream[:m]) else 0
            if ok >= self.consecutive:
                return m, pred, "learned"
            m += self.check_every
        return n_max, self.plain(stream[:n_max]), "plain"

    def run(self, engine, batch=256, seed=0):
        """Live loop on an engine (null_batch): batches until the rule stops;
        returns (values used, thresholds, source)."""
        from . import synth
        values, done, pred, ok, next_check = [], 0, None, 0, max(self.m_min, self.m0 + 2)
        while len(values) < self.n_max:
            pb = synth.permutations(engine.n, batch, seed + done, "cuda")
            values.extend(engine.null_batch(pb).cpu().numpy().ravel().tolist())
            done += batch
            if pred is None and len(values) >= self.m0:
                pred = self.model.predict_critical(values[:self.m0])
            while pred is not None and next_check <= min(len(values), self.n_max - 1):
                ok = ok + 1 if self.agrees(pred, values[:next_check]) else 0
                if ok >= self.consecutive:
                    return next_check, pred, "learned"
                next_check += self.check_every
        return self.n_max, self.plain(values), "plain"


def generate_nulls(make_config, seeds, n_values, batch=256, precision="fp16x3"):
    """Long max-TFCE nulls (pooled +/- values in permutation order) from the
    engine, one per seed; make_config(seed) returns a synth config dict."""
    from . import synth
    from .engine import PermutationEngine
    out = []
    for s in seeds:
        cfg = make_config(int(s))
        eng = PermutationEngine(cfg["Y"], cfg["X"], cfg["surfaces"], cfg["offsets"], cfg.get("H", 2.0),
                                cfg.get("E", 1.0), contrast=cfg["contrast"], precision=precision,
                                settings=cfg.get("settings"))
        n_perm = (n_values + 1) // 2
        null = eng.run(n_perm, batch=batch, seed=int(s) * 7919).ravel()[:n_values]
        out.append(null)
        del eng
    return out


def family_nulls(config, seeds, n_perm, batch=512, precision="fp16x3", log=None):
    """Long max-TFCE nulls (pooled +/- values in permutation order, 2 * n_perm
    each) of one synthetic workload family, one per seed: ``config(seed)``
    returns a synth config dict whose surfaces do not depend on the seed, so
    one engine serves every seed (``PermutationEngine.rebind``)."""
    import time
    from .engine import PermutationEngine
    eng = None
    out = []
    for j, s in enumerate(seeds):
        t0 = time.time()
        cfg = config(int(s))
        if eng is None:
            eng = PermutationEngine(cfg["Y"], cfg["X"], cfg["surfaces"], cfg["offsets"], cfg.get("H", 2.0),
                                    cfg.get("E", 1.0), contrast=cfg["contrast"], precision=precision,
                                    settings=cfg.get("settings"))
        else:
            eng.rebind(cfg["Y"], cfg["X"], contrast=cfg["contrast"])
        t1 = time.time()
        out.append(eng.run(n_perm, batch=batch, seed=int(s) * 7919).ravel().astype(np.float64))
        if log is not None:
            log(f"[study] {cfg.get('name', 'workload')} seed {int(s)} ({j + 1}/{len(seeds)}): data {t1 - t0:.1f}s, "
                f"{n_perm} permutations {time.time() - t1:.1f}s")
        del cfg
    del eng
    return out


# ----------------------------------------------------------------------------
# Workloads for training and evaluating the tail model
# ----------------------------------------------------------------------------
class WorkloadNull:
    """One permutation workload whose max-TFCE null can be drawn in batches.

    kind "glm": t of the last design column, +t and -t, two values per
    permutation (the order vertex/voxel randomise appends them);
    kind "med": type-M Sobel z of a mediation model, positive only, one value
    per permutation (vertex_tfce_mediation_randomise.py)."""

    def __init__(self, kind, Y, X=None, surfaces=None, offsets=None, settings=None, x=None, y=None,
                 precision="fp16x3"):
        from .engine import PermutationEngine
        self.kind = kind
        if kind == "glm":
            self.n = X.shape[0]
            self.engine = PermutationEngine(Y, X, surfaces, offsets, contrast=X.shape[1] - 1,
                                            precision=precision, settings=settings)
        else:
            from .randomise import MediationModel, _in_layout
            from .tfce import Adjacency, SurfaceSet
            self.n = len(x)
            surfaces, Y, _ = _in_layout(list(surfaces), offsets, Y)
            self.model = MediationModel(Y, x, y, "M", precision=precision)
            self.surfaces = SurfaceSet([Adjacency(c, s[0], s[1]) for c, s in zip(surfaces, settings)], offsets)

    def values(self, n_perm, seed=0, batch=256, log=None):
        import time
        from . import synth
        out, done = [], 0
        t0 = tl = time.time()
        while done < n_perm:
            if log is not None and time.time() - tl > 20:
                tl = time.time()
                log(f"[study]   {done}/{n_perm} permutations ({tl - t0:.0f}s)")
            P = min(batch, n_perm - done)
            pb = synth.permutations(self.n, P, seed + done, "cuda")
            if self.kind == "glm":
                out.append(self.engine.null_batch(pb).cpu().numpy().ravel())
            else:
                z = self.model.z(pb)
                out.append(self.surfaces.batch(z, 1).max(0).values[:, 0].cpu().numpy())
            done += P
        return np.concatenate(out).astype(np.float64)


def random_workload(rng, kind=None, precision="fp16x3", log=None):
    """A small random workload: vertex GLM (icosphere order 4-6, one or two
    hemispheres), voxel GLM (ellipsoid mask, 26-connectivity, E = 0.5) or
    vertex mediation; n, smoothness and covariates drawn from rng."""
    import torch
    from . import synth
    from .adjacency import voxel_csr
    if kind is None:
        kind = str(rng.choice(["vertex", "voxel", "med"], p=[0.45, 0.3, 0.25]))
    seed = int(rng.integers(1 << 30))
    g = synth._gen(seed, "cuda")
    n = int(rng.integers(40, 501))
    if log is not None:
        log(f"[study]   workload {kind}, n={n}")
    if kind in ("vertex", "med"):
        order = int(rng.choice([4, 5, 6, 7], p=[0.1, 0.3, 0.3, 0.3]))
        csr = synth.icosphere_csr(order)
        Vh = csr[0].shape[0] - 1
        hemis = int(rng.choice([1, 2]))
        S = synth.smoothing_operator(*csr, device="cuda")
        passes = int(rng.integers(1, 9))
        Y = torch.cat([synth.standardize(synth.mesh_smooth(torch.randn((n, Vh), generator=g, device="cuda"), S,
                                                           passes)) for _ in range(hemis)], 1).contiguous()
        surfaces, offsets, settings = [csr] * hemis, [h * Vh for h in range(hemis)], [(2.0, 1.0)] * hemis
        if kind == "med":
            x, y = rng.standard_normal(n), rng.standard_normal(n)
            return WorkloadNull("med", Y, surfaces=surfaces, offsets=offsets, settings=settings, x=x, y=y,
                                precision=precision)
    else:
        semi = tuple(float(s) for s in rng.uniform(7, 15, 3))
        shape = tuple(int(2 * s + 5) for s in semi)
        center = tuple((sh - 1) / 2.0 for sh in shape)
        mask = synth.ellipsoid_mask(shape, center, semi)
        vol = synth.gaussian_smooth3d(torch.randn((n,) + shape, generator=g, device="cuda"),
                                      float(rng.uniform(1.5, 4.0)))
        Y = synth.standardize(vol[:, torch.from_numpy(mask).cuda()]).contiguous()
        surfaces, offsets, settings = [voxel_csr(mask, 26)], [0], [(2.0, 0.5)]
    X = synth.design(rng, n, int(rng.integers(1, 5)), binary=bool(rng.random() < 0.5))
    return WorkloadNull("glm", Y, X=X, surfaces=surfaces, offsets=offsets, settings=settings, precision=precision)


def c3_workload(seed, precision="fp16x3"):
    """Two order-7 icospheres (327,684 vertices), n = 500, k = 6."""
    from . import synth
    cfg = synth.config3(seed=seed, device="cuda")
    return WorkloadNull("glm", cfg["Y"], X=cfg["X"], surfaces=cfg["surfaces"], offsets=cfg["offsets"],
                        settings=[(cfg["H"], cfg["E"])] * len(cfg["surfaces"]), precision=precision)


def c5_workload(seed, precision="fp16x3"):
    """Mediation on two order-7 icospheres, n = 300, three covariates
    regressed out of the mediator images first."""
    import torch
    from . import synth
    cfg = synth.config5(seed=seed, device="cuda")
    C = torch.as_tensor(np.column_stack([np.ones(cfg["n"]), cfg["covars"]]), dtype=torch.float64).cuda()
    M = cfg["M"].double()
    M = (M - C @ torch.linalg.lstsq(C, M).solution).float().contiguous()
    return WorkloadNull("med", M, surfaces=cfg["surfaces"], offsets=cfg["offsets"],
                        settings=[(cfg["H"], cfg["E"])] * len(cfg["surfaces"]), x=cfg["x"], y=cfg["y"],
                        precision=precision)


def study(n_train=96, n_train_perm=5000, n_eval=6, n_eval_perm=10000, m0=500, delta=0.0075, steps=6000, seed=0,
          n_random=0, log=print, streams=None):
    """Train the tail model on nulls of the voxel GLM (91 x 109 x 91 grid,
    ~232k voxels, n = 200) and the two-hemisphere GLM (327,684 vertices,
    n = 500) workloads with random seeds - half each, ``n_train_perm``
    permutations per null (2 values each) - plus ``n_random`` random small
    workloads; then evaluate on held-out seeds of the two-hemisphere GLM and
    of the mediation workload (327,684 vertices, n = 300) against their own
    ``n_eval_perm``-permutation plain runs.  Returns a dict of per-null
    results."""
    import time
    from . import synth
    rng = np.random.default_rng(seed)
    train = []
    half = n_train // 2
    seeds_c2 = [int(x) for x in rng.integers(1, 1 << 30, half)]
    seeds_c3 = [int(x) for x in rng.integers(1, 1 << 30, n_train - half)]
    train += family_nulls(lambda s: synth.config2(seed=s, device="cuda"), seeds_c2, n_train_perm, log=log)
    train += family_nulls(lambda s: synth.config3(seed=s, device="cuda"), seeds_c3, n_train_perm, batch=256,
                          log=log)
    for i in range(n_random):
        wl = random_workload(rng, log=log)
        train.append(wl.values(n_train_perm, seed=int(rng.integers(1 << 30)), log=log))
        log(f"[study] random training null {i + 1}/{n_random}: {wl.kind}")
        del wl
    X, Y = make_examples(train, m0=m0, per_null=128, seed=seed + 1)
    model = TailMLP(seed=seed + 2)
    losses = model.train(X, Y, steps=steps, lr=2e-3, chunk=500, log=log)
    log(f"[study] trained on {len(X)} examples: loss {losses[:50].mean():.4f} -> {losses[-200:].mean():.4f}")
    rule = LearnedStop(model, m0=m0, n_max=None, delta=delta)
    out = {"train_nulls": len(train), "train_perm": n_train_perm, "train_seeds": {"c2": seeds_c2, "c3": seeds_c3},
           "m0": m0, "delta": delta, "steps": steps, "eval": [],
           "rule": dict(m0=m0, delta=delta, m_min=rule.m_min, consecutive=rule.consecutive,
                        check_every=rule.check_every)}
    kept = {}
    for name, make in (("c3_glm_327k", c3_workload), ("c5_mediation_327k", c5_workload)):
        for e in range(n_eval):
            ts = time.time()
            wl = make(20261300 + 17 * e)
            log(f"[study] {name} #{e}: set-up {time.time() - ts:.1f}s")
            ts = time.time()
            vals = wl.values(n_eval_perm, seed=1000003 * (e + 1), log=log)
            dt = time.time() - ts
            del wl
            per_perm = 2 if name.startswith("c3") else 1
            kept["%s_%d" % (name, e)] = vals.astype(np.float32)
            truth = np.array([critical_value(vals, a) for a in ALPHAS])
            pred = model.predict_critical(vals[:m0])
            emp = np.array([critical_value(vals[:m0], a) for a in ALPHAS])
            rule.n_max = len(vals)
            used, thr, src = rule.simulate(vals)
            rec = dict(workload=name, seed=e, values=len(vals), seconds=round(dt, 2),
                       truth=truth.tolist(), pred_at_m0=pred.tolist(), emp_at_m0=emp.tolist(),
                       rel_err_pred=(np.abs(pred - truth) / truth).tolist(),
                       rel_err_emp=(np.abs(emp - truth) / truth).tolist(),
                       stop_values=int(used), stop_perms=int(used // per_perm), stop_source=src,
                       stop_rel_err=(np.abs(thr - truth) / truth).tolist())
            out["eval"].append(rec)
            log(f"[study] {name} #{e}: pred@{m0} err {np.round(rec['rel_err_pred'], 4)} emp@{m0} err "
                f"{np.round(rec['rel_err_emp'], 4)}; stop after {rec['stop_perms']} perms ({src}), err "
                f"{np.round(rec['stop_rel_err'], 4)}")
    if streams:
        np.savez(streams, **kept)
    return out


def summary(res):
    """Markdown summary of a ``study`` result."""
    ev = res["eval"]
    lines = ["| workload | null | perms used | source | rel. err. of reported 95% / 99% | pred@500 err 95% / 99% | "
             "empirical@500 err 95% / 99% |", "|---|---|---|---|---|---|---|"]
    for r in ev:
        full = r["values"] // (2 if r["workload"].startswith("c3") else 1)
        lines.append(f"| {r['workload']} | {r['seed']} | {r['stop_perms']} / {full} | {r['stop_source']} | "
                     f"{r['stop_rel_err'][0]:.4f} / {r['stop_rel_err'][1]:.4f} | "
                     f"{r['rel_err_pred'][0]:.4f} / {r['rel_err_pred'][1]:.4f} | "
                     f"{r['rel_err_emp'][0]:.4f} / {r['rel_err_emp'][1]:.4f} |")
    pe = np.array([r["rel_err_pred"] for r in ev])
    ee = np.array([r["rel_err_emp"] for r in ev])
    se = np.array([r["stop_rel_err"] for r in ev])
    frac = np.mean([r["stop_perms"] / (r["values"] // (2 if r["workload"].startswith("c3") else 1)) for r in ev])
    lines += ["", f"- mean rel. error of the prediction from {res['m0']} values: 95% {pe[:, 0].mean():.4f}, "
                  f"99% {pe[:, 1].mean():.4f}",
              f"- mean rel. error of the empirical quantile of the same values: 95% {ee[:, 0].mean():.4f}, "
              f"99% {ee[:, 1].mean():.4f}",
              f"- max rel. error of the reported threshold: 95% {se[:, 0].max():.4f}, 99% {se[:, 1].max():.4f}",
              f"- stopped early on {sum(r['stop_source'] == 'learned' for r in ev)} of {len(ev)} nulls; "
              f"permutations used: {100 * frac:.1f}% of the plain runs"]
    return "\n".join(lines)


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="learned-stop study on synthetic workloads")
    ap.add_argument("--train", type=int, default=96)
    ap.add_argument("--train-perm", type=int, default=5000)
    ap.add_argument("--random", type=int, default=0, help="extra random small training workloads")
    ap.add_argument("--eval", type=int, default=6)
    ap.add_argument("--eval-perm", type=int, default=10000)
    ap.add_argument("--delta", type=float, default=0.0075)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--out", default=None)
    ap.add_argument("--summary", default=None, help="markdown summary file")
    ap.add_argument("--streams", default=None, help="npz file for the evaluation null streams")
    args = ap.parse_args()
    res = study(args.train, args.train_perm, args.eval, args.eval_perm, delta=args.delta, steps=args.steps,
                n_random=args.random, log=lambda *a: print(*a, flush=True), streams=args.streams)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
    if args.summary:
        with open(args.summary, "w") as f:
            f.write(summary(res) + "\n")
    print(summary(res), flush=True)
