"""Learned early stop of the permutation loop.

The reference always runs the requested number of permutations and reads the
FWER critical value off the sorted null (tools/plot_permutations.py:71-85:
``sorted(null, reverse=True)[int(N * alpha)]``).  Here a small tail model
predicts the 95% and 99% critical values from the first ``m0`` null values, and
the loop may stop early only when that prediction agrees with the running
empirical critical values within ``delta`` (relative) for both levels; otherwise
it runs to ``n_max`` values and reports the plain empirical result.  The plain
run is the fallback and the judge of the model's error.

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

    def train(self, X, Y, steps=3000, batch=128, lr=2e-3, weight_decay=1e-5, seed=0, want_grad0=False):
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
        _lib.check(_lib.lib().tfce_mlp_train(_lib.ptr(self.params), _lib.ptr(self.m), _lib.ptr(self.v),
                                             _lib.ptr(Xd), _lib.ptr(Yd), _lib.ptr(idx_d), steps, b, self.t,
                                             float(lr), float(weight_decay), _lib.ptr(loss), _lib.ptr(g0),
                                             _lib.stream_ptr()))
        self.t += steps
        out = loss.cpu().numpy()
        return (out, g0.cpu().numpy(), idx) if want_grad0 else out

    def predict_critical(self, first_values):
        f, mu, sd = features(first_values)
        y = self.predict(f[None, :])[0]
        return np.exp(mu + sd * y.astype(np.float64))


class LearnedStop:
    """Stopping rule over a stream of null values (pooled +/- maxima in
    permutation order)."""

    def __init__(self, model, m0=500, n_max=10000, delta=0.01, m_min=1000):
        self.model, self.m0, self.n_max, self.delta, self.m_min = model, m0, n_max, delta, m_min

    def decide(self, values):
        """(stop, thresholds, source) for the values seen so far."""
        values = np.asarray(values, dtype=np.float64).ravel()
        if len(values) >= self.n_max:
            return True, np.array([critical_value(values[:self.n_max], a) for a in ALPHAS]), "plain"
        if len(values) < max(self.m0, self.m_min):
            return False, None, None
        pred = self.model.predict_critical(values[:self.m0])
        emp = np.array([critical_value(values, a) for a in ALPHAS])
        if np.all(np.abs(pred - emp) <= self.delta * emp):
            return True, pred, "learned"
        return False, None, None

    def simulate(self, stream, check_every=512):
        """Replay a recorded null stream: (values used, thresholds, source)."""
        stream = np.asarray(stream, dtype=np.float64).ravel()
        m = max(self.m0, self.m_min)
        while True:
            stop, thr, src = self.decide(stream[:min(m, self.n_max)])
            if stop:
                return min(m, self.n_max), thr, src
            m += check_every

    def run(self, engine, batch=256, seed=0):
        """Live loop on a PermutationEngine: batches until the rule stops."""
        from . import synth
        values = []
        done = 0
        while True:
            P = batch
            pb = synth.permutations(engine.n, P, seed + done, "cuda")
            values.extend(engine.null_batch(pb).cpu().numpy().ravel().tolist())
            done += P
            stop, thr, src = self.decide(values)
            if stop:
                return len(values), thr, src


def generate_nulls(make_config, seeds, n_values, batch=256, precision="bf16x3"):
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
