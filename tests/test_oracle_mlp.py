"""The FP64 tail-model oracle: analytic gradient vs central differences."""
import numpy as np

from oracle import mlp


def test_analytic_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(mlp.N_PARAMS) * 0.2
    X = rng.standard_normal((17, mlp.IN))
    Y = rng.standard_normal((17, mlp.OUT))
    g = mlp.grad(p, X, Y)
    idx = np.concatenate([rng.choice(mlp.N_PARAMS, 60, replace=False), [mlp.N_PARAMS - 1, 0, 2047, 2048]])
    num = mlp.numeric_grad(p, X, Y, idx)
    np.testing.assert_allclose(g[idx], num, rtol=1e-5, atol=1e-9)
    assert mlp.N_PARAMS == 6402
