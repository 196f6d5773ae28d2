"""GPU permuted OLS t-values (tfce_ols_perm_t) against the FP64 oracle.

Error metrics (t has no natural scale near 0, so two are used):
  norm  = |t_gpu - t64|_2 / |t64|_2 over the whole batch;
  point = max |t_gpu - t64| / |t64| over entries with |t64| >= 1.
Bars: FP32, FP16x3 (the drivers' default: hi/lo split-FP16 of power-of-two
scaled operands, 3 products) and BF16x6 (3-piece split, 6 products) <= 1e-5 on
both; BF16x3 (hi/lo split, 3 products: 16 significant bits per operand, about
2^-17 relative) holds norm <= 2e-5 only, and no driver uses it by default.
"""
import sys

import numpy as np
import pytest

from oracle import ols as ools

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    return _t


def make(n, k, V, seed, offset=0.0):
    rng = np.random.default_rng(seed)
    X = np.column_stack([np.ones(n)] + [rng.standard_normal(n) for _ in range(k - 2)]
                        + [(rng.random(n) < 0.5).astype(float)])
    Y = rng.standard_normal((n, V)) + 0.4 * X[:, -1:] * (rng.random(V) < 0.2) + offset
    return X.astype(np.float32).astype(np.float64), Y.astype(np.float32)


def errors(t, ref):
    t = t.astype(np.float64)
    norm = np.linalg.norm(t - ref) / np.linalg.norm(ref)
    big = np.abs(ref) >= 1
    point = np.max(np.abs(t - ref)[big] / np.abs(ref[big]))
    return norm, point


BARS = {"fp32": (1e-5, 1e-5), "bf16x6": (1e-5, 1e-5), "fp16x3": (1e-5, 1e-5), "bf16x3": (2e-5, None)}


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16x6", "fp16x3"])
@pytest.mark.parametrize("n,k,V,P", [(100, 4, 3000, 37), (160, 6, 2500, 20), (75, 5, 1030, 9)])
def test_perm_t_matches_fp64(torch, prec, n, k, V, P):
    from tfce_mediation_amd.ols import PermutedOLS
    say(f"[gpu] ols {prec} n={n} k={k} V={V} P={P}")
    X, Y = make(n, k, V, seed=n + k)
    rng = np.random.default_rng(7)
    perms = np.stack([rng.permutation(n) for _ in range(P)]).astype(np.int32)
    con = list(range(1, k))
    model = PermutedOLS(torch.from_numpy(Y).cuda(), X, contrasts=con, precision=prec)
    t = model.t(torch.from_numpy(perms).cuda()).cpu().numpy()
    ref = ools.perm_t(X, Y, perms, con)
    norm, point = errors(t, ref)
    say(f"    norm={norm:.2e} point={point:.2e}")
    nb, pb = BARS[prec]
    assert norm <= nb
    if pb is not None:
        assert point <= pb


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp16x3"])
def test_partial_permutation_and_offset(torch, prec):
    """-v mode (only columns 2..3 move: per-permutation Gram factors) and a
    large per-element mean (centring)."""
    from tfce_mediation_amd.ols import PermutedOLS
    n, k, V, P = 90, 5, 1500, 12
    say(f"[gpu] ols partial permutation {prec}")
    X, Y = make(n, k, V, seed=3, offset=40.0)
    rng = np.random.default_rng(2)
    perms = np.stack([rng.permutation(n) for _ in range(P)]).astype(np.int32)
    con = [1, 2, 3, 4]
    model = PermutedOLS(torch.from_numpy(Y).cuda(), X, contrasts=con, perm_columns=[2, 3], precision=prec)
    t = model.t(torch.from_numpy(perms).cuda()).cpu().numpy()
    ref = ools.perm_t(X, Y, perms, con, perm_columns=[2, 3])
    norm, point = errors(t, ref)
    say(f"    norm={norm:.2e} point={point:.2e}")
    assert norm <= BARS[prec][0] * 2


def test_identity_is_tval_int_with_intercept(torch):
    from tfce_mediation_amd.ols import PermutedOLS
    n, k, V = 50, 3, 700
    say("[gpu] ols identity incl. intercept contrast")
    X, Y = make(n, k, V, seed=11, offset=3.0)
    model = PermutedOLS(torch.from_numpy(Y).cuda(), X, contrasts=[0, 1, 2], precision="fp32")
    t = model.t().cpu().numpy()[:, 0]
    ref = ools.tval_int(X, np.linalg.inv(X.T @ X), Y.astype(np.float64), n, k, V)
    np.testing.assert_allclose(t, ref, rtol=2e-5, atol=1e-5)


def test_constant_element_gives_nan_like_reference(torch):
    from tfce_mediation_amd.ols import PermutedOLS
    n, k, V = 40, 3, 64
    X, Y = make(n, k, V, seed=4)
    Y[:, 5] = 2.0
    model = PermutedOLS(torch.from_numpy(Y).cuda(), X, contrasts=[1, 2], precision="bf16x3")
    t = model.t().cpu().numpy()
    assert np.all(np.isnan(t[:, 0, 5]))
    assert np.all(np.isfinite(np.delete(t[:, 0, :], 5, axis=1)))


def test_fp16x3_scaling_covers_wide_ranges(torch):
    """fp16 pieces of power-of-two scaled operands: data of magnitude 1e4 and
    1e-4 and a covariate of magnitude 1e3 keep the FP32-class bar."""
    from tfce_mediation_amd.ols import PermutedOLS
    n, k, V, P = 120, 4, 2000, 16
    say("[gpu] ols fp16x3 scaling")
    X, Y = make(n, k, V, seed=21)
    X[:, 1] *= 1000.0
    Y[:, : V // 2] *= 1e4
    Y[:, V // 2:] *= 1e-4
    rng = np.random.default_rng(5)
    perms = np.stack([rng.permutation(n) for _ in range(P)]).astype(np.int32)
    con = [1, 2, 3]
    t = PermutedOLS(torch.from_numpy(Y).cuda(), X, contrasts=con, precision="fp16x3").t(
        torch.from_numpy(perms).cuda()).cpu().numpy()
    ref = ools.perm_t(X, Y, perms, con)
    norm, point = errors(t, ref)
    say(f"    norm={norm:.2e} point={point:.2e}")
    assert norm <= 1e-5 and point <= 1e-5


@pytest.mark.parametrize("prec", ["bf16x3", "fp16x3", "bf16x6"])
def test_every_tile_gives_identical_t(torch, prec):
    """The contraction's workgroup tiles (autotuner choices) sum each output
    in the same k order: identical t-values, including a P that leaves a
    partial row tile."""
    from tfce_mediation_amd import _lib
    from tfce_mediation_amd.ols import PermutedOLS
    n, k, V, P = 130, 6, 5000, 70
    say(f"[gpu] ols tiles {prec}")
    X, Y = make(n, k, V, seed=31)
    rng = np.random.default_rng(9)
    perms = torch.from_numpy(np.stack([rng.permutation(n) for _ in range(P)]).astype(np.int32)).cuda()
    model = PermutedOLS(torch.from_numpy(Y).cuda(), X, contrasts=[1, 5], precision=prec)
    lib = _lib.lib()
    outs = []
    try:
        for tile in (0, 1, 2, 3):
            _lib.check(lib.tfce_ols_set_tile(tile))
            outs.append(model.t(perms).cpu().numpy())
    finally:
        _lib.check(lib.tfce_ols_set_tile(0))
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


@pytest.mark.parametrize("prec", ["fp32", "fp16x3", "bf16x3"])
def test_partial_f_matches_fp64(torch, prec):
    """glm_typeI's per-variable F (tfce_ols_perm_partial_f) against the FP64
    oracle: blocks of widths 1, 2 and 3 after the intercept, one covariate
    (k = 8, the largest design),
    every column permuted (shared Gram factor); bars as for t (F ~ t^2
    doubles the relative error: norm <= 2e-5, point <= 2e-5 on F >= 1 for the
    FP32-class modes)."""
    from tfce_mediation_amd.ols import PermutedOLS
    n, V, P = 110, 2600, 21
    say(f"[gpu] ols partial F {prec}")
    rng = np.random.default_rng(41)
    x1 = rng.standard_normal(n)
    grp = rng.integers(0, 3, n)
    d2 = np.column_stack([(grp == 1), (grp == 2)]).astype(float)
    b3 = rng.standard_normal((n, 3))
    cov = rng.standard_normal((n, 1))
    X = np.column_stack([np.ones(n), x1, d2, b3, cov]).astype(np.float32).astype(np.float64)
    Y = (rng.standard_normal((n, V)) + 0.5 * x1[:, None] * (rng.random(V) < 0.3)
         + 0.6 * d2[:, :1] * (rng.random(V) < 0.3)).astype(np.float32)
    perms = np.stack([rng.permutation(n) for _ in range(P)]).astype(np.int32)
    blocks = [(1, 1), (2, 2), (4, 3)]
    model = PermutedOLS(torch.from_numpy(Y).cuda(), X, contrasts=[1], precision=prec)
    f = model.partial_f(torch.from_numpy(perms).cuda(), blocks).cpu().numpy()
    ref = ools.perm_partial_f(X, Y, perms, blocks)
    norm = np.linalg.norm(f - ref) / np.linalg.norm(ref)
    big = ref >= 1
    point = np.max(np.abs(f - ref)[big] / ref[big])
    say(f"    norm={norm:.2e} point={point:.2e}")
    if prec == "bf16x3":
        assert norm <= 4e-5
    else:
        assert norm <= 2e-5 and point <= 2e-5
    # one-column block: F = t^2 of the same column from the t path
    t = model.t(torch.from_numpy(perms).cuda()).cpu().numpy()[0].astype(np.float64)
    np.testing.assert_allclose(f[0], t ** 2, rtol=2e-5, atol=1e-6)


def test_partial_f_with_fixed_columns(torch):
    """Partial permutation (only columns 1..3 move): per-permutation factors;
    the block of fixed covariates and a permuted block against the oracle."""
    from tfce_mediation_amd.ols import PermutedOLS
    n, V, P = 90, 1500, 13
    say("[gpu] ols partial F, fixed covariates")
    rng = np.random.default_rng(42)
    X = np.column_stack([np.ones(n), rng.standard_normal((n, 3)), rng.standard_normal((n, 2))])
    X = X.astype(np.float32).astype(np.float64)
    Y = (rng.standard_normal((n, V)) + 0.4 * X[:, 1:2] + 0.3 * X[:, 4:5]).astype(np.float32)
    perms = np.stack([rng.permutation(n) for _ in range(P)]).astype(np.int32)
    blocks = [(1, 3), (4, 2)]
    model = PermutedOLS(torch.from_numpy(Y).cuda(), X, contrasts=[1], perm_columns=[1, 2, 3], precision="fp16x3")
    f = model.partial_f(torch.from_numpy(perms).cuda(), blocks).cpu().numpy()
    ref = ools.perm_partial_f(X, Y, perms, blocks, perm_columns=[1, 2, 3])
    norm = np.linalg.norm(f - ref) / np.linalg.norm(ref)
    big = ref >= 1
    point = np.max(np.abs(f - ref)[big] / ref[big])
    say(f"    norm={norm:.2e} point={point:.2e}")
    assert norm <= 2e-5 and point <= 2e-5
