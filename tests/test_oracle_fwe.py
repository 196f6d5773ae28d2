"""Pins the FWER restatements with hand-made nulls and closed forms."""
import numpy as np
import pytest

from oracle import fwe as ofwe


def test_left_lookup_closed_form():
    null = np.array([3.0, 1.0, 2.0, 5.0, 4.0])
    vals = np.array([0.5, 1.0, 1.5, 3.0, 4.5, 6.0, 0.0, -1.0])
    got = ofwe.fwep_map(vals, null)
    # idx = #(null < v); p = (idx-1)/N floored at 0; v <= 0 -> 0
    assert list(got) == [0.0, 0.0, 0.0, 0.2, 0.6, 0.8, 0.0, 0.0]


def test_nearest_lookup_closed_form():
    s = np.array([1.0, 2.0, 4.0, 8.0])
    p = ofwe.p_array_of(4)
    assert ofwe.find_nearest(s, 2.9, p) == 0.25   # nearer to 2 -> idx-1
    assert ofwe.find_nearest(s, 3.1, p) == 0.5    # nearer to 4 -> idx
    assert ofwe.find_nearest(s, 3.0, p) == 0.5    # tie -> idx
    assert ofwe.find_nearest(s, 9.0, p) == 0.75   # beyond the null
    assert ofwe.find_nearest(s, 0.0, p) == 0.0    # idx 0 compares with s[-1] (wraps), stays at 0


def test_mfwer_single_surface_reduces_to_nearest_on_standardised_logs():
    rng = np.random.default_rng(0)
    null = rng.lognormal(3, 0.4, 200)
    img = np.zeros((50, 3))
    img[:, 1] = rng.lognormal(3, 0.6, 50)
    img[:, 2] = rng.lognormal(2.5, 0.6, 50)
    pos, neg = ofwe.apply_mfwer([img], 1, [0], 200, 1, lambda s, c: null, [0, 50], [1], [2])
    ln = np.log(null)
    z = (ln - ln.mean()) / ln.std() + 10
    zs = np.sort(z)
    for col, out in ((1, pos), (2, neg)):
        v = np.log(img[:, col])
        m = v > 0
        e = np.zeros(50)
        e[m] = (v[m] - ln.mean()) / ln.std() + 10
        exp = [ofwe.find_nearest(zs, k, ofwe.p_array_of(200)) for k in e]
        np.testing.assert_allclose(out[:, 0], exp)


def test_perm_max_scaling():
    t = np.array([1.0, 3.0, -2.0], np.float32)
    tf = np.array([10.0, 50.0, 0.0], np.float32)
    assert ofwe.perm_max_voxel(tf, t) == np.float32(50.0) * np.float32(np.float32(3.0) / np.float32(100))


# Three surfaces of 10, 100 and 1000 elements: the log-mask-size weights of
# tm_func.py:418-423 are log(10^j)/mean = 0.5, 1.0 and 1.5 exactly.  The log
# nulls are affine images of these patterns (one multiset, so one std), and
# the surface whose weighted standardised value is largest in each
# permutation is worked out by hand below.
_R = np.array([[2, -2, 1, 0, -1, 0],     # surface 0, weight 0.5
               [0, 1, 2, -2, 0, -1],     # surface 1, weight 1.0
               [1, 0, -2, 0, 2, -1]],    # surface 2, weight 1.5
              dtype=np.float64)
# weighted z * w (z = r / sigma):  perm 0: 1.0, 0, 1.5 -> surface 2;
# perm 1: -1, 1, 0 -> 1; perm 2: 0.5, 2, -3 -> 1; perm 3: 0, -2, 0 -> 0 (first
# of a tie); perm 4: -0.5, 0, 3 -> 2; perm 5: 0, -1, -1.5 -> 0.
_W_WINNER = [2, 1, 1, 0, 2, 0]
# unweighted z: perm 0: 2, 0, 1 -> 0; perm 1: 1; perm 2: 1; perm 3: 0 (tie
# 0 vs 0); perm 4: 2; perm 5: 0
_U_WINNER = [0, 1, 1, 0, 2, 0]


def _mfwer_case(weight):
    from scipy import stats
    pos_arr = [0, 10, 110, 1110]
    lognull = [0.3 * _R[s] + 3.0 + s for s in range(3)]
    nulls = [np.exp(l) for l in lognull]
    rng = np.random.default_rng(4)
    img = np.zeros((1110, 3))
    img[:, 1] = np.exp(rng.normal(3.5, 0.5, 1110))
    img[:, 2] = np.exp(rng.normal(3.0, 0.5, 1110))
    img[5, 1] = 0.5  # log < 0: masked out, p = 0 (lookup of 0)
    pos, neg = ofwe.apply_mfwer([img], 1, [0, 1, 2], 6, 3, lambda s, c: nulls[s], pos_arr, [1], [2],
                                weight=weight)
    z = [stats.zscore(_R[s]) for s in range(3)]  # population std, as np.std
    winners = _W_WINNER if weight == "logmasksize" else _U_WINNER
    maxes = np.sort([z[winners[i]][i] + 10 for i in range(6)])
    p = np.arange(6) / 6.0
    for col, out in ((1, pos), (2, neg)):
        exp = np.zeros(1110)
        for s in range(3):
            a, b = pos_arr[s], pos_arr[s + 1]
            lv = np.log(img[a:b, col])
            m = lv > 0
            e = np.zeros(b - a)
            e[m] = (lv[m] - lognull[s].mean()) / lognull[s].std() + 10
            exp[a:b] = [ofwe.find_nearest(maxes, k, p) for k in e]
        np.testing.assert_allclose(out[:, 0], exp, rtol=0, atol=1e-12)
    return maxes


def test_mfwer_three_surfaces_logmasksize_weights():
    m_w = _mfwer_case("logmasksize")
    m_u = _mfwer_case(None)
    # the weighting picks another surface in permutation 0 (z 1/sigma instead of 2/sigma)
    assert not np.allclose(m_w, m_u)
