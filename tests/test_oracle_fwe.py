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
