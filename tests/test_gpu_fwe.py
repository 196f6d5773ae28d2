"""GPU FWER lookups and apply_mfwer against the loop restatements (exact)."""
import numpy as np
import pytest

from oracle import fwe as ofwe

pytestmark = pytest.mark.gpu


def test_fwe_left_matches_reference_loop():
    from tfce_mediation_amd import fwe
    rng = np.random.default_rng(1)
    null = np.round(rng.lognormal(5, 0.5, 999), 4)
    vals = np.concatenate([rng.lognormal(5, 0.8, 5000), np.sort(null)[::50], [0.0, -3.0]]).astype(np.float32)
    got = fwe.calculate_fwep(vals, null)
    exp = ofwe.fwep_map(vals, null)
    np.testing.assert_array_equal(got, exp)


def test_fwe_nearest_matches_reference_loop():
    from tfce_mediation_amd import fwe
    rng = np.random.default_rng(2)
    null = np.sort(rng.normal(10, 1, 500))
    vals = np.concatenate([rng.normal(10, 1.5, 3000), null[::25], [0.0]])
    got = fwe.fwe_p64(vals, null)
    p = ofwe.p_array_of(500)
    exp = np.array([ofwe.find_nearest(null, v, p) for v in vals])
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("weight", [None, "logmasksize"])
def test_apply_mfwer_matches_reference_loop(weight):
    from tfce_mediation_amd import fwe
    rng = np.random.default_rng(3)
    sizes = [40, 120, 75]
    pos_arr = [0] + list(np.cumsum(sizes))
    V = pos_arr[-1]
    img = np.zeros((V, 3))
    img[:, 1] = rng.lognormal(3, 1.0, V) * (rng.random(V) < 0.7)
    img[:, 2] = rng.lognormal(3, 1.0, V) * (rng.random(V) < 0.7)
    nulls = {s: rng.lognormal(3 + 0.2 * s, 0.5, 300 + 10 * s) for s in range(3)}
    num_perm = min(len(v) for v in nulls.values())
    assert fwe.lowest_length(nulls.values()) == num_perm
    cb = lambda s, c: nulls[s]
    got = fwe.apply_mfwer([img], 1, [0, 1, 2], num_perm, 3, cb, pos_arr, [1], [2], weight=weight)
    exp = ofwe.apply_mfwer([img], 1, [0, 1, 2], num_perm, 3, cb, pos_arr, [1], [2], weight=weight)
    np.testing.assert_array_equal(got[0], exp[0])
    np.testing.assert_array_equal(got[1], exp[1])
