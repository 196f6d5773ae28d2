"""Step-1 hand-off loader (no GPU): numeric arrays load as saved; object
arrays (ragged adjacency) are refused unless unpickling is asked for."""
import numpy as np
import pytest

from tfce_mediation_amd import step2


def test_load_temp_refuses_pickled_arrays_by_default(tmp_path):
    np.save(tmp_path / "num_subjects", 12)
    obj = np.empty(3, dtype=object)
    obj[0], obj[1], obj[2] = [1, 2], [0], [0]
    np.save(tmp_path / "adjac", obj)
    v = step2.load_temp(str(tmp_path), ["num_subjects"])
    assert int(v["num_subjects"]) == 12
    with pytest.raises(ValueError, match="allow_pickle"):
        step2.load_temp(str(tmp_path), ["adjac"])
    v = step2.load_temp(str(tmp_path), ["adjac"], allow_pickle=True)
    assert [list(r) for r in v["adjac"]] == [[1, 2], [0], [0]]


def test_glm_design_blocks():
    """glm_typeI's design (pyfunc.py:2298-2311): intercept, each exog variable
    as one block of its columns, covariates last."""
    from tfce_mediation_amd.randomise import glm_design
    n = 5
    x = np.arange(n, dtype=float)
    g = np.column_stack([np.ones(n), np.zeros(n)])
    X, blocks = glm_design(n, [x, g], dmy_covariates=np.full((n, 1), 7.0))
    assert X.shape == (n, 5)
    assert blocks == [(1, 1), (2, 2)]
    np.testing.assert_array_equal(X[:, 0], 1.0)
    np.testing.assert_array_equal(X[:, 1], x)
    np.testing.assert_array_equal(X[:, 4], 7.0)
    X2, blocks2 = glm_design(n, [g])
    assert X2.shape == (n, 3) and blocks2 == [(1, 2)]


def test_tm_glm_folder_refuses_pickles_by_default(tmp_path):
    """A tmtemp_GLM folder holds np.save'd object arrays (adjacency, a None
    covariate array): step 2 refuses them before any GPU work unless
    unpickling is asked for."""
    d = tmp_path / "tmtemp_GLM_volume"
    d.mkdir()
    n = 6
    for k, v in dict(exog_flat=np.zeros((n, 1)), exog_shape=np.array([1]), varnames=np.array(["x"]),
                     gstat=np.array("f"), data=np.zeros((n, 3), np.float32), optstfce=np.array([2.0, 1.0]),
                     dmy_covariates=None).items():
        np.save(d / k, v)
    obj = np.empty(3, dtype=object)
    obj[0], obj[1], obj[2] = [1], [0, 2], [1]
    np.save(d / "adjac", obj)
    with pytest.raises(ValueError, match="allow_pickle"):
        step2.tm_glm(1, 2, root=str(tmp_path))


def test_tm_mediation_folder_refuses_pickles_by_default(tmp_path):
    d = tmp_path / "tmtemp_mediation_volume"
    d.mkdir()
    n = 6
    for k, v in dict(dmy_leftvar=np.arange(n, dtype=float), dmy_rightvar=np.ones(n), medtype=np.array("M"),
                     data=np.zeros((n, 3), np.float32), optstfce=np.array([2.0, 1.0]),
                     dmy_covariates=None).items():
        np.save(d / k, v)
    obj = np.empty(3, dtype=object)
    obj[0], obj[1], obj[2] = [1], [0, 2], [1]
    np.save(d / "adjac", obj)
    with pytest.raises(ValueError, match="allow_pickle"):
        step2.tm_mediation(1, 2, root=str(tmp_path))


def test_type_y_path_a_with_covariates_matches_oracle():
    """Type Y's path A with covariates (tm_models_randomise.py:482-488): the
    t of left in right[perm] ~ [1, left[perm], covariates], solved on the
    host per permutation, against row 1 of the oracle's glm_typeI t."""
    import types
    from oracle import ols as ools
    from tfce_mediation_amd.randomise import MediationModel, permutation_indices
    rng = np.random.default_rng(3)
    n = 30
    for wx in (1, 2):
        left = rng.standard_normal((n, wx)) if wx > 1 else rng.standard_normal(n)
        right = 0.5 * left.reshape(n, -1)[:, 0] + rng.standard_normal(n)
        cov = rng.standard_normal((n, 2))
        Xa = np.column_stack([np.ones(n), left.reshape(n, wx), cov])
        stub = types.SimpleNamespace(ta_fit=(Xa, right, wx), n=n)
        perms = permutation_indices(n, range(5), seed=8)
        got = MediationModel.path_a_t(stub, perms)
        for p, perm in enumerate(perms):
            exp = ools.glm_typeI(right[perm], [left[perm]], dmy_covariates=cov)[2][1]
            assert abs(got[p] - exp) <= 1e-10 * abs(exp)
