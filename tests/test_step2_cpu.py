"""Step-1 hand-off loader (no GPU): numeric arrays load as saved; object
arrays (ragged adjacency) are refused unless unpickling is asked for; the
step-1 writers save what the reference's step-1 scripts save."""
import os

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


# the variables each step-1 script saves (STEP_1_vertex_tfce_multiple_regression.py:248-266,
# STEP_1_vertex_tfce_mediation.py:247-277, STEP_0_load_volumes.py:89-96 without the nibabel
# header, STEP_1_voxel_tfce_multiple_regression.py:88-116, STEP_1_voxel_tfce_mediation.py:90-95)
VERTEX_FILES = {"pred_x", "num_subjects", "all_vertex", "num_vertex", "num_vertex_lh", "num_vertex_rh",
                "bin_mask_lh", "bin_mask_rh", "affine_mask_lh", "affine_mask_rh", "adjac_lh", "adjac_rh",
                "merge_y", "optstfce", "vdensity_lh", "vdensity_rh"}
VERTEX_MED_FILES = {"pred_x", "depend_y", "num_subjects", "num_vertex", "num_vertex_lh", "num_vertex_rh",
                    "all_vertex", "bin_mask_lh", "bin_mask_rh", "adjac_lh", "adjac_rh", "optstfce", "vdensity_lh",
                    "vdensity_rh", "merge_y"}
VOLUME_FILES = {"raw_nonzero", "affine_mask", "data_mask", "num_voxel", "num_subjects"}
VOXEL_FILES = VOLUME_FILES | {"covars", "adjac", "pred_x", "ancova", "optstfce", "raw_nonzero_corr"}
VOXEL_MED_FILES = VOLUME_FILES | {"pred_x", "depend_y", "adjac", "medtype", "optstfce", "raw_nonzero_corr"}


def _names(d):
    return {f[:-4] for f in os.listdir(d) if f.endswith(".npy")}


def test_handoff_writers_round_trip(tmp_path):
    """The step-1 writers save the reference's variables under its names, and
    step 2's reader gets them back (the adjacency only with allow_pickle)."""
    from tfce_mediation_amd import adjacency, step1, step2, synth
    rp, col = synth.icosphere_csr(2)
    V = len(rp) - 1
    rng = np.random.default_rng(4)
    m_lh, m_rh = rng.random(V) < 0.9, rng.random(V) < 0.8
    n = 12
    y = rng.standard_normal((n, int(m_lh.sum() + m_rh.sum())))
    x, dep = rng.standard_normal((n, 2)), rng.standard_normal(n)
    vd = adjacency.vertex_density(rp)
    d = step1.save_vertex_handoff("area", x, y, (rp, col), (rp, col), m_lh, m_rh, [2, 1], vd, vd, root=str(tmp_path))
    assert _names(d) == VERTEX_FILES
    with pytest.raises(ValueError, match="allow_pickle"):
        step2.load_temp(d, ["adjac_lh"])
    v = step2.load_temp(d, sorted(VERTEX_FILES), allow_pickle=True)
    assert v["merge_y"].dtype == np.float32 and np.array_equal(v["merge_y"], y.astype(np.float32))
    assert int(v["num_subjects"]) == n and int(v["all_vertex"]) == V
    assert int(v["num_vertex_lh"]) == m_lh.sum() and int(v["num_vertex"]) == y.shape[1]
    assert [list(r) for r in v["adjac_rh"]] == [col[rp[i]:rp[i + 1]].tolist() for i in range(V)]
    np.testing.assert_array_equal(v["vdensity_lh"], vd)
    np.testing.assert_array_equal(v["optstfce"], [2, 1])
    d = step1.save_vertex_mediation_handoff("area", x[:, 0], dep, y, (rp, col), (rp, col), m_lh, m_rh, [2, 1],
                                            root=str(tmp_path))
    assert _names(d) == VERTEX_MED_FILES
    np.testing.assert_array_equal(step2.load_temp(d, ["depend_y"])["depend_y"], dep)

    mask = synth.ellipsoid_mask((7, 8, 6), (3, 3.5, 2.5), (3, 3.5, 2.5))
    vrp, vcol = adjacency.voxel_csr(mask, 26)
    nv = int(mask.sum())
    yv = rng.standard_normal((n, nv))
    d = step1.save_voxel_handoff(yv, x, (vrp, vcol), mask, [2, 0.5], ancova=0, covars=rng.standard_normal(n),
                                 root=str(tmp_path / "vox"))
    assert _names(d) == VOXEL_FILES
    v = step2.load_temp(d, ["raw_nonzero_corr", "num_subjects", "num_voxel", "ancova", "data_mask"])
    assert v["raw_nonzero_corr"].shape == (nv, n) and int(v["num_voxel"]) == nv and int(v["ancova"]) == 0
    assert np.array_equal(v["data_mask"] > 0.99, mask)
    d = step1.save_voxel_mediation_handoff(yv, x[:, 0], dep, "M", (vrp, vcol), mask, [2, 0.5],
                                           root=str(tmp_path / "voxmed"))
    assert _names(d) == VOXEL_MED_FILES
    assert str(step2.load_temp(d, ["medtype"])["medtype"]) == "M"
