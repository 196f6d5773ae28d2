"""The tm-models GLM (glm_typeI, pyfunc.py:2282-2404, and the GLM branch of
tm_models_randomise.py:182-243) on the GPU against the FP64 oracle: the
single-call drop-in, the batched permutation driver's t and per-variable F
nulls against the oracle loop (FP64 statistics rounded to float32, the
oracle TFCE), and step 2 from a ``tmtemp_GLM_*`` folder laid out as the
reference's step 1 saves it (written by this test, so it may be unpickled)."""
import os

import numpy as np
import pytest

from oracle import ols as ools
from oracle import tfce as otfce
from tfce_mediation_amd import adjacency, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _exog(n, seed):
    rng = np.random.default_rng(seed)
    grp = rng.integers(0, 3, n)
    return rng, [rng.standard_normal(n), np.column_stack([grp == 1, grp == 2]).astype(float)], \
        rng.standard_normal((n, 1))


def test_glm_typeI_dropin_matches_oracle():
    from tfce_mediation_amd import pyfunc
    say("[gpu] pyfunc.glm_typeI")
    n, V = 64, 900
    rng, exog, cov = _exog(n, 1)
    Y = (rng.standard_normal((n, V)) + 0.5 * exog[0][:, None] * (rng.random(V) < 0.3)).astype(np.float32)
    perm = rng.permutation(n)
    fm, fv, t = ools.glm_typeI(Y, exog, dmy_covariates=cov, rand_array=perm)
    gfm, gfv, gt = pyfunc.glm_typeI(Y, exog, dmy_covariates=cov, output_tvalues=True, verbose=False,
                                    rand_array=perm)
    np.testing.assert_allclose(gfm, fm, rtol=1e-9)
    np.testing.assert_allclose(gfv, fv, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(gt, t, rtol=1e-6)
    # t only, with p-values; F only, with reduced residuals
    gt2, gp = pyfunc.glm_typeI(Y, exog, dmy_covariates=cov, output_fvalues=False, output_tvalues=True,
                               output_pvalues=True, verbose=False)
    _, _, t0 = ools.glm_typeI(Y, exog, dmy_covariates=cov)
    np.testing.assert_allclose(gt2, t0, rtol=1e-6)
    assert gp.shape == t0.shape and np.all((gp >= 0) & (gp <= 1))
    f3, fv3, resid = pyfunc.glm_typeI(Y, exog, output_reduced_residuals=True, verbose=False)
    X = np.column_stack([np.ones(n), exog[0], exog[1]])
    coef = np.linalg.lstsq(X, Y.astype(np.float64), rcond=None)[0]
    np.testing.assert_allclose(resid, Y - X @ coef, atol=1e-9)


@pytest.mark.parametrize("gstat", ["t", "f", "ft"])
def test_tm_glm_voxel_null_matches_oracle_loop(tmp_path, gstat):
    from tfce_mediation_amd import randomise
    say(f"[gpu] randomise.tm_glm voxel gstat={gstat}")
    mask = synth.ellipsoid_mask((12, 12, 11), (5.5, 5.5, 5), (5, 5, 4.5))
    rp, col = adjacency.voxel_csr(mask, 26)
    V = int(mask.sum())
    n = 48
    rng, exog, cov = _exog(n, 2)
    Y = (rng.standard_normal((n, V)) + 0.6 * exog[1][:, :1] * (rng.random(V) < 0.4)).astype(np.float32)
    perms = randomise.permutation_indices(n, range(1, 7), seed=3)
    nulls = randomise.tm_glm(Y, exog, cov, gstat, [2.0, 1.0], perms, adjac=(rp, col), varnames=["age", "group"],
                             out_dir=str(tmp_path), precision="fp32")
    names = []
    if gstat != "f":
        names += ["Tstat_con1", "Tstat_con2", "Tstat_con3"]
    if gstat != "t":
        names += ["Fstat_age", "Fstat_group"]
    assert sorted(nulls) == sorted(names)
    for p, perm in enumerate(perms):
        _, fv, t = ools.glm_typeI(Y, exog, dmy_covariates=cov, rand_array=perm)
        for name in names:
            if name.startswith("T"):
                j = int(name[-1])
                maps = [t[j].astype(np.float32), -t[j].astype(np.float32)]
                got = nulls[name][p]
            else:
                maps = [fv[["Fstat_age", "Fstat_group"].index(name)].astype(np.float32)]
                got = [nulls[name][p]]
            for g, z in zip(got, maps):
                enh, _, _ = otfce.tfce(rp, col, z, 2.0, 1.0, exact_pow=True)
                exp = enh.max() * (z.max() / np.float32(100))
                assert abs(g - exp) <= 1e-4 * abs(exp), (name, p, g, exp)
    for name in names:
        vals = np.loadtxt(os.path.join(tmp_path, "perm_%s_TFCE_maxVoxel.csv" % name))
        np.testing.assert_allclose(vals, np.asarray(nulls[name]).ravel(), atol=5e-5)


def _ragged(row_ptr, col):
    out = np.empty(len(row_ptr) - 1, dtype=object)
    for i in range(len(row_ptr) - 1):
        out[i] = col[row_ptr[i]:row_ptr[i + 1]].tolist()
    return out


def test_tm_glm_from_step1_folders(tmp_path):
    from tfce_mediation_amd import randomise, step2
    say("[gpu] step2.tm_glm from tmtemp_GLM_<surface> and tmtemp_GLM_volume")
    n = 36
    rng, exog, cov = _exog(n, 4)
    exog_flat = np.column_stack([exog[0], exog[1]])
    common = dict(exog_flat=exog_flat, exog_shape=np.array([1, 2]), varnames=np.array(["age", "group"]),
                  gstat=np.array("ft"), optstfce=np.array([2.0, 1.0]))
    # surface: two hemispheres of an order-3 icosphere with masks
    rp, col = synth.icosphere_csr(3)
    Vh = len(rp) - 1
    m_lh, m_rh = rng.random(Vh) < 0.9, rng.random(Vh) < 0.85
    nv = int(m_lh.sum() + m_rh.sum())
    data = rng.standard_normal((n, nv)).astype(np.float32)
    vd = adjacency.vertex_density(rp)
    d = os.path.join(tmp_path, "tmtemp_GLM_area")
    os.makedirs(d)
    for k, v in dict(common, data=data, dmy_covariates=cov, mask_lh=m_lh, mask_rh=m_rh,
                     num_vertex_lh=int(m_lh.sum()), adjac_lh=_ragged(rp, col), adjac_rh=_ragged(rp, col),
                     vdensity_lh=vd, vdensity_rh=vd).items():
        np.save(os.path.join(d, k), v)
    with pytest.raises(ValueError, match="allow_pickle"):
        step2.tm_glm(1, 4, surface="area", root=str(tmp_path))
    got = step2.tm_glm(1, 4, surface="area", root=str(tmp_path), seed=2, allow_pickle=True)
    perms = randomise.permutation_indices(n, range(1, 5), seed=2)
    ref = randomise.tm_glm(data, exog, cov, "ft", [2.0, 1.0], perms, adjac_lh=_ragged(rp, col),
                           adjac_rh=_ragged(rp, col), mask_lh=m_lh, mask_rh=m_rh, vdensity_lh=vd, vdensity_rh=vd,
                           varnames=["age", "group"])
    assert sorted(got) == sorted(ref)
    for name in ref:
        np.testing.assert_array_equal(got[name], ref[name])
        assert os.path.exists(os.path.join(tmp_path, "output_GLM_area", "perm_GLM",
                                           "perm_%s_TFCE_maxVertex.csv" % name))
    # volume, no covariates (np.save of None)
    mask = synth.ellipsoid_mask((10, 11, 10), (4.5, 5, 4.5), (4, 4.5, 4))
    adj = adjacency.create_adjac_voxel(mask, mask, int(mask.sum()), 26)
    obj = np.empty(len(adj), dtype=object)
    for i, l in enumerate(adj):
        obj[i] = l
    vdata = rng.standard_normal((n, int(mask.sum()))).astype(np.float32)
    d = os.path.join(tmp_path, "tmtemp_GLM_volume")
    os.makedirs(d)
    for k, v in dict(common, data=vdata, dmy_covariates=None, adjac=obj).items():
        np.save(os.path.join(d, k), v)
    got = step2.tm_glm(0, 3, root=str(tmp_path), seed=5, allow_pickle=True)
    perms = randomise.permutation_indices(n, range(0, 4), seed=5)
    ref = randomise.tm_glm(vdata, exog, None, "ft", [2.0, 1.0], perms, adjac=adj, varnames=["age", "group"])
    for name in ref:
        np.testing.assert_array_equal(got[name], ref[name])
        assert os.path.exists(os.path.join(tmp_path, "output_GLM_volume", "perm_GLM",
                                           "perm_%s_TFCE_maxVoxel.csv" % name))
