"""The tm-models mediation (the mediation branch of
tm_models_randomise.py:430-510: glm_typeI t-values with covariates, Aroian z
of calc_indirect, pyfunc.py:2678-2709) on the GPU against the FP64 oracle
loop (row 1 of oracle glm_typeI's t for each path, z rounded to float32, the
oracle TFCE), and step 2 from a ``tmtemp_mediation_*`` folder laid out as
the reference's step 1 saves it (written by this test, so it may be
unpickled)."""
import os

import numpy as np
import pytest

from oracle import ols as ools
from oracle import tfce as otfce
from tfce_mediation_amd import adjacency, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _vars(n, rng, left_cols=1, right_cols=1):
    def var(w):
        if w == 1:
            return rng.standard_normal(n)
        g = rng.integers(0, w + 1, n)
        return np.column_stack([g == j + 1 for j in range(w)]).astype(float)
    return var(left_cols), var(right_cols), rng.standard_normal((n, 2))


def oracle_z(Y, left, right, medtype, cov, perm):
    """Row 1 of glm_typeI's t for each path of tm_models_randomise.py:436-498
    with the permuted variables, and their Aroian z (FP64)."""
    perm = np.asarray(perm)
    lp = left[perm]
    if medtype == "I":
        ta = ools.glm_typeI(Y, [lp], dmy_covariates=cov)[2][1]
        tb = ools.glm_typeI(Y, [lp, right], dmy_covariates=cov)[2][1]
    elif medtype == "M":
        ta = ools.glm_typeI(Y, [lp], dmy_covariates=cov)[2][1]
        tb = ools.glm_typeI(Y, [right, lp], dmy_covariates=cov)[2][1]
    else:
        rp_ = right[perm]
        ta = ools.glm_typeI(rp_, [lp], dmy_covariates=cov)[2][1]
        tb = ools.glm_typeI(Y, [rp_, lp], dmy_covariates=cov)[2][1]
    return ools.sobel_from_t(ta, tb, "aroian")


def voxel_max(rp, col, z, H=2.0, E=1.0):
    z = np.where(np.isfinite(z), z, 0).astype(np.float32)
    enh, _, _ = otfce.tfce(rp, col, z, H, E, exact_pow=True)
    return enh.max() * (z.max() / np.float32(100))


def vertex_max(z, rp, col, m_lh, m_rh, vd_lh, vd_rh, H=2.0, E=1.0):
    """write_perm_maxTFCE_vertex (pyfunc.py:107-119) over two hemispheres of
    the same mesh."""
    nv = int(m_lh.sum())
    best = []
    for zh, m, vd in ((z[:nv], m_lh, vd_lh), (z[nv:], m_rh, vd_rh)):
        full = np.zeros(len(m), np.float32)
        full[m] = zh
        enh, _, _ = otfce.tfce(rp, col, full, H, E, exact_pow=True)
        best.append((enh * (full.max() / np.float32(100)) * vd).max())
    return max(best)


@pytest.mark.parametrize("medtype,with_cov", [("I", True), ("M", True), ("Y", True), ("Y", False), ("M", False)])
def test_tm_mediation_voxel_null_matches_oracle_loop(tmp_path, medtype, with_cov):
    from tfce_mediation_amd import randomise
    say(f"[gpu] randomise.tm_mediation voxel medtype={medtype} covariates={with_cov}")
    mask = synth.ellipsoid_mask((12, 12, 11), (5.5, 5.5, 5), (5, 5, 4.5))
    rp, col = adjacency.voxel_csr(mask, 26)
    V = int(mask.sum())
    n = 40
    rng = np.random.default_rng(7)
    left, right, cov = _vars(n, rng)
    if not with_cov:
        cov = None
    on = rng.random(V) < 0.4
    Y = (rng.standard_normal((n, V)) + 0.8 * left[:, None] * on + 0.8 * right[:, None] * on).astype(np.float32)
    perms = randomise.permutation_indices(n, range(1, 7), seed=11)
    null = randomise.tm_mediation(Y, left, right, medtype, cov, [2.0, 1.0], perms, adjac=(rp, col),
                                  out_dir=str(tmp_path), precision="fp32")
    assert null.shape == (6,)
    for p, perm in enumerate(perms):
        exp = voxel_max(rp, col, oracle_z(Y, left, right, medtype, cov, perm))
        assert abs(null[p] - exp) <= 1e-4 * abs(exp), (p, null[p], exp)
    vals = np.loadtxt(os.path.join(tmp_path, "perm_Zstat_%s_TFCE_maxVoxel.csv" % medtype))
    np.testing.assert_allclose(vals, null, atol=5e-5)


@pytest.mark.parametrize("medtype", ["I", "M"])
def test_tm_mediation_factor_columns(medtype):
    """Dummy-coded variables of several columns (row 1 is the first column
    of the first variable after the intercept), which take the unfused path."""
    from tfce_mediation_amd import randomise
    say(f"[gpu] randomise.tm_mediation factor columns medtype={medtype}")
    mask = synth.ellipsoid_mask((10, 11, 10), (4.5, 5, 4.5), (4, 4.5, 4))
    rp, col = adjacency.voxel_csr(mask, 26)
    V = int(mask.sum())
    n = 45
    rng = np.random.default_rng(8)
    left, right, cov = _vars(n, rng, left_cols=2, right_cols=2)
    Y = (rng.standard_normal((n, V)) + 0.9 * left[:, :1] * (rng.random(V) < 0.5)).astype(np.float32)
    perms = randomise.permutation_indices(n, range(3, 7), seed=1)
    null = randomise.tm_mediation(Y, left, right, medtype, cov[:, :1], [2.0, 1.0], perms, adjac=(rp, col),
                                  precision="fp32")
    for p, perm in enumerate(perms):
        exp = voxel_max(rp, col, oracle_z(Y, left, right, medtype, cov[:, :1], perm))
        assert abs(null[p] - exp) <= 1e-4 * abs(exp), (p, null[p], exp)


@pytest.mark.parametrize("medtype", ["I", "M"])
def test_fused_paths_with_covariates_match_separate_contractions(medtype):
    import torch
    from tfce_mediation_amd import randomise
    say(f"[gpu] MediationModel fused vs separate with covariates medtype={medtype}")
    n, V = 50, 3000
    rng = np.random.default_rng(9)
    left, right, cov = _vars(n, rng)
    Y = rng.standard_normal((n, V)).astype(np.float32)
    perms = torch.as_tensor(randomise.permutation_indices(n, range(20), seed=4)).cuda()
    fused = randomise.MediationModel(Y, left, right, medtype, covariates=cov)
    sep = randomise.MediationModel(Y, left, right, medtype, covariates=cov, fused=False)
    assert fused.fused and not sep.fused
    a, b = fused.z(perms).cpu().numpy(), sep.z(perms).cpu().numpy()
    np.testing.assert_array_equal(a, b)
    for p in (0, 7, 19):
        exp = oracle_z(Y, left, right, medtype, cov, perms[p].cpu().numpy())
        np.testing.assert_allclose(a[p], exp, rtol=2e-4, atol=2e-5)


def _ragged(row_ptr, col):
    out = np.empty(len(row_ptr) - 1, dtype=object)
    for i in range(len(row_ptr) - 1):
        out[i] = col[row_ptr[i]:row_ptr[i + 1]].tolist()
    return out


def test_tm_mediation_from_step1_folders(tmp_path):
    from tfce_mediation_amd import randomise, step2
    say("[gpu] step2.tm_mediation from tmtemp_mediation_<surface> and tmtemp_mediation_volume")
    n = 36
    rng = np.random.default_rng(10)
    left, right, cov = _vars(n, rng)
    # surface, type M with covariates: two masked hemispheres of an order-3 icosphere
    rp, col = synth.icosphere_csr(3)
    Vh = len(rp) - 1
    m_lh, m_rh = rng.random(Vh) < 0.9, rng.random(Vh) < 0.85
    nv = int(m_lh.sum() + m_rh.sum())
    on = rng.random(nv) < 0.5
    data = (rng.standard_normal((n, nv)) + 0.9 * left[:, None] * on + 0.9 * right[:, None] * on).astype(np.float32)
    vd = adjacency.vertex_density(rp)
    d = os.path.join(tmp_path, "tmtemp_mediation_area")
    os.makedirs(d)
    for k, v in dict(dmy_leftvar=left, dmy_rightvar=right, medtype=np.array("M"), data=data, dmy_covariates=cov,
                     optstfce=np.array([2.0, 1.0]), mask_lh=m_lh, mask_rh=m_rh, num_vertex_lh=int(m_lh.sum()),
                     adjac_lh=_ragged(rp, col), adjac_rh=_ragged(rp, col), vdensity_lh=vd,
                     vdensity_rh=vd).items():
        np.save(os.path.join(d, k), v)
    with pytest.raises(ValueError, match="allow_pickle"):
        step2.tm_mediation(1, 4, surface="area", root=str(tmp_path))
    got = step2.tm_mediation(1, 4, surface="area", root=str(tmp_path), seed=2, allow_pickle=True,
                             precision="fp32")
    perms = randomise.permutation_indices(n, range(1, 5), seed=2)
    for p, perm in enumerate(perms):
        z = oracle_z(data, left, right, "M", cov, perm).astype(np.float32)
        exp = vertex_max(z, rp, col, m_lh, m_rh, vd, vd)
        assert abs(got[p] - exp) <= 1e-4 * abs(exp), (p, got[p], exp)
    vals = np.loadtxt(os.path.join(tmp_path, "output_mediation_area", "perm_mediation",
                                   "perm_Zstat_M_TFCE_maxVertex.csv"))
    np.testing.assert_allclose(vals, got, atol=5e-5)
    # volume, type Y, no covariates (np.save of None)
    mask = synth.ellipsoid_mask((10, 11, 10), (4.5, 5, 4.5), (4, 4.5, 4))
    adj = adjacency.create_adjac_voxel(mask, mask, int(mask.sum()), 26)
    obj = np.empty(len(adj), dtype=object)
    for i, l in enumerate(adj):
        obj[i] = l
    vdata = rng.standard_normal((n, int(mask.sum()))).astype(np.float32)
    d = os.path.join(tmp_path, "tmtemp_mediation_volume")
    os.makedirs(d)
    for k, v in dict(dmy_leftvar=left, dmy_rightvar=right, medtype=np.array("Y"), data=vdata, dmy_covariates=None,
                     optstfce=np.array([2.0, 1.0]), adjac=obj).items():
        np.save(os.path.join(d, k), v)
    got = step2.tm_mediation(0, 3, root=str(tmp_path), seed=5, allow_pickle=True)
    perms = randomise.permutation_indices(n, range(0, 4), seed=5)
    ref = randomise.tm_mediation(vdata, left, right, "Y", None, [2.0, 1.0], perms, adjac=adj)
    np.testing.assert_array_equal(got, ref)
    assert os.path.exists(os.path.join(tmp_path, "output_mediation_volume", "perm_mediation",
                                       "perm_Zstat_Y_TFCE_maxVoxel.csv"))
