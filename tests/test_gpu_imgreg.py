"""Element-wise regressor and VIF image of the step-1 scripts
(pyfunc.image_regression / image_reg_VIF, the -v branch of
STEP_1_{vertex,voxel}_tfce_multiple_regression.py) on the GPU against the
oracle (oracle/imgreg.py, pinned in tests/test_oracle_imgreg.py).

Bars: t-values rtol 1e-6 (+ atol 1e-6 for t near 0) of the oracle's float64
per-element loop, both stored as float32; zeros where the regressor is flat or
holds a NaN; VIF rtol 1e-8 (both FP64 fits, the sum of squares formed in y's
own type the same way); TFCE images within 1 float ulp of the oracle's TFCE of
the same float32 statistic image, then the same numpy rescale.
"""
import os

import numpy as np
import pytest

from oracle import imgreg as oimg
from oracle import tfce as otfce
from tfce_mediation_amd import adjacency, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    return _t


def case(seed, V, n, p, c):
    rng = np.random.default_rng(seed)
    pred = rng.standard_normal((n, p))
    cov = rng.standard_normal((n, c)) if c else None
    img = (rng.standard_normal((V, n)) * 2 + 1).astype(np.float32)
    y = (0.5 * img + 0.3 * pred[:, :1].T + rng.standard_normal((V, n))).astype(np.float32)
    img[1] = 3.0               # flat
    img[2, 5] = np.nan         # NaN
    img[3] = (rng.standard_normal(n) * 0.002).astype(np.float32)  # std below 0.01
    return y, img, pred, cov


@pytest.mark.parametrize("V,n,p,c", [(3000, 60, 2, 2), (700, 2048, 2, 4), (500, 12, 1, 0)])
def test_image_regression_matches_oracle(torch, V, n, p, c):
    """Shared columns [1, pred, covars] plus the element's regressor; the
    2,048-subject case holds a 114,688-byte design in LDS."""
    from tfce_mediation_amd import pyfunc
    say(f"[gpu] image_regression V={V} n={n} pred={p} covars={c}")
    y, img, pred, cov = case(V + n, V, n, p, c)
    Tval, Timg = pyfunc.image_regression(y, img, pred, cov, verbose=False)
    Tref, Iref = oimg.image_regression(y, img, pred, cov)
    assert Tval.dtype == np.float32 and Tval.shape == Tref.shape and Timg.shape == Iref.shape
    np.testing.assert_allclose(Timg, Iref, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(Tval, Tref, rtol=1e-6, atol=1e-6)
    for i in (1, 2, 3):
        assert Timg[i] == 0 and np.all(Tval[i] == 0)


def test_image_regression_one_dimensional_pred_and_normalize(torch):
    """A one-dimensional pred_x keeps every column in Tval; normalize z-scales
    the regressor array along its first axis first."""
    from tfce_mediation_amd import pyfunc
    say("[gpu] image_regression 1-D pred_x, normalize")
    y, img, pred, cov = case(9, 800, 40, 1, 2)
    img[2, 5] = 0.0
    Tval, Timg = pyfunc.image_regression(y, img, pred[:, 0], cov, normalize=True, verbose=False)
    Tref, Iref = oimg.image_regression(y, img, pred[:, 0], cov, normalize=True)
    assert Tval.shape == (800, 5)
    np.testing.assert_allclose(Tval, Tref, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(Timg, Iref, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("dtype,n,k", [(np.float32, 80, 4), (np.float64, 80, 4), (np.float32, 2048, 7)])
def test_vif_matches_oracle(torch, dtype, n, k):
    from tfce_mediation_amd import pyfunc
    say(f"[gpu] image_reg_VIF {np.dtype(dtype).name} n={n} regressors={k}")
    rng = np.random.default_rng(n + k)
    V = 2500
    reg = rng.standard_normal((n, k))
    y = (reg @ rng.standard_normal((k, V)) * 0.5 + rng.standard_normal((n, V)) + 10).astype(dtype)
    y[:, 7] = 4.0  # RSS = TSS = 0: NaN, set to 0
    got = pyfunc.image_reg_VIF(y, reg)
    exp = oimg.image_reg_VIF(y, reg)
    assert got.dtype == np.float64 and got[7] == 0 and exp[7] == 0
    np.testing.assert_allclose(got, exp, rtol=1e-8)


def test_step1_vertex_regressor_images(torch, tmp_path):
    """STEP_1_vertex...:276-345 order and content: tstat_imgcovar and
    negtstat_imgcovar (TFCE) and VIF_imgcovar (statistic only) of both
    hemispheres, then tstat_con / negtstat_con from the image regression."""
    from tfce_mediation_amd import step1
    say("[gpu] step1.vertex_regression_images with a vertex-wise regressor, ico3 x2")
    row_ptr, col = synth.icosphere_csr(3)
    V = len(row_ptr) - 1
    rng = np.random.default_rng(21)
    m_lh, m_rh = rng.random(V) < 0.9, rng.random(V) < 0.85
    nl, nr = int(m_lh.sum()), int(m_rh.sum())
    n = 30
    pred = rng.standard_normal((n, 2))
    cov = rng.standard_normal((n, 1))
    img_x = rng.standard_normal((n, nl + nr)).astype(np.float32)
    y = (0.6 * img_x + 0.4 * pred[:, :1] + rng.standard_normal((n, nl + nr))).astype(np.float32)
    lists = [col[row_ptr[i]:row_ptr[i + 1]].tolist() for i in range(V)]
    out = step1.vertex_regression_images(y, pred, lists, lists, m_lh, m_rh, [2, 1], surface="area",
                                         out_dir=str(tmp_path), img_x=img_x, covars=cov)
    Tref, Iref = oimg.image_regression(y.T, img_x.T, pred, cov)
    vif = oimg.image_reg_VIF(y, np.column_stack((pred, cov)))
    parts = {"lh": slice(0, nl), "rh": slice(nl, nl + nr)}
    masks = {"lh": m_lh, "rh": m_rh}
    expect = {"tstat_imgcovar": Iref, "negtstat_imgcovar": -Iref, "VIF_imgcovar": vif}
    for j in (1, 2):
        expect["tstat_con%d" % j] = Tref[:, j]
        expect["negtstat_con%d" % j] = -Tref[:, j]
    assert set(name for name, _ in out) == set(expect)
    for (name, hemi), (stat_img, tfce_img) in out.items():
        m = masks[hemi]
        np.testing.assert_allclose(stat_img[m], expect[name][parts[hemi]].astype(np.float32), rtol=1e-6, atol=1e-6)
        assert np.all(stat_img[~m] == 0)
        assert os.path.exists(os.path.join(tmp_path, "%s_area_%s.npy" % (name, hemi)))
        has_tfce = os.path.exists(os.path.join(tmp_path, "%s_area_%s_TFCE.npy" % (name, hemi)))
        if name == "VIF_imgcovar":
            assert tfce_img is None and not has_tfce
            continue
        assert has_tfce
        masked = stat_img[m].astype(np.float64)
        enh, st, _ = otfce.tfce(row_ptr, col, stat_img, 2.0, 1.0, exact_pow=True)
        exp = (enh * (masked[np.isfinite(masked)].max() / 100)).astype(np.float32)
        d = np.abs(tfce_img.view(np.int32).astype(np.int64) - exp.view(np.int32).astype(np.int64))
        assert d.max() <= 1, (name, hemi)
    lines = [l.split(",")[0] for l in open(os.path.join(tmp_path, "max_TFCE_contrast_values.csv")).read().split()]
    assert lines[:4] == ["tstat_imgcovar_area_lh", "tstat_imgcovar_area_rh", "negtstat_imgcovar_area_lh",
                         "negtstat_imgcovar_area_rh"]
    assert lines[4] == "tstat_con1_area_lh" and len(lines) == 4 + 8


def test_step1_voxel_regressor_images(torch, tmp_path):
    """STEP_1_voxel...:162-212: tstat_imgcovar, negtstat_imgcovar (TFCE),
    VIF_imgcovar of the regressor (statistic only), then the contrasts with
    NaN t set to 0."""
    from tfce_mediation_amd import step1
    mask = synth.ellipsoid_mask((13, 15, 11), (6, 7, 5), (5.5, 6.5, 4.5))
    row_ptr, col = adjacency.voxel_csr(mask, 26)
    nv = int(mask.sum())
    say(f"[gpu] step1.voxel_regression_images with a voxel-wise regressor, voxels={nv}")
    n = 26
    rng = np.random.default_rng(8)
    pred = rng.standard_normal((n, 1))
    cov = rng.standard_normal((n, 2))
    image_x = rng.standard_normal((n, nv)).astype(np.float32)
    image_x[:, 10] = 1.0  # flat regressor: zeros
    y = (0.5 * image_x + rng.standard_normal((n, nv))).astype(np.float32)
    out = step1.voxel_regression_images(y, pred, (row_ptr, col), mask, [2, 0.5], out_dir=str(tmp_path),
                                        image_x=image_x, covars=cov)
    Tref, Iref = oimg.image_regression(y.T, image_x.T, pred, cov)
    vif = oimg.image_reg_VIF(image_x, np.column_stack((pred, cov)))
    np.testing.assert_allclose(out["tstat_imgcovar"][0], Iref, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(out["negtstat_imgcovar"][0], -Iref, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(out["VIF_imgcovar"][0], vif.astype(np.float32), rtol=1e-6)
    assert out["VIF_imgcovar"][1] is None
    np.testing.assert_allclose(out["tstat_con1"][0], Tref[:, 1], rtol=1e-6, atol=1e-6)
    assert out["tstat_con1"][0][10] == 0
    lines = [l.split(",")[0] for l in open(os.path.join(tmp_path, "max_TFCE_contrast_values.csv")).read().split()]
    assert lines == ["tstat_imgcovar", "negtstat_imgcovar", "tstat_con1", "negtstat_con1"]
