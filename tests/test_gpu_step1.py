"""Observed-statistic images (step1) and the batched F-statistic null against
the oracle.

Bars: float32 statistic images rtol 1e-6 of numpy FP64 tval_int / calcF;
TFCE images within 1 float ulp (E = 1, H = 2: the GPU's TFCE is bit-exact, the
ulp is the rescale by the FP64 maximum of the statistic, which the test only
has after the float32 cast) or 4 ulps (E = 0.5) of the oracle's TFCE of the
same float32 statistic image, then the same numpy rescale; 1 - p(FWE) equal
to the oracle's left-neighbour lookup.  sqrt(F) of the FP32 contraction
within rtol 1e-5 + atol 5e-6 of numpy FP64 calcF on the permuted design
(split-bf16x3: 2e-5 + 1e-5); the absolute part is there because F's numerator
TSS - SSE is a difference of sums of squares, so the float32 rounding of the
projections is an absolute error in F, large relative to a small F.  The
TFCE null of sqrt(F) within rtol 1e-4 (median 1e-6) of the oracle's
per-permutation loop.
"""
import os

import numpy as np
import pytest

from oracle import fwe as ofwe
from oracle import ols as ools
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


def ulp_diff(a, b):
    ia = a.astype(np.float32).view(np.int32).astype(np.int64)
    ib = b.astype(np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def smooth_rows(row_ptr, col, n, seed, passes=3):
    rng = np.random.default_rng(seed)
    V = len(row_ptr) - 1
    Y = rng.standard_normal((n, V))
    deg = np.diff(row_ptr).astype(np.int64)
    src = np.repeat(np.arange(V), deg)
    for _ in range(passes):
        S = np.zeros_like(Y)
        np.add.at(S.T, src, Y.T[col])
        Y = (Y + S) / (deg + 1)
    return ((Y - Y.mean(0)) / Y.std(0)).astype(np.float32)


def oracle_scaled_tfce(row_ptr, col, image, stat_masked, H, E, density=1):
    enh, st, _ = otfce.tfce(row_ptr, col, image, H, E, exact_pow=True)
    assert st >= 0
    return (enh * (stat_masked[np.isfinite(stat_masked)].max() / 100) * density).astype(np.float32)


def test_vertex_regression_images(torch, tmp_path):
    from tfce_mediation_amd import step1
    say("[gpu] step1.vertex_regression_images ico3 x2, n=30, k=3")
    row_ptr, col = synth.icosphere_csr(3)
    V = len(row_ptr) - 1
    rng = np.random.default_rng(11)
    m_lh, m_rh = rng.random(V) < 0.9, rng.random(V) < 0.85
    n = 30
    Y = np.concatenate([smooth_rows(row_ptr, col, n, 1)[:, m_lh], smooth_rows(row_ptr, col, n, 2)[:, m_rh]], 1)
    X = np.column_stack([np.ones(n), rng.standard_normal(n), (rng.random(n) < 0.5).astype(float)])
    Y[:, :40] += 0.8 * X[:, 2:3].astype(np.float32)
    lists = [col[row_ptr[i]:row_ptr[i + 1]].tolist() for i in range(V)]
    vd = adjacency.vertex_density(row_ptr)
    out = step1.vertex_regression_images(Y, X[:, 1:], lists, lists, m_lh, m_rh, [2, 1], vd, vd,
                                         surface="area", out_dir=str(tmp_path))
    invXX = np.linalg.inv(X.T @ X)
    t_ref = ools.tval_int(X, invXX, Y.astype(np.float64), n, 3, Y.shape[1])
    nl = int(m_lh.sum())
    assert len(out) == 2 * 2 * 2
    for (name, hemi), (stat_img, tfce_img) in out.items():
        tnum = int(name[-1])
        sgn = -1 if name.startswith("neg") else 1
        part = slice(0, nl) if hemi == "lh" else slice(nl, Y.shape[1])
        m = m_lh if hemi == "lh" else m_rh
        np.testing.assert_allclose(stat_img[m], (sgn * t_ref[tnum, part]).astype(np.float32), rtol=1e-6, atol=1e-7)
        assert np.all(stat_img[~m] == 0)
        # TFCE part: the oracle on the very statistic image the GPU saw
        masked = stat_img[m].astype(np.float64)
        exp = oracle_scaled_tfce(row_ptr, col, stat_img, masked, 2.0, 1.0, vd)
        bad = np.nonzero(ulp_diff(tfce_img, exp) > 1)[0]
        assert bad.size == 0, (name, hemi, bad[:5], tfce_img[bad[:5]], exp[bad[:5]])
        for suffix in ("_TFCE.npy", ".npy"):
            assert os.path.exists(os.path.join(tmp_path, "%s_area_%s%s" % (name, hemi, suffix)))
    lines = open(os.path.join(tmp_path, "max_TFCE_contrast_values.csv")).read().split()
    assert len(lines) == 8 and lines[0].startswith("tstat_con1_area_lh,")
    got = {l.split(",")[0]: float(l.split(",")[1]) for l in lines}
    assert got["tstat_con2_area_lh"] == pytest.approx(float(out[("tstat_con2", "lh")][1].max()), abs=1e-6)


@pytest.mark.parametrize("ancova", [False, True])
def test_voxel_regression_images(torch, tmp_path, ancova):
    from tfce_mediation_amd import step1
    mask = synth.ellipsoid_mask((15, 17, 13), (7, 8, 6), (6, 7, 5.5))
    row_ptr, col = adjacency.voxel_csr(mask, 26)
    nv = int(mask.sum())
    say(f"[gpu] step1.voxel_regression_images voxels={nv} ancova={ancova}")
    n = 24
    rng = np.random.default_rng(5)
    y = smooth_rows(row_ptr, col, n, 9)
    X = np.column_stack([np.ones(n), rng.standard_normal(n), rng.standard_normal(n)])
    y[:, :200] += 0.7 * X[:, 1:2].astype(np.float32)
    out = step1.voxel_regression_images(y, X[:, 1:], (row_ptr, col), mask, [2, 0.5], out_dir=str(tmp_path),
                                        ancova=ancova)
    if ancova:
        F = ools.calcF(X, y.astype(np.float64), n, 3)
        F[F < 0] = 0
        stats = {"fstat": np.sqrt(F)}
    else:
        t = ools.tval_int(X, np.linalg.inv(X.T @ X), y.astype(np.float64), n, 3, nv)
        t[np.isnan(t)] = 0
        stats = {}
        for tnum in (1, 2):
            stats["tstat_con%d" % tnum] = t[tnum]
            stats["negtstat_con%d" % tnum] = -t[tnum]
    assert sorted(out) == sorted(stats)
    for name, (stat, tfce_vals) in out.items():
        np.testing.assert_allclose(stat, stats[name].astype(np.float32), rtol=1e-6, atol=1e-6)
        img = stat.astype(np.float32)
        enh, st, _ = otfce.tfce(row_ptr, col, img, 2.0, 0.5, exact_pow=True)
        exp = enh * (img.max() / 100)
        assert ulp_diff(tfce_vals, exp).max() <= 4, name
        assert os.path.exists(os.path.join(tmp_path, "%s_TFCE.npy" % name))
        vol = np.load(os.path.join(tmp_path, "%s.npy" % name))
        np.testing.assert_array_equal(vol[mask], stat)


@pytest.mark.parametrize("medtype", ["M", "I"])
def test_vertex_mediation_images(torch, tmp_path, medtype):
    """STEP_1_vertex_tfce_mediation.py:279-288: SobelZ_<medtype> of both
    hemispheres, the statistic against the oracle's calc_sobelz, the TFCE
    image within 1 ulp of the oracle's TFCE of the same float32 image."""
    from tfce_mediation_amd import step1
    say(f"[gpu] step1.vertex_mediation_images {medtype} ico3 x2")
    row_ptr, col = synth.icosphere_csr(3)
    V = len(row_ptr) - 1
    rng = np.random.default_rng(31)
    m_lh, m_rh = rng.random(V) < 0.9, rng.random(V) < 0.8
    n = 40
    x = rng.standard_normal(n)
    Y = np.concatenate([smooth_rows(row_ptr, col, n, 3)[:, m_lh], smooth_rows(row_ptr, col, n, 4)[:, m_rh]], 1)
    Y[:, :60] += 0.6 * x[:, None].astype(np.float32)
    dep = 0.5 * Y[:, :60].mean(1) + rng.standard_normal(n)
    lists = [col[row_ptr[i]:row_ptr[i + 1]].tolist() for i in range(V)]
    out = step1.vertex_mediation_images(Y, x, dep, medtype, lists, lists, m_lh, m_rh, [2, 1], surface="area",
                                        out_dir=str(tmp_path))
    z = ools.calc_sobelz(medtype, x, dep, Y.astype(np.float64), n, Y.shape[1])
    nl = int(m_lh.sum())
    assert sorted(out) == [("SobelZ_%s" % medtype, "lh"), ("SobelZ_%s" % medtype, "rh")]
    for (name, hemi), (stat_img, tfce_img) in out.items():
        m = m_lh if hemi == "lh" else m_rh
        part = slice(0, nl) if hemi == "lh" else slice(nl, Y.shape[1])
        np.testing.assert_allclose(stat_img[m], z[part].astype(np.float32), rtol=1e-6)
        masked = stat_img[m].astype(np.float64)
        exp = oracle_scaled_tfce(row_ptr, col, stat_img, masked, 2.0, 1.0)
        assert ulp_diff(tfce_img, exp).max() <= 1, hemi
        for suffix in ("_TFCE.npy", ".npy"):
            assert os.path.exists(os.path.join(tmp_path, "%s_area_%s%s" % (name, hemi, suffix)))
    lines = open(os.path.join(tmp_path, "max_TFCE_contrast_values.csv")).read().split()
    assert [l.split(",")[0] for l in lines] == ["SobelZ_%s_area_lh" % medtype, "SobelZ_%s_area_rh" % medtype]


def test_voxel_mediation_images(torch, tmp_path):
    """STEP_1_voxel_tfce_mediation.py:97-104 (type M, 26-connected, E = 0.5)."""
    from tfce_mediation_amd import step1
    mask = synth.ellipsoid_mask((15, 17, 13), (7, 8, 6), (6, 7, 5.5))
    row_ptr, col = adjacency.voxel_csr(mask, 26)
    nv = int(mask.sum())
    say(f"[gpu] step1.voxel_mediation_images M voxels={nv}")
    n = 28
    rng = np.random.default_rng(6)
    x = rng.standard_normal(n)
    y = smooth_rows(row_ptr, col, n, 12)
    y[:, :150] += 0.7 * x[:, None].astype(np.float32)
    dep = 0.6 * y[:, :150].mean(1) + rng.standard_normal(n)
    out = step1.voxel_mediation_images(y, x, dep, "M", (row_ptr, col), mask, [2, 0.5], out_dir=str(tmp_path))
    z = ools.calc_sobelz("M", x, dep, y.astype(np.float64), n, nv)
    stat, tfce_vals = out["SobelZ_M"]
    np.testing.assert_allclose(stat, z.astype(np.float32), rtol=1e-6)
    img = stat.astype(np.float32)
    enh, st, _ = otfce.tfce(row_ptr, col, img, 2.0, 0.5, exact_pow=True)
    assert ulp_diff(tfce_vals, enh * (img.max() / 100)).max() <= 4
    assert os.path.exists(os.path.join(tmp_path, "SobelZ_M_TFCE.npy"))


def test_fwep_image(torch):
    from tfce_mediation_amd import step1
    say("[gpu] step1.fwep_image")
    rng = np.random.default_rng(2)
    img = np.where(rng.random(500) < 0.7, rng.gamma(2.0, 3.0, 500), 0.0).astype(np.float32)
    null = rng.gamma(2.0, 4.0, 200)
    p, nl = step1.fwep_image(img, null, neglog10=True)
    exp = np.zeros(500)
    sel = img > 0
    exp[sel] = ofwe.fwep_map(img[sel], null)
    np.testing.assert_array_equal(p, exp)
    assert np.all(p[~sel] == 0) and np.all(nl[~sel] == 0)
    np.testing.assert_allclose(nl[sel], -np.log10(1 - exp[sel]))


@pytest.mark.parametrize("precision,rtol,atol", [("fp32", 1e-5, 5e-6), ("bf16x3", 2e-5, 1e-5)])
def test_perm_fstat_matches_calcF(torch, precision, rtol, atol):
    from tfce_mediation_amd import randomise
    from tfce_mediation_amd.ols import PermutedOLS, total_ss
    say(f"[gpu] PermutedOLS.fstat {precision}")
    rng = np.random.default_rng(8)
    n, V = 60, 3000
    X = np.column_stack([np.ones(n), rng.standard_normal(n), rng.standard_normal(n), rng.random(n) < 0.5])
    Y = (rng.standard_normal((n, V)) + 0.3).astype(np.float32)
    Y[:, :500] += 0.5 * X[:, 1:2].astype(np.float32)
    perms = randomise.permutation_indices(n, range(1, 21), seed=3)
    model = PermutedOLS(Y, X, precision=precision)
    got = model.fstat(torch.as_tensor(perms).cuda(), total_ss(Y)).cpu().numpy()
    for p, perm in enumerate(perms):
        F = ools.calcF(X[perm], Y.astype(np.float64), n, 4)
        F[F < 0] = 0
        np.testing.assert_allclose(got[p], np.sqrt(F), rtol=rtol, atol=atol)


def test_voxel_randomise_fstat_null(torch, tmp_path):
    from tfce_mediation_amd import randomise
    mask = synth.ellipsoid_mask((15, 17, 13), (7, 8, 6), (6, 7, 5.5))
    row_ptr, col = adjacency.voxel_csr(mask, 26)
    n = 30
    rng = np.random.default_rng(4)
    y = smooth_rows(row_ptr, col, n, 12)
    X = np.column_stack([np.ones(n), rng.standard_normal(n), rng.standard_normal(n)])
    perms = randomise.permutation_indices(n, range(1, 11), seed=6)
    say(f"[gpu] randomise.voxel_regress ancova voxels={len(row_ptr)-1}")
    nulls = randomise.voxel_regress(y, X[:, 1:], (row_ptr, col), [2, 0.5], perms, out_dir=str(tmp_path),
                                    precision="fp32", ancova=True)
    got = nulls["fstat"]
    exp = np.zeros_like(got)
    for p, perm in enumerate(perms):
        F = ools.calcF(X[perm], y.astype(np.float64), n, 3)
        F[F < 0] = 0
        s = np.sqrt(F).astype(np.float32)
        enh, _, _ = otfce.tfce(row_ptr, col, s, 2.0, 0.5, exact_pow=True)
        exp[p] = ofwe.perm_max_voxel(enh, s)
    rel = np.abs(got - exp) / np.abs(exp)
    say(f"    max rel {rel.max():.2e} median {np.median(rel):.2e}")
    assert rel.max() < 1e-4 and np.median(rel) < 1e-6
    vals = np.loadtxt(os.path.join(tmp_path, "perm_fstat_TFCE_maxVoxel.csv"))
    np.testing.assert_allclose(vals, got, atol=5e-5)
