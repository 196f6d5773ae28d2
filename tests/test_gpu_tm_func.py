"""The mmr / mmr-lr drop-ins of tm_func (tm_func.py:54-378) and the mmr-lr
mode of the batched driver, against the oracle.

Observed maps: the drop-ins compute t (or z) in FP64 on the GPU; those must
match the numpy FP64 oracle to rtol 1e-6 after float32 rounding, and the TFCE
maps they return must equal the oracle's TFCE of the returned float32
statistic with the reference's rescale (bit for bit at E = 1, 4 ulp at
E = 0.5).  Permutation mode (perm_seed given, tm_func.py:147-149): the values
appended to the null files ("%f") against the oracle loop on the same
permutation.  mmr-lr's thresholds are per surface (each surface's own image
and maximum, tm_func.py:160-174), which the batched driver reproduces with
``mmr(merged=False)``.
"""
import os

import numpy as np
import pytest

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


def ulp(a, b):
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    b = np.ascontiguousarray(b, np.float32).reshape(-1)
    return int(np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64)).max())


def scaled_tfce(csr, img, H, E):
    """oracle TFCE of a full float32 image, as the reference rescales it
    later: returns the raw enhancement."""
    enh, _, _ = otfce.tfce(csr[0], csr[1], np.ascontiguousarray(img, np.float32), H, E, exact_pow=True)
    return enh


def design(n, k, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.standard_normal(n) for _ in range(k - 2)] + [(rng.random(n) < 0.5).astype(float)])


def smooth(row_ptr, col, n, seed):
    rng = np.random.default_rng(seed)
    V = len(row_ptr) - 1
    Y = rng.standard_normal((n, V))
    deg = np.diff(row_ptr).astype(np.int64)
    src = np.repeat(np.arange(V), deg)
    for _ in range(3):
        S = np.zeros_like(Y)
        np.add.at(S.T, src, Y.T[col])
        Y = (Y + S) / (deg + 1)
    return ((Y - Y.mean(0)) / Y.std(0)).astype(np.float32)


def test_low_ram_calculate_tfce(torch, tmp_path):
    from tfce_mediation_amd import tm_func
    from tfce_mediation_amd.tfce import CreateAdjSet
    say("[gpu] tm_func.low_ram_calculate_tfce")
    rp, col = synth.icosphere_csr(4)
    V = len(rp) - 1
    mask = np.random.default_rng(1).random(V) < 0.9
    n = 36
    data = smooth(rp, col, n, 2)[:, mask]
    pred = design(n, 4, 3)
    vd = adjacency.vertex_density(rp)[mask]
    op = CreateAdjSet(2.0, 1.0, (rp, col))
    tv, tp, tn = tm_func.low_ram_calculate_tfce(data, mask.astype(np.float32), pred, op, vd)
    X = np.column_stack([np.ones(n), pred])
    t64 = ools.tval_int(X, np.linalg.inv(X.T @ X), data.astype(np.float64), n, X.shape[1], data.shape[1])[1:]
    np.testing.assert_allclose(tv, t64.astype(np.float32), rtol=1e-6, atol=1e-6)
    for c in range(tv.shape[0]):
        img = np.zeros(V, np.float32)
        img[mask] = tv[c]
        for sg, got in ((1, tp[c]), (-1, tn[c])):
            x = (sg * img).astype(np.float32)
            exp = scaled_tfce((rp, col), x, 2.0, 1.0)[mask] * (x.max() / np.float32(100)) * vd
            np.testing.assert_array_equal(got, exp.astype(np.float32))
    # permutation mode: np.random.seed(perm_number + perm_seed), files in output_dir
    for p in (3, 4):
        tm_func.low_ram_calculate_tfce(data, mask.astype(np.float32), pred, op, vd, set_surf_count=2,
                                       perm_number=p, randomise=True, output_dir=str(tmp_path), perm_seed=100)
    for c in range(tv.shape[0]):
        vals = np.loadtxt(os.path.join(tmp_path, "perm_maxTFCE_surf2_tcon%d.csv" % (c + 1)))
        assert vals.shape == (4,)
        for i, p in enumerate((3, 4)):
            perm = synth.reference_permutation(p, 100, n)
            t = ools.perm_t(X, data, [perm], [c + 1])[0, 0].astype(np.float32)
            img = np.zeros(V, np.float32)
            img[mask] = t
            for k, sg in enumerate((1, -1)):
                x = (sg * img).astype(np.float32)
                exp = (scaled_tfce((rp, col), x, 2.0, 1.0)[mask] * (x.max() / np.float32(100)) * vd).max()
                assert abs(vals[2 * i + k] - exp) <= 1e-4 * abs(exp) + 2e-6, (c, p, k, vals[2 * i + k], exp)


def test_calc_mixed_tfce_groups(torch):
    from tfce_mediation_amd import tm_func
    from tfce_mediation_amd.tfce import CreateAdjSet
    say("[gpu] tm_func.calc_mixed_tfce")
    surfs = [synth.icosphere_csr(3), adjacency.voxel_csr(synth.ellipsoid_mask((11, 12, 10), (5, 5.5, 4.5),
                                                                              (4.5, 5, 4)), 26),
             synth.icosphere_csr(2)]
    sizes = [len(s[0]) - 1 for s in surfs]
    masks = [np.ones((sz, 1, 1), bool) for sz in sizes]
    pos = tm_func.create_position_array(masks)
    n = 30
    rng = np.random.default_rng(5)
    Y = rng.standard_normal((n, sum(sizes))).astype(np.float32)
    pred = design(n, 3, 6)
    settings = np.array([0, 1, 0])
    he = {0: (2.0, 1.0), 1: (2.0, 0.5)}
    groups = {g: [i for i in range(3) if settings[i] == g] for g in (0, 1)}
    ops, merged = {}, {}
    for g, idx in groups.items():
        merged[g] = adjacency.merge_csr([surfs[i] for i in idx])
        ops[g] = CreateAdjSet(*he[g], merged[g])
    tv, tp, tn = tm_func.calc_mixed_tfce(settings, Y, masks, pos, 1, pred, ops)
    X = np.column_stack([np.ones(n), pred])
    t64 = ools.tval_int(X, np.linalg.inv(X.T @ X), Y.astype(np.float64), n, 3, Y.shape[1])[1:]
    np.testing.assert_allclose(tv, t64.astype(np.float32), rtol=1e-6, atol=1e-6)
    for g, idx in groups.items():
        cols = np.concatenate([np.arange(pos[i], pos[i + 1]) for i in idx])
        for c in range(2):
            img = tv[c][cols].astype(np.float32)
            offs = np.cumsum([0] + [sizes[i] for i in idx])
            for sg, got in ((1, tp), (-1, tn)):
                x = (sg * img).astype(np.float32)
                enh = scaled_tfce(merged[g], x, *he[g])
                for j, i in enumerate(idx):
                    a, b = offs[j], offs[j + 1]
                    exp = (enh[a:b] * (x[a:b].max() / np.float32(100))).astype(np.float32)
                    gv = got[c, pos[i]:pos[i + 1]].astype(np.float32)
                    if he[g][1] == 1.0:
                        np.testing.assert_array_equal(gv, exp)
                    else:
                        assert ulp(gv, exp) <= 4


def test_mediation_drop_ins(torch, tmp_path):
    from tfce_mediation_amd import tm_func
    from tfce_mediation_amd.tfce import CreateAdjSet
    say("[gpu] tm_func.calculate_mediation_tfce / low_ram_calculate_mediation_tfce")
    surfs = [synth.icosphere_csr(3), synth.icosphere_csr(3)]
    sizes = [len(s[0]) - 1 for s in surfs]
    masks = [np.ones((sz, 1, 1), bool) for sz in sizes]
    pos = tm_func.create_position_array(masks)
    full = tm_func.create_full_mask(masks)
    mrp, mcol = adjacency.merge_csr(surfs)
    n = 48
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    M = np.concatenate([smooth(*surfs[0], n, 8), smooth(*surfs[1], n, 9)], 1)
    M[:, :60] += (0.6 * x[:, None]).astype(np.float32)
    op = CreateAdjSet(2.0, 1.0, (mrp, mcol))
    z, tz = tm_func.calculate_mediation_tfce("M", M, masks, x, y, op, 1, pos, full)
    z64 = ools.calc_sobelz("M", x, y, M.astype(np.float64), n, M.shape[1])
    np.testing.assert_allclose(z, z64.astype(np.float32), rtol=2e-6, atol=1e-6)
    enh = scaled_tfce((mrp, mcol), z, 2.0, 1.0)
    for s in range(2):
        a, b = pos[s], pos[s + 1]
        np.testing.assert_array_equal(tz[a:b], (enh[a:b] * (z[a:b].max() / np.float32(100))).astype(np.float32))
    # low-RAM form on one surface with a mask; then permutation mode
    rp, col = surfs[0]
    mask = np.random.default_rng(2).random(sizes[0]) < 0.8
    data = M[:, :sizes[0]][:, mask]
    op1 = CreateAdjSet(2.0, 1.0, (rp, col))
    zz, tzz = tm_func.low_ram_calculate_mediation_tfce("I", data, mask.astype(np.float32), x, y, op1, 1)
    z64 = ools.calc_sobelz("I", x, y, data.astype(np.float64), n, data.shape[1])
    np.testing.assert_allclose(zz, z64.astype(np.float32), rtol=2e-6, atol=1e-6)
    img = np.zeros(sizes[0], np.float32)
    img[mask] = zz
    np.testing.assert_array_equal(tzz, (scaled_tfce((rp, col), img, 2.0, 1.0)[mask]
                                        * (zz.max() / np.float32(100))).astype(np.float32))
    for p in (1, 2, 3):
        tm_func.low_ram_calculate_mediation_tfce("M", data, mask.astype(np.float32), x, y, op1, 1,
                                                 set_surf_count=0, perm_number=p, randomise=True,
                                                 output_dir=str(tmp_path), perm_seed=7)
    vals = np.loadtxt(os.path.join(tmp_path, "perm_maxTFCE_surf0_M_zstat.csv"))
    for i, p in enumerate((1, 2, 3)):
        perm = synth.reference_permutation(p, 7, n)
        zp = ools.calc_sobelz("M", x[perm], y, data.astype(np.float64), n, data.shape[1]).astype(np.float32)
        img = np.zeros(sizes[0], np.float32)
        img[mask] = zp
        exp = (scaled_tfce((rp, col), img, 2.0, 1.0)[mask] * (zp.max() / np.float32(100))).max()
        assert abs(vals[i] - exp) <= 1e-4 * abs(exp) + 2e-6, (p, vals[i], exp)


def test_mmr_lr_per_surface_thresholds(torch, tmp_path):
    """mmr(merged=False): each surface's own threshold sequence and rescale,
    the loop of low_ram_calculate_tfce over surfaces (tm_func.py:144-185)."""
    from tfce_mediation_amd import randomise
    say("[gpu] randomise.mmr(merged=False)")
    surfs = [synth.icosphere_csr(3), synth.icosphere_csr(2),
             adjacency.voxel_csr(synth.ellipsoid_mask((10, 10, 9), (4.5, 4.5, 4), (4, 4, 3.5)), 26)]
    settings = [(2.0, 1.0), (2.0, 1.0), (2.0, 0.5)]
    sizes = [len(s[0]) - 1 for s in surfs]
    offs = list(np.cumsum([0] + sizes[:-1]))
    n = 32
    rng = np.random.default_rng(11)
    Y = rng.standard_normal((n, sum(sizes))).astype(np.float32)
    pred = design(n, 3, 12)
    X = np.column_stack([np.ones(n), pred])
    perms = randomise.permutation_indices(n, range(1, 7), seed=50)
    res = randomise.mmr(Y, pred, surfs, offs, settings, perms, merged=False, precision="fp32",
                        out_dir=str(tmp_path))
    worst = 0.0
    for c in (1, 2):
        for p, perm in enumerate(perms):
            t = ools.perm_t(X, Y, [perm], [c])[0, 0].astype(np.float32)
            for s in range(3):
                a, b = offs[s], offs[s] + sizes[s]
                for k, sg in enumerate((1, -1)):
                    x = (sg * t[a:b]).astype(np.float32)
                    exp = (scaled_tfce(surfs[s], x, *settings[s]) * (x.max() / np.float32(100))).max()
                    got = res[(s, c)][p, k]
                    rel = abs(got - exp) / abs(exp)
                    worst = max(worst, rel)
                    assert rel <= 1e-4, (c, p, s, k, got, exp)
    say(f"    max rel vs the oracle loop: {worst:.2e}")
    vals = np.loadtxt(os.path.join(tmp_path, "perm_maxTFCE_surf1_tcon2.csv"))
    np.testing.assert_allclose(vals, res[(1, 2)].ravel(), atol=1e-6)
