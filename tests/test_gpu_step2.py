"""Step 2 from step-1 hand-off directories laid out as the reference's step 1
writes them (np.save of the same variables, ragged adjacency as object
arrays; written by this test, so they may be unpickled): the files appended
where the reference appends them, equal to the batched drivers called
directly, and the voxel mediation driver against the oracle loop."""
import os

import numpy as np
import pytest

from oracle import ols as ools
from oracle import tfce as otfce
from tfce_mediation_amd import adjacency, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def ragged(row_ptr, col):
    lists = [col[row_ptr[i]:row_ptr[i + 1]].tolist() for i in range(len(row_ptr) - 1)]
    out = np.empty(len(lists), dtype=object)
    for i, l in enumerate(lists):
        out[i] = l
    return out


def save(d, **arrays):
    os.makedirs(d, exist_ok=True)
    for k, v in arrays.items():
        np.save(os.path.join(d, k), v)


def test_vertex_regression_from_step1_directory(tmp_path):
    import torch
    from tfce_mediation_amd import randomise, step2
    say("[gpu] step2.vertex_regression from python_temp_<surface>")
    rp, col = synth.icosphere_csr(3)
    V = len(rp) - 1
    rng = np.random.default_rng(1)
    m_lh, m_rh = rng.random(V) < 0.9, rng.random(V) < 0.85
    n = 30
    merge_y = rng.standard_normal((n, int(m_lh.sum() + m_rh.sum()))).astype(np.float32)
    pred_x = np.column_stack([rng.standard_normal(n), (rng.random(n) < 0.5).astype(float)])
    vd = adjacency.vertex_density(rp)
    save(os.path.join(tmp_path, "python_temp_area"), pred_x=pred_x, num_subjects=n, all_vertex=V,
         num_vertex=merge_y.shape[1], num_vertex_lh=int(m_lh.sum()), num_vertex_rh=int(m_rh.sum()),
         bin_mask_lh=m_lh, bin_mask_rh=m_rh, adjac_lh=ragged(rp, col), adjac_rh=ragged(rp, col),
         merge_y=merge_y, optstfce=np.array([2.0, 1.0]), vdensity_lh=vd, vdensity_rh=vd)
    with pytest.raises(ValueError, match="allow_pickle"):
        step2.vertex_regression("area", 1, 8, root=str(tmp_path))
    got = step2.vertex_regression("area", 1, 8, root=str(tmp_path), seed=3, allow_pickle=True)
    perms = randomise.permutation_indices(n, range(1, 9), seed=3)
    ref = randomise.vertex_regress(merge_y, pred_x, ragged(rp, col), ragged(rp, col), m_lh, m_rh, [2.0, 1.0],
                                   perms, vdensity_lh=vd, vdensity_rh=vd)
    for c in (1, 2):
        np.testing.assert_array_equal(got[c], ref[c])
        vals = np.loadtxt(os.path.join(tmp_path, "output_area", "perm_Tstat_area",
                                       "perm_tstat_con%d_TFCE_maxVertex.csv" % c))
        np.testing.assert_allclose(vals, got[c].ravel(), atol=5e-5)
    torch.cuda.synchronize()


def test_voxel_regression_and_ancova_from_step1_directory(tmp_path):
    from tfce_mediation_amd import randomise, step2
    say("[gpu] step2.voxel_regression from python_temp (t and F branches)")
    mask = synth.ellipsoid_mask((12, 13, 11), (5.5, 6, 5), (5, 5.5, 4.5))
    adj = adjacency.create_adjac_voxel(mask, mask, int(mask.sum()), 26)
    rng = np.random.default_rng(2)
    n = 28
    y = rng.standard_normal((int(mask.sum()), n)).astype(np.float32)
    pred_x = np.column_stack([rng.standard_normal(n), rng.standard_normal(n)])
    obj = np.empty(len(adj), dtype=object)
    for i, l in enumerate(adj):
        obj[i] = l
    for ancova in (0, 1):
        root = os.path.join(tmp_path, "a%d" % ancova)
        save(os.path.join(root, "python_temp"), num_voxel=y.shape[0], num_subjects=n, raw_nonzero_corr=y,
             pred_x=pred_x, adjac=obj, ancova=ancova, optstfce=np.array([2.0, 0.5]))
        got = step2.voxel_regression(2, 5, root=root, seed=9, allow_pickle=True)
        numbers = range(2, 11) if ancova else range(2, 6)
        perms = randomise.permutation_indices(n, numbers, seed=9)
        ref = randomise.voxel_regress(y.T, pred_x, adj, [2.0, 0.5], perms, ancova=bool(ancova))
        for key in ref:
            np.testing.assert_array_equal(got[key], ref[key])
        name = "perm_fstat_TFCE_maxVoxel.csv" if ancova else "perm_tstat_con1_TFCE_maxVoxel.csv"
        assert os.path.exists(os.path.join(root, "output", "perm_Tstat", name))


def test_voxel_mediation_driver_and_directory(tmp_path):
    from tfce_mediation_amd import randomise, step2
    say("[gpu] randomise.voxel_mediation / step2.voxel_mediation")
    mask = synth.ellipsoid_mask((12, 12, 11), (5.5, 5.5, 5), (5, 5, 4.5))
    rp, col = adjacency.voxel_csr(mask, 26)
    V = int(mask.sum())
    rng = np.random.default_rng(4)
    n = 40
    x = rng.standard_normal(n)
    dy = rng.standard_normal(n)
    M = (rng.standard_normal((n, V)) + 0.5 * x[:, None] * (rng.random(V) < 0.3)).astype(np.float32)
    perms = randomise.permutation_indices(n, range(0, 6), seed=1)
    for mt in ("M", "I"):
        null = randomise.voxel_mediation(M, x, dy, mt, (rp, col), [2.0, 1.0], perms, precision="fp32")
        for p, perm in enumerate(perms):
            z = ools.calc_sobelz(mt, x[perm], dy, M.astype(np.float64), n, V).astype(np.float32)
            enh, _, _ = otfce.tfce(rp, col, z, 2.0, 1.0, exact_pow=True)
            exp = enh.max() * (z.max() / np.float32(100))
            assert abs(null[p] - exp) <= 1e-4 * abs(exp), (mt, p, null[p], exp)
    obj = np.empty(V, dtype=object)
    for i in range(V):
        obj[i] = col[rp[i]:rp[i + 1]].tolist()
    save(os.path.join(tmp_path, "python_temp"), num_voxel=V, num_subjects=n, raw_nonzero_corr=M.T.copy(),
         pred_x=x, depend_y=dy, adjac=obj, medtype="M", optstfce=np.array([2.0, 1.0]))
    got = step2.voxel_mediation("M", 0, 5, root=str(tmp_path), seed=1, allow_pickle=True)
    ref = randomise.voxel_mediation(M, x, dy, "M", (rp, col), [2.0, 1.0], perms)
    np.testing.assert_array_equal(got, ref)
    vals = np.loadtxt(os.path.join(tmp_path, "output_med_M", "perm_SobelZ", "perm_Zstat_M_TFCE_maxVoxel.csv"))
    np.testing.assert_allclose(vals, got, atol=5e-5)


def test_vertex_mediation_from_step1_directory(tmp_path):
    from tfce_mediation_amd import randomise, step2
    say("[gpu] step2.vertex_mediation from python_temp_med_<surface>")
    rp, col = synth.icosphere_csr(3)
    V = len(rp) - 1
    rng = np.random.default_rng(6)
    m = rng.random(V) < 0.9
    n = 36
    merge_y = rng.standard_normal((n, 2 * int(m.sum()))).astype(np.float32)
    x, dy = rng.standard_normal(n), rng.standard_normal(n)
    save(os.path.join(tmp_path, "python_temp_med_area"), pred_x=x, depend_y=dy, num_subjects=n,
         num_vertex=merge_y.shape[1], num_vertex_lh=int(m.sum()), num_vertex_rh=int(m.sum()), all_vertex=V,
         bin_mask_lh=m, bin_mask_rh=m, adjac_lh=ragged(rp, col), adjac_rh=ragged(rp, col),
         optstfce=np.array([2.0, 1.0]), vdensity_lh=1, vdensity_rh=1, merge_y=merge_y)
    got = step2.vertex_mediation("area", "I", 0, 4, root=str(tmp_path), seed=2, allow_pickle=True)
    perms = randomise.permutation_indices(n, range(0, 5), seed=2)
    ref = randomise.vertex_mediation(merge_y, x, dy, "I", ragged(rp, col), ragged(rp, col), m, m, [2.0, 1.0], perms)
    np.testing.assert_array_equal(got, ref)
    assert os.path.exists(os.path.join(tmp_path, "output_med_area", "perm_SobelZ_I",
                                       "perm_Zstat_I_TFCE_maxVertex.csv"))


def test_mmr_from_a_tmi_file(tmp_path):
    """A TMI file (masked mesh + volume mask, faces, data) through
    tm_io.mmr_inputs into randomise.mmr (merged thresholds: the masked
    vertices' zeros enter the merged image's maximum) against the oracle on
    the reference's merged full-mask image (tm_func.py:54-98)."""
    from tfce_mediation_amd import randomise, tm_io
    say("[gpu] mmr from a TMI file")
    rng = np.random.default_rng(8)
    verts, faces = synth.icosphere(3)
    keep = rng.random(len(verts)) < 0.8
    vol = synth.ellipsoid_mask((10, 11, 9), (4.5, 5, 4), (4.2, 4.6, 3.8))
    nl, nv = int(keep.sum()), int(vol.sum())
    n = 30
    data = rng.standard_normal((nl + nv, n)).astype(np.float32)
    name = tm_io.write_tm_filetype(str(tmp_path / "mmr"), image_array=data, masking_array=[keep[:, None, None], vol],
                                   vertex_array=[verts], face_array=[faces])
    merge_y, surfs, offs, vd, zero = tm_io.mmr_inputs(tm_io.read_tm_filetype(name, verbose=False))
    X = np.column_stack([np.ones(n), rng.standard_normal(n), (rng.random(n) < 0.5).astype(float)])
    perms = randomise.permutation_indices(n, range(1, 5), seed=4)
    res = randomise.mmr(merge_y, X[:, 1:], surfs, offs, [(2.0, 1.0)] * 2, perms, merged=True, zero_in_image={0: True},
                        vdensity=vd, precision="fp32")
    # the reference's merged image: full mesh (zeros where masked) then the volume
    rp_full, col_full = adjacency.faces_to_csr(faces, len(verts))
    mrp, mcol = adjacency.merge_csr([(rp_full, col_full), adjacency.voxel_csr(vol, 26)])
    fullmask = np.concatenate([keep, np.ones(nv, bool)])
    pos = [0, nl, nl + nv]
    for c in (1, 2):
        for p, perm in enumerate(perms):
            t = ools.perm_t(X, merge_y, [perm], [c])[0, 0].astype(np.float32)
            for k, sg in enumerate((1, -1)):
                img = np.zeros(len(fullmask), np.float32)
                img[fullmask] = sg * t
                enh, _, _ = otfce.tfce(mrp, mcol, img, 2.0, 1.0, exact_pow=True)
                ein, tin = enh[fullmask], img[fullmask]
                for s in range(2):
                    a, b = pos[s], pos[s + 1]
                    exp = (ein[a:b] * (tin[a:b].max() / np.float32(100)) * vd[a:b]).max()
                    got = res[(s, c)][p, k]
                    assert abs(got - exp) <= 1e-4 * abs(exp) + 1e-6, (c, p, k, s, got, exp)


def test_vertex_regression_with_exchange_blocks(tmp_path):
    """-e: permutations within exchange blocks (pyfunc.rand_blocks,
    pyfunc.py:2733-2757) read from the block file, end to end on the GPU; the
    null equals the driver's on the same permutations, which keep every
    subject inside its block, and matches the oracle loop."""
    from tfce_mediation_amd import randomise, step2
    say("[gpu] step2.vertex_regression with exchange blocks")
    rp, col = synth.icosphere_csr(3)
    V = len(rp) - 1
    rng = np.random.default_rng(8)
    m_lh = m_rh = np.ones(V, bool)
    n = 24
    merge_y = rng.standard_normal((n, 2 * V)).astype(np.float32)
    pred_x = (rng.random(n) < 0.5).astype(float)[:, None]
    blocks = np.repeat(np.array(["a", "b", "c"]), 8)
    np.savetxt(os.path.join(tmp_path, "blocks.csv"), blocks, fmt="%s")
    save(os.path.join(tmp_path, "python_temp_thick"), pred_x=pred_x, num_subjects=n, all_vertex=V,
         num_vertex=2 * V, num_vertex_lh=V, num_vertex_rh=V, bin_mask_lh=m_lh, bin_mask_rh=m_rh,
         adjac_lh=ragged(rp, col), adjac_rh=ragged(rp, col), merge_y=merge_y, optstfce=np.array([2.0, 1.0]),
         vdensity_lh=np.ones(V, np.float32), vdensity_rh=np.ones(V, np.float32))
    got = step2.vertex_regression("thick", 0, 5, root=str(tmp_path), seed=4, allow_pickle=True,
                                  exchangeblock=os.path.join(tmp_path, "blocks.csv"), precision="fp32")
    perms = randomise.permutation_indices(n, range(0, 6), seed=4, blocks=blocks)
    for perm in perms:
        assert sorted(perm.tolist()) == list(range(n))
        for lab in "abc":  # a block's positions take their subjects from one block
            assert len(set(blocks[perm[blocks == lab]])) == 1
    ref = randomise.vertex_regress(merge_y, pred_x, ragged(rp, col), ragged(rp, col), m_lh, m_rh, [2.0, 1.0],
                                   perms, precision="fp32")
    np.testing.assert_array_equal(got[1], ref[1])
    X = np.column_stack([np.ones(n), pred_x])
    for p, perm in enumerate(perms):
        t = ools.perm_t(X, merge_y, [perm], [1])[0, 0].astype(np.float32)
        for k, s in enumerate((1, -1)):
            best = 0.0
            for h in range(2):
                img = (s * t[h * V:(h + 1) * V]).astype(np.float32)
                enh, _, _ = otfce.tfce(rp, col, img, 2.0, 1.0, exact_pow=True)
                best = max(best, float(enh.max() * (img.max() / np.float32(100))))
            assert abs(got[1][p, k] - best) <= 1e-4 * abs(best), (p, k, got[1][p, k], best)


def _tmi_temp(root, n, rng, mediation=False):
    """A tmi_temp folder as tm_mmr_rand_low_ram.py:150-201 saves it: a masked
    order-3 icosphere (vertex image: full-surface mask and adjacency, data
    and vdensity of the in-mask vertices) and a voxel image (adjacency over
    its voxels, all-true mask)."""
    import argparse
    d = os.path.join(root, "tmi_temp")
    os.makedirs(d, exist_ok=True)
    rp, col = synth.icosphere_csr(3)
    V = len(rp) - 1
    m0 = rng.random(V) < 0.85
    vmask = synth.ellipsoid_mask((9, 10, 9), (4, 4.5, 4), (3.6, 4.1, 3.6))
    adj1 = adjacency.create_adjac_voxel(vmask, vmask, int(vmask.sum()), 26)
    a1 = np.empty(len(adj1), dtype=object)
    for i, l in enumerate(adj1):
        a1[i] = l
    deg0 = np.diff(rp).astype(float)
    vd0 = (1 - deg0 / deg0.max() + deg0.mean() / deg0.max()).astype(np.float32)[m0]
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    np.savetxt(os.path.join(root, "x.csv"), x, delimiter=",")
    np.savetxt(os.path.join(root, "y.csv"), y, delimiter=",")
    data0 = (rng.standard_normal((n, int(m0.sum()))) + 0.6 * x[:, None] * (rng.random(int(m0.sum())) < 0.3))
    data1 = rng.standard_normal((n, len(adj1))) + 0.5 * x[:, None]
    for s, (adj, mask, data, vd) in enumerate([(ragged(rp, col), m0, data0, vd0),
                                               (a1, np.ones(len(adj1), bool), data1, np.array([1.0]))]):
        np.save(os.path.join(d, "%d_adjacency_temp" % s), adj)
        np.save(os.path.join(d, "%d_mask_temp" % s), mask)
        np.save(os.path.join(d, "%d_data_temp" % s), data.astype(np.float32))
        np.save(os.path.join(d, "%d_vdensity_temp" % s), np.asarray(vd, np.float32))
    opts = argparse.Namespace(input=None if mediation else ["x.csv"], tfce=[2.0, 1.0, 2.0, 0.5],
                              assigntfcesettings=[0, 1], inputmediation=["M", "x.csv", "y.csv"] if mediation else None)
    np.save(os.path.join(d, "opts"), np.array(opts, dtype=object))
    return {0: (ragged(rp, col), m0, data0.astype(np.float32), vd0, (2.0, 1.0)),
            1: (a1, np.ones(len(adj1), bool), data1.astype(np.float32), np.array([1.0], np.float32), (2.0, 0.5))}, x, y


@pytest.mark.parametrize("mediation", [False, True])
def test_mmr_lr_from_tmi_temp(tmp_path, mediation):
    """mmr-lr's permutation branch (tm_mmr_rand_low_ram_parallel.py:174-212)
    from a tmi_temp folder, batched, against the per-permutation drop-in
    (tm_func.low_ram_calculate_tfce / _mediation_tfce, themselves checked
    against the oracle), file by file."""
    from tfce_mediation_amd import step2, tm_func
    from tfce_mediation_amd.tfce import CreateAdjSet
    say(f"[gpu] step2.mmr_lr mediation={mediation}")
    rng = np.random.default_rng(12)
    n = 32
    surf, x, y = _tmi_temp(str(tmp_path), n, rng, mediation)
    with pytest.raises(ValueError, match="allow_pickle"):
        step2.mmr_lr(0, 1, 3, os.path.join(tmp_path, "out"), 5, root=str(tmp_path))
    out, ref = os.path.join(tmp_path, "out"), os.path.join(tmp_path, "ref")
    os.makedirs(ref)
    for s, (adj, mask, data, vd, (H, E)) in surf.items():
        step2.mmr_lr(s, 1, 6, out, 5, root=str(tmp_path), allow_pickle=True, precision="fp32")
        calc = CreateAdjSet(H, E, adj)
        for p in range(1, 7):
            if mediation:
                tm_func.low_ram_calculate_mediation_tfce("M", data, mask, x, y, calc, vd, set_surf_count=s,
                                                         perm_number=p, randomise=True, output_dir=ref,
                                                         perm_seed=5)
            else:
                tm_func.low_ram_calculate_tfce(data, mask, x, calc, vd, set_surf_count=s, perm_number=p,
                                               randomise=True, output_dir=ref, perm_seed=5)
    names = sorted(os.listdir(ref))
    assert names == sorted(os.listdir(out)) and len(names) == 2
    for name in names:
        got = np.loadtxt(os.path.join(out, name))
        exp = np.loadtxt(os.path.join(ref, name))
        assert got.shape == exp.shape == ((6,) if mediation else (12,))
        np.testing.assert_allclose(got, exp, rtol=1e-4, atol=1e-6)


def test_vertex_regression_from_step1_writer(tmp_path):
    """The directory step1.save_vertex_handoff writes drives step 2 to the
    nulls of the batched driver called directly."""
    from tfce_mediation_amd import randomise, step1, step2
    say("[gpu] step2.vertex_regression from step1.save_vertex_handoff")
    rp, col = synth.icosphere_csr(3)
    V = len(rp) - 1
    rng = np.random.default_rng(2)
    m_lh, m_rh = rng.random(V) < 0.85, rng.random(V) < 0.9
    n = 26
    merge_y = rng.standard_normal((n, int(m_lh.sum() + m_rh.sum()))).astype(np.float32)
    pred_x = rng.standard_normal((n, 2))
    vd = adjacency.vertex_density(rp)
    step1.save_vertex_handoff("thickness", pred_x, merge_y, (rp, col), (rp, col), m_lh, m_rh, [2, 1], vd, vd,
                              root=str(tmp_path))
    got = step2.vertex_regression("thickness", 1, 6, root=str(tmp_path), seed=5, allow_pickle=True)
    perms = randomise.permutation_indices(n, range(1, 7), seed=5)
    ref = randomise.vertex_regress(merge_y, pred_x, ragged(rp, col), ragged(rp, col), m_lh, m_rh, [2.0, 1.0],
                                   perms, vdensity_lh=vd, vdensity_rh=vd)
    for c in (1, 2):
        np.testing.assert_array_equal(got[c], ref[c])
