"""TMI container (tm_io.py:72-444 layout): a file built byte by byte in this
test is read back in the right orientation; the writer's header lines and
payload are the ones the format defines; a full round trip; the hand-off to
randomise.mmr (mmr_inputs) against inputs built by hand.  No GPU."""
import numpy as np
import pytest

from oracle import graphs
from tfce_mediation_amd import adjacency, synth, tm_io


def test_reads_a_hand_built_file(tmp_path):
    # data 3 elements x 2 subjects, stored transposed (subject-major), then a 2x2x1 mask
    data = np.array([[1, 2], [3, 4], [5, 6]], np.float32)
    mask = np.array([[[1], [0]], [[1], [1]]], bool)
    head = ("tmi\nformat binary_little_endian 1.1\ncomment hand made\n"
            "element data_array\ndtype float32\nnbytes 24\ndatashape 3 2\n"
            "element masking_array\ndtype uint8\nnbytes 4\nnmasked 3\nmaskshape 2 2 1\nmaskname m0\n"
            "history mode_add 0 1 1 0 0 0\nend_header\n").encode()
    payload = np.array([1, 3, 5, 2, 4, 6], "<f4").tobytes() + bytes([1, 1, 0, 1])  # mask in Fortran order
    p = tmp_path / "hand.tmi"
    p.write_bytes(head + payload)
    el, img, masks, names, aff, v, f, sn, adj, hist, cols = tm_io.read_tm_filetype(str(p), verbose=False)
    assert el == ["data_array", "masking_array"]
    np.testing.assert_array_equal(img[0], data)
    np.testing.assert_array_equal(masks[0], mask)
    assert names == ["m0"] and hist == ["history mode_add 0 1 1 0 0 0"]


def test_writer_header_and_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    verts, faces = synth.icosphere(1)
    vmask = (rng.random(len(verts)) < 0.8)[:, None, None]
    vol = synth.ellipsoid_mask((6, 7, 5), (2.5, 3, 2), (2.5, 2.8, 2))
    V = int(vmask.sum() + vol.sum())
    data = rng.standard_normal((V, 4)).astype(np.float32)
    adj_v = [sorted(s) for s in graphs.create_adjac_vertex(verts, faces)]
    adj_m = graphs.create_adjac_voxel(vol, vol.astype(np.float32), int(vol.sum()), 26)
    aff = np.eye(4, dtype=np.float32)
    name = tm_io.write_tm_filetype(str(tmp_path / "t"), image_array=data, masking_array=[vmask, vol],
                                   maskname=["lh", "vol"], affine_array=[aff, aff], vertex_array=[verts],
                                   face_array=[faces], surfname=["ico1"], adjacency_array=[adj_v, adj_m],
                                   columnids=np.array([10, 11, 12, 13]))
    text = open(name, "rb").read()
    head = text[:text.index(b"end_header\n")].decode().split("\n")
    assert head[:2] == ["tmi", "format binary_little_endian 1.1"]
    i = head.index("element data_array")
    assert head[i + 1:i + 4] == ["dtype float32", "nbytes %d" % (V * 4 * 4), "datashape %d 4" % V]
    i = head.index("element masking_array")
    assert head[i + 1:i + 6] == ["dtype uint8", "nbytes %d" % len(verts), "nmasked %d" % int(vmask.sum()),
                                 "maskshape %d 1 1" % len(verts), "maskname lh"]
    assert head.count("element adjacency_object") == 2
    assert head[-2:] == ["history mode_add 0 1 2 2 1 2", ""]
    el, img, masks, names, affs, vv, ff, sn, adj, hist, cols = tm_io.read_tm_filetype(name, verbose=False)
    np.testing.assert_array_equal(img[0], data)
    np.testing.assert_array_equal(masks[0], vmask)
    np.testing.assert_array_equal(masks[1], vol)
    np.testing.assert_array_equal(vv[0], verts.astype(np.float32))
    np.testing.assert_array_equal(ff[0], faces)
    np.testing.assert_array_equal(affs[1], aff)
    assert names == ["lh", "vol"] and sn == ["ico1"] and list(cols[0]) == [10, 11, 12, 13]
    assert adj == [None, None]  # pickled objects are not read unless asked for
    adj = tm_io.read_tm_filetype(name, verbose=False, allow_pickle=True)[8]
    assert [list(r) for r in adj[0]] == adj_v and [list(r) for r in adj[1]] == adj_m


def test_mmr_inputs_match_hand_built(tmp_path):
    rng = np.random.default_rng(1)
    verts, faces = synth.icosphere(2)
    keep = rng.random(len(verts)) < 0.85
    vol = synth.ellipsoid_mask((7, 8, 6), (3, 3.5, 2.5), (3, 3.3, 2.4))
    V = int(keep.sum() + vol.sum())
    data = rng.standard_normal((V, 9)).astype(np.float32)
    name = tm_io.write_tm_filetype(str(tmp_path / "m"), image_array=data, masking_array=[keep[:, None, None], vol],
                                   vertex_array=[verts], face_array=[faces])
    tmi = tm_io.read_tm_filetype(name, verbose=False)
    merge_y, surfs, offs, vd, zero = tm_io.mmr_inputs(tmi)
    np.testing.assert_array_equal(merge_y, data.T)
    assert offs == [0, int(keep.sum())] and zero == [True, False]
    rp, col = adjacency.faces_to_csr(faces, len(verts))
    exp_lh = adjacency.induced_subgraph(rp, col, keep)
    exp_vol = adjacency.voxel_csr(vol, 26)
    for got, exp in zip(surfs, (exp_lh, exp_vol)):
        np.testing.assert_array_equal(got[0], exp[0])
        np.testing.assert_array_equal(got[1], exp[1])
    deg = np.diff(rp)[keep].astype(np.float64)
    np.testing.assert_allclose(vd[:int(keep.sum())], 1 - deg / deg.max() + deg.mean() / deg.max(), rtol=1e-6)
    with pytest.raises(ValueError, match="not a TFCE_mediation image"):
        (tmp_path / "bad.tmi").write_bytes(b"ply\n")
        tm_io.read_tm_filetype(str(tmp_path / "bad.tmi"))
