"""Vectorised adjacency builders vs loop restatements of the reference
(oracle/graphs.py follows pyfunc.py:37-76 line by line)."""
import numpy as np
import pytest

from oracle import graphs
from tfce_mediation_amd import adjacency, synth


def test_icosphere_counts():
    for o in range(5):
        v, f = synth.icosphere(o)
        assert v.shape[0] == 10 * 4 ** o + 2
        assert f.shape[0] == 20 * 4 ** o
        rp, col = adjacency.faces_to_csr(f, v.shape[0])
        deg = np.diff(rp)
        assert col.shape[0] == 60 * 4 ** o  # 2 * (30 * 4^o) directed edges
        assert (deg == 5).sum() == 12 and np.all((deg == 5) | (deg == 6))
        assert np.allclose(np.linalg.norm(v, axis=1), 1.0)


def test_vertex_adjacency_matches_reference_loop():
    v, f = synth.icosphere(2)
    ref = graphs.create_adjac_vertex(v, f)
    rp, col = adjacency.faces_to_csr(f, v.shape[0])
    for i, s in enumerate(ref):
        assert sorted(s) == col[rp[i]:rp[i + 1]].tolist()
    rp2, col2 = adjacency.to_csr(ref)
    assert np.array_equal(rp2, rp)


@pytest.mark.parametrize("dirtype", [26, 6])
@pytest.mark.parametrize("case", ["cube", "ellipsoid", "touching_walls", "random"])
def test_voxel_adjacency_matches_reference_loop(dirtype, case):
    rng = np.random.default_rng(1)
    if case == "cube":
        mask = np.ones((3, 3, 3), bool)
    elif case == "ellipsoid":
        mask = synth.ellipsoid_mask((9, 11, 8), (4, 5, 3.5), (4, 5, 3.4))
    elif case == "touching_walls":
        mask = np.ones((4, 5, 4), bool)
        mask[0, 0, 0] = False
    else:
        mask = rng.random((6, 7, 5)) < 0.6
    data_mask = mask.astype(np.float32)
    ref = graphs.create_adjac_voxel(mask, data_mask, int(mask.sum()), dirtype)
    rp, col = adjacency.voxel_csr(mask, dirtype)
    assert rp.shape[0] - 1 == len(ref)
    for i, lst in enumerate(ref):
        assert col[rp[i]:rp[i + 1]].tolist() == lst, (case, i)
    assert ref[0] == []


def test_merge_and_subgraph():
    a = synth.icosphere_csr(1)
    b = synth.icosphere_csr(0)
    rp, col = adjacency.merge_csr([a, b])
    assert rp.shape[0] - 1 == 42 + 12
    lists = adjacency.merge_adjacency_array([0, 1], [graphs.create_adjac_vertex(*synth.icosphere(1)),
                                                    graphs.create_adjac_vertex(*synth.icosphere(0))])
    for i in range(54):
        assert sorted(lists[i]) == sorted(col[rp[i]:rp[i + 1]].tolist())
    keep = np.zeros(54, bool)
    keep[:42] = True
    srp, scol = adjacency.induced_subgraph(rp, col, keep)
    assert np.array_equal(srp, a[0]) and np.array_equal(scol, a[1])


def test_vertex_density_formula():
    rp, col = synth.icosphere_csr(2)
    vd = adjacency.vertex_density(rp)
    deg = np.diff(rp).astype(float)
    assert vd.dtype == np.float32
    assert np.allclose(vd[deg == 6], 1 - 1 + deg.mean() / 6)
    assert np.allclose(vd[deg == 5], 1 - 5 / 6 + deg.mean() / 6)
