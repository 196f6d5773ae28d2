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


def _bandwidth(rp, col):
    src = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return int(np.abs(src - np.asarray(col)).max()) if len(col) else 0


def test_locality_order_is_a_band_reducing_permutation():
    """tfce_locality_order (host code of the library): a permutation; a
    shuffled path comes back with bandwidth 1; on an icosphere the bandwidth is
    within 25% of scipy's reverse Cuthill-McKee; components stay contiguous."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    from tfce_mediation_amd.adjacency import locality_order, relabel_csr, to_csr
    from tfce_mediation_amd import synth
    rng = np.random.default_rng(0)
    n = 500
    lab = rng.permutation(n)
    edges = [(lab[i], lab[i + 1]) for i in range(n - 1)]
    adj = [[] for _ in range(n)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    rp, col = to_csr(adj)
    order = locality_order(rp, col)
    assert np.array_equal(np.sort(order), np.arange(n))
    rp2, col2 = relabel_csr(rp, col, order)
    assert _bandwidth(rp2, col2) == 1
    # the relabelled graph is the same graph
    src = np.repeat(np.arange(n), np.diff(rp))
    e1 = sorted(zip(order[src].tolist(), order[np.asarray(col)].tolist()))
    src2 = np.repeat(np.arange(n), np.diff(rp2))
    assert e1 == sorted(zip(src2.tolist(), col2.tolist()))
    rp, col = synth.icosphere_csr(5)
    V = len(rp) - 1
    rp2, col2 = relabel_csr(rp, col, locality_order(rp, col))
    r = reverse_cuthill_mckee(csr_matrix((np.ones(len(col)), col, rp), shape=(V, V)), symmetric_mode=True)
    inv = np.empty(V, np.int64)
    inv[r] = np.arange(V)
    rp3, col3 = relabel_csr(rp, col, inv)
    assert _bandwidth(rp2, col2) <= 1.25 * _bandwidth(rp3, col3)
    assert _bandwidth(rp2, col2) < _bandwidth(rp, col) / 10
    # two components: each occupies a contiguous id range
    adj = [[1], [0, 2], [1], [4], [3]]
    o = locality_order(*to_csr(adj))
    assert sorted(o[:3].tolist()) in ([0, 1, 2], [2, 3, 4])


# ---------------------------------------------------------------------------
# Hand-derived lists (pyfunc.py:48-76 read as rules, not as code): labels are
# C-order ranks of the mask voxels; with 26-connectivity a voxel lists every
# mask voxel at Chebyshev distance 1 except label 0; with 6-connectivity only a
# voxel whose 3x3x3 window is not clipped by the volume lists anything (its
# face neighbours except label 0); label 0 lists nothing.
def _brute(mask, dirtype):
    idx = np.argwhere(mask)  # C order = label order
    out = []
    dims = np.array(mask.shape)
    for a, pa in enumerate(idx):
        lst = []
        if a != 0 and (dirtype == 26 or (np.all(pa > 0) and np.all(pa < dims - 1))):
            for b, pb in enumerate(idx):
                d = np.abs(pa - pb)
                if b == 0 or b == a:
                    continue
                if (dirtype == 26 and d.max() == 1) or (dirtype == 6 and d.sum() == 1):
                    lst.append(b)
        out.append(lst)
    return out


def _both(mask, dirtype):
    ref = graphs.create_adjac_voxel(mask, mask.astype(np.float32), int(mask.sum()), dirtype)
    rp, col = adjacency.voxel_csr(mask, dirtype)
    mine = [col[rp[i]:rp[i + 1]].tolist() for i in range(len(rp) - 1)]
    return ref, mine


def test_voxel_adjacency_hand_derived_cube3():
    mask = np.ones((3, 3, 3), bool)
    for dirtype in (26, 6):
        ref, mine = _both(mask, dirtype)
        assert ref == mine == _brute(mask, dirtype)
    ref26, _ = _both(mask, 26)
    assert ref26[0] == []
    assert ref26[13] == [i for i in range(1, 27) if i != 13]  # centre: everyone but itself and 0
    assert ref26[26] == [13, 14, 16, 17, 22, 23, 25]          # far corner (2,2,2)
    assert ref26[1] == [2, 3, 4, 5, 9, 10, 11, 12, 13, 14]    # (0,0,1): 0 is never a neighbour
    assert all(0 not in lst for lst in ref26)
    ref6, _ = _both(mask, 6)
    assert ref6[13] == [4, 10, 12, 14, 16, 22]
    assert all(lst == [] for i, lst in enumerate(ref6) if i != 13)  # every other voxel touches a wall


def test_voxel_adjacency_hand_derived_cube4_one_removed():
    mask = np.ones((4, 4, 4), bool)
    mask[1, 2, 1] = False  # C-order index 25: labels past it shift down by one

    def lab(x, y, z):
        i = 16 * x + 4 * y + z
        assert i != 25
        return i - (i > 25)
    for dirtype in (26, 6):
        ref, mine = _both(mask, dirtype)
        assert len(ref) == 63
        assert ref == mine == _brute(mask, dirtype)
    ref26, _ = _both(mask, 26)
    ref6, _ = _both(mask, 6)
    assert ref26[0] == [] and ref6[0] == []
    assert ref26[lab(0, 0, 1)] == [2, 4, 5, 6, 16, 17, 18, 20, 21, 22]
    assert ref6[lab(1, 1, 1)] == sorted([lab(0, 1, 1), lab(2, 1, 1), lab(1, 0, 1), lab(1, 1, 0), lab(1, 1, 2)])
    assert ref6[lab(2, 2, 2)] == sorted([lab(1, 2, 2), lab(3, 2, 2), lab(2, 1, 2), lab(2, 3, 2), lab(2, 2, 1),
                                         lab(2, 2, 3)])
    interior = {lab(x, y, z) for x in (1, 2) for y in (1, 2) for z in (1, 2) if (x, y, z) != (1, 2, 1)}
    assert len(interior) == 7
    assert all((lst != []) == (i in interior) for i, lst in enumerate(ref6))
    # the removed voxel's neighbours lose it; the centre-ish voxel (2,2,2) lists 25 others
    assert len(ref26[lab(2, 2, 2)]) == 25
    assert all(0 not in lst for lst in ref26 + ref6)


def test_vertex_adjacency_hand_derived_octahedron():
    """pyfunc.py:37-46 on an octahedron: every vertex is joined to the four
    vertices that are not its antipode (0-1, 2-3, 4-5 are the antipodal pairs)."""
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], float)
    f = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4], [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]])
    want = [[2, 3, 4, 5], [2, 3, 4, 5], [0, 1, 4, 5], [0, 1, 4, 5], [0, 1, 2, 3], [0, 1, 2, 3]]
    assert [sorted(s) for s in graphs.create_adjac_vertex(v, f)] == want
    rp, col = adjacency.faces_to_csr(f, 6)
    assert [col[rp[i]:rp[i + 1]].tolist() for i in range(6)] == want


def test_vertex_adjacency_degenerate_face_keeps_the_repeated_vertex():
    """pyfunc.py:37-46 adds faces[i, 1] to faces[i, 0]'s set even when they are
    the same vertex: a face (3, 3, 4) puts 3 in its own set.  Vertices no face
    names keep empty rows."""
    v = np.zeros((6, 3))
    f = np.array([[0, 1, 2], [3, 3, 4], [2, 1, 0]])
    ref = graphs.create_adjac_vertex(v, f)
    assert sorted(ref[3]) == [3, 4] and sorted(ref[4]) == [3] and ref[5] == set()
    rp, col = adjacency.faces_to_csr(f, 6)
    assert [col[rp[i]:rp[i + 1]].tolist() for i in range(6)] == [sorted(s) for s in ref]
