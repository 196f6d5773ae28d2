"""GPU adjacency builders (tfce_adj_voxel_csr / tfce_adj_mesh_csr) against
the host builders and the loop restatements of pyfunc.py:37-76
(oracle/graphs.py): integer work, so equal entry for entry."""
import numpy as np
import pytest

from oracle import graphs
from tfce_mediation_amd import adjacency, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _masks():
    rng = np.random.default_rng(5)
    out = {
        "cube3": np.ones((3, 3, 3), bool),
        "ellipsoid": synth.ellipsoid_mask((21, 25, 19), (10, 12, 9), (9, 11, 8.5)),
        "walls": np.ones((6, 7, 5), bool),
        "random": rng.random((9, 10, 8)) < 0.55,
        "one_voxel": np.zeros((4, 4, 4), bool),
        "line": np.zeros((1, 1, 9), bool),
    }
    out["one_voxel"][2, 1, 3] = True
    out["line"][0, 0, 2:7] = True
    return out


@pytest.mark.parametrize("dirtype", [26, 6])
@pytest.mark.parametrize("name", sorted(_masks()))
def test_voxel_csr_on_gpu_matches_host_and_reference_loop(name, dirtype):
    mask = _masks()[name]
    say(f"[gpu] voxel adjacency {name} dirtype={dirtype}")
    rp, col = adjacency.voxel_csr_device(mask.astype(np.float32) * 3.0, dirtype)
    rp, col = rp.cpu().numpy(), col.cpu().numpy()
    hrp, hcol = adjacency.voxel_csr(mask, dirtype)
    np.testing.assert_array_equal(rp, hrp)
    np.testing.assert_array_equal(col, hcol)
    if mask.sum() <= 5000:
        ref = graphs.create_adjac_voxel(mask, mask.astype(np.float32), int(mask.sum()), dirtype)
        assert [col[rp[i]:rp[i + 1]].tolist() for i in range(len(rp) - 1)] == ref


def test_voxel_csr_on_gpu_full_brain_grid():
    """The 91 x 109 x 91 grid with the ellipsoidal brain mask (~232k voxels)."""
    mask = synth.ellipsoid_mask()
    say(f"[gpu] voxel adjacency C2 grid V={int(mask.sum())}")
    for dirtype in (26, 6):
        rp, col = adjacency.voxel_csr_device(mask, dirtype)
        hrp, hcol = adjacency.voxel_csr(mask, dirtype)
        np.testing.assert_array_equal(rp.cpu().numpy(), hrp)
        np.testing.assert_array_equal(col.cpu().numpy(), hcol)


@pytest.mark.parametrize("order", [0, 1, 3, 5, 7])
def test_mesh_csr_on_gpu_matches_host(order):
    v, f = synth.icosphere(order)
    say(f"[gpu] mesh adjacency icosphere order {order}")
    rng = np.random.default_rng(order)
    f = f[rng.permutation(len(f))]
    rp, col = adjacency.faces_to_csr_device(f, v.shape[0])
    hrp, hcol = adjacency.faces_to_csr(f, v.shape[0])
    np.testing.assert_array_equal(rp.cpu().numpy(), hrp)
    np.testing.assert_array_equal(col.cpu().numpy(), hcol)
    if order <= 3:
        ref = graphs.create_adjac_vertex(v, f)
        rpn, coln = rp.cpu().numpy(), col.cpu().numpy()
        assert [coln[rpn[i]:rpn[i + 1]].tolist() for i in range(len(rpn) - 1)] == [sorted(s) for s in ref]


def test_mesh_csr_on_gpu_edge_cases():
    import torch
    from tfce_mediation_amd._lib import TfceHipError
    say("[gpu] mesh adjacency edge cases")
    f = np.array([[0, 1, 2], [3, 3, 4], [2, 1, 0]])
    rp, col = adjacency.faces_to_csr_device(f, 6)
    ref = graphs.create_adjac_vertex(np.zeros((6, 3)), f)
    rpn, coln = rp.cpu().numpy(), col.cpu().numpy()
    assert [coln[rpn[i]:rpn[i + 1]].tolist() for i in range(6)] == [sorted(s) for s in ref]
    rp, col = adjacency.faces_to_csr_device(np.zeros((0, 3), np.int32), 4)
    assert rp.cpu().numpy().tolist() == [0, 0, 0, 0, 0] and col.numel() == 0
    with pytest.raises(TfceHipError, match="outside"):
        adjacency.faces_to_csr_device(np.array([[0, 1, 7]]), 5)
    torch.cuda.synchronize()
