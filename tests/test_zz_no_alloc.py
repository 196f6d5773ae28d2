"""tfce_batch_max_noalloc (no device allocation, static per-GPU scratch)
against tfce_batch_max through the C ABI: identical maxima, bit for bit, on
consecutive calls whose second input is smaller, with and without vdensity
and one-sided sign masks; a request whose records do not fit the scratch is
refused before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def smooth_maps(row_ptr, col, P, seed):
    rng = np.random.default_rng(seed)
    V = len(row_ptr) - 1
    x = rng.standard_normal((P, V))
    deg = np.diff(row_ptr).astype(np.int64)
    src = np.repeat(np.arange(V), deg)
    for _ in range(3):
        s = np.zeros_like(x)
        np.add.at(s.T, src, x.T[col])
        x = (x + s) / (deg + 1)
    return (x / x.std(1, keepdims=True)).astype(np.float32)


def both(lib, adj, maps, sign_mask, vd):
    import torch
    from tfce_mediation_amd import _lib
    P = maps.shape[0]
    a = torch.full((P, 2), -7.0, dtype=torch.float32, device="cuda")
    b = torch.full((P, 2), -9.0, dtype=torch.float32, device="cuda")
    _lib.check(lib.tfce_batch_max(adj.handle, _lib.ptr(maps), P, sign_mask, _lib.ptr(vd), _lib.ptr(a),
                                  _lib.stream_ptr()))
    _lib.check(lib.tfce_batch_max_noalloc(adj.handle, _lib.ptr(maps), P, sign_mask, _lib.ptr(vd), _lib.ptr(b),
                                          _lib.stream_ptr()))
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_matches_batch_max():
    import torch
    from tfce_mediation_amd import _lib, adjacency, synth
    from tfce_mediation_amd.tfce import Adjacency
    lib = _lib.lib()
    cases = [(6, 200, 3, False), (4, 37, 2, True), (7, 16, 3, False), (5, 64, 1, True)]
    for order, P, sign_mask, use_vd in cases:
        print(f"[gpu] no-alloc batch max: order-{order} icosphere, P={P}, signs {sign_mask}, vdensity {use_vd}",
              flush=True)
        rp, col = synth.icosphere_csr(order)
        adj = Adjacency((rp, col), 2.0, 1.0)
        maps = torch.from_numpy(smooth_maps(rp, col, P, order * 100 + P)).cuda()
        vd = torch.from_numpy(adjacency.vertex_density(rp)).cuda() if use_vd else None
        a, b = both(lib, adj, maps, sign_mask, vd)
        np.testing.assert_array_equal(a, b)
        assert np.all(a[:, 0] > 0) if sign_mask & 1 else np.all(a[:, 0] == 0)


def test_noalloc_refuses_oversized_request():
    import torch
    from tfce_mediation_amd import _lib, synth
    from tfce_mediation_amd.tfce import Adjacency
    lib = _lib.lib()
    rp, col = synth.icosphere_csr(3)
    adj = Adjacency((rp, col), 2.0, 1.0)
    P = 20000  # 40,000 items of 8.2 KB records: more than the 256 MiB scratch
    print(f"[gpu] no-alloc refusal P={P}", flush=True)
    maps = torch.zeros((P, len(rp) - 1), dtype=torch.float32, device="cuda")
    out = torch.zeros((P, 2), dtype=torch.float32, device="cuda")
    rc = lib.tfce_batch_max_noalloc(adj.handle, _lib.ptr(maps), P, 3, None, _lib.ptr(out), _lib.stream_ptr())
    assert rc == -1000
    assert b"scratch" in lib.tfce_last_error()
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0
