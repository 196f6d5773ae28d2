"""GPU parity of the TFCE sweep (libtfcehip) against the C oracle.

Bar: bit-exact float32 against the oracle in correctly-rounded-power mode for
H=2 with E=1 (every increment is one correctly rounded product); E=0.5 / 0.67
use sqrt / pow on the GPU where the oracle calls glibc pow, so those allow
4 ulp.  Against the reference-identical oracle (libm powf for T^H, one ulp off
for ~0.06% of thresholds) the bar is 4 ulp.  Component partitions must be
identical at every checked level.
"""
import sys

import numpy as np
import pytest

from oracle import tfce as otfce
from tfce_mediation_amd import adjacency, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    return _t


def smooth_map(row_ptr, col, seed, passes=3, shift=0.3, V=None):
    rng = np.random.default_rng(seed)
    V = len(row_ptr) - 1
    x = rng.standard_normal(V)
    deg = np.diff(row_ptr).astype(np.int64)
    src = np.repeat(np.arange(V), deg)
    for _ in range(passes):
        s = np.bincount(src, weights=x[col], minlength=V)
        x = (x + s) / (deg + 1)
    return (x / x.std() + shift).astype(np.float32)


def ulp_diff(a, b):
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    return np.abs(ai - bi)


def check_close(got, ref, E, H):
    if E == 1.0 and H in (1.0, 2.0):
        bad = np.nonzero(got.view(np.uint32) != ref.view(np.uint32))[0]
        assert bad.size == 0, (bad[:5], got[bad[:5]], ref[bad[:5]])
    else:
        assert ulp_diff(got, ref).max() <= 4


def mesh_with_extra_edges(order, n_extra, hops, seed):
    """Icosphere plus random symmetric edges between vertices a few mesh hops
    apart: irregular degrees (rows wider than 8), as on native meshes."""
    rp, col = synth.icosphere_csr(order)
    V = len(rp) - 1
    rng = np.random.default_rng(seed)
    nbrs = [set(col[rp[v]:rp[v + 1]].tolist()) for v in range(V)]
    for _ in range(n_extra):
        a = int(rng.integers(V))
        b = a
        for _ in range(hops):
            b = int(rng.choice(sorted(nbrs[b])))
        if b != a:
            nbrs[a].add(b)
            nbrs[b].add(a)
    return adjacency.to_csr([sorted(x) for x in nbrs])


def shuffled_rows(row_ptr, col, seed):
    """The same graph with every row's neighbours in random order."""
    rng = np.random.default_rng(seed)
    col = col.copy()
    for v in range(len(row_ptr) - 1):
        rng.shuffle(col[row_ptr[v]:row_ptr[v + 1]])
    return row_ptr, col


GRAPHS = {
    "ico3": lambda: synth.icosphere_csr(3),
    # rows of at most six neighbours in no particular order: the library puts
    # them in ring order; masked: open rings (boundary vertices) and short rows
    "ico4_shuffled": lambda: shuffled_rows(*synth.icosphere_csr(4), 7),
    "ico5_masked": lambda: adjacency.induced_subgraph(*synth.icosphere_csr(5),
                                                      np.random.default_rng(3).random(10242) < 0.85),
    # a few vertices of degree 7 and 8: rows of width 8 that are not ring-packed
    "ico4_w8": lambda: mesh_with_extra_edges(4, 60, 2, 4),
    # in-plane 8-connected grid (width 8, not ring-packed; voxel 0 isolated)
    "grid8": lambda: adjacency.voxel_csr(np.ones((1, 30, 40), bool), 26),
    # rows of width 12: two groups of 8 in the neighbour-position rows
    "ico4_w12": lambda: mesh_with_extra_edges(4, 3000, 2, 1),
    # rows of width 20, padded to 32: rows of joined positions found with the
    # full row adjacency (the 26-connected volumes' path) on an irregular mesh
    "ico4_w20": lambda: mesh_with_extra_edges(4, 9000, 2, 2),
    "ico5": lambda: synth.icosphere_csr(5),
    "vox26": lambda: adjacency.voxel_csr(synth.ellipsoid_mask((21, 25, 19), (10, 12, 9), (9, 11, 8.5)), 26),
    "vox6_walls": lambda: adjacency.voxel_csr(np.ones((7, 8, 6), bool), 6),
}


@pytest.mark.parametrize("gname", sorted(GRAPHS))
@pytest.mark.parametrize("HE", [(2.0, 1.0), (2.0, 0.5), (2.0, 0.67)])
def test_run_matches_oracle(torch, gname, HE):
    from tfce_mediation_amd.tfce import CreateAdjSet
    H, E = HE
    row_ptr, col = GRAPHS[gname]()
    say(f"[gpu] CreateAdjSet.run {gname} V={len(row_ptr)-1} H={H} E={E}")
    op = CreateAdjSet(H, E, (row_ptr, col))
    for seed in range(3):
        img = smooth_map(row_ptr, col, seed)
        if gname == "vox6_walls":
            img = img + np.float32(1e-4) * np.arange(len(img), dtype=np.float32) / len(img)
        for sgn in (1, -1):
            x = (sgn * img).astype(np.float32)
            enh = np.zeros_like(x)
            op.run(x, enh)
            ref, st, _ = otfce.tfce(row_ptr, col, x, H, E, exact_pow=True)
            assert st > 0
            check_close(enh, ref, E, H)
            refpow, _, _ = otfce.tfce(row_ptr, col, x, H, E)
            assert ulp_diff(enh, refpow).max() <= 4


def test_run_accumulates_and_degenerate(torch):
    from tfce_mediation_amd.tfce import CreateAdjSet
    row_ptr, col = synth.icosphere_csr(4)
    V = len(row_ptr) - 1
    say("[gpu] accumulate / degenerate ico4")
    op = CreateAdjSet(2, 1, (row_ptr, col))
    img = smooth_map(row_ptr, col, 11)
    base = np.linspace(-1, 1, V).astype(np.float32)
    enh = base.copy()
    op.run(img, enh)
    ref, _, _ = otfce.tfce(row_ptr, col, img, 2.0, 1.0, enhn=base.copy(), exact_pow=True)
    check_close(enh, ref, 1.0, 2.0)
    enh = base.copy()
    op.run(np.zeros(V, np.float32), enh)  # max == 0: the reference never returns; defined as no-op
    assert np.array_equal(enh, base)
    op.run(-np.abs(img) - 0.5, enh)  # all negative: nothing is active
    assert np.array_equal(enh, base)


@pytest.mark.parametrize("gname,E", [("ico3", 1.0), ("ico4_shuffled", 1.0), ("ico5", 0.5), ("ico5_masked", 1.0),
                                     ("ico4_w8", 1.0), ("vox26", 0.5)])
def test_pair_sweep_matches_single_image_sweeps(torch, gname, E):
    """Both signs of a map in one workgroup (ring-packed surfaces, MAX mode)
    against the one-image-per-workgroup sweep (knob off), bit for bit, and
    against the oracle; surfaces that are not ring-packed (ico4_w8, vox26)
    must take the single-image sweep either way.  The LDS-resident sweep is
    off so that these small surfaces reach the global-memory sweeps."""
    from tfce_mediation_amd import _lib
    from tfce_mediation_amd.tfce import Adjacency
    row_ptr, col = GRAPHS[gname]()
    V = len(row_ptr) - 1
    P = 37
    say(f"[gpu] paired vs single-image sweep {gname} V={V} P={P} E={E}")
    maps = np.stack([smooth_map(row_ptr, col, 900 + p, shift=0.05 * (p % 3 - 1)) for p in range(P)])
    maps[3] = np.abs(maps[3])   # the -image has nothing above 0
    maps[4] = -np.abs(maps[4])  # the +image is degenerate
    adj = Adjacency((row_ptr, col), 2.0, E)
    d = torch.from_numpy(np.ascontiguousarray(maps, np.float32)).cuda()
    lib = _lib.lib()
    try:
        _lib.check(lib.tfce_set_lds_sweep(0))
        _lib.check(lib.tfce_set_pair_sweep(1))
        pair = adj.batch_max(d, 3).cpu().numpy()
        _lib.check(lib.tfce_set_pair_sweep(0))
        single = adj.batch_max(d, 3).cpu().numpy()
    finally:
        _lib.check(lib.tfce_set_pair_sweep(1))
        _lib.check(lib.tfce_set_lds_sweep(1))
    np.testing.assert_array_equal(pair.view(np.uint32), single.view(np.uint32))
    for p in (0, 3, 4, 17, 36):
        for k, sgn in enumerate((1, -1)):
            x = (sgn * maps[p]).astype(np.float32)
            if x.max() <= 0:
                assert pair[p, k] == 0
                continue
            enh, _, _ = otfce.tfce(row_ptr, col, x, 2.0, E, exact_pow=True)
            exp = np.float32(enh.max() * np.float32(x.max() / np.float32(100)))
            if E == 1.0:
                assert pair[p, k] == exp, (p, k, pair[p, k], exp)
            else:
                assert ulp_diff(np.float32([pair[p, k]]), np.float32([exp])).max() <= 4


def test_pair_sweep_two_hemispheres_and_shared_thresholds(torch):
    """Two order-7 hemispheres in one set (the headline shape) and the
    shared-threshold mode of the merged mmr: paired and single-image sweeps
    give the same maxima bit for bit."""
    from tfce_mediation_amd import _lib
    from tfce_mediation_amd.tfce import Adjacency, SurfaceSet
    row_ptr, col = synth.icosphere_csr(7)
    V = len(row_ptr) - 1
    P = 8
    say(f"[gpu] paired sweep, two hemispheres V={2 * V} P={P}")
    maps = np.concatenate([np.stack([smooth_map(row_ptr, col, 40 + 2 * p + h, passes=5, shift=0.0)
                                     for p in range(P)]) for h in range(2)], axis=1)
    adj = Adjacency((row_ptr, col), 2.0, 1.0)
    surfs = SurfaceSet([adj, adj], [0, V])
    d = torch.from_numpy(np.ascontiguousarray(maps, np.float32)).cuda()
    lib = _lib.lib()
    outs = {}
    try:
        for on in (1, 0):
            _lib.check(lib.tfce_set_pair_sweep(on))
            outs[on] = (surfs.batch(d, 3).cpu().numpy(),
                        surfs.batch(d, 3, shared_thresholds=True).cpu().numpy())
    finally:
        _lib.check(lib.tfce_set_pair_sweep(1))
    for a_, b_ in zip(outs[1], outs[0]):
        np.testing.assert_array_equal(a_.view(np.uint32), b_.view(np.uint32))
    for p in (0, 5):
        for h in range(2):
            for k, sgn in enumerate((1, -1)):
                x = (sgn * maps[p, h * V:(h + 1) * V]).astype(np.float32)
                enh, _, _ = otfce.tfce(row_ptr, col, x, 2.0, 1.0, exact_pow=True)
                assert outs[1][0][h, p, k] == np.float32(enh.max() * np.float32(x.max() / np.float32(100)))


@pytest.mark.parametrize("gname", sorted(GRAPHS))
def test_whole_gpu_single_image_matches_oracle(torch, gname):
    """tfce_run over the whole GPU (cooperative level sweep, forced on for
    every size) against the oracle, at three grid sizes; E = 1 bit-exact,
    E = 0.5 within 4 ulp; a degenerate image adds nothing."""
    from tfce_mediation_amd import _lib
    from tfce_mediation_amd.tfce import CreateAdjSet
    row_ptr, col = GRAPHS[gname]()
    lib = _lib.lib()
    say(f"[gpu] whole-GPU tfce_run {gname} V={len(row_ptr)-1}")
    try:
        for grid in (0, 1, 7):
            _lib.check(lib.tfce_set_single_image_mode(1, grid))
            for E in (1.0, 0.5):
                op = CreateAdjSet(2.0, E, (row_ptr, col))
                for seed in range(2):
                    x = smooth_map(row_ptr, col, 40 + seed) * np.float32(1 - 2 * seed)
                    if gname == "vox6_walls":
                        x = x + np.float32(1e-4) * np.arange(len(x), dtype=np.float32) / len(x)
                    enh = np.zeros_like(x)
                    op.run(x, enh)
                    ref, st, _ = otfce.tfce(row_ptr, col, x, 2.0, E, exact_pow=True)
                    assert st > 0
                    check_close(enh, ref, E, 2.0)
        base = np.linspace(-1, 1, len(row_ptr) - 1).astype(np.float32)
        enh = base.copy()
        op.run(np.zeros_like(base), enh)
        assert np.array_equal(enh, base)
    finally:
        _lib.check(lib.tfce_set_single_image_mode(-1, 0))


def test_whole_gpu_single_image_large_surfaces(torch):
    """The default mode takes the whole-GPU sweep for surfaces of 32,768
    elements and more: an order-7 hemisphere (163,842 vertices) and a
    26-connected volume, against the oracle and against the one-workgroup
    sweep, bit for bit."""
    from tfce_mediation_amd import _lib
    from tfce_mediation_amd.tfce import CreateAdjSet
    lib = _lib.lib()
    cases = [("ico7", synth.icosphere_csr(7)),
             ("vox26_big", adjacency.voxel_csr(synth.ellipsoid_mask((46, 52, 44), (22.5, 25.5, 21.5),
                                                                    (21, 24, 20)), 26))]
    for name, (row_ptr, col) in cases:
        V = len(row_ptr) - 1
        assert V >= 32768
        say(f"[gpu] whole-GPU tfce_run {name} V={V}")
        op = CreateAdjSet(2.0, 1.0, (row_ptr, col))
        x = smooth_map(row_ptr, col, 77, passes=5, shift=0.1)
        enh = np.zeros_like(x)
        op.run(x, enh)
        ref, st, _ = otfce.tfce(row_ptr, col, x, 2.0, 1.0, exact_pow=True)
        check_close(enh, ref, 1.0, 2.0)
        try:
            _lib.check(lib.tfce_set_single_image_mode(0, 0))
            one = np.zeros_like(x)
            op.run(x, one)
        finally:
            _lib.check(lib.tfce_set_single_image_mode(-1, 0))
        assert np.array_equal(one.view(np.uint32), enh.view(np.uint32))


def canon(labels):
    out = np.full(labels.shape, -1, np.int64)
    act = labels >= 0
    rep = labels[act]
    idx = np.nonzero(act)[0]
    mins = {}
    for r, v in zip(rep, idx):
        if r not in mins or v < mins[r]:
            mins[r] = v
    out[idx] = [mins[r] for r in rep]
    return out


@pytest.mark.parametrize("gname", ["ico5", "ico5_masked", "ico4_w8", "vox26", "vox6_walls"])
def test_partitions_equal_at_every_level(torch, gname):
    from tfce_mediation_amd.tfce import Adjacency
    row_ptr, col = GRAPHS[gname]()
    V = len(row_ptr) - 1
    adj = Adjacency((row_ptr, col), 2.0, 1.0)
    img = smooth_map(row_ptr, col, 5, shift=0.0)
    if gname == "vox6_walls":
        img = img + np.float32(1e-4) * np.arange(V, dtype=np.float32) / V
    lv = otfce.levels(img.max())
    d_img = torch.from_numpy(img).cuda()
    for level in range(0, len(lv), 7):
        say(f"[gpu] labels {gname} level {level}")
        enh = torch.zeros(V, dtype=torch.float32, device="cuda")
        lab = torch.full((V,), -1, dtype=torch.int32, device="cuda")
        st = torch.zeros(1, dtype=torch.int32, device="cuda")
        adj.run_device(d_img, enh, accumulate=False, labels_level=level, labels=lab, status=st)
        _, _, ref = otfce.tfce(row_ptr, col, img, 2.0, 1.0, labels_level=level)
        assert int(st.item()) == len(lv)
        np.testing.assert_array_equal(canon(lab.cpu().numpy()), ref)


@pytest.mark.parametrize("gname,E", [("ico5", 1.0), ("ico5_masked", 1.0), ("ico4_w8", 1.0), ("grid8", 0.5),
                                     ("ico4_w20", 1.0), ("vox26", 0.5)])
def test_batch_max_matches_oracle(torch, gname, E):
    from tfce_mediation_amd.tfce import Adjacency
    row_ptr, col = GRAPHS[gname]()
    V = len(row_ptr) - 1
    P = 24
    say(f"[gpu] batch_max {gname} P={P} E={E}")
    maps = np.stack([smooth_map(row_ptr, col, 100 + p, shift=0.0) for p in range(P)])
    adj = Adjacency((row_ptr, col), 2.0, E)
    d = torch.from_numpy(maps).cuda()
    out = adj.batch_max(d, 3).cpu().numpy()
    vd = adjacency.vertex_density(row_ptr)
    outv = adj.batch_max(d, 3, vdensity=torch.from_numpy(vd).cuda()).cpu().numpy()
    for p in range(P):
        for k, sgn in enumerate((1, -1)):
            x = (sgn * maps[p]).astype(np.float32)
            enh, _, _ = otfce.tfce(row_ptr, col, x, 2.0, E, exact_pow=True)
            s = np.float32(x.max() / np.float32(100))
            exp = np.float32(enh.max() * s)
            expv = ((enh * s) * vd).max()
            if E == 1.0:
                assert out[p, k] == exp and outv[p, k] == expv
            else:
                assert ulp_diff(np.float32([out[p, k]]), np.float32([exp])).max() <= 4
                assert ulp_diff(np.float32([outv[p, k]]), np.float32([expv])).max() <= 4
    one = adj.batch_max(d, 2).cpu().numpy()
    assert np.all(one[:, 0] == 0) and np.array_equal(one[:, 1], out[:, 1])


def test_large_hemisphere_maps(torch):
    from tfce_mediation_amd.tfce import Adjacency
    row_ptr, col = synth.icosphere_csr(7)
    V = len(row_ptr) - 1
    P = 6
    say(f"[gpu] order-7 hemisphere V={V} P={P}")
    maps = np.stack([smooth_map(row_ptr, col, 300 + p, passes=5, shift=0.0) for p in range(P)])
    adj = Adjacency((row_ptr, col), 2.0, 1.0)
    out = adj.batch_max(torch.from_numpy(maps).cuda(), 3).cpu().numpy()
    for p in range(P):
        for k, sgn in enumerate((1, -1)):
            x = (sgn * maps[p]).astype(np.float32)
            enh, _, _ = otfce.tfce(row_ptr, col, x, 2.0, 1.0, exact_pow=True)
            assert out[p, k] == np.float32(enh.max() * np.float32(x.max() / np.float32(100)))


@pytest.mark.parametrize("gname,E", [("ico3", 1.0), ("ico4_shuffled", 1.0), ("ico5", 1.0), ("ico5_masked", 1.0),
                                     ("ico4_w8", 1.0), ("ico4_w12", 1.0), ("ico4_w20", 0.5), ("grid8", 0.5),
                                     ("vox26", 0.5), ("vox6_walls", 1.0)])
def test_lds_sweep_matches_global_sweep(torch, gname, E):
    """Surfaces of at most 13,248 elements run in the LDS-resident sweep; the
    global-memory sweep (knob off) must give the same maxima bit for bit, and
    both match the oracle (asymmetric lists always take the global sweep)."""
    from tfce_mediation_amd import _lib
    from tfce_mediation_amd.tfce import Adjacency
    row_ptr, col = GRAPHS[gname]()
    V = len(row_ptr) - 1
    P = 40
    say(f"[gpu] LDS vs global sweep {gname} V={V} P={P} E={E}")
    maps = np.stack([smooth_map(row_ptr, col, 500 + p, shift=0.0) for p in range(P)])
    if gname == "vox6_walls":
        maps = maps + np.float32(1e-4) * np.arange(V, dtype=np.float32) / V
    adj = Adjacency((row_ptr, col), 2.0, E)
    d = torch.from_numpy(np.ascontiguousarray(maps, np.float32)).cuda()
    lib = _lib.lib()
    try:
        _lib.check(lib.tfce_set_lds_sweep(1))
        lds = adj.batch_max(d, 3).cpu().numpy()
        one = adj.batch_max(d, 2).cpu().numpy()
        _lib.check(lib.tfce_set_lds_sweep(0))
        glob = adj.batch_max(d, 3).cpu().numpy()
    finally:
        _lib.check(lib.tfce_set_lds_sweep(1))
    np.testing.assert_array_equal(lds, glob)
    assert np.all(one[:, 0] == 0) and np.array_equal(one[:, 1], lds[:, 1])
    for p in range(0, P, 5):
        for k, sgn in enumerate((1, -1)):
            x = (sgn * maps[p]).astype(np.float32)
            enh, _, _ = otfce.tfce(row_ptr, col, x, 2.0, E, exact_pow=True)
            exp = np.float32(enh.max() * np.float32(x.max() / np.float32(100)))
            if E == 1.0:
                assert lds[p, k] == exp, (p, k, lds[p, k], exp)
            else:
                assert ulp_diff(np.float32([lds[p, k]]), np.float32([exp])).max() <= 4


def test_mixed_set_splits_between_sweeps(torch):
    """A set with surfaces above and below the LDS limit: the large ones in
    the global sweep, the small ones in the LDS sweep, per-surface maxima
    equal to the all-global launch and to the oracle."""
    from tfce_mediation_amd import _lib
    from tfce_mediation_amd.tfce import Adjacency, SurfaceSet
    graphs = [synth.icosphere_csr(6), synth.icosphere_csr(3), GRAPHS["vox26"](), synth.icosphere_csr(5),
              GRAPHS["vox6_walls"]()]
    sizes = [len(g[0]) - 1 for g in graphs]
    offs = np.cumsum([0] + sizes[:-1])
    width = int(sum(sizes))
    P = 12
    say(f"[gpu] mixed LDS/global set sizes={sizes} P={P}")
    rng = np.random.default_rng(3)
    maps = np.zeros((P, width), np.float32)
    for i, (g, o) in enumerate(zip(graphs, offs)):
        for p in range(P):
            maps[p, o:o + sizes[i]] = smooth_map(g[0], g[1], 700 + 13 * p + i, shift=0.0)
    Es = [1.0, 1.0, 0.5, 1.0, 1.0]
    ss = SurfaceSet([Adjacency(g, 2.0, e) for g, e in zip(graphs, Es)], offs)
    d = torch.from_numpy(maps).cuda()
    lib = _lib.lib()
    try:
        _lib.check(lib.tfce_set_lds_sweep(1))
        got = ss.batch(d, 3).cpu().numpy()
        _lib.check(lib.tfce_set_lds_sweep(0))
        ref = ss.batch(d, 3).cpu().numpy()
    finally:
        _lib.check(lib.tfce_set_lds_sweep(1))
    np.testing.assert_array_equal(got, ref)
    for i, (g, o) in enumerate(zip(graphs, offs)):
        for p in range(0, P, 4):
            for k, sgn in enumerate((1, -1)):
                x = (sgn * maps[p, o:o + sizes[i]]).astype(np.float32)
                enh, _, _ = otfce.tfce(g[0], g[1], x, 2.0, Es[i], exact_pow=True)
                exp = np.float32(enh.max() * np.float32(x.max() / np.float32(100)))
                if Es[i] == 1.0:
                    assert got[i, p, k] == exp, (i, p, k)
                else:
                    assert ulp_diff(np.float32([got[i, p, k]]), np.float32([exp])).max() <= 4


def test_small_mixed_set_with_vdensity(torch):
    """Per-element weights (the FULL sweep) on a set of small surfaces, ring-
    packed and not: every surface takes the global sweep with the kernel
    copy for mixed layouts, maxima equal to the oracle."""
    from tfce_mediation_amd.tfce import Adjacency, SurfaceSet
    graphs = [synth.icosphere_csr(3), GRAPHS["vox26"](), GRAPHS["ico4_w8"]()]
    sizes = [len(g[0]) - 1 for g in graphs]
    offs = np.cumsum([0] + sizes[:-1])
    P = 6
    say(f"[gpu] small mixed set with vdensity sizes={sizes}")
    maps = np.zeros((P, int(sum(sizes))), np.float32)
    vd = np.concatenate([adjacency.vertex_density(g[0]) for g in graphs]).astype(np.float32)
    for i, (g, o) in enumerate(zip(graphs, offs)):
        for p in range(P):
            maps[p, o:o + sizes[i]] = smooth_map(g[0], g[1], 900 + 7 * p + i, shift=0.0)
    ss = SurfaceSet([Adjacency(g, 2.0, 1.0) for g in graphs], offs)
    got = ss.batch(torch.from_numpy(maps).cuda(), 3, vdensity=torch.from_numpy(vd).cuda()).cpu().numpy()
    for i, (g, o) in enumerate(zip(graphs, offs)):
        for p in range(P):
            for k, sgn in enumerate((1, -1)):
                x = (sgn * maps[p, o:o + sizes[i]]).astype(np.float32)
                enh, _, _ = otfce.tfce(g[0], g[1], x, 2.0, 1.0, exact_pow=True)
                s = np.float32(x.max() / np.float32(100))
                assert got[i, p, k] == ((enh * s) * vd[o:o + sizes[i]]).max(), (i, p, k)


@pytest.mark.parametrize("E", [1.0, 0.5])
def test_wide_table_joined_rows_with_overflow(torch, E):
    """26-connected volumes take rows of at most four joined positions per
    element; an element with more (a voxel whose earlier neighbours fall in
    separate groups) is joined from its table row instead.  A lattice of
    high voxels at even coordinates gives every all-odd voxel eight separate
    earlier neighbours (the cube corners around it), and noise gives ties
    and other groupings; the maxima and the FULL-mode enhancement must match
    the oracle, with the LDS-resident sweep off so the global sweep runs."""
    from tfce_mediation_amd import _lib
    from tfce_mediation_amd.tfce import Adjacency
    mask = synth.ellipsoid_mask((24, 26, 22), (11.5, 12.5, 10.5), (11, 12, 10))
    row_ptr, col = adjacency.voxel_csr(mask, 26)
    V = len(row_ptr) - 1
    coords = np.argwhere(mask)
    lattice = np.all(coords % 2 == 0, axis=1)
    rng = np.random.default_rng(31)
    P = 6
    maps = []
    for p in range(P):
        x = rng.standard_normal(V) * (0.05 if p % 2 == 0 else 1.0)
        x[lattice] += 3.0 + 0.5 * p
        x = np.round(x, 1 if p == 4 else 6)  # p = 4: many equal values (same-level neighbours)
        maps.append(x.astype(np.float32))
    maps = np.stack(maps)
    say(f"[gpu] joined-position rows with overflow V={V} P={P} E={E}")
    adj = Adjacency((row_ptr, col), 2.0, E)
    d = torch.from_numpy(np.ascontiguousarray(maps)).cuda()
    lib = _lib.lib()
    try:
        _lib.check(lib.tfce_set_lds_sweep(0))
        got = adj.batch_max(d, 3).cpu().numpy()
        full = []
        for p in range(2):
            e = torch.zeros(V, dtype=torch.float32, device="cuda")
            adj.run_device(d[p].contiguous(), e)
            full.append(e.cpu().numpy())
    finally:
        _lib.check(lib.tfce_set_lds_sweep(1))
    for p in range(P):
        for k, sgn in enumerate((1, -1)):
            x = (sgn * maps[p]).astype(np.float32)
            if x.max() <= 0:
                assert got[p, k] == 0
                continue
            enh, _, _ = otfce.tfce(row_ptr, col, x, 2.0, E, exact_pow=True)
            exp = np.float32(enh.max() * np.float32(x.max() / np.float32(100)))
            check_close(np.float32([got[p, k]]), np.float32([exp]), E, 2.0)
    for p in range(2):
        enh, _, _ = otfce.tfce(row_ptr, col, maps[p], 2.0, E, exact_pow=True)
        check_close(full[p], enh.astype(np.float32), E, 2.0)
