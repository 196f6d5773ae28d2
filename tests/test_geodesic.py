"""Geodesic-distance adjacency (adjacency.pyx:44-106, lib/geodesic): the
library's exact window propagation (host code, no GPU) against closed forms
and against the reference's own exact algorithm, compiled from its vendored
headers into oracle/_ref/libgeoref.so where the reference tree is present."""
import numpy as np
import pytest

from oracle import geodesic as ogeo
from tfce_mediation_amd import adjacency, synth
from tfce_mediation_amd._lib import TfceHipError


def grid(nx, ny, h=1.0):
    xs, ys = np.meshgrid(np.arange(nx) * h, np.arange(ny) * h)
    # float32 coordinates, as the library takes them (adjacency.pyx:44-46)
    v = np.column_stack([xs.ravel(), ys.ravel(), np.zeros(nx * ny)]).astype(np.float32).astype(np.float64)
    f = []
    for y in range(ny - 1):
        for x in range(nx - 1):
            a = y * nx + x
            f += [[a, a + 1, a + nx + 1], [a, a + nx + 1, a + nx]]
    return v, np.array(f)


def bumpy_sphere(order, radius, noise, seed):
    v, f = synth.icosphere(order)
    rng = np.random.default_rng(seed)
    v = np.asarray(v) * radius * (1 + noise * rng.standard_normal((len(v), 1)))
    return v, np.asarray(f)


def as_dict(i, j, d):
    return {(int(a), int(b)): float(x) for a, b, x in zip(i, j, d)}


def test_flat_grid_distances_are_euclidean():
    """On a flat, convex sheet every geodesic is the straight segment: each
    pair closer than R, and only those, with its Euclidean distance."""
    v, f = grid(9, 7, 0.7)
    R = 2.3
    got = as_dict(*adjacency.geodesic_pairs(v, f, R))
    n = len(v)
    exp = {}
    for a in range(n):
        for b in range(a + 1, n):
            d = float(np.linalg.norm(v[a] - v[b]))
            if d < R:
                exp[(a, b)] = d
    assert sorted(got) == sorted(exp)
    for k, d in exp.items():
        assert abs(got[k] - d) <= 1e-12 * max(1.0, d)


def test_convex_sphere_neighbours_at_edge_length():
    """On a convex polyhedron the geodesic between the ends of an edge is
    the edge, and no pair is closer than its straight-line distance."""
    v, f = synth.icosphere(3)
    v = (np.asarray(v) * 10.0).astype(np.float32).astype(np.float64)
    i, j, d = adjacency.geodesic_pairs(v, f, 2.0)
    got = as_dict(i, j, d)
    rp, col = synth.icosphere_csr(3)
    for a in range(len(v)):
        for b in col[rp[a]:rp[a + 1]]:
            if a < b:
                e = float(np.linalg.norm(v[a] - v[b]))
                if e < 2.0:
                    assert abs(got[(a, int(b))] - e) <= 1e-12 * e
    chord = np.linalg.norm(v[i] - v[j], axis=1)
    assert np.all(d >= chord * (1 - 1e-12))


@pytest.mark.skipif(not ogeo.have_reference(), reason="the reference's geodesic library is not built here")
@pytest.mark.parametrize("case", ["bumpy", "open_cap", "wide_radius"])
def test_matches_reference_exact_algorithm(case):
    """Same pairs and distances as lib/geodesic's exact algorithm on meshes
    with saddle vertices (a bumpy sphere), with a boundary (an open cap:
    boundary vertices bend geodesics too) and with a radius of several
    edges."""
    if case == "bumpy":
        v, f = bumpy_sphere(3, 10.0, 0.05, 1)
        R = 3.0
    elif case == "open_cap":
        v, f = bumpy_sphere(3, 10.0, 0.05, 2)
        keep = v[f].mean(1)[:, 2] > -1.0
        f = f[keep]
        used = np.unique(f)
        remap = -np.ones(len(v), int)
        remap[used] = np.arange(len(used))
        v, f = v[used], remap[f]
        R = 3.0
    else:
        v, f = bumpy_sphere(4, 10.0, 0.03, 3)
        R = 4.0
    got = as_dict(*adjacency.geodesic_pairs(v, f, R))
    ref = as_dict(*ogeo.ref_pairs(v, f, R))
    # pairs within rounding of R may fall either side
    near = {k for k, d in list(got.items()) + list(ref.items()) if abs(d - R) <= 1e-9}
    assert sorted(set(got) - near) == sorted(set(ref) - near)
    for k in set(got) & set(ref):
        assert abs(got[k] - ref[k]) <= 1e-12 * max(1.0, ref[k]), (k, got[k], ref[k])


def test_adjacency_lists_per_threshold():
    """adjacency.compute's return: one list of neighbour lists per threshold,
    symmetric, in increasing order, nested as the thresholds are."""
    v, f = bumpy_sphere(3, 10.0, 0.03, 4)
    i, j, d = adjacency.geodesic_pairs(v, f, 3.0)
    adj2, adj3 = adjacency.geodesic_adjacency(v, f, [2.0, 3.0])
    for t, adj in ((2.0, adj2), (3.0, adj3)):
        assert len(adj) == len(v)
        pairs = {(a, b) for a, row in enumerate(adj) for b in row}
        exp = {(int(a), int(b)) for a, b, x in zip(i, j, d) if x < t}
        exp |= {(b, a) for a, b in exp}
        assert pairs == exp
        assert all(row == sorted(row) for row in adj)
    assert all(set(r2) <= set(r3) for r2, r3 in zip(adj2, adj3))


def test_rejects_bad_faces():
    v, f = grid(3, 3)
    f = f.copy()
    f[0, 0] = 99
    with pytest.raises(TfceHipError):
        adjacency.geodesic_pairs(v, f, 1.5)
    with pytest.raises(TfceHipError):
        adjacency.geodesic_pairs(v, grid(3, 3)[1], -1.0)


def test_one_source_matches_the_full_run():
    """compute_distance_parallel (adjacency.pyx:82-106) for one vertex: the
    full run's pairs of that source."""
    v, f = bumpy_sphere(3, 10.0, 0.03, 5)
    i, j, d = adjacency.geodesic_pairs(v, f, 3.0)
    for s in (0, 17, len(v) - 2):
        got = adjacency.compute_distance_parallel(v, f, 3.0, s)
        exp = [[int(a), int(b), float(x)] for a, b, x in zip(i, j, d) if a == s]
        assert got == exp
