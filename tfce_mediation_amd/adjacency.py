"""Adjacency builders and CSR conversion for the TFCE operator.

* ``create_adjac_vertex`` - triangle-edge adjacency, as ``pyfunc.py:37-46``
  (each face adds its other two vertices to each of its vertices' sets, so a
  face that repeats a vertex puts that vertex in its own set).
* ``create_adjac_voxel`` - 6/26-neighbourhood adjacency of a voxel mask with the
  reference's quirks (``pyfunc.py:48-76``): the first in-mask voxel (label 0) is
  nobody's neighbour and has an empty list; with 6-connectivity a voxel whose
  3x3x3 window is clipped by the volume edge keeps no neighbours.
* ``to_csr`` - ragged adjacency (list of sets/lists, object array from
  ``np.load``) or (row_ptr, col) -> int32 CSR.
* ``merge_csr`` / ``merge_adjacency_array`` - concatenate surfaces with index
  offsets (``tm_func.py:521-540``).
* ``vertex_density`` - the degree-based weight of
  ``STEP_1_vertex_tfce_multiple_regression.py:165-173`` and
  ``tm_multimodality_multisurface_regression.py:449-462``.
All of these are vectorised host code run once per analysis;
``voxel_csr_device`` and ``faces_to_csr_device`` build the same CSR on the GPU
(``tfce_adj_voxel_csr`` / ``tfce_adj_mesh_csr``) from a device-resident mask or
face list.
"""
import numpy as np


def faces_to_csr(faces, n_vertices):
    faces = np.asarray(faces, dtype=np.int64)
    a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
    src = np.concatenate([a, a, b, b, c, c])
    dst = np.concatenate([b, c, a, c, a, b])
    key = np.unique(src * n_vertices + dst)
    src = (key // n_vertices).astype(np.int64)
    dst = (key % n_vertices).astype(np.int32)
    row_ptr = np.zeros(n_vertices + 1, dtype=np.int64)
    np.add.at(row_ptr, src + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.int32)
    return row_ptr, dst


def create_adjac_vertex(vertices, faces):
    """Ragged (list of sets) form, as ``pyfunc.py:37-46`` returns it."""
    row_ptr, col = faces_to_csr(faces, np.asarray(vertices).shape[0])
    return [set(col[row_ptr[i]:row_ptr[i + 1]].tolist()) for i in range(len(row_ptr) - 1)]


_OFFSETS26 = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)
              if (dx, dy, dz) != (0, 0, 0)]
_OFFSETS6 = [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]


def voxel_csr(mask, dirtype=26):
    """CSR of ``create_adjac_voxel(mask, mask, mask.sum(), dirtype)``.

    Elements are the in-mask voxels in C order (``np.where`` order); rows are
    sorted ascending like the reference's ``sorted(i)``.
    """
    mask = np.asarray(mask, dtype=bool)
    X, Y, Z = mask.shape
    ind = np.nonzero(mask)
    n = ind[0].size
    label = np.full(mask.shape, -1, dtype=np.int64)
    label[ind] = np.arange(n)
    x, y, z = (np.asarray(i, dtype=np.int64) for i in ind)
    offsets = _OFFSETS26 if int(dirtype) == 26 else _OFFSETS6
    if int(dirtype) == 6:
        interior = (x > 0) & (x < X - 1) & (y > 0) & (y < Y - 1) & (z > 0) & (z < Z - 1)
    srcs, dsts = [], []
    for dx, dy, dz in offsets:
        nx, ny, nz = x + dx, y + dy, z + dz
        ok = (nx >= 0) & (nx < X) & (ny >= 0) & (ny < Y) & (nz >= 0) & (nz < Z)
        if int(dirtype) == 6:
            ok &= interior
        src = np.nonzero(ok)[0]
        nb = label[nx[ok], ny[ok], nz[ok]]
        good = nb > 0  # label 0 (first voxel) and out-of-mask are never neighbours
        srcs.append(src[good])
        dsts.append(nb[good])
    src = np.concatenate(srcs)
    dst = np.concatenate(dsts)
    keep = src != 0  # adjacency[0] = []
    src, dst = src[keep], dst[keep]
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row_ptr, src + 1, 1)
    return np.cumsum(row_ptr).astype(np.int32), dst.astype(np.int32)


def create_adjac_voxel(data_index, data_mask, num_voxel, dirtype=26):
    """Ragged form with the reference signature (``pyfunc.py:48-76``)."""
    mask = np.asarray(data_index, dtype=bool)
    row_ptr, col = voxel_csr(mask, dirtype)
    assert row_ptr.shape[0] - 1 == int(num_voxel)
    return [col[row_ptr[i]:row_ptr[i + 1]].tolist() for i in range(int(num_voxel))]


def to_csr(adjacency):
    """Ragged adjacency or (row_ptr, col) -> (row_ptr int32[V+1], col int32[nnz])."""
    if isinstance(adjacency, tuple) and len(adjacency) == 2:
        row_ptr = np.ascontiguousarray(adjacency[0], dtype=np.int32)
        col = np.ascontiguousarray(adjacency[1], dtype=np.int32)
        return row_ptr, col
    if hasattr(adjacency, "tocsr"):
        m = adjacency.tocsr()
        return (np.ascontiguousarray(m.indptr, dtype=np.int32),
                np.ascontiguousarray(m.indices, dtype=np.int32))
    lists = [np.fromiter((int(u) for u in row), dtype=np.int64) for row in adjacency]
    lengths = np.array([len(r) for r in lists], dtype=np.int64)
    row_ptr = np.zeros(len(lists) + 1, dtype=np.int64)
    np.cumsum(lengths, out=row_ptr[1:])
    col = np.concatenate(lists) if lists and row_ptr[-1] > 0 else np.zeros(0, dtype=np.int64)
    return row_ptr.astype(np.int32), col.astype(np.int32)


def merge_csr(csrs):
    """Block-diagonal concatenation of CSR graphs (tm_func.py:521-540)."""
    rows, cols, off_v, off_e = [np.zeros(1, dtype=np.int64)], [], 0, 0
    for rp, c in csrs:
        rp = np.asarray(rp, dtype=np.int64)
        rows.append(rp[1:] + off_e)
        cols.append(np.asarray(c, dtype=np.int64) + off_v)
        off_v += rp.shape[0] - 1
        off_e += int(rp[-1])
    return (np.concatenate(rows).astype(np.int32),
            np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32))


def merge_adjacency_array(adjacent_range, adjacency_array):
    """Ragged merge with offsets, same result as ``tm_func.py:521-540``."""
    csrs = [to_csr(adjacency_array[e]) for e in adjacent_range]
    rp, c = merge_csr(csrs)
    return [c[rp[i]:rp[i + 1]].tolist() for i in range(rp.shape[0] - 1)]


def induced_subgraph(row_ptr, col, keep):
    """CSR restricted to the elements where ``keep`` is true, re-indexed."""
    keep = np.asarray(keep, dtype=bool)
    newid = np.full(keep.shape[0], -1, dtype=np.int64)
    newid[keep] = np.arange(int(keep.sum()))
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    src = np.repeat(np.arange(keep.shape[0]), np.diff(row_ptr))
    dst = np.asarray(col, dtype=np.int64)
    ok = keep[src] & keep[dst]
    s, d = newid[src[ok]], newid[dst[ok]]
    rp = np.zeros(int(keep.sum()) + 1, dtype=np.int64)
    np.add.at(rp, s + 1, 1)
    return np.cumsum(rp).astype(np.int32), d.astype(np.int32)


def vertex_density(row_ptr):
    """1 - deg/max(deg) + mean(deg)/max(deg), float32 (float64 arithmetic)."""
    deg = np.diff(np.asarray(row_ptr, dtype=np.int64)).astype(np.float64)
    return np.array((1 - (deg / deg.max()) + (deg.mean() / deg.max())), dtype=np.float32)


def locality_order(row_ptr, col):
    """new_of_old[v]: reverse Cuthill-McKee order of the adjacency, computed
    by the library's host code (tfce_locality_order)."""
    from . import _lib
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    V = row_ptr.shape[0] - 1
    out = np.empty(V, dtype=np.int32)
    lib = _lib.load_library()
    rc = lib.tfce_locality_order(row_ptr.ctypes.data, col.ctypes.data if col.size else None, V, out.ctypes.data)
    if rc != 0:
        raise _lib.TfceHipError(f"tfce_locality_order: {lib.tfce_last_error().decode()}")
    return out


def relabel_csr(row_ptr, col, new_of_old):
    """The same graph with element v renamed new_of_old[v]; each new list is
    sorted."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    col = np.asarray(col, dtype=np.int64)
    new_of_old = np.asarray(new_of_old, dtype=np.int64)
    V = row_ptr.shape[0] - 1
    src = new_of_old[np.repeat(np.arange(V), np.diff(row_ptr))]
    dst = new_of_old[col]
    o = np.lexsort((dst, src))
    rp = np.zeros(V + 1, dtype=np.int32)
    np.cumsum(np.bincount(src, minlength=V), out=rp[1:])
    return rp, dst[o].astype(np.int32)


def is_symmetric(row_ptr, col):
    """Every listed pair (v, u), u != v, is also listed as (u, v)."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    col = np.asarray(col, dtype=np.int64)
    V = row_ptr.shape[0] - 1
    src = np.repeat(np.arange(V), np.diff(row_ptr))
    keep = src != col
    fwd = np.unique(src[keep] * V + col[keep])
    bwd = np.unique(col[keep] * V + src[keep])
    return fwd.shape == bwd.shape and bool(np.array_equal(fwd, bwd))


def locality_layout(surfaces, offsets, width):
    """Element layout for the batched sweep: every symmetric surface is
    relabelled with its locality order (neighbours close in memory); others
    keep their order (their tie rule depends on the caller's indices).
    Returns (relabelled CSRs, cols) with cols[new column] = old map column, so
    maps[:, cols] are the maps in the new layout.  Relabelling a symmetric
    graph changes no partition, no TFCE value and no maximum."""
    cols = np.arange(width, dtype=np.int64)
    out, cache = [], {}
    for (rp, col), off in zip(surfaces, offsets):
        key = (id(rp), id(col))
        if key not in cache:
            V = len(rp) - 1
            if is_symmetric(rp, col):
                new_of_old = locality_order(rp, col)
                old_of_new = np.empty(V, dtype=np.int64)
                old_of_new[new_of_old] = np.arange(V)
                cache[key] = (relabel_csr(rp, col, new_of_old), old_of_new)
            else:
                cache[key] = ((rp, col), np.arange(V, dtype=np.int64))
        csr, old_of_new = cache[key]
        out.append(csr)
        cols[off:off + len(old_of_new)] = off + old_of_new
    return out, cols


def voxel_csr_device(mask, dirtype=26, stream=None):
    """``voxel_csr`` on the GPU: (row_ptr, col) int32 device tensors of the
    in-mask voxels of ``mask`` (numpy or torch, any dtype; nonzero = in)."""
    import ctypes
    import torch
    from . import _lib
    m = torch.as_tensor(np.asarray(mask) != 0 if not isinstance(mask, torch.Tensor) else mask != 0)
    m = m.to(device="cuda", dtype=torch.uint8).contiguous()
    X, Y, Z = (int(d) for d in m.shape)
    lib = _lib.lib()
    V, nnz = ctypes.c_int32(), ctypes.c_int64()
    _lib.check(lib.tfce_adj_voxel_csr(_lib.ptr(m), X, Y, Z, int(dirtype), None, None, ctypes.byref(V),
                                      ctypes.byref(nnz), _lib.stream_ptr(stream)))
    rp = torch.empty(V.value + 1, dtype=torch.int32, device="cuda")
    col = torch.empty(max(nnz.value, 1), dtype=torch.int32, device="cuda")
    _lib.check(lib.tfce_adj_voxel_csr(_lib.ptr(m), X, Y, Z, int(dirtype), _lib.ptr(rp), _lib.ptr(col),
                                      ctypes.byref(V), ctypes.byref(nnz), _lib.stream_ptr(stream)))
    return rp, col[:nnz.value]


def faces_to_csr_device(faces, n_vertices, stream=None):
    """``faces_to_csr`` on the GPU: (row_ptr, col) int32 device tensors."""
    import ctypes
    import torch
    from . import _lib
    f = torch.as_tensor(np.asarray(faces) if not isinstance(faces, torch.Tensor) else faces)
    f = f.to(device="cuda", dtype=torch.int32).contiguous().reshape(-1, 3)
    lib = _lib.lib()
    nnz = ctypes.c_int64()
    F = int(f.shape[0])
    _lib.check(lib.tfce_adj_mesh_csr(_lib.ptr(f), F, int(n_vertices), None, None, ctypes.byref(nnz),
                                     _lib.stream_ptr(stream)))
    rp = torch.empty(int(n_vertices) + 1, dtype=torch.int32, device="cuda")
    col = torch.empty(max(nnz.value, 1), dtype=torch.int32, device="cuda")
    _lib.check(lib.tfce_adj_mesh_csr(_lib.ptr(f), F, int(n_vertices), _lib.ptr(rp), _lib.ptr(col),
                                     ctypes.byref(nnz), _lib.stream_ptr(stream)))
    return rp, col[:nnz.value]


def geodesic_pairs(vertices, faces, max_distance, n_threads=0, sources=None):
    """(i, j, distance) arrays of every vertex pair i < j of a triangle mesh
    closer than max_distance along the surface, in (i, j) order: the exact
    geodesic distances adjacency.pyx:44-106 takes from lib/geodesic, computed
    by the library's window propagation on host threads (one-time set-up
    work, like the reference's).  sources: (begin, end) range of i (default
    every vertex)."""
    import ctypes
    from . import _lib
    v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float32).reshape(-1, 3))
    f = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).reshape(-1, 3))
    lib = _lib.host_lib()
    h = ctypes.c_void_p()
    b, e = (0, 0) if sources is None else (int(sources[0]), int(sources[1]))
    if sources is not None and e <= b:
        return np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.float64)
    _lib.check(lib.tfce_geodesic_pairs(v.ctypes.data, len(v), f.ctypes.data, len(f), float(max_distance), b, e,
                                       int(n_threads), ctypes.byref(h)))
    try:
        n = ctypes.c_int64()
        _lib.check(lib.tfce_geodesic_count(h, ctypes.byref(n)))
        i = np.empty(n.value, np.int32)
        j = np.empty(n.value, np.int32)
        d = np.empty(n.value, np.float64)
        _lib.check(lib.tfce_geodesic_copy(h, i.ctypes.data, j.ctypes.data, d.ctypes.data))
    finally:
        lib.tfce_geodesic_destroy(h)
    return i, j, d


def geodesic_adjacency(vertices, faces, thresholds, n_threads=0):
    """adjacency.compute(v, f, thresholds) (adjacency.pyx:44-79): for each
    threshold, the list per vertex of the vertices closer than it along the
    surface, in increasing order (the order the reference's loops append
    them in)."""
    thresholds = np.atleast_1d(np.asarray(thresholds, dtype=np.float64))
    nv = int(np.asarray(vertices).reshape(-1, 3).shape[0])
    i, j, d = geodesic_pairs(vertices, faces, float(thresholds.max()), n_threads)
    out = []
    for t in thresholds:
        keep = d < t
        a = np.concatenate([i[keep], j[keep]])
        b = np.concatenate([j[keep], i[keep]])
        order = np.lexsort((b, a))
        a, b = a[order], b[order]
        starts = np.searchsorted(a, np.arange(nv + 1))
        out.append([b[starts[v]:starts[v + 1]].tolist() for v in range(nv)])
    return out


def compute_distance_parallel(v, f, threshold, i):
    """adjacency.compute_distance_parallel (adjacency.pyx:82-106): [[i, j,
    distance]] for every vertex j > i closer than threshold to vertex i."""
    a, b, d = geodesic_pairs(v, f, float(np.max(threshold)), n_threads=1, sources=(int(i), int(i) + 1))
    return [[int(x), int(y), float(z)] for x, y, z in zip(a, b, d)]
