"""Loop restatements of the reference's adjacency builders and a brute-force
component finder.  TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``create_adjac_vertex``  follows ``pyfunc.py:37-46`` (set per vertex, the
  three edges of every face in both directions).
* ``create_adjac_voxel``   follows ``pyfunc.py:48-76`` line by line: labels
  stored in a float copy of the mask, 3x3x3 window clipped at the volume
  edges, ``local_area > 0`` (so label 0 is never a neighbour), 6-connectivity
  keeps only the voxel itself at a clipped window, self removed, lists sorted,
  ``adjacency[0] = []``.
* ``bfs_components``       brute-force connected components of the vertices
  with ``image > T`` (the closed definition of a TFCE cluster for a symmetric
  adjacency), labelled by the smallest member index.
* ``tfce_bruteforce``      TFCE from BFS components in float64 (no shared code
  with the C oracle): sum over the reference's float32 levels of
  size^E * T^H for every active vertex.
"""
import numpy as np


def create_adjac_vertex(vertices, faces):
    adjacency = [set([]) for _ in range(np.asarray(vertices).shape[0])]
    for i in range(faces.shape[0]):
        a, b, c = (int(faces[i, 0]), int(faces[i, 1]), int(faces[i, 2]))
        adjacency[a].add(b); adjacency[a].add(c)
        adjacency[b].add(a); adjacency[b].add(c)
        adjacency[c].add(a); adjacency[c].add(b)
    return adjacency


_SIX = np.array([0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 1, 1, 1, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0]).reshape(3, 3, 3)


def create_adjac_voxel(data_index, data_mask, num_voxel, dirtype=26):
    ind = np.where(data_index)
    dm = np.zeros_like(data_mask, dtype=np.float32)
    x_dim, y_dim, z_dim = data_mask.shape
    adjacency = [set([]) for _ in range(num_voxel)]
    label = 0
    for x, y, z in zip(ind[0], ind[1], ind[2]):
        dm[x, y, z] = label
        label += 1
    for x, y, z in zip(ind[0], ind[1], ind[2]):
        xMin, xMax = max(x - 1, 0), min(x + 1, x_dim - 1)
        yMin, yMax = max(y - 1, 0), min(y + 1, y_dim - 1)
        zMin, zMax = max(z - 1, 0), min(z + 1, z_dim - 1)
        local_area = dm[xMin:xMax + 1, yMin:yMax + 1, zMin:zMax + 1]
        if int(dirtype) == 6:
            if local_area.shape != (3, 3, 3):
                local_area = np.array(dm[x, y, z])
            else:
                local_area = local_area * _SIX
        cV = int(dm[x, y, z])
        for j in np.atleast_1d(local_area[local_area > 0]):
            adjacency[cV].add(int(j))
    out = [[u for u in sorted(s) if u != index] for index, s in enumerate(adjacency)]
    out[0] = []
    return out


def bfs_components(lists, image, T):
    V = len(lists)
    active = np.asarray(image) > T
    label = np.full(V, -1, dtype=np.int64)
    for s in range(V):
        if not active[s] or label[s] >= 0:
            continue
        comp = [s]
        label[s] = s
        head = 0
        while head < len(comp):
            v = comp[head]
            head += 1
            for u in lists[v]:
                if active[u] and label[u] < 0:
                    label[u] = s
                    comp.append(u)
    return label


def reference_levels(max_t, min_t=0.0, delta_t=0.0):
    """The reference's float32 threshold sequence (fast_tfce.hpp:34-39)."""
    T = np.float32(max_t)
    dT = np.float32(delta_t)
    if dT == 0:
        dT = np.float32(T / np.float32(100))
    out = []
    if not (T >= np.float32(min_t)):
        return out
    if not (dT > 0):
        return None
    while T >= np.float32(min_t):
        out.append(T)
        T = np.float32(T - dT)
    return out


def tfce_bruteforce(lists, image, H, E):
    image = np.asarray(image, dtype=np.float32)
    levels = reference_levels(image.max())
    enh = np.zeros(len(lists), dtype=np.float64)
    if not levels:
        return enh
    for T in levels:
        lab = bfs_components(lists, image, T)
        act = lab >= 0
        if not act.any():
            continue
        sizes = np.bincount(lab[act], minlength=len(lists))
        enh[act] += sizes[lab[act]].astype(np.float64) ** float(np.float32(E)) * float(T) ** float(np.float32(H))
    return enh
