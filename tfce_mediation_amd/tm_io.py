"""Drop-in for the TMI container of ``tfce_mediation.tm_io`` (tm_io.py:72-444):
the multimodal, multisurface file the mmr / mmr-lr tools read their data,
masks, surfaces and adjacency from, and the hand-off of that file to the
batched drivers.

* ``read_tm_filetype``  tm_io.py:284-444: the text header (``tmi``, ``format
  binary_little_endian | ascii``, then per element ``element`` / ``dtype`` /
  ``nbytes`` / shape lines, ``history`` lines, ``end_header``) and the
  payload, which ends the file: each element's ``nbytes`` in header order,
  arrays stored transposed (Fortran order of the shapes the header gives).
  Same 11-tuple as the reference.
* ``write_tm_filetype`` tm_io.py:72-282, binary form: the same header lines and
  transposed arrays.  The payload is written in header order (data, masks,
  affines, vertex/face pairs, adjacency objects, column ids), the order the
  reader consumes it in; the reference writer puts column ids before the
  adjacency objects while its header lists them after, which its own reader
  cannot read back when a file holds both.
* ``mmr_inputs``        what tm_multimodality_multisurface_regression.py:410-480
  derives from the file for the statistics (merge_y = image_array[0].T, the
  position array, the vertex-density weights) in the form
  ``randomise.mmr`` takes: per surface the CSR of its adjacency restricted to
  its mask's elements, and its offset in merge_y.

Adjacency objects are pickled Python lists in the reference's files
(``dtype python_object``); unpickling runs code from the file, so they are
read only with ``allow_pickle=True``, which a caller should pass only for
files it trusts.  Without it they come back as None and ``mmr_inputs``
rebuilds the adjacency from the surfaces' faces or the volume masks.
"""
import os
import pickle

import numpy as np

from .adjacency import create_adjac_vertex, create_adjac_voxel, induced_subgraph, to_csr

TMI_VERSION = "1.1"  # tm_filetype_version of the files this writer produces


def read_tm_filetype(tm_file, verbose=True, allow_pickle=False):
    """(element, image_array, masking_array, maskname, affine_array,
    vertex_array, face_array, surfname, adjacency_array, tmi_history,
    columnids) of a TMI file."""
    filesize = os.stat(tm_file).st_size
    element, dtypes, nbytes, nmasked = [], [], [], []
    datashape, maskshape, maskname, affineshape, vertexshape, faceshape = [], [], [], [], [], []
    surfname, adjlength, history = [], [], []
    listlength = 0
    with open(tm_file, "rb") as f:
        first = f.readline().decode("UTF-8").split()
        if not first or first[0] != "tmi":
            raise ValueError("%s is not a TFCE_mediation image (no 'tmi' line)" % tm_file)
        fmt = f.readline().decode("UTF-8").split()
        if not fmt or fmt[0] != "format":
            raise ValueError("%s: unknown file format line" % tm_file)
        filetype = fmt[1]
        while True:
            line = f.readline()
            if not line:
                raise ValueError("%s: header has no end_header line" % tm_file)
            w = line.decode("UTF-8").split()
            if not w:
                continue
            key = w[0]
            if key == "end_header":
                break
            if key == "element":
                element.append(w[1])
            elif key == "dtype":
                dtypes.append(w[1])
            elif key == "nbytes":
                nbytes.append(int(w[1]))
            elif key == "datashape":
                datashape.append((int(w[1]), int(w[2])))
            elif key == "nmasked":
                nmasked.append(int(w[1]))
            elif key == "maskshape":
                maskshape.append((int(w[1]), int(w[2]), int(w[3])))
            elif key == "maskname":
                maskname.append(w[1])
            elif key == "affineshape":
                affineshape.append((int(w[1]), int(w[2])))
            elif key == "vertexshape":
                vertexshape.append((int(w[1]), int(w[2])))
            elif key == "surfname":
                surfname.append(w[1])
            elif key == "faceshape":
                faceshape.append((int(w[1]), int(w[2])))
            elif key == "adjlength":
                adjlength.append(int(w[1]))
            elif key == "history":
                history.append(" ".join(w))
            elif key == "listlength":
                listlength = int(w[1])
        img, masks, affines, verts, faces, adjs, cols = [], [], [], [], [], [], []
        cnt = dict(mask=0, affine=0, vertex=0, face=0, adjacency=0)
        if filetype.startswith("binary"):
            order = "<" if filetype == "binary_little_endian" else ">"
            position = filesize - sum(nbytes)
            for e, name in enumerate(element):
                f.seek(position)
                if verbose:
                    print("reading %s" % name)
                if name == "adjacency_object":
                    if allow_pickle:
                        obj = pickle.load(f)
                        adjs.append(np.array(obj[:adjlength[cnt["adjacency"]]], dtype=object))
                    else:
                        adjs.append(None)
                    cnt["adjacency"] += 1
                    position += nbytes[e]
                    continue
                dt = np.dtype(dtypes[e]).newbyteorder(order)
                arr = np.frombuffer(f.read(nbytes[e]), dtype=dt).astype(np.dtype(dtypes[e]))
                position += nbytes[e]
                if name == "data_array":
                    nv, ns = datashape[0]
                    img.append(arr[:nv * ns].reshape(ns, nv).T)
                elif name == "masking_array":
                    sh = maskshape[cnt["mask"]]
                    m = arr[:sh[0] * sh[1] * sh[2]].reshape(sh[2], sh[1], sh[0]).T
                    masks.append(np.array(m, dtype=bool))
                    cnt["mask"] += 1
                elif name == "affine":
                    sh = affineshape[cnt["affine"]]
                    affines.append(arr[:sh[0] * sh[1]].reshape(sh[1], sh[0]).T)
                    cnt["affine"] += 1
                elif name == "vertex":
                    sh = vertexshape[cnt["vertex"]]
                    verts.append(arr[:sh[0] * sh[1]].reshape(sh[1], sh[0]).T)
                    cnt["vertex"] += 1
                elif name == "face":
                    sh = faceshape[cnt["face"]]
                    faces.append(arr[:sh[0] * sh[1]].reshape(sh[1], sh[0]).T)
                    cnt["face"] += 1
                elif name == "column_id":
                    cols.append(arr[:listlength])
        elif filetype == "ascii":
            for e, name in enumerate(element):
                if name == "data_array":
                    nv, ns = datashape[0]
                    img.append(np.array([f.readline().split() for _ in range(nv)], dtype=np.float32))
                elif name == "masking_array":
                    idx = np.array([f.readline().split() for _ in range(nmasked[cnt["mask"]])], dtype=np.int64)
                    m = np.zeros(maskshape[cnt["mask"]], dtype=bool)
                    if idx.size:
                        m[idx[:, 0], idx[:, 1], idx[:, 2]] = True
                    masks.append(m)
                    cnt["mask"] += 1
                elif name == "affine":
                    sh = affineshape[cnt["affine"]]
                    affines.append(np.array([f.readline().split() for _ in range(sh[0])], dtype=np.float32))
                    cnt["affine"] += 1
                elif name == "vertex":
                    sh = vertexshape[cnt["vertex"]]
                    verts.append(np.array([f.readline().split() for _ in range(sh[0])], dtype=np.float32))
                    cnt["vertex"] += 1
                elif name == "face":
                    sh = faceshape[cnt["face"]]
                    faces.append(np.array([f.readline().split() for _ in range(sh[0])], dtype=np.int32))
                    cnt["face"] += 1
                elif name == "column_id":
                    cols.append(np.array([f.readline().strip() for _ in range(listlength)], dtype=dtypes[e]))
        else:
            raise ValueError("unknown TMI file type: %s" % filetype)
    return element, img, masks, maskname, affines, verts, faces, surfname, adjs, history, cols


def write_tm_filetype(outname, image_array=None, masking_array=(), maskname=(), affine_array=(), vertex_array=(),
                      face_array=(), surfname=(), adjacency_array=(), columnids=None, tmi_history=None,
                      timestamp=0):
    """Binary TMI file of the given parts; the history line records what was
    added (``history mode_add <timestamp> <data> <masks> <affines>
    <surfaces> <adjacency objects>``).  Returns the file name."""
    if not outname.endswith("tmi"):
        outname += ".tmi"
    masks = [np.asarray(m, dtype=bool) for m in masking_array]
    affines = [np.asarray(a, dtype=np.float32) for a in affine_array]
    verts = [np.asarray(v, dtype=np.float32) for v in vertex_array]
    faces = [np.asarray(t, dtype=np.uint32) for t in face_array]
    adjs = [list(a) for a in adjacency_array]
    history = list(tmi_history or [])
    head = ["tmi", "format binary_%s_endian %s" % ("little", TMI_VERSION), "comment made with TFCE_mediation"]
    payload = []
    if image_array is not None:
        img = np.asarray(image_array, dtype=np.float32)
        nv, ns = (img.shape[0], 1) if img.ndim == 1 else img.shape
        head += ["element data_array", "dtype float32", "nbytes %d" % img.nbytes, "datashape %d %d" % (nv, ns)]
        payload.append(np.ascontiguousarray(img.reshape(nv, ns).T).astype("<f4").tobytes())
    for i, m in enumerate(masks):
        head += ["element masking_array", "dtype uint8", "nbytes %d" % m.nbytes, "nmasked %d" % int(m.sum()),
                 "maskshape %d %d %d" % m.shape, "maskname %s" % (maskname[i] if i < len(maskname) else "unknown")]
        payload.append(np.ascontiguousarray((m * 1).T).astype(np.uint8).tobytes())
    for a in affines:
        head += ["element affine", "dtype float32", "nbytes %d" % a.nbytes, "affineshape %d %d" % a.shape]
        payload.append(np.ascontiguousarray(a.T).astype("<f4").tobytes())
    for i, (v, t) in enumerate(zip(verts, faces)):
        head += ["surfname %s" % (surfname[i] if i < len(surfname) else "unknown"),
                 "element vertex", "dtype float32", "nbytes %d" % v.nbytes, "vertexshape %d %d" % v.shape,
                 "element face", "dtype uint32", "nbytes %d" % t.nbytes, "faceshape %d %d" % t.shape]
        payload.append(np.ascontiguousarray(v.T).astype("<f4").tobytes())
        payload.append(np.ascontiguousarray(t.T).astype("<u4").tobytes())
    for a in adjs:
        blob = pickle.dumps(a, -1)
        head += ["element adjacency_object", "dtype python_object", "nbytes %d" % len(blob), "adjlength %d" % len(a)]
        payload.append(blob)
    if columnids is not None:
        c = np.asarray(columnids)
        head += ["element column_id", "dtype %s" % c.dtype, "nbytes %d" % c.nbytes, "listlength %d" % len(c)]
        payload.append(c.tobytes())
    history.append("history mode_add %d %d %d %d %d %d" % (timestamp, 1 if image_array is not None else 0, len(masks),
                                                         len(affines), len(verts), len(adjs)))
    head += history + ["end_header"]
    with open(outname, "wb") as f:
        f.write(("\n".join(head) + "\n").encode("UTF-8"))
        for p in payload:
            f.write(p)
    return outname


def mmr_inputs(tmi, adjacency=None, weight=True):
    """From the 11-tuple of ``read_tm_filetype``: (merge_y [n, V], per-surface
    CSR restricted to its mask, offsets, vdensity [V] or None, implicit_zero
    per surface) for ``randomise.mmr``.  ``adjacency``: ragged lists per mask
    (over the mask's full element set, as the file stores them) to use instead
    of the file's; when neither is there, vertex surfaces are rebuilt from the
    file's faces (surface i for mask i) and volume masks with 26-connectivity.
    vdensity follows tm_multimodality_multisurface_regression.py:449-462."""
    _, img, masks, _, _, verts, faces, _, adjs, _, _ = tmi
    merge_y = np.ascontiguousarray(np.asarray(img[0], dtype=np.float32).T)
    offsets, surfaces, zero, vd = [], [], [], []
    off = 0
    for i, m in enumerate(masks):
        vertex_mask = m.shape[2] == 1
        lists = None
        if adjacency is not None:
            lists = adjacency[i]
        elif i < len(adjs) and adjs[i] is not None:
            lists = adjs[i]
        elif vertex_mask:
            lists = create_adjac_vertex(verts[i], faces[i])
        else:
            lists = create_adjac_voxel(m, m, int(m.sum()), 26)
        rp, col = to_csr([sorted(int(u) for u in row) for row in lists])
        deg = np.diff(rp).astype(np.float64)
        if vertex_mask:
            keep = m[:, 0, 0]
            rp, col = induced_subgraph(rp, col, keep)
            deg = deg[keep]
            zero.append(not bool(keep.all()))
        else:
            zero.append(False)
        nel = len(rp) - 1
        surfaces.append((rp, col))
        offsets.append(off)
        off += nel
        if weight and deg.size:
            vd.append((1 - deg / deg.max() + deg.mean() / deg.max()).astype(np.float32))
    if off != merge_y.shape[1]:
        raise ValueError("the masks hold %d elements, the data array %d" % (off, merge_y.shape[1]))
    vdensity = np.concatenate(vd) if weight and vd else None
    return merge_y, surfaces, offsets, vdensity, zero
