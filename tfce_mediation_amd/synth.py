"""Synthetic workloads for the permutation hot path (seeded, no downloads).

* ``icosphere(order)``      - subdivided icosahedron, V = 10*4^order + 2,
                              triangle-edge CSR (degree 5 at the 12 original
                              vertices, 6 elsewhere).
* ``ellipsoid_mask(...)``   - boolean voxel mask; CSR via adjacency.voxel_csr.
* ``mesh_smooth``           - k passes of x_v <- mean(x_v and its neighbours).
* ``gaussian_smooth3d``     - separable Gaussian (FWHM in voxels).
* ``config1`` .. ``config5`` - the five workloads of this project's benchmark
  table (small vertex GLM, voxel GLM, two-hemisphere order-7 GLM, uneven
  multi-modal set, mediation).  ``n_perm`` and sizes can be overridden so
  tests run reduced copies of the same generators.
Heavy arrays can be generated directly on a torch device (``device=``).
"""
import numpy as np

from .adjacency import faces_to_csr, voxel_csr

_ICO_V = np.array([(-1, 1.618033988749895, 0), (1, 1.618033988749895, 0),
                   (-1, -1.618033988749895, 0), (1, -1.618033988749895, 0),
                   (0, -1, 1.618033988749895), (0, 1, 1.618033988749895),
                   (0, -1, -1.618033988749895), (0, 1, -1.618033988749895),
                   (1.618033988749895, 0, -1), (1.618033988749895, 0, 1),
                   (-1.618033988749895, 0, -1), (-1.618033988749895, 0, 1)])
_ICO_F = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9),
                   (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2),
                   (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
                   (8, 6, 7), (9, 8, 1)], dtype=np.int64)


def icosphere(order):
    """(vertices [V,3] float64 on the unit sphere, faces [F,3] int64)."""
    v = _ICO_V / np.linalg.norm(_ICO_V, axis=1, keepdims=True)
    f = _ICO_F.copy()
    for _ in range(int(order)):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // nv, uk % nv
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1) + nv  # midpoints of (01, 12, 20) per face
        ab, bc, ca = m[0], m[1], m[2]
        f = np.concatenate([np.stack([f[:, 0], ab, ca], 1), np.stack([f[:, 1], bc, ab], 1),
                            np.stack([f[:, 2], ca, bc], 1), np.stack([ab, bc, ca], 1)])
        v = np.concatenate([v, mid])
    return v, f


def icosphere_csr(order):
    v, f = icosphere(order)
    return faces_to_csr(f, v.shape[0])


def ellipsoid_mask(shape=(91, 109, 91), center=(45, 54, 45), semi=(36, 44, 35)):
    g = np.ogrid[:shape[0], :shape[1], :shape[2]]
    r = sum(((gi - c) / s) ** 2 for gi, c, s in zip(g, center, semi))
    return r <= 1.0


def bfs_ball(row_ptr, col, seed_vertex, size):
    """First ``size`` vertices reached by BFS from ``seed_vertex``."""
    seen = np.zeros(row_ptr.shape[0] - 1, dtype=bool)
    out = [int(seed_vertex)]
    seen[seed_vertex] = True
    head = 0
    while head < len(out) and len(out) < size:
        v = out[head]
        head += 1
        for u in col[row_ptr[v]:row_ptr[v + 1]]:
            if not seen[u]:
                seen[u] = True
                out.append(int(u))
                if len(out) >= size:
                    break
    return np.array(out, dtype=np.int64)


def _torch():
    import torch
    return torch


def smoothing_operator(row_ptr, col, device="cpu"):
    """Sparse (A + I) with rows normalised (torch CSR, float32)."""
    torch = _torch()
    V = row_ptr.shape[0] - 1
    deg = np.diff(row_ptr.astype(np.int64))
    rows = np.concatenate([np.repeat(np.arange(V), deg), np.arange(V)])
    cols = np.concatenate([col.astype(np.int64), np.arange(V)])
    w = np.concatenate([1.0 / (deg + 1.0)[np.repeat(np.arange(V), deg)], 1.0 / (deg + 1.0)])
    idx = torch.from_numpy(np.stack([rows, cols]))
    S = torch.sparse_coo_tensor(idx, torch.from_numpy(w.astype(np.float32)), (V, V)).coalesce()
    return S.to(device)


def mesh_smooth(Y, S, passes):
    """Y [n, V] (torch) smoothed along V with the row-normalised operator S."""
    torch = _torch()
    Z = Y.t().contiguous()
    for _ in range(passes):
        Z = torch.sparse.mm(S, Z)
    return Z.t().contiguous()


def gaussian_smooth3d(vol, fwhm):
    """vol [n, X, Y, Z] torch float32, separable Gaussian of the given FWHM
    (voxels), zero padded.  Written as shifted weighted sums (elementwise
    kernels only): a library convolution would compile a kernel per new
    volume shape on first use, which takes tens of seconds."""
    torch = _torch()
    sigma = fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    r = int(np.ceil(3 * sigma))
    t = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (t / sigma) ** 2)
    k = k / k.sum()
    out = vol
    for dim in (1, 2, 3):
        x = out.movedim(dim, -1)
        xp = torch.nn.functional.pad(x, (r, r))
        acc = torch.zeros_like(x)
        L = x.shape[-1]
        for j in range(2 * r + 1):
            acc.add_(xp[..., j:j + L], alpha=float(k[j]))
        out = acc.movedim(-1, dim)
    return out.contiguous()

def standardize(Y):
    """Per-element zero mean, unit variance over subjects (torch, dim 0)."""
    m = Y.mean(0, keepdim=True)
    s = Y.std(0, keepdim=True).clamp_min(1e-12)
    return (Y - m) / s


def _gen(seed, device):
    torch = _torch()
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def design(rng, n, n_cov, binary=True):
    cols = [np.ones(n)] + [rng.standard_normal(n) for _ in range(n_cov)]
    if binary:
        g = (rng.random(n) < 0.5).astype(np.float64)
        if g.sum() in (0, n):
            g[0] = 1.0 - g[0]
        cols.append(g)
    return np.column_stack(cols).astype(np.float32).astype(np.float64)


def config1(seed=20261203, n=100, order=5, n_perm=100, device="cpu"):
    """Vertex GLM: order-5 icosphere (10,242 V), n=100, X=[1,c1,c2,g]."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    row_ptr, col = icosphere_csr(order)
    V = row_ptr.shape[0] - 1
    X = design(rng, n, 2)
    noise = torch.randn((n, V), generator=_gen(seed, device), device=device)
    S = smoothing_operator(row_ptr, col, device)
    Y = standardize(mesh_smooth(noise, S, 5))
    patch = bfs_ball(row_ptr, col, int(rng.integers(V)), min(300, V))
    g = torch.tensor(X[:, -1], dtype=torch.float32, device=device)
    Y[:, torch.from_numpy(patch).to(device)] += 0.5 * g[:, None]
    return dict(name="c1_vertex_10k", surfaces=[(row_ptr, col)], offsets=[0], V=V, n=n, X=X,
                Y=Y.float().contiguous(), contrast=X.shape[1] - 1, H=2.0, E=1.0, n_perm=n_perm,
                seed=seed, implicit_zero=[False])


def config2(seed=20261204, n=200, shape=(91, 109, 91), center=(45, 54, 45), semi=(36, 44, 35),
            n_perm=1000, device="cpu"):
    """Voxel GLM: ellipsoid mask in a 91x109x91 grid, 26-connectivity, n=200, k=5."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    mask = ellipsoid_mask(shape, center, semi)
    row_ptr, col = voxel_csr(mask, 26)
    V = int(mask.sum())
    X = design(rng, n, 3)
    noise = torch.randn((n,) + tuple(shape), generator=_gen(seed, device), device=device)
    vol = gaussian_smooth3d(noise, 3.0)
    m = torch.from_numpy(mask).to(device)
    Y = standardize(vol[:, m])
    g = torch.tensor(X[:, -1], dtype=torch.float32, device=device)
    idx = np.nonzero(mask)
    coords = np.stack(idx, 1)
    for _ in range(2):
        c = coords[rng.integers(V)]
        inball = np.nonzero(((coords - c) ** 2).sum(1) <= 16)[0]
        Y[:, torch.from_numpy(inball).to(device)] += 0.4 * g[:, None]
    return dict(name="c2_voxel_%dk" % (V // 1000), surfaces=[(row_ptr, col)], offsets=[0], V=V, n=n,
                X=X, Y=Y.float().contiguous(), contrast=X.shape[1] - 1, H=2.0, E=0.5, n_perm=n_perm,
                seed=seed, mask=mask, implicit_zero=[False])


def config3(seed=20261205, n=500, order=7, n_perm=5000, batch=256, device="cpu"):
    """Two-hemisphere GLM: 2 x order-7 icosphere (327,684 V), n=500, k=6."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    row_ptr, col = icosphere_csr(order)
    Vh = row_ptr.shape[0] - 1
    X = design(rng, n, 4)
    g = torch.tensor(X[:, -1], dtype=torch.float32, device=device)
    S = smoothing_operator(row_ptr, col, device)
    halves = []
    gen = _gen(seed, device)
    for h in range(2):
        noise = torch.randn((n, Vh), generator=gen, device=device)
        halves.append(standardize(mesh_smooth(noise, S, 5)))
    Y = torch.cat(halves, 1)
    for _ in range(5):
        h = int(rng.integers(2))
        size = int(rng.integers(100, 2001))
        d = float(rng.uniform(0.2, 0.6))
        patch = bfs_ball(row_ptr, col, int(rng.integers(Vh)), size) + h * Vh
        Y[:, torch.from_numpy(patch).to(device)] += d * g[:, None]
    return dict(name="c3_2hemi_327k", surfaces=[(row_ptr, col), (row_ptr, col)], offsets=[0, Vh],
                V=2 * Vh, n=n, X=X, Y=Y.float().contiguous(), contrast=X.shape[1] - 1, H=2.0, E=1.0,
                n_perm=n_perm, batch=batch, seed=seed, implicit_zero=[False, False])


def config4(seed=20261206, n=350, n_mesh=40, n_vox=10, mesh_orders=(4, 5, 6, 7),
            vox_range=(20000, 230000), n_perm=2000, device="cpu"):
    """Uneven multimodal set: meshes of orders 4-7 and ellipsoid voxel masks."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    surfaces, settings, offsets, kinds = [], [], [], []
    cache = {}
    off = 0
    orders = rng.choice(mesh_orders, size=n_mesh)
    for o in orders:
        if o not in cache:
            cache[o] = icosphere_csr(int(o))
        surfaces.append(cache[o])
        settings.append((2.0, 1.0))
        kinds.append("mesh")
        offsets.append(off)
        off += cache[o][0].shape[0] - 1
    for _ in range(n_vox):
        target = rng.uniform(*vox_range)
        # ellipsoid volume 4/3 pi abc with axes ratio 1 : 1.2 : 0.95
        a = (target / (4.0 / 3.0 * np.pi * 1.2 * 0.95)) ** (1.0 / 3.0)
        semi = (a, 1.2 * a, 0.95 * a)
        shape = tuple(int(2 * s + 5) for s in semi)
        center = tuple((sh - 1) / 2.0 for sh in shape)
        mask = ellipsoid_mask(shape, center, semi)
        surfaces.append(voxel_csr(mask, 26))
        settings.append((2.0, 0.5))
        kinds.append("voxel")
        offsets.append(off)
        off += int(mask.sum())
    V = off
    X = design(rng, n, 2)
    Y = torch.randn((n, V), generator=_gen(seed, device), device=device)
    return dict(name="c4_multimodal", surfaces=surfaces, offsets=offsets, settings=settings, kinds=kinds,
                V=V, n=n, X=X, Y=Y.contiguous(), contrast=X.shape[1] - 1, n_perm=n_perm, seed=seed,
                implicit_zero=[False] * len(surfaces))


def config5(seed=20261207, n=300, order=7, n_perm=10000, device="cpu"):
    """Mediation: X -> M (image, a-path 0.3 on a 1,500-vertex patch) -> Y (scalar)."""
    torch = _torch()
    rng = np.random.default_rng(seed)
    row_ptr, col = icosphere_csr(order)
    Vh = row_ptr.shape[0] - 1
    x = rng.standard_normal(n)
    covars = rng.standard_normal((n, 3))
    S = smoothing_operator(row_ptr, col, device)
    gen = _gen(seed, device)
    halves = [standardize(mesh_smooth(torch.randn((n, Vh), generator=gen, device=device), S, 5))
              for _ in range(2)]
    M = torch.cat(halves, 1)
    patch = bfs_ball(row_ptr, col, int(rng.integers(Vh)), 1500)
    xt = torch.tensor(x, dtype=torch.float32, device=device)
    M[:, torch.from_numpy(patch).to(device)] += 0.3 * xt[:, None]
    y = 0.25 * M[:, torch.from_numpy(patch).to(device)].mean(1).double().cpu().numpy() \
        + rng.standard_normal(n)
    return dict(name="c5_mediation_327k", surfaces=[(row_ptr, col), (row_ptr, col)], offsets=[0, Vh],
                V=2 * Vh, n=n, x=x, y=y, covars=covars, M=M.float().contiguous(), H=2.0, E=1.0,
                n_perm=n_perm, seed=seed, implicit_zero=[False, False])


def permutations(n, P, seed, device="cpu"):
    """P random permutations of range(n) as int32 [P, n] (torch)."""
    torch = _torch()
    g = _gen(seed, device)
    return torch.argsort(torch.rand((P, n), generator=g, device=device), dim=1).to(torch.int32)


def reference_permutation(perm_number, perm_seed, n):
    """The permutation mmr-lr draws for one permutation number
    (np.random.seed(perm_number + perm_seed), tm_func.py:147-149)."""
    rs = np.random.RandomState(int(perm_number) + int(perm_seed))
    return rs.permutation(list(range(n)))
