#!/usr/bin/env python
"""Permutation throughput of the TFCE_mediation hot path on one MI355X.

Workload (default, c3): two order-7 icospheres (2 x 163,842 = 327,684
vertices), n = 500 subjects, design [1, 4 N(0,1) covariates, binary group]
(k = 6), contrast on the group column, t two-tailed (+t and -t), TFCE H=2 E=1
with the reference step max/100, max over hemispheres per permutation.
Synthetic smoothed data with five effect patches, generated on the GPU.
`--config c1|c2|c4|c5` measures the other workloads of the project's table the
same way (c5: mediation, +z only).

A step is one batch of P permutations through the whole path (P = 256 at
c3; by default the engine's batch size, which gives the TFCE launch four
items per compute unit: 512 for c1, c2 and c5): the permuted-design
contraction on MFMA (split precision, 3 products: FP16x3 by default, the
split scheme of the precision target that meets the 1e-5 t bar; BF16x3 and
plain FP32 are timed beside it) with the fused t epilogue, then one
persistent TFCE launch over every (surface, permutation, sign) item, then
the max over surfaces.  Inputs are resident in HBM before the timed region.

Printed: one JSON line with the metric, the plain FP32 per-permutation GPU path
measured the same way, the same steps in the other precisions and replayed
from a captured hipGraph, the autotuner's launch shapes, the statistic's error
against FP64, a roofline entry for the dominant kernel (TFCE sweep: touch-once
algorithmic bytes over the HIP-event time of the sweep inside the timed
region, beside the counter-measured traffic), and the CPU baseline (reference
TFCE build + the reference's numpy statistic code path) timed on this host's
cores over a bounded sample, with its agreement with the GPU null.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
N_THR = 100

WORKLOADS = {
    "c1": ("c1: order-5 icosphere (10,242 vertices), n=100, k=4, contrast=group, TFCE H=2 E=1 dh=max/100, "
           "+/- t", "synthetic (smoothed Gaussian, one 300-vertex effect patch; random permutations)"),
    "c2": ("c2: 91x109x91 grid, ellipsoid mask, 26-connectivity, n=200, k=5, contrast=last, TFCE H=2 E=0.5, "
           "+/- t", "synthetic (Gaussian noise, FWHM 3 voxels, two r=4 effect spheres; random permutations)"),
    "c3": ("c3: 2 x order-7 icosphere (327,684 vertices), n=500, k=6, contrast=group, TFCE H=2 E=1 "
           "dh=max/100, +/- t, max over hemispheres",
           "synthetic (smoothed Gaussian on order-7 icospheres, 5 effect patches; random permutations)"),
    "c4": ("c4: 40 icospheres of orders 4-7 + 10 ellipsoid voxel masks, n=350, k=4, mesh H=2 E=1, voxel "
           "H=2 E=0.5, +/- t, max over surfaces", "synthetic (N(0,1) data; random permutations)"),
    "c5": ("c5: mediation on 2 x order-7 icosphere (327,684 vertices), n=300, type M Aroian z, 3 covariates "
           "regressed out, TFCE H=2 E=1, +z only, max over hemispheres",
           "synthetic (smoothed Gaussian mediator with a 1,500-vertex a-path patch; random permutations)"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ----------------------------------------------------------------------------
# CPU baseline worker (spawned process, single-threaded BLAS)
# ----------------------------------------------------------------------------
def _cpu_worker(args):
    ypath, xpath, ppath, csr_path, perm_ids, settings, offsets, contrast, med = args
    import numpy as np
    from oracle import ols as ools
    from oracle import tfce as otfce
    Y = np.load(ypath, mmap_mode="r")
    X = np.load(xpath)
    perms = np.load(ppath)
    z = np.load(csr_path)
    csrs = [(z["rp%d" % i], z["col%d" % i]) for i in range(len(settings))]
    use_ref = otfce.have_reference()
    n = X.shape[0]
    V = Y.shape[1]

    def beta_se(Xd):
        # calc_beta_se as the reference runs it: float64 products on the
        # float32 data and the per-element se_of_slope loop (cynumstats.pyx:47-74)
        k = Xd.shape[1]
        invXX = np.linalg.inv(np.dot(Xd.T, Xd))
        a = ools.cy_lin_lstsqr_mat(Xd, Y)
        sigma2 = np.sum((Y - np.dot(Xd, a)) ** 2, axis=0) / (n - k)
        se = np.zeros((k, V), dtype=np.float32)
        for j in range(V):
            se[:, j] = np.sqrt(np.diag(sigma2[j] * invXX))
        return a, se

    out = []
    t0 = time.time()
    for p in perm_ids:
        if med:
            # vertex_tfce_mediation_randomise.py:82-91, type M, Aroian:
            # x permuted, y fixed; path A y_img ~ [1, x], path B y_img ~ [1, y, x]
            x, yv = X[:, 0], X[:, 1]
            xp = x[perms[p]]
            a_a, se_a = beta_se(np.column_stack([np.ones(n), xp]))
            a_b, se_b = beta_se(np.column_stack([np.ones(n), yv, xp]))
            ta, tb = a_a[1] / se_a[1], a_b[1] / se_b[1]
            stat = 1 / np.sqrt((1 / (tb ** 2)) + (1 / (ta ** 2)) + (1 / (ta ** 2 * tb ** 2)))
            signs = (1.0,)
        else:
            a, se = beta_se(X[perms[p]])
            stat = (a / se)[contrast]
            signs = (1.0, -1.0)
        vals = []
        for sgn in signs:
            best = None
            for (rp, col), (H, E), off in zip(csrs, settings, offsets):
                Vs = len(rp) - 1
                img = np.ascontiguousarray(sgn * stat[off:off + Vs], dtype=np.float32)
                if use_ref:
                    enh = otfce.ref_tfce(rp, col, img, H, E)
                else:
                    enh, _, _ = otfce.tfce(rp, col, img, H, E)
                m = enh.max() * (img.max() / np.float32(100))
                best = m if best is None else max(best, m)
            vals.append(float(best))
        out.append(vals)
    return time.time() - t0, out, use_ref


def host_cores():
    """(os.cpu_count(), CPUs this process may run on, CPUs of the cgroup quota or None)."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = nproc
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    return nproc, aff, quota


def cpu_baseline(cfg, perms_np, n_workers, per_worker, settings, med=False):
    import numpy as np
    tmp = tempfile.mkdtemp(prefix="tfce_cpu_")
    ypath = os.path.join(tmp, "Y.npy")
    np.save(ypath, cfg["Y"].cpu().numpy())
    xpath = os.path.join(tmp, "X.npy")
    # mediation: the columns are x and the outcome y
    np.save(xpath, np.column_stack([cfg["x"], cfg["y"]]) if med else cfg["X"])
    ppath = os.path.join(tmp, "perms.npy")
    np.save(ppath, perms_np)
    csr_path = os.path.join(tmp, "csr.npz")
    arrays = {}
    for i, (rp, col) in enumerate(cfg["surfaces"]):
        arrays["rp%d" % i], arrays["col%d" % i] = rp, col
    np.savez(csr_path, **arrays)
    jobs = [(ypath, xpath, ppath, csr_path, list(range(w * per_worker, (w + 1) * per_worker)), settings,
             list(cfg["offsets"]), cfg.get("contrast", 0), med) for w in range(n_workers)]
    env_keep = {k: os.environ.get(k) for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in env_keep:
        os.environ[k] = "1"
    try:
        ctx = mp.get_context("spawn")
        t0 = time.time()
        with ctx.Pool(n_workers) as pool:
            res = pool.map(_cpu_worker, jobs)
        wall = time.time() - t0
    finally:
        for k, v in env_keep.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        for f in (ypath, xpath, ppath, csr_path):
            try:
                os.remove(f)
            except OSError:
                pass
        try:
            os.rmdir(tmp)
        except OSError:
            pass
    n_done = n_workers * per_worker
    worker_times = [r[0] for r in res]
    values = [v for r in res for v in r[1]]
    return dict(perms=n_done, wall=wall, per_core_rate=per_worker / (sum(worker_times) / len(worker_times)),
                aggregate_rate=n_done / max(worker_times), values=values, used_reference=res[0][2])


def _measured_traffic(workload, images):
    """HBM bytes per launch of the workload's sweep kernel over `images`
    images, from the committed rocprofv3 counter summary (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated, measured over the images its line names and
    scaled per image): (bytes, summary path, kernel name), or (None, None,
    None)."""
    path = os.path.join(ROOT, "profiles", "%s_tfce_hbm_traffic.md" % workload)
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 5 and parts[0] == "traffic_bytes_per_launch" and parts[2] == workload:
                    return int(parts[3]) * images // int(parts[4]), os.path.relpath(path, ROOT), parts[1]
    except OSError:
        pass
    return None, None, None


def make_workload(name, precision):
    """(cfg, engine, sign_mask, per-surface (H, E), credited contraction flops per permutation)."""
    import numpy as np
    import torch
    from tfce_mediation_amd import synth
    from tfce_mediation_amd.engine import MediationEngine, PermutationEngine
    cfg = getattr(synth, "config%d" % int(name[1]))(device="cuda")
    n_surf = len(cfg["surfaces"])
    settings = cfg.get("settings") or [(cfg["H"], cfg["E"])] * n_surf
    if name == "c5":
        C = torch.as_tensor(np.column_stack([np.ones(cfg["n"]), cfg["covars"]]), dtype=torch.float64).cuda()
        M = cfg["M"].double()
        M = (M - C @ torch.linalg.lstsq(C, M).solution).float().contiguous()
        cfg["Y"] = M
        eng = MediationEngine(M, cfg["x"], cfg["y"], cfg["surfaces"], cfg["offsets"], cfg["H"], cfg["E"],
                              precision=precision)
        return cfg, eng, 1, settings, 2.0 * (2 + 3) * cfg["n"] * cfg["V"]
    eng = PermutationEngine(cfg["Y"], cfg["X"], cfg["surfaces"], cfg["offsets"], cfg.get("H", 2.0),
                            cfg.get("E", 1.0), contrast=cfg["contrast"], precision=precision, settings=settings)
    return cfg, eng, 3, settings, 2.0 * cfg["X"].shape[1] * cfg["n"] * cfg["V"]


def plain_engine(name, cfg, settings, precision="fp32"):
    from tfce_mediation_amd.engine import MediationEngine, PermutationEngine
    if name == "c5":
        return MediationEngine(cfg["Y"], cfg["x"], cfg["y"], cfg["surfaces"], cfg["offsets"], cfg["H"], cfg["E"],
                               precision=precision)
    return PermutationEngine(cfg["Y"], cfg["X"], cfg["surfaces"], cfg["offsets"], contrast=cfg["contrast"],
                             precision=precision, settings=settings)


def active_counts(x):
    """Active elements per row of the float32 device maps x [P, V] at the last
    level of the reference's sequence (T0 = max, T -= fl(T0/100) while T >= 0,
    active iff x > T): int64 numpy [P]."""
    import numpy as np
    import torch
    T0 = x.max(1).values.cpu().numpy().astype(np.float32)
    dT = (T0 / np.float32(100)).astype(np.float32)
    last = np.full_like(T0, np.inf)
    T = T0.copy()
    live = (T0 >= 0) & (dT > 0)
    while live.any():
        last[live] = T[live]
        T = np.where(live, (T - dT).astype(np.float32), T)
        live &= T >= 0
    return (x > torch.as_tensor(last, device=x.device)[:, None]).sum(1).cpu().numpy()


def fp64_sobel(cfg, perm):
    """Aroian z of the mediation workload (type M: x permuted, y fixed) from the
    FP64 single-design path, in map columns."""
    import numpy as np
    from tfce_mediation_amd.cynumstats import _direct
    n = cfg["n"]
    xp = np.asarray(cfg["x"], dtype=np.float64)[perm]
    one = np.ones(n)
    ta = _direct(np.column_stack([one, xp]), cfg["Y"], t=True)["t"][1]
    tb = _direct(np.column_stack([one, np.asarray(cfg["y"], dtype=np.float64), xp]), cfg["Y"], t=True)["t"][1]
    return 1 / np.sqrt((1 / (tb ** 2)) + (1 / (ta ** 2)) + (1 / (ta ** 2 * tb ** 2)))


# ----------------------------------------------------------------------------
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=None,
                    help="permutations per step (default: the engine's default_batch())")
    ap.add_argument("--precision", default="fp16x3",
                    help="contraction precision of the headline (fp16x3: hi/lo split-FP16, 3 products; "
                         "bf16x3: hi/lo split-BF16, 3 products; fp32)")
    ap.add_argument("--plain-perms", type=int, default=16)
    ap.add_argument("--cpu-workers", type=int, default=None,
                    help="CPU baseline processes (default: the CPUs this process may use, at most 16, "
                         "the host share of one GPU)")
    ap.add_argument("--cpu-perms-per-worker", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default contraction tile and sweep occupancy")
    ap.add_argument("--tile", type=int, default=None,
                    help="contraction tile (0: 256x128, 1: 512x64, 2: 512x128, 3: 256x256) instead of autotuning")
    ap.add_argument("--wgs-per-cu", type=int, default=None,
                    help="sweep workgroups per CU instead of autotuning")
    ap.add_argument("--lds-sweep", type=int, default=None, choices=(0, 1),
                    help="surfaces of at most 13,248 elements in the LDS-resident sweep (1) or the global one (0) "
                         "instead of autotuning")
    ap.add_argument("--pair-sweep", type=int, default=None, choices=(0, 1),
                    help="both signs of a map in one workgroup over ring-packed surfaces (1) or one image per "
                         "workgroup (0) instead of autotuning")
    ap.add_argument("--pipeline", action="store_true",
                    help="run the contraction of step s+1 on a second stream during the sweep of step s "
                         "(engine.pipeline) instead of one after the other on one stream")
    ap.add_argument("--no-precision-check", action="store_true",
                    help="skip the t-vs-FP64 error check and the FP16x3 timing")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the hipGraph replay measurement of the same steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="copy a buffer of this many MiB once before the timed region, so that a "
                         "counter run (FETCH_SIZE / WRITE_SIZE) holds a plain copy of known size")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding of permutation ranges is design-only here", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from tfce_mediation_amd import synth
    from tfce_mediation_amd.tfce import Adjacency

    torch.cuda.set_device(0)
    t_setup = time.time()
    name = args.config
    cfg, eng, sign_mask, settings, flops_per_perm = make_workload(name, args.precision)
    n_signs = 2 if sign_mask == 3 else 1
    torch.cuda.synchronize()
    log(f"[bench] {name}: data and engine ready in {time.time() - t_setup:.1f}s: V={cfg['V']} n={cfg['n']} "
        f"surfaces={len(cfg['surfaces'])}")
    P = args.batch if args.batch is not None else eng.default_batch()
    n = cfg["n"]
    perm_batches = [synth.permutations(n, P, 1000 + s, "cuda") for s in range(args.warmup + args.steps)]

    # ---- launch shapes picked on this workload (not cached: measured here) ----
    tuned = None
    if (args.tile is not None or args.wgs_per_cu is not None or args.lds_sweep is not None
            or args.pair_sweep is not None):
        from tfce_mediation_amd import autotune
        tuned = {"gemm_tile": 0 if args.tile is None else args.tile,
                 "tfce_wgs_per_cu": 4 if args.wgs_per_cu is None else args.wgs_per_cu,
                 "lds_sweep": 1 if args.lds_sweep is None else args.lds_sweep,
                 "pair_sweep": 1 if args.pair_sweep is None else args.pair_sweep}
        autotune.apply(tuned)
    elif not args.no_autotune:
        from tfce_mediation_amd import autotune
        tuned = autotune.tune(eng, P, cache=None, log=log)

    # ---- warmup -------------------------------------------------------------
    for s in range(args.warmup):
        eng.null_batch(perm_batches[s])
    torch.cuda.synchronize()
    log("[bench] warmup done")

    if args.copy_calibration:
        src = torch.ones(args.copy_calibration << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy: {args.copy_calibration} MiB read, {args.copy_calibration} MiB written")
        del src, dst

    # ---- timed region -------------------------------------------------------
    # each step's contraction then its sweep on one stream (--pipeline: the
    # contraction of step s+1 on a second stream during the sweep of step s)
    stream = torch.cuda.current_stream()
    timing = {}
    timed = perm_batches[args.warmup:]
    if args.pipeline:
        eng.pipeline(perm_batches[:args.warmup])  # the second stream's scratch grows here
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if not args.pipeline:
        for perms in timed:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e2 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            t = eng.tvalues(perms)
            e1.record(stream)
            per = eng.surfaces.batch(t, sign_mask, vdensity=eng.vdensity)
            e2.record(stream)
            per.max(0).values
            timing.setdefault("contraction", []).append((e0, e1))
            timing.setdefault("sweep", []).append((e1, e2))
    else:
        eng.pipeline(timed, timing=timing)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    gemm_ms = float(np.mean([a.elapsed_time(b) for a, b in timing["contraction"]]))
    tfce_ms = float(np.mean([a.elapsed_time(b) for a, b in timing["sweep"]]))
    value = args.steps * P / elapsed
    mode = "pipelined" if args.pipeline else "sequential"
    log(f"[bench] batched {args.precision} ({mode}): {value:.1f} perms/s  ({elapsed / args.steps * 1e3:.2f} ms/step; "
        f"contraction {gemm_ms:.2f} ms, TFCE {tfce_ms:.2f} ms)")

    # ---- the same steps replayed from a captured hipGraph ---------------------
    graph_rate = None
    if not args.no_graph:
        run = eng.graph(P)
        same = np.array_equal(run(perm_batches[args.warmup]).cpu().numpy(),
                              eng.null_batch(perm_batches[args.warmup]).cpu().numpy())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(args.steps):
            run(perm_batches[args.warmup + s])
        torch.cuda.synchronize()
        graph_rate = args.steps * P / (time.perf_counter() - t0)
        log(f"[bench] hipGraph replay: {graph_rate:.1f} perms/s (replay output equals eager: {same})")
        if not same:
            graph_rate = None
        del run

    # ---- t against FP64; the other split scheme and batched FP32 -------------
    # t of 4 permutations of this workload (engine layout) against the FP64
    # single-design path (tval_int semantics on X[perm], cynumstats) mapped
    # back to the engine's columns (c5: the Sobel z against the FP64 path
    # model); then the same timed steps with the contraction in the other
    # split scheme and in plain FP32 (same engine, same sweep).
    t_err = {}
    alt_rate = {}
    if not args.no_precision_check:
        from tfce_mediation_amd.cynumstats import _direct
        perms_chk = perm_batches[args.warmup][:4]
        cols = eng.cols if eng.cols is not None else np.arange(eng.V)
        if name == "c5":
            ref = np.stack([fp64_sobel(cfg, p)[cols] for p in perms_chk.cpu().numpy()])
        else:
            X = np.asarray(cfg["X"], dtype=np.float64)
            ref = np.stack([_direct(X[p], cfg["Y"], t=True)["t"][eng.contrast][cols]
                            for p in perms_chk.cpu().numpy()])

        def terr(tg):
            tg = tg.astype(np.float64)
            big = np.abs(ref) >= 1
            return {"norm": float(np.linalg.norm(tg - ref) / np.linalg.norm(ref)),
                    "point_abs_ge_1": float(np.max(np.abs(tg - ref)[big] / np.abs(ref[big])))}
        t_err[args.precision] = terr(eng.tvalues(perms_chk).cpu().numpy())
        for prec in ("fp16x3", "bf16x3", "fp32"):
            if prec == args.precision:
                continue
            _, eng2, _, _, _ = make_workload(name, prec)
            if prec != "fp32":
                t_err[prec] = terr(eng2.tvalues(perms_chk).cpu().numpy())
            for s in range(args.warmup):
                eng2.null_batch(perm_batches[s])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for perms in timed:
                eng2.null_batch(perms)
            torch.cuda.synchronize()
            alt_rate[prec] = args.steps * P / (time.perf_counter() - t0)
            del eng2
        log(f"[bench] error vs FP64 (4 permutations): {t_err}; batched steps in other precisions: "
            + ", ".join(f"{k} {v:.1f} perms/s" for k, v in alt_rate.items()))

    # ---- plain FP32 per-permutation path ------------------------------------
    # the reference's call structure: one t map per permutation, then one
    # CreateAdjSet.run per (sign, surface) and the write_perm_maxTFCE rescale
    plain = None
    if not args.no_plain:
        eng32 = plain_engine(name, cfg, settings)
        adjs = [Adjacency(csr, h, e) for csr, (h, e) in zip(eng32.csrs, settings)]
        offs = [int(o) for o in cfg["offsets"]]
        enh = torch.empty(max(a.V for a in adjs), dtype=torch.float32, device="cuda")
        signs = (1.0, -1.0) if n_signs == 2 else (1.0,)

        def plain_perm(perm):
            t = eng32.tvalues(perm[None])[0]
            vals = []
            for sgn in signs:
                best = None
                for adj, off in zip(adjs, offs):
                    img = (sgn * t[off:off + adj.V]).contiguous()
                    out_s = enh[:adj.V]
                    adj.run_device(img, out_s, accumulate=False)
                    m = out_s.max() * (img.max() / 100.0)
                    best = m if best is None else torch.maximum(best, m)
                vals.append(best)
            return torch.stack(vals)

        pp = perm_batches[-1][: args.plain_perms]
        plain_perm(pp[0])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for i in range(args.plain_perms):
            plain_perm(pp[i])
        torch.cuda.synchronize()
        plain = args.plain_perms / (time.perf_counter() - t1)
        log(f"[bench] plain fp32 per-permutation path: {plain:.2f} perms/s")
        del eng32

    # ---- CPU baseline ---------------------------------------------------------
    nproc, n_aff, quota = host_cores()
    share = min(n_aff, 16) if quota is None else max(1, min(n_aff, int(quota), 16))
    nw = args.cpu_workers if args.cpu_workers is not None else share
    if name == "c4":
        nw = min(nw, 4)  # each process holds FP64 temporaries of the 3.5 M-element data (about 10 GB)
    cpu = None
    if not args.no_cpu:
        perms_np = perm_batches[args.warmup].cpu().numpy()
        log(f"[bench] CPU baseline: {nw} single-threaded workers x {args.cpu_perms_per_worker} permutation(s) "
            f"(os.cpu_count() {nproc}, affinity {n_aff}, cgroup quota {quota})")
        cpu = cpu_baseline(cfg, perms_np, nw, args.cpu_perms_per_worker, settings, med=(name == "c5"))
        log(f"[bench] CPU: aggregate {cpu['aggregate_rate']:.3f} perms/s, per core {cpu['per_core_rate']:.3f}")
        # the same permutations through the GPU path: agreement of the null values
        gpu_vals = eng.null_batch(perm_batches[args.warmup][: cpu["perms"]]).cpu().numpy()
        cpu_vals = np.array(cpu["values"])
        rel = np.abs(gpu_vals - cpu_vals) / np.maximum(np.abs(cpu_vals), 1e-12)
        cpu["max_rel_diff_null"] = float(rel.max())
        cpu["median_rel_diff_null"] = float(np.median(rel))
        log(f"[bench] GPU vs CPU null values: max rel diff {rel.max():.2e}, median {np.median(rel):.2e}")

    # ---- roofline of the dominant kernel (TFCE sweep) -----------------------
    # Touch-once algorithmic bytes of the incremental sweep, per image (one
    # surface, one permutation, one sign) with V elements, A of them active at
    # the last level and d = nnz / V neighbours on average:
    #   read the map 4V; write and read each element's activation position 8V;
    #   read each active element's neighbour list 4dA and gather its
    #   neighbours' positions 4dA; write and read one union-find node
    #   {parent, size, value} per active element 24A.
    #   B = 12 V + (8 d + 24) A
    # A is counted on the maps of the first timed step with the reference's
    # float32 level sequence (the last threshold >= 0).
    maps = eng.tvalues(perm_batches[args.warmup])
    signs = (1.0, -1.0) if n_signs == 2 else (1.0,)
    touch_bytes = 0.0
    level_model_bytes = 0.0
    images = 0
    for (rp, col), off in zip(eng.csrs, cfg["offsets"]):
        V_s, nnz_s = len(rp) - 1, len(col)
        d = nnz_s / V_s
        for sgn in signs:
            x = sgn * maps[:, int(off):int(off) + V_s]
            A = active_counts(x)
            touch_bytes += float((12.0 * V_s + (8.0 * d + 24.0) * A).sum())
            level_model_bytes += float(P * (N_THR * (4 * (V_s + 1) + 4 * nnz_s + 8 * V_s) + 8 * V_s))
            images += P
    del maps
    achieved = touch_bytes / (tfce_ms * 1e-3) / 1e9
    traffic, traffic_src, traffic_kernel = _measured_traffic(name, images)
    gemm_flops = flops_per_perm * P
    workload, data = WORKLOADS[name]
    if name == "c3":
        metric = "permutations/s at 327,684 vertices, n=500: batched split-BF16 vs plain FP32"
    else:
        metric = f"permutations/s ({name}): batched split-precision vs plain FP32"
    prec_text = {"fp16x3": "fp16x3 (hi/lo split-FP16 of power-of-two scaled operands, 3 MFMA products, fp32 "
                           "accumulate; fp64 SSE/t epilogue; fp32 TFCE sums, int32 union-find)",
                 "bf16x3": "bf16x3 (hi/lo split-BF16, 3 MFMA products, fp32 accumulate; fp64 SSE/t epilogue; "
                           "fp32 TFCE sums, int32 union-find)"}
    out = {
        "metric": metric,
        "value": round(value, 2),
        "unit": "permutations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec_text.get(args.precision, args.precision),
        "data": data,
        "config": {
            "workload": workload,
            "permutations_per_step": P,
            "precision": args.precision,
            "plain_fp32_perms_per_s": None if plain is None else round(plain, 3),
            "plain_path": "per permutation, the reference's call structure: one FP32 t map, then one "
                          "single-image CreateAdjSet.run per surface and sign (tfce_run: the whole-GPU "
                          "level sweep for surfaces of 32,768 elements and more, one workgroup below)",
            "batched_over_plain": None if plain is None else round(value / plain, 2),
            "batched_fp32_perms_per_s": None if "fp32" not in alt_rate else round(alt_rate["fp32"], 1),
            "batched_over_batched_fp32": None if "fp32" not in alt_rate else round(value / alt_rate["fp32"], 3),
            "bf16x3_perms_per_s": None if "bf16x3" not in alt_rate else round(alt_rate["bf16x3"], 1),
            "fp16x3_perms_per_s": None if "fp16x3" not in alt_rate else round(alt_rate["fp16x3"], 1),
            "hipgraph_replay_perms_per_s": None if graph_rate is None else round(graph_rate, 1),
            "stat_rel_err_vs_fp64": t_err or None,
            "autotuned": None if tuned is None else {"gemm_tile": ["256x128", "512x64", "512x128", "256x256"][tuned["gemm_tile"]],
                                                     "tfce_wgs_per_cu": tuned["tfce_wgs_per_cu"],
                                                     "lds_sweep": tuned.get("lds_sweep", 1),
                                                     "pair_sweep": tuned.get("pair_sweep", 1)},
            "schedule": mode + (" (contraction of step s+1 on a second stream during the sweep of step s)"
                                if args.pipeline else " (contraction then sweep, one stream)"),
            "contraction_ms_per_step": round(gemm_ms, 3),
            "tfce_ms_per_step": round(tfce_ms, 3),
            "contraction_tflops": round(gemm_flops / (gemm_ms * 1e-3) / 1e12, 2),
            "contraction_frac_of_dense_peak": round(gemm_flops / (gemm_ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
        },
        "roofline": {
            "kernel": (traffic_kernel or "tfce_batch_kernel") + " (persistent TFCE sweep)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "traffic_over_algorithmic": None if traffic is None else round(traffic / touch_bytes, 3),
            "traffic_gbs": None if traffic is None else round(traffic / (tfce_ms * 1e-3) / 1e9, 1),
            "traffic_frac_of_peak": None if traffic is None else round(traffic / (tfce_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": int(touch_bytes),
            "images_per_launch": images,
            "byte_model": "touch-once: B = 12 V + (8 d + 24) A per image (V elements, A active at the last "
                          "level, d neighbours per element)",
            "level_sweep_model_frac": round(level_model_bytes / (tfce_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 3),
        },
        "cpu_baseline": None if cpu is None else {
            "value": round(cpu["aggregate_rate"], 4),
            "unit": "permutations/s",
            "cores": nw,
            "host_nproc": nproc,
            "host_affinity_cpus": n_aff,
            "host_cgroup_cpus": quota,
            "kind": "reference" if cpu["used_reference"] else "port",
            "per_core": round(cpu["per_core_rate"], 4),
            "sample": f"{cpu['perms']} permutations of the same workload, {args.cpu_perms_per_worker} per "
                      f"single-threaded process: " + ("the reference's numpy calc_beta_se code path for both "
                      "mediation paths (float64, per-element se_of_slope loop) and the Aroian z"
                      if name == "c5" else "the reference's numpy tval_int code path (float64, per-element "
                      "se_of_slope loop)") + " + its own fast_tfce.hpp (oracle/_ref) for every sign and "
                      "surface + write_perm_maxTFCE scaling; shell echo excluded",
            "max_rel_diff_vs_gpu_null": cpu.get("max_rel_diff_null"),
            "median_rel_diff_vs_gpu_null": cpu.get("median_rel_diff_null"),
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
