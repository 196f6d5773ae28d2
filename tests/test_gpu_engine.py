"""The batched engine's element layout: holding the surfaces in the library's
locality order must not change a single null value (max-TFCE is invariant
under relabelling a symmetric graph, and each element's t-value is computed
by the same arithmetic wherever its column sits)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_locality_layout_leaves_the_null_unchanged():
    import torch
    from tfce_mediation_amd import synth
    from tfce_mediation_amd.engine import PermutationEngine
    print("[gpu] engine locality layout", flush=True)
    for make in (lambda: synth.config1(seed=5, n=80, order=5, device="cuda"),
                 lambda: synth.config2(seed=6, n=60, shape=(30, 34, 30), center=(14.5, 16.5, 14.5),
                                       semi=(12, 14, 12), device="cuda")):
        cfg = make()
        perms = synth.permutations(cfg["n"], 64, 11, "cuda")
        out = []
        for locality in (False, True):
            eng = PermutationEngine(cfg["Y"], cfg["X"], cfg["surfaces"], cfg["offsets"], cfg["H"], cfg["E"],
                                    contrast=cfg["contrast"], precision="bf16x3", locality=locality)
            out.append(eng.null_batch(perms).cpu().numpy())
            if locality:
                assert eng.cols is not None
                t = eng.tvalues(perms[:4]).cpu().numpy()
        np.testing.assert_array_equal(out[0], out[1])
        # the t-values are the same numbers in a permuted column order
        eng0 = PermutationEngine(cfg["Y"], cfg["X"], cfg["surfaces"], cfg["offsets"], cfg["H"], cfg["E"],
                                 contrast=cfg["contrast"], precision="bf16x3", locality=False)
        t0 = eng0.tvalues(perms[:4]).cpu().numpy()
        np.testing.assert_array_equal(t0[:, eng.cols], t)
        torch.cuda.synchronize()


def test_graph_replay_matches_eager():
    """The captured batch replayed several times, with eager batches and
    synchronisations between the replays, gives the eager null every time."""
    import torch
    from tfce_mediation_amd import synth
    from tfce_mediation_amd.engine import PermutationEngine
    print("[gpu] engine hipGraph replay", flush=True)
    cfg = synth.config1(seed=9, n=64, order=5, device="cuda")
    eng = PermutationEngine(cfg["Y"], cfg["X"], cfg["surfaces"], cfg["offsets"], cfg["H"], cfg["E"],
                            contrast=cfg["contrast"], precision="bf16x3")
    P = 48
    run = eng.graph(P)
    for rep in range(4):
        perms = synth.permutations(cfg["n"], P, 100 + rep, "cuda")
        got = run(perms).clone()
        torch.cuda.synchronize()
        ref = eng.null_batch(perms)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy())


def test_pipeline_matches_batch_by_batch():
    """Contraction of batch b+1 on a second stream during the sweep of batch b:
    every batch's null equals null_batch on it, including a short last batch,
    for the t engine (+/-t) and the mediation engine (+z)."""
    import torch
    from tfce_mediation_amd import synth
    from tfce_mediation_amd.engine import MediationEngine, PermutationEngine
    print("[gpu] engine pipeline", flush=True)
    cfg = synth.config1(seed=12, n=72, order=5, device="cuda")
    eng = PermutationEngine(cfg["Y"], cfg["X"], cfg["surfaces"], cfg["offsets"], cfg["H"], cfg["E"],
                            contrast=cfg["contrast"], precision="bf16x3")
    x = cfg["X"][:, -1].cpu().numpy() if isinstance(cfg["X"], torch.Tensor) else np.asarray(cfg["X"])[:, -1]
    rng = np.random.default_rng(4)
    y = 0.3 * x + rng.standard_normal(len(x))
    med = MediationEngine(cfg["Y"], x, y, cfg["surfaces"], cfg["offsets"], cfg["H"], cfg["E"],
                          precision="bf16x3")
    for e in (eng, med):
        batches = [synth.permutations(cfg["n"], P, 40 + i, "cuda") for i, P in enumerate((64, 64, 64, 37))]
        timing = {}
        outs = e.pipeline(batches, timing=timing)
        torch.cuda.synchronize()
        assert len(timing["sweep"]) == len(timing["contraction"]) == 4
        for pb, got in zip(batches, outs):
            ref = e.null_batch(pb)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy())
        for pipelined in (False, True):
            full = e.run(sum(int(b.shape[0]) for b in batches), batch=64, perms=torch.cat(batches).contiguous(),
                         pipelined=pipelined)
            np.testing.assert_array_equal(full, np.concatenate([o.cpu().numpy() for o in outs]))


def test_default_batch_fills_the_compute_units():
    """Four TFCE items per compute unit, in multiples of 64 permutations
    between 256 and 1024."""
    import torch
    from tfce_mediation_amd.tfce import default_batch
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    print(f"[gpu] default_batch with {cus} compute units", flush=True)
    for items in (1, 2, 3, 4, 50):
        b = default_batch(items)
        assert b % 64 == 0 and 256 <= b <= 1024
        assert items * b >= min(4 * cus, items * 1024)
        assert b == 256 or items * (b - 64) < 4 * cus


def test_graph_replay_after_scratch_grows_on_capture_stream():
    """A graph holds the scratch of the stream it was captured on.  A larger
    eager batch on that same stream grows the scratch afterwards; the
    library keeps the captured buffers alive, so later replays still give
    the eager null."""
    import torch
    from tfce_mediation_amd import synth
    from tfce_mediation_amd.engine import PermutationEngine
    print("[gpu] engine hipGraph replay after the capture stream's scratch grows", flush=True)
    cfg = synth.config1(seed=19, n=60, order=4, device="cuda")
    eng = PermutationEngine(cfg["Y"], cfg["X"], cfg["surfaces"], cfg["offsets"], cfg["H"], cfg["E"],
                            contrast=cfg["contrast"])
    P = 40
    run = eng.graph(P)
    big = synth.config1(seed=20, n=60, order=6, device="cuda")
    eng_big = PermutationEngine(big["Y"], big["X"], big["surfaces"], big["offsets"], big["H"], big["E"],
                                contrast=big["contrast"])
    with torch.cuda.stream(run.stream):
        grown = eng_big.null_batch(synth.permutations(big["n"], 300, 3, "cuda"))
    run.stream.synchronize()
    assert grown.shape == (300, 2)
    for rep in range(3):
        perms = synth.permutations(cfg["n"], P, 200 + rep, "cuda")
        got = run(perms).clone()
        torch.cuda.synchronize()
        ref = eng.null_batch(perms)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy(), ref.cpu().numpy())
