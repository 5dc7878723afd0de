"""Batched burst pairs (s1ifg_batch_*): every pair of a batch of uneven
bursts gives the same bits as the single-pair entry points, eagerly and on
every replay of the captured HIP graph, for every launch shape."""

import numpy as np
import pytest

from tests.scenes import scene

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NAMES = ("small", "odd", "small_misreg", "wide")
ALPHA = 0.5
SEED = 11


@pytest.fixture(scope="module")
def ctx():
    from sentinel_gpu_processor_amd import _native as N
    c = N.Context(0)
    yield c
    c.sync()
    c.close()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _bits(t):
    torch.cuda.synchronize()
    a = t.cpu().numpy()
    return a.view(np.uint32) if a.dtype != np.complex64 else a.view(np.uint64)


@pytest.fixture(scope="module")
def pairs(ctx):
    from sentinel_gpu_processor_amd import _native as N
    out = []
    for name in NAMES:
        sc = scene(name)
        out.append(dict(
            g_ref=N.Geometry(ctx, sc["orbit_ref"], sc["burst_ref"]),
            g_sec=N.Geometry(ctx, sc["orbit_sec"], sc["burst_sec"]),
            grid=sc["grid"], dem_t=_dev(sc["dem"]["data"]), dem=sc["dem"],
            ref_slc=_dev(sc["ref_slc"]), sec_slc=_dev(sc["sec_slc"]), offpoly=sc["offpoly"]))
    return out


def _single(ctx, p, interp, gold_mode, win, rl=1, cl=4):
    from sentinel_gpu_processor_amd import kernels as K
    dr = K.deramp(ctx, p["g_ref"], p["ref_slc"])
    ds = K.deramp(ctx, p["g_sec"], p["sec_slc"])
    ifg, pr, ps = K.pair(ctx, p["g_ref"], p["g_sec"], p["grid"], p["dem_t"], p["dem"], dr, ds,
                         interp, rl, cl, offpoly=p["offpoly"])
    filt = K.goldstein(ctx, ifg, alpha=ALPHA, mode=gold_mode)
    coh = K.coherence(ctx, ifg, pr, ps, win=win, fill="hash", seed=SEED)
    return dict(ifg=ifg, pow_ref=pr, pow_sec=ps, filt=filt, coh=coh)


def _assert_same(batch, want, what):
    for k, (got, exp) in enumerate(zip(batch.outputs(), want)):
        for key in ("ifg", "pow_ref", "pow_sec", "filt", "coh"):
            a, b = _bits(got[key]), _bits(exp[key])
            assert np.array_equal(a, b), f"{what}: pair {k} {key} differs in {np.count_nonzero(a != b)} values"


@pytest.mark.parametrize("interp,gold_mode,win,rl,cl", [("sinc", "fp16_split", 5, 1, 4),
                                                        ("bilinear", "fp32", 1, 1, 4),
                                                        ("sinc", "fp32", 3, 1, 4),
                                                        ("sinc", "fp16_split", 5, 2, 2)])
def test_batch_bit_identical_to_single_pairs(ctx, pairs, interp, gold_mode, win, rl, cl):
    from sentinel_gpu_processor_amd.batch import PairBatch
    print(f"batch parity {interp} {gold_mode} win={win} looks {rl} x {cl}", flush=True)
    want = [_single(ctx, p, interp, gold_mode, win, rl, cl) for p in pairs]
    b = PairBatch(ctx, pairs, interp=interp, rl=rl, cl=cl, alpha=ALPHA, gold_mode=gold_mode,
                  coh_win=win, coh_fill="hash", seed=SEED)
    b.run(graph=False)
    _assert_same(b, want, "eager")
    # graph replays, with eager work and synchronisation in between
    for rep in range(3):
        for o in b.outputs():
            for t in o.values():
                t.zero_()
        b.run(graph=True)
        torch.cuda.synchronize()
        _ = torch.ones(1024, device="cuda").sum().item()
        _assert_same(b, want, f"graph replay {rep}")
    # the raw bursts are untouched
    for name, p in zip(NAMES, pairs):
        assert np.array_equal(_bits(p["ref_slc"]), _bits(_dev(scene(name)["ref_slc"])))
    b.close()


def _single_raw(ctx, p, win):
    from sentinel_gpu_processor_amd import kernels as K
    ifg, pr, ps = K.pair(ctx, p["g_ref"], p["g_sec"], p["grid"], p["dem_t"], p["dem"],
                         p["ref_slc"], p["sec_slc"], "sinc", 1, 4, offpoly=p["offpoly"],
                         raw=True)
    filt = K.goldstein(ctx, ifg, alpha=ALPHA, mode="fp16_split")
    coh = K.coherence(ctx, ifg, pr, ps, win=win, fill="hash", seed=SEED)
    return dict(ifg=ifg, pow_ref=pr, pow_sec=ps, filt=filt, coh=coh)


def test_batch_raw_bit_identical_to_single_pairs(ctx, pairs):
    """raw=True: no deramp pass, the pair kernel deramps the raw bursts as it
    stages them; each pair equals kernels.pair(raw=True) + Goldstein +
    coherence, eagerly, on graph replays and for two launch shapes (one
    block per tile, persistent blocks on the tile queue)."""
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd.batch import PairBatch
    want = [_single_raw(ctx, p, 5) for p in pairs]
    b = PairBatch(ctx, pairs, interp="sinc", rl=1, cl=4, alpha=ALPHA, gold_mode="fp16_split",
                  coh_win=5, coh_fill="hash", seed=SEED, raw=True)
    assert all(o.get("ref_deramped") is None for o in b.pairs)
    b.run(graph=False)
    _assert_same(b, want, "raw eager")
    for rep in range(3):
        for o in b.outputs():
            for t in o.values():
                t.zero_()
        b.run(graph=True)
        torch.cuda.synchronize()
        _ = torch.ones(1024, device="cuda").sum().item()
        _assert_same(b, want, f"raw graph replay {rep}")
    for shape in ({"pair_blocks_per_cu": 2, "gold_blocks_per_cu": 1},
                  {"pair_blocks_per_cu": N.BLOCK_PER_ITEM, "gold_blocks_per_cu": N.BLOCK_PER_ITEM}):
        b.launch = shape
        b.run(graph=True)
        _assert_same(b, want, f"raw shape {shape}")
    for name, p in zip(NAMES, pairs):
        assert np.array_equal(_bits(p["ref_slc"]), _bits(_dev(scene(name)["ref_slc"])))
    b.close()
    # raw bursts need the sinc tile kernel
    with pytest.raises(N.S1IfgError, match="raw bursts"):
        PairBatch(ctx, pairs[:1], interp="bilinear", raw=True)


def test_batch_launch_shapes_do_not_change_results(ctx, pairs):
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd.batch import PairBatch
    want = [_single(ctx, p, "sinc", "fp16_split", 5) for p in pairs]
    b = PairBatch(ctx, pairs, interp="sinc", gold_mode="fp16_split", coh_win=5, seed=SEED)
    for shape in ({"pair_block_x": 32, "pair_blocks_per_cu": 1, "gold_blocks_per_cu": 1,
                   "stream_blocks_per_cu": 1},
                  {"pair_block_x": 128, "pair_blocks_per_cu": 16, "gold_blocks_per_cu": 4,
                   "stream_blocks_per_cu": 32},
                  {"pair_block_x": 256, "pair_blocks_per_cu": 2},
                  {"pair_blocks_per_cu": N.BLOCK_PER_ITEM, "gold_blocks_per_cu": N.BLOCK_PER_ITEM,
                   "stream_blocks_per_cu": N.BLOCK_PER_ITEM}):
        b.launch = shape
        got = b.launch
        for k, v in shape.items():
            assert got[k] == v
        b.run(graph=True)
        _assert_same(b, want, f"shape {shape}")
    b.close()


def test_batch_autotune(ctx, pairs, tmp_path, monkeypatch):
    from sentinel_gpu_processor_amd import autotune
    from sentinel_gpu_processor_amd.batch import PairBatch
    monkeypatch.setenv("S1IFG_TUNE_CACHE", str(tmp_path / "tune.json"))
    want = [_single(ctx, p, "sinc", "fp16_split", 5) for p in pairs]
    b = PairBatch(ctx, pairs, interp="sinc", gold_mode="fp16_split", coh_win=5, seed=SEED)
    shape, ms, trace = autotune.tune(b, reps=1)
    assert ms > 0 and len(trace) > 5
    assert b.launch == shape
    # cached on the second call
    shape2, ms2, trace2 = autotune.tune(b, reps=1)
    assert shape2 == shape and trace2 == []
    b.run(graph=True)
    _assert_same(b, want, "tuned")
    b.close()


def test_batch_rejects_bad_input(ctx, pairs):
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd.batch import PairBatch
    p = dict(pairs[0])
    p["ref_deramped"] = p["ref_slc"]          # would deramp the raw burst in place
    with pytest.raises(N.S1IfgError, match="alias"):
        PairBatch(ctx, [p])
    with pytest.raises(N.S1IfgError, match="window"):
        PairBatch(ctx, pairs[:1], coh_win=17)
    with pytest.raises(N.S1IfgError, match="pair_block_x"):
        PairBatch(ctx, pairs[:1], launch={"pair_block_x": 48})
