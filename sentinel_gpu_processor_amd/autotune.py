"""Launch-shape autotuner of the batched pair kernels.

The tunable shape is the pair-kernel tile width and the persistent grid of
each kernel (blocks per CU); results do not depend on it, so a shape can be
picked by time alone.  The search is coordinate descent from the built-in
defaults over each field's candidates, timing whole batch runs with device
events on the context stream.  Results are cached in a JSON file keyed by
the device, the processing options and the batch's size class.
"""

import gc
import json
import os
import time

import torch

from . import _native as N

CANDIDATES = {
    "pair_block_x": (32, 64, 128),
    "pair_blocks_per_cu": (3, 4, 6, 8, 16, N.BLOCK_PER_ITEM),
    "gold_blocks_per_cu": (1, 2, 3, 4),
    "stream_blocks_per_cu": (4, 8, 16, 32, 64, N.BLOCK_PER_ITEM),
}


def cache_path():
    return os.environ.get("S1IFG_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "s1ifg", "tune.json"))


def cache_key(batch, extra=""):
    props = torch.cuda.get_device_properties(batch.ctx.device)
    px = sum(p["grid"]["nrow"] * p["grid"]["ncol"] for p in batch.pairs)
    size_class = int(round(float(px)).bit_length())  # log2 bucket of target pixels
    return (f"{props.name}|cu{props.multi_processor_count}|n{batch.n}|px2^{size_class}|"
            f"{batch.key}{extra}")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, table):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + f".tmp{os.getpid()}"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def time_batch(batch, reps=2, graph=False):
    """Mean device time (ms) of `reps` runs after one warm-up run (Python's
    garbage collector held off meanwhile: a collection pass between the
    launches would leave the device idle and skew the comparison)."""
    stream = batch.ctx.stream
    batch.run(graph=graph)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    gc_was_on = gc.isenabled()
    gc.disable()
    try:
        e0.record(stream)
        for _ in range(reps):
            batch.run(graph=graph)
        e1.record(stream)
        e1.synchronize()
    finally:
        if gc_was_on:
            gc.enable()
    return e0.elapsed_time(e1) / reps


def tune(batch, reps=2, use_cache=True, log=None):
    """Pick the launch shape of `batch` (sets it) and return
    (shape, ms, trace)."""
    path = cache_path()
    key = cache_key(batch)
    table = _load(path) if use_cache else {}
    if key in table:
        batch.launch = table[key]["shape"]
        return table[key]["shape"], table[key]["ms"], []
    t0 = time.time()
    best = dict(batch.launch)
    best_ms = time_batch(batch, reps)
    trace = [(dict(best), best_ms)]
    for field, values in CANDIDATES.items():
        for v in values:
            if v == best[field]:
                continue
            shape = dict(best, **{field: v})
            batch.launch = shape
            ms = time_batch(batch, reps)
            trace.append((shape, ms))
            if log:
                log(f"tune {shape} {ms:.3f} ms")
            if ms < best_ms:
                best, best_ms = shape, ms
        batch.launch = best
    if use_cache:
        table = _load(path)
        table[key] = {"shape": best, "ms": best_ms, "tuned_s": round(time.time() - t0, 2)}
        _save(path, table)
    return best, best_ms, trace
