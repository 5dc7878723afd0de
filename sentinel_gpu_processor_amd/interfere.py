"""Drop-in for ``s1proc.interfere``: burst matching, per-burst
interferograms, stitching into IW interferograms (interfere.py:16-747).

The per-burst crossmul + looks runs in-process on the GPU through
``daemon.CrossmulDriver`` (the reference launches its crossmul_daemon
executable and parses its stdout, interfere.py:337-533); the file contract
(.gslc in, per-burst .int out, OK / FAIL accounting, skip-existing resume)
is unchanged.  Stitching (interfere.py:81-180) runs on the GPU
(s1ifg_stitch_subswath, s1ifg_merge_subswath).
"""

import glob
import os
from collections import defaultdict
from typing import List, Optional, Tuple

import numpy as np

from . import geocoordinates
from . import kernels as K
from .runtime import context, device
from .sario import BurstGroup, CroppedImage, Subswath

try:
    from tqdm.auto import tqdm
except Exception:  # pragma: no cover
    def tqdm(x, **k):
        return x

import logging

logger = logging.getLogger(__name__)


def match_bursts(ref_subswath: Subswath, sec_subswath: Subswath) -> List[Tuple[str, str]]:
    """Pair each reference burst with the secondary burst minimising
    |dtop| + |dbottom| (interfere.py:16-45)."""
    pairs = []
    if ref_subswath.is_empty() or sec_subswath.is_empty():
        return pairs
    st = np.array([b.top for b in sec_subswath.bursts])
    sb = np.array([b.bottom for b in sec_subswath.bursts])
    for rb in ref_subswath.bursts:
        best = int(np.argmin(np.abs(st - rb.top) + np.abs(sb - rb.bottom)))
        pairs.append((rb.data, sec_subswath.bursts[best].data))
    return pairs


def interfere_subswath(ref_subswath, sec_subswath, ifg_path) -> List[Tuple[str, str, str]]:
    """(ref, sec, out) triples of one subswath (interfere.py:48-78)."""
    out = []
    for a, b in match_bursts(ref_subswath, sec_subswath):
        n1 = os.path.splitext(os.path.basename(a))[0]
        n2 = os.path.splitext(os.path.basename(b))[0]
        out.append((a, b, os.path.join(ifg_path, n1 + "_" + n2 + ".int")))
    return out


def stitch_subswath(burst_ifgs, out_float):
    """Mosaic the burst interferograms of one subswath, removing the phase
    offset at scene boundaries (interfere.py:81-150).  The bursts are read
    on the host; the mosaic (overlap mean + median offsets, replacement) is
    built on the GPU by s1ifg_stitch_subswath.  The burst files are removed
    afterwards, as in the reference."""
    import torch  # noqa: F401  (device tensors)
    subswath = Subswath(burst_ifgs)
    left, top, right, bottom = subswath.bounds()
    dtype = np.float32 if out_float else np.complex64
    dev = device()
    bursts, prev = [], None
    for burst in subswath.bursts:
        burst.dtype = dtype
        words = os.path.basename(burst.data).split("_")
        uid = (words[1] + words[2], words[7] + words[8])
        data = burst.load_data(burst.left, burst.top, burst.right, burst.bottom)
        bursts.append((burst.left, burst.top, burst.right, burst.bottom,
                       K.from_numpy(np.ascontiguousarray(data, dtype=dtype), dev),
                       prev is not None and uid != prev))
        prev = uid
    mos, _ = K.stitch_subswath(context(), bursts, out_float=out_float,
                               bounds=(left, top, right, bottom))
    ifg = CroppedImage(subswath.nrow0, subswath.ncol0, left, top, right, bottom,
                       K.to_numpy(mos), dtype=dtype)
    for f in burst_ifgs:
        os.remove(f)
    return ifg


def stitch(burst_pairs, outfile, out_float):
    """Mosaic the three subswaths into one raw (headerless) image
    (interfere.py:153-180); the merge runs on the GPU
    (s1ifg_merge_subswath)."""
    import torch
    subs = []
    for i in range(1, 4):
        bp = [b for b in burst_pairs if (f"iw{i}" in b) and os.path.exists(b)]
        if bp:
            subs.append(stitch_subswath(bp, out_float))
    if not subs:
        logger.warning(f"Empty subswaths for {outfile}")
        return
    nrow0, ncol0 = subs[0].nrow0, subs[0].ncol0
    dtype = np.float32 if out_float else np.complex64
    ctx = context()
    full = torch.zeros((nrow0, ncol0), dtype=torch.float32 if out_float else torch.complex64,
                       device=device())
    for s in subs:
        K.merge_subswath(ctx, K.from_numpy(np.ascontiguousarray(s.data, dtype=dtype), device()),
                         full, s.top, s.left, out_float)
    mm = np.memmap(outfile, dtype=dtype, mode="w+", shape=(nrow0, ncol0))
    mm[:] = K.to_numpy(full)
    mm.flush()


def interfere_single_scene(ref_burst_group, sec_burst_group, ifg_path):
    """Triples for all three subswaths (interfere.py:183-211)."""
    out = []
    for i in range(3):
        out.extend(interfere_subswath(ref_burst_group.subswaths[i], sec_burst_group.subswaths[i],
                                      ifg_path))
    return out


def parse_fname(fn):
    b = os.path.basename(fn)
    return b[0:8], b[9:20]


def _reorder_tasks_for_cache_locality(task_items):
    """Order tasks so consecutive ones share an image (interfere.py:221-334):
    drain every remaining task touching the cached image, then continue from
    the other image of the last drained task; cold-start from the first
    remaining task at dead ends.  ref/sec columns are never swapped."""
    n = len(task_items)
    if n <= 1:
        return list(task_items)
    adj = defaultdict(set)
    for i, (r, s, _o) in enumerate(task_items):
        adj[r].add(i)
        adj[s].add(i)
    remaining = set(range(n))
    out = []

    def take(i):
        remaining.discard(i)
        r, s, _ = task_items[i]
        adj[r].discard(i)
        adj[s].discard(i)
        out.append(task_items[i])

    i0 = min(remaining)
    take(i0)
    cached = task_items[i0][0]
    while remaining:
        batch = [i for i in list(adj.get(cached, ())) if i in remaining]
        if batch:
            for i in batch:
                take(i)
            lr, ls, _ = out[-1]
            cached = ls if lr == cached else lr
        else:
            i = min(remaining)
            take(i)
            cached = task_items[i][0]
    return out


def _run_crossmul_daemon(task_items, rowlook, collook, out_float, io_workers, cpu_workers,
                         gpu_workers, streams_per_gpu, max_slots, verbose) -> Tuple[int, int]:
    """Run every (ref, sec, out) task through the GPU crossmul driver;
    returns (succeeded, failed) as the reference parses them from the
    daemon's SUMMARY line (interfere.py:337-533)."""
    from .daemon import CrossmulDriver
    tasks = _reorder_tasks_for_cache_locality(task_items)
    io = io_workers if io_workers and io_workers > 0 else min(4, os.cpu_count() or 1)
    slots = max_slots if max_slots and max_slots > 0 else 4
    streams = streams_per_gpu if streams_per_gpu and streams_per_gpu > 0 else 2
    with open(os.devnull, "w") as devnull:
        drv = CrossmulDriver(rowlook, collook, out_float, io, slots, verbose,
                             out=None if verbose else devnull, streams=streams)
        return drv.run(tasks)


def interfere(img_pair_file: str, slc_path: str, rscfile: str, small_rsc_file: str,
              ifg_path: str = "igrams", rowlook: int = 1, collook: int = 1, out_float: bool = False,
              io_workers: int = -1, cpu_workers: int = -1, gpu_workers: int = -1,
              streams_per_gpu: int = -1, max_slots: int = -1, verbose: bool = False):
    """Form interferograms for every date pair in ``img_pair_file``
    (interfere.py:536-704)."""
    os.makedirs(ifg_path, exist_ok=True)
    rsc = geocoordinates.GeoCoordinates(rscfile)
    rsclook = rsc.take_look(rowlook, collook)
    if os.path.dirname(small_rsc_file):
        os.makedirs(os.path.dirname(small_rsc_file), exist_ok=True)
    rsclook.save_as_rsc(small_rsc_file)
    ref_dates, sec_dates, groups = [], [], {}
    with open(os.path.join(ifg_path, img_pair_file), "r") as f:
        for line in f:
            w = line.split()
            if len(w) < 2:
                continue
            ref_dates.append(w[0])
            sec_dates.append(w[1])
            for d in (w[0], w[1]):
                if d not in groups:
                    groups[d] = BurstGroup(glob.glob(os.path.join(slc_path, f"{d}*.gslc")))
    burst_pairs, pair_map = [], {}
    for rd, sd in zip(ref_dates, sec_dates):
        outfile = os.path.join(ifg_path, f"{rd}_{sd}.int")
        if os.path.exists(outfile):
            continue
        bp = interfere_single_scene(groups[rd], groups[sd], ifg_path)
        pair_map[outfile] = [b[2] for b in bp]
        if not all(os.path.exists(b[2]) for b in bp):
            burst_pairs.extend(bp)
    if not burst_pairs:
        logger.warning("Did not find any burst pairs.")
        for outfile in pair_map:
            stitch(pair_map[outfile], outfile, out_float)
        return
    tuning = [io_workers, cpu_workers, gpu_workers, streams_per_gpu, max_slots]
    auto = all(v == -1 for v in tuning)
    manual = all(isinstance(v, int) and v > 0 for v in tuning)
    if not (auto or manual):
        import sys
        print("[FATAL] Invalid parameter configuration. Please either omit all tuning arguments "
              "for hardware-managed auto-tuning, or provide the complete set of parameters "
              "(--io-workers, --cpu-workers, --streams-per-gpu, --gpu-workers, --max-slots). "
              "Partial overrides are not allowed.", file=sys.stderr)
        raise SystemExit(1)
    tasks = [t for t in burst_pairs if not os.path.exists(t[2])]
    ok, failed = _run_crossmul_daemon(tasks, rowlook, collook, out_float, io_workers, cpu_workers,
                                      gpu_workers, streams_per_gpu, max_slots, verbose)
    if failed > 0:
        logger.warning("%d burst pair(s) failed.", failed)
    for outfile in tqdm(pair_map, desc="stitching"):
        stitch(pair_map[outfile], outfile, out_float)


def run_interfere(config: str = "config.yaml", verbose: bool = False):
    """interfere() with the config file's io / proc settings (interfere.py:707-747)."""
    from .config import load_config
    cfg = load_config(config)
    io, p = cfg.io, cfg.proc

    def opt(v):
        return v if v is not None else -1

    interfere(img_pair_file=io.img_pair_file, slc_path=io.slc_path, rscfile=io.rsc_file,
              small_rsc_file=io.multilook_rsc_file, ifg_path=io.ifg_path, rowlook=p.rowlook,
              collook=p.collook, out_float=False, io_workers=opt(p.io_workers),
              cpu_workers=opt(p.cpu_workers), gpu_workers=opt(p.gpu_workers),
              streams_per_gpu=opt(p.streams_per_gpu), max_slots=opt(p.max_slots), verbose=verbose)
