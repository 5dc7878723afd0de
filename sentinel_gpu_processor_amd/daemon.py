"""Burst-pair crossmul driver with the file contract of the reference's
``crossmul_daemon`` (crossmul_daemon.cu:566-1192, 1744-1912), on one GPU.

Contract kept from the reference:
- tasks: one ``ref.gslc sec.gslc out.int`` per line; malformed lines are
  skipped with a warning (crossmul_daemon.cu:1501-1524);
- overlap of the two windows = intersection snapped inward to look
  multiples; empty overlap fails the task (crossmul_daemon.cu:606-632);
- output: int32[64] header [nrow0/rl, ncol0/cl, left/cl, top/rl, right/cl,
  bottom/rl] + cf32 looks (float phase with --out-float)
  (crossmul_daemon.cu:660-666, 1134-1192);
- stdout: ``OK <out>``, ``PROGRESS done failed total elapsed_s``,
  ``SUMMARY ok failed total elapsed_s``; stderr:
  ``[crossmul_daemon] FAIL <out> - <reason>``; exit status 1 if any task
  failed (crossmul_daemon.cu:1911);
- all-or-nothing tuning arguments (crossmul_daemon.cu:1762-1785).

Design: an I/O thread pool reads the needed rows of both files ahead of the
GPU (bounded by the slot count).  The GPU stage has ``streams`` lanes
(the reference's per-GPU streams and task slots, crossmul_daemon.cu:127-261,
921-975): each lane owns a HIP stream (default priority), a library
context bound to it, page-locked host staging buffers and device buffers,
all grown on demand and reused.  A task on a lane is: copy the row blocks
into the lane's page-locked buffers, asynchronous host-to-device copies,
the fused crossmul + looks kernel directly on the file layout (the column
crop is a pointer offset with the file's row stride), the optional phase
kernel, an asynchronous device-to-host copy of the looks, and an event.
The loop polls the lanes' events (as the reference polls its streams) and
hands finished looks to a writer thread, so one task's transfers overlap
another's kernel.  A two-entry image cache serves consecutive tasks that
share an image (the task order from
``interfere._reorder_tasks_for_cache_locality`` maximises that).
"""

import argparse
import collections
import concurrent.futures as cf
import ctypes
import os
import sys
import threading
import time

import numpy as np

from .sario import NHEAD, read_header, save_cropped


class TaskFail(Exception):
    pass


def overlap(h1, h2, rowlook, collook):
    """Intersection snapped inward to look multiples (crossmul_daemon.cu:612-623)."""
    h1 = [int(v) for v in h1[:6]]
    h2 = [int(v) for v in h2[:6]]
    left = max(h1[2], h2[2])
    left = (left + collook - 1) // collook * collook
    right = min(h1[4], h2[4]) // collook * collook
    top = max(h1[3], h2[3])
    top = (top + rowlook - 1) // rowlook * rowlook
    bottom = min(h1[5], h2[5]) // rowlook * rowlook
    if bottom - top <= 0 or right - left <= 0:
        return None
    return left, top, right, bottom


def read_tasks(path, err=sys.stderr):
    tasks = []
    with open(path) as f:
        for line in f:
            if not line.strip():
                continue
            w = line.split()
            if len(w) < 3:
                print(f"[crossmul_daemon] WARNING: skipping malformed task line: {line.rstrip()}",
                      file=err)
                continue
            tasks.append((w[0], w[1], w[2]))
    return tasks


class _ImageCache:
    """Small LRU of decoded images keyed by path (the reference keeps one
    cached image, crossmul_daemon.cu:668-764)."""

    def __init__(self, size=2):
        self.size = size
        self.d = collections.OrderedDict()
        self.lock = threading.Lock()

    def get(self, key, loader):
        with self.lock:
            if key in self.d:
                self.d.move_to_end(key)
                return self.d[key]
        v = loader()
        with self.lock:
            self.d[key] = v
            while len(self.d) > self.size:
                self.d.popitem(last=False)
        return v


def _read_rows(path, h, r0, r1):
    ncol = int(h[4] - h[2])
    n = (r1 - r0) * ncol
    with open(path, "rb") as f:
        f.seek(NHEAD * 4 + (r0 - int(h[3])) * ncol * 8)
        d = np.fromfile(f, dtype=np.complex64, count=n)
    if d.size != n:
        raise TaskFail(f"I/O error: {path} truncated ({d.size} of {n} values)")
    return d.reshape(r1 - r0, ncol)


class CrossmulDriver:
    def __init__(self, rowlook=1, collook=1, out_float=False, io_workers=2, max_slots=4,
                 verbose=False, out=None, err=None, progress_every=2.0, streams=2):
        self.rl = int(rowlook)
        self.cl = int(collook)
        self.out_float = bool(out_float)
        self.io_workers = max(1, int(io_workers))
        self.max_slots = max(1, int(max_slots))
        self.streams = max(1, int(streams))
        self.verbose = verbose
        self.out = out or sys.stdout
        self.err = err or sys.stderr
        self.progress_every = progress_every
        self.cache = _ImageCache(2)

    def _emit(self, msg, err=False):
        print(msg, file=self.err if err else self.out, flush=True)

    def _prepare(self, task):
        ref, sec, out = task
        try:
            h1 = read_header(ref)
        except Exception:
            raise TaskFail(f"cannot read header: {ref}")
        try:
            h2 = read_header(sec)
        except Exception:
            raise TaskFail(f"cannot read header: {sec}")
        ov = overlap(h1, h2, self.rl, self.cl)
        if ov is None:
            raise TaskFail("empty overlap region")
        left, top, right, bottom = ov
        try:
            fa = self.cache.get(ref, lambda: _read_rows(ref, h1, int(h1[3]), int(h1[5])))
            fb = self.cache.get(sec, lambda: _read_rows(sec, h2, int(h2[3]), int(h2[5])))
            a = fa[top - int(h1[3]):bottom - int(h1[3])]
            b = fb[top - int(h2[3]):bottom - int(h2[3])]
        except TaskFail:
            raise
        except Exception as e:
            raise TaskFail(f"I/O error: {e}")
        header = [int(h1[0]) // self.rl, int(h1[1]) // self.cl, left // self.cl, top // self.rl,
                  right // self.cl, bottom // self.rl]
        return dict(out=out, a=a, b=b, la=left - int(h1[2]), lb=left - int(h2[2]),
                    nrow=bottom - top, ncol=right - left, header=header)

    def run(self, tasks, ctx=None):
        """Process every task; returns (succeeded, failed).  ctx: the device
        is taken from it (default: the runtime's); the lanes have their own
        contexts and streams."""
        import torch
        from .runtime import context
        base = ctx or context()
        lanes = [_Lane(base.device) for _ in range(self.streams)]
        n = len(tasks)
        self.ok = self.failed = 0
        t0 = time.time()
        last = t0
        self._writes = collections.deque()
        with cf.ThreadPoolExecutor(self.io_workers) as io, cf.ThreadPoolExecutor(1) as wr:
            self._wr = wr
            pending = collections.deque()
            it = iter(tasks)
            for _ in range(max(self.max_slots, self.streams)):
                t = next(it, None)
                if t is None:
                    break
                pending.append((t, io.submit(self._prepare, t)))
            while pending or any(l.busy for l in lanes):
                progressed = False
                for lane in lanes:  # retire finished lanes (cudaStreamQuery-style)
                    if lane.busy and lane.done():
                        self._retire(lane)
                        progressed = True
                for lane in lanes:  # start ready tasks on free lanes, in task order
                    if lane.busy or not pending or not pending[0][1].done():
                        continue
                    task, fut = pending.popleft()
                    t = next(it, None)
                    if t is not None:
                        pending.append((t, io.submit(self._prepare, t)))
                    progressed = True
                    try:
                        lane.launch(task, fut.result(), self.rl, self.cl, self.out_float)
                    except TaskFail as e:
                        self._fail(task, str(e))
                    except Exception as e:  # GPU / launch errors fail the task, not the run
                        lane.reset()
                        self._fail(task, f"GPU error: {e}")
                self._collect(block=False)
                if not progressed:
                    busy = [l for l in lanes if l.busy]
                    if busy:
                        busy[0].event.synchronize()
                    elif pending:
                        cf.wait([pending[0][1]], return_when=cf.FIRST_COMPLETED)
                now = time.time()
                if now - last >= self.progress_every:
                    last = now
                    self._emit(f"PROGRESS {self.ok + self.failed} {self.failed} {n} {int(now - t0)}")
            self._collect(block=True)
        for lane in lanes:
            lane.close()
        self._emit(f"SUMMARY {n - self.failed} {self.failed} {n} {int(time.time() - t0)}")
        return n - self.failed, self.failed

    def _fail(self, task, why):
        self.failed += 1
        self._emit(f"[crossmul_daemon] FAIL {task[2]} - {why}", err=True)

    def _retire(self, lane):
        task, p, data = lane.take()
        h = p["header"]
        self._writes.append((task, self._wr.submit(save_cropped, p["out"], data, h[0], h[1], h[2],
                                                   h[3], h[4], h[5])))

    def _collect(self, block):
        while self._writes and (block or self._writes[0][1].done()):
            wt, wf = self._writes.popleft()
            try:
                wf.result()
                self.ok += 1
                self._emit(f"OK {wt[2]}")
            except Exception as e:
                self._fail(wt, f"write error: {e}")


class _Lane:
    """One GPU lane: a stream, a context bound to it, page-locked staging
    and device buffers (grown on demand), and the task in flight."""

    def __init__(self, device):
        import torch
        from . import _native as N
        self.dev = torch.device(f"cuda:{device}")
        self.stream = torch.cuda.Stream(device=self.dev)
        self.ctx = N.Context(device, stream=self.stream)
        self.host = {}
        self.devbuf = {}
        self.busy = False
        self.event = None

    def _buf(self, store, key, n, dtype, pinned):
        import torch
        b = store.get(key)
        if b is None or b.numel() < n or b.dtype != dtype:
            if pinned:
                b = torch.empty(max(n, 1), dtype=dtype, pin_memory=True)
            else:
                with torch.cuda.stream(self.stream):
                    b = torch.empty(max(n, 1), dtype=dtype, device=self.dev)
            store[key] = b
        return b[:n]

    def launch(self, task, p, rl, cl, out_float):
        import torch
        from . import _native as N
        from . import kernels as K
        a, b = p["a"], p["b"]
        nr, nc = p["nrow"] // rl, p["ncol"] // cl
        ha = self._buf(self.host, "a", a.size, torch.complex64, True)
        hb = self._buf(self.host, "b", b.size, torch.complex64, True)
        np.copyto(ha.numpy().reshape(a.shape), a)
        np.copyto(hb.numpy().reshape(b.shape), b)
        da = self._buf(self.devbuf, "a", a.size, torch.complex64, False)
        db = self._buf(self.devbuf, "b", b.size, torch.complex64, False)
        ifg = self._buf(self.devbuf, "ifg", nr * nc, torch.complex64, False).view(nr, nc)
        odt = torch.float32 if out_float else torch.complex64
        ho = self._buf(self.host, "out", nr * nc, odt, True)
        with torch.cuda.stream(self.stream):
            da.copy_(ha, non_blocking=True)
            db.copy_(hb, non_blocking=True)
            pa = ctypes.c_void_p(da.data_ptr() + 8 * p["la"])
            pb = ctypes.c_void_p(db.data_ptr() + 8 * p["lb"])
            N.check(self.ctx.lib.s1ifg_crossmul_looks(
                self.ctx.handle, pa, a.shape[1], pb, b.shape[1], p["nrow"], p["ncol"], rl, cl,
                N.ptr(ifg), None, None))
            if out_float:
                ph = self._buf(self.devbuf, "ph", nr * nc, torch.float32, False).view(nr, nc)
                res = K.phase(self.ctx, ifg, out=ph)
            else:
                res = ifg
            ho.copy_(res.reshape(-1), non_blocking=True)
            self.event = torch.cuda.Event()
            self.event.record(self.stream)
        self.task, self.p, self.shape, self.out = task, p, (nr, nc), ho
        self.busy = True

    def done(self):
        return self.event.query()

    def take(self):
        self.event.synchronize()
        data = self.out.numpy().reshape(self.shape).copy()
        self.busy = False
        return self.task, self.p, data

    def reset(self):
        if self.event is not None:
            self.event.synchronize()
        self.busy = False

    def close(self):
        self.stream.synchronize()
        self.ctx.close()


def main(argv=None):
    ap = argparse.ArgumentParser(prog="crossmul_daemon")
    ap.add_argument("--rowlook", type=int, default=1)
    ap.add_argument("--collook", type=int, default=1)
    ap.add_argument("--io-workers", type=int, default=-1)
    ap.add_argument("--cpu-workers", type=int, default=-1)
    ap.add_argument("--gpu-workers", type=int, default=-1)
    ap.add_argument("--streams-per-gpu", type=int, default=-1)
    ap.add_argument("--max-slots", type=int, default=-1)
    ap.add_argument("--tasks-file", default="")
    ap.add_argument("--out-float", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    tune = [a.io_workers, a.cpu_workers, a.streams_per_gpu, a.max_slots, a.gpu_workers]
    if not (all(v == -1 for v in tune) or all(v > 0 for v in tune)):
        print("[FATAL] Invalid parameter configuration. Please either omit all tuning arguments "
              "for hardware-managed auto-tuning, or provide the complete set of parameters "
              "(--io-workers, --cpu-workers, --streams-per-gpu, --gpu-workers, --max-slots). "
              "Partial overrides are not allowed.", file=sys.stderr)
        return 1
    if not a.tasks_file:
        print("[crossmul_daemon] ERROR: --tasks-file <path> is required.", file=sys.stderr)
        return 1
    if not os.path.exists(a.tasks_file):
        print(f"[crossmul_daemon] ERROR: cannot open tasks file: {a.tasks_file}", file=sys.stderr)
        return 1
    tasks = read_tasks(a.tasks_file)
    io_workers = a.io_workers if a.io_workers > 0 else min(4, os.cpu_count() or 1)
    slots = a.max_slots if a.max_slots > 0 else 4
    streams = a.streams_per_gpu if a.streams_per_gpu > 0 else 2
    drv = CrossmulDriver(a.rowlook, a.collook, a.out_float, io_workers, slots, a.verbose,
                         streams=streams)
    _, failed = drv.run(tasks)
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
