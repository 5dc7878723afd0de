"""Batched burst pairs: all bursts of an IW pair (or the pairs of a stack)
through one launch sequence of persistent kernels, optionally replayed as
one HIP graph (s1ifg_batch_* in include/s1ifg.h).

Each pair's products are bit-identical to running kernels.deramp (x2),
kernels.pair, kernels.goldstein and kernels.coherence on that pair alone
(with raw=True: kernels.pair(raw=True), no deramp pass); the
batch exists to remove the per-burst launch and synchronisation overhead of
the reference's loops (geo2rdr.cu:552-664, crossmul_daemon.cu:921-1117) and
to fill the GPU with bursts of uneven size.
"""

import ctypes

import numpy as np
import torch

from . import _native as N


class PairBatch:
    """A batch of burst pairs with every device buffer bound at creation.

    pairs: list of dicts with keys
      g_ref, g_sec      Geometry of the reference / secondary burst
      grid              target grid dict (nrow, ncol, lat0, lon0, dlat, dlon)
      dem_t, dem        DEM tensor (float32 or int16) and its grid dict
      ref_slc, sec_slc  raw bursts, complex64 (lines x samples); None when
                        ref_deramped / sec_deramped hold the deramped burst
      dh, offpoly       optional height error (float32, grid shape) and the
                        secondary's 6-coefficient offset polynomial
    Output buffers (ref_deramped, sec_deramped, ifg, pow_ref, pow_sec, filt,
    coh) are allocated unless given in the dict.
    raw=True: the pair kernel deramps the raw bursts as it stages them (no
    deramp pass, no deramped buffers; sinc interpolation, rl * cl <= 4).
    """

    def __init__(self, ctx, pairs, interp="sinc", rl=1, cl=4, alpha=0.5,
                 gold_mode="fp16_split", coh_win=5, coh_fill="hash", seed=0, launch=None,
                 filt=True, coh=True, raw=False):
        self.ctx = ctx
        self.lib = ctx.lib
        self.handle = None
        self.pairs = []
        n = len(pairs)
        if n < 1:
            raise ValueError("empty batch")
        tasks = (N.PairTask * n)()
        fill = {"clamp": N.COH_FILL_CLAMP, "nan": N.COH_FILL_NAN, "hash": N.COH_FILL_HASH}[coh_fill]
        for k, p in enumerate(pairs):
            br, bs = p["g_ref"].burst, p["g_sec"].burst
            q = dict(p)
            for side, b in (("ref", br), ("sec", bs)):
                shape = (b["lines"], b["samples"])
                src = q.get(f"{side}_slc")
                if raw and src is None:
                    raise ValueError(f"pair {k}: raw=True needs {side}_slc")
                if q.get(f"{side}_deramped") is None and not raw:
                    if src is None:
                        raise ValueError(f"pair {k}: {side}_slc or {side}_deramped is needed")
                    q[f"{side}_deramped"] = torch.empty_like(src)
                for name in (f"{side}_slc", f"{side}_deramped"):
                    t = q.get(name)
                    if t is not None and (t.dtype != torch.complex64 or tuple(t.shape) != shape):
                        raise ValueError(f"pair {k}: {name} must be complex64 {shape}")
            dev = (q["ref_slc"] if raw else q["ref_deramped"]).device
            g = p["grid"]
            nr, nc = g["nrow"] // rl, g["ncol"] // cl
            if q.get("ifg") is None:
                q["ifg"] = torch.empty((nr, nc), dtype=torch.complex64, device=dev)
            if q.get("pow_ref") is None:
                q["pow_ref"] = torch.empty((nr, nc), dtype=torch.float32, device=dev)
            if q.get("pow_sec") is None:
                q["pow_sec"] = torch.empty((nr, nc), dtype=torch.float32, device=dev)
            if filt and q.get("filt") is None:
                q["filt"] = torch.empty((nr, nc), dtype=torch.complex64, device=dev)
            if coh and q.get("coh") is None:
                q["coh"] = torch.empty((nr, nc), dtype=torch.float32, device=dev)
            t = tasks[k]
            t.g_ref = p["g_ref"].handle.value
            t.g_sec = p["g_sec"].handle.value
            t.grid = N.make_grid(g)
            t.dem = N.make_dem(p["dem_t"], p["dem"])
            t.dh = _ptrval(q.get("dh"))
            if q.get("offpoly") is not None:
                t.offpoly_sec[:] = [float(v) for v in np.asarray(q["offpoly"], np.float64).reshape(6)]
                t.use_offpoly = 1
            t.ref_slc = _ptrval(q.get("ref_slc"))
            t.sec_slc = _ptrval(q.get("sec_slc"))
            t.ref_deramped = _ptrval(q.get("ref_deramped"))
            t.sec_deramped = _ptrval(q.get("sec_deramped"))
            t.ifg_ml = _ptrval(q["ifg"])
            t.pow_ref = _ptrval(q["pow_ref"])
            t.pow_sec = _ptrval(q["pow_sec"])
            t.filt = _ptrval(q.get("filt") if filt else None)
            t.coh = _ptrval(q.get("coh") if coh else None)
            self.pairs.append(q)
        o = N.BatchOpts(interp=N.INTERP[interp], rl=rl, cl=cl, alpha=alpha,
                        gold_mode=N.GOLD_MODE[gold_mode], coh_win=coh_win, coh_fill=fill,
                        coh_seed=seed, raw=int(bool(raw)))
        for key, v in (launch or {}).items():
            setattr(o, key, int(v))
        h = ctypes.c_void_p()
        N.check(self.lib.s1ifg_batch_create(ctx.handle, tasks, n, ctypes.byref(o),
                                            ctypes.byref(h)))
        self.handle = h
        self.n = n
        self.key = (f"{interp}|{rl}x{cl}|{gold_mode}|coh{coh_win}|f{int(filt)}c{int(coh)}"
                    + ("|raw" if raw else ""))

    # ------------------------------------------------------------------
    def run(self, graph=True):
        """Enqueue the batch on the context stream (as one graph launch when
        graph=True)."""
        N.check(self.lib.s1ifg_batch_run(self.handle, 1 if graph else 0))

    @property
    def launch(self):
        o = N.BatchOpts()
        N.check(self.lib.s1ifg_batch_get_opts(self.handle, ctypes.byref(o)))
        return {k: getattr(o, k) for k in N.LAUNCH_FIELDS}

    @launch.setter
    def launch(self, shape):
        o = N.BatchOpts()
        N.check(self.lib.s1ifg_batch_get_opts(self.handle, ctypes.byref(o)))
        for k, v in shape.items():
            if k not in N.LAUNCH_FIELDS:
                raise KeyError(k)
            setattr(o, k, int(v))
        N.check(self.lib.s1ifg_batch_set_launch(self.handle, ctypes.byref(o)))

    def scratch_bytes(self):
        v = ctypes.c_int64()
        N.check(self.lib.s1ifg_batch_bytes(self.handle, ctypes.byref(v)))
        return v.value

    def outputs(self):
        keys = ("ifg", "pow_ref", "pow_sec", "filt", "coh")
        return [{k: p.get(k) for k in keys} for p in self.pairs]

    def close(self):
        if self.handle:
            self.lib.s1ifg_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptrval(t):
    p = N.ptr(t)
    return None if p is None else p.value
