"""Interferogram stack: the pairs of a multi-date stack (by default the
sequential pairs (k, k+1)) through one batch of persistent kernels, each
date's burst deramped once however many pairs use it (raw=True: no deramp
pass; the pair kernel deramps each date's samples as it stages them).

The reference forms a stack's interferograms pair by pair through files
(interfere.py:536-700 reads the pair list; the crossmul daemon processes one
task per burst pair).  Here every date stays resident in HBM: raw burst and
deramped burst per date, products per pair.
"""

import torch

from . import _native as N
from .batch import PairBatch


def sequential_pairs(n_dates):
    return [(k, k + 1) for k in range(n_dates - 1)]


class Stack:
    """dates: list of dicts (orbit, burst, slc) sharing one target grid and
    DEM; slc is a complex64 device tensor (lines x samples).  offpolys:
    optional {pair index: 6 coefficients} for the secondary of that pair."""

    def __init__(self, ctx, dates, grid, dem_t, dem, pairs=None, interp="sinc", rl=1, cl=4,
                 alpha=0.5, gold_mode="fp16_split", coh_win=5, coh_fill="clamp", seed=0,
                 offpolys=None, launch=None, raw=False):
        self.ctx = ctx
        self.pairs = list(pairs) if pairs is not None else sequential_pairs(len(dates))
        self.geoms = [N.Geometry(ctx, d["orbit"], d["burst"]) for d in dates]
        self.raw = [d["slc"] for d in dates]
        used = sorted({k for p in self.pairs for k in p})
        self.deramped = {} if raw else {k: torch.empty_like(self.raw[k]) for k in used}
        tasks = []
        for i, (a, b) in enumerate(self.pairs):
            tasks.append(dict(g_ref=self.geoms[a], g_sec=self.geoms[b], grid=grid,
                              dem_t=dem_t, dem=dem, ref_slc=self.raw[a], sec_slc=self.raw[b],
                              ref_deramped=self.deramped.get(a), sec_deramped=self.deramped.get(b),
                              offpoly=(offpolys or {}).get(i)))
        self.batch = PairBatch(ctx, tasks, interp=interp, rl=rl, cl=cl, alpha=alpha,
                               gold_mode=gold_mode, coh_win=coh_win, coh_fill=coh_fill,
                               seed=seed, launch=launch, raw=raw)

    def run(self, graph=True):
        """Enqueue every pair of the stack (one graph launch when graph=True)."""
        self.batch.run(graph=graph)

    def products(self):
        """Per pair: dict(ifg, pow_ref, pow_sec, filt, coh) device tensors."""
        return self.batch.outputs()

    def close(self):
        self.batch.close()
        for g in self.geoms:
            g.close()
