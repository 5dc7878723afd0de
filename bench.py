"""Burst-pair interferogram formation throughput on one MI355X.

A step is one pass of the hot path over one synthetic burst pair that is
already resident in HBM: deramp of both bursts, geo2rdr + resampling +
reramp + flattening of both, crossmul with 1x4 looks, Goldstein filter
(alpha 0.5, 32/8) and 5x5 windowed coherence.

Forms:
  plain  per-stage kernels (deramp x2, geocode x2 to full-resolution
         geocoded SLCs, crossmul+looks, Goldstein FP32, coherence)
  fused  deramp x2, fused pair kernel (geocoded SLCs never written),
         Goldstein (FP16-split MFMA when built, else FP32), coherence

Prints one JSON line (rank 0).  Timing uses HIP events on the stream the
library launches on.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "burst-pair output Mpix/s + % of HBM BW peak, plain vs batched/FP16-split"
HBM_PEAK = 8.0e12


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--config", default="c2", choices=["c1", "c2"])
    p.add_argument("--interp", default="sinc", choices=["bilinear", "sinc"])
    p.add_argument("--forms", default="plain,fused")
    p.add_argument("--gold-mode", default="auto", choices=["auto", "fp32", "fp16_split"])
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-sample", default="c1")
    return p.parse_args()


def stage_bytes(L, N, dem_px, rl, cl, dem_bytes=4):
    """Algorithmic (compulsory) HBM bytes per launch of each stage."""
    n = L * N
    nml = (L // rl) * (N // cl)
    return {
        "deramp": 16 * n,                       # read + write one burst cf32
        "geocode": 8 * n + 8 * n + dem_bytes * dem_px,  # burst read once, target write
        "pair": 16 * n + dem_bytes * dem_px + 16 * nml,  # two bursts, DEM, ifg+2 powers
        "crossmul_looks": 16 * n + 16 * nml,
        "goldstein": 16 * nml,
        "coherence": 16 * nml + 4 * nml,
    }


def b_alg(L, N, dem_px, rl, cl, dem_bytes=4):
    n = L * N
    nml = (L // rl) * (N // cl)
    return 16 * n + dem_bytes * dem_px + 20 * nml


def run_gpu(args):
    import numpy as np
    import torch

    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    from sentinel_gpu_processor_amd import synthetic as S

    dev = torch.device("cuda:0")
    cfg = S.CONFIGS[args.config]
    sc = S.make_scene(lines=cfg["lines"], samples=cfg["samples"], dem_shape=cfg["dem"],
                      dem_kind=cfg["dem_kind"], hmax=cfg["hmax"], seed=0,
                      misreg=cfg["offset"], misreg_ramp=(0.2, 0.1) if cfg["offset"] else (0, 0),
                      fidelity="fast", device=dev)
    L, Nr = cfg["lines"], cfg["samples"]
    rl, cl = 1, 4
    ctx = N.Context(0)
    gr = N.Geometry(ctx, sc["orbit_ref"], sc["burst_ref"])
    gs = N.Geometry(ctx, sc["orbit_sec"], sc["burst_sec"])
    dem = sc["dem"]
    dem_t = torch.from_numpy(dem["data"]).to(dev)
    ref = sc["ref_slc"]
    sec = sc["sec_slc"]
    offpoly = sc["offpoly"]
    grid = sc["grid"]
    dref = torch.empty_like(ref)
    dsec = torch.empty_like(sec)
    g_r = torch.empty((grid["nrow"], grid["ncol"]), dtype=torch.complex64, device=dev)
    g_s = torch.empty_like(g_r)
    nml = (grid["nrow"] // rl, grid["ncol"] // cl)
    ifg = torch.empty(nml, dtype=torch.complex64, device=dev)
    pr = torch.empty(nml, dtype=torch.float32, device=dev)
    ps = torch.empty(nml, dtype=torch.float32, device=dev)
    filt = torch.empty_like(ifg)
    coh = torch.empty(nml, dtype=torch.float32, device=dev)

    gold_mode = args.gold_mode
    if gold_mode == "auto":
        try:
            K.goldstein(ctx, ifg[:64, :64].contiguous(), 0.5, mode="fp16_split")
            gold_mode = "fp16_split"
        except N.S1IfgError:
            gold_mode = "fp32"

    def plain(ev):
        ev.mark("deramp")
        K.deramp(ctx, gr, ref, out=dref)
        K.deramp(ctx, gs, sec, out=dsec)
        ev.mark("geocode")
        K.geocode(ctx, gr, grid, dem_t, dem, dref, args.interp, out=g_r)
        K.geocode(ctx, gs, grid, dem_t, dem, dsec, args.interp, offpoly=offpoly, out=g_s)
        ev.mark("crossmul_looks")
        ifg_, _, _ = crossmul_into(K, ctx, g_r, g_s, rl, cl, ifg, pr, ps)
        ev.mark("goldstein")
        K.goldstein(ctx, ifg, 0.5, mode="fp32", out=filt)
        ev.mark("coherence")
        K.coherence(ctx, ifg, pr, ps, win=5, out=coh)
        ev.mark(None)

    def fused(ev):
        ev.mark("deramp")
        K.deramp(ctx, gr, ref, out=dref)
        K.deramp(ctx, gs, sec, out=dsec)
        ev.mark("pair")
        K.pair(ctx, gr, gs, grid, dem_t, dem, dref, dsec, args.interp, rl, cl,
               offpoly=offpoly, out=(ifg, pr, ps))
        ev.mark("goldstein")
        K.goldstein(ctx, ifg, 0.5, mode=gold_mode, out=filt)
        ev.mark("coherence")
        K.coherence(ctx, ifg, pr, ps, win=5, out=coh)
        ev.mark(None)

    forms = {"plain": plain, "fused": fused}
    results = {}
    for name in [f for f in args.forms.split(",") if f]:
        fn = forms[name]
        for _ in range(args.warmup):
            fn(NoEvents())
        torch.cuda.synchronize()
        ev = StageEvents()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        start.record()
        for _ in range(args.steps):
            fn(ev)
        stop.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = start.elapsed_time(stop) / args.steps
        results[name] = dict(ms_per_step=ms, wall_s=wall, stages=ev.averages(args.steps))
        print(f"[bench] {name}: {ms:.3f} ms/step  stages(ms/step) "
              + " ".join(f"{k}={v:.3f}" for k, v in results[name]["stages"].items()),
              file=sys.stderr, flush=True)
    results["_gold_mode"] = gold_mode
    results["_shape"] = (L, Nr, dem["data"].size, rl, cl)
    return results


def crossmul_into(K, ctx, a, b, rl, cl, ifg, pr, ps):
    from sentinel_gpu_processor_amd import _native as N
    nrow, ncol = a.shape
    N.check(ctx.lib.s1ifg_crossmul_looks(ctx.handle, N.ptr(a), ncol, N.ptr(b), ncol, nrow, ncol,
                                         rl, cl, N.ptr(ifg), N.ptr(pr), N.ptr(ps)))
    return ifg, pr, ps


class NoEvents:
    def mark(self, name):
        pass


class StageEvents:
    """Events recorded between stages on the current stream (the stream the
    library context is bound to)."""

    def __init__(self):
        import torch
        self.torch = torch
        self.marks = []

    def mark(self, name):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        self.marks.append((name, e))

    def averages(self, steps):
        tot = {}
        for (n0, e0), (_, e1) in zip(self.marks[:-1], self.marks[1:]):
            if n0 is None:
                continue
            tot[n0] = tot.get(n0, 0.0) + e0.elapsed_time(e1)
        return {k: v / steps for k, v in tot.items()}


def cpu_baseline(sample):
    """FP64 NumPy oracle chain on a bounded sample (single process)."""
    import numpy as np

    from oracle import geom as G
    from oracle import goldstein as GS
    from oracle import ifg as I
    from sentinel_gpu_processor_amd import synthetic as S

    cfg = S.CONFIGS[sample]
    sc = S.make_scene(lines=cfg["lines"], samples=cfg["samples"], dem_shape=cfg["dem"],
                      hmax=cfg["hmax"], seed=0, fidelity="fast", device="cuda:0")
    ref = sc["ref_slc"].cpu().numpy()
    sec = sc["sec_slc"].cpu().numpy()
    t0 = time.perf_counter()
    b = sc["burst_ref"]
    h = G.target_heights(sc["grid"], sc["dem"])
    su = G.deramp_setup(sc["orbit_ref"], b)
    ss = G.deramp_setup(sc["orbit_sec"], b)
    dr = G.deramp(ref, su).astype(np.complex64)
    ds = G.deramp(sec, ss).astype(np.complex64)
    o_r, _ = G.reproject(sc["orbit_ref"], b, su, sc["grid"], h, dr)
    o_s, _ = G.reproject(sc["orbit_sec"], b, ss, sc["grid"], h, ds)
    ifg = I.cpx_looks(I.conj_mul(o_r, o_s), 1, 4)
    p1 = I.cpx_looks(np.abs(o_r) ** 2, 1, 4).real
    p2 = I.cpx_looks(np.abs(o_s) ** 2, 1, 4).real
    GS.goldstein(ifg, 32, 0.5)
    I.coherence_window(ifg, p1, p2, 5)
    dt = time.perf_counter() - t0
    npx = cfg["lines"] * cfg["samples"]
    return dict(value=npx / dt / 1e6, unit="Mpix/s", cores=1, kind="port",
                sample=f"{sample}: one {cfg['lines']}x{cfg['samples']} burst pair through the "
                       f"FP64 NumPy oracle chain (bilinear, 1x4 looks, Goldstein 32/0.5, 5x5 "
                       f"coherence) in {dt:.1f} s")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(burst pairs shard across GPUs by process, see DESIGN.md)"}))
        sys.exit(2)
    res = run_gpu(args)
    L, Nr, dem_px, rl, cl = res["_shape"]
    npx = L * Nr
    balg = b_alg(L, Nr, dem_px, rl, cl)
    sb = stage_bytes(L, Nr, dem_px, rl, cl)
    forms = {}
    for name in ("plain", "fused"):
        if name not in res:
            continue
        r = res[name]
        ms = r["ms_per_step"]
        forms[name] = dict(ms_per_step=round(ms, 4), mpix_s=round(npx / (ms * 1e-3) / 1e6, 1),
                           hbm_frac=round(balg / (ms * 1e-3) / HBM_PEAK, 4),
                           stages_ms={k: round(v, 4) for k, v in r["stages"].items()})
    best = "fused" if "fused" in forms else "plain"
    fb = forms[best]
    r = res[best]
    # dominant kernel: the stage with the largest share of the step
    dom = max(r["stages"], key=r["stages"].get)
    per_launch = {"deramp": 2, "geocode": 2}.get(dom, 1)
    t_launch = r["stages"][dom] / per_launch * 1e-3
    achieved = sb[dom] / t_launch / 1e9
    roof = dict(kernel=dom, bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK / 1e9,
                unit="GB/s", frac=round(achieved * 1e9 / HBM_PEAK, 4), traffic=None,
                alg_bytes_per_launch=sb[dom])
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args.cpu_sample)
        except Exception as e:  # reported, never fatal for the GPU number
            cpu = dict(error=repr(e))
    out = {
        "metric": METRIC,
        "value": fb["mpix_s"],
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fb["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 geometry)",
        "data": "synthetic",
        "config": {"workload": f"{args.config}: {L}x{Nr} complex64 burst pair, DEM {dem_px} px "
                               f"f32, looks 1x4, {args.interp} resampling, Goldstein 32/8 "
                               f"alpha 0.5 ({res['_gold_mode']}), 5x5 coherence",
                   "form": best, "interp": args.interp},
        "hbm_frac_path": fb["hbm_frac"],
        "b_alg_bytes": balg,
        "forms": forms,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
