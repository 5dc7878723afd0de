"""Tensor-level wrappers of the libs1ifg entry points.

Every function takes and returns device tensors (complex64 / float32 /
float64) and enqueues the HIP kernels on the context's stream.
"""

import ctypes

import numpy as np
import torch

from . import _native as N


def _new(shape, dtype, like_device):
    return torch.empty(shape, dtype=dtype, device=like_device)


def _need(t, dtype, shape=None, name="tensor"):
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} shape {tuple(t.shape)} != {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def deramp(ctx, geom, slc, out=None):
    """TOPS deramp of one burst (geo2rdr.cu:138-158); out may be slc."""
    b = geom.burst
    _need(slc, torch.complex64, (b["lines"], b["samples"]), "burst")
    if out is None:
        out = torch.empty_like(slc)
    N.check(ctx.lib.s1ifg_deramp(ctx.handle, geom.handle, N.ptr(slc), N.ptr(out)))
    return out


def deramp_ci16(ctx, geom, raw, out=None):
    """TOPS deramp of one burst given as interleaved CInt16 (an int16 tensor
    of shape (lines, samples, 2): re, im), written as complex64; the
    Sentinel-1 measurement format read without a float copy
    (sentinel_scene.py:164-207 + geo2rdr.cu:138-158)."""
    b = geom.burst
    _need(raw, torch.int16, (b["lines"], b["samples"], 2), "CInt16 burst")
    if out is None:
        out = torch.empty((b["lines"], b["samples"]), dtype=torch.complex64, device=raw.device)
    _need(out, torch.complex64, (b["lines"], b["samples"]), "deramped burst")
    N.check(ctx.lib.s1ifg_deramp_ci16(ctx.handle, geom.handle, N.ptr(raw), N.ptr(out)))
    return out


def _dem_args(dem_t, dem):
    return N.make_dem(dem_t, dem)


def geo2rdr(ctx, geom, grid, dem_t, dem, dh=None, offpoly=None):
    """Per-target-pixel (rgoff, azoff, range) in FP64 (validation mode)."""
    g = N.make_grid(grid)
    d = _dem_args(dem_t, dem)
    shape = (grid["nrow"], grid["ncol"])
    rg = _new(shape, torch.float64, dem_t.device)
    az = _new(shape, torch.float64, dem_t.device)
    rng = _new(shape, torch.float64, dem_t.device)
    if dh is not None:
        _need(dh, torch.float32, shape, "dh")
    _keep, pp = N.offpoly_arg(offpoly)
    N.check(ctx.lib.s1ifg_geo2rdr(ctx.handle, geom.handle, ctypes.byref(g), ctypes.byref(d),
                                  N.ptr(dh), pp, N.ptr(rg), N.ptr(az), N.ptr(rng)))
    return rg, az, rng


def geocode(ctx, geom, grid, dem_t, dem, deramped, interp="bilinear", dh=None, offpoly=None,
            out=None):
    """Geocode a deramped burst onto the target grid (geo2rdr.cu:170-248)."""
    b = geom.burst
    _need(deramped, torch.complex64, (b["lines"], b["samples"]), "deramped burst")
    g = N.make_grid(grid)
    d = _dem_args(dem_t, dem)
    shape = (grid["nrow"], grid["ncol"])
    if out is None:
        out = _new(shape, torch.complex64, deramped.device)
    if dh is not None:
        _need(dh, torch.float32, shape, "dh")
    _keep, pp = N.offpoly_arg(offpoly)
    N.check(ctx.lib.s1ifg_geocode(ctx.handle, geom.handle, ctypes.byref(g), ctypes.byref(d),
                                  N.ptr(dh), pp, N.ptr(deramped), N.INTERP[interp], N.ptr(out)))
    return out


def valid_rows(ctx, img):
    """(first, last_exclusive) non-zero rows (geo2rdr.cu:86-128, 273-299)."""
    _need(img, torch.complex64, None, "image")
    fl = torch.empty(2, dtype=torch.int32, device=img.device)
    N.check(ctx.lib.s1ifg_valid_rows(ctx.handle, N.ptr(img), img.shape[0], img.shape[1],
                                     N.ptr(fl)))
    f = fl.cpu().numpy()
    return int(f[0]), int(f[1])


def crossmul_looks(ctx, ref, sec, rl, cl, with_power=True, out=None):
    """Fused conj(ref)*sec + look means (sario.cu:199-286)."""
    _need(ref, torch.complex64, None, "ref")
    _need(sec, torch.complex64, tuple(ref.shape), "sec")
    nrow, ncol = ref.shape
    shape = (nrow // rl, ncol // cl)
    ifg = out if out is not None else _new(shape, torch.complex64, ref.device)
    pr = _new(shape, torch.float32, ref.device) if with_power else None
    ps = _new(shape, torch.float32, ref.device) if with_power else None
    N.check(ctx.lib.s1ifg_crossmul_looks(ctx.handle, N.ptr(ref), ncol, N.ptr(sec), ncol, nrow,
                                         ncol, rl, cl, N.ptr(ifg), N.ptr(pr), N.ptr(ps)))
    return ifg, pr, ps


def looks(ctx, img, rl, cl):
    """Complex look means (sario.cu:241-286)."""
    _need(img, torch.complex64, None, "image")
    nrow, ncol = img.shape
    out = _new((nrow // rl, ncol // cl), torch.complex64, img.device)
    N.check(ctx.lib.s1ifg_looks(ctx.handle, N.ptr(img), ncol, nrow, ncol, rl, cl, N.ptr(out)))
    return out


def phase(ctx, z, out=None):
    """atan2 of a complex image (sario.cu:231-239)."""
    _need(z, torch.complex64, None, "image")
    if out is None:
        out = _new(tuple(z.shape), torch.float32, z.device)
    _need(out, torch.float32, tuple(z.shape), "phase")
    N.check(ctx.lib.s1ifg_phase(ctx.handle, N.ptr(z), z.numel(), N.ptr(out)))
    return out


def amplitude(ctx, slc, rl, cl):
    """sqrt of power looks (coherence.py:46-48; 1x1 quirk sqrt|s|)."""
    _need(slc, torch.complex64, None, "slc")
    nrow, ncol = slc.shape
    amp = _new((nrow // rl, ncol // cl), torch.float32, slc.device)
    N.check(ctx.lib.s1ifg_amplitude(ctx.handle, N.ptr(slc), ncol, nrow, ncol, rl, cl, N.ptr(amp)))
    return amp


_FILL = {"clamp": N.COH_FILL_CLAMP, "nan": N.COH_FILL_NAN, "hash": N.COH_FILL_HASH}


def coherence(ctx, ifg_ml, pow_ref, pow_sec, win=1, fill="clamp", seed=0, out=None):
    """Look-cell (win=1, coherence.py:142) or windowed coherence."""
    _need(ifg_ml, torch.complex64, None, "ifg")
    _need(pow_ref, torch.float32, tuple(ifg_ml.shape), "pow_ref")
    _need(pow_sec, torch.float32, tuple(ifg_ml.shape), "pow_sec")
    nr, nc = ifg_ml.shape
    coh = out if out is not None else _new((nr, nc), torch.float32, ifg_ml.device)
    N.check(ctx.lib.s1ifg_coherence(ctx.handle, N.ptr(ifg_ml), N.ptr(pow_ref), N.ptr(pow_sec),
                                    nr, nc, win, _FILL[fill], seed, N.ptr(coh)))
    return coh


def stitch_subswath(ctx, bursts, out_float=False, bounds=None, offsets=False):
    """Mosaic of one subswath's burst interferograms on the device
    (interfere.py:81-150).

    bursts: (left, top, right, bottom, data, estimate) per burst, sorted by
    top + bottom; data complex64 (float32 phase when out_float) of the
    burst's bounds; estimate: its ids differ from the previous burst's.
    bounds: the subswath box, default the first / last bursts' box
    (sario.py:230-237).  Returns (mosaic, (left, top, right, bottom)) and,
    with offsets=True, the float64 phase offset applied to each burst."""
    if not bursts:
        raise ValueError("no bursts")
    dt = torch.float32 if out_float else torch.complex64
    if bounds is None:
        f, l_ = bursts[0], bursts[-1]
        bounds = (min(f[0], l_[0]), min(f[1], l_[1]), max(f[2], l_[2]), max(f[3], l_[3]))
    left, top, right, bottom = (int(v) for v in bounds)
    arr = (N.StitchBurst * len(bursts))()
    dev = None
    for i, (bl, bt, br, bb, data, est) in enumerate(bursts):
        _need(data, dt, (int(bb) - int(bt), int(br) - int(bl)), f"burst {i}")
        dev = data.device
        arr[i] = N.StitchBurst(N.ptr(data), int(bl), int(bt), int(br), int(bb), int(bool(est)), 0)
    mosaic = _new((bottom - top, right - left), dt, dev)
    offs = _new((len(bursts),), torch.float64, dev) if offsets else None
    N.check(ctx.lib.s1ifg_stitch_subswath(ctx.handle, arr, len(bursts), int(bool(out_float)),
                                          N.ptr(mosaic), left, top, right, bottom,
                                          N.ptr(offs) if offsets else None))
    if offsets:
        return mosaic, (left, top, right, bottom), offs
    return mosaic, (left, top, right, bottom)


def merge_subswath(ctx, src, dst, top, left, out_float=False):
    """dst[top:, left:] takes src where src has data (interfere.py:170-179)."""
    dt = torch.float32 if out_float else torch.complex64
    _need(src, dt, None, "subswath")
    _need(dst, dt, None, "image")
    h, w = src.shape
    if top < 0 or left < 0 or top + h > dst.shape[0] or left + w > dst.shape[1]:
        raise ValueError(f"subswath {h} x {w} at ({top}, {left}) outside {tuple(dst.shape)}")
    if not dst.is_contiguous():
        raise ValueError("image must be contiguous")
    base = N.ptr(dst).value + (top * dst.shape[1] + left) * dst.element_size()
    N.check(ctx.lib.s1ifg_merge_subswath(ctx.handle, N.ptr(src), h, w, int(bool(out_float)),
                                         ctypes.c_void_p(base), dst.shape[1]))
    return dst


def goldstein(ctx, ifg, alpha=0.5, n_win=32, mode="fp32", out=None):
    """Goldstein filter (goldstein.cu:87-408), deterministic overlap-add."""
    _need(ifg, torch.complex64, None, "ifg")
    nr, nc = ifg.shape
    out = out if out is not None else torch.empty_like(ifg)
    N.check(ctx.lib.s1ifg_goldstein(ctx.handle, N.ptr(ifg), nr, nc, n_win, float(alpha),
                                    N.GOLD_MODE[mode], N.ptr(out)))
    return out


def pair(ctx, g_ref, g_sec, grid, dem_t, dem, ref_deramped, sec_deramped, interp="bilinear",
         rl=1, cl=4, dh=None, offpoly=None, out=None, noalloc=False, raw=False):
    """Fused burst pair -> (ifg_ml, pow_ref, pow_sec) on the look grid.
    noalloc=True calls the entry point that allocates no device memory
    (grid tables in the library's fixed array; pass `out` too); raw=True
    takes the raw bursts and deramps their samples as they are read
    (s1ifg_pair_raw)."""
    g = N.make_grid(grid)
    d = _dem_args(dem_t, dem)
    br, bs = g_ref.burst, g_sec.burst
    _need(ref_deramped, torch.complex64, (br["lines"], br["samples"]), "ref burst")
    _need(sec_deramped, torch.complex64, (bs["lines"], bs["samples"]), "sec burst")
    shape = (grid["nrow"] // rl, grid["ncol"] // cl)
    if dh is not None:
        _need(dh, torch.float32, (grid["nrow"], grid["ncol"]), "dh")
    if out is None:
        out = (_new(shape, torch.complex64, ref_deramped.device),
               _new(shape, torch.float32, ref_deramped.device),
               _new(shape, torch.float32, ref_deramped.device))
    ifg, pr, ps = out
    _keep, pp = N.offpoly_arg(offpoly)
    if noalloc and raw:
        raise ValueError("the allocation-free pair takes deramped bursts")
    fn = (ctx.lib.s1ifg_pair_noalloc if noalloc else
          ctx.lib.s1ifg_pair_raw if raw else ctx.lib.s1ifg_pair)
    N.check(fn(ctx.handle, g_ref.handle, g_sec.handle, ctypes.byref(g), ctypes.byref(d),
               N.ptr(dh), pp, N.ptr(ref_deramped), N.ptr(sec_deramped), N.INTERP[interp], rl, cl,
               N.ptr(ifg), N.ptr(pr), N.ptr(ps)))
    return ifg, pr, ps


def pair_backward(ctx, g_ref, g_sec, grid, dem_t, dem, ref_deramped, sec_deramped, g_ifg,
                  g_pr=None, g_ps=None, interp="bilinear", rl=1, cl=4, dh=None, offpoly=None,
                  want_dh=True):
    """Adjoint of pair(): (dL/d offpoly as float64[6] device tensor,
    dL/d dh as float32 target-grid tensor or None)."""
    g = N.make_grid(grid)
    d = _dem_args(dem_t, dem)
    shape = (grid["nrow"] // rl, grid["ncol"] // cl)
    for t, name, dt in ((g_ifg, "g_ifg", torch.complex64), (g_pr, "g_pr", torch.float32),
                        (g_ps, "g_ps", torch.float32)):
        if t is not None:
            _need(t, dt, shape, name)
    if dh is not None:
        _need(dh, torch.float32, (grid["nrow"], grid["ncol"]), "dh")
    dev = ref_deramped.device
    g_off = _new((6,), torch.float64, dev)
    g_dh = _new((grid["nrow"], grid["ncol"]), torch.float32, dev) if want_dh else None
    _keep, pp = N.offpoly_arg(offpoly)
    N.check(ctx.lib.s1ifg_pair_backward(
        ctx.handle, g_ref.handle, g_sec.handle, ctypes.byref(g), ctypes.byref(d), N.ptr(dh), pp,
        N.ptr(ref_deramped), N.ptr(sec_deramped), N.INTERP[interp], rl, cl, N.ptr(g_ifg),
        N.ptr(g_pr), N.ptr(g_ps), N.ptr(g_off), N.ptr(g_dh)))
    return g_off, g_dh


def coherence_backward(ctx, ifg_ml, pow_ref, pow_sec, g_coh, win=1):
    """Adjoint of coherence(): (g_ifg complex64, g_pow_ref, g_pow_sec)."""
    _need(ifg_ml, torch.complex64, None, "ifg")
    shape = tuple(ifg_ml.shape)
    _need(pow_ref, torch.float32, shape, "pow_ref")
    _need(pow_sec, torch.float32, shape, "pow_sec")
    _need(g_coh, torch.float32, shape, "g_coh")
    gi = _new(shape, torch.complex64, ifg_ml.device)
    g1 = _new(shape, torch.float32, ifg_ml.device)
    g2 = _new(shape, torch.float32, ifg_ml.device)
    N.check(ctx.lib.s1ifg_coherence_backward(ctx.handle, N.ptr(ifg_ml), N.ptr(pow_ref),
                                             N.ptr(pow_sec), shape[0], shape[1], win,
                                             N.ptr(g_coh), N.ptr(gi), N.ptr(g1), N.ptr(g2)))
    return gi, g1, g2


def orbit_eval(ctx, geom, t):
    """Orbit position/velocity/acceleration at device times t (float64)."""
    _need(t, torch.float64, None, "t")
    n = t.numel()
    pos = _new((n, 3), torch.float64, t.device)
    vel = _new((n, 3), torch.float64, t.device)
    acc = _new((n, 3), torch.float64, t.device)
    N.check(ctx.lib.s1ifg_orbit_eval(ctx.handle, geom.handle, N.ptr(t), n, N.ptr(pos),
                                     N.ptr(vel), N.ptr(acc)))
    return pos, vel, acc


def to_numpy(t):
    return t.detach().cpu().numpy()


def from_numpy(a, device="cuda:0"):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


# ------------------------------------------------------- PS phase similarity
def ps_scan_offsets(rdmin, rdmax):
    """Neighbour offsets (dr, dc) in the reference's exploration order
    (phase_similarity.cu:295-366), an int32 numpy array of shape (n, 2)."""
    lib = N.load()
    n = lib.s1ifg_ps_scan_offsets(int(rdmin), int(rdmax), None, 0)
    N.check(min(n, 0))
    out = np.empty((n, 2), np.int32)
    if n:
        N.check(min(lib.s1ifg_ps_scan_offsets(int(rdmin), int(rdmax),
                                              out.ctypes.data_as(ctypes.c_void_p), n), 0))
    return out


def _ps_stack(stack):
    _need(stack, torch.complex64, None, "interferogram stack")
    if stack.dim() != 3:
        raise ValueError("stack must be (nifg, nrow, ncol)")
    return stack.shape


def _ps_mask(ps, shape):
    _need(ps, torch.int32, shape, "candidate mask")


def ps_normalize(ctx, stack, out=None):
    """Unit phasors, 0 where |z| <= 1e-8 (phase_similarity.cu:116-134)."""
    _need(stack, torch.complex64, None, "stack")
    if out is None:
        out = torch.empty_like(stack)
    _need(out, torch.complex64, stack.shape, "output")
    N.check(ctx.lib.s1ifg_ps_normalize(ctx.handle, N.ptr(stack), N.ptr(out), stack.numel()))
    return out


def ps_update(ctx, ps, sim, threshold, overwrite=True):
    """In place: ps = 1 where sim >= threshold, else 0 if overwrite
    (phase_similarity.cu:145-157)."""
    _need(ps, torch.int32, None, "candidate mask")
    _need(sim, torch.float32, ps.shape, "similarity")
    N.check(ctx.lib.s1ifg_ps_update(ctx.handle, N.ptr(ps), N.ptr(sim), float(threshold),
                                    int(bool(overwrite)), ps.numel()))
    return ps


def ps_median_similarity(ctx, stack_norm, ps, nneigh, rdmin, rdmax, out=None):
    nifg, nrow, ncol = _ps_stack(stack_norm)
    _ps_mask(ps, (nrow, ncol))
    if out is None:
        out = torch.empty((nrow, ncol), dtype=torch.float32, device=ps.device)
    _need(out, torch.float32, (nrow, ncol), "output")
    N.check(ctx.lib.s1ifg_ps_median_similarity(ctx.handle, N.ptr(stack_norm), nifg, nrow, ncol,
                                               N.ptr(ps), int(nneigh), int(rdmin), int(rdmax),
                                               N.ptr(out)))
    return out


def ps_maximum_similarity(ctx, stack_norm, ps, nneigh, rdmin, rdmax, out=None):
    nifg, nrow, ncol = _ps_stack(stack_norm)
    _ps_mask(ps, (nrow, ncol))
    if out is None:
        out = torch.empty((nrow, ncol), dtype=torch.float32, device=ps.device)
    _need(out, torch.float32, (nrow, ncol), "output")
    N.check(ctx.lib.s1ifg_ps_maximum_similarity(ctx.handle, N.ptr(stack_norm), nifg, nrow, ncol,
                                                N.ptr(ps), int(nneigh), int(rdmin), int(rdmax),
                                                N.ptr(out)))
    return out


def ps_similarity(ctx, stack, ps, nneigh, rdmin, rdmax, med_sim_th, med=None, mx=None):
    """The phase_similarity executable's sequence on a resident stack
    (phase_similarity.cu:375-548): returns (med_sim, max_sim); ps (int32,
    device) is updated in place to the candidates whose median similarity
    reaches med_sim_th."""
    nifg, nrow, ncol = _ps_stack(stack)
    _ps_mask(ps, (nrow, ncol))
    if med is None:
        med = torch.empty((nrow, ncol), dtype=torch.float32, device=ps.device)
    if mx is None:
        mx = torch.empty((nrow, ncol), dtype=torch.float32, device=ps.device)
    _need(med, torch.float32, (nrow, ncol), "med_sim")
    _need(mx, torch.float32, (nrow, ncol), "max_sim")
    N.check(ctx.lib.s1ifg_ps_similarity(ctx.handle, N.ptr(stack), nifg, nrow, ncol, N.ptr(ps),
                                        int(nneigh), int(rdmin), int(rdmax), float(med_sim_th),
                                        N.ptr(med), N.ptr(mx)))
    return med, mx
