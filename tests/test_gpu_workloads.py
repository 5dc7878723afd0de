"""GPU parity at the batched, differentiable and stack workloads' sizes.

C3  bursts of the synthetic IW pair (sizes 1450-1530 x 20000-25500 from the
    seed), one per subswath, through one batched graph: each pair's products
    against the oracle on crops of its grid.
C4  the 1500 x 25000 pair with loss = -mean(5 x 5 coherence) and a height
    error dh ~ N(0, 5 m): the gradient with respect to the 6 offset-
    polynomial coefficients and dh against central finite differences of the
    forward at full size (1e-3 px, 1e-2 m), and against the FP64 oracle
    gradient on a crop.
C5  the 20-date stack of 1500 x 25000 bursts with atmospheric screens,
    19 sequential pairs in one batch: every pair against the oracle on a
    crop (phase, coherence and Goldstein errors within 1e-3)."""

import numpy as np
import pytest

from oracle import adjoint as A
from oracle import geom as G
from oracle import goldstein as GS
from oracle import ifg as I
from sentinel_gpu_processor_amd import synthetic as S
from tests.configs import c2_scene, crop_grid, oracle_pair_crop, window_deramped
from tests.scenes import phase_rms

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from sentinel_gpu_processor_amd import _native as N
    c = N.Context(0)
    yield c
    c.sync()
    c.close()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _geoms(ctx, orbit_ref, burst_ref, orbit_sec, burst_sec):
    from sentinel_gpu_processor_amd import _native as N
    return N.Geometry(ctx, orbit_ref, burst_ref), N.Geometry(ctx, orbit_sec, burst_sec)


def goldstein_interior(filt, ifg, r0, c0, nr, nc):
    """Goldstein output of a full image over rows [r0, r0 + nr) x columns
    [c0, c0 + nc) (r0, c0, nr, nc multiples of 8, 32 away from the image
    edges) against the oracle filter of the sub-image with a 32-pixel
    border: the sub-image's patch grid is the full image's there, so the
    patches covering the region are the same.  Returns (got, want, mask)."""
    sub = ifg[r0 - 32:r0 + nr + 32, c0 - 32:c0 + nc + 32]
    want, wsum = GS.goldstein(sub.astype(np.complex64), 32, 0.5)
    want = want[32:-32, 32:-32]
    m = wsum[32:-32, 32:-32] > 0
    return filt[r0:r0 + nr, c0:c0 + nc], want, m


def check_pair_crop(sc, prods, ref_h, sec_h, crop, interp="sinc"):
    """ifg / powers / coherence / Goldstein of a batch (or single) pair's
    products against the oracle over a crop (r0, c0, nr, nc) of its target
    grid (c0, nc multiples of 32)."""
    r0, c0, nr, nc = crop
    g = crop_grid(sc["grid"], r0, c0, nr, nc)
    oi, o1, o2, _, _ = oracle_pair_crop(sc, ref_h, sec_h, g, interp)
    sl = np.s_[r0:r0 + nr, c0 // 4:(c0 + nc) // 4]
    ifg = _np(prods["ifg"])
    pr, ps = _np(prods["pow_ref"])[sl], _np(prods["pow_sec"])[sl]
    m = (o1 > 0) & (o2 > 0)
    assert m.mean() > 0.2
    assert np.count_nonzero((pr > 0) != (o1 > 0)) <= 8
    big = m & (np.abs(oi) > 1e-3 * np.abs(oi).max())
    assert phase_rms(ifg[sl], oi, big) < 1e-4
    coh = _np(prods["coh"])[sl]
    cw = I.coherence_window(oi, o1, o2, 5)
    inner = np.zeros_like(m)
    inner[2:-2, 2:-2] = True  # the 5 x 5 window inside the crop
    ok = inner & m & np.isfinite(cw) & np.isfinite(coh)
    assert ok.mean() > 0.2
    assert np.max(np.abs(coh[ok] - cw[ok])) < 1e-3
    # Goldstein over the crop's interior, from the full filtered image
    fr0, fc0 = max(r0, 32), c0 // 4 + 32
    fnr = min(r0 + nr, ifg.shape[0] - 32) - fr0
    fnr -= fnr % 8
    fnc = nc // 4 - 64
    got, want, fm = goldstein_interior(_np(prods["filt"]), ifg, fr0, fc0, fnr, fnc)
    # against the oracle filter of the GPU interferogram (the filter stage),
    # and of the oracle's own interferogram on the same pixels (the chain)
    fm &= np.abs(want) > 1e-3 * np.abs(want).max()
    assert phase_rms(got, want, fm) < 3e-4
    o_sub = np.zeros_like(ifg)
    o_sub[sl] = oi
    _, want2, fm2 = goldstein_interior(_np(prods["filt"]), o_sub, fr0, fc0, fnr, fnc)
    rows = np.s_[fr0 - r0:fr0 - r0 + fnr, 32:32 + fnc]
    fm2 &= big[rows] & (np.abs(want2) > 1e-3 * np.abs(want2).max())
    # the oracle's image holds zeros beyond the crop: keep pixels whose
    # covering patches (origins within 31 pixels) lie inside the crop, or
    # at the image's own edge
    rr = np.arange(fr0, fr0 + fnr)[:, None]
    cc = np.arange(fc0, fc0 + fnc)[None, :]
    inside = (((rr - 31 >= r0) | (r0 == 0)) & ((rr + 31 < r0 + nr) | (r0 + nr == ifg.shape[0]))
              & (cc - 31 >= c0 // 4) & (cc + 31 < (c0 + nc) // 4))
    fm2 &= inside
    assert fm2.sum() > 1000
    assert phase_rms(got, want2, fm2) < 1e-3


# ---------------------------------------------------------------------- C3
def test_c3_batch_matches_oracle(ctx):
    """Three bursts of the IW pair (one per subswath, sizes from the seed),
    one batched graph, each pair against the oracle on a 1024-column crop
    of its grid (all lines)."""
    from sentinel_gpu_processor_amd.batch import PairBatch
    sel = (4, 13, 22)
    scs = S.make_iw(seed=0, fidelity="fast", device="cuda:0", select=sel)
    sizes = S.iw_sizes(0)
    tasks, keep = [], []
    for k, sc in zip(sel, scs):
        assert (sc["burst_ref"]["lines"], sc["burst_ref"]["samples"]) == sizes[k]
        assert 1450 <= sizes[k][0] <= 1530 and 20000 <= sizes[k][1] <= 25500
        gr, gs = _geoms(ctx, sc["orbit_ref"], sc["burst_ref"], sc["orbit_sec"], sc["burst_sec"])
        dem_t = _dev(sc["dem"]["data"])
        keep.append((gr, gs, dem_t))
        tasks.append(dict(g_ref=gr, g_sec=gs, grid=sc["grid"], dem_t=dem_t, dem=sc["dem"],
                          ref_slc=sc["ref_slc"], sec_slc=sc["sec_slc"], offpoly=sc["offpoly"]))
    b = PairBatch(ctx, tasks, interp="sinc", rl=1, cl=4, alpha=0.5, gold_mode="fp16_split",
                  coh_win=5, coh_fill="nan")
    b.run(graph=True)
    b.run(graph=True)  # a replay gives the same products
    outs = b.outputs()
    for sc, prods in zip(scs, outs):
        N = sc["grid"]["ncol"]
        c0 = (N // 2) // 64 * 64
        check_pair_crop(sc, prods, _np(sc["ref_slc"]), _np(sc["sec_slc"]),
                        (0, c0, sc["grid"]["nrow"], 1024))
    b.close()


# ---------------------------------------------------------------------- C4
@pytest.fixture(scope="module")
def c4(ctx):
    from sentinel_gpu_processor_amd import kernels as K
    from sentinel_gpu_processor_amd.adjoint import interior_cells
    sc = c2_scene()
    gr, gs = _geoms(ctx, sc["orbit_ref"], sc["burst_ref"], sc["orbit_sec"], sc["burst_sec"])
    dem_t = _dev(sc["dem"]["data"])
    dref = K.deramp(ctx, gr, sc["ref_slc"])
    dsec = K.deramp(ctx, gs, sc["sec_slc"])
    g = torch.Generator(device="cuda:0")
    g.manual_seed(4)
    grid = sc["grid"]
    dh = (torch.randn(grid["nrow"], grid["ncol"], device="cuda:0", generator=g) * 5.0)
    p = np.asarray(sc["offpoly"], np.float64)
    mask = interior_cells(ctx, gr, gs, grid, dem_t, sc["dem"], dref, dsec, "sinc", p, dh, 1, 4, 5)
    return dict(sc=sc, gr=gr, gs=gs, dem_t=dem_t, dref=dref, dsec=dsec, dh=dh.contiguous(), p=p,
                mask=mask)


def _loss_and_grad(ctx, C, p, dh, grid=None, mask=None, want_grad=True):
    from sentinel_gpu_processor_amd import kernels as K
    sc = C["sc"]
    grid = grid or sc["grid"]
    mask = C["mask"] if mask is None else mask
    cnt = int(mask.sum())
    ifg, pr, ps = K.pair(ctx, C["gr"], C["gs"], grid, C["dem_t"], sc["dem"], C["dref"],
                         C["dsec"], "sinc", 1, 4, dh=dh, offpoly=p)
    coh = K.coherence(ctx, ifg, pr, ps, win=5)
    loss = -float(coh.double()[mask].sum()) / cnt
    if not want_grad:
        return loss
    gcoh = torch.where(mask, torch.tensor(-1.0 / cnt, device="cuda:0"),
                       torch.tensor(0.0, device="cuda:0")).float().contiguous()
    gi, g1, g2 = K.coherence_backward(ctx, ifg, pr, ps, gcoh, win=5)
    g_off, g_dh = K.pair_backward(ctx, C["gr"], C["gs"], grid, C["dem_t"], sc["dem"], C["dref"],
                                  C["dsec"], gi, g1, g2, "sinc", 1, 4, dh=dh, offpoly=p)
    return loss, _np(g_off), g_dh


def test_c4_gradient_matches_finite_differences(ctx, c4):
    p, dh = c4["p"], c4["dh"]
    assert int(c4["mask"].sum()) > 1_000_000
    _, gp, gdh = _loss_and_grad(ctx, c4, p, dh)
    fd = np.zeros(6)
    for k in range(6):
        d = np.zeros(6)
        d[k] = 1e-3
        fd[k] = (_loss_and_grad(ctx, c4, p + d, dh, want_grad=False)
                 - _loss_and_grad(ctx, c4, p - d, dh, want_grad=False)) / 2e-3
    scale = np.abs(fd).max()
    assert np.all(np.abs(gp - fd) <= 5e-3 * np.maximum(np.abs(fd), 0.05 * scale)), (gp, fd)
    g = torch.Generator(device="cuda:0")
    g.manual_seed(5)
    delta = torch.randn(dh.shape, device="cuda:0", generator=g)
    fd_h = (_loss_and_grad(ctx, c4, p, dh + 1e-2 * delta, want_grad=False)
            - _loss_and_grad(ctx, c4, p, dh - 1e-2 * delta, want_grad=False)) / 2e-2
    an_h = float((gdh.double() * delta.double()).sum())
    assert abs(an_h - fd_h) <= 5e-3 * abs(fd_h), (an_h, fd_h)


def test_c4_gradient_matches_oracle_on_crop(ctx, c4):
    """The same loss on a 512 x 1024 crop of the C2 grid (the crop as the
    target grid, dh its part of the full field): GPU gradient against the
    FP64 oracle gradient (oracle/adjoint.py)."""
    sc = c4["sc"]
    r0, c0, nr, nc = 500, 12032, 512, 1024
    g = crop_grid(sc["grid"], r0, c0, nr, nc)
    dh_c = c4["dh"][r0:r0 + nr, c0:c0 + nc].contiguous()
    dh_h = _np(dh_c).astype(np.float64)
    h = G.target_heights(g, sc["dem"]) + dh_h
    ref_h, sec_h = _np(sc["ref_slc"]), _np(sc["sec_slc"])
    su, dr, _ = window_deramped(sc["orbit_ref"], sc["burst_ref"], ref_h, g, h)
    ss, ds, _ = window_deramped(sc["orbit_sec"], sc["burst_sec"], sec_h, g, h, c4["p"])
    osc = dict(sc, grid=g)
    F = A.pair_forward(osc, (su, ss), h, dr, ds, "sinc", c4["p"], 1, 4)
    mask = A.interior_mask(F, 1, 4, 5)
    assert mask.mean() > 0.3
    loss_o, gp_o, gh_o = A.grad_neg_mean_coherence(osc, (su, ss), h, dr, ds, "sinc", c4["p"], 1,
                                                   4, 5, mask)
    loss, gp, gh = _loss_and_grad(ctx, c4, c4["p"], dh_c, grid=g,
                                  mask=torch.from_numpy(mask).cuda())
    gh = _np(gh)
    assert loss == pytest.approx(loss_o, rel=1e-3)
    assert np.linalg.norm(gp - gp_o) <= 1e-2 * np.linalg.norm(gp_o), (gp, gp_o)
    assert np.max(np.abs(gh - gh_o)) <= 1e-2 * np.max(np.abs(gh_o))
    c = np.sum(gh * gh_o) / np.sqrt(np.sum(gh * gh) * np.sum(gh_o * gh_o))
    assert c > 0.9999


# ---------------------------------------------------------------------- C5
def test_c5_stack_matches_oracle(ctx):
    """20 dates of 1500 x 25000 with atmospheric screens, 19 sequential
    pairs through one batched graph; every pair against the oracle on a
    512 x 512 crop (RMS phase, coherence and Goldstein errors <= 1e-3)."""
    from sentinel_gpu_processor_amd.stack import Stack
    c = S.CONFIGS["c2"]
    st = S.make_stack(n_dates=20, lines=c["lines"], samples=c["samples"], dem_shape=c["dem"],
                      dem_kind=c["dem_kind"], hmax=c["hmax"], seed=0, fidelity="fast",
                      device="cuda:0")
    dem_t = _dev(st["dem"]["data"])
    run = Stack(ctx, st["dates"], st["grid"], dem_t, st["dem"], interp="sinc", rl=1, cl=4,
                gold_mode="fp16_split", coh_win=5, coh_fill="nan")
    run.run(graph=True)
    prods = run.products()
    assert len(prods) == 19
    host = {}
    r0, c0, nr, nc = 480, 12288, 512, 512
    for k, (a, b) in enumerate(run.pairs):
        for i in (a, b):
            if i not in host:
                host[i] = _np(st["dates"][i]["slc"])
        sc = dict(orbit_ref=st["dates"][a]["orbit"], burst_ref=st["dates"][a]["burst"],
                  orbit_sec=st["dates"][b]["orbit"], burst_sec=st["dates"][b]["burst"],
                  grid=st["grid"], dem=st["dem"], offpoly=None)
        print(f"c5 pair {k}: dates {a}-{b}", flush=True)
        check_pair_crop(sc, prods[k], host[a], host[b], (r0, c0, nr, nc))
        host.pop(a)
    run.close()
