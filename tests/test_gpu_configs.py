"""GPU parity at the benchmark workloads' sizes.

C1  the CPU-runnable crop, whole (512 x 2048, Gaussian hill 0-1500 m):
    every stage against the FP64 oracle chain.
C2  the 1500 x 25000 pair on the fractal 0-2500 m DEM with misregistration:
    the full-size pair kernel (with its tile classes counted) against the
    oracle on a 1500 x 3072 crop, the FP16-split Goldstein filter and the
    5 x 5 coherence of that crop against the oracle, and the whole pair
    through size-independent properties (validity equals the geo2rdr
    offsets' window, coherence in [0, 1], the injected phase recovered).
"""

import numpy as np
import pytest

from oracle import goldstein as GS
from oracle import ifg as I
from tests.configs import c2_scene, crop_grid, oracle_pair_crop
from tests.scenes import oracle_chain, phase_rms, scene
from tests.test_oracle_tops import closed_form_phase

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


def _geoms(ctx, sc):
    from sentinel_gpu_processor_amd import _native as N
    return (N.Geometry(ctx, sc["orbit_ref"], sc["burst_ref"]),
            N.Geometry(ctx, sc["orbit_sec"], sc["burst_sec"]))


# ---------------------------------------------------------------------- C1
@pytest.mark.parametrize("interp", ["bilinear", "sinc"])
def test_c1_chain_matches_oracle(ctx, interp):
    """Deramp, the fused pair, Goldstein (both modes) and coherence (look
    cell and 5 x 5) of the whole C1 pair; errors against the FP64 oracle
    chain within 1e-3 rad / 1e-3 (with tighter stage bounds)."""
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("c1")
    oc = oracle_chain("c1", interp)
    gr, gs = _geoms(ctx, sc)
    dem = sc["dem"]
    dem_t = _dev(dem["data"])
    dref = K.deramp(ctx, gr, _dev(sc["ref_slc"]))
    dsec = K.deramp(ctx, gs, _dev(sc["sec_slc"]))
    for d, want in ((dref, oc["dr"]), (dsec, oc["ds"])):
        assert np.max(np.abs(_np(d) - want)) / np.max(np.abs(want)) < 2e-6
    ifg_t, pr_t, ps_t = K.pair(ctx, gr, gs, sc["grid"], dem_t, dem, dref, dsec, interp, 1, 4)
    ifg, pr, ps = _np(ifg_t), _np(pr_t), _np(ps_t)
    assert ifg.shape == (512, 512)
    m = (oc["p1"] > 0) & (oc["p2"] > 0)
    assert m.mean() > 0.3
    assert np.count_nonzero((pr > 0) != (oc["p1"] > 0)) <= 4
    big = m & (np.abs(oc["ifg"]) > 1e-3 * np.abs(oc["ifg"]).max())
    assert phase_rms(ifg, oc["ifg"], big) < 1e-4
    cg = np.abs(ifg) / np.sqrt(pr * ps + 1e-30)
    cw = np.abs(oc["ifg"]) / np.sqrt(oc["p1"] * oc["p2"] + 1e-30)
    assert np.max(np.abs(cg[m] - cw[m])) < 1e-4
    # Goldstein: GPU input through both modes against the oracle filter of the
    # oracle's own interferogram (stage and chain error together)
    want_f, wsum = GS.goldstein(oc["ifg"].astype(np.complex64), 32, 0.5)
    fm = (wsum > 0) & big
    for mode in ("fp32", "fp16_split"):
        f = _np(K.goldstein(ctx, ifg_t, 0.5, mode=mode))
        assert phase_rms(f, want_f, fm & (np.abs(want_f) > 1e-3 * np.abs(want_f).max())) < 1e-3
    # coherence: look cell and 5 x 5 against the oracle chain's products
    for win in (1, 5):
        coh = _np(K.coherence(ctx, ifg_t, pr_t, ps_t, win=win, fill="nan"))
        if win == 1:
            want = I.coherence_lookcell(oc["ifg"], np.sqrt(oc["p1"]), np.sqrt(oc["p2"]),
                                        fill="mask")
        else:
            want = I.coherence_window(oc["ifg"], oc["p1"], oc["p2"], 5)
        ok = np.isfinite(want) & np.isfinite(coh) & m
        assert ok.mean() > 0.3
        assert np.max(np.abs(coh[ok] - want[ok])) < 1e-3
        assert np.sqrt(np.mean((coh[ok] - want[ok]) ** 2)) < 1e-4


# ---------------------------------------------------------------------- C2
@pytest.fixture(scope="module")
def c2(ctx):
    """The C2 pair deramped on the GPU and its full-size fused products,
    with the pair kernel's tile classes."""
    from sentinel_gpu_processor_amd import kernels as K
    from sentinel_gpu_processor_amd import _native as N
    sc = c2_scene()
    gr, gs = _geoms(ctx, sc)
    dem = sc["dem"]
    dem_t = _dev(dem["data"])
    dref = K.deramp(ctx, gr, sc["ref_slc"])
    dsec = K.deramp(ctx, gs, sc["sec_slc"])
    stats = torch.zeros(N.TILE_STATS, dtype=torch.int64, device="cuda:0")
    ctx.set_tile_stats(stats)
    try:
        ifg, pr, ps = K.pair(ctx, gr, gs, sc["grid"], dem_t, dem, dref, dsec, "sinc", 1, 4,
                             offpoly=sc["offpoly"])
        st = _np(stats).tolist()
    finally:
        ctx.set_tile_stats(None)
    return dict(sc=sc, gr=gr, gs=gs, dem_t=dem_t, dref=dref, dsec=dsec, ifg=ifg, pr=pr, ps=ps,
                stats=st)


C2_CROP = (0, 0, 1500, 3072)  # rows, columns: the grid's west edge, outside corners included


def test_c2_tile_classes(c2):
    """At C2 the pair kernel meets every tile class: skipped (outside a
    burst), tap window staged in LDS, window beyond the LDS budget (global
    reads; only raw bursts shear such windows)."""
    tiles, skipped, staged, glob, empty, sheared = c2["stats"]
    print(f"c2 tiles {tiles}: burst-tiles skipped {skipped}, staged {staged}, global {glob}, "
          f"empty {empty}")
    grid = c2["sc"]["grid"]
    assert tiles == -(-grid["nrow"] // 16) * -(-(grid["ncol"] // 4) // 16)
    assert skipped + staged + glob + empty == 2 * tiles
    assert sheared == 0
    assert skipped > 0.2 * 2 * tiles
    assert staged > 0.05 * 2 * tiles
    assert glob > 0


def test_c2_crop_matches_oracle(ctx, c2):
    """The full-size kernel's products on a 1500 x 3072 crop (the first 48
    tile columns, where tiles are skipped, staged and read from global
    memory) against the oracle on that crop; the same crop as its own grid
    gives the same bits and its own tile classes include all three."""
    from sentinel_gpu_processor_amd import kernels as K
    from sentinel_gpu_processor_amd import _native as N
    sc = c2["sc"]
    r0, c0, nr, nc = C2_CROP
    g = crop_grid(sc["grid"], r0, c0, nr, nc)
    ifg = _np(c2["ifg"])[r0:r0 + nr, c0 // 4:(c0 + nc) // 4]
    pr = _np(c2["pr"])[r0:r0 + nr, c0 // 4:(c0 + nc) // 4]
    ps = _np(c2["ps"])[r0:r0 + nr, c0 // 4:(c0 + nc) // 4]
    stats = torch.zeros(N.TILE_STATS, dtype=torch.int64, device="cuda:0")
    ctx.set_tile_stats(stats)
    try:
        out = K.pair(ctx, c2["gr"], c2["gs"], g, c2["dem_t"], sc["dem"], c2["dref"], c2["dsec"],
                     "sinc", 1, 4, offpoly=sc["offpoly"])
        st = _np(stats).tolist()
    finally:
        ctx.set_tile_stats(None)
    for a, b in zip(out, (ifg, pr, ps)):
        assert np.array_equal(_np(a), b)
    assert st[1] > 0 and st[2] > 0 and st[3] > 0, st
    ref_h = _np(sc["ref_slc"])
    sec_h = _np(sc["sec_slc"])
    oi, o1, o2, _, _ = oracle_pair_crop(sc, ref_h, sec_h, g, "sinc")
    m = (o1 > 0) & (o2 > 0)
    assert m.mean() > 0.2
    # no skipped tile holds a pixel the oracle finds valid, and conversely
    assert np.count_nonzero((pr > 0) != (o1 > 0)) <= 8
    assert np.count_nonzero((ps > 0) != (o2 > 0)) <= 8
    big = m & (np.abs(oi) > 1e-3 * np.abs(oi).max())
    assert phase_rms(ifg, oi, big) < 1e-4
    cg = np.abs(ifg) / np.sqrt(pr * ps + 1e-30)
    cw = np.abs(oi) / np.sqrt(o1 * o2 + 1e-30)
    assert np.max(np.abs(cg[m] - cw[m])) < 1e-4
    # FP16-split Goldstein and 5 x 5 coherence of the crop
    f = _np(K.goldstein(ctx, _dev(ifg), 0.5, mode="fp16_split"))
    want_f, wsum = GS.goldstein(oi.astype(np.complex64), 32, 0.5)
    fm = (wsum > 0) & big & (np.abs(want_f) > 1e-3 * np.abs(want_f).max())
    assert phase_rms(f, want_f, fm) < 1e-3
    coh = _np(K.coherence(ctx, _dev(ifg), _dev(pr), _dev(ps), win=5, fill="nan"))
    want_c = I.coherence_window(oi, o1, o2, 5)
    ok = np.isfinite(want_c) & np.isfinite(coh) & m
    assert np.max(np.abs(coh[ok] - want_c[ok])) < 1e-3


def test_c2_full_size_properties(ctx, c2):
    """The whole 1500 x 25000 pair: finite products; a look cell has data
    exactly where all its pixels' sampling positions (GPU geo2rdr, the
    offsets checked against the oracle elsewhere) lie in both bursts' sinc
    validity windows -- up to cells on the window's edge; 5 x 5 coherence
    in [0, 1]."""
    from sentinel_gpu_processor_amd import kernels as K
    sc = c2["sc"]
    grid = sc["grid"]
    pr, ps = c2["pr"], c2["ps"]
    assert bool(torch.isfinite(torch.view_as_real(c2["ifg"])).all())
    coh = K.coherence(ctx, c2["ifg"], pr, ps, win=5)
    assert float(coh.min()) >= 0.0 and float(coh.max()) <= 1.0
    nrow, ncol = grid["nrow"], grid["ncol"]
    cells = torch.ones((nrow, ncol // 4), dtype=torch.bool, device="cuda:0")
    for g, b, op in ((c2["gr"], sc["burst_ref"], None), (c2["gs"], sc["burst_sec"], sc["offpoly"])):
        rg, az, _ = K.geo2rdr(ctx, g, grid, c2["dem_t"], sc["dem"], offpoly=op)
        v = ((rg >= b["first_valid_sample"] + 3) & (rg < b["last_valid_sample"] - 4)
             & (az >= b["first_valid_line"] + 5) & (az < b["last_valid_line"] - 5))
        del rg, az
        cells &= v[:, :ncol // 4 * 4].reshape(nrow, ncol // 4, 4).all(dim=2)
        del v
    has = (pr > 0) & (ps > 0)
    assert float(cells.float().mean()) > 0.3
    # cells with every pixel valid have data; cells with data have at least
    # one valid pixel in each burst (a partly valid cell is the edge case)
    assert bool(has[cells].all())
    assert int((has & ~cells).sum()) < 0.01 * int(cells.sum())


def test_c2_full_size_physics(ctx):
    """The whole 1500 x 25000 pair recovers what the generator put in.  With
    a zero baseline the two bursts see every point at the same range, so the
    interferogram phase is exactly the residual of the two bursts' reramps
    at their sampling positions (the secondary is misregistered), in closed
    form from the straight-line orbits; checked on blocks of 5 x 5 sampled
    look cells over the full 1500 x 6250 look grid.  (With a baseline the
    fast generator's topographic phase is sampled on a coarse grid, too
    coarse for the fractal DEM's fringes.)"""
    from sentinel_gpu_processor_amd import kernels as K
    sc = c2_scene(baseline=(0.0, 0.0))
    gr, gs = _geoms(ctx, sc)
    dem_t = _dev(sc["dem"]["data"])
    dref = K.deramp(ctx, gr, sc["ref_slc"])
    dsec = K.deramp(ctx, gs, sc["sec_slc"])
    ifg, pr, ps = K.pair(ctx, gr, gs, sc["grid"], dem_t, sc["dem"], dref, dsec, "sinc", 1, 4,
                         offpoly=sc["offpoly"])
    del dref, dsec
    grid = sc["grid"]
    ifg = _np(ifg)
    has = (_np(pr) > 0) & (_np(ps) > 0)
    # prediction at the centre of every 5th look cell in both directions
    rs = np.arange(2, grid["nrow"], 5)
    cs = np.arange(2, grid["ncol"] // 4, 5)
    sub = dict(grid, nrow=len(rs), ncol=len(cs), lat0=grid["lat0"] + rs[0] * grid["dlat"],
               dlat=5 * grid["dlat"], lon0=grid["lon0"] + (4 * cs[0] + 1.5) * grid["dlon"],
               dlon=20 * grid["dlon"])
    lr, ls = sc["line_orbits"]
    ph_r, _, _ = closed_form_phase(lr, sc["burst_ref"], sub, sc["dem"])
    ph_s, _, _ = closed_form_phase(ls, sc["burst_sec"], sub, sc["dem"], sc["offpoly"])
    resid = ph_s - ph_r  # the flattening terms are equal: zero baseline
    mk = has[np.ix_(rs, cs)]
    B = 10  # blocks of 10 x 10 sampled look cells (50 x 50 look cells apart)
    nb_r, nb_c = mk.shape[0] // B, mk.shape[1] // B

    def blocks(x):
        return x[:nb_r * B, :nb_c * B].reshape(nb_r, B, nb_c, B).sum(axis=(1, 3))

    cnt = blocks(mk.astype(np.int64))
    sel = cnt == B * B
    assert sel.sum() > 300
    ang = np.abs(np.angle(blocks(ifg[np.ix_(rs, cs)] * np.exp(-1j * resid) * mk)[sel]))
    print(f"c2 physics: {sel.sum()} blocks, median {np.median(ang):.4f} rad, "
          f"99% {np.quantile(ang, 0.99):.4f} rad")
    assert np.median(ang) < 0.04, np.median(ang)
    assert np.quantile(ang, 0.99) < 0.2, np.quantile(ang, 0.99)
    assert np.ptp(resid[mk]) > 1.0  # the residual the check discriminates
    # without the residual (a reramp that cancelled between the bursts) the
    # check fails
    ang0 = np.abs(np.angle(blocks(ifg[np.ix_(rs, cs)] * mk)[sel]))
    assert np.quantile(ang0, 0.99) > 0.5


def test_c2_raw_pair_matches_deramped(ctx, c2):
    """The raw-burst pair (samples deramped as staged; a window beyond the
    LDS budget sheared along the tile's skew and staged when that fits, its
    samples deramped per burst column; per tap on the tiles still read from
    global memory) over the whole C2 pair against deramp + the pair kernel,
    and against the oracle on the crop.  Most of the windows the deramped
    kernel reads from global memory fit once sheared."""
    from sentinel_gpu_processor_amd import kernels as K
    from sentinel_gpu_processor_amd import _native as N
    sc = c2["sc"]
    stats = torch.zeros(N.TILE_STATS, dtype=torch.int64, device="cuda:0")
    ctx.set_tile_stats(stats)
    try:
        ifg, pr, ps = K.pair(ctx, c2["gr"], c2["gs"], sc["grid"], c2["dem_t"], sc["dem"],
                             sc["ref_slc"], sc["sec_slc"], "sinc", 1, 4, offpoly=sc["offpoly"],
                             raw=True)
        st = _np(stats).tolist()
    finally:
        ctx.set_tile_stats(None)
    tiles, skipped, staged, glob, empty, sheared = st
    print(f"c2 raw tiles {tiles}: burst-tiles skipped {skipped}, staged {staged}, "
          f"sheared {sheared}, global {glob}, empty {empty}")
    assert skipped + staged + sheared + glob + empty == 2 * tiles
    assert (skipped, staged + sheared + glob, empty) == (
        c2["stats"][1], c2["stats"][2] + c2["stats"][3], c2["stats"][4])
    assert sheared > 0 and glob < 0.2 * c2["stats"][3]
    d = (ifg - c2["ifg"]).abs().max().item()
    assert d <= 2e-5 * c2["ifg"].abs().max().item()
    assert bool(((pr > 0) == (c2["pr"] > 0)).all())
    r0, c0, nr, nc = C2_CROP
    g = crop_grid(sc["grid"], r0, c0, nr, nc)
    oi, o1, o2, _, _ = oracle_pair_crop(sc, _np(sc["ref_slc"]), _np(sc["sec_slc"]), g, "sinc")
    m = (o1 > 0) & (o2 > 0) & (np.abs(oi) > 1e-3 * np.abs(oi).max())
    got = _np(ifg)[r0:r0 + nr, c0 // 4:(c0 + nc) // 4]
    assert phase_rms(got, oi, m) < 1e-4
