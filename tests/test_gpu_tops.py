"""GPU parity with non-degenerate TOPS metadata: range-dependent FM rate and
Doppler centroid, records that change with time, the burst centre between
records (geo2rdr.cu:465-549 set-up; the deramp, geo2rdr.cu:138-158; the
reramp, geo2rdr.cu:210-219), checked against the FP64 oracle and against
closed forms the oracle does not enter."""

import numpy as np
import pytest

from oracle import geom as G
from sentinel_gpu_processor_amd import synthetic as S
from tests.scenes import oracle_chain, phase_rms, scene
from tests.test_oracle_tops import closed_form_phase, ramp_burst

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


@pytest.mark.parametrize("shift", [0.0, -20.0, 20.0])
def test_geom_setup_varying(ctx, shift):
    """Set-up scalars and the per-sample kt / etaref vectors with the centre
    between, after and before the records."""
    from sentinel_gpu_processor_amd import _native as N
    sc = scene("small_tops")
    b = dict(sc["burst_ref"])
    b["fm"], b["dc"] = S.tops_records(43200.0, b["fm"][1, 1], "varying", shift=shift)
    g = N.Geometry(ctx, sc["orbit_ref"], b)
    q = g.query()
    su = G.deramp_setup(sc["orbit_ref"], b)
    assert q["ks"] == pytest.approx(su["ks"], rel=1e-12)
    assert np.allclose(q["fm"], su["fm"], rtol=1e-14, atol=0)
    assert np.allclose(q["dc"], su["dc"], rtol=1e-14, atol=0)
    assert q["etac0"] == pytest.approx(su["etac0"], rel=1e-12)
    kt, er = g.vectors()
    assert np.max(np.abs(kt - su["kt"])) <= 1e-12 * np.abs(su["kt"]).max()
    assert np.max(np.abs(er - su["etaref"])) <= 1e-15
    assert np.ptp(kt) > 0.5 and np.ptp(er) > 1e-4


def test_deramp_varying(ctx):
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("small_tops")
    gr, gs = _geoms(ctx, sc)
    oc = oracle_chain("small_tops")
    for g, slc, want in ((gr, sc["ref_slc"], oc["dr"]), (gs, sc["sec_slc"], oc["ds"])):
        d = _np(K.deramp(ctx, g, _dev(slc)))
        assert np.max(np.abs(d - want)) / np.max(np.abs(want)) < 2e-6


@pytest.mark.parametrize("interp", ["bilinear", "sinc"])
def test_geocode_varying(ctx, interp):
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("small_tops")
    gr, gs = _geoms(ctx, sc)
    oc = oracle_chain("small_tops", interp)
    dem = sc["dem"]
    dem_t = _dev(dem["data"])
    dref = K.deramp(ctx, gr, _dev(sc["ref_slc"]))
    dsec = K.deramp(ctx, gs, _dev(sc["sec_slc"]))
    o_r = _np(K.geocode(ctx, gr, sc["grid"], dem_t, dem, dref, interp))
    o_s = _np(K.geocode(ctx, gs, sc["grid"], dem_t, dem, dsec, interp, offpoly=sc["offpoly"]))
    for got, want in ((o_r, oc["gr"]), (o_s, oc["gs"])):
        vg, vw = got != 0, want != 0
        assert np.count_nonzero(vg != vw) <= 2
        both = vg & vw
        assert both.mean() > 0.3
        err = np.abs(got[both] - want[both]) / np.abs(want[both]).max()
        assert err.max() < 2e-5
        assert phase_rms(got, want, both & (np.abs(want) > 1e-2)) < 1e-4


@pytest.mark.parametrize("interp", ["bilinear", "sinc"])
def test_pair_varying(ctx, interp):
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("small_tops")
    gr, gs = _geoms(ctx, sc)
    oc = oracle_chain("small_tops", interp)
    dem = sc["dem"]
    dem_t = _dev(dem["data"])
    dref = K.deramp(ctx, gr, _dev(sc["ref_slc"]))
    dsec = K.deramp(ctx, gs, _dev(sc["sec_slc"]))
    ifg, pr, ps = K.pair(ctx, gr, gs, sc["grid"], dem_t, dem, dref, dsec, interp, 1, 4,
                         offpoly=sc["offpoly"])
    ifg, pr, ps = _np(ifg), _np(pr), _np(ps)
    m = (oc["p1"] > 0) & (oc["p2"] > 0) & (np.abs(oc["ifg"]) > 1e-3 * np.abs(oc["ifg"]).max())
    assert m.mean() > 0.2
    assert phase_rms(ifg, oc["ifg"], m) < 1e-4
    assert np.count_nonzero((pr > 0) != (oc["p1"] > 0)) <= 2
    assert np.count_nonzero((ps > 0) != (oc["p2"] > 0)) <= 2


@pytest.mark.parametrize("interp", ["bilinear", "sinc"])
@pytest.mark.parametrize("which", ["ref", "sec"])
def test_reramp_closed_form_gpu(ctx, interp, which):
    """A burst holding only its TOPS ramp, deramped and geocoded on the GPU
    over 1024 lines (the ramps span +-1.05 s), against the closed form
    exp(i (4 pi R / lambda + pi kt (eta - etaref)^2)); the secondary with
    its offset polynomial."""
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    sc = S.make_scene(lines=1024, samples=512, dem_shape=(96, 64), hmax=300.0, seed=0,
                      misreg=(0.3, 0.05), misreg_ramp=(0.2, 0.1), tops="varying",
                      fidelity="none")
    k = 0 if which == "ref" else 1
    orbit = (sc["orbit_ref"], sc["orbit_sec"])[k]
    lorb = sc["line_orbits"][k]
    b = (sc["burst_ref"], sc["burst_sec"])[k]
    op = sc["offpoly"] if k else None
    g = N.Geometry(ctx, orbit, b)
    dz = K.deramp(ctx, g, _dev(ramp_burst(lorb, b)))
    dem = sc["dem"]
    out = _np(K.geocode(ctx, g, sc["grid"], _dev(dem["data"]), dem, dz, interp, offpoly=op))
    ph, _, _ = closed_form_phase(lorb, b, sc["grid"], dem, op)
    v = out != 0
    assert v.mean() > 0.3
    assert np.allclose(np.abs(out[v]), 1, atol=1e-5)
    d = np.angle(out[v] * np.exp(-1j * ph[v]))
    assert np.max(np.abs(d)) < 2e-4
    assert np.sqrt(np.mean(d * d)) < 3e-5


def test_long_misreg_physics(ctx):
    """1024 misregistered lines with varying TOPS rates: the interferogram
    phase equals the injected topographic phase plus the residual of the two
    bursts' reramps at their sampling positions, both from closed forms
    (straight-line zero Doppler, the generator's rates).  That residual
    spans more than a radian here, so a reramp of the wrong sign or time
    origin fails."""
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("long_tops")
    gr, gs = _geoms(ctx, sc)
    dem = sc["dem"]
    dem_t = _dev(dem["data"])
    dref = K.deramp(ctx, gr, _dev(sc["ref_slc"]))
    dsec = K.deramp(ctx, gs, _dev(sc["sec_slc"]))
    ifg, pr, ps = K.pair(ctx, gr, gs, sc["grid"], dem_t, dem, dref, dsec, "sinc", 1, 4,
                         offpoly=sc["offpoly"])
    ifg, pr, ps = _np(ifg), _np(pr), _np(ps)
    grid = sc["grid"]
    lr, ls = sc["line_orbits"]
    ph_r, _, _ = closed_form_phase(lr, sc["burst_ref"], grid, dem)
    ph_s, _, _ = closed_form_phase(ls, sc["burst_sec"], grid, dem, sc["offpoly"])
    r = np.arange(grid["nrow"], dtype=np.float64)[:, None]
    c = np.arange(grid["ncol"], dtype=np.float64)[None, :]
    lat = grid["lat0"] + r * grid["dlat"] + 0 * c
    lon = grid["lon0"] + c * grid["dlon"] + 0 * r
    X = S.llh2xyz(lat, lon, S.dem_height(dem, lat, lon))
    _, Rr = lr.zero_doppler(X)
    _, Rs = ls.zero_doppler(X)
    lam = sc["burst_ref"]["wavelength"]
    resid = (ph_s - 4 * np.pi * Rs / lam) - (ph_r - 4 * np.pi * Rr / lam)
    topo = 2 * np.pi * S.dem_height(dem, lat, lon) / sc["amb_height"]
    nc = ifg.shape[1]

    def looks(ph):
        return np.exp(1j * ph)[:, :nc * 4].reshape(grid["nrow"], nc, 4).mean(axis=2)

    m = (pr > 0) & (ps > 0)
    assert m.mean() > 0.3
    for pred, should_pass in ((topo + resid, True), (topo, False), (topo - resid, False)):
        e = ifg * np.conj(looks(pred)) * m
        # mean residual angle over blocks of 32 x 32 look cells
        nb_r, nb_c = ifg.shape[0] // 32, nc // 32
        blk = e[:nb_r * 32, :nb_c * 32].reshape(nb_r, 32, nb_c, 32).sum(axis=(1, 3))
        cnt = m[:nb_r * 32, :nb_c * 32].reshape(nb_r, 32, nb_c, 32).sum(axis=(1, 3))
        ang = np.abs(np.angle(blk[cnt > 512]))
        if should_pass:
            assert ang.max() < 0.25, ang.max()
        else:
            assert ang.max() > 0.5, ang.max()
    valid = m[:, :, None].repeat(4, axis=2).reshape(grid["nrow"], nc * 4)
    assert np.ptp(resid[:, :nc * 4][valid]) > 1.0


@pytest.mark.parametrize("interp", ["bilinear", "sinc"])
@pytest.mark.parametrize("name", ["small_misreg", "small_tops"])
def test_pair_raw_matches_oracle(ctx, interp, name):
    """s1ifg_pair_raw (raw bursts, deramped as they are read) against the
    oracle chain and against deramp + s1ifg_pair."""
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene(name)
    gr, gs = _geoms(ctx, sc)
    oc = oracle_chain(name, interp)
    dem = sc["dem"]
    dem_t = _dev(dem["data"])
    ref, sec = _dev(sc["ref_slc"]), _dev(sc["sec_slc"])
    ifg, pr, ps = (_np(t) for t in K.pair(ctx, gr, gs, sc["grid"], dem_t, dem, ref, sec, interp,
                                          1, 4, offpoly=sc["offpoly"], raw=True))
    m = (oc["p1"] > 0) & (oc["p2"] > 0) & (np.abs(oc["ifg"]) > 1e-3 * np.abs(oc["ifg"]).max())
    assert m.mean() > 0.2
    assert phase_rms(ifg, oc["ifg"], m) < 1e-4
    assert np.count_nonzero((pr > 0) != (oc["p1"] > 0)) <= 2
    dref, dsec = K.deramp(ctx, gr, ref), K.deramp(ctx, gs, sec)
    i2, p2, s2 = (_np(t) for t in K.pair(ctx, gr, gs, sc["grid"], dem_t, dem, dref, dsec, interp,
                                         1, 4, offpoly=sc["offpoly"]))
    assert np.max(np.abs(ifg - i2)) <= 2e-5 * np.abs(i2).max()
    assert np.max(np.abs(pr - p2)) <= 2e-5 * p2.max()


@pytest.mark.parametrize("coarse", [(8, 4), (2, 1)])
@pytest.mark.parametrize("raw", [False, True])
def test_pair_coarse_grid_global_taps(ctx, raw, coarse):
    """A target grid 8 x 4 times coarser than the bursts' sampling: one
    wavefront's rows of a tile span a tap window beyond the LDS budget, so
    the taps are read from global memory (deramped per tap when raw).  2 x 1
    times coarser: on raw bursts some of those windows fit the budget once
    sheared along the tile's skew (staged sheared, deramped per burst column
    as staged).  The products still match the oracle on that grid."""
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    from tests.configs import oracle_pair_crop
    sc = scene("small_misreg")
    fr, fc = coarse
    g = dict(sc["grid"], nrow=sc["grid"]["nrow"] // fr, ncol=sc["grid"]["ncol"] // fc,
             dlat=sc["grid"]["dlat"] * fr, dlon=sc["grid"]["dlon"] * fc)
    gr, gs = _geoms(ctx, sc)
    dem = sc["dem"]
    ref, sec = _dev(sc["ref_slc"]), _dev(sc["sec_slc"])
    if not raw:
        ref, sec = K.deramp(ctx, gr, ref), K.deramp(ctx, gs, sec)
    stats = torch.zeros(N.TILE_STATS, dtype=torch.int64, device="cuda:0")
    ctx.set_tile_stats(stats)
    try:
        out = K.pair(ctx, gr, gs, g, _dev(dem["data"]), dem, ref, sec, "sinc", 1, 4,
                     offpoly=sc["offpoly"], raw=raw)
        st = _np(stats).tolist()
    finally:
        ctx.set_tile_stats(None)
    assert st[3] > 0, st
    if raw and coarse == (2, 1):
        assert st[5] > 0, st
    if not raw:
        assert st[5] == 0, st
    ifg, pr, ps = (_np(t) for t in out)
    oi, o1, o2, _, _ = oracle_pair_crop(sc, sc["ref_slc"], sc["sec_slc"], g, "sinc")
    m = (o1 > 0) & (o2 > 0) & (np.abs(oi) > 1e-3 * np.abs(oi).max())
    assert m.mean() > 0.2
    assert phase_rms(ifg, oi, m) < 1e-4
    assert np.count_nonzero((pr > 0) != (o1 > 0)) <= 2
