"""GPU parity: every libs1ifg stage against the FP64 oracle on the same
synthetic inputs (called through the C ABI via ctypes)."""

import numpy as np
import pytest

from oracle import geom as G
from oracle import goldstein as GS
from oracle import ifg as I
from tests.scenes import oracle_chain, phase_rms, scene

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from sentinel_gpu_processor_amd import _native as N
    c = N.Context(0)
    yield c
    c.sync()
    c.close()


def _geoms(ctx, sc):
    from sentinel_gpu_processor_amd import _native as N
    return (N.Geometry(ctx, sc["orbit_ref"], sc["burst_ref"]),
            N.Geometry(ctx, sc["orbit_sec"], sc["burst_sec"]))


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_geom_setup_matches_oracle(ctx):
    sc = scene("small")
    gr, _ = _geoms(ctx, sc)
    q = gr.query()
    su = G.deramp_setup(sc["orbit_ref"], sc["burst_ref"])
    assert q["ks"] == pytest.approx(su["ks"], rel=1e-12)
    assert np.allclose(q["fm"], su["fm"], rtol=1e-14, atol=0)
    assert np.allclose(q["dc"], su["dc"], rtol=1e-14, atol=0)
    assert q["etac0"] == pytest.approx(su["etac0"], rel=1e-12, abs=1e-15)


def test_orbit_eval_matches_hermite(ctx):
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("small")
    gr, _ = _geoms(ctx, sc)
    t_orb, pos, vel = sc["orbit_ref"]
    tq = np.linspace(t_orb[0] - 0.5, t_orb[-1] + 0.5, 257)
    p, v, a = K.orbit_eval(ctx, gr, _dev(tq))
    xo, vo = G.intp_orbit(t_orb, pos, vel, tq)
    assert np.max(np.abs(_np(p) - xo)) < 1e-6
    # polynomial vs Lagrange-form evaluation: rounding only (m/s)
    assert np.max(np.abs(_np(v) - vo)) < 1e-6


def test_deramp_matches_oracle(ctx):
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("small")
    gr, _ = _geoms(ctx, sc)
    oc = oracle_chain("small")
    d = _np(K.deramp(ctx, gr, _dev(sc["ref_slc"])))
    ref = oc["dr"]
    assert np.max(np.abs(d - ref)) / np.max(np.abs(ref)) < 2e-6
    # in place gives the same bits as out of place
    t = _dev(sc["ref_slc"])
    K.deramp(ctx, gr, t, out=t)
    assert np.array_equal(_np(t), d)


def test_deramp_odd_width(ctx):
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("odd")
    gr, _ = _geoms(ctx, sc)
    su = G.deramp_setup(sc["orbit_ref"], sc["burst_ref"])
    d = _np(K.deramp(ctx, gr, _dev(sc["ref_slc"])))
    ref = G.deramp(sc["ref_slc"], su)
    assert np.max(np.abs(d - ref)) / np.max(np.abs(ref)) < 2e-6


@pytest.mark.parametrize("name", ["small", "small_misreg"])
def test_geo2rdr_offsets(ctx, name):
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene(name)
    gr, gs = _geoms(ctx, sc)
    oc = oracle_chain(name)
    dem = sc["dem"]
    dem_t = _dev(dem["data"])
    rg, az, rng = K.geo2rdr(ctx, gr, sc["grid"], dem_t, dem)
    assert np.max(np.abs(_np(rg) - oc["geo_r"]["rgoff"])) < 2e-6
    assert np.max(np.abs(_np(az) - oc["geo_r"]["azoff"])) < 2e-6
    assert np.max(np.abs(_np(rng) - oc["geo_r"]["rng"])) < 5e-6
    rg, az, rng = K.geo2rdr(ctx, gs, sc["grid"], dem_t, dem, offpoly=sc["offpoly"])
    drg, daz = G.offset_poly_shift(sc["offpoly"], oc["geo_s"]["rgoff"], oc["geo_s"]["azoff"],
                                   sc["burst_sec"])
    assert np.max(np.abs(_np(rg) - (oc["geo_s"]["rgoff"] + drg))) < 2e-6
    assert np.max(np.abs(_np(az) - (oc["geo_s"]["azoff"] + daz))) < 2e-6


def test_geo2rdr_curved_orbit(ctx):
    """An orbit bent by quadratic to quartic terms: its interval polynomials
    have higher-order terms, so the Newton iteration's orbit node table
    holds more nodes per interval.  Offsets equal the oracle's (Hermite
    evaluation at every step) as on the straight orbits."""
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("small")
    t, pos, vel = sc["orbit_ref"]
    e = t - t.mean()
    A = np.array([3.5, -2.0, -2.5])     # m/s^2 / 2
    B = np.array([2e-3, 1e-3, -3e-3])   # m/s^3 / 6
    C = np.array([1e-4, -2e-4, 1e-4])   # m/s^4 / 24
    p2 = pos + A * e[:, None] ** 2 + B * e[:, None] ** 3 + C * e[:, None] ** 4
    v2 = vel + 2 * A * e[:, None] + 3 * B * e[:, None] ** 2 + 4 * C * e[:, None] ** 3
    orb = (t, p2, v2)
    g = N.Geometry(ctx, orb, sc["burst_ref"])
    dem = sc["dem"]
    rg, az, rng = K.geo2rdr(ctx, g, sc["grid"], _dev(dem["data"]), dem)
    o = G.geo2rdr_grid(orb, sc["burst_ref"], sc["grid"], G.target_heights(sc["grid"], dem))
    assert np.max(np.abs(_np(rg) - o["rgoff"])) < 2e-6
    assert np.max(np.abs(_np(az) - o["azoff"])) < 2e-6
    assert np.max(np.abs(_np(rng) - o["rng"])) < 5e-6


def _bent(orbit, A, B, C):
    t, pos, vel = orbit
    e = t - t.mean()
    p2 = pos + A * e[:, None] ** 2 + B * e[:, None] ** 3 + C * e[:, None] ** 4
    v2 = vel + 2 * A * e[:, None] + 3 * B * e[:, None] ** 2 + 4 * C * e[:, None] ** 3
    return (t, p2, v2)


@pytest.mark.parametrize("raw", [False, True])
def test_pair_curved_orbits(ctx, raw):
    """The fused pair kernel on bent orbits (about 9 m/s^2 of acceleration,
    as a circular orbit's, plus cubic and quartic terms): each look cell's
    first pixel starts its Newton solve at the time the tile's corner
    solutions predict, the other pixels at the first pixel's root; the
    products equal the oracle's, whose Newton iteration starts at the burst
    mid-time."""
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    from tests.configs import oracle_pair_crop
    sc = dict(scene("small_misreg"))
    sc["orbit_ref"] = _bent(sc["orbit_ref"], np.array([3.5, -2.0, -2.5]),
                            np.array([2e-3, 1e-3, -3e-3]), np.array([1e-4, -2e-4, 1e-4]))
    sc["orbit_sec"] = _bent(sc["orbit_sec"], np.array([3.4, -2.1, -2.5]),
                            np.array([-1e-3, 2e-3, -2e-3]), np.array([-1e-4, 1e-4, 2e-4]))
    gr = N.Geometry(ctx, sc["orbit_ref"], sc["burst_ref"])
    gs = N.Geometry(ctx, sc["orbit_sec"], sc["burst_sec"])
    dem = sc["dem"]
    ref, sec = _dev(sc["ref_slc"]), _dev(sc["sec_slc"])
    if not raw:
        ref, sec = K.deramp(ctx, gr, ref), K.deramp(ctx, gs, sec)
    ifg, pr, ps = (_np(t) for t in K.pair(ctx, gr, gs, sc["grid"], _dev(dem["data"]), dem, ref,
                                         sec, "sinc", 1, 4, offpoly=sc["offpoly"], raw=raw))
    oi, o1, o2, _, _ = oracle_pair_crop(sc, sc["ref_slc"], sc["sec_slc"], sc["grid"], "sinc")
    m = (o1 > 0) & (o2 > 0) & (np.abs(oi) > 1e-3 * np.abs(oi).max())
    assert m.mean() > 0.2
    assert phase_rms(ifg, oi, m) < 1e-4
    assert np.max(np.abs(pr[m] - o1[m]) / o1[m]) < 1e-4
    assert np.count_nonzero((pr > 0) != (o1 > 0)) <= 2
    assert np.count_nonzero((ps > 0) != (o2 > 0)) <= 2


@pytest.mark.parametrize("raw", [False, True])
@pytest.mark.parametrize("interp,rl,cl", [("sinc", 2, 2), ("sinc", 1, 1), ("sinc", 4, 1),
                                          ("sinc", 1, 3), ("sinc", 2, 3), ("bilinear", 2, 2)])
def test_pair_look_shapes(ctx, raw, interp, rl, cl):
    """Look shapes other than 1 x 4: the tile kernel with the look shape
    at run time (rl * cl <= 4, sinc), the per-pixel kernel (rl * cl > 4,
    and bilinear), deramped and raw bursts, against the oracle's looks."""
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    from tests.configs import oracle_pair_crop
    sc = scene("small_misreg")
    gr, gs = _geoms(ctx, sc)
    dem = sc["dem"]
    ref, sec = _dev(sc["ref_slc"]), _dev(sc["sec_slc"])
    if not raw:
        ref, sec = K.deramp(ctx, gr, ref), K.deramp(ctx, gs, sec)
    print(f"pair looks {rl} x {cl} {interp} raw={raw}", flush=True)
    ifg, pr, ps = (_np(t) for t in K.pair(ctx, gr, gs, sc["grid"], _dev(dem["data"]), dem, ref,
                                         sec, interp, rl, cl, offpoly=sc["offpoly"], raw=raw))
    oi, o1, o2, _, _ = oracle_pair_crop(sc, sc["ref_slc"], sc["sec_slc"], sc["grid"], interp,
                                        rl, cl)
    assert ifg.shape == oi.shape
    m = (o1 > 0) & (o2 > 0) & (np.abs(oi) > 1e-3 * np.abs(oi).max())
    assert m.mean() > 0.2
    assert phase_rms(ifg, oi, m) < 1e-4
    assert np.max(np.abs(pr[m] - o1[m]) / o1[m]) < 1e-4
    assert np.count_nonzero((pr > 0) != (o1 > 0)) <= 2 * rl * cl
    assert np.count_nonzero((ps > 0) != (o2 > 0)) <= 2 * rl * cl


@pytest.mark.parametrize("interp", ["bilinear", "sinc"])
@pytest.mark.parametrize("name", ["small", "small_misreg"])
def test_geocode_matches_oracle(ctx, interp, name):
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene(name)
    gr, gs = _geoms(ctx, sc)
    oc = oracle_chain(name, interp)
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


def test_valid_rows(ctx):
    from sentinel_gpu_processor_amd import kernels as K
    img = np.zeros((40, 33), np.complex64)
    assert K.valid_rows(ctx, _dev(img)) == (0, 0)
    img[7, 32] = 1e-30
    img[22, 0] = 1j
    assert K.valid_rows(ctx, _dev(img)) == G.nonzero_rows(img) == (7, 23)


@pytest.mark.parametrize("rl,cl", [(1, 4), (3, 5), (1, 1), (2, 1)])
def test_crossmul_looks(ctx, rl, cl):
    from sentinel_gpu_processor_amd import kernels as K
    rng = np.random.default_rng(3)
    a = (rng.standard_normal((61, 203)) + 1j * rng.standard_normal((61, 203))).astype(np.complex64)
    b = (rng.standard_normal((61, 203)) + 1j * rng.standard_normal((61, 203))).astype(np.complex64)
    ifg, pr, ps = K.crossmul_looks(ctx, _dev(a), _dev(b), rl, cl)
    want = I.cpx_looks(I.conj_mul(a, b), rl, cl)
    assert np.max(np.abs(_np(ifg) - want)) < 1e-5 * np.abs(want).max() + 1e-6
    assert np.allclose(_np(pr), I.cpx_looks(np.abs(a) ** 2, rl, cl).real, rtol=1e-5, atol=1e-6)
    assert np.allclose(_np(ps), I.cpx_looks(np.abs(b) ** 2, rl, cl).real, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("rl,cl", [(1, 4), (2, 3), (1, 1)])
def test_amplitude(ctx, rl, cl):
    from sentinel_gpu_processor_amd import kernels as K
    rng = np.random.default_rng(4)
    a = (rng.standard_normal((40, 91)) + 1j * rng.standard_normal((40, 91))).astype(np.complex64)
    amp = _np(K.amplitude(ctx, _dev(a), rl, cl))
    assert np.allclose(amp, I.amplitude(a, rl, cl), rtol=2e-6, atol=1e-7)


@pytest.mark.parametrize("win", [1, 5])
def test_coherence(ctx, win):
    from sentinel_gpu_processor_amd import kernels as K
    oc = oracle_chain("small")
    ifg = oc["ifg"].astype(np.complex64)
    p1 = oc["p1"].astype(np.float32)
    p2 = oc["p2"].astype(np.float32)
    coh = _np(K.coherence(ctx, _dev(ifg), _dev(p1), _dev(p2), win=win, fill="nan"))
    if win == 1:
        want = I.coherence_lookcell(ifg.astype(np.complex128), np.sqrt(p1.astype(np.float64)),
                                    np.sqrt(p2.astype(np.float64)), fill="mask")
    else:
        want = I.coherence_window(ifg, p1, p2, win)
    ok = np.isfinite(want) & np.isfinite(coh)
    assert ok.mean() > 0.2
    assert np.max(np.abs(coh[ok] - want[ok])) < 2e-5


# tolerances: FP32 FFT path ~1e-6 relative; FP16 hi/lo split DFT-GEMM path keeps
# ~22 bits per operand and applies the smoothing table through its two
# leading eigen-terms (residual 1.3e-6 of the table maximum)
GOLD_TOL = {"fp32": (5e-5, 1e-4), "fp16_split": (2e-4, 3e-4)}


@pytest.mark.parametrize("mode", ["fp32", "fp16_split"])
@pytest.mark.parametrize("shape", [(64, 96), (77, 130), (32, 32), (70, 605)])
def test_goldstein(ctx, shape, mode):
    from sentinel_gpu_processor_amd import kernels as K
    rng = np.random.default_rng(5)
    yy, xx = np.mgrid[0:shape[0], 0:shape[1]]
    x = (np.exp(1j * (0.05 * xx + 0.02 * yy * yy / shape[0])) * (1 + 0.3 * rng.standard_normal(shape))
         + 0.4 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))).astype(np.complex64)
    got = _np(K.goldstein(ctx, _dev(x), alpha=0.5, mode=mode))
    want, wsum = GS.goldstein(x, 32, 0.5)
    m = wsum > 0
    assert np.array_equal(got[~m], np.zeros(np.count_nonzero(~m), np.complex64))
    err = np.abs(got[m] - want[m]) / np.abs(want[m]).max()
    tol_rel, tol_ph = GOLD_TOL[mode]
    assert err.max() < tol_rel
    assert phase_rms(got, want, m) < tol_ph


@pytest.mark.parametrize("mode", ["fp32", "fp16_split"])
def test_goldstein_scale_and_zero_patches(ctx, mode):
    """Large dynamic range (1e6) and an all-zero region: the per-patch power
    of two scaling and the mean <= 1e-8 branch (goldstein.cu:170-175)."""
    from sentinel_gpu_processor_amd import kernels as K
    rng = np.random.default_rng(8)
    x = (rng.standard_normal((96, 128)) + 1j * rng.standard_normal((96, 128))).astype(np.complex64)
    x[:, :40] *= 1e6
    x[50:, 70:] = 0
    got = _np(K.goldstein(ctx, _dev(x), alpha=0.7, mode=mode))
    want, wsum = GS.goldstein(x, 32, 0.7)
    m = wsum > 0
    tol_rel, tol_ph = GOLD_TOL[mode]
    for sl in (np.s_[:, :8], np.s_[:, 72:]):
        g, w, mm = got[sl], want[sl], m[sl]
        big = mm & (np.abs(w) > 1e-3 * np.abs(w[mm]).max())
        assert np.max(np.abs(g[mm] - w[mm])) < tol_rel * np.abs(w[mm]).max()
        assert phase_rms(g, w, big) < tol_ph
    assert np.all(want[82:, 104:] == 0)
    assert np.all(got[82:, 104:] == 0)


def test_goldstein_zero_patches_passed_over(ctx):
    """The FP16-split pass skips patches whose components are all +0 or -0
    (no-data regions, about half of a burst's look grid).  Against the FP32
    path, which filters every patch: the same zero / non-finite pattern
    (a NaN or an infinity inside an otherwise zero region still spreads over
    its patches), values within the split tolerance elsewhere; and against
    the oracle where the oracle is finite."""
    from sentinel_gpu_processor_amd import kernels as K
    rng = np.random.default_rng(11)
    x = (rng.standard_normal((160, 224)) + 1j * rng.standard_normal((160, 224))).astype(np.complex64)
    x[:72, :96] = 0                       # zero corner: every patch inside it is passed over
    x[:72, :40] = -0.0 + 0j               # negative zeros are zeros too
    x[100:, 150:] = 0
    x[120, 200] = np.float32(np.nan)      # a NaN in a zero region
    x[20, 20] = complex(np.inf, 0.0)      # an infinity in the zero corner
    got = _np(K.goldstein(ctx, _dev(x), alpha=0.5, mode="fp16_split"))
    f32 = _np(K.goldstein(ctx, _dev(x), alpha=0.5, mode="fp32"))
    fin_g = np.isfinite(got.real) & np.isfinite(got.imag)
    fin_f = np.isfinite(f32.real) & np.isfinite(f32.imag)
    assert np.array_equal(fin_g, fin_f)
    assert (~fin_g).any() and fin_g.any()
    assert np.array_equal(got[fin_g] == 0, f32[fin_g] == 0)
    assert (got[fin_g] == 0).any()
    err = np.abs(got[fin_g] - f32[fin_g]).max() / np.abs(f32[fin_g]).max()
    assert err < GOLD_TOL["fp16_split"][0] + GOLD_TOL["fp32"][0]
    want, wsum = GS.goldstein(x, 32, 0.5)
    m = (wsum > 0) & np.isfinite(want.real) & np.isfinite(want.imag)
    assert np.array_equal(fin_g[wsum > 0], m[wsum > 0])
    big = m & (np.abs(want) > 1e-3 * np.abs(want[m]).max())
    assert phase_rms(got, want, big) < GOLD_TOL["fp16_split"][1]


@pytest.mark.parametrize("mode", ["fp32", "fp16_split"])
def test_goldstein_alpha0_identity(ctx, mode):
    from sentinel_gpu_processor_amd import kernels as K
    rng = np.random.default_rng(6)
    x = (rng.standard_normal((64, 80)) + 1j * rng.standard_normal((64, 80))).astype(np.complex64)
    got = _np(K.goldstein(ctx, _dev(x), alpha=0.0, mode=mode))
    _, wsum = GS.goldstein(x, 32, 0.0)
    m = wsum > 0
    assert np.max(np.abs(got[m] - x[m])) < {"fp32": 1e-5, "fp16_split": 1e-4}[mode] * np.abs(x).max()


@pytest.mark.parametrize("mode", ["fp32", "fp16_split"])
def test_goldstein_deterministic(ctx, mode):
    from sentinel_gpu_processor_amd import kernels as K
    rng = np.random.default_rng(7)
    x = _dev((rng.standard_normal((96, 160)) + 1j * rng.standard_normal((96, 160))).astype(np.complex64))
    a = _np(K.goldstein(ctx, x, mode=mode))
    b = _np(K.goldstein(ctx, x, mode=mode))
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("interp", ["bilinear", "sinc"])
@pytest.mark.parametrize("name", ["small", "small_misreg"])
def test_pair_matches_oracle(ctx, interp, name):
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene(name)
    gr, gs = _geoms(ctx, sc)
    oc = oracle_chain(name, interp)
    dem = sc["dem"]
    dem_t = _dev(dem["data"])
    dref = K.deramp(ctx, gr, _dev(sc["ref_slc"]))
    dsec = K.deramp(ctx, gs, _dev(sc["sec_slc"]))
    ifg, pr, ps = K.pair(ctx, gr, gs, sc["grid"], dem_t, dem, dref, dsec, interp, 1, 4,
                         offpoly=sc["offpoly"])
    ifg, pr, ps = _np(ifg), _np(pr), _np(ps)
    m = (oc["p1"] > 0) & (oc["p2"] > 0) & (np.abs(oc["ifg"]) > 1e-3 * np.abs(oc["ifg"]).max())
    assert phase_rms(ifg, oc["ifg"], m) < 1e-4
    cg = np.abs(ifg) / np.sqrt(pr * ps + 1e-30)
    cw = np.abs(oc["ifg"]) / np.sqrt(oc["p1"] * oc["p2"] + 1e-30)
    assert np.max(np.abs(cg[m] - cw[m])) < 1e-4
    # the fused path equals geocode + crossmul_looks to FP32 rounding
    o_r = K.geocode(ctx, gr, sc["grid"], dem_t, dem, dref, interp)
    o_s = K.geocode(ctx, gs, sc["grid"], dem_t, dem, dsec, interp, offpoly=sc["offpoly"])
    ifg2, pr2, ps2 = K.crossmul_looks(ctx, o_r, o_s, 1, 4)
    assert np.max(np.abs(_np(ifg2) - ifg)) <= 1e-6 * np.abs(ifg).max()


def test_scene_physics_coherence(ctx):
    """Processing the synthetic pair recovers the injected phase and a
    coherence near gamma0 (what the scene construction fixes)."""
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("small")
    gr, gs = _geoms(ctx, sc)
    dem = sc["dem"]
    dem_t = _dev(dem["data"])
    dref = K.deramp(ctx, gr, _dev(sc["ref_slc"]))
    dsec = K.deramp(ctx, gs, _dev(sc["sec_slc"]))
    ifg, pr, ps = K.pair(ctx, gr, gs, sc["grid"], dem_t, dem, dref, dsec, "sinc", 1, 4)
    coh = _np(K.coherence(ctx, ifg, pr, ps, win=5))
    ifg = _np(ifg)
    h = G.target_heights(sc["grid"], dem)[:, ::4][:, :ifg.shape[1]]
    res = ifg * np.exp(-2j * np.pi * h / sc["amb_height"])
    m = (_np(pr) > 0) & (_np(ps) > 0)
    assert 0.55 < np.median(coh[m]) < 0.9
    assert abs(np.angle(np.mean(res[m]))) < 0.1
