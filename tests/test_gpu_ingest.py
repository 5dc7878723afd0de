"""GPU tests of the reference's production input types: the int16 DEM
(geo2rdr.cu:171, 199; short int *dem) through geocoding, the fused pair and
its tile-skip test, and the geo2rdr executable's whole input contract --
SQLite parameter table, orbit / FM-rate / Doppler text files, an int16 DEM
with its .rsc, burst data streamed as on stdin -- into .gslc files
(geo2rdr.cu:316-651), each against the FP64 oracle."""

import io
import sqlite3

import numpy as np
import pytest

from oracle import geom as G
from tests.scenes import dem_i16, oracle_chain, phase_rms, scene

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


@pytest.mark.parametrize("interp", ["bilinear", "sinc"])
@pytest.mark.parametrize("name", ["small_misreg", "small_tops"])
def test_int16_dem_geocode_and_pair(ctx, interp, name):
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene(name)
    dem = dem_i16(sc["dem"])
    dem_t = _dev(dem["data"])
    assert dem_t.dtype == torch.int16
    gr, gs = _geoms(ctx, sc)
    oc = oracle_chain(name, interp, dem16=True)
    dref = K.deramp(ctx, gr, _dev(sc["ref_slc"]))
    dsec = K.deramp(ctx, gs, _dev(sc["sec_slc"]))
    o_s = _np(K.geocode(ctx, gs, sc["grid"], dem_t, dem, dsec, interp, offpoly=sc["offpoly"]))
    vg, vw = o_s != 0, oc["gs"] != 0
    assert np.count_nonzero(vg != vw) <= 2
    both = vg & vw
    assert phase_rms(o_s, oc["gs"], both & (np.abs(oc["gs"]) > 1e-2)) < 1e-4
    ifg, pr, ps = K.pair(ctx, gr, gs, sc["grid"], dem_t, dem, dref, dsec, interp, 1, 4,
                         offpoly=sc["offpoly"])
    ifg, pr, ps = _np(ifg), _np(pr), _np(ps)
    m = (oc["p1"] > 0) & (oc["p2"] > 0) & (np.abs(oc["ifg"]) > 1e-3 * np.abs(oc["ifg"]).max())
    assert m.mean() > 0.2
    assert phase_rms(ifg, oc["ifg"], m) < 1e-4
    assert np.count_nonzero((pr > 0) != (oc["p1"] > 0)) <= 2
    # the heights are the rounded ones: the float DEM gives a different product
    of = oracle_chain(name, interp)
    assert phase_rms(of["ifg"], oc["ifg"], m) > 1e-4


def _write_reference_inputs(d, sc, extra_burst_dt):
    """The geo2rdr executable's inputs for the scene's reference burst plus a
    second burst extra_burst_dt seconds later: SQLite `file` table
    (sql_mod.py:114-126 layout, geo2rdr.cu:333-376 keys), orbit text
    (orbit.cu:147-168), FM / DC records (sario.cu:80-103), int16 DEM + .rsc."""
    from sentinel_gpu_processor_amd.geocoordinates import GeoCoordinates
    b = sc["burst_ref"]
    t, pos, vel = sc["orbit_ref"]
    with open(d / "orbit.txt", "w") as f:
        f.write(f"{len(t)}\n")
        for i in range(len(t)):
            f.write(" ".join(repr(float(v)) for v in (t[i], *pos[i], *vel[i])) + " 0 0 0\n")
    for name, rec in (("fmrate.txt", b["fm"]), ("dc.txt", b["dc"])):
        with open(d / name, "w") as f:
            f.write(f"{len(rec)}\n")
            for r in rec:
                f.write(" ".join(repr(float(v)) for v in r) + "\n")
    dem = dem_i16(sc["dem"])
    dem["data"].tofile(str(d / "dem.dem"))
    ny, nx = dem["data"].shape
    rsc = GeoCoordinates(rscparams=dict(nlat=ny, nlon=nx, latmax=dem["lat0"], lonmin=dem["lon0"],
                                        dlat=dem["dlat"], dlon=dem["dlon"]))
    rsc.save_as_rsc(str(d / "dem.dem.rsc"))
    vals = dict(demfile="dem.dem", rscfile="dem.dem.rsc", samplesPerBurst=b["samples"],
                linesPerBurst=b["lines"], azimuthBursts=2, prf=repr(b["prf"]),
                wvl=repr(b["wavelength"]), hmin=0.0, hmax=300.0,
                rangeSamplingRate=repr(b["range_sampling_rate"]),
                slantRangeTime=repr(b["slant_range_time"]),
                azimuthSteeringRate=repr(b["azimuth_steering_rate"]),
                azimuthTimeInterval=repr(b["azimuth_time_interval"]),
                radarFrequency=repr(b["radar_frequency"]), orbinfo="orbit.txt",
                fmrateinfo="fmrate.txt", dcinfo="dc.txt")
    for k, dt in ((1, 0.0), (2, extra_burst_dt)):
        vals[f"azimuthTimeSeconds{k}"] = repr(b["t_start"] + dt)
        vals[f"firstValidLine{k}"] = b["first_valid_line"]
        vals[f"lastValidLine{k}"] = b["last_valid_line"]
        vals[f"firstValidSample{k}"] = b["first_valid_sample"]
        vals[f"lastValidSample{k}"] = b["last_valid_sample"]
    db = sqlite3.connect(str(d / "params.db"))
    db.execute("CREATE TABLE file (name TEXT UNIQUE, value TEXT, units TEXT, type TEXT, "
               "comments TEXT)")
    for k, v in vals.items():
        db.execute("INSERT INTO file (name, value) VALUES (?, ?)", (k, str(v)))
    db.commit()
    db.close()
    return dem, rsc


def test_geo2rdr_from_db_matches_oracle(tmp_path):
    """Two bursts (the second 0.06 s later, on the same data) from the
    database, int16 DEM and a byte stream, into .gslc files whose headers
    and pixels match the oracle's reproject on the DEM windows."""
    from sentinel_gpu_processor_amd import geo2rdr as GR
    from sentinel_gpu_processor_amd.sario import CroppedImage
    sc = scene("small_tops")
    dt = 0.06
    dem, rsc = _write_reference_inputs(tmp_path, sc, dt)
    slc = np.ascontiguousarray(sc["ref_slc"], np.complex64)
    stream = io.BytesIO(slc.tobytes() + slc.tobytes())
    files = GR.geo2rdr_from_db(str(tmp_path / "params.db"), str(tmp_path / "out"), stream=stream)
    assert [f.rsplit("_", 1)[1] for f in files] == ["0.gslc", "1.gslc"]
    torb, xx, vv = sc["orbit_ref"]
    for k, fn in enumerate(files):
        b = dict(sc["burst_ref"])
        b["t_start"] += k * dt
        t0, t1 = b["t_start"], b["t_start"] + (b["lines"] - 1) / b["prf"]
        r0 = b["slant_range_time"] * G.SOL / 2
        r1 = r0 + (b["samples"] - 1) * G.SOL / 2 / b["range_sampling_rate"]
        ll = G.bounds(t0, t1, r0, r1, 0.0, 300.0, torb, xx, vv)
        left, top, right, bottom = GR.dem_window(ll, rsc)
        grid = dict(nrow=bottom - top, ncol=right - left, lat0=rsc.latmax + rsc.dlat * top,
                    lon0=rsc.lonmin + rsc.dlon * left, dlat=rsc.dlat, dlon=rsc.dlon)
        h = dem["data"][top:bottom, left:right].astype(np.float64)
        su = G.deramp_setup(sc["orbit_ref"], b)
        want, _ = G.reproject(sc["orbit_ref"], b, su, grid, h,
                              G.deramp(slc, su).astype(np.complex64))
        f, l = G.nonzero_rows(want)
        ci = CroppedImage.from_file(fn, load_data=True)
        assert (ci.nrow0, ci.ncol0) == (rsc.nlat, rsc.nlon)
        assert (ci.left, ci.top, ci.right, ci.bottom) == (left, top + f, right, top + l)
        got, w = ci.data, want[f:l]
        m = (got != 0) & (w != 0)
        assert m.mean() > 0.3
        assert np.count_nonzero((got != 0) != (w != 0)) <= 2
        assert np.max(np.abs(got[m] - w[m])) < 2e-5 * np.abs(w[m]).max()


def test_geo2rdr_from_db_short_stream(tmp_path):
    """Fewer bytes than the bursts need: an error, as read_binary_stdin's
    short read (geo2rdr.cu:67-75), not a silent partial output."""
    from sentinel_gpu_processor_amd import geo2rdr as GR
    sc = scene("small_tops")
    _write_reference_inputs(tmp_path, sc, 0.06)
    slc = np.ascontiguousarray(sc["ref_slc"], np.complex64)
    with pytest.raises(EOFError):
        GR.geo2rdr_from_db(str(tmp_path / "params.db"), str(tmp_path / "out"),
                           stream=io.BytesIO(slc.tobytes()))


def _ci16(slc, peak=3000.0):
    """Complex samples quantised to interleaved CInt16 (re, im)."""
    s = peak / np.abs(slc).max()
    q = np.stack([np.round(slc.real * s), np.round(slc.imag * s)], -1)
    return np.clip(q, -32768, 32767).astype(np.int16)


@pytest.mark.parametrize("odd", [False, True])
@pytest.mark.parametrize("aligned", [True, False])
@pytest.mark.parametrize("name", ["small_tops", "small_misreg"])
def test_deramp_ci16(ctx, name, odd, aligned):
    """CInt16 bursts deramped straight into complex float: bit for bit the
    complex-float deramp of the converted samples (8-byte vector path and
    the per-sample path: odd widths, misaligned input), and against the
    oracle's deramp (geo2rdr.cu:138-158) of the converted samples."""
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene(name)
    b = dict(sc["burst_ref"])
    raw = _ci16(sc["ref_slc"])
    if odd:
        b["samples"] -= 1
        b["last_valid_sample"] = min(b["last_valid_sample"], b["samples"] - 1)
        raw = np.ascontiguousarray(raw[:, :-1])
    L, Nr = b["lines"], b["samples"]
    print(f"[deramp_ci16] {name} {L}x{Nr} aligned={aligned}", flush=True)
    g = N.Geometry(ctx, sc["orbit_ref"], b)
    if aligned:
        raw_t = _dev(raw)
    else:
        buf = torch.zeros(raw.size + 2, dtype=torch.int16, device="cuda")
        raw_t = buf[2:].view(L, Nr, 2)
        raw_t.copy_(_dev(raw))
        assert raw_t.data_ptr() % 8 == 4
    conv = (raw[..., 0].astype(np.float32) + 1j * raw[..., 1].astype(np.float32)).astype(np.complex64)
    got = K.deramp_ci16(ctx, g, raw_t)
    want_dev = K.deramp(ctx, g, _dev(conv))
    got_np, want_np = _np(got), _np(want_dev)
    assert np.array_equal(got_np.view(np.uint32), want_np.view(np.uint32))
    su = G.deramp_setup(sc["orbit_ref"], b)
    ora = G.deramp(conv, su)
    assert np.max(np.abs(got_np - ora)) <= 2e-6 * np.abs(ora).max()
    g.close()


def test_deramp_ci16_rejects_bad_buffers(ctx):
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("small_tops")
    b = sc["burst_ref"]
    L, Nr = b["lines"], b["samples"]
    g = N.Geometry(ctx, sc["orbit_ref"], b)
    buf = torch.zeros(L * Nr * 8, dtype=torch.uint8, device="cuda")
    raw = buf[:L * Nr * 4].view(torch.int16).view(L, Nr, 2)
    out = buf.view(torch.complex64).view(L, Nr)
    with pytest.raises(N.S1IfgError, match="overlaps"):
        K.deramp_ci16(ctx, g, raw, out=out)
    with pytest.raises(ValueError):
        K.deramp_ci16(ctx, g, torch.zeros((L, Nr + 1, 2), dtype=torch.int16, device="cuda"))
    with pytest.raises(TypeError):
        K.deramp_ci16(ctx, g, torch.zeros((L, Nr, 2), dtype=torch.float32, device="cuda"))
    g.close()


@pytest.mark.parametrize("stored", [True, False])
def test_geo2rdr_from_tiff_matches_float_stream(tmp_path, stored):
    """The scene step end to end: two bursts of a CInt16 measurement TIFF
    inside a zip (stored and deflated) through geo2rdr_from_tiff give the
    same .gslc bytes as the complex-float stream of the converted samples
    through geo2rdr_from_db (sentinel_scene.py:164-207 -> geo2rdr.cu
    --stdin), and a TIFF whose size disagrees with the database is refused."""
    import zipfile

    from sentinel_gpu_processor_amd import geo2rdr as GR
    from tests.tiffgen import write_ci16_tiff
    sc = scene("small_tops")
    _write_reference_inputs(tmp_path, sc, 0.06)
    raw = _ci16(sc["ref_slc"])
    two = np.concatenate([raw, raw])
    blob = write_ci16_tiff(None, two, rps=1)
    zp = tmp_path / "S1A_IW_SLC.zip"
    member = "S1A_IW_SLC.SAFE/measurement/s1a-iw1-slc-vv-001.tiff"
    with zipfile.ZipFile(zp, "w", zipfile.ZIP_STORED if stored else zipfile.ZIP_DEFLATED) as z:
        z.writestr(member, blob)
    db = str(tmp_path / "params.db")
    got = GR.geo2rdr_from_tiff(db, str(tmp_path / "tif"), str(zp), member)
    conv = (two[..., 0].astype(np.float32) + 1j * two[..., 1].astype(np.float32)).astype(np.complex64)
    want = GR.geo2rdr_from_db(db, str(tmp_path / "flt"), stream=io.BytesIO(conv.tobytes()))
    assert [f.rsplit("_", 1)[1] for f in got] == ["0.gslc", "1.gslc"]
    assert [f.rsplit("_", 1)[1] for f in want] == ["0.gslc", "1.gslc"]
    for a, b in zip(got, want):
        with open(a, "rb") as fa, open(b, "rb") as fb:
            assert fa.read() == fb.read()
    bad = tmp_path / "short.tiff"
    bad.write_bytes(write_ci16_tiff(None, raw, rps=1))
    with pytest.raises(RuntimeError, match="do not match DB metadata"):
        GR.geo2rdr_from_tiff(db, str(tmp_path / "bad"), str(bad))
