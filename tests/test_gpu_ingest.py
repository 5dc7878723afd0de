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
