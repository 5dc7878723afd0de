"""Back-geocoding of Sentinel-1 TOPS bursts onto a DEM lat/lon grid: the
``geo2rdr`` executable's pipeline (geo2rdr.cu:316-734), in-process on the GPU.

Per burst (geo2rdr.cu:466-651):
  footprint for heights [hmin, hmax]  -> DEM window (left, top, right, bottom)
  deramp (HIP)  -> geo2rdr + interpolation + reramp + flattening on the
  DEM-window grid (HIP)  -> first/last non-zero rows (HIP)  -> crop ->
  ``<prefix>_burst_<k>.gslc`` with header
  [dem nlat, dem nlon, left, top + first, right, top + last].

Inputs come either as Python objects (``geocode_bursts``) or, as for the
reference executable, from its SQLite ``file`` table plus the orbit /
FM-rate / Doppler text files it names (``geo2rdr_from_db``;
sql_mod.cpp:17-85, orbit.cu:147-168, sario.cu:80-103).
"""

import os
import sqlite3

import numpy as np

from . import _native as N
from . import kernels as K
from .geocoordinates import GeoCoordinates
from .runtime import context, device
from .sario import save_cropped

SOL = 299792458.0


# ------------------------------------------------------------ reference inputs
def get_param(db, name, table="file"):
    """Value of ``name`` in the SQLite parameter table (sql_mod.cpp:17-67)."""
    cur = db.execute(f"SELECT value FROM {table} WHERE name = ?", (name,))
    row = cur.fetchone()
    if row is None:
        raise KeyError(f"parameter {name!r} not in table {table!r}")
    return row[0]


def read_orbit_ascii(path):
    """n, then n rows of t x y z vx vy vz (3 ignored) (orbit.cu:147-168)."""
    with open(path) as f:
        tok = f.read().split()
    n = int(tok[0])
    v = np.array(tok[1:1 + 10 * n], dtype=np.float64).reshape(n, 10)
    return v[:, 0].copy(), v[:, 1:4].copy(), v[:, 4:7].copy()


def read_polynomials(path):
    """n, then n records (t, t0, c0, c1, c2) (sario.cu:80-103)."""
    with open(path) as f:
        tok = f.read().split()
    n = int(tok[0])
    return np.array(tok[1:1 + 5 * n], dtype=np.float64).reshape(n, 5)


def bursts_from_db(dbname):
    """Swath-level parameters and per-burst metadata from the reference's
    SQLite table (geo2rdr.cu:333-384)."""
    db = sqlite3.connect(dbname)
    try:
        g = lambda k: get_param(db, k)
        sw = dict(demfile=g("demfile"), rscfile=g("rscfile"), samples=int(g("samplesPerBurst")),
                  lines=int(g("linesPerBurst")), nbursts=int(g("azimuthBursts")),
                  prf=float(g("prf")), wvl=float(g("wvl")), hmin=float(g("hmin")),
                  hmax=float(g("hmax")), range_sampling_rate=float(g("rangeSamplingRate")),
                  slant_range_time=float(g("slantRangeTime")),
                  azimuth_steering_rate=float(g("azimuthSteeringRate")),
                  azimuth_time_interval=float(g("azimuthTimeInterval")),
                  radar_frequency=float(g("radarFrequency")), orbinfo=g("orbinfo"),
                  fmrateinfo=g("fmrateinfo"), dcinfo=g("dcinfo"))
        try:
            sw["slc_file"] = g("slc_file")
        except KeyError:
            sw["slc_file"] = ""
        per = []
        for k in range(1, sw["nbursts"] + 1):
            per.append(dict(t_start=float(g(f"azimuthTimeSeconds{k}")),
                            first_valid_line=int(g(f"firstValidLine{k}")),
                            last_valid_line=int(g(f"lastValidLine{k}")),
                            first_valid_sample=int(g(f"firstValidSample{k}")),
                            last_valid_sample=int(g(f"lastValidSample{k}"))))
    finally:
        db.close()
    return sw, per


def burst_dict(sw, b, fm, dc):
    """Merge swath and burst metadata into the s1ifg burst description."""
    return dict(lines=sw["lines"], samples=sw["samples"], first_valid_line=b["first_valid_line"],
                last_valid_line=b["last_valid_line"], first_valid_sample=b["first_valid_sample"],
                last_valid_sample=b["last_valid_sample"], t_start=b["t_start"], prf=sw["prf"],
                azimuth_time_interval=sw["azimuth_time_interval"],
                slant_range_time=sw["slant_range_time"],
                range_sampling_rate=sw["range_sampling_rate"], wavelength=sw["wvl"],
                radar_frequency=sw["radar_frequency"],
                azimuth_steering_rate=sw["azimuth_steering_rate"], fm=fm, dc=dc)


# ----------------------------------------------------------------- pipeline
def dem_window(latlon, rsc):
    """DEM sub-window of a footprint (geo2rdr.cu:569-586), or None if empty."""
    latmin, latmax, lonmin, lonmax = latlon
    left = int((lonmin - rsc.lonmin) / rsc.dlon)
    top = int((latmax - rsc.latmax) / rsc.dlat)
    right = int((lonmax - rsc.lonmin) / rsc.dlon + 1)
    bottom = int((latmin - rsc.latmax) / rsc.dlat + 1)
    left, top = max(0, left), max(0, top)
    right, bottom = min(int(rsc.nlon), right), min(int(rsc.nlat), bottom)
    if left >= right or top >= bottom:
        return None
    return left, top, right, bottom


def geocode_bursts(orbit, bursts, burst_data, dem, rsc, slcoutfile, hmin, hmax,
                   interp="bilinear"):
    """Geocode every burst and write ``<slcoutfile>_burst_<k>.gslc``.

    orbit: (t, pos, vel); bursts: list of s1ifg burst dicts; burst_data:
    iterable of complex64 (lines x samples) arrays, or int16 (lines x samples
    x 2) CInt16 arrays as stored in the measurement TIFF, in burst order; dem: 2-D
    int16 / float32 array on the grid ``rsc`` (GeoCoordinates).  Returns the
    list of files written (bursts outside the DEM or all-zero are skipped,
    as geo2rdr.cu:583-584, 632-633).
    """
    ctx = context()
    dev = device()
    dem = np.asarray(dem)
    written = []
    for k, (b, data) in enumerate(zip(bursts, burst_data)):
        geom = N.Geometry(ctx, orbit, b)
        win = dem_window(geom.bounds(hmin, hmax), rsc)
        if win is None:
            continue
        left, top, right, bottom = win
        sub = np.ascontiguousarray(dem[top:bottom, left:right])
        if sub.dtype not in (np.int16, np.float32):
            sub = sub.astype(np.float32)
        if sub.shape[0] < 2 or sub.shape[1] < 2:
            continue
        grid = {"nrow": bottom - top, "ncol": right - left,
                "lat0": rsc.latmax + rsc.dlat * top, "lon0": rsc.lonmin + rsc.dlon * left,
                "dlat": rsc.dlat, "dlon": rsc.dlon}
        demd = dict(grid)
        dem_t = K.from_numpy(sub, dev)
        arr = np.asarray(data)
        if arr.dtype == np.int16:
            # CInt16 measurement samples: converted inside the deramp
            raw = K.from_numpy(arr.reshape(b["lines"], b["samples"], 2), dev)
            z = K.deramp_ci16(ctx, geom, raw)
            del raw
        else:
            z = K.from_numpy(arr.astype(np.complex64, copy=False).reshape(b["lines"],
                                                                        b["samples"]), dev)
            K.deramp(ctx, geom, z, out=z)
        out = K.geocode(ctx, geom, grid, dem_t, demd, z, interp)
        first, last = K.valid_rows(ctx, out)
        if first == last:
            continue
        rows = K.to_numpy(out[first:last])
        fn = f"{slcoutfile}_burst_{k}.gslc"
        save_cropped(fn, rows, int(rsc.nlat), int(rsc.nlon), left, top + first, right, top + last)
        written.append(fn)
        geom.close()
    return written


def _db_inputs(dbname):
    sw, per = bursts_from_db(dbname)
    root = os.path.dirname(os.path.abspath(dbname))
    res = lambda p: p if os.path.isabs(p) else os.path.join(root, p)
    orbit = read_orbit_ascii(res(sw["orbinfo"]))
    fm = read_polynomials(res(sw["fmrateinfo"]))
    dc = read_polynomials(res(sw["dcinfo"]))
    rsc = GeoCoordinates(res(sw["rscfile"]))
    dem = np.memmap(res(sw["demfile"]), dtype=np.int16, mode="r",
                    shape=(int(rsc.nlat), int(rsc.nlon)))
    bursts = [burst_dict(sw, b, fm, dc) for b in per]
    return sw, res, orbit, rsc, dem, bursts


def geo2rdr_from_db(dbname, slcoutfile, slcinfile="", stream=None, interp="bilinear"):
    """The ``geo2rdr db slcout [slcin] [--stdin]`` workflow
    (geo2rdr.cu:316-734): parameters from the database, burst data from
    ``stream`` (a binary file object, like stdin) or from the SLC file."""
    sw, res, orbit, rsc, dem, bursts = _db_inputs(dbname)
    n = sw["lines"] * sw["samples"]

    def data():
        src = stream
        path = slcinfile or (res(sw["slc_file"]) if sw["slc_file"] else "")
        f = src if src is not None else open(path, "rb")
        try:
            for _ in bursts:
                buf = f.read(8 * n)
                if len(buf) != 8 * n:  # geo2rdr.cu:67-75
                    raise EOFError(f"expected {8 * n} bytes of burst data, got {len(buf)}")
                yield np.frombuffer(buf, dtype=np.complex64)
        finally:
            if src is None:
                f.close()

    return geocode_bursts(orbit, bursts, data(), dem, rsc, slcoutfile, sw["hmin"], sw["hmax"],
                          interp)


def geo2rdr_from_tiff(dbname, slcoutfile, zip_path, tiff_name=None, interp="bilinear"):
    """The scene step that streams a measurement TIFF into ``geo2rdr db
    slcout --stdin`` (sentinel_scene.py:164-207, 419-437), in one process:
    parameters from the database, bursts read from the TIFF (``tiff_name``
    inside the zip ``zip_path``, or ``zip_path`` itself when ``tiff_name`` is
    None) as CInt16 and converted on the GPU in the deramp.  Raises
    RuntimeError when the TIFF's size disagrees with the database, as the
    reference does.  Returns the .gslc files written."""
    from .tiff import stream_tiff_from_zip
    sw, _, orbit, rsc, dem, bursts = _db_inputs(dbname)
    data = stream_tiff_from_zip(zip_path, tiff_name, sw["samples"], sw["lines"], len(bursts))
    return geocode_bursts(orbit, bursts, data, dem, rsc, slcoutfile, sw["hmin"], sw["hmax"],
                          interp)
