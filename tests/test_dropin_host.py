"""Host-side logic of the drop-in layer (no GPU): file formats, task
ordering, burst matching, overlap rules, stitching, config, reference
parameter files and the daemon's argument contract."""

import os
import sqlite3
import subprocess
import sys

import numpy as np
import pytest

from oracle import ifg as OI
from sentinel_gpu_processor_amd import daemon as D
from sentinel_gpu_processor_amd import interfere as IF
from sentinel_gpu_processor_amd import sario as SA
from sentinel_gpu_processor_amd.geocoordinates import GeoCoordinates

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cropped_image_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    d = (rng.standard_normal((5, 7)) + 1j * rng.standard_normal((5, 7))).astype(np.complex64)
    f = str(tmp_path / "a.gslc")
    SA.save_cropped(f, d, 40, 50, 10, 3, 17, 8)
    raw = np.fromfile(f, dtype=np.int32, count=64)
    assert list(raw[:6]) == [40, 50, 10, 3, 17, 8] and np.all(raw[6:] == 0)
    ci = SA.CroppedImage.from_file(f, load_data=True)
    assert (ci.nrow, ci.ncol) == (5, 7) and np.array_equal(ci.data, d)
    lazy = SA.CroppedImage.from_file(f)
    full = lazy.load_data(8, 1, 20, 12)  # window larger than the data
    want = np.zeros((11, 12), np.complex64)
    want[2:7, 2:9] = d
    assert np.array_equal(full, want)
    assert np.array_equal(lazy.load_data(10, 3, 17, 8), d)


def test_truncated_file_raises(tmp_path):
    f = str(tmp_path / "t.gslc")
    SA.save_cropped(f, np.ones((4, 4), np.complex64), 10, 10, 0, 0, 4, 4)
    with open(f, "r+b") as fh:
        fh.truncate(64 * 4 + 8 * 5)
    with pytest.raises(OSError):
        SA.CroppedImage.from_file(f, load_data=True)
    with open(f, "r+b") as fh:
        fh.truncate(10)
    with pytest.raises(OSError):
        SA.read_header(f)


def test_subswath_sorting_and_bounds(tmp_path):
    files = []
    for k, (t, b) in enumerate([(40, 60), (0, 25), (20, 45)]):
        f = str(tmp_path / f"20200101_x_iw1_{k}.gslc")
        SA.save_cropped(f, np.ones((b - t, 3), np.complex64), 100, 10, 2, t, 5, b)
        files.append(f)
    s = SA.Subswath(files)
    assert [b.top for b in s.bursts] == [0, 20, 40]
    assert s.bounds() == [2, 0, 5, 60]
    g = SA.BurstGroup(files)
    assert g.subswaths[0].size == 3 and g.subswaths[1].is_empty()


def test_match_bursts_matches_oracle(tmp_path):
    def mk(name, spans):
        fs = []
        for k, (t, b) in enumerate(spans):
            f = str(tmp_path / f"{name}_{k}.gslc")
            SA.save_cropped(f, np.ones((b - t, 2), np.complex64), 200, 10, 0, t, 2, b)
            fs.append(f)
        return SA.Subswath(fs)

    ref = mk("r", [(0, 30), (28, 60), (58, 90)])
    sec = mk("s", [(1, 31), (59, 92), (27, 61), (120, 150)])
    pairs = IF.match_bursts(ref, sec)
    idx = OI.match_bursts([(b.top, b.bottom) for b in ref.bursts],
                          [(b.top, b.bottom) for b in sec.bursts])
    assert [p[1] for p in pairs] == [sec.bursts[i].data for i in idx]


def test_overlap_rules_match_oracle():
    cases = [((100, 100, 10, 3, 50, 40), (100, 100, 12, 5, 47, 44), 1, 4),
             ((100, 100, 10, 3, 50, 40), (100, 100, 12, 5, 47, 44), 3, 5),
             ((100, 100, 0, 0, 10, 10), (100, 100, 20, 20, 30, 30), 1, 1),   # disjoint
             ((100, 100, 0, 0, 50, 50), (100, 100, 10, 10, 20, 20), 2, 3),   # nested
             ((100, 100, 0, 0, 3, 3), (100, 100, 1, 1, 3, 3), 4, 4)]          # < one look
    for h1, h2, rl, cl in cases:
        o = D.overlap(h1, h2, rl, cl)
        w = OI.overlap(h1, h2, rl, cl)
        if w is None:
            assert o is None
        else:
            assert o == (w["left"], w["top"], w["right"], w["bottom"])
            assert o[0] % cl == 0 and o[2] % cl == 0 and o[1] % rl == 0 and o[3] % rl == 0


def test_reorder_tasks_for_cache_locality():
    t = [("A", "B", "1"), ("C", "D", "2"), ("A", "C", "3"), ("B", "E", "4"), ("D", "F", "5")]
    out = IF._reorder_tasks_for_cache_locality(t)
    assert sorted(out) == sorted(t)
    # seed is task 0 (cache A); tasks touching A are drained (0 then 2), the
    # other image of the last drained task (C) becomes the cache key, ...
    assert out[0] == t[0] and out[1] == t[2]
    assert out[2] == t[1]  # touches C
    assert out[3] == t[4]  # touches D
    assert out[4] == t[3]  # cold start
    # ref / sec never swapped
    assert all(o in t for o in out)


@pytest.mark.gpu
def test_stitch_subswath_recovers_offsets(tmp_path):
    """Bursts from two scenes with a known phase offset across the scene
    boundary: the stitched mosaic is continuous (interfere.py:102-143), and
    equals the oracle's restatement."""
    rng = np.random.default_rng(1)
    nrow0, ncol0 = 60, 40
    base = np.exp(1j * rng.uniform(-np.pi, np.pi, (nrow0, ncol0))).astype(np.complex64)
    spans = [(0, 25), (20, 45), (40, 60)]
    offs = [0.0, 0.0, 0.7]  # third burst comes from another scene
    files, bursts, ids = [], [], []
    for k, ((t, b), d) in enumerate(zip(spans, offs)):
        scene = "s1" if k < 2 else "s2"
        name = f"2020_{scene}_x_a_b_c_d_{scene}_y_{k}_iw1.int"
        data = (base[t:b] * np.exp(1j * d)).astype(np.complex64)
        f = str(tmp_path / name)
        SA.save_cropped(f, data, nrow0, ncol0, 0, t, ncol0, b)
        files.append(f)
        bursts.append(((nrow0, ncol0, 0, t, ncol0, b), data))
        w = name.split("_")
        ids.append((w[1] + w[2], w[7] + w[8]))
    want, _, _ = OI.stitch_subswath(bursts, ids)
    got = IF.stitch_subswath(files, False)
    assert np.allclose(got.data, want, atol=1e-6)
    assert np.allclose(np.angle(got.data * np.conj(base)), 0, atol=1e-5)
    assert not any(os.path.exists(f) for f in files)


def test_geocoordinates_take_look_and_round_trip(tmp_path):
    g = GeoCoordinates(rscparams=dict(nlat=100, nlon=203, latmax=34.5, lonmin=-118.2,
                                      dlat=-0.0002, dlon=0.0003))
    f = str(tmp_path / "a.rsc")
    g.save_as_rsc(f)
    h = GeoCoordinates(f)
    assert (h.nlat, h.nlon) == (100, 203)
    assert h.latmax == pytest.approx(34.5) and h.dlon == pytest.approx(0.0003)
    t = h.take_look(2, 4)
    assert (t.nlat, t.nlon) == (50, 50)
    assert t.latmax == pytest.approx(34.5 - 0.0002 * 0.5)
    assert t.lonmin == pytest.approx(-118.2 + 0.0003 * 1.5)
    assert h.ll2xy(34.5 - 0.0002 * 10.5, -118.2 + 0.0003 * 7.5) == (10, 7)


def test_config_loading(tmp_path):
    f = tmp_path / "config.yaml"
    f.write_text("io:\n  slc_path: slc\n  ifg_path: igrams\n  unknown_key: 1\n"
                 "proc:\n  rowlook: 2\n  collook: 8\n  io_workers: 3\n"
                 "filter:\n  enable: true\n  parameters:\n    goldstein_alpha: 0.7\n"
                 "tropo:\n  enable: false\n")
    from sentinel_gpu_processor_amd.config import load_config
    c = load_config(str(f))
    assert c.io.slc_path == str(tmp_path / "slc")
    assert (c.proc.rowlook, c.proc.collook, c.proc.io_workers) == (2, 8, 3)
    assert c.filter.enable and c.filter.parameters.goldstein_alpha == 0.7
    assert c.filter.parameters.window_size == 32


def test_daemon_task_file_and_tuning_contract(tmp_path):
    tf = tmp_path / "tasks.txt"
    tf.write_text("a b c\n\nbad line\nd e f\n")
    assert D.read_tasks(str(tf)) == [("a", "b", "c"), ("d", "e", "f")]
    env = dict(os.environ, PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-m", "sentinel_gpu_processor_amd.daemon", "--rowlook", "1",
                        "--io-workers", "2", "--tasks-file", str(tf)], capture_output=True,
                       text=True, env=env, cwd=ROOT)
    assert r.returncode == 1 and "Partial overrides are not allowed" in r.stderr
    r = subprocess.run([sys.executable, "-m", "sentinel_gpu_processor_amd.daemon"],
                       capture_output=True, text=True, env=env, cwd=ROOT)
    assert r.returncode == 1 and "--tasks-file" in r.stderr


def test_interfere_partial_tuning_is_fatal(tmp_path):
    from sentinel_gpu_processor_amd.geocoordinates import GeoCoordinates as GC
    slc = tmp_path / "slc"
    slc.mkdir()
    ig = tmp_path / "igrams"
    ig.mkdir()
    GC(rscparams=dict(nlat=20, nlon=20, latmax=1, lonmin=1, dlat=-0.1, dlon=0.1)).save_as_rsc(
        str(tmp_path / "a.rsc"))
    for d in ("20200101", "20200113"):
        SA.save_cropped(str(slc / f"{d}_x_iw1_0.gslc"), np.ones((10, 20), np.complex64), 20, 20, 0,
                        0, 20, 10)
    (ig / "pairs").write_text("20200101 20200113\n")
    with pytest.raises(SystemExit):
        IF.interfere("pairs", str(slc), str(tmp_path / "a.rsc"), str(tmp_path / "small.rsc"),
                     str(ig), 1, 1, io_workers=2)


def test_reference_parameter_files(tmp_path):
    from sentinel_gpu_processor_amd import geo2rdr as GR
    (tmp_path / "orb.txt").write_text("2\n1.0 2 3 4 5 6 7 0 0 0\n2.0 3 4 5 6 7 8 0 0 0\n")
    t, p, v = GR.read_orbit_ascii(str(tmp_path / "orb.txt"))
    assert list(t) == [1.0, 2.0] and list(p[1]) == [3, 4, 5] and list(v[0]) == [5, 6, 7]
    (tmp_path / "fm.txt").write_text("1\n10.0 0.005 -2300 1 2\n")
    assert GR.read_polynomials(str(tmp_path / "fm.txt")).tolist() == [[10.0, 0.005, -2300, 1, 2]]
    db = sqlite3.connect(str(tmp_path / "p.db"))
    db.execute("CREATE TABLE file (name TEXT UNIQUE, value TEXT, units TEXT, type TEXT, comments TEXT)")
    vals = dict(demfile="dem", rscfile="dem.rsc", samplesPerBurst=100, linesPerBurst=50,
                azimuthBursts=1, prf=486.5, wvl=0.0554658, hmin=-100, hmax=2000,
                rangeSamplingRate=64345238.0, slantRangeTime=0.0053, azimuthSteeringRate=1.59,
                azimuthTimeInterval=0.0020556, radarFrequency=5.405e9, orbinfo="orb.txt",
                fmrateinfo="fm.txt", dcinfo="fm.txt", azimuthTimeSeconds1=100.5,
                firstValidLine1=3, lastValidLine1=47, firstValidSample1=5, lastValidSample1=95)
    for k, v in vals.items():
        db.execute("INSERT INTO file (name, value) VALUES (?, ?)", (k, str(v)))
    db.commit()
    db.close()
    sw, per = GR.bursts_from_db(str(tmp_path / "p.db"))
    assert sw["samples"] == 100 and sw["lines"] == 50 and sw["hmax"] == 2000.0
    assert per == [dict(t_start=100.5, first_valid_line=3, last_valid_line=47,
                        first_valid_sample=5, last_valid_sample=95)]


def test_dem_window_rule():
    from sentinel_gpu_processor_amd import geo2rdr as GR
    rsc = GeoCoordinates(rscparams=dict(nlat=100, nlon=100, latmax=10.0, lonmin=20.0,
                                        dlat=-0.125, dlon=0.125))
    # geo2rdr.cu:575-586 (truncating int conversions, then clamping)
    assert GR.dem_window((8.0, 9.5, 20.25, 21.0), rsc) == (2, 4, 9, 17)
    assert GR.dem_window((-5.0, -4.0, 20.25, 21.0), rsc) is None
    assert GR.dem_window((8.0, 10.5, 19.0, 21.0), rsc) == (0, 0, 9, 17)
