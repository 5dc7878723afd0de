"""GPU tests of the drop-in modules: geo2rdr .gslc production, interfere()
end to end, the crossmul daemon contract, amplitude / coherence products
and Goldstein file filtering, each against the FP64 oracle."""

import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import geom as G
from oracle import goldstein as GS
from oracle import ifg as OI
from tests.scenes import phase_rms, scene

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _dem_rsc(sc):
    from sentinel_gpu_processor_amd.geocoordinates import GeoCoordinates
    d = sc["dem"]
    ny, nx = d["data"].shape
    return GeoCoordinates(rscparams=dict(nlat=ny, nlon=nx, latmax=d["lat0"], lonmin=d["lon0"],
                                         dlat=d["dlat"], dlon=d["dlon"]))


@pytest.fixture(scope="module")
def gslc(tmp_path_factory):
    """Both bursts of the small scene geocoded onto (a window of) its DEM grid."""
    from sentinel_gpu_processor_amd import geo2rdr as GR
    sc = scene("small")
    rsc = _dem_rsc(sc)
    d = tmp_path_factory.mktemp("slc")
    names = {}
    for tag, date in (("ref", "20200101"), ("sec", "20200113")):
        prefix = str(d / f"{date}_S1A_IW_SLC_a_b_c_iw1")
        files = GR.geocode_bursts(sc[f"orbit_{tag}"], [sc[f"burst_{tag}"]], [sc[f"{tag}_slc"]],
                                  sc["dem"]["data"], rsc, prefix, 0.0, 300.0)
        assert len(files) == 1
        names[tag] = files[0]
    return sc, rsc, d, names


def test_geocode_bursts_matches_oracle(gslc):
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import geo2rdr as GR
    from sentinel_gpu_processor_amd.runtime import context
    from sentinel_gpu_processor_amd.sario import CroppedImage
    sc, rsc, d, names = gslc
    b = sc["burst_ref"]
    geom = N.Geometry(context(), sc["orbit_ref"], b)
    ll = geom.bounds(0.0, 300.0)
    torb, xx, vv = sc["orbit_ref"]
    t0 = b["t_start"]
    t1 = t0 + (b["lines"] - 1) / b["prf"]
    r0 = b["slant_range_time"] * G.SOL / 2
    r1 = r0 + (b["samples"] - 1) * G.SOL / 2 / b["range_sampling_rate"]
    want_ll = G.bounds(t0, t1, r0, r1, 0.0, 300.0, torb, xx, vv)
    assert np.max(np.abs(np.array(ll) - np.array(want_ll))) < 1e-9
    left, top, right, bottom = GR.dem_window(ll, rsc)
    ci = CroppedImage.from_file(names["ref"], load_data=True)
    assert (ci.nrow0, ci.ncol0, ci.left, ci.right) == (rsc.nlat, rsc.nlon, left, right)
    grid = dict(nrow=bottom - top, ncol=right - left, lat0=rsc.latmax + rsc.dlat * top,
                lon0=rsc.lonmin + rsc.dlon * left, dlat=rsc.dlat, dlon=rsc.dlon)
    h = sc["dem"]["data"][top:bottom, left:right].astype(np.float64)
    su = G.deramp_setup(sc["orbit_ref"], b)
    want, _ = G.reproject(sc["orbit_ref"], b, su, grid, h,
                          G.deramp(sc["ref_slc"], su).astype(np.complex64))
    f, l = G.nonzero_rows(want)
    assert (ci.top, ci.bottom) == (top + f, top + l)
    got = ci.data
    w = want[f:l]
    m = (got != 0) & (w != 0)
    assert np.count_nonzero((got != 0) != (w != 0)) <= 2
    assert np.max(np.abs(got[m] - w[m])) < 2e-5 * np.abs(w[m]).max()


def test_interfere_end_to_end(gslc, tmp_path):
    from sentinel_gpu_processor_amd import interfere as IF
    from sentinel_gpu_processor_amd.sario import CroppedImage
    sc, rsc, d, names = gslc
    rsc.save_as_rsc(str(tmp_path / "dem.rsc"))
    ig = tmp_path / "igrams"
    ig.mkdir()
    (ig / "pairs").write_text("20200101 20200113\n")
    IF.interfere("pairs", str(d), str(tmp_path / "dem.rsc"), str(tmp_path / "small.rsc"), str(ig),
                 rowlook=2, collook=3)
    out = np.fromfile(str(ig / "20200101_20200113.int"), dtype=np.complex64)
    nr, nc = rsc.nlat // 2, rsc.nlon // 3
    out = out.reshape(nr, nc)
    a = CroppedImage.from_file(names["ref"], load_data=True)
    b = CroppedImage.from_file(names["sec"], load_data=True)
    want_ml, hdr = OI.crossmul_pair(a.data, [a.nrow0, a.ncol0, a.left, a.top, a.right, a.bottom],
                                    b.data, [b.nrow0, b.ncol0, b.left, b.top, b.right, b.bottom],
                                    2, 3)
    full = np.zeros((nr, nc), np.complex128)
    full[hdr[3]:hdr[5], hdr[2]:hdr[4]] = want_ml
    m = np.abs(full) > 1e-3 * np.abs(full).max()
    assert m.mean() > 0.1
    assert np.max(np.abs(out - full)) < 1e-5 * np.abs(full).max()
    assert phase_rms(out, full, m) < 1e-5
    # resume: a second call finds the output and does nothing
    IF.interfere("pairs", str(d), str(tmp_path / "dem.rsc"), str(tmp_path / "small.rsc"), str(ig),
                 rowlook=2, collook=3)


def test_daemon_contract_fail_lines(gslc, tmp_path):
    from sentinel_gpu_processor_amd.sario import save_cropped
    sc, rsc, d, names = gslc
    far = str(tmp_path / "far.gslc")
    save_cropped(far, np.ones((4, 4), np.complex64), rsc.nlat, rsc.nlon, 0, 0, 4, 4)
    tf = tmp_path / "tasks.txt"
    ok_out = str(tmp_path / "ok.int")
    tf.write_text(f"{names['ref']} {names['sec']} {ok_out}\n"
                  f"{names['ref']} /nonexistent.gslc {tmp_path / 'missing.int'}\n"
                  f"{names['ref']} {far} {tmp_path / 'disjoint.int'}\n")
    env = dict(os.environ, PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-m", "sentinel_gpu_processor_amd.daemon", "--rowlook", "1",
                        "--collook", "4", "--tasks-file", str(tf)], capture_output=True, text=True,
                       env=env, cwd=ROOT, timeout=300)
    assert r.returncode == 1, r.stderr
    assert f"OK {ok_out}" in r.stdout
    assert "SUMMARY 1 2 3" in r.stdout
    assert "cannot read header" in r.stderr and "empty overlap region" in r.stderr
    assert os.path.exists(ok_out)


def test_amplitude_mosaic_and_coherence_files(gslc, tmp_path):
    from sentinel_gpu_processor_amd import coherence as CO
    from sentinel_gpu_processor_amd import interfere as IF
    from sentinel_gpu_processor_amd.sario import CroppedImage, readc
    sc, rsc, d, names = gslc
    amp_dir = tmp_path / "amp"
    CO.multilook_amp(str(d), str(amp_dir), 1, 4)
    nr, nc = rsc.nlat, rsc.nlon // 4
    for tag, date in (("ref", "20200101"), ("sec", "20200113")):
        amp = np.fromfile(str(amp_dir / f"{date}.amp"), dtype=np.float32).reshape(nr, nc)
        ci = CroppedImage.from_file(names[tag], load_data=True)
        l0 = (ci.left + 3) // 4
        r0 = ci.right // 4
        want = OI.amplitude(ci.data[:, l0 * 4 - ci.left:r0 * 4 - ci.left], 1, 4)
        assert np.allclose(amp[ci.top:ci.bottom, l0:r0], want, rtol=1e-5, atol=1e-7)
    ig = tmp_path / "ig"
    ig.mkdir()
    rsc.save_as_rsc(str(tmp_path / "dem.rsc"))
    (ig / "pairs").write_text("20200101 20200113\n")
    IF.interfere("pairs", str(d), str(tmp_path / "dem.rsc"), str(tmp_path / "s.rsc"), str(ig), 1, 4)
    CO.coherence(str(ig), str(amp_dir), nr, nc, 1, 4)
    cc = readc(str(ig / "20200101_20200113.cc"), nc)
    ifg = np.fromfile(str(ig / "20200101_20200113.int"), dtype=np.complex64).reshape(nr, nc)
    a1 = np.fromfile(str(amp_dir / "20200101.amp"), dtype=np.float32).reshape(nr, nc)
    a2 = np.fromfile(str(amp_dir / "20200113.amp"), dtype=np.float32).reshape(nr, nc)
    want = OI.coherence_lookcell(ifg.astype(np.complex128), a1.astype(np.float64),
                                 a2.astype(np.float64), fill="mask")
    ok = np.isfinite(want)
    assert np.max(np.abs(cc.imag[ok] - want[ok])) < 1e-5
    raw = OI.coherence_lookcell(ifg.astype(np.complex128), a1.astype(np.float64),
                                a2.astype(np.float64), fill="none")
    bad = cc.imag[~ok]
    # > 1 in FP64: filled with the deterministic hash value in [0, 1), or,
    # where FP32 rounding lands the kernel's value on exactly 1.0, left as is
    edge = (bad == 1.0) & (raw[~ok] - 1 < 1e-6)
    assert np.all(((bad >= 0) & (bad < 1)) | edge)
    assert np.allclose(cc.real, a1 * a2, rtol=1e-6)


def test_goldstein_file_filter(tmp_path):
    from sentinel_gpu_processor_amd import phase_correction as PC
    rng = np.random.default_rng(9)
    x = (np.exp(1j * 0.1 * np.arange(80))[None, :] * np.ones((48, 1))
         + 0.5 * (rng.standard_normal((48, 80)) + 1j * rng.standard_normal((48, 80))))
    x = x.astype(np.complex64)
    f = str(tmp_path / "a.int")
    x.tofile(f)
    lst = tmp_path / "list"
    lst.write_text(f + "\n")
    outs = PC.goldstein_filter(str(lst), 48, 80, 32, 0.5, mode="fp32")
    assert outs == [f + ".filt"]
    got = np.fromfile(outs[0], dtype=np.complex64).reshape(48, 80)
    want, wsum = GS.goldstein(x, 32, 0.5)
    m = wsum > 0
    assert phase_rms(got, want, m) < 1e-4


def test_sario_looks_on_gpu():
    from sentinel_gpu_processor_amd import sario as SA
    rng = np.random.default_rng(10)
    a = (rng.standard_normal((30, 41)) + 1j * rng.standard_normal((30, 41))).astype(np.complex64)
    assert np.allclose(SA.cpxlooks(a, 3, 4), OI.cpx_looks(a, 3, 4), rtol=1e-5, atol=1e-6)
    assert np.allclose(SA.powlooks(a, 3, 4), OI.powlooks(a, 3, 4), rtol=1e-5, atol=1e-6)
    assert np.allclose(SA.powlooks(a, 1, 1), OI.powlooks(a, 1, 1), rtol=1e-5, atol=1e-6)


def test_daemon_lanes_match_single_lane(gslc, tmp_path):
    """The driver with 1 and with 3 stream lanes (overlapped transfers and
    kernels, crossmul_daemon.cu:921-975): the same bytes in every output,
    complex and phase, with a failing task in the middle and windows of
    different sizes so that the lanes' buffers grow."""
    import io

    from sentinel_gpu_processor_amd.daemon import CrossmulDriver
    from sentinel_gpu_processor_amd.sario import CroppedImage, save_cropped
    sc, rsc, d, names = gslc
    a = CroppedImage.from_file(names["ref"], load_data=True)
    tasks = []
    for k in range(7):
        cut = str(tmp_path / f"cut{k}.gslc")
        dr = 3 * k
        save_cropped(cut, a.data[dr:], a.nrow0, a.ncol0, a.left, a.top + dr, a.right, a.bottom)
        tasks.append((names["ref"], cut, f"OUT{k}.int"))
    tasks.insert(3, (names["ref"], "/nonexistent.gslc", "OUTbad.int"))
    for out_float in (False, True):
        res = {}
        for streams in (1, 3):
            od = tmp_path / f"s{streams}_{int(out_float)}"
            od.mkdir()
            tk = [(r, s_, str(od / o)) for r, s_, o in tasks]
            so, se = io.StringIO(), io.StringIO()
            drv = CrossmulDriver(1, 4, out_float, 2, 4, out=so, err=se, streams=streams)
            assert drv.run(tk) == (7, 1)
            assert "SUMMARY 7 1 8" in so.getvalue() and "FAIL" in se.getvalue()
            res[streams] = {o: open(str(od / o), "rb").read() for _, _, o in tasks
                            if o != "OUTbad.int"}
        assert res[1] == res[3]
