"""The measurement-TIFF reader (sentinel_gpu_processor_amd/tiff.py) on
files written by tests/tiffgen.py: every supported layout read back
exactly, zip members stored and deflated, the reference's dimension check
(sentinel_scene.py:183-192) and the formats it refuses."""

import zipfile

import numpy as np
import pytest

from sentinel_gpu_processor_amd import tiff as T
from tests.tiffgen import write_ci16_tiff


def _data(h=37, w=23, seed=1):
    return np.random.default_rng(seed).integers(-32768, 32768, (h, w, 2)).astype(np.int16)


LAYOUTS = [
    dict(),
    dict(order=">"),
    dict(rps=5),
    dict(rps=7, shuffle=True),
    dict(big=True, rps=4),
    dict(big=True, order=">", rps=37),
    dict(rps=100),
    dict(tile=(16, 16)),
    dict(tile=(32, 16), order=">", big=True),
    dict(encoding="pair", rps=3),
]


@pytest.mark.parametrize("kw", LAYOUTS, ids=lambda kw: ",".join(f"{k}={v}" for k, v in kw.items()) or "plain")
def test_layouts_read_back(tmp_path, kw):
    d = _data()
    fn = write_ci16_tiff(str(tmp_path / "m.tiff"), d, **kw)
    with T.open_measurement(fn) as tif:
        assert (tif.width, tif.height) == (23, 37)
        for r0, n in ((0, 37), (0, 1), (5, 9), (36, 1), (13, 0), (30, 7)):
            got = tif.read_rows(r0, n)
            assert got.dtype == np.int16 and got.dtype.isnative
            np.testing.assert_array_equal(got, d[r0:r0 + n])
        with pytest.raises(T.TiffError):
            tif.read_rows(30, 8)


@pytest.mark.parametrize("comp", [zipfile.ZIP_STORED, zipfile.ZIP_DEFLATED])
def test_zip_members_and_bursts(tmp_path, comp):
    d = _data(h=3 * 11, w=20, seed=2)
    blob = write_ci16_tiff(None, d, rps=4)
    zp = tmp_path / "S1X.zip"
    with zipfile.ZipFile(zp, "w", compression=comp) as z:
        z.writestr("S1X.SAFE/manifest.safe", b"<xml/>")
        z.writestr("S1X.SAFE/measurement/s1a-iw1-slc-vv.tiff", blob)
    bursts = list(T.stream_tiff_from_zip(str(zp), "S1X.SAFE/measurement/s1a-iw1-slc-vv.tiff",
                                         20, 11, 3))
    assert len(bursts) == 3
    for k, b in enumerate(bursts):
        np.testing.assert_array_equal(b, d[11 * k:11 * (k + 1)])
    # a plain file through the same call
    fn = tmp_path / "m.tiff"
    fn.write_bytes(blob)
    got = np.concatenate(list(T.stream_tiff_from_zip(str(fn), None, 20, 11, 3)))
    np.testing.assert_array_equal(got, d)


def test_dimension_check_matches_reference_message(tmp_path):
    fn = write_ci16_tiff(str(tmp_path / "m.tiff"), _data(h=20, w=8))
    with pytest.raises(RuntimeError, match=r"TIFF dimensions \(8x20\) do not match DB metadata "
                                           r"\(nrange=8, lines_per_burst=7, nbursts=3, "
                                           r"expected_h=21\)"):
        next(T.stream_tiff_from_zip(fn, None, 8, 7, 3))
    with pytest.raises(RuntimeError, match="do not match"):
        next(T.stream_tiff_from_zip(fn, None, 9, 10, 2))


@pytest.mark.parametrize("kw,msg", [
    (dict(compression=5), "compressed"),
    (dict(sample_format=6), "complex int16"),
    (dict(encoding="pair", sample_format=3), "complex int16"),
])
def test_refused_formats(tmp_path, kw, msg):
    fn = write_ci16_tiff(str(tmp_path / "m.tiff"), _data(h=4, w=4), **kw)
    with pytest.raises(T.TiffError, match=msg):
        T.open_measurement(fn)


def test_not_a_tiff_and_truncated(tmp_path):
    p = tmp_path / "x.tiff"
    p.write_bytes(b"PK\x03\x04" + b"\0" * 20)
    with pytest.raises(T.TiffError, match="not a TIFF"):
        T.open_measurement(str(p))
    blob = write_ci16_tiff(None, _data(h=16, w=8), rps=2)
    # strip data cut off: the IFD sits after the pixels, so drop bytes in
    # the middle by pointing a strip past the end instead
    good = tmp_path / "g.tiff"
    good.write_bytes(blob)
    with T.open_measurement(str(good)) as tif:
        tif.offsets[3] = len(blob)
        with pytest.raises(T.TiffError, match="outside the file"):
            tif.read_rows(6, 2)
