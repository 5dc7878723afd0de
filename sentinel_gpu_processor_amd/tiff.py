"""Reader of Sentinel-1 SLC measurement TIFFs (complex int16 rasters),
plain or inside the product's .zip, streamed one burst at a time.

The reference opens the measurement through GDAL's /vsizip/ driver and
streams each burst as complex float to the geo2rdr executable's stdin
(sentinel_scene.py:164-207).  GDAL is not part of this package: this module
parses the TIFF structure itself and returns each burst in its stored
format, interleaved CInt16, so that the burst crosses PCIe at 4 bytes per
sample and is converted on the GPU inside the deramp (s1ifg_deramp_ci16).

Supported: classic TIFF and BigTIFF, either byte order, uncompressed,
strips (contiguous runs of strips are read in one call) or tiles, and the
two encodings of a complex int16 pixel (SampleFormat 5 with 32 bits per
sample and one sample per pixel, or SampleFormat 2 with two 16-bit samples
per pixel, chunky).  Anything else raises ``TiffError``.
"""

import os
import struct
import zipfile

import numpy as np

# tag numbers (TIFF 6.0)
_WIDTH, _LENGTH, _BITS, _COMPRESSION = 256, 257, 258, 259
_STRIP_OFFSETS, _SPP, _ROWS_PER_STRIP, _STRIP_COUNTS = 273, 277, 278, 279
_PLANAR, _TILE_W, _TILE_L, _TILE_OFFSETS, _TILE_COUNTS = 284, 322, 323, 324, 325
_SAMPLE_FORMAT = 339
# field type -> (struct code, size)
_FIELD = {1: ("B", 1), 3: ("H", 2), 4: ("I", 4), 6: ("b", 1), 8: ("h", 2), 9: ("i", 4),
          16: ("Q", 8), 17: ("q", 8)}


class TiffError(ValueError):
    pass


class _Window:
    """A byte range [base, base + size) of an open file."""

    def __init__(self, f, base=0, size=None, close=None):
        self.f, self.base, self.close_fn = f, base, close
        if size is None:
            f.seek(0, os.SEEK_END)
            size = f.tell() - base
        self.size = size

    def read_at(self, off, n):
        if off < 0 or off + n > self.size:
            raise TiffError(f"byte range [{off}, {off + n}) outside the file ({self.size} bytes)")
        self.f.seek(self.base + off)
        b = self.f.read(n)
        if len(b) != n:
            raise TiffError(f"short read at {off}: {len(b)} of {n} bytes")
        return b

    def close(self):
        if self.close_fn is not None:
            self.close_fn()
            self.close_fn = None


class TiffImage:
    """First image (IFD) of a complex int16 TIFF.

    ``width`` x ``height`` pixels; ``read_rows(r0, n)`` returns rows
    [r0, r0 + n) as an int16 array of shape (n, width, 2) (re, im), in the
    host's byte order."""

    def __init__(self, src):
        self._w = src
        head = src.read_at(0, 16 if src.size >= 16 else 8)
        bo = head[:2]
        if bo == b"II":
            self.e = "<"
        elif bo == b"MM":
            self.e = ">"
        else:
            raise TiffError("not a TIFF file (byte-order mark)")
        magic = struct.unpack(self.e + "H", head[2:4])[0]
        if magic == 42:
            self.big = False
            ifd = struct.unpack(self.e + "I", head[4:8])[0]
        elif magic == 43:
            self.big = True
            osz, _, ifd = struct.unpack(self.e + "HHQ", head[4:16])
            if osz != 8:
                raise TiffError("BigTIFF with offsets that are not 8 bytes")
        else:
            raise TiffError(f"not a TIFF file (magic {magic})")
        tags = self._ifd(ifd)

        def one(tag, default=None):
            v = tags.get(tag)
            if v is None:
                if default is None:
                    raise TiffError(f"TIFF tag {tag} missing")
                return default
            return v[0]

        self.width = int(one(_WIDTH))
        self.height = int(one(_LENGTH))
        if one(_COMPRESSION, 1) != 1:
            raise TiffError(f"compressed TIFF (Compression = {one(_COMPRESSION)}); only "
                            "uncompressed measurement rasters are supported")
        bits = tags.get(_BITS, (1,))
        spp = int(one(_SPP, 1))
        fmt = tags.get(_SAMPLE_FORMAT, (1,))
        if spp == 1 and bits[0] == 32 and fmt[0] == 5:
            pass  # complex int16, one 32-bit sample per pixel
        elif spp == 2 and all(b == 16 for b in bits) and all(f == 2 for f in fmt):
            if one(_PLANAR, 1) != 1:
                raise TiffError("planar (separate) complex samples are not supported")
        else:
            raise TiffError(f"not a complex int16 raster (BitsPerSample {bits}, "
                            f"SamplesPerPixel {spp}, SampleFormat {fmt})")
        self.dtype = np.dtype(self.e + "i2")
        self.row_bytes = 4 * self.width
        if _TILE_OFFSETS in tags:
            self.tw, self.tl = int(one(_TILE_W)), int(one(_TILE_L))
            self.offsets = np.asarray(tags[_TILE_OFFSETS], np.int64)
            self.counts = np.asarray(tags[_TILE_COUNTS], np.int64)
            self.tiles_across = -(-self.width // self.tw)
            need = self.tiles_across * -(-self.height // self.tl)
            if len(self.offsets) < need or np.any(self.counts[:need] < 4 * self.tw * self.tl):
                raise TiffError("tile table does not cover the image")
        else:
            self.tw = None
            self.rps = min(int(one(_ROWS_PER_STRIP, self.height)), self.height)
            self.offsets = np.asarray(tags[_STRIP_OFFSETS], np.int64)
            self.counts = np.asarray(tags[_STRIP_COUNTS], np.int64)
            need = -(-self.height // self.rps)
            if len(self.offsets) < need:
                raise TiffError("strip table does not cover the image")
            rows = np.minimum(self.rps, self.height - self.rps * np.arange(need))
            if np.any(self.counts[:need] < rows * self.row_bytes):
                raise TiffError("strip shorter than its rows")

    def _ifd(self, off):
        e = self.e
        if self.big:
            n = struct.unpack(e + "Q", self._w.read_at(off, 8))[0]
            ent, base, inline = 20, off + 8, 8
        else:
            n = struct.unpack(e + "H", self._w.read_at(off, 2))[0]
            ent, base, inline = 12, off + 2, 4
        raw = self._w.read_at(base, n * ent)
        tags = {}
        for k in range(n):
            rec = raw[k * ent:(k + 1) * ent]
            if self.big:
                tag, typ, cnt = struct.unpack(e + "HHQ", rec[:12])
                val = rec[12:20]
            else:
                tag, typ, cnt = struct.unpack(e + "HHI", rec[:8])
                val = rec[8:12]
            if typ not in _FIELD:
                continue  # ASCII, rationals, ...: nothing this reader needs
            code, size = _FIELD[typ]
            nbytes = size * cnt
            if nbytes > inline:
                voff = struct.unpack(e + ("Q" if self.big else "I"), val)[0]
                val = self._w.read_at(voff, nbytes)
            tags[tag] = struct.unpack(e + code * cnt, val[:nbytes])
        return tags

    def read_rows(self, r0, n):
        if r0 < 0 or n < 0 or r0 + n > self.height:
            raise TiffError(f"rows [{r0}, {r0 + n}) outside the image (height {self.height})")
        out = np.empty((n, self.width, 2), self.dtype)
        if self.tw is None:
            self._strips(r0, n, out)
        else:
            self._tiles(r0, n, out)
        if self.e == ">":
            out = out.astype(np.int16)
        return out

    def _strips(self, r0, n, out):
        rb, rps = self.row_bytes, self.rps
        flat = out.reshape(n, rb // 2)
        s0, s1 = r0 // rps, (r0 + n - 1) // rps if n else r0 // rps - 1
        s = s0
        while s <= s1:
            # run of strips stored back to back: one read
            e = s
            while (e < s1 and self.offsets[e + 1] == self.offsets[e] + rps * rb
                   and self.counts[e] >= rps * rb):
                e += 1
            a = max(r0, s * rps)
            b = min(r0 + n, (e + 1) * rps, self.height)
            off = int(self.offsets[s]) + (a - s * rps) * rb
            buf = self._w.read_at(off, (b - a) * rb)
            flat[a - r0:b - r0] = np.frombuffer(buf, self.dtype).reshape(b - a, rb // 2)
            s = e + 1

    def _tiles(self, r0, n, out):
        tw, tl = self.tw, self.tl
        for tr in range(r0 // tl, (r0 + n - 1) // tl + 1 if n else 0):
            a, b = max(r0, tr * tl), min(r0 + n, (tr + 1) * tl)
            for tc in range(self.tiles_across):
                k = tr * self.tiles_across + tc
                buf = self._w.read_at(int(self.offsets[k]), 4 * tw * tl)
                tile = np.frombuffer(buf, self.dtype).reshape(tl, tw, 2)
                c0, c1 = tc * tw, min(self.width, (tc + 1) * tw)
                out[a - r0:b - r0, c0:c1] = tile[a - tr * tl:b - tr * tl, :c1 - c0]

    def close(self):
        self._w.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def open_measurement(path, member=None):
    """Open a measurement TIFF: a file on disk, or ``member`` of the zip
    archive ``path``.  A stored (uncompressed) member is read in place; a
    deflated one through the zip module's stream."""
    if member is None:
        f = open(path, "rb")
        return TiffImage(_Window(f, close=f.close))
    z = zipfile.ZipFile(path)
    try:
        info = z.getinfo(member)
        if info.compress_type == zipfile.ZIP_STORED:
            f = open(path, "rb")
            f.seek(info.header_offset)
            lh = f.read(30)
            if lh[:4] != b"PK\x03\x04":
                f.close()
                raise TiffError(f"bad local header for {member} in {path}")
            nlen, xlen = struct.unpack("<HH", lh[26:30])
            base = info.header_offset + 30 + nlen + xlen
            z.close()
            return TiffImage(_Window(f, base, info.file_size, close=f.close))
        m = z.open(member)

        def close():
            m.close()
            z.close()

        return TiffImage(_Window(m, 0, info.file_size, close=close))
    except BaseException:
        z.close()
        raise


def stream_tiff_from_zip(zip_path, tiff_name, nrange, lines_per_burst, nbursts):
    """Bursts of a measurement TIFF, one (lines_per_burst, nrange, 2) int16
    array at a time, after the reference's dimension check against the
    database (sentinel_scene.py:164-207).  ``tiff_name`` None: ``zip_path``
    is the TIFF itself."""
    with open_measurement(zip_path, tiff_name) as tif:
        w, h = tif.width, tif.height
        expected_h = lines_per_burst * nbursts
        if w != nrange or h != expected_h:
            raise RuntimeError(
                f"TIFF dimensions ({w}x{h}) do not match DB metadata "
                f"(nrange={nrange}, lines_per_burst={lines_per_burst}, "
                f"nbursts={nbursts}, expected_h={expected_h})")
        for k in range(nbursts):
            yield tif.read_rows(k * lines_per_burst, lines_per_burst)
