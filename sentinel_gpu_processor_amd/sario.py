"""Drop-in for the ``s1proc.sario`` pieces on the interferogram path:
cropped-image files (.gslc / per-burst .int: int32[64] header + row-major
data), subswath / burst groups, rmg files and look helpers.

File formats follow sario.hpp:108-123 and sario.py:19-261, 379-420:
header[0:6] = nrow0, ncol0, left, top, right, bottom (the window of the full
grid the data cover); the payload is (bottom-top) x (right-left) values.
The look helpers run on the GPU (libs1ifg); the file handling is host code.
"""

import os
from typing import List

import numpy as np

NHEAD = 64  # sario.py:19, sario.hpp:14


def read_header(filename):
    with open(filename, "rb") as f:
        h = np.fromfile(f, dtype=np.int32, count=NHEAD)
    if h.size != NHEAD:
        raise OSError(f"{filename}: truncated header ({h.size} of {NHEAD} words)")
    return h


def save_cropped(filename, data, nrow0, ncol0, left, top, right, bottom, extra=None):
    """Write int32[64] header + data (sario.hpp:108-123).  Raises on error
    (the reference prints and returns, sario.hpp:115-118)."""
    data = np.ascontiguousarray(data)
    if data.shape != (bottom - top, right - left):
        raise ValueError(f"data shape {data.shape} != window {(bottom - top, right - left)}")
    h = np.zeros(NHEAD, np.int32)
    h[:6] = [nrow0, ncol0, left, top, right, bottom]
    if extra is not None:
        h[6:6 + len(extra)] = extra
    tmp = filename + ".part"
    with open(tmp, "wb") as f:
        h.tofile(f)
        data.tofile(f)
    os.replace(tmp, filename)


class CroppedImage:
    """A window [top, bottom) x [left, right) of an nrow0 x ncol0 grid
    (sario.py:23-215)."""

    def __init__(self, nrow0, ncol0, left, top, right, bottom, data, dtype=np.complex64):
        self.nrow0 = int(nrow0)
        self.ncol0 = int(ncol0)
        self.left = int(left)
        self.top = int(top)
        self.right = int(right)
        self.bottom = int(bottom)
        self.data = data
        self.nrow = self.bottom - self.top
        self.ncol = self.right - self.left
        self.dtype = dtype

    @classmethod
    def from_file(cls, filename, load_data=False, dtype=np.complex64):
        h = read_header(filename)
        nrow0, ncol0, left, top, right, bottom = (int(v) for v in h[:6])
        data = filename
        if load_data:
            n = (bottom - top) * (right - left)
            with open(filename, "rb") as f:
                f.seek(NHEAD * 4)
                d = np.fromfile(f, count=n, dtype=dtype)
            if d.size != n:
                raise OSError(f"{filename}: truncated payload ({d.size} of {n} values)")
            data = d.reshape(bottom - top, right - left)
        return cls(nrow0, ncol0, left, top, right, bottom, data, dtype)

    def save(self, filename):
        save_cropped(filename, self.data, self.nrow0, self.ncol0, self.left, self.top,
                     self.right, self.bottom)

    def resample(self, left, top, right, bottom):
        """Copy into another window, zero outside the overlap (sario.py:85-118)."""
        if self.data is None or isinstance(self.data, str):
            raise RuntimeError("Data are not loaded for this CroppedImage")
        res = np.zeros((bottom - top, right - left), dtype=self.data.dtype)
        ol, orr = max(left, self.left), min(right, self.right)
        ot, ob = max(top, self.top), min(bottom, self.bottom)
        if ol < orr and ot < ob:
            res[ot - top:ob - top, ol - left:orr - left] = \
                self.data[ot - self.top:ob - self.top, ol - self.left:orr - self.left]
        return res

    def load_data(self, left=None, top=None, right=None, bottom=None):
        """Load (part of) the image, as sario.py:120-189."""
        left = 0 if left is None else left
        top = 0 if top is None else top
        right = self.ncol0 if right is None else right
        bottom = self.nrow0 if bottom is None else bottom
        if isinstance(self.data, np.ndarray):
            if (left, top, right, bottom) == (self.left, self.top, self.right, self.bottom):
                return self.data
            return self.resample(left, top, right, bottom)
        if isinstance(self.data, str):
            ot, ob = max(top, self.top), min(bottom, self.bottom)
            ob = max(ob, ot)
            with open(self.data, "rb") as f:
                f.seek(NHEAD * 4 + (ot - self.top) * self.ncol * np.dtype(self.dtype).itemsize)
                d = np.fromfile(f, count=(ob - ot) * self.ncol, dtype=self.dtype)
            d = d.reshape(ob - ot, self.ncol)
            if (left, top, right, bottom) == (self.left, self.top, self.right, self.bottom):
                return d
            tmp = CroppedImage(self.nrow0, self.ncol0, self.left, ot, self.right, ob, d)
            return tmp.resample(left, top, right, bottom)
        return None

    def __str__(self):
        return (f"nrow0: {self.nrow0}\nncol0: {self.ncol0}\nleft: {self.left}\n"
                f"top: {self.top}\nright: {self.right}\nbottom: {self.bottom}\n"
                f"nrow: {self.nrow}\nncol: {self.ncol}")


class Subswath:
    """Bursts of one subswath sorted by top + bottom (sario.py:218-253)."""

    def __init__(self, burst_files: List[str]):
        self.bursts = [CroppedImage.from_file(s) for s in burst_files]
        self.size = len(self.bursts)
        self.nrow0 = self.bursts[0].nrow0 if self.size else 0
        self.ncol0 = self.bursts[0].ncol0 if self.size else 0
        self.sort()

    def bounds(self):
        if self.is_empty():
            return None
        b0, b1 = self.bursts[0], self.bursts[-1]
        return [min(b0.left, b1.left), min(b0.top, b1.top), max(b0.right, b1.right),
                max(b0.bottom, b1.bottom)]

    def sort(self):
        key = np.array([b.top + b.bottom for b in self.bursts], dtype=int)
        self.bursts = [self.bursts[i] for i in np.argsort(key, kind="stable")]

    def is_empty(self):
        return self.size == 0


class BurstGroup:
    """Bursts of one date split by subswath name iw1/iw2/iw3 (sario.py:256-261)."""

    def __init__(self, burst_files: List[str]):
        self.subswaths = [Subswath([s for s in burst_files if f"iw{k}" in s]) for k in (1, 2, 3)]


def readc(f, nrg):
    """rmg reader (sario.py:379-384): rows of [real nrg | imag nrg] float32."""
    data = np.fromfile(f, dtype="float32")
    naz = len(data) // nrg // 2
    data = data.reshape(naz, 2 * nrg)
    return data[:, :nrg] + 1j * data[:, nrg:]


def savec(c, filename):
    """rmg writer (sario.py:387-394)."""
    naz, nrg = c.shape
    out = np.zeros((naz, 2 * nrg), dtype="float32")
    out[:, :nrg] = np.real(c)
    out[:, nrg:] = np.imag(c)
    out.tofile(filename)


def cpxlooks(imgbg, rowlook, collook):
    """Look means of a complex (or real) image on the GPU (sario.py:397-413)."""
    from . import kernels as K
    from .runtime import context, device
    a = np.asarray(imgbg)
    real = not np.iscomplexobj(a)
    t = K.from_numpy(a.astype(np.complex64), device())
    out = K.to_numpy(K.looks(context(), t, rowlook, collook))
    return out.real.astype(np.float32) if real else out


def powlooks(imgbg, rowlook, collook):
    """Power looks on the GPU (sario.py:416-420): mean |s|^2 over the look
    cell; at 1x1 the reference returns |s|."""
    from . import kernels as K
    from .runtime import context, device
    t = K.from_numpy(np.asarray(imgbg, np.complex64), device())
    amp = K.to_numpy(K.amplitude(context(), t, rowlook, collook))
    if rowlook == 1 and collook == 1:
        return amp * amp  # amp = sqrt|s|
    return amp * amp
