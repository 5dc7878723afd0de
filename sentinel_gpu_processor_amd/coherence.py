"""Drop-in for ``s1proc.coherence``: multilooked amplitude mosaics (.amp) and
the look-cell coherence product (.cc, rmg layout) (coherence.py:14-180).

The pixel work (power looks, coherence) runs in HIP kernels; the mosaic
bookkeeping is host code.  The reference fills coherence values > 1 with
unseeded random numbers (coherence.py:143-144); here the fill is a
deterministic hash of the pixel index (``fill="hash"``, seed 0) by
default, or clamping / NaN on request.
"""

import glob
import os
from typing import Sequence

import numpy as np

from . import kernels as K
from .runtime import context, device
from .sario import CroppedImage, savec

try:
    from tqdm import tqdm
except Exception:  # pragma: no cover
    def tqdm(x, **k):
        return x


def multilook(slc_list: Sequence[str], outfile: str, rowlook: int, collook: int):
    """Amplitude mosaic sqrt(power looks) of the .gslc bursts of one date on
    the look-aligned grid; later bursts fill only pixels still 0
    (coherence.py:14-55)."""
    ctx = context()
    mm = None
    for i, f in enumerate(slc_list):
        ci = CroppedImage.from_file(f, load_data=True)
        left_sm = (ci.left + collook - 1) // collook
        top_sm = (ci.top + rowlook - 1) // rowlook
        right_sm = ci.right // collook
        bottom_sm = ci.bottom // rowlook
        if mm is None:
            mm = np.memmap(outfile, dtype=np.float32, mode="w+",
                           shape=(ci.nrow0 // rowlook, ci.ncol0 // collook))
        r0 = top_sm * rowlook - ci.top
        r1 = bottom_sm * rowlook - ci.top
        c0 = left_sm * collook - ci.left
        c1 = right_sm * collook - ci.left
        if r1 <= r0 or c1 <= c0:
            continue
        block = np.ascontiguousarray(ci.data[r0:r1, c0:c1])
        amp = K.to_numpy(K.amplitude(ctx, K.from_numpy(block, device()), rowlook, collook))
        if i == 0:
            mm[top_sm:bottom_sm, left_sm:right_sm] = amp
        else:
            old = mm[top_sm:bottom_sm, left_sm:right_sm]
            rep = (old == 0) & (amp != 0)
            old[rep] = amp[rep]
    if mm is not None:
        mm.flush()


def multilook_amp(slc_dir: str, amp_dir: str = "amp", rowlook: int = 1, collook: int = 1):
    """One .amp per date in slc_dir (coherence.py:58-83)."""
    os.makedirs(amp_dir, exist_ok=True)
    slc_list = glob.glob(os.path.join(slc_dir, "*.gslc"))
    dates = sorted(np.unique([os.path.basename(s)[0:8] for s in slc_list]))
    for d in tqdm(dates, desc="multilook"):
        out = os.path.join(amp_dir, f"{d}.amp")
        if os.path.exists(out):
            continue
        multilook(glob.glob(os.path.join(slc_dir, f"{d}*.gslc")), out, rowlook, collook)


def run_multilook_amp(config: str = "config.yaml"):
    from .config import load_config
    cfg = load_config(config)
    multilook_amp(slc_dir=cfg.io.slc_path, amp_dir=cfg.io.amp_path, rowlook=cfg.proc.rowlook,
                  collook=cfg.proc.collook)


def coherence_image(ifg, amp1, amp2, fill="hash", seed=0):
    """|ifg| / (amp1 + 1e-12) / (amp2 + 1e-12) on the GPU (coherence.py:142)."""
    ctx = context()
    dev = device()
    ifg = np.ascontiguousarray(ifg, np.complex64)
    p1 = np.ascontiguousarray(amp1, np.float32) ** 2
    p2 = np.ascontiguousarray(amp2, np.float32) ** 2
    coh = K.coherence(ctx, K.from_numpy(ifg, dev), K.from_numpy(p1, dev), K.from_numpy(p2, dev),
                      win=1, fill=fill, seed=seed)
    return K.to_numpy(coh)


def coherence(ifg_dir: str, amp_dir: str, nrow: int, ncol: int, rowlook: int, collook: int,
              fill: str = "hash", seed: int = 0):
    """A .cc file (rmg: [amp1*amp2 | coherence]) for every .int in ifg_dir
    (coherence.py:109-147; rowlook / collook are unused, as there)."""
    os.makedirs(amp_dir, exist_ok=True)
    prev = None
    amp1 = None
    for f in tqdm(sorted(glob.glob(os.path.join(ifg_dir, "*.int"))), desc="coherence"):
        out = f.replace(".int", ".cc")
        ifg = np.fromfile(f, dtype=np.complex64)
        b = os.path.basename(f)
        d1, d2 = b[0:8], b[9:17]
        if d1 != prev:
            amp1 = np.fromfile(os.path.join(amp_dir, f"{d1}.amp"), dtype=np.float32)
            prev = d1
        amp2 = np.fromfile(os.path.join(amp_dir, f"{d2}.amp"), dtype=np.float32)
        c = coherence_image(ifg.reshape(nrow, ncol), amp1.reshape(nrow, ncol),
                            amp2.reshape(nrow, ncol), fill=fill, seed=seed)
        savec((amp1 * amp2).reshape(nrow, ncol) + 1j * c, out)


def run_coherence(config: str = "config.yaml"):
    from .config import load_config
    from .geocoordinates import GeoCoordinates
    cfg = load_config(config)
    rsc = GeoCoordinates(cfg.io.multilook_rsc_file)
    multilook_amp(slc_dir=cfg.io.slc_path, amp_dir=cfg.io.amp_path, rowlook=cfg.proc.rowlook,
                  collook=cfg.proc.collook)
    coherence(ifg_dir=cfg.io.ifg_path, amp_dir=cfg.io.amp_path, nrow=rsc.nlat, ncol=rsc.nlon,
              rowlook=cfg.proc.rowlook, collook=cfg.proc.collook)
