"""Goldstein filtering of wrapped interferogram files: the ``goldstein``
executable's contract (goldstein.cu:410-524) and the filter step of
``s1proc.phase_correction`` (phase_correction.py:71-85), in-process on the
GPU.  The tropospheric correction of phase_correction.py is out of scope.
"""

import os

import numpy as np

from . import kernels as K
from .runtime import context, device


def process_input(in_str):
    """A single ``.int`` file, or a text file listing inputs (goldstein.cu:60-81)."""
    if in_str.endswith(".int"):
        return [in_str], True
    if not os.path.exists(in_str):
        return [], False
    with open(in_str) as f:
        return [l.rstrip("\n") for l in f if l.strip()], False


def goldstein_filter(in_str, nrow, ncol, n_win=32, alpha=0.5, out_str="", mode="fp16_split"):
    """Filter every input (raw cf32 nrow x ncol); outputs go to
    ``<in><out_str>`` (default ``.filt``), or ``filtered.int`` for a single
    ``.int`` argument (goldstein.cu:469-493).  Returns the output paths."""
    files, is_ifg = process_input(in_str)
    outs = []
    if not files:
        print("Input list is empty, nothing to filter.")
        return outs
    ctx = context()
    dev = device()
    for f in files:
        if is_ifg and not out_str:
            out = "filtered.int"
        else:
            out = f + (out_str or ".filt")
        x = np.fromfile(f, dtype=np.complex64, count=nrow * ncol)
        if x.size != nrow * ncol:
            raise OSError(f"{f}: expected {nrow * ncol} complex values, found {x.size}")
        y = K.goldstein(ctx, K.from_numpy(x.reshape(nrow, ncol), dev), alpha, n_win, mode)
        K.to_numpy(y).tofile(out)
        outs.append(out)
    return outs


def phase_correction(ifg_path=None, config="config.yaml", verbose=False, mode="fp16_split"):
    """Goldstein-filter every .int of ifg_path when ``filter.enable`` is set
    (phase_correction.py:12-85, filter branch)."""
    import glob
    from .config import load_config
    from .geocoordinates import GeoCoordinates
    cfg = load_config(config)
    if ifg_path is None:
        ifg_path = cfg.io.ifg_path
    files = sorted(glob.glob(os.path.join(ifg_path, "*.int")))
    if not os.path.exists(cfg.io.multilook_rsc_file):
        GeoCoordinates(cfg.io.rsc_file).take_look(cfg.proc.rowlook, cfg.proc.collook) \
            .save_as_rsc(cfg.io.multilook_rsc_file)
    rsc = GeoCoordinates(cfg.io.multilook_rsc_file)
    outs = []
    if cfg.filter.enable:
        p = cfg.filter.parameters
        os.makedirs("temp_filter", exist_ok=True)
        lst = os.path.join("temp_filter", "ifg_list")
        with open(lst, "w") as f:
            f.write("\n".join(files))
        outs = goldstein_filter(lst, rsc.nlat, rsc.nlon, p.window_size, p.goldstein_alpha,
                                mode=mode)
    return outs
