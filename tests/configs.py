"""Benchmark-sized synthetic workloads for the GPU parity tests, and the
oracle on crops of them (test infrastructure).

The full-size scenes are generated on the GPU (``fidelity="fast"``); the
oracle runs on a crop of the target grid, reading only the window of each
burst the crop's pixels sample, so its cost follows the crop, not the
burst."""

import functools

import numpy as np

from oracle import geom as G
from oracle import ifg as I
from sentinel_gpu_processor_amd import synthetic as S


@functools.lru_cache(maxsize=2)
def c2_scene(tops="constant", baseline=(120.0, 30.0)):
    """The 1500 x 25000 pair: fractal DEM 2000 x 3000 (0-2500 m), secondary
    misregistered by (0.3, 0.05) px plus ramps (0.2, 0.1), SLCs on cuda:0."""
    c = S.CONFIGS["c2"]
    return S.make_scene(lines=c["lines"], samples=c["samples"], dem_shape=c["dem"],
                        dem_kind=c["dem_kind"], hmax=c["hmax"], seed=0, misreg=c["offset"],
                        misreg_ramp=(0.2, 0.1), fidelity="fast", device="cuda:0", tops=tops,
                        baseline=baseline)


def crop_grid(grid, r0, c0, nr, nc):
    return dict(grid, nrow=nr, ncol=nc, lat0=grid["lat0"] + r0 * grid["dlat"],
                lon0=grid["lon0"] + c0 * grid["dlon"])


def _window_deramp(slc, setup, rg, az, valid):
    """The deramped burst (complex64, full shape) filled only over the
    window the valid sampling positions (rg, az) read with an 8-tap kernel
    (zeros elsewhere)."""
    L, N = len(setup["eta"]), len(setup["kt"])
    out = np.zeros((L, N), np.complex64)
    if not valid.any():
        return out
    a0 = max(int(np.floor(az[valid].min())) - 4, 0)
    a1 = min(int(np.floor(az[valid].max())) + 6, L)
    r0 = max(int(np.floor(rg[valid].min())) - 4, 0)
    r1 = min(int(np.floor(rg[valid].max())) + 6, N)
    sub = dict(setup, eta=setup["eta"][a0:a1], kt=setup["kt"][r0:r1],
               etaref=setup["etaref"][r0:r1])
    data = np.asarray(slc[a0:a1, r0:r1])
    out[a0:a1, r0:r1] = G.deramp(data, sub).astype(np.complex64)
    return out


def window_deramped(orbit, burst, slc_host, grid, heights, offpoly=None):
    """(setup, deramped, geo): the oracle's set-up, the burst deramped over
    the window a (crop) grid's pixels sample (complex64, full shape, zeros
    elsewhere) and the grid's geo2rdr solution."""
    setup = G.deramp_setup(orbit, burst)
    geo = G.geo2rdr_grid(orbit, burst, grid, heights)
    drg, daz = G.offset_poly_shift(offpoly, geo["rgoff"], geo["azoff"], burst)
    rg, az = geo["rgoff"] + drg, geo["azoff"] + daz
    valid = ((rg >= burst["first_valid_sample"]) & (rg < burst["last_valid_sample"])
             & (az >= burst["first_valid_line"]) & (az < burst["last_valid_line"]))
    return setup, _window_deramp(slc_host, setup, rg, az, valid), geo


def oracle_geocode_crop(orbit, burst, slc_host, grid, heights, interp, offpoly=None):
    """The oracle's deramp + reproject of one burst onto a (crop) grid."""
    setup, dz, geo = window_deramped(orbit, burst, slc_host, grid, heights, offpoly)
    out, _ = G.reproject(orbit, burst, setup, grid, heights, dz, interp=interp, offpoly=offpoly,
                         geo=geo)
    return out


def oracle_pair_crop(sc, ref_host, sec_host, grid, interp, rl=1, cl=4):
    """Oracle pair products on a crop grid: (ifg_ml, p1, p2, gr, gs)."""
    h = G.target_heights(grid, sc["dem"])
    gr = oracle_geocode_crop(sc["orbit_ref"], sc["burst_ref"], ref_host, grid, h, interp)
    gs = oracle_geocode_crop(sc["orbit_sec"], sc["burst_sec"], sec_host, grid, h, interp,
                             sc["offpoly"])
    ifg = I.cpx_looks(I.conj_mul(gr, gs), rl, cl)
    p1 = I.cpx_looks(np.abs(gr) ** 2, rl, cl).real
    p2 = I.cpx_looks(np.abs(gs) ** 2, rl, cl).real
    return ifg, p1, p2, gr, gs
