"""Shared synthetic scenes and oracle products for the tests (cached per
process)."""

import functools

import numpy as np

from oracle import geom as G
from oracle import ifg as I
from sentinel_gpu_processor_amd import synthetic as S


@functools.lru_cache(maxsize=None)
def scene(name="small", seed=0):
    if name == "small":
        return S.make_scene(lines=128, samples=512, dem_shape=(64, 64), hmax=300.0, seed=seed)
    if name == "small_misreg":
        return S.make_scene(lines=128, samples=512, dem_shape=(64, 64), hmax=300.0, seed=seed,
                            misreg=(0.3, 0.05), misreg_ramp=(0.2, 0.1))
    if name == "odd":
        return S.make_scene(lines=97, samples=301, dem_shape=(40, 50), hmax=200.0, seed=seed)
    if name == "wide":
        return S.make_scene(lines=150, samples=640, dem_shape=(48, 80), hmax=250.0, seed=seed,
                            misreg=(-0.2, 0.1))
    if name == "small_tops":
        # range-dependent FM rate and Doppler centroid, time-varying records
        return S.make_scene(lines=128, samples=512, dem_shape=(64, 64), hmax=300.0, seed=seed,
                            misreg=(0.3, 0.05), misreg_ramp=(0.2, 0.1), tops="varying")
    if name == "long_tops":
        # 1024 lines: the TOPS ramps span +-1.05 s of azimuth time
        return S.make_scene(lines=1024, samples=512, dem_shape=(96, 64), hmax=300.0, seed=seed,
                            misreg=(0.3, 0.05), misreg_ramp=(0.2, 0.1), tops="varying")
    if name == "c1":
        return S.make_scene(**_c1_kwargs(), seed=seed)
    raise KeyError(name)


def _c1_kwargs():
    c = S.CONFIGS["c1"]
    return dict(lines=c["lines"], samples=c["samples"], dem_shape=c["dem"], hmax=c["hmax"])


def dem_i16(dem):
    """The DEM as the reference stores it: int16 metres (rounded)."""
    return dict(dem, data=np.round(np.asarray(dem["data"], np.float64)).astype(np.int16))


@functools.lru_cache(maxsize=None)
def oracle_chain(name="small", interp="bilinear", seed=0, rl=1, cl=4, dem16=False):
    sc = scene(name, seed)
    b = sc["burst_ref"]
    dem = dem_i16(sc["dem"]) if dem16 else sc["dem"]
    h = G.target_heights(sc["grid"], dem)
    su = G.deramp_setup(sc["orbit_ref"], b)
    ss = G.deramp_setup(sc["orbit_sec"], sc["burst_sec"])
    dr = G.deramp(sc["ref_slc"], su)
    ds = G.deramp(sc["sec_slc"], ss)
    gr, geo_r = G.reproject(sc["orbit_ref"], b, su, sc["grid"], h, dr.astype(np.complex64),
                            interp=interp)
    gs, geo_s = G.reproject(sc["orbit_sec"], sc["burst_sec"], ss, sc["grid"], h,
                            ds.astype(np.complex64), interp=interp, offpoly=sc["offpoly"])
    ifg = I.cpx_looks(I.conj_mul(gr, gs), rl, cl)
    p1 = I.cpx_looks(np.abs(gr) ** 2, rl, cl).real
    p2 = I.cpx_looks(np.abs(gs) ** 2, rl, cl).real
    return dict(h=h, su=su, ss=ss, dr=dr, ds=ds, gr=gr, gs=gs, geo_r=geo_r, geo_s=geo_s,
                ifg=ifg, p1=p1, p2=p2)


def phase_rms(a, b, mask=None):
    d = np.angle(a * np.conj(b))
    if mask is not None:
        d = d[mask]
    return float(np.sqrt(np.mean(d * d)))
