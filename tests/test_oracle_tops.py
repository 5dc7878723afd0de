"""Pins of the oracle's TOPS terms with range-dependent FM rate and Doppler
centroid and time-varying records (geo2rdr.cu:465-549 set-up, the reramp
of geo2rdr.cu:210-219):

- the record interpolation (closest_sample rules, geo2rdr.cu:485-527)
  against numpy's linear interpolation with end clamping, with the burst
  centre before, between and after the records and on a record;
- kt / etaref per range sample against the synthetic generator's own
  formulas (numpy polynomial evaluation);
- the reramp and flattening phase in closed form: a burst holding only its
  TOPS ramp, exp(+i pi kt (eta - etaref)^2), deramps to 1, and its geocoded
  value at a target point is exp(i (4 pi R / lambda + pi kt(tau)
  (eta - etaref(tau))^2)) at that point's zero-Doppler time and range, which
  the straight-line orbit gives in closed form."""

import numpy as np
import pytest

from oracle import geom as G
from sentinel_gpu_processor_amd import synthetic as S


def _scene(tops="varying", misreg=None, lines=128, samples=512, seed=0):
    return S.make_scene(lines=lines, samples=samples, dem_shape=(64, 64), hmax=300.0, seed=seed,
                        misreg=misreg, misreg_ramp=(0.2, 0.1) if misreg else (0.0, 0.0),
                        tops=tops, fidelity="none")


@pytest.mark.parametrize("shift", [0.0, -20.0, 20.0, 0.9, 1.3])
def test_record_interpolation_matches_numpy(shift):
    """shift -20 / +20 s: the centre after the last / before the first
    record; 0.9 and 1.3 put it exactly on an FM / Doppler record."""
    tc = 43200.0
    fm, dc = S.tops_records(tc, 5.3e-3, "varying", shift=-shift)
    for rec in (fm, dc):
        got = G.interp_record(rec, tc)
        want = S.record_at(rec, tc)
        assert np.allclose(got, want, rtol=1e-14, atol=0)
    if shift == 0.0:
        # between records: not equal to any single record
        for rec in (fm, dc):
            v = G.interp_record(rec, tc)
            assert not any(np.allclose(v, r[1:]) for r in rec)


@pytest.mark.parametrize("shift", [0.0, -20.0, 20.0])
def test_deramp_setup_varying_matches_generator(shift):
    sc = _scene()
    b = dict(sc["burst_ref"])
    b["fm"], b["dc"] = S.tops_records(43200.0, b["fm"][1, 1], "varying", shift=shift)
    su = G.deramp_setup(sc["orbit_ref"], b)
    kt, etaref, eta = S.deramp_coefficients(sc["line_orbits"][0], b)
    assert np.max(np.abs(su["kt"] - kt)) < 1e-9 * np.abs(kt).max()
    assert np.max(np.abs(su["etaref"] - etaref)) < 1e-12
    assert np.array_equal(su["eta"], eta)
    # the terms are not degenerate: kt and etaref vary across the swath
    assert np.ptp(kt) > 0.5
    assert np.ptp(etaref) > 1e-4


def ramp_burst(line_orbit, burst):
    """A burst holding only its TOPS ramp (complex64)."""
    return np.exp(1j * S.ramp(line_orbit, burst)).astype(np.complex64)


def closed_form_phase(line_orbit, burst, grid, dem, offpoly=None):
    """Phase of the geocoded ramp burst at every target pixel, from the
    straight-line zero-Doppler solution and the generator's TOPS rates;
    also the sampling positions (rg, az) with the offset polynomial."""
    r = np.arange(grid["nrow"], dtype=np.float64)[:, None]
    c = np.arange(grid["ncol"], dtype=np.float64)[None, :]
    lat = grid["lat0"] + r * grid["dlat"] + 0 * c
    lon = grid["lon0"] + c * grid["dlon"] + 0 * r
    h = S.dem_height(dem, lat, lon)
    X = S.llh2xyz(lat, lon, h)
    t, R = line_orbit.zero_doppler(X)
    R0 = burst["slant_range_time"] * S.SOL / 2
    dr = S.SOL / 2 / burst["range_sampling_rate"]
    dtaz = 1.0 / burst["prf"]
    rg = (R - R0) / dr
    az = (t - burst["t_start"]) / dtaz
    drg = daz = 0.0
    if offpoly is not None:
        p = offpoly
        xs, ys = rg / burst["samples"], az / burst["lines"]
        drg = p[0] + p[1] * xs + p[2] * ys
        daz = p[3] + p[4] * xs + p[5] * ys
    tau = 2 * (R + drg * dr) / S.SOL
    eta = -burst["lines"] * dtaz / 2 + (t - burst["t_start"]) + daz * dtaz
    kt, etaref = S.tops_rates(line_orbit, burst, tau)
    ph = 4 * np.pi / burst["wavelength"] * R + np.pi * kt * (eta - etaref) ** 2
    return ph, rg + drg, az + daz


@pytest.mark.parametrize("interp", ["bilinear", "sinc"])
@pytest.mark.parametrize("tops,misreg", [("varying", None), ("varying", (0.3, 0.05)),
                                         ("constant", (0.3, 0.05))])
def test_reramp_closed_form(interp, tops, misreg):
    sc = _scene(tops, misreg)
    k = 1 if misreg else 0  # the secondary carries the offset polynomial
    orbit = (sc["orbit_ref"], sc["orbit_sec"])[k]
    lorb = sc["line_orbits"][k]
    b = (sc["burst_ref"], sc["burst_sec"])[k]
    op = sc["offpoly"] if k else None
    su = G.deramp_setup(orbit, b)
    dz = G.deramp(ramp_burst(lorb, b), su)
    assert np.max(np.abs(dz - 1)) < 1e-6  # the ramp is what the deramp removes
    h = G.target_heights(sc["grid"], sc["dem"])
    out, _ = G.reproject(orbit, b, su, sc["grid"], h, dz.astype(np.complex64), interp=interp,
                         offpoly=op)
    ph, rg, az = closed_form_phase(lorb, b, sc["grid"], sc["dem"], op)
    v = out != 0
    assert v.mean() > 0.3
    assert np.allclose(np.abs(out[v]), 1, atol=2e-6)
    d = np.angle(out[v] * np.exp(-1j * ph[v]))
    assert np.max(np.abs(d)) < 1e-5
    # the reramp term is large here: a wrong sign or time origin would show
    eta_span = b["lines"] / b["prf"] / 2
    assert np.pi * 1700 * eta_span ** 2 > 50
