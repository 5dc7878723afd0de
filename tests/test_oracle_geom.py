"""Pins of the FP64 geometry oracle: the reference's own compiled host code
(golden vectors), closed forms and invariants."""

import os

import numpy as np
import pytest

from oracle import geom as G

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ref_geometry.npz"))
A, E2 = G.RA, G.RE2


def test_llh2xyz_known_answers():
    assert np.allclose(G.llh2xyz(0, 0, 0), [A, 0, 0], atol=1e-9)
    assert np.allclose(G.llh2xyz(90, 0, 0), [0, 0, A * np.sqrt(1 - E2)], atol=1e-6)
    assert np.allclose(G.llh2xyz(0, 90, 1234.5), [0, A + 1234.5, 0], atol=1e-6)


def test_llh2xyz_matches_reference_build():
    llh = GOLD["llh_in"]
    got = G.llh2xyz(llh[:, 0], llh[:, 1], llh[:, 2])
    assert np.max(np.abs(got - GOLD["llh2xyz_out"])) < 1e-8


def test_xyz2llh_matches_reference_build_and_round_trips():
    got = G.xyz2llh(GOLD["llh2xyz_out"])
    ref = GOLD["xyz2llh_out"]
    assert np.max(np.abs(got[:, :2] - ref[:, :2])) < 1e-11
    assert np.max(np.abs(got[:, 2] - ref[:, 2])) < 1e-7
    llh = GOLD["llh_in"][3:]
    ok = np.abs(llh[:, 0]) < 89
    back = G.llh2xyz(got[3:, 0], got[3:, 1], got[3:, 2])
    assert np.max(np.linalg.norm(back[ok] - GOLD["llh2xyz_out"][3:][ok], axis=1)) < 2e-6


def test_intp_orbit_matches_reference_build():
    t, pos, vel = GOLD["orb_t"], GOLD["orb_pos"], GOLD["orb_vel"]
    x, v = G.intp_orbit(t, pos, vel, GOLD["intp_t"])
    assert np.max(np.abs(x - GOLD["intp_pos"])) < 1e-6
    assert np.max(np.abs(v - GOLD["intp_vel"])) < 1e-9


def test_hermite_reproduces_degree7_polynomials():
    rng = np.random.default_rng(1)
    c = rng.standard_normal((8, 3))
    t = np.arange(10, dtype=float) * 1.0

    def f(tt):
        return sum(c[k] * (tt[..., None] - 4.2) ** k for k in range(8))

    def df(tt):
        return sum(k * c[k] * (tt[..., None] - 4.2) ** (k - 1) for k in range(1, 8))

    tq = np.linspace(1.0, 7.9, 97)
    x, v = G.intp_orbit(t, f(t), df(t), tq)
    assert np.max(np.abs(x - f(tq))) < 1e-6 * np.max(np.abs(f(tq)))
    assert np.max(np.abs(v - df(tq))) < 1e-6 * np.max(np.abs(df(tq)))


def test_hermite_interpolates_knots_and_stencil_rule():
    t, pos, vel = GOLD["orb_t"], GOLD["orb_pos"], GOLD["orb_vel"]
    x, v = G.intp_orbit(t, pos, vel, t)
    assert np.max(np.abs(x - pos)) < 1e-6
    # truncated index clamped to [1, n-3]
    assert G.stencil_index(t, t[0] - 3.0) == 1
    assert G.stencil_index(t, t[-1] + 2.0) == len(t) - 3
    assert G.stencil_index(t, t[5] + 9.99) == 5


def test_circular_orbit_accuracy():
    from tests.golden.make_ref_vectors import circular_orbit
    t, pos, vel = GOLD["orb_t"], GOLD["orb_pos"], GOLD["orb_vel"]
    tq = np.linspace(t[1], t[-3], 200)
    x, v = G.intp_orbit(t, pos, vel, tq)
    xa, va = circular_orbit(tq)
    assert np.max(np.linalg.norm(x - xa, axis=1)) < 1e-3  # 10 s nodes, 7 km orbit radius
    assert np.max(np.linalg.norm(v - va, axis=1)) < 1e-5


def test_orbitrangetime_matches_reference_build():
    t, pos, vel = GOLD["orb_t"], GOLD["orb_pos"], GOLD["orb_vel"]
    tl, dr, it = G.orbitrangetime(t, pos, vel, GOLD["ort_X"], GOLD["ort_tmid"][0],
                                  GOLD["ort_xmid"], GOLD["ort_vmid"])
    assert np.max(np.abs(tl - GOLD["ort_t"])) < 1e-9
    assert np.max(np.abs(dr - GOLD["ort_dr"])) < 1e-6
    # curved orbit: the approximate derivative converges linearly (several steps)
    assert it.max() > 3


def test_orbitrangetime_straight_line_closed_form():
    rng = np.random.default_rng(2)
    x0 = np.array([-2.5e6, -4.6e6, 4.3e6])
    v = np.array([1200.0, 2300.0, 7000.0])
    t = np.arange(20.0)
    pos = x0 + t[:, None] * v
    vel = np.repeat(v[None], 20, 0)
    # targets whose zero-Doppler time lies inside the state-vector span
    ttrue = rng.uniform(3, 16, 50)
    off = rng.uniform(-5e5, 5e5, (50, 3))
    off -= (off @ v)[:, None] * v / (v @ v)
    X = x0 + ttrue[:, None] * v + off
    xm, vm = G.intp_orbit(t, pos, vel, 9.5)
    tl, dr, it = G.orbitrangetime(t, pos, vel, X, 9.5, xm, vm)
    tstar = ((X - x0) @ v) / (v @ v)
    assert np.max(np.abs(tl - tstar)) < 1e-9
    R = np.linalg.norm(X - (x0 + tstar[:, None] * v), axis=1)
    assert np.max(np.abs(np.linalg.norm(dr, axis=1) - R)) < 1e-6
    # zero Doppler to rounding (relative to |dr||v|)
    assert np.max(np.abs(np.sum(dr * v, axis=1)) / (np.linalg.norm(dr, axis=1) * 7.5e3)) < 1e-11
    assert it.max() <= 2


def test_bounds_matches_reference_build():
    a = GOLD["bounds_args"]
    t, pos, vel = GOLD["orb_t"], GOLD["orb_pos"], GOLD["orb_vel"]
    got = G.bounds(a[0], a[1], a[2], a[3], a[4], a[5], t, pos, vel)
    assert np.max(np.abs(np.array(got) - GOLD["bounds_out"])) < 1e-9


def test_closest_sample_and_record_interpolation():
    rec = np.array([[0.0, 1, 10, 0, 0], [10.0, 2, 20, 0, 0], [20.0, 3, 30, 0, 0]])
    assert G.closest_sample(rec[:, 0], 4.9) == 0
    assert G.closest_sample(rec[:, 0], 5.0) == 0  # ties: first minimum
    assert G.closest_sample(rec[:, 0], 5.1) == 1
    assert np.allclose(G.interp_record(rec, -5.0), rec[0, 1:])        # before first
    assert np.allclose(G.interp_record(rec, 25.0), rec[2, 1:])        # after last
    assert np.allclose(G.interp_record(rec, 12.5), [2.25, 22.5, 0, 0])  # idx=1, t > rec
    assert np.allclose(G.interp_record(rec, 7.5), [1.75, 17.5, 0, 0])   # idx=1, t < rec
    assert np.allclose(G.interp_record(rec, 10.0), rec[1, 1:])        # exact record


def _scene():
    from tests.scenes import scene
    return scene("small")


def test_deramp_setup_closed_form():
    sc = _scene()
    b = sc["burst_ref"]
    su = G.deramp_setup(sc["orbit_ref"], b)
    vs = 7500.0
    ks = 2 * vs * b["radar_frequency"] / G.SOL * b["azimuth_steering_rate"] * np.pi / 180
    assert su["ks"] == pytest.approx(ks, rel=1e-12)
    ka = -2300.0
    assert np.allclose(su["kt"], ka * ks / (ka - ks), rtol=1e-13)
    assert np.all(su["etaref"] == 0)
    L = b["lines"]
    assert su["eta"][0] == pytest.approx(-L * b["azimuth_time_interval"] / 2)


def test_deramp_preserves_modulus_and_inverts_ramp():
    sc = _scene()
    su = G.deramp_setup(sc["orbit_ref"], sc["burst_ref"])
    z = sc["ref_slc"].astype(np.complex128)
    d = G.deramp(z, su)
    assert np.allclose(np.abs(d), np.abs(z), rtol=1e-12)
    L, N = z.shape
    ramp = np.exp(1j * G.reramp_phase_of_data(su, np.arange(L), np.arange(N)))
    w = np.exp(1j * np.arange(N) * 0.1)[None, :] * np.ones((L, 1))
    assert np.allclose(G.deramp(w * ramp, su), w, atol=1e-9)


def test_reproject_bilinear_exact_on_bilinear_field():
    sc = _scene()
    b = sc["burst_ref"]
    su = G.deramp_setup(sc["orbit_ref"], b)
    h = G.target_heights(sc["grid"], sc["dem"])
    L, N = b["lines"], b["samples"]
    ii, jj = np.mgrid[0:L, 0:N].astype(np.float64)
    coef = (0.3 + 0.1j, 0.01 - 0.02j, 0.004 + 0.001j, 1e-5 + 2e-5j)
    field = lambda a, r: coef[0] + coef[1] * a + coef[2] * r + coef[3] * a * r
    z = field(ii, jj).astype(np.complex64)
    out, geo = G.reproject(sc["orbit_ref"], b, su, sc["grid"], h, z)
    v = out != 0
    assert v.mean() > 0.3
    dt = 1.0 / b["prf"]
    L2 = b["lines"]
    # undo the phase factor: flattening minus reramp
    fm = su["fm"]
    rp = geo["rng"][v]
    ka = fm[1] + (rp * 2 / G.SOL - fm[0]) * (fm[2] + fm[3] * (rp * 2 / G.SOL - fm[0]))
    kt = ka * su["ks"] / (ka - su["ks"])
    eta = -L2 * dt / 2 + geo["tline"][v] - b["t_start"]
    ph = 4 * np.pi / b["wavelength"] * rp + np.pi * kt * eta * eta
    val = out[v] * np.exp(-1j * ph)
    want = field(geo["azoff"][v], geo["rgoff"][v])
    assert np.max(np.abs(val - want)) < 2e-6 * np.max(np.abs(want))


def test_sinc_taps():
    w = G.sinc_taps(np.array([0.0]))[0]
    assert np.allclose(w, [0, 0, 0, 1, 0, 0, 0, 0], atol=1e-15)
    f = np.linspace(0, 0.999, 50)
    assert np.allclose(G.sinc_taps(f).sum(axis=-1), 1.0)
    # symmetry: taps at f and 1-f are mirror images
    a = G.sinc_taps(np.array([0.3]))[0]
    b = G.sinc_taps(np.array([0.7]))[0]
    assert np.allclose(a[::-1], b, atol=1e-12)


def test_sinc_reproduces_bandlimited_exponentials():
    n = np.arange(64)
    for fr in (0.05, 0.2, 0.35):
        x = np.exp(2j * np.pi * fr * n)
        pos = 20 + np.linspace(0, 1, 33)[:-1]
        i = np.floor(pos).astype(int)
        w = G.sinc_taps(pos - i)
        got = np.array([np.dot(w[k], x[i[k] - 3:i[k] + 5]) for k in range(len(pos))])
        want = np.exp(2j * np.pi * fr * pos)
        # truncation/window bound of the 8-tap Hann kernel grows with frequency
        tol = {0.05: 3e-3, 0.2: 4e-3, 0.35: 1.3e-1}[fr]
        assert np.max(np.abs(got - want)) < tol
