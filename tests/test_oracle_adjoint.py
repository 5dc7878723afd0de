"""Pins of the FP64 adjoint (oracle/adjoint.py): its gradients of
loss = -mean(coherence) equal central finite differences of the FP64
forward, per offset-polynomial coefficient and along random directions of
the height-error field (no GPU)."""

import functools

import numpy as np
import pytest

from oracle import adjoint as A
from oracle import geom as G
from sentinel_gpu_processor_amd import synthetic as S


def test_sinc_tap_derivative():
    f = np.linspace(0.0, 0.999, 37)
    w, dw = A.sinc_taps_deriv(f)
    assert np.allclose(w, G.sinc_taps(f), atol=1e-15)
    e = 1e-6
    fd = (G.sinc_taps(f + e) - G.sinc_taps(f - e)) / (2 * e)
    assert np.max(np.abs(dw - fd)) < 1e-8


@functools.lru_cache(maxsize=None)
def _case(seed=0):
    sc = S.make_scene(lines=64, samples=256, dem_shape=(32, 48), hmax=200.0, seed=seed,
                      misreg=(0.3, 0.05), misreg_ramp=(0.2, 0.1))
    setups = (G.deramp_setup(sc["orbit_ref"], sc["burst_ref"]),
              G.deramp_setup(sc["orbit_sec"], sc["burst_sec"]))
    dr = G.deramp(sc["ref_slc"], setups[0]).astype(np.complex64)
    ds = G.deramp(sc["sec_slc"], setups[1]).astype(np.complex64)
    h0 = G.target_heights(sc["grid"], sc["dem"])
    dh = np.random.default_rng(seed + 7).normal(0.0, 5.0, h0.shape).astype(np.float32)
    return sc, setups, dr, ds, h0, dh


@pytest.mark.parametrize("interp,win", [("sinc", 5), ("sinc", 1)])
def test_gradient_matches_finite_differences(interp, win):
    sc, setups, dr, ds, h0, dh = _case()
    p = np.asarray(sc["offpoly"], np.float64)
    h = h0 + dh.astype(np.float64)

    F = A.pair_forward(sc, setups, h, dr, ds, interp, p, 1, 4)
    mask = A.interior_mask(F, 1, 4, win)
    assert mask.mean() > 0.15

    def loss(pp, hh):
        return A.loss_neg_mean_coherence(sc, setups, hh, dr, ds, interp, pp, 1, 4, win,
                                         mask, round_blend=False)[0]

    l0, gp, gh = A.grad_neg_mean_coherence(sc, setups, h, dr, ds, interp, p, 1, 4, win, mask,
                                           round_blend=False)
    assert np.isfinite(l0) and -1 < l0 < 0
    e_p, e_h, tol = 1e-3, 1e-2, 2e-3       # 1e-3 px / 1e-2 m steps
    for k in range(6):
        d = np.zeros(6)
        d[k] = e_p
        fd = (loss(p + d, h) - loss(p - d, h)) / (2 * e_p)
        assert gp[k] == pytest.approx(fd, rel=tol, abs=2e-6), (k, gp[k], fd)
    rng = np.random.default_rng(1)
    for _ in range(2):
        delta = rng.standard_normal(h.shape)
        fd = (loss(p, h + e_h * delta) - loss(p, h - e_h * delta)) / (2 * e_h)
        got = float(np.sum(gh * delta))
        assert got == pytest.approx(fd, rel=tol, abs=2e-6), (got, fd)


def test_bilinear_burst_derivatives_away_from_kinks():
    """The bilinear-specific part of the adjoint (interpolation slopes and
    the reramp at the shifted position): d out / d(range, azimuth shift) of
    the secondary burst against central differences at every pixel whose
    sampling position is not within the step of an integer (bilinear
    interpolation has kinks there; the chain above the burst output is the
    same code as for sinc)."""
    sc, setups, dr, ds, h0, dh = _case()
    p = np.asarray(sc["offpoly"], np.float64)
    h = h0 + dh.astype(np.float64)
    T = A.burst_terms(sc["orbit_sec"], sc["burst_sec"], setups[1], sc["grid"], h, ds,
                      "bilinear", p, round_blend=False)
    geo = G.geo2rdr_grid(sc["orbit_sec"], sc["burst_sec"], sc["grid"], h)
    drg, daz = G.offset_poly_shift(p, geo["rgoff"], geo["azoff"], sc["burst_sec"])
    pos = {0: geo["rgoff"] + drg, 3: geo["azoff"] + daz}
    # 1e-4 px: large against the FP64 rounding of the ~1e8 rad flattening
    # phase (~3e-8 rad), small against the pixel
    e = 1e-4
    for k, dpos in ((0, T["I_rg"] + 1j * T["out"] * (-T["R_rp"]) * T["dmrg"]),
                    (3, T["I_az"] + 1j * T["out"] * (-T["R_eta"]) * T["dtaz"])):
        d = np.zeros(6)
        d[k] = e
        op, _ = G.reproject(sc["orbit_sec"], sc["burst_sec"], setups[1], sc["grid"], h, ds,
                            "bilinear", p + d, round_blend=False)
        om, _ = G.reproject(sc["orbit_sec"], sc["burst_sec"], setups[1], sc["grid"], h, ds,
                            "bilinear", p - d, round_blend=False)
        fd = (op - om) / (2 * e)
        frac = pos[k] - np.floor(pos[k])
        ok = T["valid"] & (op != 0) & (om != 0) & (frac > 2 * e) & (frac < 1 - 2 * e)
        assert ok.mean() > 0.4
        err = np.abs(fd[ok] - dpos[ok]) / np.abs(fd[ok]).max()
        assert err.max() < 5e-4, (k, err.max())
