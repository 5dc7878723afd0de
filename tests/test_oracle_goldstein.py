"""Pins of the Goldstein restatement (oracle/goldstein.py) by properties the
algorithm fixes (no GPU): alpha = 0 is the identity, plane waves keep their
phase, the filter commutes with complex scaling, the smoothing is a circular
correlation with the table, the patch grid, and zero / constant patches."""

import numpy as np
import pytest

from oracle import goldstein as GS


def test_table_properties():
    t = GS.GAUSS7
    assert t.shape == (7, 7)
    assert np.allclose(t, t.T) and np.allclose(t, t[::-1, ::-1])
    assert t.sum() == pytest.approx(1.2102, abs=2e-4)   # not 1 (goldstein.cu:43-52)
    assert t[3, 3] == pytest.approx(0.221141, abs=1e-7)


def test_smoothing_is_circular_correlation_with_table():
    n = 32
    for (ky, kx) in ((0, 0), (5, 30), (31, 2)):
        Z = np.zeros((n, n), complex)
        Z[ky, kx] = 1.0
        S = GS.smooth_mag(Z)
        want = np.zeros((n, n))
        for dy in range(-3, 4):
            for dx in range(-3, 4):
                want[(ky - dy) % n, (kx - dx) % n] += GS.GAUSS7[dy + 3, dx + 3]
        assert np.allclose(S, want, atol=1e-15)


def test_patch_grid_counts_and_clamp():
    oy, ox = GS.patch_grid(100, 75, 32)
    assert len(oy) == int(np.ceil(100 / 8)) - 3 and len(ox) == int(np.ceil(75 / 8)) - 3
    assert oy[0] == 0 and oy[-1] == 100 - 32 and ox[-1] == 75 - 32
    assert np.all(np.diff(oy)[:-1] == 8)
    oy, ox = GS.patch_grid(32, 32, 32)
    assert list(oy) == [0] and list(ox) == [0]


def test_weights_additive_triangle():
    w = GS.weights(32)
    assert w[0, 0] == 0 and w[15, 15] == 30 and w[16, 16] == 30 and w[31, 0] == 0
    assert np.allclose(w, w.T)


def test_alpha0_is_identity():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((80, 96)) + 1j * rng.standard_normal((80, 96))
    out, wsum = GS.goldstein(x, 32, 0.0)
    m = wsum > 0
    assert np.max(np.abs(out[m] - x[m])) < 1e-12
    assert np.all(out[~m] == 0)


def test_plane_wave_keeps_phase_and_scaling_commutes():
    n = 96
    yy, xx = np.mgrid[0:n, 0:n]
    x = np.exp(2j * np.pi * (5 * xx + 3 * yy) / 32.0)      # periodic in every patch
    out, wsum = GS.goldstein(x, 32, 0.7)
    m = wsum > 0
    assert np.max(np.abs(np.angle(out[m] * np.conj(x[m])))) < 1e-10
    rng = np.random.default_rng(1)
    y = rng.standard_normal((64, 64)) + 1j * rng.standard_normal((64, 64))
    c = 3.7 * np.exp(0.9j)
    a, _ = GS.goldstein(y, 32, 0.5)
    b, _ = GS.goldstein(c * y, 32, 0.5)
    assert np.max(np.abs(b - c * a)) < 1e-10 * np.abs(b).max()


def test_filter_patch_weights_where_positive():
    """Z * (S / mean)^alpha where the weight is positive; a zero patch
    (mean <= 1e-8) is left unchanged (goldstein.cu:125-140, 166-175)."""
    rng = np.random.default_rng(2)
    P = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
    F = GS.filter_patches(P[None], 0.5)[0]
    Z = np.fft.fft2(P)
    S = GS.smooth_mag(Z)
    W = (S / (S.sum() / 1024)) ** 0.5
    assert np.allclose(F, np.fft.ifft2(Z * W) * 1024, rtol=1e-12, atol=1e-9)
    Z0 = GS.filter_patches(np.zeros((1, 32, 32), complex), 0.5)
    assert np.all(Z0 == 0)
