"""FP64 NumPy restatement of the Goldstein filter (goldstein.cu:42-408), test
infrastructure.

cuFFT (closed library, not in the reference tree) is replaced by numpy.fft:
cuFFT FORWARD is the unnormalised e^{-i} DFT = numpy.fft.fft2, cuFFT INVERSE
is the unnormalised e^{+i} DFT = numpy.fft.ifft2 * n^2; the reference rescales
by 1/n^2 in the overlap-add (goldstein.cu:267).

Parity pins (tests/test_oracle_goldstein.py): alpha = 0 is the identity where
the weight sum is positive; plane waves keep their phase; the filter commutes
with complex scaling; the smoothing is the circular correlation of |Z| with the
table (impulse responses); table symmetry and sum; patch-grid counts and
clamping; additive triangular weights; zero patches stay zero.  The FFT itself
is numpy.fft (parity with cuFFT unpinned: no reference test or vector holds
cuFFT output).

Deliberate departures (masked in parity): the reference never zeroes its
output buffer (goldstein.cu:484), so pixels with weight sum 0 hold garbage
there; here they are 0.
"""

import numpy as np

# goldstein.cu:43-52 (project data: the reference's 7x7 smoothing table)
GAUSS7 = np.array([
    0.000004, 0.000088, 0.000543, 0.000948, 0.000543, 0.000088, 0.000004,
    0.000088, 0.001915, 0.011776, 0.020584, 0.011776, 0.001915, 0.000088,
    0.000543, 0.011776, 0.072410, 0.126584, 0.072410, 0.011776, 0.000543,
    0.000948, 0.020584, 0.126584, 0.221141, 0.126584, 0.020584, 0.000948,
    0.000543, 0.011776, 0.072410, 0.126584, 0.072410, 0.011776, 0.000543,
    0.000088, 0.001915, 0.011776, 0.020584, 0.011776, 0.001915, 0.000088,
    0.000004, 0.000088, 0.000543, 0.000948, 0.000543, 0.000088, 0.000004,
], dtype=np.float32).astype(np.float64).reshape(7, 7)


def patch_grid(nrow, ncol, n_win):
    """goldstein.cu:334-342 and 308-313: counts and clamped origins."""
    inc = n_win // 4
    pi = int(np.ceil(np.float32(nrow) / np.float32(inc))) - 3
    pj = int(np.ceil(np.float32(ncol) / np.float32(inc))) - 3
    pi = max(pi, 1)
    pj = max(pj, 1)
    oy = np.minimum(np.arange(pi) * inc, nrow - n_win)
    ox = np.minimum(np.arange(pj) * inc, ncol - n_win)
    return oy, ox


def smooth_mag(Z):
    """goldstein.cu:87-123: circular 7x7 correlation of |Z| with GAUSS7."""
    M = np.abs(Z)
    S = np.zeros_like(M)
    for ky in range(-3, 4):
        for kx in range(-3, 4):
            S += GAUSS7[ky + 3, kx + 3] * np.roll(np.roll(M, -ky, axis=-2), -kx, axis=-1)
    return S


def filter_patches(P, alpha):
    """Steps 2-6 on a stack of patches P (..., n, n) -> filtered (unnormalised
    inverse transform, i.e. n^2 times the spatial patch)."""
    n = P.shape[-1]
    Z = np.fft.fft2(P)
    S = smooth_mag(Z)
    mean = S.reshape(S.shape[:-2] + (-1,)).sum(axis=-1) / (n * n)
    ok = mean > 1e-8
    S = np.where(ok[..., None, None], S / np.where(ok, mean, 1.0)[..., None, None], 0.0)
    W = np.power(S, alpha)
    Z = np.where(W > 0, Z * W, Z)
    return np.fft.ifft2(Z) * (n * n)


def weights(n):
    """goldstein.cu:257-259: w = wx + wy (additive triangular)."""
    t = np.arange(n)
    w1 = np.where(t < n // 2, t, n - 1 - t).astype(np.float64)
    return w1[None, :] + w1[:, None]


def goldstein(img, n_win=32, alpha=0.5):
    """Whole-image filter (goldstein.cu:328-408); returns (out, wsum)."""
    img = np.asarray(img, np.complex128)
    nrow, ncol = img.shape
    oy, ox = patch_grid(nrow, ncol, n_win)
    w = weights(n_win)
    out = np.zeros((nrow, ncol), np.complex128)
    wsum = np.zeros((nrow, ncol))
    norm = 1.0 / (n_win * n_win)
    for y0 in oy:
        P = np.stack([img[y0:y0 + n_win, x0:x0 + n_win] for x0 in ox])
        F = filter_patches(P, alpha)
        for k, x0 in enumerate(ox):
            out[y0:y0 + n_win, x0:x0 + n_win] += F[k] * w * norm
            wsum[y0:y0 + n_win, x0:x0 + n_win] += w
    ok = wsum > 0
    out[ok] /= wsum[ok]
    out[~ok] = 0
    return out, wsum
