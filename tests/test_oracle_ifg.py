"""Pins of the crossmul / looks / coherence restatement (oracle/ifg.py)
against closed forms, brute force on tiny arrays and statistical known
answers (no GPU)."""

import numpy as np
import pytest

from oracle import ifg as I


def test_conj_mul_phase_is_secondary_minus_reference():
    rng = np.random.default_rng(0)
    pa = rng.uniform(-np.pi, np.pi, (5, 7))
    pb = rng.uniform(-np.pi, np.pi, (5, 7))
    ma, mb = rng.uniform(0.5, 2, (2, 5, 7))
    c = I.conj_mul(ma * np.exp(1j * pa), mb * np.exp(1j * pb))
    d = np.angle(c * np.exp(-1j * (pb - pa)))
    assert np.max(np.abs(d)) < 1e-12
    assert np.allclose(np.abs(c), ma * mb, rtol=1e-13)
    # the reference's real/imaginary formulas (sario.cu:199-211) written out
    a = (rng.standard_normal((3, 4)) + 1j * rng.standard_normal((3, 4)))
    b = (rng.standard_normal((3, 4)) + 1j * rng.standard_normal((3, 4)))
    c = I.conj_mul(a, b)
    for i in range(3):
        for j in range(4):
            re = a[i, j].real * b[i, j].real + a[i, j].imag * b[i, j].imag
            im = a[i, j].real * b[i, j].imag - a[i, j].imag * b[i, j].real
            assert c[i, j] == pytest.approx(complex(re, im), rel=1e-14, abs=1e-14)


def test_cpx_looks_brute_force_constant_and_ramp():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((7, 13)) + 1j * rng.standard_normal((7, 13))
    for rl, cl in ((1, 1), (1, 4), (2, 3), (3, 2)):
        got = I.cpx_looks(x, rl, cl)
        assert got.shape == (7 // rl, 13 // cl)
        for r in range(got.shape[0]):
            for c in range(got.shape[1]):
                want = x[r * rl:(r + 1) * rl, c * cl:(c + 1) * cl].mean()
                assert got[r, c] == pytest.approx(want, rel=1e-13, abs=1e-14)
    assert np.allclose(I.cpx_looks(np.full((6, 8), 2 - 3j), 2, 4), 2 - 3j)
    ramp = np.arange(12, dtype=float)[None, :].repeat(4, 0)
    assert np.allclose(I.cpx_looks(ramp, 1, 4).real[0], [1.5, 5.5, 9.5])


def test_overlap_cases():
    # disjoint
    assert I.overlap((10, 20, 0, 0, 5, 5), (10, 20, 6, 0, 10, 5), 1, 1) is None
    # nested, aligned
    ov = I.overlap((100, 200, 0, 0, 200, 100), (100, 200, 8, 4, 100, 60), 2, 4)
    assert (ov["left"], ov["top"], ov["right"], ov["bottom"]) == (8, 4, 100, 60)
    assert ov["header"] == [50, 50, 2, 2, 25, 30]
    # unaligned edges snap inward to look multiples
    ov = I.overlap((100, 200, 3, 5, 197, 99), (100, 200, 1, 2, 199, 98), 3, 4)
    assert (ov["left"], ov["right"]) == (4, 196)
    assert (ov["top"], ov["bottom"]) == (6, 96)
    assert ov["nrow"] == 90 and ov["ncol"] == 192


def test_powlooks_quirk_and_amplitude():
    rng = np.random.default_rng(2)
    s = (rng.standard_normal((4, 8)) + 1j * rng.standard_normal((4, 8))).astype(np.complex64)
    # 1 x 1 looks: powlooks is |s|, so the amplitude is sqrt|s| (sario.py:417-418)
    assert np.allclose(I.amplitude(s, 1, 1), np.sqrt(np.abs(s)), rtol=1e-6)
    # otherwise sqrt(mean |s|^2) over the cell
    a = I.amplitude(s, 2, 4)
    for r in range(2):
        for c in range(2):
            cell = s[2 * r:2 * r + 2, 4 * c:4 * c + 4].astype(np.complex128)
            assert a[r, c] == pytest.approx(np.sqrt(np.mean(np.abs(cell) ** 2)), rel=1e-12)
    assert np.allclose(I.amplitude(np.full((4, 8), 3 + 4j), 2, 2), 5.0)


def test_coherence_identical_slcs_and_bound():
    rng = np.random.default_rng(3)
    s = rng.standard_normal((16, 32)) + 1j * rng.standard_normal((16, 32))
    ifg = I.cpx_looks(I.conj_mul(s, s), 2, 4)
    amp = I.amplitude(s, 2, 4)
    g = I.coherence_lookcell(ifg, amp, amp, fill="none")
    assert np.allclose(g, 1.0, atol=1e-10)
    t = rng.standard_normal((16, 32)) + 1j * rng.standard_normal((16, 32))
    g = I.coherence_lookcell(I.cpx_looks(I.conj_mul(s, t), 2, 4), amp, I.amplitude(t, 2, 4),
                             fill="none")
    assert np.all(g <= 1 + 1e-12)          # Cauchy-Schwarz


def test_coherence_statistical_known_answer():
    """s2 = g0 s1 + sqrt(1 - g0^2) n over N = 64 looks: the sample coherence
    is close to g0 (bias ~ (1 - g0^2)^2 / (2 N g0) at this size)."""
    rng = np.random.default_rng(4)
    g0 = 0.8
    shape = (256, 512)
    s1 = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)
    n = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)
    s2 = g0 * s1 + np.sqrt(1 - g0 ** 2) * n
    ifg = I.cpx_looks(I.conj_mul(s1, s2), 8, 8)
    g = I.coherence_lookcell(ifg, I.amplitude(s1, 8, 8), I.amplitude(s2, 8, 8), fill="none")
    assert abs(np.mean(g) - g0) < 0.01
    assert np.std(g) < 0.05


def test_coherence_window_brute_force():
    rng = np.random.default_rng(5)
    ifg = rng.standard_normal((9, 11)) + 1j * rng.standard_normal((9, 11))
    p1 = rng.uniform(0.5, 2, (9, 11))
    p2 = rng.uniform(0.5, 2, (9, 11))
    for win in (1, 3, 5):
        g = I.coherence_window(ifg, p1, p2, win)
        h = win // 2
        for r in range(9):
            for c in range(11):
                sl = np.s_[max(r - h, 0):r + h + 1, max(c - h, 0):c + h + 1]
                want = abs(ifg[sl].sum()) / np.sqrt(p1[sl].sum() * p2[sl].sum())
                assert g[r, c] == pytest.approx(want, rel=1e-12)
    # identical SLCs -> 1 for any window
    s = rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))
    assert np.allclose(I.coherence_window(s * np.conj(s), np.abs(s) ** 2, np.abs(s) ** 2, 5), 1)


def test_cc_rmg_layout():
    a1 = np.full((2, 3), 2.0)
    a2 = np.full((2, 3), 3.0)
    coh = np.arange(6, dtype=float).reshape(2, 3) / 10
    m = I.cc_rmg(a1, a2, coh)
    assert m.shape == (2, 6) and m.dtype == np.float32
    assert np.all(m[:, :3] == 6.0) and np.allclose(m[:, 3:], coh)


def test_match_bursts_first_minimum_wins():
    ref = [(0, 100), (90, 190), (180, 280)]
    sec = [(2, 101), (88, 190), (91, 189), (500, 600)]
    # |dtop| + |dbottom|: burst 1 ties between sec 1 (2 + 0) and sec 2 (1 + 1)
    assert I.match_bursts(ref, sec) == [0, 1, 2]
