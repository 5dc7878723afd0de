"""Pins of the oracle's stitch restatement (interfere.py:81-180) by closed
forms: injected per-burst phase offsets are recovered, the median step
rejects a minority of outliers, a two-pixel overlap gives the mean of the
two angles, offsets carry over to bursts whose ids match, phase images are
shifted and wrapped into [-pi, pi], and pixels holding data are kept."""

import numpy as np

from oracle import ifg as OI


def _field(rng, nrow, ncol):
    return np.exp(1j * rng.uniform(-np.pi, np.pi, (nrow, ncol))).astype(np.complex64)


def _bursts(base, spans, offs, ids, ncol):
    nrow0 = base.shape[0]
    out = []
    for (t, b), d in zip(spans, offs):
        data = (base[t:b] * np.exp(1j * d)).astype(np.complex64)
        out.append(((nrow0, ncol, 0, t, ncol, b), data))
    return out, ids


def test_offsets_recovered():
    rng = np.random.default_rng(3)
    base = _field(rng, 70, 30)
    spans = [(0, 25), (20, 45), (40, 70)]
    offs = [0.3, -1.1, 2.0]
    bursts, ids = _bursts(base, spans, offs, [("a", "b"), ("c", "d"), ("e", "f")], 30)
    mos, bounds, got = OI.stitch_subswath(bursts, ids)
    assert bounds == (0, 0, 30, 70)
    # every pixel equals the base field rotated by burst 0's offset
    assert np.allclose(mos, base * np.exp(1j * 0.3), atol=2e-6)
    assert np.allclose(got, [0.0, -1.1 - 0.3, 2.0 - 0.3], atol=2e-6)


def test_median_rejects_minority():
    """60 % of the overlap at phase difference 0.5, 40 % at 2.5: the mean is
    pulled towards 2.5, the median of the residuals brings it back to 0.5."""
    ncol = 50
    old = np.zeros((10, ncol), np.complex64)
    old[:4] = 1.0
    new = np.ones((10, ncol), np.complex64) * np.exp(1j * 0.5).astype(np.complex64)
    new[:4, 30:] = np.exp(1j * 2.5)
    bursts = [((10, ncol, 0, 0, ncol, 4), old[:4]), ((10, ncol, 0, 0, ncol, 10), new)]
    mos, _, offs = OI.stitch_subswath(bursts, [("a", "b"), ("c", "d")])
    m = np.angle(np.mean(np.concatenate([np.full(120, np.exp(0.5j)), np.full(80, np.exp(2.5j))])))
    assert abs(m - 0.5) > 0.3  # the mean alone is biased
    assert abs(offs[1] - 0.5) < 1e-6
    assert np.allclose(mos[4:], 1.0, atol=1e-6)


def test_two_pixel_overlap_mean_of_angles():
    a, b = 0.2, 0.9
    old = np.zeros((2, 2), np.complex64)
    old[0] = 1.0
    new = np.array([[np.exp(1j * a), np.exp(1j * b)], [1.0, 1.0]], np.complex64)
    bursts = [((2, 2, 0, 0, 2, 1), old[:1]), ((2, 2, 0, 0, 2, 2), new)]
    _, _, offs = OI.stitch_subswath(bursts, [("a", "b"), ("c", "d")])
    # mean angle (a+b)/2, residuals +-(b-a)/2, median of two = their mean = 0
    assert abs(offs[1] - (a + b) / 2) < 1e-6


def test_offset_carries_over_matching_ids():
    rng = np.random.default_rng(4)
    base = _field(rng, 60, 20)
    spans = [(0, 25), (20, 45), (40, 60)]
    # bursts 1 and 2 share ids (same scene pair): no new estimate for burst 2
    bursts, ids = _bursts(base, spans, [0.0, 0.8, 0.8], [("a", "b"), ("c", "d"), ("c", "d")], 20)
    mos, _, offs = OI.stitch_subswath(bursts, ids)
    assert np.allclose(offs, [0.0, 0.8, 0.8], atol=2e-6)
    assert np.allclose(mos, base, atol=2e-6)
    # with a different offset on burst 2 the carried offset leaves a step
    bursts, ids = _bursts(base, spans, [0.0, 0.8, 1.5], [("a", "b"), ("c", "d"), ("c", "d")], 20)
    mos, _, offs = OI.stitch_subswath(bursts, ids)
    assert np.allclose(np.angle(mos[45:] * np.conj(base[45:])), 0.7, atol=2e-6)


def test_phase_images_shift_and_wrap():
    rng = np.random.default_rng(5)
    ph = rng.uniform(-np.pi, np.pi, (40, 16)).astype(np.float32)
    d = 2.5
    b1 = np.angle(np.exp(1j * (ph[15:40] + d))).astype(np.float32)
    bursts = [((40, 16, 0, 0, 16, 20), ph[:20]), ((40, 16, 0, 15, 16, 40), b1)]
    mos, _, offs = OI.stitch_subswath(bursts, [("a", "b"), ("c", "d")], out_float=True)
    assert abs(offs[1] - d) < 2e-5
    assert mos.dtype == np.float32
    assert np.all(np.abs(mos) <= np.float32(np.pi))
    assert np.allclose(np.angle(np.exp(1j * (mos - ph))), 0, atol=2e-5)


def test_pixels_with_data_are_kept_and_merge():
    old = np.full((4, 4), 2 + 0j, np.complex64)
    new = np.full((4, 4), 3 + 0j, np.complex64)
    new[:, 0] = 0
    bursts = [((8, 4, 0, 0, 4, 4), old), ((8, 4, 0, 2, 4, 6), new)]
    mos, bounds, offs = OI.stitch_subswath(bursts, [("a", "b"), ("a", "b")])
    assert bounds == (0, 0, 4, 6) and offs == [0.0, 0.0]
    assert np.all(mos[:4] == 2) and np.all(mos[4:6, 1:] == 3) and np.all(mos[4:6, 0] == 0)
    full = OI.merge_subswaths([(bounds, mos), ((2, 0, 4, 2), np.full((2, 2), 5, np.complex64))],
                              8, 4)
    assert np.all(full[0:2, 2:4] == 5) and np.all(full[0:2, 0:2] == 2) and np.all(full[6:] == 0)
