"""Pins of the PS phase-similarity oracle (oracle/psps.py) and of its C
restatement (oracle/c/s1cpu.c): the neighbour scan order against rings
traced by hand from phase_similarity.cu:295-366 and against invariants,
the similarity passes against an independent vectorised brute force and
closed-form special cases."""

import numpy as np
import pytest

from oracle import psps as P


def _ring(x, y):
    if x == 0 or y == 0:
        return [(x, y), (-x, -y), (y, x), (-y, -x)][:4]
    pts = [(x, y), (-x, -y), (x, -y), (-x, y)]
    if x != y:
        pts += [(y, x), (-y, -x), (y, -x), (-y, x)]
    return pts


def _axis(r):
    return [(r, 0), (-r, 0), (0, r), (0, -r)]


# Rings 1-4 traced by hand from the reference's steps (midpoint circle of
# radius r from (r, 0); an unvisited point inside at the same y is visited
# first and the walk steps back out): ring 4 picks up (3, 2), which the
# circles of radius 3 and 4 both miss.
HAND = {
    1: _axis(1) + _ring(1, 1),
    2: _axis(2) + _ring(2, 1),
    3: _axis(3) + _ring(3, 1) + _ring(2, 2),
    4: _axis(4) + _ring(4, 1) + _ring(3, 2) + _ring(4, 2) + _ring(3, 3),
}


def test_scan_offsets_hand_traced():
    want = [p for r in (1, 2, 3, 4) for p in HAND[r]]
    assert P.scan_offsets(0, 5).tolist() == [list(p) for p in want]
    assert P.scan_offsets(2, 5).tolist() == [list(p) for p in HAND[3] + HAND[4]]
    assert P.scan_offsets(0, 2).tolist() == [list(p) for p in HAND[1]]
    assert len(P.scan_offsets(0, 1)) == 0 and len(P.scan_offsets(3, 3)) == 0


@pytest.mark.parametrize("rdmax", [2, 3, 7, 12, 25, 41])
def test_scan_offsets_invariants(rdmax):
    off = P.scan_offsets(0, rdmax)
    pts = [tuple(p) for p in off.tolist()]
    assert len(set(pts)) == len(pts)  # each offset once
    s = set(pts)
    assert (0, 0) not in s
    # 8-fold symmetric set, every point inside the square of the last ring
    for x, y in pts:
        for q in ((-x, y), (x, -y), (y, x)):
            assert q in s
        assert max(abs(x), abs(y)) <= rdmax - 1
    # hole-free: with (x, y), every point closer to both axes is present
    for x, y in pts:
        for a in range(abs(x) + 1):
            for b in range(abs(y) + 1):
                if (a, b) != (0, 0):
                    assert (a, b) in s
    # and the union is a digital disc of radius ~rdmax - 1
    d = np.hypot(off[:, 0], off[:, 1])
    assert d.max() < rdmax - 0.5 + 0.75
    for x in range(-rdmax, rdmax + 1):
        for y in range(-rdmax, rdmax + 1):
            if 0 < np.hypot(x, y) <= rdmax - 1.5:
                assert (x, y) in s
    # the reference sizes its table as sum over rings of 8 r
    assert len(pts) <= sum(8 * r for r in range(1, rdmax))


def test_scan_offsets_rdmin_drops_inner_rings():
    for rdmin in (0, 1, 3, 6):
        full = P.scan_offsets(0, 12)
        part = P.scan_offsets(rdmin, 12)
        inner = sum(len(P.scan_offsets(0, r + 1)) - len(P.scan_offsets(0, r))
                    for r in range(1, rdmin + 1))
        np.testing.assert_array_equal(part, full[inner:])


def _stack(nifg, h, w, seed, frac=0.5):
    rng = np.random.default_rng(seed)
    z = (rng.standard_normal((nifg, h, w)) + 1j * rng.standard_normal((nifg, h, w))).astype(np.complex64)
    z[:, rng.random((h, w)) < 0.05] = 0  # some all-zero pixels
    ps = (rng.random((h, w)) < frac).astype(np.int32)
    return z, ps


def _brute(stack, ps, nneigh, offsets, kind):
    """All pair similarities per offset, first-N selection by cumulative
    counts, np.sort order statistics: vectorised, no per-pixel loops."""
    a = stack.astype(np.complex128)
    mag = np.abs(a)
    z = np.where(mag > 1e-8, a / np.where(mag > 0, mag, 1), 0)
    nifg, h, w = z.shape
    R = int(np.abs(offsets).max()) if len(offsets) else 0
    zp = np.zeros((nifg, h + 2 * R, w + 2 * R), complex)
    zp[:, R:R + h, R:R + w] = z
    pp = np.zeros((h + 2 * R, w + 2 * R), bool)
    pp[R:R + h, R:R + w] = ps != 0
    K = len(offsets)
    sims = np.empty((K, h, w))
    valid = np.empty((K, h, w), bool)
    for k, (dr, dc) in enumerate(offsets):
        nb = zp[:, R + dr:R + dr + h, R + dc:R + dc + w]
        sims[k] = np.mean((z * nb.conj()).real, axis=0)
        valid[k] = pp[R + dr:R + dr + h, R + dc:R + dc + w]
    take = valid & (np.cumsum(valid, axis=0) <= max(nneigh, 1))
    cnt = take.sum(axis=0)
    if kind == "median":
        v = np.sort(np.where(take, sims, np.inf), axis=0)
        idx = np.minimum(cnt // 2, max(K - 1, 0))
        med = np.take_along_axis(v, idx[None], 0)[0] if K else np.zeros((h, w))
        return np.where((ps != 0) & (cnt > 0), med, 0)
    mx = np.max(np.where(take, sims, -np.inf), axis=0) if K else np.full((h, w), -np.inf)
    return np.where(ps != 0, 1.0, np.where(cnt > 0, mx, 0.0))


@pytest.mark.parametrize("nneigh,rdmin,rdmax", [(5, 0, 4), (12, 1, 6), (1, 0, 3), (40, 0, 5)])
def test_similarity_against_brute_force(nneigh, rdmin, rdmax):
    z, ps = _stack(4, 13, 17, seed=nneigh + rdmax)
    offs = P.scan_offsets(rdmin, rdmax)
    zn = P.normalize(z)
    med = P.median_similarity(zn, ps, nneigh, offs)
    np.testing.assert_allclose(med, _brute(z, ps, nneigh, offs, "median"), atol=2e-6)
    mx = P.maximum_similarity(zn, ps, nneigh, offs)
    np.testing.assert_allclose(mx, _brute(z, ps, nneigh, offs, "max"), atol=2e-6)
    ps2 = P.update_ps(ps, med, 0.1)
    assert set(np.unique(ps2)) <= {0, 1}
    np.testing.assert_array_equal(ps2, (med >= np.float32(0.1)).astype(np.int32))
    keep = P.update_ps(ps, med, 0.1, overwrite=False)
    np.testing.assert_array_equal(keep, np.maximum(ps, ps2))


def test_similarity_closed_forms():
    rng = np.random.default_rng(3)
    nifg, h, w = 6, 9, 11
    phase = rng.uniform(-np.pi, np.pi, nifg)
    amp = rng.uniform(0.5, 3.0, (nifg, h, w))
    z = (amp * np.exp(1j * phase)[:, None, None]).astype(np.complex64)
    z[:, :, 5:] *= -1  # opposite phases right of column 5
    ps = np.ones((h, w), np.int32)
    med, mx, ps2 = P.similarity(z, ps, 8, 0, 3, 0.5)
    # identical phases: similarity 1; across the sign change: -1
    left_only = med[:, :3]
    np.testing.assert_allclose(left_only, 1.0, atol=1e-6)
    assert np.all(np.abs(np.abs(med) - 1) < 1e-6)
    np.testing.assert_array_equal(ps2, (med > 0).astype(np.int32))
    assert np.all(mx[ps2 == 1] == 1)
    assert np.all(np.abs(np.abs(mx) - 1) < 1e-6)
    # isolated candidates have no neighbours: median 0; a pixel with no
    # candidate in reach has maximum similarity 0
    ps = np.zeros((h, w), np.int32)
    ps[0, 0] = ps[8, 10] = 1
    offs = P.scan_offsets(0, 4)
    zn = P.normalize(z)
    med = P.median_similarity(zn, ps, 8, offs)
    assert np.all(med == 0)
    mx = P.maximum_similarity(zn, ps, 8, offs)
    assert mx[0, 0] == 1 and mx[8, 10] == 1 and mx[4, 5] == 0
    np.testing.assert_allclose(mx[1, 1], 1.0, atol=1e-6)   # same phase as (0, 0)
    np.testing.assert_allclose(mx[7, 9], 1.0, atol=1e-6)   # both right of column 5


def test_normalize():
    z = np.array([3 + 4j, 0, 1e-9 + 0j, -2j, np.nan], np.complex64)
    n = P.normalize(z)
    np.testing.assert_allclose(n[:4], [0.6 + 0.8j, 0, 0, -1j], atol=1e-7)
    assert n[4] == 0


def test_c_restatement_matches():
    import os

    from oracle import cport as C
    if not os.path.exists(C.LIB):
        pytest.skip("oracle/libs1cpu.so not built (make -C oracle all)")
    for rdmin, rdmax in ((0, 5), (2, 5), (0, 1), (0, 0), (3, 30), (0, 64)):
        np.testing.assert_array_equal(C.ps_scan(rdmin, rdmax), P.scan_offsets(rdmin, rdmax))
    z, ps = _stack(5, 21, 26, seed=11, frac=0.4)
    for nneigh, rdmin, rdmax, th in ((7, 0, 5, 0.05), (20, 1, 8, -0.02), (1, 0, 2, 0.0)):
        want = P.similarity(z, ps, nneigh, rdmin, rdmax, th)
        got = C.ps_similarity(z, ps, nneigh, rdmin, rdmax, th)
        np.testing.assert_allclose(got[0], want[0], atol=1e-6)
        near = np.abs(want[0] - th) < 1e-5
        np.testing.assert_array_equal(got[2][~near], want[2][~near])
        if not near.any():
            np.testing.assert_allclose(got[1], want[1], atol=1e-6)
    np.testing.assert_allclose(C.ps_normalize(z), P.normalize(z), atol=1e-7)
