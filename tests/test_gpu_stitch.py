"""GPU stitching (s1ifg_stitch_subswath / s1ifg_merge_subswath) against the
oracle's restatement of interfere.py:81-180: mosaics and per-burst offsets
of complex and phase-image subswaths with noisy overlaps, holes, matching
ids (carried offsets), empty overlaps, odd / even / single-pixel overlap
counts, and the merge of subswaths into the full image.  Values are float32
products of float32 trigonometry; the tolerance (2e-5 rad / 2e-5 absolute
on unit-modulus pixels) covers one-ulp differences of sin / cos / atan2 and
the FP64 (device) against complex64 pairwise (NumPy) mean."""

import numpy as np
import pytest

from oracle import ifg as OI

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
TOL = 2e-5


def _make(rng, ncol, spans, offs, ids, out_float, cols=None, holes=0.05, noise=0.3):
    nrow0 = max(b for _, b in spans)
    base = rng.uniform(-np.pi, np.pi, (nrow0, ncol))
    bursts, lst = [], []
    for k, ((t, b), d) in enumerate(zip(spans, offs)):
        l, r = cols[k] if cols else (0, ncol)
        ph = base[t:b, l:r] + d + rng.normal(0, noise, (b - t, r - l))
        if out_float:
            data = np.angle(np.exp(1j * ph)).astype(np.float32)
        else:
            data = (rng.uniform(0.5, 2.0, ph.shape) * np.exp(1j * ph)).astype(np.complex64)
        data[rng.random(data.shape) < holes] = 0
        bursts.append(((nrow0, ncol, l, t, r, b), data))
        lst.append((l, t, r, b, data))
    return bursts, lst


def _gpu(lst, ids, out_float):
    from sentinel_gpu_processor_amd import kernels as K
    from sentinel_gpu_processor_amd.runtime import context
    ctx = context()
    est = [True] + [ids[i] != ids[i - 1] for i in range(1, len(ids))]
    dev = [(l, t, r, b, torch.from_numpy(d).to("cuda:0"), e)
           for (l, t, r, b, d), e in zip(lst, est)]
    mos, bounds, offs = K.stitch_subswath(ctx, dev, out_float=out_float, offsets=True)
    torch.cuda.synchronize()
    return mos.cpu().numpy(), bounds, offs.cpu().numpy()


def _check(bursts, lst, ids, out_float):
    want, wb, woffs = OI.stitch_subswath(bursts, ids, out_float=out_float)
    got, gb, goffs = _gpu(lst, ids, out_float)
    assert tuple(gb) == tuple(wb)
    assert np.allclose(goffs, woffs, atol=TOL, rtol=0), (goffs, woffs)
    if out_float:
        # a pixel within TOL of +-pi may wrap to the other end
        diff = np.angle(np.exp(1j * (got.astype(np.float64) - want)))
        assert np.all(np.abs(diff) < TOL)
        assert np.array_equal(got == 0, want == 0)
    else:
        assert np.array_equal(got.real == 0, want.real == 0)
        assert np.allclose(got, want, atol=TOL * 4, rtol=0)
    return woffs


@pytest.mark.parametrize("out_float", [False, True])
def test_stitch_noisy_overlaps(out_float):
    rng = np.random.default_rng(11)
    spans = [(0, 130), (110, 250), (231, 380), (360, 500), (470, 600)]
    offs = [0.0, 0.9, -2.2, 3.0, 1.4]
    ids = [("a", str(k)) for k in range(5)]
    bursts, lst = _make(rng, 700, spans, offs, ids, out_float)
    woffs = _check(bursts, lst, ids, out_float)
    assert len(set(np.round(woffs, 3))) == 5


@pytest.mark.parametrize("out_float", [False, True])
def test_stitch_carry_empty_and_ragged(out_float):
    """Bursts 1-2 share ids (carried offset); burst 3 does not touch the
    mosaic (empty overlap: the offset carries); ragged column extents."""
    rng = np.random.default_rng(12)
    spans = [(0, 60), (50, 120), (110, 170), (175, 230), (220, 300)]
    cols = [(0, 333), (5, 320), (10, 333), (0, 300), (3, 333)]
    offs = [0.5, -0.7, -0.7, 2.0, 0.1]
    ids = [("a", "0"), ("a", "1"), ("a", "1"), ("a", "3"), ("a", "4")]
    bursts, lst = _make(rng, 333, spans, offs, ids, out_float, cols=cols)
    woffs = _check(bursts, lst, ids, out_float)
    assert woffs[2] == woffs[1] and woffs[3] == woffs[2]


@pytest.mark.parametrize("out_float", [False, True])
def test_stitch_three_bursts_on_a_row(out_float):
    """Burst 2 reaches rows of burst 0 as well as burst 1's: the boundaries
    are then estimated one after the other against the mosaic (the other
    tests have each burst meet only its predecessor, which lets all
    boundaries be estimated at once)."""
    rng = np.random.default_rng(15)
    spans = [(0, 100), (40, 140), (90, 200), (185, 260)]
    offs = [0.2, 1.3, -2.0, 0.7]
    ids = [("t", str(k)) for k in range(4)]
    bursts, lst = _make(rng, 240, spans, offs, ids, out_float)
    _check(bursts, lst, ids, out_float)


@pytest.mark.parametrize("out_float", [False, True])
def test_stitch_even_offsets(out_float):
    """Even mosaic width with even, non-zero burst column offsets: the
    16-byte complex kernels (placement of burst 0 at column 2 of the box,
    later bursts at columns 4 and 6), and the same layout in phase images."""
    rng = np.random.default_rng(14)
    spans = [(0, 70), (60, 140), (128, 200), (190, 260)]
    cols = [(2, 334), (4, 330), (6, 334), (0, 334)]
    offs = [0.3, -1.1, 2.4, -0.6]
    ids = [("e", str(k)) for k in range(4)]
    bursts, lst = _make(rng, 334, spans, offs, ids, out_float, cols=cols)
    _check(bursts, lst, ids, out_float)


@pytest.mark.parametrize("n_ov", [1, 2, 3, 64, 65])
def test_stitch_small_overlap_counts(n_ov):
    """Overlaps of 1, 2, 3, 64 and 65 pixels: single, even and odd medians."""
    rng = np.random.default_rng(100 + n_ov)
    ncol = 80
    old = (np.exp(1j * rng.uniform(-3, 3, (4, ncol)))).astype(np.complex64)
    old[3, n_ov:] = 0
    new = (np.exp(1j * rng.uniform(-3, 3, (6, ncol)))).astype(np.complex64)
    bursts = [((9, ncol, 0, 0, ncol, 4), old), ((9, ncol, 0, 3, ncol, 9), new)]
    lst = [(0, 0, ncol, 4, old), (0, 3, ncol, 9, new)]
    _check(bursts, lst, [("a", "b"), ("c", "d")], False)


def test_stitch_iw_sized_subswath():
    """Nine 400 x 3000 bursts with 60-row overlaps (bit pattern of the
    selection at ~180,000 overlap pixels per boundary)."""
    rng = np.random.default_rng(13)
    spans = [(k * 340, k * 340 + 400) for k in range(9)]
    offs = rng.uniform(-3, 3, 9)
    ids = [("x", str(k)) for k in range(9)]
    bursts, lst = _make(rng, 3000, spans, offs, ids, False, holes=0.01)
    _check(bursts, lst, ids, False)


@pytest.mark.parametrize("out_float", [False, True])
def test_merge_subswaths(out_float):
    from sentinel_gpu_processor_amd import kernels as K
    from sentinel_gpu_processor_amd.runtime import context
    rng = np.random.default_rng(14)
    dt = np.float32 if out_float else np.complex64
    subs = []
    for (l, t, r, b) in ((0, 0, 50, 40), (40, 5, 90, 45), (80, 2, 120, 38)):
        d = (rng.standard_normal((b - t, r - l)) + (0 if out_float else 1j) *
             rng.standard_normal((b - t, r - l))).astype(dt)
        d[rng.random(d.shape) < 0.2] = 0
        subs.append(((l, t, r, b), d))
    want = OI.merge_subswaths(subs, 50, 130, out_float)
    ctx = context()
    dst = torch.zeros((50, 130), dtype=torch.float32 if out_float else torch.complex64,
                      device="cuda:0")
    for (l, t, r, b), d in subs:
        K.merge_subswath(ctx, torch.from_numpy(d).to("cuda:0"), dst, t, l, out_float)
    assert np.array_equal(dst.cpu().numpy(), want)


def test_stitch_rejects_burst_outside_bounds():
    from sentinel_gpu_processor_amd import kernels as K
    from sentinel_gpu_processor_amd.runtime import context
    a = torch.ones((10, 10), dtype=torch.complex64, device="cuda:0")
    b = torch.ones((10, 12), dtype=torch.complex64, device="cuda:0")
    c = torch.ones((10, 10), dtype=torch.complex64, device="cuda:0")
    with pytest.raises(Exception, match="outside"):
        K.stitch_subswath(context(), [(0, 0, 10, 10, a, True), (0, 5, 12, 15, b, True),
                                      (0, 10, 10, 20, c, True)])


@pytest.mark.parametrize("out_float", [False, True])
def test_dropin_stitch_files(tmp_path, out_float):
    """interfere.stitch on per-burst .int files of two subswaths (the
    reference's file contract, interfere.py:153-180) against the oracle's
    stitch + merge; the burst files are removed."""
    import os

    from sentinel_gpu_processor_amd import interfere as IF
    from sentinel_gpu_processor_amd import sario as SA
    rng = np.random.default_rng(15)
    nrow0, ncol0 = 90, 120
    dt = np.float32 if out_float else np.complex64
    files, subs = [], []
    for sw, (l, r) in ((1, (0, 70)), (2, (60, 120))):
        spans = [(0, 35), (30, 62), (57, 90)]
        bursts, ids = [], []
        for k, (t, b) in enumerate(spans):
            ph = rng.uniform(-3, 3, (b - t, r - l)) + 0.4 * k
            data = (np.angle(np.exp(1j * ph)) if out_float else np.exp(1j * ph)).astype(dt)
            data[rng.random(data.shape) < 0.05] = 0
            name = f"x_d1_s{k}_a_b_c_d_d2_s{k}_iw{sw}.int"
            f = str(tmp_path / name)
            SA.save_cropped(f, data, nrow0, ncol0, l, t, r, b)
            files.append(f)
            bursts.append(((nrow0, ncol0, l, t, r, b), data))
            w = name.split("_")
            ids.append((w[1] + w[2], w[7] + w[8]))
        mos, bnd, _ = OI.stitch_subswath(bursts, ids, out_float=out_float)
        subs.append((bnd, mos))
    want = OI.merge_subswaths(subs, nrow0, ncol0, out_float)
    out = str(tmp_path / "pair.int")
    IF.stitch(files, out, out_float)
    got = np.fromfile(out, dtype=dt).reshape(nrow0, ncol0)
    if out_float:
        assert np.all(np.abs(np.angle(np.exp(1j * (got.astype(np.float64) - want)))) < TOL)
    else:
        assert np.allclose(got, want, atol=4 * TOL, rtol=0)
    assert not any(os.path.exists(f) for f in files)


def test_stitch_deterministic():
    """Two runs give bit-identical mosaics and offsets (fixed-order FP64
    partial sums, integer histograms)."""
    rng = np.random.default_rng(16)
    spans = [(k * 90, k * 90 + 120) for k in range(6)]
    ids = [("x", str(k)) for k in range(6)]
    bursts, lst = _make(rng, 1000, spans, rng.uniform(-3, 3, 6), ids, False)
    a = _gpu(lst, ids, False)
    b = _gpu(lst, ids, False)
    assert np.array_equal(a[0].view(np.uint64), b[0].view(np.uint64))
    assert np.array_equal(a[2], b[2])
