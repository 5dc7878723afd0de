"""GPU parity of the PS phase-similarity path (phase_similarity.cu:116-548)
against the float oracles: the NumPy restatement (oracle/psps.py, pinned by
brute force) on tiny stacks and the C restatement (oracle/c/s1cpu_ps.c) on
larger ones.  Both follow the reference's float sequence, as the kernels
do, so the similarities and the updated candidate masks must agree bit
for bit."""

import numpy as np
import pytest

from oracle import psps as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from sentinel_gpu_processor_amd import _native as N
    c = N.Context(0)
    yield c
    c.sync()
    c.close()


@pytest.fixture(scope="module")
def C():
    import os

    from oracle import cport
    if not os.path.exists(cport.LIB):
        pytest.skip("oracle/libs1cpu.so not built")
    return cport


def _gpu(ctx, stack, ps, nneigh, rdmin, rdmax, th):
    from sentinel_gpu_processor_amd import kernels as K
    print(f"[ps] stack {stack.shape} nneigh={nneigh} rd=({rdmin},{rdmax}) th={th}", flush=True)
    st = torch.from_numpy(np.ascontiguousarray(stack)).cuda()
    pst = torch.from_numpy(np.ascontiguousarray(ps, np.int32)).cuda()
    med, mx = K.ps_similarity(ctx, st, pst, nneigh, rdmin, rdmax, th)
    torch.cuda.synchronize()
    return med.cpu().numpy(), mx.cpu().numpy(), pst.cpu().numpy()


def _random(nifg, h, w, seed, frac):
    rng = np.random.default_rng(seed)
    z = (rng.standard_normal((nifg, h, w)) + 1j * rng.standard_normal((nifg, h, w))).astype(np.complex64)
    z[:, rng.random((h, w)) < 0.05] = 0
    ps = (rng.random((h, w)) < frac).astype(np.int32)
    return z, ps


def _same(got, want):
    np.testing.assert_array_equal(got[2], want[2])
    for g, w in zip(got[:2], want[:2]):
        assert np.array_equal(g.view(np.uint32), w.view(np.uint32)), np.max(np.abs(g - w))


@pytest.mark.parametrize("nneigh,rdmin,rdmax,th", [(5, 0, 4, 0.1), (12, 1, 6, 0.0),
                                                   (1, 0, 3, 0.2), (40, 0, 5, -0.1)])
def test_tiny_against_numpy_oracle(ctx, nneigh, rdmin, rdmax, th):
    z, ps = _random(4, 13, 17, seed=nneigh + rdmax, frac=0.5)
    _same(_gpu(ctx, z, ps, nneigh, rdmin, rdmax, th), P.similarity(z, ps, nneigh, rdmin, rdmax, th))


CASES = [
    # nifg, nrow, ncol, nneigh, rdmin, rdmax, th: register series (<= 32
    # interferograms) and chunked ones, each neighbour-list size, nneigh 0
    (7, 97, 131, 20, 0, 8, 0.3),
    (33, 50, 64, 33, 0, 6, 0.5),
    (40, 61, 77, 70, 2, 12, 0.2),
    (1, 30, 45, 1, 0, 3, 0.05),
    (5, 64, 64, 0, 1, 4, 0.1),
    (8, 40, 40, 128, 0, 16, 0.25),
    (65, 23, 29, 10, 0, 5, 0.4),
    (12, 70, 90, 50, 0, 9, 0.2),    # tiled, 64 neighbours, series in registers
    (19, 48, 80, 32, 0, 16, 0.3),   # tiled, offsets beyond the staged halo
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[:3])) + f"-n{c[3]}")
def test_against_c_oracle(ctx, C, case):
    from sentinel_gpu_processor_amd import synthetic as S
    nifg, h, w, nneigh, rdmin, rdmax, th = case
    z, ps, _ = S.ps_stack(nifg, h, w, seed=h + w, ps_frac=0.2, extra_frac=0.15)
    z[:, 3, 5] = 0  # a pixel without signal: unit phasor 0
    got = _gpu(ctx, z, ps, nneigh, rdmin, rdmax, th)
    want = C.ps_similarity(z, ps, nneigh, rdmin, rdmax, th)
    _same(got, want)
    assert 0 < got[2].sum() < got[2].size


def test_masks_all_and_none(ctx, C):
    z, _ = _random(6, 33, 47, seed=2, frac=0.3)
    for ps, th in ((np.ones((33, 47), np.int32), 0.05), (np.zeros((33, 47), np.int32), 0.05),
                   (np.zeros((33, 47), np.int32), 0.0), (np.ones((33, 47), np.int32), 2.0)):
        got = _gpu(ctx, z, ps, 9, 0, 5, th)
        _same(got, C.ps_similarity(z, ps, 9, 0, 5, th))
    # no candidates and a positive threshold: nothing anywhere
    got = _gpu(ctx, z, np.zeros((33, 47), np.int32), 9, 0, 5, 0.05)
    assert not got[0].any() and not got[1].any() and not got[2].any()
    # threshold 0: every median (0 included) passes, every pixel is PS
    got = _gpu(ctx, z, np.zeros((33, 47), np.int32), 9, 0, 5, 0.0)
    assert got[2].all() and np.all(got[1] == 1)


def test_tile_capacity_overflow(ctx, C):
    """Every pixel a candidate with 19 interferograms and rdmax 16: a tile's
    window holds more eligible series than its LDS slots, so part of them
    (and the neighbours past the halo) are read from global memory; the
    result must not change."""
    from sentinel_gpu_processor_amd import synthetic as S
    z, _, _ = S.ps_stack(19, 60, 70, seed=21)
    for ps, th in ((np.ones((60, 70), np.int32), 0.3), (np.ones((60, 70), np.int32), -2.0)):
        got = _gpu(ctx, z, ps, 32, 0, 16, th)
        _same(got, C.ps_similarity(z, ps, 32, 0, 16, th))


def test_stage_entry_points(ctx, C):
    """normalize, median, update and maximum as the reference's separate
    kernels (phase_similarity.cu:116-283) against the oracle's."""
    from sentinel_gpu_processor_amd import kernels as K
    from sentinel_gpu_processor_amd import synthetic as S
    z, ps, _ = S.ps_stack(9, 45, 70, seed=4)
    st = torch.from_numpy(z).cuda()
    zn = K.ps_normalize(ctx, st)
    zn_np = zn.cpu().numpy()
    want_zn = C.ps_normalize(z)
    assert np.array_equal(zn_np.view(np.uint32), want_zn.view(np.uint32))
    assert np.array_equal(P.normalize(z).view(np.uint32), want_zn.view(np.uint32))
    off = C.ps_scan(1, 7)
    pst = torch.from_numpy(ps).cuda()
    med = K.ps_median_similarity(ctx, zn, pst, 15, 1, 7).cpu().numpy()
    want_med = C.ps_median(want_zn, ps, 15, off)
    assert np.array_equal(med.view(np.uint32), want_med.view(np.uint32))
    for ow in (True, False):
        p2 = pst.clone()
        K.ps_update(ctx, p2, torch.from_numpy(med).cuda(), 0.3, overwrite=ow)
        np.testing.assert_array_equal(p2.cpu().numpy(), C.ps_update(ps, med, 0.3, ow))
    mx = K.ps_maximum_similarity(ctx, zn, pst, 15, 1, 7).cpu().numpy()
    want_mx = C.ps_max(want_zn, ps, 15, off)
    assert np.array_equal(mx.view(np.uint32), want_mx.view(np.uint32))
    # in-place normalisation, as the reference does it
    st2 = st.clone()
    K.ps_normalize(ctx, st2, out=st2)
    assert torch.equal(st2, zn)


def test_realistic_stack_selects_scatterers(ctx, C):
    """A 19-interferogram stack of 600 x 1600 pixels: bit-identical to the C
    oracle, and the selection keeps the stable scatterers and drops the
    random candidates."""
    from sentinel_gpu_processor_amd import synthetic as S
    z, ps, truth = S.ps_stack(19, 600, 1600, seed=7, ps_frac=0.15, extra_frac=0.1)
    got = _gpu(ctx, z, ps, 32, 0, 16, 0.5)
    _same(got, C.ps_similarity(z, ps, 32, 0, 16, 0.5))
    sel = got[2] == 1
    assert sel[truth].mean() > 0.95
    assert sel[(ps == 1) & ~truth].mean() < 0.05


def test_repeat_and_table_switch(ctx):
    z, ps = _random(5, 40, 60, seed=9, frac=0.3)
    a = _gpu(ctx, z, ps, 10, 0, 7, 0.1)
    b = _gpu(ctx, z, ps, 10, 2, 9, 0.1)
    c = _gpu(ctx, z, ps, 10, 0, 7, 0.1)
    _same(a, c)
    _same(b, P.similarity(z, ps, 10, 2, 9, 0.1))


def test_errors(ctx):
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    st = torch.zeros((3, 8, 9), dtype=torch.complex64, device="cuda")
    ps = torch.zeros((8, 9), dtype=torch.int32, device="cuda")
    with pytest.raises(N.S1IfgError, match="nneigh"):
        K.ps_similarity(ctx, st, ps, 129, 0, 4, 0.1)
    with pytest.raises(N.S1IfgError, match="rdmax"):
        K.ps_similarity(ctx, st, ps, 8, 0, 5000, 0.1)
    with pytest.raises(ValueError):
        K.ps_similarity(ctx, st, torch.zeros((8, 8), dtype=torch.int32, device="cuda"), 8, 0, 4, 0.1)
    with pytest.raises(TypeError):
        K.ps_similarity(ctx, st, ps.float(), 8, 0, 4, 0.1)


def test_dropin_files(tmp_path, monkeypatch, C):
    """_ps_selection (psps.py:11-78) on files: the list file under temp/,
    raw interferograms and the int32 candidate file in, the two float32
    similarity files out, equal to the oracle's."""
    from sentinel_gpu_processor_amd import psps
    from sentinel_gpu_processor_amd import synthetic as S
    z, ps, _ = S.ps_stack(6, 70, 90, seed=12)
    names = []
    for k in range(6):
        fn = str(tmp_path / f"20200101_2020{k + 2:02d}01.int")
        z[k].tofile(fn)
        names.append(fn)
    ps.tofile(str(tmp_path / "ps.int32"))
    monkeypatch.chdir(tmp_path)
    psps._ps_selection(names, 70, 90, str(tmp_path / "ps.int32"), "med.f32", "max.f32", 12, 0, 6,
                       0.4, 0.8)
    assert (tmp_path / "temp").is_dir()
    lst = list((tmp_path / "temp").iterdir())
    assert len(lst) == 1 and lst[0].read_text().split()[:3] == ["6", "70", "90"]
    med = np.fromfile("med.f32", np.float32).reshape(70, 90)
    mx = np.fromfile("max.f32", np.float32).reshape(70, 90)
    want = C.ps_similarity(z, ps, 12, 0, 6, 0.4)
    assert np.array_equal(med.view(np.uint32), want[0].view(np.uint32))
    assert np.array_equal(mx.view(np.uint32), want[1].view(np.uint32))
    assert psps.main([]) == 0
