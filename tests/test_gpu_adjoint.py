"""Adjoint kernels (s1ifg_pair_backward, s1ifg_coherence_backward) and the
torch.autograd wrapper against the FP64 oracle gradient and against
central finite differences of the GPU forward."""

import numpy as np
import pytest

from oracle import adjoint as A
from tests.test_oracle_adjoint import _case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from sentinel_gpu_processor_amd import _native as N
    c = N.Context(0)
    yield c
    c.sync()
    c.close()


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.fixture(scope="module")
def setup(ctx):
    from sentinel_gpu_processor_amd import _native as N
    sc, setups, dr, ds, h0, dh = _case()
    gr = N.Geometry(ctx, sc["orbit_ref"], sc["burst_ref"])
    gs = N.Geometry(ctx, sc["orbit_sec"], sc["burst_sec"])
    return dict(sc=sc, setups=setups, dr=dr, ds=ds, h0=h0, dh=dh, gr=gr, gs=gs,
                dem_t=_dev(sc["dem"]["data"]), dr_t=_dev(dr), ds_t=_dev(ds), dh_t=_dev(dh))


def _oracle(S, interp, win):
    sc = S["sc"]
    p = np.asarray(sc["offpoly"], np.float64)
    h = S["h0"] + S["dh"].astype(np.float64)
    F = A.pair_forward(sc, S["setups"], h, S["dr"], S["ds"], interp, p, 1, 4)
    mask = A.interior_mask(F, 1, 4, win)
    loss, gp, gh = A.grad_neg_mean_coherence(sc, S["setups"], h, S["dr"], S["ds"], interp, p, 1,
                                             4, win, mask)
    return mask, loss, gp, gh


def _gpu_loss_grad(ctx, S, interp, win, mask, p=None):
    from sentinel_gpu_processor_amd import adjoint as AD
    sc = S["sc"]
    p = np.asarray(sc["offpoly"], np.float64) if p is None else p
    off = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    dh = S["dh_t"].clone().requires_grad_(True)
    coh = AD.pair_coherence(ctx, S["gr"], S["gs"], sc["grid"], S["dem_t"], sc["dem"],
                            S["dr_t"], S["ds_t"], off, dh, interp=interp, rl=1, cl=4, win=win)
    m = torch.from_numpy(mask).cuda()
    loss = -coh.double()[m].mean()
    loss.backward()
    torch.cuda.synchronize()
    return float(loss), off.grad.numpy().copy(), dh.grad.cpu().numpy()


@pytest.mark.parametrize("interp,win", [("sinc", 5), ("bilinear", 5), ("sinc", 1)])
def test_gradients_match_oracle(ctx, setup, interp, win):
    print(f"adjoint vs oracle: {interp} win={win}", flush=True)
    mask, l_or, gp_or, gh_or = _oracle(setup, interp, win)
    loss, gp, gh = _gpu_loss_grad(ctx, setup, interp, win, mask)
    assert loss == pytest.approx(l_or, rel=1e-3)
    assert np.linalg.norm(gp - gp_or) <= 1e-2 * np.linalg.norm(gp_or), (gp, gp_or)
    assert np.max(np.abs(gh - gh_or)) <= 1e-2 * np.max(np.abs(gh_or))
    # direction of the height gradient field
    c = np.sum(gh * gh_or) / np.sqrt(np.sum(gh * gh) * np.sum(gh_or * gh_or))
    assert c > 0.9999


def test_offpoly_gradient_matches_gpu_finite_differences(ctx, setup):
    """Central differences (1e-3 px) of the GPU forward itself."""
    from sentinel_gpu_processor_amd import kernels as K
    S = setup
    sc = S["sc"]
    mask, _, _, _ = _oracle(S, "sinc", 5)
    m = torch.from_numpy(mask).cuda()
    p = np.asarray(sc["offpoly"], np.float64)

    def loss(pp):
        ifg, pr, ps = K.pair(ctx, S["gr"], S["gs"], sc["grid"], S["dem_t"], sc["dem"], S["dr_t"],
                             S["ds_t"], "sinc", 1, 4, dh=S["dh_t"], offpoly=pp)
        coh = K.coherence(ctx, ifg, pr, ps, win=5)
        return -float(coh.double()[m].mean())

    _, gp, _ = _gpu_loss_grad(ctx, S, "sinc", 5, mask)
    for k in range(6):
        e = 1e-3
        d = np.zeros(6)
        d[k] = e
        fd = (loss(p + d) - loss(p - d)) / (2 * e)
        assert gp[k] == pytest.approx(fd, rel=1e-2, abs=2e-5), (k, gp[k], fd)


def test_backward_is_deterministic_and_matches_direct_call(ctx, setup):
    from sentinel_gpu_processor_amd import kernels as K
    S = setup
    sc = S["sc"]
    ifg, pr, ps = K.pair(ctx, S["gr"], S["gs"], sc["grid"], S["dem_t"], sc["dem"], S["dr_t"],
                         S["ds_t"], "sinc", 1, 4, dh=S["dh_t"], offpoly=sc["offpoly"])
    g = torch.full(ifg.shape, -1.0 / ifg.numel(), device="cuda")
    outs = []
    for _ in range(2):
        gi, g1, g2 = K.coherence_backward(ctx, ifg, pr, ps, g, win=5)
        go, gd = K.pair_backward(ctx, S["gr"], S["gs"], sc["grid"], S["dem_t"], sc["dem"],
                                 S["dr_t"], S["ds_t"], gi, g1, g2, "sinc", 1, 4, dh=S["dh_t"],
                                 offpoly=sc["offpoly"])
        torch.cuda.synchronize()
        outs.append((go.cpu().numpy(), gd.cpu().numpy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    # the autograd path gives the same numbers for loss = -mean(coherence)
    from sentinel_gpu_processor_amd import adjoint as AD
    off = torch.tensor(np.asarray(sc["offpoly"], np.float64), requires_grad=True)
    dh = S["dh_t"].clone().requires_grad_(True)
    coh = AD.pair_coherence(ctx, S["gr"], S["gs"], sc["grid"], S["dem_t"], sc["dem"], S["dr_t"],
                            S["ds_t"], off, dh, interp="sinc", win=5)
    (-coh.mean()).backward()
    torch.cuda.synchronize()
    assert np.allclose(off.grad.numpy(), outs[0][0], rtol=1e-6, atol=1e-12)
    assert np.allclose(dh.grad.cpu().numpy(), outs[0][1], rtol=1e-5, atol=1e-12)


def test_coherence_backward_matches_oracle(ctx):
    from sentinel_gpu_processor_amd import kernels as K
    rng = np.random.default_rng(3)
    shape = (37, 53)
    ifg = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    p1 = (np.abs(ifg) * rng.uniform(1.1, 2.0, shape)).astype(np.float32)
    p2 = (np.abs(ifg) * rng.uniform(1.1, 2.0, shape)).astype(np.float32)
    ifg[:4, :6] = 0
    p1[:4, :6] = 0
    p2[:4, :6] = 0
    g = rng.standard_normal(shape).astype(np.float32)
    for win in (1, 3, 5):
        gi, g1, g2 = K.coherence_backward(ctx, _dev(ifg), _dev(p1), _dev(p2), _dev(g), win=win)
        torch.cuda.synchronize()
        wi, w1, w2 = A.coherence_backward(ifg.astype(np.complex128), p1.astype(np.float64),
                                          p2.astype(np.float64), win, g.astype(np.float64))
        for got, want in ((gi, wi), (g1, w1), (g2, w2)):
            got = got.cpu().numpy()
            assert np.max(np.abs(got - want)) <= 2e-5 * np.max(np.abs(want)) + 1e-7, win
