"""The C/OpenMP restatement (oracle/c/s1cpu.c), used for the CPU timing
beside the GPU, against the FP64 NumPy oracle on the test scenes: the
set-up, deramp, reproject (bilinear and sinc, with the offset polynomial and
varying TOPS rates), crossmul + looks, Goldstein and coherence."""

import numpy as np
import pytest

from oracle import cport as C
from oracle import geom as G
from oracle import goldstein as GS
from oracle import ifg as I
from tests.scenes import oracle_chain, phase_rms, scene


@pytest.fixture(scope="module", autouse=True)
def built():
    import os
    if not os.path.exists(C.LIB):
        pytest.skip("oracle/libs1cpu.so not built (make -C oracle all)")


@pytest.mark.parametrize("name", ["small_misreg", "small_tops"])
def test_setup_and_deramp(name):
    sc = scene(name)
    for orbit, b in ((sc["orbit_ref"], sc["burst_ref"]), (sc["orbit_sec"], sc["burst_sec"])):
        kt, er, scal = C.deramp_setup(orbit, b)
        su = G.deramp_setup(orbit, b)
        assert np.allclose(kt, su["kt"], rtol=1e-13, atol=0)
        assert np.allclose(er, su["etaref"], rtol=0, atol=1e-15)
        assert scal[0] == pytest.approx(su["ks"], rel=1e-13)
        assert np.allclose(scal[1:5], su["fm"], rtol=1e-14) and np.allclose(scal[5:9], su["dc"])
    kt, er, _ = C.deramp_setup(sc["orbit_ref"], sc["burst_ref"])
    d = C.deramp(sc["burst_ref"], kt, er, sc["ref_slc"])
    want = oracle_chain(name)["dr"]
    assert np.max(np.abs(d - want)) <= 1e-6 * np.abs(want).max()


@pytest.mark.parametrize("interp", ["bilinear", "sinc"])
@pytest.mark.parametrize("name", ["small_misreg", "small_tops"])
def test_reproject_and_pair(interp, name):
    sc = scene(name)
    oc = oracle_chain(name, interp)
    out = C.pair_chain(sc, sc["ref_slc"], sc["sec_slc"], interp)
    for got, want in ((out["gr"], oc["gr"]), (out["gs"], oc["gs"])):
        assert np.count_nonzero((got != 0) != (want != 0)) <= 2
        both = (got != 0) & (want != 0)
        assert both.mean() > 0.3
        assert np.max(np.abs(got[both] - want[both])) <= 2e-6 * np.abs(want).max()
    m = (oc["p1"] > 0) & (oc["p2"] > 0) & (np.abs(oc["ifg"]) > 1e-3 * np.abs(oc["ifg"]).max())
    assert phase_rms(out["ifg"], oc["ifg"], m) < 1e-5
    assert np.allclose(out["p1"], oc["p1"], rtol=1e-5, atol=1e-7)


def test_goldstein_and_coherence():
    rng = np.random.default_rng(5)
    yy, xx = np.mgrid[0:77, 0:130]
    x = (np.exp(1j * (0.05 * xx + 0.02 * yy * yy / 77)) * (1 + 0.3 * rng.standard_normal(xx.shape))
         + 0.4 * (rng.standard_normal(xx.shape) + 1j * rng.standard_normal(xx.shape)))
    x = x.astype(np.complex64)
    for alpha in (0.5, 0.0, 0.8):
        got = C.goldstein(x, alpha)
        want, wsum = GS.goldstein(x, 32, alpha)
        m = wsum > 0
        assert np.array_equal(got[~m], np.zeros(np.count_nonzero(~m), np.complex64))
        assert np.max(np.abs(got[m] - want[m])) < 2e-5 * np.abs(want[m]).max()
    p1 = np.abs(x) ** 2 + 0.1
    p2 = np.abs(x[::-1]) ** 2 + 0.1
    for win in (1, 5):
        got = C.coherence(x, p1, p2, win)
        want = I.coherence_window(x, p1.astype(np.float32), p2.astype(np.float32), win)
        assert np.max(np.abs(got - want)) < 1e-6
