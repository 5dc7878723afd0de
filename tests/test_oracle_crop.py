"""The oracle on a crop of the target grid, reading only the window of each
burst the crop samples (tests/configs.py), equals the crop of the oracle on
the whole grid: the GPU tests at benchmark sizes rely on it."""

import numpy as np
import pytest

from tests.configs import crop_grid, oracle_pair_crop
from tests.scenes import oracle_chain, scene


@pytest.mark.parametrize("interp", ["bilinear", "sinc"])
def test_crop_oracle_equals_full_oracle(interp):
    sc = scene("small_misreg")
    oc = oracle_chain("small_misreg", interp)
    r0, c0, nr, nc = 16, 64, 96, 256
    g = crop_grid(sc["grid"], r0, c0, nr, nc)
    ifg, p1, p2, gr, gs = oracle_pair_crop(sc, sc["ref_slc"], sc["sec_slc"], g, interp)
    assert np.count_nonzero(gr) > 0.3 * gr.size
    np.testing.assert_allclose(gr, oc["gr"][r0:r0 + nr, c0:c0 + nc], rtol=0, atol=5e-6)
    np.testing.assert_allclose(gs, oc["gs"][r0:r0 + nr, c0:c0 + nc], rtol=0, atol=5e-6)
    sl = np.s_[r0:r0 + nr, c0 // 4:(c0 + nc) // 4]
    np.testing.assert_allclose(ifg, oc["ifg"][sl], rtol=0, atol=5e-6)
    np.testing.assert_allclose(p1, oc["p1"][sl], rtol=0, atol=5e-6)
