"""s1ifg_pair_noalloc, the fused pair that allocates no device memory (its
grid tables in the library's fixed device array), against s1ifg_pair on the
same inputs: bit-identical products on two calls in a row, the second on a
smaller grid; the context's workspace stays empty; a grid whose tables
exceed the fixed array is refused on the host."""

import numpy as np
import pytest

from tests.configs import crop_grid
from tests.scenes import scene

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _np(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("interp", ["sinc", "bilinear"])
def test_pair_noalloc_matches_pair(interp):
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("small_misreg")
    ctx = N.Context(0)   # a fresh context: no workspace yet
    ctx_ref = N.Context(0)
    gr = N.Geometry(ctx, sc["orbit_ref"], sc["burst_ref"])
    gs = N.Geometry(ctx, sc["orbit_sec"], sc["burst_sec"])
    dem = sc["dem"]
    dem_t = _dev(dem["data"])
    dref = K.deramp(ctx_ref, gr, _dev(sc["ref_slc"]))
    dsec = K.deramp(ctx_ref, gs, _dev(sc["sec_slc"]))
    grids = [sc["grid"], crop_grid(sc["grid"], 8, 64, 100, 320)]  # second call smaller
    for grid in grids:
        print(f"pair_noalloc {interp} grid {grid['nrow']} x {grid['ncol']}", flush=True)
        shape = (grid["nrow"], grid["ncol"] // 4)
        out = (torch.full(shape, np.nan, dtype=torch.complex64, device="cuda:0"),
               torch.full(shape, np.nan, dtype=torch.float32, device="cuda:0"),
               torch.full(shape, np.nan, dtype=torch.float32, device="cuda:0"))
        K.pair(ctx, gr, gs, grid, dem_t, dem, dref, dsec, interp, 1, 4,
               offpoly=sc["offpoly"], out=out, noalloc=True)
        want = K.pair(ctx_ref, gr, gs, grid, dem_t, dem, dref, dsec, interp, 1, 4,
                      offpoly=sc["offpoly"])
        for a, b in zip(out, want):
            a, b = _np(a), _np(b)
            assert np.isfinite(a).all()
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        assert np.count_nonzero(_np(out[1])) > 0.2 * _np(out[1]).size
        assert ctx.workspace_bytes() == 0  # nothing allocated on the no-alloc context
    gr.close()
    gs.close()
    ctx.close()
    ctx_ref.close()


def test_pair_noalloc_refuses_large_grids():
    """Tables beyond 2 MiB: refused before any launch, with the limit in the
    message."""
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    sc = scene("small_misreg")
    ctx = N.Context(0)
    gr = N.Geometry(ctx, sc["orbit_ref"], sc["burst_ref"])
    gs = N.Geometry(ctx, sc["orbit_sec"], sc["burst_sec"])
    dem = sc["dem"]
    dem_t = _dev(dem["data"])
    z = torch.zeros((sc["burst_ref"]["lines"], sc["burst_ref"]["samples"]), dtype=torch.complex64,
                    device="cuda:0")
    big = dict(sc["grid"], nrow=16, ncol=90000)  # 24 * 90000 B of column table > 2 MiB
    shape = (big["nrow"], big["ncol"] // 4)
    out = (torch.empty(shape, dtype=torch.complex64, device="cuda:0"),
           torch.empty(shape, dtype=torch.float32, device="cuda:0"),
           torch.empty(shape, dtype=torch.float32, device="cuda:0"))
    with pytest.raises(N.S1IfgError) as e:
        K.pair(ctx, gr, gs, big, dem_t, dem, z, z, "sinc", 1, 4, out=out, noalloc=True)
    assert e.value.code == N.S1IFG_ERANGE
    assert "2097152" in str(e.value)
    assert ctx.workspace_bytes() == 0
    ctx.close()
