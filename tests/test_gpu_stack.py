"""20-date stack (19 sequential pairs) with per-date atmospheric phase
screens through one batch: per-pair RMS phase and coherence error against
the FP64 oracle (on a crop of each pair), Goldstein as FP16 split on MFMA,
and bit-identity of every pair with the single-pair entry points."""

import numpy as np
import pytest

from oracle import geom as G
from oracle import goldstein as GS
from oracle import ifg as I
from tests.scenes import phase_rms

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

N_DATES = 20
LINES, SAMPLES = 128, 1024
CROP = (32, 96, 128, 768)     # rows r0:r1, columns c0:c1 of the target grid (look aligned)


@pytest.fixture(scope="module")
def ctx():
    from sentinel_gpu_processor_amd import _native as N
    c = N.Context(0)
    yield c
    c.sync()
    c.close()


@pytest.fixture(scope="module")
def stack(ctx):
    from sentinel_gpu_processor_amd import synthetic as S
    from sentinel_gpu_processor_amd.stack import Stack
    st = S.make_stack(n_dates=N_DATES, lines=LINES, samples=SAMPLES, dem_shape=(64, 96),
                      dem_kind="hill", hmax=300.0, seed=3, fidelity="fast", device="cuda:0")
    dem_t = torch.from_numpy(st["dem"]["data"]).cuda()
    run = Stack(ctx, st["dates"], st["grid"], dem_t, st["dem"], interp="sinc", rl=1, cl=4,
                gold_mode="fp16_split", coh_win=5)
    run.run(graph=True)
    torch.cuda.synchronize()
    yield st, run, dem_t
    run.close()


def _bits(t):
    a = t.cpu().numpy()
    return a.view(np.uint64) if a.dtype == np.complex64 else a.view(np.uint32)


def test_stack_pairs_bit_identical_to_single_pairs(ctx, stack):
    from sentinel_gpu_processor_amd import kernels as K
    st, run, dem_t = stack
    assert len(run.pairs) == N_DATES - 1
    assert len(run.deramped) == N_DATES          # each date deramped once
    prods = run.products()
    for i in (0, 7, N_DATES - 2):
        a, b = run.pairs[i]
        dr = K.deramp(ctx, run.geoms[a], run.raw[a])
        ds = K.deramp(ctx, run.geoms[b], run.raw[b])
        ifg, pr, ps = K.pair(ctx, run.geoms[a], run.geoms[b], st["grid"], dem_t, st["dem"], dr,
                             ds, "sinc", 1, 4)
        filt = K.goldstein(ctx, ifg, 0.5, mode="fp16_split")
        coh = K.coherence(ctx, ifg, pr, ps, win=5)
        torch.cuda.synchronize()
        for key, want in (("ifg", ifg), ("pow_ref", pr), ("pow_sec", ps), ("filt", filt),
                          ("coh", coh)):
            assert np.array_equal(_bits(prods[i][key]), _bits(want)), (i, key)


def test_stack_errors_against_oracle_per_pair(stack):
    """RMS interferometric phase error, coherence error and Goldstein
    (FP16 split) phase error of every pair on a crop, against the FP64
    oracle chain; all below 1e-3."""
    from sentinel_gpu_processor_amd import _native as N
    from sentinel_gpu_processor_amd import kernels as K
    st, run, dem_t = stack
    r0, r1, c0, c1 = CROP
    g = dict(st["grid"])
    g["lat0"] = g["lat0"] + r0 * g["dlat"]
    g["lon0"] = g["lon0"] + c0 * g["dlon"]
    g["nrow"], g["ncol"] = r1 - r0, c1 - c0
    h = G.target_heights(g, st["dem"])
    prods = run.products()
    setups, der = {}, {}

    def date(k):
        if k not in setups:
            d = st["dates"][k]
            setups[k] = G.deramp_setup(d["orbit"], d["burst"])
            der[k] = G.deramp(d["slc"].cpu().numpy(), setups[k]).astype(np.complex64)
        return setups[k], der[k]

    ctx = run.ctx
    rows = []
    for i, (a, b) in enumerate(run.pairs):
        sa, da = date(a)
        sb, db = date(b)
        oa, _ = G.reproject(st["dates"][a]["orbit"], st["dates"][a]["burst"], sa, g, h, da, "sinc")
        ob, _ = G.reproject(st["dates"][b]["orbit"], st["dates"][b]["burst"], sb, g, h, db, "sinc")
        ifg_o = I.cpx_looks(I.conj_mul(oa, ob), 1, 4)
        p1 = I.cpx_looks(np.abs(oa) ** 2, 1, 4).real
        p2 = I.cpx_looks(np.abs(ob) ** 2, 1, 4).real
        coh_o = I.coherence_window(ifg_o, p1, p2, 5)
        ifg_g = prods[i]["ifg"][r0:r1, c0 // 4:c1 // 4].contiguous()
        coh_g = prods[i]["coh"][r0:r1, c0 // 4:c1 // 4].cpu().numpy()
        ig = ifg_g.cpu().numpy()
        m = (p1 > 0) & (p2 > 0) & (np.abs(ifg_o) > 1e-3 * np.abs(ifg_o).max())
        e_ph = phase_rms(ig, ifg_o, m)
        inner = np.zeros_like(m)
        inner[2:-2, 2:-2] = True
        e_coh = float(np.max(np.abs(coh_g[inner & m] - coh_o[inner & m])))
        f_g = K.goldstein(ctx, ifg_g, 0.5, mode="fp16_split").cpu().numpy()
        f_o, wsum = GS.goldstein(ifg_o, 32, 0.5)
        mf = (wsum > 0) & (np.abs(f_o) > 1e-3 * np.abs(f_o).max())
        e_gold = phase_rms(f_g, f_o, mf)
        rows.append((i, a, b, e_ph, e_coh, e_gold))
    print("\npair dates  phase_rms_rad  coh_max_err  goldstein_fp16split_phase_rms_rad")
    for i, a, b, e1, e2, e3 in rows:
        print(f"{i:4d} {a:2d}-{b:2d}  {e1:.3e}  {e2:.3e}  {e3:.3e}")
    for i, a, b, e1, e2, e3 in rows:
        assert e1 < 1e-3 and e2 < 1e-3 and e3 < 1e-3, (i, e1, e2, e3)
