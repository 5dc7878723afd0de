"""Differentiable burst pair: torch.autograd.Function wrappers around the
hand-written adjoint kernels (s1ifg_pair_backward, s1ifg_coherence_backward).

Gradients flow to the secondary's 6-coefficient offset polynomial
(d_rg = p0 + p1 rg/N + p2 az/L, d_az = p3 + p4 rg/N + p5 az/L, applied to the
sampling position and the reramp) and to the per-target-pixel height error
dh added to the DEM height.  The backward pass recomputes the geometry and
the sinc taps (and their derivatives) instead of storing them.  No reference
counterpart: the reference has no gradient path.

    offpoly = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    dh = torch.zeros(grid_shape, device="cuda", requires_grad=True)
    coh = pair_coherence(ctx, g_ref, g_sec, grid, dem_t, dem, ref_d, sec_d,
                         offpoly, dh, win=5)
    loss = -coh.mean()
    loss.backward()          # offpoly.grad, dh.grad
"""

import torch

from . import kernels as K


class PairFunction(torch.autograd.Function):
    """(offpoly[6], dh) -> (ifg_ml, pow_ref, pow_sec) of the fused pair."""

    @staticmethod
    def forward(fctx, offpoly, dh, ctx, g_ref, g_sec, grid, dem_t, dem, ref_d, sec_d, interp,
                rl, cl):
        op = offpoly.detach().to("cpu", torch.float64).numpy()
        dh_ = None if dh is None else dh.detach().contiguous()
        ifg, pr, ps = K.pair(ctx, g_ref, g_sec, grid, dem_t, dem, ref_d, sec_d, interp, rl, cl,
                             dh=dh_, offpoly=op)
        fctx.args = (ctx, g_ref, g_sec, grid, dem_t, dem, ref_d, sec_d, interp, rl, cl, op)
        fctx.dh = dh_
        fctx.offpoly_meta = (offpoly.device, offpoly.dtype)
        fctx.has_dh = dh is not None
        return ifg, pr, ps

    @staticmethod
    def backward(fctx, g_ifg, g_pr, g_ps):
        ctx, g_ref, g_sec, grid, dem_t, dem, ref_d, sec_d, interp, rl, cl, op = fctx.args
        c = lambda t: None if t is None else t.contiguous()
        want_dh = fctx.has_dh and fctx.needs_input_grad[1]
        g_off, g_dh = K.pair_backward(ctx, g_ref, g_sec, grid, dem_t, dem, ref_d, sec_d,
                                      c(g_ifg), c(g_pr), c(g_ps), interp, rl, cl, dh=fctx.dh,
                                      offpoly=op, want_dh=want_dh)
        dev, dt = fctx.offpoly_meta
        g_off = g_off.to(dev, dt) if fctx.needs_input_grad[0] else None
        return (g_off, g_dh) + (None,) * 11


class CoherenceFunction(torch.autograd.Function):
    """(ifg_ml, pow_ref, pow_sec) -> coherence (look-cell for win = 1,
    windowed otherwise)."""

    @staticmethod
    def forward(fctx, ifg, pr, ps, ctx, win, fill, seed):
        coh = K.coherence(ctx, ifg.detach(), pr.detach(), ps.detach(), win=win, fill=fill,
                          seed=seed)
        fctx.save_for_backward(ifg, pr, ps)
        fctx.ctx, fctx.win = ctx, win
        return coh

    @staticmethod
    def backward(fctx, g):
        ifg, pr, ps = fctx.saved_tensors
        gi, g1, g2 = K.coherence_backward(fctx.ctx, ifg.detach(), pr.detach(), ps.detach(),
                                          g.contiguous(), win=fctx.win)
        return gi, g1, g2, None, None, None, None


def pair_products(ctx, g_ref, g_sec, grid, dem_t, dem, ref_d, sec_d, offpoly, dh=None,
                  interp="sinc", rl=1, cl=4):
    """Differentiable (ifg_ml, pow_ref, pow_sec)."""
    return PairFunction.apply(offpoly, dh, ctx, g_ref, g_sec, grid, dem_t, dem, ref_d, sec_d,
                              interp, rl, cl)


def pair_coherence(ctx, g_ref, g_sec, grid, dem_t, dem, ref_d, sec_d, offpoly, dh=None,
                   interp="sinc", rl=1, cl=4, win=5, fill="clamp", seed=0):
    """Differentiable coherence of the burst pair."""
    ifg, pr, ps = pair_products(ctx, g_ref, g_sec, grid, dem_t, dem, ref_d, sec_d, offpoly, dh,
                                interp, rl, cl)
    return CoherenceFunction.apply(ifg, pr, ps, ctx, win, fill, seed)


def interior_cells(ctx, g_ref, g_sec, grid, dem_t, dem, ref_d, sec_d, interp, offpoly, dh, rl,
                   cl, win, margin=2):
    """Multilooked pixels whose coherence window sees only target pixels
    valid in both bursts with `margin` valid pixels around them (device bool
    tensor): finite differences of the loss there cannot cross a validity
    edge, where the products jump."""
    import torch.nn.functional as F
    a = K.geocode(ctx, g_ref, grid, dem_t, dem, ref_d, interp, dh=dh)
    b = K.geocode(ctx, g_sec, grid, dem_t, dem, sec_d, interp, dh=dh, offpoly=offpoly)
    bad = ((a == 0) | (b == 0)).float()[None, None]
    del a, b
    bad = F.max_pool2d(bad, 2 * margin + 1, stride=1, padding=margin)[0, 0] > 0
    nr, nc = grid["nrow"] // rl, grid["ncol"] // cl
    cell = ~bad[:nr * rl, :nc * cl].reshape(nr, rl, nc, cl).any(dim=3).any(dim=1)
    h = win // 2
    badc = F.max_pool2d((~cell).float()[None, None], 2 * h + 1, stride=1, padding=h)[0, 0] > 0
    cell = ~badc
    cell[:h, :] = False
    cell[-h:, :] = False
    cell[:, :h] = False
    cell[:, -h:] = False
    return cell
