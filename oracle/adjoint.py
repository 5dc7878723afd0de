"""FP64 NumPy adjoint of the fused burst pair and of coherence (test
infrastructure).  No reference counterpart: the reference has no gradient
path.  The forward it differentiates is oracle/geom.py reproject (sampling
position and reramp at the offset-polynomial-shifted position, flattening at
the geometric range) followed by crossmul and look means (oracle/ifg.py)
and the windowed or look-cell coherence.

Parity pins (tests/test_oracle_adjoint.py): central finite differences of the
FP64 forward, per offset-polynomial coefficient and along random directions
of the height-error field; the sinc tap derivative against finite
differences of the taps.

Conventions: the gradient of a real loss with respect to a complex value z
is dL/dRe z + i dL/dIm z (the convention torch uses for complex tensors).
"""

import numpy as np

from oracle import geom as G
from oracle import ifg as I

SOL = G.SOL


# ---------------------------------------------------------------- sinc taps
def sinc_taps_deriv(frac):
    """(w, dw/dfrac) of oracle.geom.sinc_taps: w_k = h(f - k) / sum_j h(f - j),
    h(x) = sinc(x) (1 + cos(pi x / 4)) / 2."""
    frac = np.asarray(frac, np.float64)
    k = np.arange(-G.SINC_HALF + 1, G.SINC_HALF + 1)
    x = frac[..., None] - k
    s = np.sinc(x)
    with np.errstate(invalid="ignore", divide="ignore"):
        ds = np.where(x == 0, 0.0, (np.cos(np.pi * x) - s) / np.where(x == 0, 1.0, x))
    win = 0.5 * (1 + np.cos(np.pi * x / 4))
    dwin = -0.5 * np.pi / 4 * np.sin(np.pi * x / 4)
    h = s * win
    dh = ds * win + s * dwin
    S = h.sum(axis=-1, keepdims=True)
    dS = dh.sum(axis=-1, keepdims=True)
    return h / S, (dh - h / S * dS) / S


# ------------------------------------------------------------- one burst
def _orbit_va(orbit, t, delta=1e-3):
    torb, xx, vv = orbit
    _, v = G.intp_orbit(torb, xx, vv, t)
    _, vp = G.intp_orbit(torb, xx, vv, t + delta)
    _, vm = G.intp_orbit(torb, xx, vv, t - delta)
    return v, (vp - vm) / (2 * delta)


def burst_terms(orbit, burst, setup, grid, heights, deramped, interp, offpoly,
                round_blend=True):
    """Forward value of one burst at every target pixel plus the partial
    derivatives the adjoint needs (all FP64, full target grid; zeros where
    invalid)."""
    geo = G.geo2rdr_grid(orbit, burst, grid, heights)
    out, _ = G.reproject(orbit, burst, setup, grid, heights, deramped, interp, offpoly, geo,
                         round_blend)
    shape = out.shape
    rng, t = geo["rng"], geo["tline"]
    rgoff0, azoff0 = geo["rgoff"], geo["azoff"]
    N, L = burst["samples"], burst["lines"]
    p = np.zeros(6) if offpoly is None else np.asarray(offpoly, np.float64)
    xs, ys = rgoff0 / N, azoff0 / L
    drg = p[0] + p[1] * xs + p[2] * ys
    daz = p[3] + p[4] * xs + p[5] * ys
    rg, az = rgoff0 + drg, azoff0 + daz
    dtaz = 1.0 / burst["prf"]
    dmrg = SOL / 2.0 / burst["range_sampling_rate"]
    valid = out != 0
    # interpolation derivatives at the sampling position (deramped data)
    z = deramped.reshape(-1).astype(np.complex128)
    I_rg = np.zeros(shape, np.complex128)
    I_az = np.zeros(shape, np.complex128)
    v = valid
    if interp == "bilinear":
        intr = rg[v].astype(np.int64)
        inta = az[v].astype(np.int64)
        fr, fa = rg[v] - intr, az[v] - inta
        i1 = inta * N + intr
        b1, b2, b3, b4 = z[i1], z[i1 + 1], z[i1 + N], z[i1 + N + 1]
        I_rg[v] = (b2 - b1) * (1 - fa) + (b4 - b3) * fa
        I_az[v] = (b3 * (1 - fr) + b4 * fr) - (b1 * (1 - fr) + b2 * fr)
    else:
        intr = np.floor(rg[v]).astype(np.int64)
        inta = np.floor(az[v]).astype(np.int64)
        wr, dwr = sinc_taps_deriv(rg[v] - intr)
        wa, dwa = sinc_taps_deriv(az[v] - inta)
        acc_r = np.zeros(intr.shape, np.complex128)
        acc_a = np.zeros(intr.shape, np.complex128)
        for ia in range(8):
            row = (inta + ia - 3) * N
            s = np.zeros(intr.shape, np.complex128)
            ds = np.zeros(intr.shape, np.complex128)
            for ir in range(8):
                zz = z[row + intr + ir - 3]
                s += wr[:, ir] * zz
                ds += dwr[:, ir] * zz
            acc_r += wa[:, ia] * ds
            acc_a += dwa[:, ia] * s
        I_rg[v] = acc_r
        I_az[v] = acc_a
    # phase and its partials (reramp at the sampling position)
    rp = rng + drg * dmrg
    fm, dc, ks = setup["fm"], setup["dc"], setup["ks"]
    tau = rp * 2 / SOL
    ka = fm[1] + (tau - fm[0]) * (fm[2] + fm[3] * (tau - fm[0]))
    dka = 2 / SOL * (fm[2] + 2 * fm[3] * (tau - fm[0]))
    fnc = dc[1] + (tau - dc[0]) * (dc[2] + dc[3] * (tau - dc[0]))
    dfnc = 2 / SOL * (dc[2] + 2 * dc[3] * (tau - dc[0]))
    kt = ka * ks / (ka - ks)
    dkt = -ks * ks / (ka - ks) ** 2 * dka
    etaref = -fnc / ka - setup["etac0"]
    detaref = -(dfnc * ka - fnc * dka) / (ka * ka)
    eta = -L * dtaz / 2.0 + t - burst["t_start"] + daz * dtaz
    ed = eta - etaref
    R_rp = -np.pi * (dkt * ed * ed - 2 * kt * ed * detaref)
    R_eta = -np.pi * kt * 2 * ed
    phase = 4 * np.pi / burst["wavelength"] * rng + np.pi * kt * ed * ed
    rot = np.exp(1j * phase)
    vel, acc = _orbit_va(orbit, t)
    lat = (grid["lat0"] + np.arange(grid["nrow"]) * grid["dlat"])[:, None]
    lon = (grid["lon0"] + np.arange(grid["ncol"]) * grid["dlon"])[None, :]
    X = G.llh2xyz(lat, lon, heights)
    dr = X - G.intp_orbit(orbit[0], orbit[1], orbit[2], t)[0]
    nvec = np.stack(np.broadcast_arrays(np.cos(np.radians(lat)) * np.cos(np.radians(lon)),
                                        np.cos(np.radians(lat)) * np.sin(np.radians(lon)),
                                        np.sin(np.radians(lat)) + 0 * lon), axis=-1)
    vn = np.sum(vel * nvec, axis=-1)
    den = np.sum(vel * vel, axis=-1) - np.sum(dr * acc, axis=-1)
    dt_dh = vn / den
    drng_dh = np.sum(dr * nvec, axis=-1) / rng
    return dict(out=out, valid=valid, I_rg=I_rg * rot, I_az=I_az * rot, R_rp=R_rp, R_eta=R_eta,
                p=p, xs=xs, ys=ys, dtaz=dtaz, dmrg=dmrg, wvl=burst["wavelength"],
                dt_dh=dt_dh, drng_dh=drng_dh, N=N, L=L)


def burst_adjoint(T, G_out):
    """Given G_out = dL/dout (complex, target grid), return (g_p[6], g_h)
    for one burst's terms T from burst_terms."""
    v = T["valid"]
    Gc = np.where(v, G_out, 0)
    gRG = np.real(np.conj(Gc) * T["I_rg"])
    gAZ = np.real(np.conj(Gc) * T["I_az"])
    gPHI = -np.imag(np.conj(Gc) * T["out"])
    gRNGf = gPHI * 4 * np.pi / T["wvl"]
    gRP = -gPHI * T["R_rp"]
    gETA = -gPHI * T["R_eta"]
    gDRG = gRG + gRP * T["dmrg"]
    gDAZ = gAZ + gETA * T["dtaz"]
    p, xs, ys = T["p"], T["xs"], T["ys"]
    g_p = np.array([gDRG.sum(), (gDRG * xs).sum(), (gDRG * ys).sum(),
                    gDAZ.sum(), (gDAZ * xs).sum(), (gDAZ * ys).sum()])
    g_rgoff = gRG + gDRG * p[1] / T["N"] + gDAZ * p[4] / T["N"]
    g_azoff = gAZ + gDRG * p[2] / T["L"] + gDAZ * p[5] / T["L"]
    g_rng = gRNGf + g_rgoff / T["dmrg"] + gRP
    g_t = g_azoff / T["dtaz"] + gETA
    g_h = g_t * T["dt_dh"] + g_rng * T["drng_dh"]
    return g_p, np.where(v, g_h, 0.0)


# ---------------------------------------------------------- pair + looks
def pair_forward(sc, setups, heights, dr, ds, interp, offpoly, rl, cl, round_blend=True):
    Tr = burst_terms(sc["orbit_ref"], sc["burst_ref"], setups[0], sc["grid"], heights, dr,
                     interp, None, round_blend)
    Ts = burst_terms(sc["orbit_sec"], sc["burst_sec"], setups[1], sc["grid"], heights, ds,
                     interp, offpoly, round_blend)
    zr, zs = Tr["out"], Ts["out"]
    ifg = I.cpx_looks(I.conj_mul(zr, zs), rl, cl)
    pr = I.cpx_looks(np.abs(zr) ** 2, rl, cl).real
    ps = I.cpx_looks(np.abs(zs) ** 2, rl, cl).real
    return dict(ifg=ifg, pr=pr, ps=ps, Tr=Tr, Ts=Ts)


def pair_backward(F, g_ifg, g_pr, g_ps, rl, cl):
    """dL/d offpoly (6) and dL/d dh (target grid) from the gradients of the
    look-mean products."""
    zr, zs = F["Tr"]["out"], F["Ts"]["out"]
    nr, nc = g_ifg.shape
    n = rl * cl
    up = lambda a: np.repeat(np.repeat(a, rl, axis=0), cl, axis=1)
    gI = np.zeros(zr.shape, np.complex128)
    gp1 = np.zeros(zr.shape)
    gp2 = np.zeros(zr.shape)
    gI[:nr * rl, :nc * cl] = up(g_ifg) / n
    gp1[:nr * rl, :nc * cl] = up(g_pr) / n
    gp2[:nr * rl, :nc * cl] = up(g_ps) / n
    G_r = np.conj(gI) * zs + 2 * gp1 * zr
    G_s = gI * zr + 2 * gp2 * zs
    _, gh_r = burst_adjoint(F["Tr"], G_r)
    gp, gh_s = burst_adjoint(F["Ts"], G_s)
    return gp, gh_r + gh_s


# -------------------------------------------------------------- coherence
def coherence_forward(ifg, p1, p2, win):
    """win = 1: the look-cell estimator |ifg| / (sqrt(p1)+1e-12) / (sqrt(p2)+1e-12);
    win >= 3: |sum_W ifg| / sqrt(sum_W p1 * sum_W p2)."""
    if win == 1:
        return I.coherence_lookcell(ifg, np.sqrt(p1), np.sqrt(p2), fill="none")
    return I.coherence_window(ifg, p1, p2, win)


def coherence_backward(ifg, p1, p2, win, g):
    """(g_ifg, g_p1, g_p2) from g = dL/dgamma; zero where gamma > 1 (filled)
    or where the denominator or |ifg| is 0."""
    if win == 1:
        a1 = np.sqrt(p1) + 1e-12
        a2 = np.sqrt(p2) + 1e-12
        mag = np.abs(ifg)
        gam = mag / a1 / a2
        ok = (gam <= 1) & (mag > 0) & (p1 > 0) & (p2 > 0)
        gi = np.where(ok, g / (a1 * a2) * ifg / np.where(mag > 0, mag, 1), 0)
        g1 = np.where(ok, -g * gam / a1 * 0.5 / np.sqrt(np.where(p1 > 0, p1, 1)), 0)
        g2 = np.where(ok, -g * gam / a2 * 0.5 / np.sqrt(np.where(p2 > 0, p2, 1)), 0)
        return gi, g1, g2
    A = I.boxsum(ifg, win)
    B = I.boxsum(p1, win)
    C = I.boxsum(p2, win)
    den = np.sqrt(B * C)
    mag = np.abs(A)
    ok = (den > 0) & (mag > 0)
    gam = np.where(ok, mag / np.where(ok, den, 1), 0)
    ok &= gam <= 1
    gA = np.where(ok, g * A / np.where(ok, mag * den, 1), 0)
    gB = np.where(ok, -0.5 * g * gam / np.where(ok, B, 1), 0)
    gC = np.where(ok, -0.5 * g * gam / np.where(ok, C, 1), 0)
    return I.boxsum(gA, win), I.boxsum(gB, win), I.boxsum(gC, win)


def loss_neg_mean_coherence(sc, setups, heights, dr, ds, interp, offpoly, rl, cl, win,
                            mask=None, round_blend=True):
    """loss = -mean(coherence) over the multilooked grid (or over `mask`)."""
    F = pair_forward(sc, setups, heights, dr, ds, interp, offpoly, rl, cl, round_blend)
    gam = coherence_forward(F["ifg"], F["pr"], F["ps"], win)
    sel = gam if mask is None else gam[mask]
    return -float(np.mean(sel)), F, gam


def grad_neg_mean_coherence(sc, setups, heights, dr, ds, interp, offpoly, rl, cl, win,
                            mask=None, round_blend=True):
    """(loss, g_offpoly[6], g_dh) of loss = -mean(coherence) (over `mask`)."""
    loss, F, gam = loss_neg_mean_coherence(sc, setups, heights, dr, ds, interp, offpoly, rl,
                                           cl, win, mask, round_blend)
    if mask is None:
        g = np.full(gam.shape, -1.0 / gam.size)
    else:
        g = np.where(mask, -1.0 / np.count_nonzero(mask), 0.0)
    gi, g1, g2 = coherence_backward(F["ifg"], F["pr"], F["ps"], win, g)
    gp, gh = pair_backward(F, gi, g1, g2, rl, cl)
    return loss, gp, gh


def interior_mask(F, rl, cl, win, margin=2):
    """Multilooked pixels whose coherence window only sees target pixels
    that are valid in both bursts with `margin` valid pixels around them, so
    that small perturbations cannot move a sampling position across a
    validity edge that the loss sees."""
    ok = (F["Tr"]["out"] != 0) & (F["Ts"]["out"] != 0)
    for _ in range(margin):
        e = ok.copy()
        e[1:, :] &= ok[:-1, :]
        e[:-1, :] &= ok[1:, :]
        e[:, 1:] &= ok[:, :-1]
        e[:, :-1] &= ok[:, 1:]
        ok = e
    nr, nc = ok.shape[0] // rl, ok.shape[1] // cl
    cell = ok[:nr * rl, :nc * cl].reshape(nr, rl, nc, cl).all(axis=(1, 3))
    h = win // 2
    if h:
        cnt = I.boxsum(cell.astype(np.int64), win)
        full = I.boxsum(np.ones_like(cell, dtype=np.int64), win)
        cell = cnt == full
        cell[:h, :] = cell[-h:, :] = False
        cell[:, :h] = cell[:, -h:] = False
    return cell
