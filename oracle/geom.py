"""FP64 NumPy restatement of the reference geometry path (test infrastructure).

Every function below follows the cited lines of the reference repository and
is vectorised over pixels; the arithmetic (order of operations, constants,
branch rules) is kept as in the source so that FP64 results agree with the
reference to rounding.

Parity pins (see tests/test_oracle_geom.py):
  llh2xyz / xyz2llh        closed-form KATs, round trip, compiled reference
                           (oracle/_ref) golden vectors
  orbithermite/intp_orbit  polynomial exactness (degree <= 7), knot values,
                           compiled reference golden vectors
  orbitrangetime           straight-orbit closed form, zero-Doppler invariant,
                           compiled reference golden vectors
  rah2ll / bounds          compiled reference golden vectors
  closest_sample / interp_record
                           branch cases (before / between / after records)
  deramp_setup             closed form for constant FM rate and zero DC
  deramp                   |z| preserved; deramp∘(inverse ramp) = identity
  reproject (bilinear)     bilinear exactness on bilinear fields; integer
                           offsets give samples; synthetic-scene coherence
  reproject (sinc)         band-limited exponentials, zero shift identity
                           (no reference counterpart)
"""

import numpy as np

RA = 6378137.0  # orbit.hpp:6-8
RE2 = 0.00669437999015  # orbit.hpp:9-11
SOL = 299792458.0  # geo2rdr.cu:26-28


# --------------------------------------------------------------------------
# A1 / A2  geodesy (orbit.cu:15-70)
# --------------------------------------------------------------------------
def llh2xyz(lat, lon, h, r_a=RA, r_e2=RE2):
    """orbit.cu:15-35.  lat/lon in degrees, h in metres; vectorised."""
    r_lat = np.asarray(lat, np.float64) * np.pi / 180
    r_lon = np.asarray(lon, np.float64) * np.pi / 180
    h = np.asarray(h, np.float64)
    coslat = np.cos(r_lat)
    coslon = np.cos(r_lon)
    sinlat = np.sin(r_lat)
    sinlon = np.sin(r_lon)
    re = r_a / np.sqrt(1.0 - r_e2 * sinlat * sinlat)
    x = (re + h) * coslat * coslon
    y = (re + h) * coslat * sinlon
    z = (re * (1 - r_e2) + h) * sinlat
    return np.stack(np.broadcast_arrays(x, y, z), axis=-1)


def xyz2llh(xyz, r_a=RA, r_e2=RE2):
    """orbit.cu:37-70 (one-step Bowring).  xyz (..., 3) -> (..., 3) lat, lon, h."""
    xyz = np.asarray(xyz, np.float64)
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    r_q2 = 1.0 / (1.0 - r_e2)
    r_q = np.sqrt(r_q2)
    r_q3 = r_q2 - 1.0
    r_b = r_a * np.sqrt(1 - r_e2)
    lon = np.arctan2(y, x)
    lon = np.where((y < 0) & (lon > 0), lon - np.pi, lon)
    lon = np.where((y > 0) & (lon < 0), lon + np.pi, lon)
    r_p = np.sqrt(x * x + y * y)
    r_tant = (z / r_p) * r_q
    r_theta = np.arctan(r_tant)
    r_tant = (z + r_q3 * r_b * np.sin(r_theta) ** 3) / (
        r_p - r_e2 * r_a * np.cos(r_theta) ** 3)
    lat = np.arctan(r_tant)
    r_re = r_a / np.sqrt(1.0 - r_e2 * np.sin(lat) * np.sin(lat))
    h = r_p / np.cos(lat) - r_re
    return np.stack([lat / np.pi * 180, lon / np.pi * 180, h], axis=-1)


# --------------------------------------------------------------------------
# A3 / A4  orbit interpolation (orbit.cu:170-245)
# --------------------------------------------------------------------------
def orbithermite(tt, xx, vv, t):
    """orbit.cu:170-222, vectorised over the leading axis.

    tt (..., 4), xx (..., 4, 3), vv (..., 4, 3), t (...) -> satx, satv (..., 3)
    """
    tt = np.asarray(tt, np.float64)
    t = np.asarray(t, np.float64)
    shape = t.shape
    li = np.ones(shape + (4,))
    a = np.zeros(shape + (4,))
    b = np.zeros(shape + (4,))
    a2 = np.zeros(shape + (4,))
    b2 = np.zeros(shape + (4,))
    for i in range(4):
        dl = np.zeros(shape)
        hdot = np.zeros(shape)
        for j in range(4):
            if i == j:
                continue
            dl = dl + 1.0 / (tt[..., i] - tt[..., j])
            li[..., i] = li[..., i] * (t - tt[..., j]) / (tt[..., i] - tt[..., j])
            p = 1.0 / (tt[..., i] - tt[..., j])
            for k in range(4):
                if k == i or k == j:
                    continue
                p = p * (t - tt[..., k]) / (tt[..., i] - tt[..., k])
            hdot = hdot + p
        l2 = li[..., i] * li[..., i]
        a[..., i] = (1 - 2 * (t - tt[..., i]) * dl) * l2
        b[..., i] = (t - tt[..., i]) * l2
        a2[..., i] = -2 * dl * l2 + (1 - 2 * (t - tt[..., i]) * dl) * li[..., i] * 2 * hdot
        b2[..., i] = l2 + (t - tt[..., i]) * li[..., i] * 2 * hdot
    satx = np.zeros(shape + (3,))
    satv = np.zeros(shape + (3,))
    for j in range(3):
        for i in range(4):
            satx[..., j] += a[..., i] * xx[..., i, j] + b[..., i] * vv[..., i, j]
            satv[..., j] += a2[..., i] * xx[..., i, j] + b2[..., i] * vv[..., i, j]
    return satx, satv


def stencil_index(torb, t):
    """orbit.cu:229-232: truncated index clamped to [1, n-3]."""
    n = len(torb)
    t = np.asarray(t, np.float64)
    iloc = np.trunc((t - torb[0]) / (torb[1] - torb[0])).astype(np.int64)
    iloc = np.maximum(iloc, 1)
    iloc = np.minimum(iloc, n - 3)
    return iloc


def intp_orbit(torb, xx, vv, t):
    """orbit.cu:224-245.  torb (n,), xx/vv (n,3), t scalar or array."""
    torb = np.asarray(torb, np.float64)
    xx = np.asarray(xx, np.float64)
    vv = np.asarray(vv, np.float64)
    t = np.asarray(t, np.float64)
    iloc = stencil_index(torb, t)
    idx = iloc[..., None] - 1 + np.arange(4)
    return orbithermite(torb[idx], xx[idx], vv[idx], t)


def orbitrangetime(torb, xx, vv, xyz, tline0, satx0, satv0, maxit=51, tol=5.0e-9):
    """orbit.cu:247-279, vectorised Newton with per-pixel convergence.

    Returns (tline, dr, niter) with dr = xyz - satx(tline_final).
    """
    xyz = np.asarray(xyz, np.float64)
    shape = xyz.shape[:-1]
    tline = np.full(shape, float(tline0))
    satx = np.broadcast_to(np.asarray(satx0, np.float64), shape + (3,)).copy()
    satv = np.broadcast_to(np.asarray(satv0, np.float64), shape + (3,)).copy()
    active = np.ones(shape, bool)
    niter = np.zeros(shape, np.int32)
    for _ in range(maxit):
        if not active.any():
            break
        ia = np.nonzero(active.ravel())[0]
        X = xyz.reshape(-1, 3)[ia]
        sx = satx.reshape(-1, 3)[ia]
        sv = satv.reshape(-1, 3)[ia]
        tl = tline.ravel()[ia]
        tprev = tl.copy()
        dr = X - sx
        fn = dr[:, 0] * sv[:, 0] + dr[:, 1] * sv[:, 1] + dr[:, 2] * sv[:, 2]
        fnprime = -(sv[:, 0] * sv[:, 0]) - sv[:, 1] * sv[:, 1] - sv[:, 2] * sv[:, 2]
        tl = tl - fn / fnprime
        nx, nv = intp_orbit(torb, xx, vv, tl)
        tline.ravel()[ia] = tl
        satx.reshape(-1, 3)[ia] = nx
        satv.reshape(-1, 3)[ia] = nv
        niter.ravel()[ia] += 1
        done = np.abs(tl - tprev) < tol
        active.ravel()[ia[done]] = False
    dr = xyz - satx
    return tline, dr, niter


# --------------------------------------------------------------------------
# A6  burst footprint (bounds.cu:237-317)
# --------------------------------------------------------------------------
def _reast(r_lat, r_a=RA, r_e2=RE2):
    return r_a / np.sqrt(1.0 - r_e2 * np.sin(r_lat) ** 2)


def _rnorth(r_lat, r_a=RA, r_e2=RE2):
    return r_a * (1.0 - r_e2) / (1.0 - r_e2 * np.sin(r_lat) ** 2) ** 1.5


def _rdir(r_lat, hdg, r_a=RA, r_e2=RE2):
    re = _reast(r_lat, r_a, r_e2)
    rn = _rnorth(r_lat, r_a, r_e2)
    return (re * rn) / (re * np.cos(hdg) ** 2 + rn * np.sin(hdg) ** 2)


def ellipsoidal_hdg(ll1, ll2, r_a=RA, r_e2=RE2):
    """bounds.cu:112-199 (Sodano inverse geodesic heading)."""
    r_e = np.sqrt(r_e2)
    r_sqrtome2 = np.sqrt(1.0 - r_e2)
    r_f = 1.0 - r_sqrtome2
    r_ep = r_e * r_f / (r_e2 - r_f)
    r_n = r_f / r_e2
    r_k1 = (16.0 * r_e2 * r_n * r_n + r_ep * r_ep) / (r_ep * r_ep)
    r_k2 = (16.0 * r_e2 * r_n * r_n) / (16.0 * r_e2 * r_n * r_n + r_ep * r_ep)
    r_k3 = (16.0 * r_e2 * r_n * r_n) / (r_ep * r_ep)
    r_k4 = (16.0 * r_n - r_ep * r_ep) / (16.0 * r_e2 * r_n * r_n + r_ep * r_ep)
    r_k5 = 16.0 / (r_e2 * (16.0 * r_e2 * r_n * r_n + r_ep * r_ep))
    rad = np.pi / 180.0
    r_tanlati = np.tan(ll1[0] * rad)
    r_tanlatf = np.tan(ll2[0] * rad)
    r_l = abs(ll2[1] * rad - ll1[1] * rad)
    r_lsign = ll2[1] * rad - ll1[1] * rad
    r_sinlon = np.sin(r_l)
    r_coslon = np.cos(r_l)
    r_tanbetai = r_sqrtome2 * r_tanlati
    r_tanbetaf = r_sqrtome2 * r_tanlatf
    r_cosbetai = 1.0 / np.sqrt(1.0 + r_tanbetai * r_tanbetai)
    r_cosbetaf = 1.0 / np.sqrt(1.0 + r_tanbetaf * r_tanbetaf)
    r_sinbetai = r_tanbetai * r_cosbetai
    r_sinbetaf = r_tanbetaf * r_cosbetaf
    r_ac = r_sinbetai * r_sinbetaf
    r_bc = r_cosbetai * r_cosbetaf
    r_cosphi = r_ac + r_bc * r_coslon
    r_sinphi = np.sign(r_sinlon) * np.sqrt(1.0 - min(r_cosphi * r_cosphi, 1.0))
    r_phi = abs(np.arctan2(r_sinphi, r_cosphi))
    if r_a * abs(r_phi) <= 1e-6:
        return 0.0
    r_ca = (r_bc * r_sinlon) / r_sinphi
    r_cb = r_ca * r_ca
    r_cc = (r_cosphi * (1.0 - r_cb)) / r_k1
    r_cd = (-2.0 * r_ac) / r_k1
    r_ce = -r_ac * r_k2
    r_cf = r_k3 * r_cc
    r_cg = r_phi * r_phi / r_sinphi
    r_x = ((r_phi * (r_k4 + r_cb) + r_sinphi * (r_cc + r_cd) + r_cg * (r_cf + r_ce))
           * r_ca) / r_k5
    r_lambda = r_l + r_x
    r_sinlam = np.sin(r_lambda)
    r_coslam = np.cos(r_lambda)
    r_t2 = (r_tanbetaf * r_cosbetai - r_coslam * r_sinbetai)
    r_sinlon = r_sinlam * np.sign(r_lsign)
    return np.arctan2(r_sinlon, r_t2)


def tcnbasis(pos, vel, look_dir="RIGHT"):
    """bounds.cu:66-92."""
    llh = xyz2llh(pos)
    r_lat = llh[0] * np.pi / 180.0
    r_lon = llh[1] * np.pi / 180.0
    r_n = np.array([-np.cos(r_lat) * np.cos(r_lon), -np.cos(r_lat) * np.sin(r_lon),
                    -np.sin(r_lat)])
    if look_dir == "RIGHT":
        r_temp = np.cross(r_n, vel)
    else:
        r_temp = np.cross(vel, r_n)
    r_c = r_temp / np.linalg.norm(r_temp)
    r_temp = np.cross(r_c, r_n)
    r_t = r_temp / np.linalg.norm(r_temp)
    return r_t, r_c, r_n


def rah2ll(rah, torb, xx, vv, start_time, end_time, look_dir="RIGHT"):
    """bounds.cu:237-286.  rah (8,3) rows of (range, time, height)."""
    xs, _ = intp_orbit(torb, xx, vv, start_time)
    xe, _ = intp_orbit(torb, xx, vv, end_time)
    xm, _ = intp_orbit(torb, xx, vv, (start_time + end_time) / 2.0)
    llhs = xyz2llh(xs)
    llhe = xyz2llh(xe)
    hdg = ellipsoidal_hdg(llhs, llhe)
    llhm = xyz2llh(xm)
    rcurv = _rdir(llhm[0] * np.pi / 180.0, hdg)
    lat = np.zeros(8)
    lon = np.zeros(8)
    for i in range(8):
        r, t, h = rah[i]
        xyz, vel = intp_orbit(torb, xx, vv, t)
        llh = xyz2llh(xyz)
        vhat = vel / np.sqrt(np.dot(vel, vel))
        that, chat, nhat = tcnbasis(xyz, vel, look_dir)
        aa = rcurv + llh[2]
        bb = rcurv + h
        costheta = 0.5 * ((aa / r + r / aa - bb / aa * bb / r))
        sintheta = np.sqrt(1. - costheta * costheta)
        gamm = costheta * r
        alpha = (0.0 * r - gamm * np.dot(nhat, vhat)) / np.dot(vhat, that)
        beta = np.sqrt((r * sintheta) ** 2 - alpha * alpha)
        xyztar = xyz + gamm * nhat + alpha * that + beta * chat
        llt = xyz2llh(xyztar)
        lat[i] = llt[0]
        lon[i] = llt[1]
    return lat, lon


def bounds(start_time, end_time, start_rng, end_rng, hmin, hmax, torb, xx, vv,
           look_dir="RIGHT"):
    """bounds.cu:288-317 -> (latmin, latmax, lonmin, lonmax)."""
    rah = np.zeros((8, 3))
    for i in range(8):
        rah[i, 0] = start_rng if i < 4 else end_rng
        rah[i, 1] = start_time if i % 4 < 2 else end_time
        rah[i, 2] = hmin if i % 2 == 0 else hmax
    lat, lon = rah2ll(rah, torb, xx, vv, start_time, end_time, look_dir)
    return lat.min(), lat.max(), lon.min(), lon.max()


# --------------------------------------------------------------------------
# A7  deramp coefficients (geo2rdr.cu:48-59, 465-549)
# --------------------------------------------------------------------------
def closest_sample(t, t0):
    """geo2rdr.cu:48-59: first index of the minimum |t[i]-t0|."""
    idx = 0
    mindist = 1e10
    for i in range(len(t)):
        dist = abs(t[i] - t0)
        if mindist > dist:
            mindist = dist
            idx = i
    return idx


def interp_record(rec, tc):
    """geo2rdr.cu:485-527.  rec (n,5) rows (t, t0, c0, c1, c2) -> (t0, c0, c1, c2)."""
    rec = np.asarray(rec, np.float64)
    n = rec.shape[0]
    idx = closest_sample(rec[:, 0], tc)
    if rec[idx, 0] < tc and idx < n - 1:
        frac = (tc - rec[idx, 0]) / (rec[idx + 1, 0] - rec[idx, 0])
        v = rec[idx, 1:] + frac * (rec[idx + 1, 1:] - rec[idx, 1:])
    elif rec[idx, 0] > tc and idx > 0:
        frac = (tc - rec[idx - 1, 0]) / (rec[idx, 0] - rec[idx - 1, 0])
        v = rec[idx - 1, 1:] + frac * (rec[idx, 1:] - rec[idx - 1, 1:])
    else:
        v = rec[idx, 1:].copy()
    return v  # (t0, c0, c1, c2)


def deramp_setup(orbit, burst):
    """geo2rdr.cu:477-542.  Returns a dict with ks, fm/dc coefficients, kt,
    etaref, eta (all FP64) and etac0."""
    torb, xx, vv = orbit
    L = burst["lines"]
    N = burst["samples"]
    ati = burst["azimuth_time_interval"]
    tc = burst["t_start"] + L * ati / 2
    _, vpoint = intp_orbit(torb, xx, vv, tc)
    vs = np.sqrt(vpoint[0] ** 2 + vpoint[1] ** 2 + vpoint[2] ** 2)
    ks = 2.0 * vs * burst["radar_frequency"] / SOL * burst["azimuth_steering_rate"] * np.pi / 180.0
    fm = interp_record(burst["fm"], tc)
    dc = interp_record(burst["dc"], tc)
    eta = -L * ati / 2. + np.arange(L) * ati
    trange = burst["slant_range_time"] + np.arange(N) / burst["range_sampling_rate"]
    dt = trange - fm[0]
    ka = fm[1] + dt * (fm[2] + fm[3] * dt)
    dt = trange - dc[0]
    fnc = dc[1] + dt * (dc[2] + dc[3] * dt)
    kt = ka * ks / (ka - ks)
    etac = -fnc / ka
    etaref = etac - etac[0]
    return dict(ks=ks, fm=fm, dc=dc, kt=kt, eta=eta, etaref=etaref, etac0=etac[0])


# --------------------------------------------------------------------------
# A8  deramp (geo2rdr.cu:138-158)
# --------------------------------------------------------------------------
def deramp(burst_data, setup):
    """z[i,j] * exp(i*phi), phi = -pi*kt[j]*(eta[i]-etaref[j])^2; complex128 out."""
    eta = setup["eta"][:, None]
    etadiff = eta - setup["etaref"][None, :]
    phase = -np.pi * setup["kt"][None, :] * etadiff * etadiff
    return burst_data.astype(np.complex128) * np.exp(1j * phase)


def reramp_phase_of_data(setup, rows, cols):
    """The ramp a TOPS SLC carries at integer samples (inverse of deramp)."""
    etadiff = setup["eta"][rows][:, None] - setup["etaref"][cols][None, :]
    return np.pi * setup["kt"][cols][None, :] * etadiff * etadiff


# --------------------------------------------------------------------------
# Sinc interpolator (no reference counterpart; the reference is bilinear only)
# --------------------------------------------------------------------------
SINC_HALF = 4  # taps at floor(x)-3 .. floor(x)+4


def sinc_taps(frac):
    """Hann-windowed sinc, 8 taps, normalised to unit DC gain.

    frac (...) in [0, 1): weight for sample floor(x)+k, k = -3..4, is
    h(frac - k) / sum, h(x) = sinc(x) * cos^2(pi x / 8).
    """
    frac = np.asarray(frac, np.float64)
    k = np.arange(-SINC_HALF + 1, SINC_HALF + 1)
    x = frac[..., None] - k
    h = np.sinc(x) * np.cos(np.pi * x / 8.0) ** 2
    return h / h.sum(axis=-1, keepdims=True)


# --------------------------------------------------------------------------
# A9  reproject (geo2rdr.cu:170-248), generalised target grid
# --------------------------------------------------------------------------
def target_heights(grid, dem):
    """Bilinear DEM height at every target pixel (FP64).

    grid: dict(nrow, ncol, lat0, lon0, dlat, dlon); dem: dict(data (float32
    2-D), lat0, lon0, dlat, dlon).  Indices are clamped to the DEM.  When the
    DEM and target grids coincide this returns the DEM samples, which is the
    reference's h = dem[i] (geo2rdr.cu:199).
    """
    r = np.arange(grid["nrow"], dtype=np.float64)
    c = np.arange(grid["ncol"], dtype=np.float64)
    lat = grid["lat0"] + r * grid["dlat"]
    lon = grid["lon0"] + c * grid["dlon"]
    d = np.asarray(dem["data"])
    ny, nx = d.shape
    fy = (lat - dem["lat0"]) / dem["dlat"]
    fx = (lon - dem["lon0"]) / dem["dlon"]
    fy = np.clip(fy, 0.0, ny - 1.0)
    fx = np.clip(fx, 0.0, nx - 1.0)
    iy = np.minimum(np.floor(fy).astype(np.int64), ny - 2)
    ix = np.minimum(np.floor(fx).astype(np.int64), nx - 2)
    wy = (fy - iy)[:, None]
    wx = (fx - ix)[None, :]
    d = d.astype(np.float64)
    h00 = d[iy][:, ix]
    h01 = d[iy][:, ix + 1]
    h10 = d[iy + 1][:, ix]
    h11 = d[iy + 1][:, ix + 1]
    return (h00 * (1 - wx) + h01 * wx) * (1 - wy) + (h10 * (1 - wx) + h11 * wx) * wy


def geo2rdr_grid(orbit, burst, grid, heights):
    """Per target pixel: (tline, R, rgoff, azoff) from the reference Newton.

    geo2rdr.cu:195-207 with tmid/xmid/vmid from geo2rdr.cu:588-596.
    """
    torb, xx, vv = orbit
    L = burst["lines"]
    dtaz = 1.0 / burst["prf"]
    tstart = burst["t_start"]
    tend = tstart + (L - 1) * dtaz
    tmid = (tstart + tend) * 0.5
    xmid, vmid = intp_orbit(torb, xx, vv, tmid)
    r = np.arange(grid["nrow"], dtype=np.float64)
    c = np.arange(grid["ncol"], dtype=np.float64)
    lat = (grid["lat0"] + r * grid["dlat"])[:, None]
    lon = (grid["lon0"] + c * grid["dlon"])[None, :]
    X = llh2xyz(lat, lon, heights)
    tline, dr, niter = orbitrangetime(torb, xx, vv, X, tmid, xmid, vmid)
    rngpix = np.sqrt(dr[..., 0] ** 2 + dr[..., 1] ** 2 + dr[..., 2] ** 2)
    rngstart = burst["slant_range_time"] * SOL / 2.0
    dmrg = SOL / 2.0 / burst["range_sampling_rate"]
    rgoff = (rngpix - rngstart) / dmrg
    azoff = (tline - tstart) / dtaz
    return dict(tline=tline, rng=rngpix, rgoff=rgoff, azoff=azoff, niter=niter)


def offset_poly_shift(offpoly, rgoff, azoff, burst):
    """Offset polynomial (no reference counterpart): shifts in pixels.

    d_rg = p0 + p1*(rgoff/samples) + p2*(azoff/lines)
    d_az = p3 + p4*(rgoff/samples) + p5*(azoff/lines)
    """
    if offpoly is None:
        return 0.0, 0.0
    p = np.asarray(offpoly, np.float64)
    x = rgoff / burst["samples"]
    y = azoff / burst["lines"]
    return p[0] + p[1] * x + p[2] * y, p[3] + p[4] * x + p[5] * y


def reproject(orbit, burst, setup, grid, heights, deramped, interp="bilinear",
              offpoly=None, geo=None, round_blend=True):
    """geo2rdr.cu:194-247 on a target grid; complex128 out (zeros where invalid).

    deramped: complex burst after deramp (complex64 as the kernel reads it).
    interp: "bilinear" (reference) or "sinc" (8-tap windowed sinc).
    offpoly: optional 6 offset-polynomial coefficients (sampling position and
    ramp only; the flattening phase keeps the geometric range).
    round_blend=False keeps the bilinear range blends in FP64 (a smooth
    function of the position, for finite-difference checks of the adjoint).
    """
    if geo is None:
        geo = geo2rdr_grid(orbit, burst, grid, heights)
    rngpix = geo["rng"]
    rgoff0 = geo["rgoff"]
    azoff0 = geo["azoff"]
    drg, daz = offset_poly_shift(offpoly, rgoff0, azoff0, burst)
    rgoff = rgoff0 + drg
    azoff = azoff0 + daz
    L = burst["lines"]
    dtaz = 1.0 / burst["prf"]
    dmrg = SOL / 2.0 / burst["range_sampling_rate"]
    fvs, lvs = burst["first_valid_sample"], burst["last_valid_sample"]
    flv, llv = burst["first_valid_line"], burst["last_valid_line"]
    if interp == "bilinear":
        valid = (rgoff > fvs) & (rgoff < lvs - 1) & (azoff >= flv + 5) & (azoff < llv - 5)
    else:
        valid = (rgoff >= fvs + 3) & (rgoff < lvs - 4) & (azoff >= flv + 5) & (azoff < llv - 5)
    out = np.zeros(rgoff.shape, np.complex128)
    if not valid.any():
        return out, geo
    rp = rngpix[valid] + (drg[valid] if np.ndim(drg) else drg) * dmrg
    fm = setup["fm"]
    dc = setup["dc"]
    dt = rp * 2 / SOL - fm[0]
    ka = fm[1] + dt * (fm[2] + fm[3] * dt)
    dt = rp * 2 / SOL - dc[0]
    fnc = dc[1] + dt * (dc[2] + dc[3] * dt)
    kt = ka * setup["ks"] / (ka - setup["ks"])
    etac = -fnc / ka
    etaref = etac - setup["etac0"]
    eta = -L * dtaz / 2. + geo["tline"][valid] - burst["t_start"]
    if np.ndim(daz):
        eta = eta + daz[valid] * dtaz
    elif daz:
        eta = eta + daz * dtaz
    etadiff = eta - etaref
    reramp = -np.pi * kt * etadiff * etadiff
    rg = rgoff[valid]
    az = azoff[valid]
    N = burst["samples"]
    z = deramped.reshape(-1)
    if interp == "bilinear":
        intr = rg.astype(np.int64)
        fracr = rg - intr
        inta = az.astype(np.int64)
        fraca = az - inta
        idx1 = inta * N + intr
        b1, b2, b3, b4 = z[idx1], z[idx1 + 1], z[idx1 + N], z[idx1 + N + 1]
        # geo2rdr.cu:232-235: the range blend is stored as float2
        c1 = b1 * (1 - fracr) + b2 * fracr
        c2 = b3 * (1 - fracr) + b4 * fracr
        if round_blend:
            c1 = _to_cf32(c1.real, c1.imag)
            c2 = _to_cf32(c2.real, c2.imag)
        res = c1.astype(np.complex128) * (1 - fraca) + c2.astype(np.complex128) * fraca
    else:
        intr = np.floor(rg).astype(np.int64)
        inta = np.floor(az).astype(np.int64)
        wr = sinc_taps(rg - intr)
        wa = sinc_taps(az - inta)
        res = np.zeros(rg.shape, np.complex128)
        for ka_ in range(8):
            row = (inta + ka_ - (SINC_HALF - 1)) * N
            acc = np.zeros(rg.shape, np.complex128)
            for kr in range(8):
                acc += wr[:, kr] * z[row + intr + kr - (SINC_HALF - 1)]
            res += wa[:, ka_] * acc
    phase = 4.0 * np.pi / burst["wavelength"] * rngpix[valid] - reramp
    out[valid] = res * np.exp(1j * phase)
    return out, geo


def _to_cf32(re, im):
    return (np.asarray(re, np.float32) + 1j * np.asarray(im, np.float32)).astype(np.complex64)


def nonzero_rows(img):
    """geo2rdr.cu:86-128, 273-299: (first, last_exclusive) or (0, 0)."""
    nz = np.nonzero(np.any(img != 0, axis=1))[0]
    if len(nz) == 0:
        return 0, 0
    return int(nz[0]), int(nz[-1]) + 1
