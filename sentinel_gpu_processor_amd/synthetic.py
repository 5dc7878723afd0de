"""Synthetic Sentinel-1 TOPS burst pairs for tests and benchmarks.

The scene is physically consistent with the processing chain:

- straight-line ECEF orbits at 693 km (reference and secondary separated by
  a baseline), right-looking, 1 s state vectors;
- a DEM (Gaussian hill or fractal) on a lat/lon grid covering the burst
  footprint, and a target lat/lon grid with the burst's pixel count;
- a band-limited circular-Gaussian reflectivity field z on the reference
  radar grid; the reference SLC is z times the TOPS azimuth ramp
  exp(+i pi kt (eta - etaref)^2) (the inverse of the deramp of
  geo2rdr.cu:138-158);
- the secondary SLC at radar pixel q' holds, for the ground point P' seen
  there, gamma0 z(ref coords of P') + sqrt(1-gamma0^2) n times
  exp(-i 4 pi (R_sec(P') - R_ref(P'))/lambda) exp(i phi(P')) times its own
  ramp, so that after geocoding and crossmul the interferogram phase is
  phi (topographic fringes of the DEM at a chosen ambiguity height plus a
  band-limited phase noise) and the coherence is about gamma0.

``fidelity="exact"`` evaluates the secondary mapping per pixel (FP64) and
resamples z with a 12-tap windowed sinc; ``fidelity="fast"`` evaluates the
mapping on a coarse grid and resamples on the GPU with torch (used for the
large benchmark bursts, where only the data volume and geometry matter);
``fidelity="none"`` builds the orbits, bursts, grid and DEM only.
"""

import math

import numpy as np

SOL = 299792458.0
RA = 6378137.0
RE2 = 0.00669437999015

WAVELENGTH = 0.0554658
RANGE_PIXEL = 2.3296
AZ_INTERVAL = 2.0556e-3
FM_RATE = -2300.0
STEERING_RATE = 1.59
ALTITUDE = 693.0e3
SPEED = 7500.0


# ---------------------------------------------------------------- geodesy
def llh2xyz(lat, lon, h):
    lat = np.radians(lat)
    lon = np.radians(lon)
    sl, cl = np.sin(lat), np.cos(lat)
    re = RA / np.sqrt(1.0 - RE2 * sl * sl)
    return np.stack(np.broadcast_arrays((re + h) * cl * np.cos(lon), (re + h) * cl * np.sin(lon),
                                        (re * (1 - RE2) + h) * sl), axis=-1)


def xyz2llh(xyz):
    """Iterative (exact to < 1e-9 m) geodetic conversion."""
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    lon = np.arctan2(y, x)
    p = np.hypot(x, y)
    lat = np.arctan2(z, p * (1 - RE2))
    for _ in range(6):
        sl = np.sin(lat)
        re = RA / np.sqrt(1 - RE2 * sl * sl)
        h = p / np.cos(lat) - re
        lat = np.arctan2(z, p * (1 - RE2 * re / (re + h)))
    sl = np.sin(lat)
    re = RA / np.sqrt(1 - RE2 * sl * sl)
    h = p / np.cos(lat) - re
    return np.degrees(lat), np.degrees(lon), h


def _frame(lat, lon):
    la, lo = np.radians(lat), np.radians(lon)
    up = np.array([np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)])
    east = np.array([-np.sin(lo), np.cos(lo), 0.0])
    north = np.cross(up, east)
    return up, east, north


# ----------------------------------------------------------------- orbit
class LineOrbit:
    """x(t) = s0 + v (t - tc)."""

    def __init__(self, s0, v, tc):
        self.s0 = np.asarray(s0, np.float64)
        self.v = np.asarray(v, np.float64)
        self.tc = float(tc)

    def pos(self, t):
        t = np.asarray(t, np.float64)
        return self.s0 + self.v * (t - self.tc)[..., None]

    def state_vectors(self, t0, n):
        t = t0 + np.arange(n, dtype=np.float64)
        return t, self.pos(t), np.repeat(self.v[None, :], n, axis=0)

    def zero_doppler(self, X):
        """Closed-form zero-Doppler time and range for a straight line."""
        vv = self.v @ self.v
        t = self.tc + ((X - self.s0) @ self.v) / vv
        d = X - self.pos(t)
        return t, np.sqrt(np.sum(d * d, axis=-1))


# ------------------------------------------------------------------- DEM
def gaussian_hill(ny, nx, hmax=1500.0, sigma_frac=0.22):
    y = (np.arange(ny) - (ny - 1) / 2) / ny
    x = (np.arange(nx) - (nx - 1) / 2) / nx
    r2 = y[:, None] ** 2 + x[None, :] ** 2
    return (hmax * np.exp(-r2 / (2 * sigma_frac ** 2))).astype(np.float32)


def fractal_dem(ny, nx, rng, hurst=0.8, hmax=2500.0):
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.fftfreq(nx)[None, :]
    k = np.sqrt(ky * ky + kx * kx)
    k[0, 0] = 1.0
    amp = k ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    ph = rng.uniform(0, 2 * np.pi, (ny, nx))
    f = np.real(np.fft.ifft2(amp * np.exp(1j * ph)))
    f = (f - f.min()) / (f.max() - f.min())
    return (f * hmax).astype(np.float32)


def dem_height(dem, lat, lon):
    d = dem["data"].astype(np.float64)
    ny, nx = d.shape
    fy = np.clip((lat - dem["lat0"]) / dem["dlat"], 0, ny - 1.0)
    fx = np.clip((lon - dem["lon0"]) / dem["dlon"], 0, nx - 1.0)
    iy = np.minimum(np.floor(fy).astype(np.int64), ny - 2)
    ix = np.minimum(np.floor(fx).astype(np.int64), nx - 2)
    wy = fy - iy
    wx = fx - ix
    return ((d[iy, ix] * (1 - wx) + d[iy, ix + 1] * wx) * (1 - wy)
            + (d[iy + 1, ix] * (1 - wx) + d[iy + 1, ix + 1] * wx) * wy)


# ------------------------------------------------------------ radar geometry
def rdr2geo(orbit, t, R, dem, iters=8):
    """Ground point on the DEM surface at azimuth time t and range R
    (right-looking), vectorised."""
    t = np.asarray(t, np.float64)
    R = np.asarray(R, np.float64)
    s = orbit.pos(t)
    vhat = orbit.v / np.linalg.norm(orbit.v)
    up = s / np.linalg.norm(s, axis=-1, keepdims=True)
    up = up - (up @ vhat)[..., None] * vhat
    up /= np.linalg.norm(up, axis=-1, keepdims=True)
    right = np.cross(vhat, up)
    rs = np.linalg.norm(s, axis=-1)

    def point(th):
        return s + R[..., None] * (-np.cos(th)[..., None] * up + np.sin(th)[..., None] * right)

    def resid(th):
        lat, lon, h = xyz2llh(point(th))
        return h - dem_height(dem, lat, lon)

    # bracket: heights from -500 m (far look angle) to DEM max + 500 m (near)
    dmax = float(np.max(dem["data"])) + 500.0

    def theta_at(hh):
        re = 6371000.0 + hh
        return np.arccos(np.clip((rs * rs + R * R - re * re) / (2 * rs * R), -1, 1))

    lo = theta_at(dmax + 20000.0)  # smaller look angle: point above the surface
    hi = theta_at(-20000.0)        # larger look angle: point below the surface
    for _ in range(44):
        mid = 0.5 * (lo + hi)
        above = resid(mid) > 0
        lo = np.where(above, mid, lo)
        hi = np.where(above, hi, mid)
    theta = 0.5 * (lo + hi)
    return point(theta)


def record_at(rec, tc):
    """Coefficients (t0, c0, c1, c2) of the FM-rate / Doppler records
    rec[k] = (t, t0, c0, c1, c2) at time tc: linear interpolation between the
    two records around tc, the first (last) record before (after) them.
    Records are sorted by time."""
    rec = np.asarray(rec, np.float64)
    return np.array([np.interp(tc, rec[:, 0], rec[:, 1 + k]) for k in range(4)])


def tops_rates(orbit, burst, tau):
    """TOPS rates of a straight-line-orbit burst at two-way range times tau:
    (kt, etaref).  ka and the Doppler centroid are the records' polynomials
    at the burst centre time; ks = 2 |v| f0 / c * steering rate;
    kt = ka ks / (ka - ks); the Doppler-centroid time etac = -fdc / ka is
    referred to its value at the first range sample."""
    L = burst["lines"]
    tc = burst["t_start"] + L * burst["azimuth_time_interval"] / 2
    vs = float(np.linalg.norm(orbit.v))
    ks = 2.0 * vs * burst["radar_frequency"] / SOL * burst["azimuth_steering_rate"] * np.pi / 180.0
    fm = record_at(burst["fm"], tc)
    dc = record_at(burst["dc"], tc)

    def etac_ka(t):
        ka = np.polyval([fm[3], fm[2], fm[1]], t - fm[0])
        fdc = np.polyval([dc[3], dc[2], dc[1]], t - dc[0])
        return -fdc / ka, ka

    etac, ka = etac_ka(np.asarray(tau, np.float64))
    etac0, _ = etac_ka(burst["slant_range_time"])
    return ka * ks / (ka - ks), etac - etac0


def tops_records(t_center, tau_c, kind="constant", shift=None):
    """FM-rate and Doppler-centroid records (t, t0, c0, c1, c2), three each,
    around a burst centred at t_center with mid-swath two-way range time
    tau_c.

    "constant": ka = -2300 Hz/s, fdc = 0, the records at t_center + 2.75 k.
    "varying": ka = c0 + c1 dt + c2 dt^2 with c0 ~ -2300 Hz/s,
    c1 ~ 4.5e5 Hz/s^2, c2 ~ -7.5e7 Hz/s^3 and fdc with c0 ~ 25 Hz,
    c1 ~ -4e4 Hz/s, c2 ~ 8e6 Hz/s^2 (dt = tau - t0), every coefficient and
    t0 different per record; the FM records are 0.9 s and the Doppler
    records -1.3 s off t_center + 2.75 k, so the burst centre falls between
    two records of each.  shift adds a time offset to every record (to put
    the centre before the first or after the last record)."""
    ks = (-1, 0, 1)
    if kind == "constant":
        fm = np.array([[t_center + k * 2.75, tau_c, FM_RATE, 0.0, 0.0] for k in ks])
        dc = np.array([[t_center + k * 2.75, tau_c, 0.0, 0.0, 0.0] for k in ks])
    elif kind == "varying":
        fm = np.array([[t_center + 0.9 + 2.75 * k, tau_c + 2e-7 * k, FM_RATE + 15.0 * k,
                        4.5e5 + 2.0e4 * k, -7.5e7 + 5.0e6 * k] for k in ks])
        dc = np.array([[t_center - 1.3 + 2.75 * k, tau_c - 1e-7 * k, 25.0 - 4.0 * k,
                        -4.0e4 + 3.0e3 * k, 8.0e6 - 1.0e6 * k] for k in ks])
    else:
        raise ValueError(f"tops kind {kind!r}")
    if shift:
        fm[:, 0] += shift
        dc[:, 0] += shift
    return fm, dc


def deramp_coefficients(orbit, burst):
    """kt[N], etaref[N], eta[L] of the burst (the quantities of
    geo2rdr.cu:477-542) for a straight-line orbit: tops_rates at the range
    samples, eta = -L ati / 2 + i ati."""
    L, N = burst["lines"], burst["samples"]
    ati = burst["azimuth_time_interval"]
    eta = -L * ati / 2. + np.arange(L) * ati
    tr = burst["slant_range_time"] + np.arange(N) / burst["range_sampling_rate"]
    kt, etaref = tops_rates(orbit, burst, tr)
    return kt, etaref, eta


def ramp(orbit, burst, rows=None, cols=None):
    kt, etaref, eta = deramp_coefficients(orbit, burst)
    if rows is not None:
        eta = eta[rows]
    if cols is not None:
        kt, etaref = kt[cols], etaref[cols]
    d = eta[:, None] - etaref[None, :]
    return np.pi * kt[None, :] * d * d


# --------------------------------------------------------------- speckle
def bandlimited_noise(rng, shape, bw_az=0.7, bw_rg=0.85):
    w = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)
    fa = np.abs(np.fft.fftfreq(shape[0]))[:, None] * 2
    fr = np.abs(np.fft.fftfreq(shape[1]))[None, :] * 2
    win = ((fa < bw_az) * 0.5 * (1 + np.cos(np.pi * np.minimum(fa / bw_az, 1)))
           * (fr < bw_rg) * 0.5 * (1 + np.cos(np.pi * np.minimum(fr / bw_rg, 1))))
    z = np.fft.ifft2(np.fft.fft2(w) * win)
    z /= np.sqrt(np.mean(np.abs(z) ** 2))
    return z


def _hann_sinc_taps(frac, half):
    k = np.arange(-half + 1, half + 1)
    x = frac[..., None] - k
    h = np.sinc(x) * np.cos(np.pi * x / (2 * half)) ** 2
    return h / h.sum(axis=-1, keepdims=True)


def sinc_resample(z, a, r, half=6, chunk=1 << 18):
    """z sampled at fractional (a, r) with a (2*half)-tap windowed sinc."""
    out = np.zeros(a.shape, np.complex128)
    af, rf = a.ravel(), r.ravel()
    of = out.ravel()
    nr = z.shape[1]
    zf = z.ravel()
    for s in range(0, af.size, chunk):
        aa = af[s:s + chunk]
        rr = rf[s:s + chunk]
        ia = np.floor(aa).astype(np.int64)
        ir = np.floor(rr).astype(np.int64)
        wa = _hann_sinc_taps(aa - ia, half)
        wr = _hann_sinc_taps(rr - ir, half)
        acc = np.zeros(aa.shape, np.complex128)
        for ka in range(2 * half):
            row = (ia + ka - half + 1) * nr
            accr = np.zeros(aa.shape, np.complex128)
            for kr in range(2 * half):
                accr += wr[:, kr] * zf[row + ir + kr - half + 1]
            acc += wa[:, ka] * accr
        of[s:s + chunk] = acc
    return out


# ----------------------------------------------------------------- scene
CONFIGS = {
    # name: lines, samples, dem shape, dem kind
    "c1": dict(lines=512, samples=2048, dem=(256, 256), dem_kind="hill", hmax=1500.0,
               offset=None),
    "c2": dict(lines=1500, samples=25000, dem=(2000, 3000), dem_kind="fractal", hmax=2500.0,
               offset=(0.3, 0.05)),
    # 27 bursts (3 subswaths x 9), sizes drawn from the seed
    "c3": dict(n_sub=3, n_burst=9, lines=(1450, 1530), samples=(20000, 25500), dem=(2000, 3000),
               dem_kind="fractal", hmax=2500.0, offset=(0.3, 0.05)),
}


def iw_sizes(seed=0, n_sub=3, n_burst=9, lines=(1450, 1530), samples=(20000, 25500)):
    """Burst sizes of a synthetic IW pair: n_sub x n_burst sizes drawn
    uniformly (integers, bounds inclusive) from the seed."""
    rng = np.random.default_rng(seed)
    n = n_sub * n_burst
    L = rng.integers(lines[0], lines[1] + 1, n)
    N = rng.integers(samples[0], samples[1] + 1, n)
    return [(int(a), int(b)) for a, b in zip(L, N)]


def make_iw(seed=0, n_sub=3, n_burst=9, lines=(1450, 1530), samples=(20000, 25500),
            dem_shape=(2000, 3000), dem_kind="fractal", hmax=2500.0, misreg=(0.3, 0.05),
            fidelity="fast", device=None, select=None):
    """Synthetic IW pair: n_sub subswaths x n_burst bursts of uneven size.
    Burst k of subswath s is centred 0.17 deg further along track than burst
    k-1 (about one burst length) and subswath s looks 50 km further in
    range than s-1; each burst pair has its own DEM window and seed.
    select: indices (s * n_burst + k) of the bursts to build (default all);
    each is the same as in the whole IW pair."""
    sizes = iw_sizes(seed, n_sub, n_burst, lines, samples)
    out = []
    for s in range(n_sub):
        for k in range(n_burst):
            if select is not None and s * n_burst + k not in select:
                continue
            L, N = sizes[s * n_burst + k]
            sc = make_scene(lines=L, samples=N, dem_shape=dem_shape, dem_kind=dem_kind,
                            hmax=hmax, seed=seed * 1000 + s * n_burst + k, misreg=misreg,
                            misreg_ramp=(0.2, 0.1) if misreg else (0.0, 0.0),
                            fidelity=fidelity, device=device,
                            lat_c=34.0 + 0.17 * (k - n_burst // 2),
                            look_range=800.0e3 + 50.0e3 * s,
                            t_center=43200.0 + 2.75 * k)
            sc["subswath"], sc["burst"] = s + 1, k + 1
            out.append(sc)
    return out


GROUND_RANGE_PIXEL = 4.0        # m, ground spacing of a range sample (approx.)
GROUND_AZIMUTH_PIXEL = 13.9     # m, ground spacing of an azimuth line (approx.)


def aps_field(lines, samples, seed, corr=2000.0, std=1.0, coarse=16):
    """Atmospheric phase screen (radians) in radar geometry: a Gaussian random
    field with Gaussian covariance exp(-r^2 / corr^2) (e-folding distance
    `corr` metres on the ground) and standard deviation `std`, built on a grid
    `coarse` pixels apart and upsampled bilinearly (numpy, float32)."""
    rng = np.random.default_rng(seed)
    cy = lines // coarse + 3
    cx = samples // coarse + 3
    w = rng.standard_normal((cy, cx))
    # white noise smoothed by a Gaussian of std corr / 2 has covariance
    # exp(-r^2 / corr^2)
    sy = corr / 2 / (GROUND_AZIMUTH_PIXEL * coarse)
    sx = corr / 2 / (GROUND_RANGE_PIXEL * coarse)
    ky = np.fft.fftfreq(cy)[:, None]
    kx = np.fft.fftfreq(cx)[None, :]
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-2 * np.pi ** 2 * ((ky * sy) ** 2 + (kx * sx) ** 2))))
    f = (f - f.mean()) / f.std() * std
    yi = np.arange(lines) / coarse
    xi = np.arange(samples) / coarse
    iy = np.floor(yi).astype(np.int64)
    ix = np.floor(xi).astype(np.int64)
    wy = (yi - iy)[:, None]
    wx = (xi - ix)[None, :]
    out = ((f[iy][:, ix] * (1 - wx) + f[iy][:, ix + 1] * wx) * (1 - wy)
           + (f[iy + 1][:, ix] * (1 - wx) + f[iy + 1][:, ix + 1] * wx) * wy)
    return out.astype(np.float32)


def ps_stack(nifg, nrow, ncol, seed=0, ps_frac=0.15, extra_frac=0.1, ps_noise=0.35,
             aps_std=1.5, device=None):
    """Interferogram stack for PS selection: nifg interferograms of
    nrow x ncol complex samples, image after image.  A fraction ps_frac of
    the pixels are stable scatterers, whose phase in interferogram k is a
    smooth screen (aps_field, aps_std radians) plus Gaussian noise of
    ps_noise radians and whose amplitude is steady; the others carry a
    uniform random phase and Rayleigh amplitude.  The candidate mask
    (int32) marks the scatterers plus a fraction extra_frac of the other
    pixels.  Returns (stack, candidates, truth): numpy arrays, or torch
    tensors on `device` (truth, the scatterer mask, stays numpy)."""
    rng = np.random.default_rng(seed)
    truth = rng.random((nrow, ncol)) < ps_frac
    cand = (truth | (rng.random((nrow, ncol)) < extra_frac)).astype(np.int32)
    stack = np.empty((nifg, nrow, ncol), np.complex64)
    amp_ps = rng.lognormal(0.0, 0.2, (nrow, ncol)).astype(np.float32)
    for k in range(nifg):
        screen = aps_field(nrow, ncol, seed * 1000 + 31 + k, corr=2000.0, std=aps_std)
        noise = np.where(truth, rng.normal(0.0, ps_noise, (nrow, ncol)),
                         rng.uniform(-np.pi, np.pi, (nrow, ncol))).astype(np.float32)
        amp = np.where(truth, amp_ps, rng.rayleigh(0.6, (nrow, ncol))).astype(np.float32)
        ph = screen + noise
        stack[k].real = amp * np.cos(ph)
        stack[k].imag = amp * np.sin(ph)
    if device is None:
        return stack, cand, truth
    import torch
    return (torch.from_numpy(stack).to(device), torch.from_numpy(cand).to(device), truth)


def make_stack(n_dates=20, lines=1500, samples=25000, dem_shape=(2000, 3000),
               dem_kind="fractal", hmax=2500.0, seed=0, gamma0=0.8, bperp_max=150.0,
               aps_corr=2000.0, aps_std=1.0, fidelity="fast", device=None):
    """Synthetic stack of n_dates acquisitions of one burst footprint.

    Date 0 is the reference geometry; date k > 0 flies a parallel orbit with
    a perpendicular baseline drawn uniformly from [-bperp_max, bperp_max]
    (|b| >= 10 m) and carries the common speckle (coherence gamma0 with date
    0), its own noise, and its own atmospheric phase screen (aps_field,
    corr metres, aps_std radians); date 0 carries a screen too.  Returns a
    dict with the per-date orbits, bursts and SLCs, the shared grid and DEM,
    the baselines and the screens."""
    rng = np.random.default_rng(seed + 7919)
    b = rng.uniform(-bperp_max, bperp_max, n_dates)
    b = np.where(np.abs(b) < 10.0, np.copysign(10.0, b), b)
    b[0] = 0.0
    screens = [aps_field(lines, samples, seed * 1000 + 500 + k, aps_corr, aps_std)
               for k in range(n_dates)]
    dates = []
    base = None
    for k in range(1, n_dates):
        sc = make_scene(lines=lines, samples=samples, dem_shape=dem_shape, dem_kind=dem_kind,
                        hmax=hmax, seed=seed, gamma0=gamma0, baseline=(float(b[k]), 0.0),
                        amb_height=500.0 * 120.0 / float(b[k]), fidelity=fidelity,
                        device=device, sec_seed=seed * 1000 + k, aps_sec=screens[k])
        if base is None:
            base = sc
            ref = sc["ref_slc"]
            if fidelity == "exact":
                ref = (ref.astype(np.complex128) * np.exp(1j * screens[0])).astype(np.complex64)
            else:
                import torch
                a0 = torch.as_tensor(screens[0], device=ref.device)
                ref = (ref * torch.polar(torch.ones_like(a0), a0)).to(torch.complex64)
            dates.append(dict(orbit=sc["orbit_ref"], burst=sc["burst_ref"], slc=ref))
        dates.append(dict(orbit=sc["orbit_sec"], burst=sc["burst_sec"], slc=sc["sec_slc"]))
        sc["ref_slc"] = None
    return dict(dates=dates, grid=base["grid"], dem=base["dem"], bperp=b, aps=screens)


def make_scene(lines=512, samples=2048, dem_shape=(256, 256), dem_kind="hill", hmax=1500.0,
               seed=0, gamma0=0.8, phase_noise=0.5, amb_height=500.0, heading=-10.0,
               look_range=800.0e3, baseline=(120.0, 30.0), misreg=None, misreg_ramp=(0.0, 0.0),
               fidelity="exact", device=None, lat_c=34.0, lon_c=-118.0, t_center=43200.0,
               margin=24, sec_seed=None, aps_sec=None, tops="constant"):
    """Build one synthetic burst pair.

    tops: "constant" (FM rate -2300 Hz/s, zero Doppler centroid, three equal
    records) or "varying" (range-dependent FM rate and Doppler centroid whose
    records change with time, the burst centre between two records): see
    tops_records.

    misreg: (range px, azimuth px) constant misregistration of the secondary
    content (corrected by the offset polynomial p = [drg, rg_ramp, 0, daz,
    0, az_ramp]); misreg_ramp: linear ramps (pixels across the burst in range
    for both).  sec_seed: separate seed of the secondary's own noise (the
    reference speckle still follows `seed`); aps_sec: phase screen (radians,
    lines x samples) multiplied into the secondary.  Returns a dict with
    orbits, bursts, grid, dem, SLCs and the offset polynomial that realigns
    the secondary.
    """
    rng = np.random.default_rng(seed)
    L, N = lines, samples
    up, east, north = _frame(lat_c, lon_c)
    Xc = llh2xyz(lat_c, lon_c, 0.0)
    psi = np.radians(heading)
    vhat = np.cos(psi) * north + np.sin(psi) * east
    rhat = np.cos(psi) * east - np.sin(psi) * north
    g = math.sqrt(look_range ** 2 - ALTITUDE ** 2)
    s0 = Xc - g * rhat + ALTITUDE * up
    v = SPEED * vhat
    orb_ref = LineOrbit(s0, v, t_center)
    look = (Xc - s0) / np.linalg.norm(Xc - s0)
    perp = np.cross(look, vhat)
    perp /= np.linalg.norm(perp)
    orb_sec = LineOrbit(s0 + baseline[0] * perp + baseline[1] * look, v, t_center)

    Rc = float(np.linalg.norm(Xc - s0))
    R0 = Rc - (N // 2) * RANGE_PIXEL
    t_start = t_center - (L - 1) / 2 * AZ_INTERVAL
    t_end = t_start + (L - 1) * AZ_INTERVAL
    n_sv = int(math.ceil(t_end - t_start)) + 12
    t0 = math.floor(t_start) - 6
    fm, dc = tops_records(t_center, 2 * Rc / SOL, tops)
    flv = max(2, L // 50)
    fvs = max(2, N // 100)
    burst = dict(lines=L, samples=N, first_valid_line=flv, last_valid_line=L - flv,
                 first_valid_sample=fvs, last_valid_sample=N - fvs, t_start=t_start,
                 prf=1.0 / AZ_INTERVAL, azimuth_time_interval=AZ_INTERVAL,
                 slant_range_time=2 * R0 / SOL, range_sampling_rate=SOL / (2 * RANGE_PIXEL),
                 wavelength=WAVELENGTH, radar_frequency=SOL / WAVELENGTH,
                 azimuth_steering_rate=STEERING_RATE, fm=fm, dc=dc)
    burst_sec = dict(burst)
    orbit_ref = orb_ref.state_vectors(t0, n_sv)
    orbit_sec = orb_sec.state_vectors(t0, n_sv)

    # footprint at h = 0 and hmax -> DEM and target grids
    tt = np.array([t_start, t_start, t_end, t_end])
    RR = np.array([R0, R0 + (N - 1) * RANGE_PIXEL] * 2)
    flat = {"data": np.zeros((2, 2), np.float32), "lat0": lat_c + 1, "lon0": lon_c - 1,
            "dlat": -2.0, "dlon": 2.0}
    lats, lons = [], []
    for hh in (0.0, hmax):
        flat["data"][:] = hh
        Xb = rdr2geo(orb_ref, tt, RR, flat)
        la, lo, _ = xyz2llh(Xb)
        lats.append(la)
        lons.append(lo)
    lats = np.concatenate(lats)
    lons = np.concatenate(lons)
    lat_min, lat_max, lon_min, lon_max = lats.min(), lats.max(), lons.min(), lons.max()
    pad_lat = 0.04 * (lat_max - lat_min)
    pad_lon = 0.04 * (lon_max - lon_min)
    dny, dnx = dem_shape
    dem = {"lat0": lat_max + pad_lat, "lon0": lon_min - pad_lon,
           "dlat": -(lat_max - lat_min + 2 * pad_lat) / (dny - 1),
           "dlon": (lon_max - lon_min + 2 * pad_lon) / (dnx - 1)}
    if dem_kind == "hill":
        dem["data"] = gaussian_hill(dny, dnx, hmax)
    else:
        dem["data"] = fractal_dem(dny, dnx, rng, hmax=hmax)
    grid = {"nrow": L, "ncol": N, "lat0": lat_max, "lon0": lon_min,
            "dlat": -(lat_max - lat_min) / (L - 1), "dlon": (lon_max - lon_min) / (N - 1)}

    # offset polynomial that realigns the misregistered secondary
    offpoly = None
    if misreg is not None:
        offpoly = np.array([misreg[0], misreg_ramp[0], 0.0, misreg[1], 0.0, misreg_ramp[1]])

    out = dict(orbit_ref=orbit_ref, orbit_sec=orbit_sec, burst_ref=burst, burst_sec=burst_sec,
               grid=grid, dem=dem, offpoly=offpoly, gamma0=gamma0, amb_height=amb_height,
               line_orbits=(orb_ref, orb_sec))
    if fidelity == "none":
        return out
    if fidelity == "exact":
        ref, sec = _slcs_exact(rng, orb_ref, orb_sec, burst, dem, gamma0, phase_noise,
                               amb_height, offpoly, margin, sec_seed)
        if aps_sec is not None:
            sec = sec * np.exp(1j * np.asarray(aps_sec, np.float64))
        out["ref_slc"] = ref.astype(np.complex64)
        out["sec_slc"] = sec.astype(np.complex64)
    else:
        out.update(_slcs_fast(rng, orb_ref, orb_sec, burst, dem, gamma0, offpoly, device,
                              sec_seed, aps_sec))
    return out


def _sec_positions(burst, offpoly, ii, jj):
    """Geometric secondary position whose content sits at sample (ii, jj)."""
    if offpoly is None:
        return ii.astype(np.float64), jj.astype(np.float64)
    p = offpoly
    xs = jj / burst["samples"]
    ys = ii / burst["lines"]
    drg = p[0] + p[1] * xs + p[2] * ys
    daz = p[3] + p[4] * xs + p[5] * ys
    return ii - daz, jj - drg


def _slcs_exact(rng, orb_ref, orb_sec, burst, dem, gamma0, phase_noise, amb_height, offpoly,
                margin, sec_seed=None):
    L, N = burst["lines"], burst["samples"]
    R0 = burst["slant_range_time"] * SOL / 2
    m = margin
    shape = (L + 2 * m, N + 2 * m)
    z = bandlimited_noise(rng, shape)
    rs = rng if sec_seed is None else np.random.default_rng(sec_seed)
    n = bandlimited_noise(rs, shape)
    eps = np.real(bandlimited_noise(rs, shape, 0.3, 0.3)) * phase_noise
    ref = z[m:m + L, m:m + N] * np.exp(1j * ramp(orb_ref, burst))
    ii, jj = np.meshgrid(np.arange(L), np.arange(N), indexing="ij")
    ia, ja = _sec_positions(burst, offpoly, ii, jj)
    t = burst["t_start"] + ia * burst["azimuth_time_interval"]
    R = R0 + ja * RANGE_PIXEL
    P = rdr2geo(orb_sec, t.ravel(), R.ravel(), dem)
    tr, Rr = orb_ref.zero_doppler(P)
    a = (tr - burst["t_start"]) / burst["azimuth_time_interval"] + m
    r = (Rr - R0) / RANGE_PIXEL + m
    lat, lon, _ = xyz2llh(P)
    h = dem_height(dem, lat, lon)
    a = np.clip(a, 6, shape[0] - 7)
    r = np.clip(r, 6, shape[1] - 7)
    zs = sinc_resample(z, a, r)
    ns = sinc_resample(n, a, r)
    es = np.real(sinc_resample(eps.astype(np.complex128), a, r))
    dR = R.ravel() - Rr
    phi = 2 * np.pi * h / amb_height + es
    sec = ((gamma0 * zs + math.sqrt(1 - gamma0 ** 2) * ns) * np.exp(-1j * 4 * np.pi * dR / WAVELENGTH)
           * np.exp(1j * phi)).reshape(L, N)
    sec = sec * np.exp(1j * ramp(orb_sec, burst))
    return ref, sec


def _slcs_fast(rng, orb_ref, orb_sec, burst, dem, gamma0, offpoly, device, sec_seed=None,
               aps_sec=None):
    import torch
    import torch.nn.functional as F
    L, N = burst["lines"], burst["samples"]
    R0 = burst["slant_range_time"] * SOL / 2
    dev = torch.device(device or "cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(rng.integers(1 << 31)))

    gen2 = gen
    if sec_seed is not None:
        gen2 = torch.Generator(device=dev)
        gen2.manual_seed(int(sec_seed))

    def noise(g=gen):
        w = torch.complex(torch.randn(L, N, device=dev, generator=g),
                          torch.randn(L, N, device=dev, generator=g)) / math.sqrt(2)
        fa = torch.fft.fftfreq(L, device=dev).abs()[:, None] * 2
        fr = torch.fft.fftfreq(N, device=dev).abs()[None, :] * 2
        win = ((fa < 0.7) * 0.5 * (1 + torch.cos(math.pi * torch.clamp(fa / 0.7, max=1)))
               * (fr < 0.85) * 0.5 * (1 + torch.cos(math.pi * torch.clamp(fr / 0.85, max=1))))
        zz = torch.fft.ifft2(torch.fft.fft2(w) * win)
        return zz / torch.sqrt(torch.mean(zz.abs() ** 2))

    z = noise()
    n2 = noise(gen2)
    # coarse secondary mapping (exact geometry), upsampled on the GPU
    cs, cr = max(L // 24, 2), max(N // 96, 2)
    ic = np.linspace(0, L - 1, cs)
    jc = np.linspace(0, N - 1, cr)
    ii, jj = np.meshgrid(ic, jc, indexing="ij")
    ia, ja = _sec_positions(burst, offpoly, ii, jj)
    t = burst["t_start"] + ia * burst["azimuth_time_interval"]
    R = R0 + ja * RANGE_PIXEL
    P = rdr2geo(orb_sec, t.ravel(), R.ravel(), dem)
    tr, Rr = orb_ref.zero_doppler(P)
    a = ((tr - burst["t_start"]) / burst["azimuth_time_interval"]).reshape(cs, cr)
    r = ((Rr - R0) / RANGE_PIXEL).reshape(cs, cr)
    dR = (R.ravel() - Rr).reshape(cs, cr)
    coarse = torch.tensor(np.stack([a, r, dR]), dtype=torch.float64, device=dev)[None]
    fine = F.interpolate(coarse, size=(L, N), mode="bilinear", align_corners=True)[0]
    gy = fine[0] / (L - 1) * 2 - 1
    gx = fine[1] / (N - 1) * 2 - 1
    grid = torch.stack([gx, gy], dim=-1)[None].float()

    def sample(f):
        two = torch.stack([f.real, f.imag])[None].float()
        s = F.grid_sample(two, grid, mode="bilinear", padding_mode="border", align_corners=True)[0]
        return torch.complex(s[0], s[1])

    zs = sample(z)
    ns = sample(n2)
    kt, etaref, eta = deramp_coefficients(orb_ref, burst)
    d = torch.tensor(eta, device=dev)[:, None] - torch.tensor(etaref, device=dev)[None, :]
    ph = (math.pi * torch.tensor(kt, device=dev)[None, :] * d * d)
    ph = torch.remainder(ph, 2 * math.pi).float()
    rmp = torch.polar(torch.ones_like(ph), ph)
    ref = (z * rmp).to(torch.complex64)
    phs = torch.remainder(-4 * math.pi * fine[2] / WAVELENGTH, 2 * math.pi).float()
    if aps_sec is not None:
        aps = torch.as_tensor(aps_sec, device=dev).float()
        phs = torch.remainder(phs + aps, 2 * math.pi)
    sec = (gamma0 * zs + math.sqrt(1 - gamma0 ** 2) * ns) * torch.polar(torch.ones_like(phs), phs)
    sec = (sec * rmp).to(torch.complex64)
    return {"ref_slc": ref.contiguous(), "sec_slc": sec.contiguous()}
