// Context, error reporting and the host FP64 set-up of a burst.
//
// The set-up restates geo2rdr.cu:386-400 and 465-549 (orbit state at the
// burst centre, ks, FM/DC coefficient interpolation by closest_sample rules,
// kt/etaref per range sample) and bounds.cu:237-317 (footprint), and builds
// the polynomial form of the 4-node Hermite orbit interpolation of
// orbit.cu:170-245 that the kernels evaluate.

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "s1ifg_internal.h"

namespace s1ifg {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_error(hipError_t e, const char *what) {
  if (e == hipErrorOutOfMemory)
    return set_error(S1IFG_ENOMEM, "%s: %s", what, hipGetErrorString(e));
  return set_error(S1IFG_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int check_ctx(s1ifg_ctx *ctx) {
  if (!ctx) return set_error(S1IFG_EINVAL, "null context");
  int dev = -1;
  S1_HIP(hipGetDevice(&dev));
  if (dev != ctx->device) S1_HIP(hipSetDevice(ctx->device));
  return S1IFG_OK;
}

int check_geom(s1ifg_ctx *ctx, const s1ifg_geom *g) {
  if (!g) return set_error(S1IFG_EINVAL, "null geometry");
  if (g->device != ctx->device)
    return set_error(S1IFG_ESTATE, "geometry built on device %d used on device %d",
                     g->device, ctx->device);
  return S1IFG_OK;
}

int ctx_workspace(s1ifg_ctx *ctx, size_t bytes, void **out) {
  if (bytes <= ctx->ws_bytes) {
    *out = ctx->ws;
    return S1IFG_OK;
  }
  if (ctx->ws_limit >= 0 && (int64_t)bytes > ctx->ws_limit)
    return set_error(S1IFG_ENOMEM, "workspace of %zu bytes exceeds the limit of %lld bytes",
                     bytes, (long long)ctx->ws_limit);
  if (ctx->ws) {
    S1_HIP(hipStreamSynchronize(ctx->stream));
    S1_HIP(hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  size_t want = bytes + bytes / 8;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) return hip_error(e, "workspace hipMalloc");
  ctx->ws = p;
  ctx->ws_bytes = want;
  *out = p;
  return S1IFG_OK;
}

int make_grid(const s1ifg_grid *grid, GridDev *out) {
  if (!grid) return set_error(S1IFG_EINVAL, "null grid");
  if (grid->nrow <= 0 || grid->ncol <= 0)
    return set_error(S1IFG_EINVAL, "grid size %d x %d", grid->nrow, grid->ncol);
  if ((int64_t)grid->nrow * grid->ncol > (int64_t)1 << 31)
    return set_error(S1IFG_ERANGE, "grid of %d x %d pixels exceeds 2^31", grid->nrow, grid->ncol);
  if (!(grid->dlat != 0.0) || !(grid->dlon != 0.0) || !std::isfinite(grid->lat0) ||
      !std::isfinite(grid->lon0))
    return set_error(S1IFG_EINVAL, "grid spacing must be non-zero and origin finite");
  out->nrow = grid->nrow;
  out->ncol = grid->ncol;
  out->lat0 = grid->lat0;
  out->lon0 = grid->lon0;
  out->dlat = grid->dlat;
  out->dlon = grid->dlon;
  return S1IFG_OK;
}

int make_dem(const s1ifg_dem *dem, DemDev *out) {
  if (!dem || !dem->data) return set_error(S1IFG_EINVAL, "null DEM");
  if (dem->dtype != S1IFG_DEM_F32 && dem->dtype != S1IFG_DEM_I16)
    return set_error(S1IFG_EINVAL, "DEM dtype %d", dem->dtype);
  if (dem->nrow < 2 || dem->ncol < 2)
    return set_error(S1IFG_EINVAL, "DEM must be at least 2 x 2 (got %d x %d)", dem->nrow, dem->ncol);
  if (!(dem->dlat != 0.0) || !(dem->dlon != 0.0))
    return set_error(S1IFG_EINVAL, "DEM spacing must be non-zero");
  out->data = dem->data;
  out->dtype = dem->dtype;
  out->nrow = dem->nrow;
  out->ncol = dem->ncol;
  out->lat0 = dem->lat0;
  out->lon0 = dem->lon0;
  out->dlat = dem->dlat;
  out->dlon = dem->dlon;
  return S1IFG_OK;
}

int make_offpoly(const double *p, OffPoly *out) {
  std::memset(out, 0, sizeof(*out));
  if (!p) return S1IFG_OK;
  for (int i = 0; i < 6; ++i) {
    if (!std::isfinite(p[i])) return set_error(S1IFG_EINVAL, "offset polynomial not finite");
    out->p[i] = p[i];
  }
  out->on = 1;
  return S1IFG_OK;
}

// ---------------------------------------------------------------- geodesy
static void llh2xyz(const double llh[3], double xyz[3]) {
  double lat = llh[0] * kPi / 180, lon = llh[1] * kPi / 180, h = llh[2];
  double sl = std::sin(lat), cl = std::cos(lat);
  double re = kRa / std::sqrt(1. - kRe2 * sl * sl);
  xyz[0] = (re + h) * cl * std::cos(lon);
  xyz[1] = (re + h) * cl * std::sin(lon);
  xyz[2] = (re * (1 - kRe2) + h) * sl;
}

static void xyz2llh(const double xyz[3], double llh[3]) {
  double x = xyz[0], y = xyz[1], z = xyz[2];
  double q2 = 1. / (1. - kRe2), q = std::sqrt(q2), q3 = q2 - 1., b = kRa * std::sqrt(1 - kRe2);
  double lon = std::atan2(y, x);
  if (y < 0 && lon > 0) lon -= kPi;
  if (y > 0 && lon < 0) lon += kPi;
  double p = std::sqrt(x * x + y * y);
  double th = std::atan((z / p) * q);
  double st = std::sin(th), ct = std::cos(th);
  double lat = std::atan((z + q3 * b * st * st * st) / (p - kRe2 * kRa * ct * ct * ct));
  double sl = std::sin(lat);
  double re = kRa / std::sqrt(1. - kRe2 * sl * sl);
  llh[0] = lat / kPi * 180;
  llh[1] = lon / kPi * 180;
  llh[2] = p / std::cos(lat) - re;
}

// ---------------------------------------------------- orbit polynomials
// Polynomial helpers on coefficient arrays (ascending powers).
static void pmul(const double *a, int na, const double *b, int nb, double *c) {
  for (int i = 0; i < na + nb - 1; ++i) c[i] = 0;
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j) c[i + j] += a[i] * b[j];
}

// Hermite interpolant on nodes s[0..3] (values x, derivatives h*v in s-units)
// -> coef[comp*8 + m].
static void hermite_poly(const double s[4], const double x[4][3], const double v[4][3],
                         double h, double coef[24]) {
  for (int i = 0; i < 24; ++i) coef[i] = 0;
  for (int i = 0; i < 4; ++i) {
    double L[4] = {1, 0, 0, 0};
    int nl = 1;
    double dl = 0;
    for (int j = 0; j < 4; ++j) {
      if (j == i) continue;
      double f[2] = {-s[j] / (s[i] - s[j]), 1.0 / (s[i] - s[j])};
      double t[4];
      pmul(L, nl, f, 2, t);
      nl += 1;
      for (int k = 0; k < nl; ++k) L[k] = t[k];
      dl += 1.0 / (s[i] - s[j]);
    }
    double L2[7];
    pmul(L, 4, L, 4, L2);
    // a_i = (1 - 2 (s - s_i) dl) L^2 ; b_i = h (s - s_i) L^2
    double fa[2] = {1 + 2 * s[i] * dl, -2 * dl};
    double fb[2] = {-s[i] * h, h};
    double A[8], B[8];
    pmul(fa, 2, L2, 7, A);
    pmul(fb, 2, L2, 7, B);
    for (int c = 0; c < 3; ++c)
      for (int m = 0; m < 8; ++m) coef[c * 8 + m] += A[m] * x[i][c] + B[m] * v[i][c];
  }
}

// Taylor coefficients of p(s0 + e) = sum_j c[j] e^j from the ascending
// power coefficients a[0..7] of p(s).
static void poly_shift(const double *a, double s0, double c[8]) {
  for (int j = 0; j < 8; ++j) {
    double sum = 0, binom = 1, pw = 1;  // C(i, j) s0^(i-j) for i = j, j+1, ...
    for (int i = j; i < 8; ++i) {
      sum += a[i] * binom * pw;
      binom = binom * (i + 1) / (i + 1 - j);
      pw *= s0;
    }
    c[j] = sum;
  }
}

// Node intervals per stencil interval for the orbit node table: the
// smallest M (4 .. 256) at which a cubic Taylor step from the nearest of
// the M + 1 nodes of an interval, |e| <= 1 / (2 M) in interval units,
// reproduces every interval polynomial to within 1e-10 m, 1e-10 m/s and
// 1e-8 m/s^2 (the sum of the omitted terms bounds the difference); 0 when
// none does.  nst intervals of coef[24] (ascending powers per component).
static int orbit_node_count(const std::vector<double> &coef, int nst, double h) {
  for (int M = 4; M <= 256; M *= 2) {
    const double e = 0.5 / M;
    bool ok = true;
    for (int k = 0; k < nst && ok; ++k)
      for (int m = 0; m <= M && ok; ++m)
        for (int comp = 0; comp < 3 && ok; ++comp) {
          double c[8];
          poly_shift(&coef[(size_t)k * 24 + comp * 8], (double)m / M, c);
          double rx = 0, rv = 0, ra = 0;
          for (int j = 4; j < 8; ++j) {
            rx += std::fabs(c[j]) * std::pow(e, j);
            rv += j * std::fabs(c[j]) * std::pow(e, j - 1) / h;
            ra += j * (j - 1) * std::fabs(c[j]) * std::pow(e, j - 2) / (h * h);
          }
          ok = rx < 1e-10 && rv < 1e-10 && ra < 1e-8;
        }
    if (ok) return M;
  }
  return 0;
}

static int stencil(const double *t, int n, double tq) {
  int il = (int)((tq - t[0]) / (t[1] - t[0]));  // orbit.cu:229 (truncation)
  il = il > 1 ? il : 1;
  il = il < n - 3 ? il : n - 3;
  return il;
}

// Evaluate the polynomial form on the host (same arithmetic as the kernels).
static void orbit_eval_host(const std::vector<double> &coef, const double *t, int n,
                            double tq, double x[3], double v[3], double a[3]) {
  int k = stencil(t, n, tq);
  double h = t[1] - t[0];
  double s = (tq - t[k]) / h;
  const double *c = &coef[(size_t)(k - 1) * 24];
  for (int comp = 0; comp < 3; ++comp) {
    const double *cc = c + comp * 8;
    double p = cc[7], dp = 0, ddp = 0;
    for (int m = 6; m >= 0; --m) {
      ddp = ddp * s + 2 * dp;
      dp = dp * s + p;
      p = p * s + cc[m];
    }
    x[comp] = p;
    v[comp] = dp / h;
    a[comp] = ddp / (h * h);
  }
}

// Longest squared time step dt^2 for which one Newton step of the
// zero-Doppler equation f(t) = (X - x(t)).v(t) = 0, with the exact
// derivative f' = -|v|^2 + (X - x).a, from a point within dt of the root
// leaves an error below 5e-10 s (a tenth of the reference's 5e-9 s
// tolerance): the step's error is f''/(2 f') dt^2 with
// |f''| <= 3 |v| |a| + |X - x| |j|, bounded over the burst time +- 3 s and
// for slant ranges up to rmax + 50 km.  The constant is floored at
// 1e-3 / s, the value of a curved orbit at 700 km, so that the rule does
// not depend on the orbit being straight.
static double warm_step_limit(const std::vector<double> &coef, const double *t, int n,
                              double t0, double t1, double rmax) {
  double vmin = INFINITY, vmax = 0, amax = 0, jmax = 0;
  const int m = 256;
  for (int i = 0; i <= m; ++i) {
    const double tq = t0 - 3.0 + (t1 - t0 + 6.0) * i / m;
    double x[3], v[3], a[3], xa[3], va[3], aa[3], xb[3], vb[3], ab[3];
    const double hh = 1e-2;
    orbit_eval_host(coef, t, n, tq, x, v, a);
    orbit_eval_host(coef, t, n, tq - hh, xa, va, aa);
    orbit_eval_host(coef, t, n, tq + hh, xb, vb, ab);
    const double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double an = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    double jn = 0;
    for (int k = 0; k < 3; ++k) jn += std::pow((ab[k] - aa[k]) / (2 * hh), 2);
    vmin = std::fmin(vmin, vn);
    vmax = std::fmax(vmax, vn);
    amax = std::fmax(amax, an);
    jmax = std::fmax(jmax, std::sqrt(jn));
  }
  const double R = rmax + 50e3;
  const double den = vmin * vmin - R * amax;
  if (!(den > 0)) return 0.0;  // no warm steps
  const double C = std::fmax(2.0 * (3 * vmax * amax + R * jmax) / (2 * den), 1e-3);
  return 5e-10 / C;
}

static int closest_sample(const double *rec, int n, double t0) {
  // geo2rdr.cu:48-59 on records (t, t0, c0, c1, c2)
  int idx = 0;
  double mindist = 1e10;
  for (int i = 0; i < n; ++i) {
    double d = std::fabs(rec[i * 5] - t0);
    if (mindist > d) {
      mindist = d;
      idx = i;
    }
  }
  return idx;
}

static void interp_record(const double *rec, int n, double tc, double out[4]) {
  // geo2rdr.cu:485-527
  int idx = closest_sample(rec, n, tc);
  const double *r0;
  const double *r1;
  double frac;
  if (rec[idx * 5] < tc && idx < n - 1) {
    r0 = rec + idx * 5;
    r1 = rec + (idx + 1) * 5;
    frac = (tc - r0[0]) / (r1[0] - r0[0]);
  } else if (rec[idx * 5] > tc && idx > 0) {
    r0 = rec + (idx - 1) * 5;
    r1 = rec + idx * 5;
    frac = (tc - r0[0]) / (r1[0] - r0[0]);
  } else {
    for (int k = 0; k < 4; ++k) out[k] = rec[idx * 5 + 1 + k];
    return;
  }
  for (int k = 0; k < 4; ++k) out[k] = r0[1 + k] + frac * (r1[1 + k] - r0[1 + k]);
}

}  // namespace s1ifg

using namespace s1ifg;

extern "C" {

int s1ifg_abi_version(void) { return S1IFG_ABI_VERSION; }

const char *s1ifg_last_error(void) { return g_last_error.c_str(); }

int s1ifg_ctx_create(int device, s1ifg_ctx **out) {
  if (!out) return set_error(S1IFG_EINVAL, "null output pointer");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return hip_error(e, "hipGetDeviceCount");
  if (device < 0 || device >= n)
    return set_error(S1IFG_EINVAL, "device %d out of range (have %d)", device, n);
  S1_HIP(hipSetDevice(device));
  s1ifg_ctx *c = new s1ifg_ctx();
  c->device = device;
  c->stream = nullptr;
  c->ws = nullptr;
  c->ws_bytes = 0;
  c->ws_limit = -1;
  c->tile_stats = nullptr;
  c->ps_cache = nullptr;
  const char *lim = std::getenv("S1IFG_WORKSPACE_LIMIT");
  if (lim && *lim) c->ws_limit = std::atoll(lim);
  *out = c;
  return S1IFG_OK;
}

void s1ifg_ctx_destroy(s1ifg_ctx *ctx) {
  if (!ctx) return;
  // Destructors report nothing: errors here leave only a leak.
  (void)hipSetDevice(ctx->device);
  if (ctx->ws) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->ws);
  }
  ps_cache_free(ctx);
  delete ctx;
}

int s1ifg_ctx_set_stream(s1ifg_ctx *ctx, void *stream) {
  if (!ctx) return set_error(S1IFG_EINVAL, "null context");
  ctx->stream = (hipStream_t)stream;
  return S1IFG_OK;
}

int s1ifg_ctx_set_tile_stats(s1ifg_ctx *ctx, uint64_t *counts) {
  if (!ctx) return set_error(S1IFG_EINVAL, "null context");
  ctx->tile_stats = reinterpret_cast<unsigned long long *>(counts);
  return S1IFG_OK;
}

int s1ifg_ctx_sync(s1ifg_ctx *ctx) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  S1_HIP(hipStreamSynchronize(ctx->stream));
  return S1IFG_OK;
}

int s1ifg_ctx_workspace(s1ifg_ctx *ctx, int64_t *bytes) {
  if (!ctx || !bytes) return set_error(S1IFG_EINVAL, "null argument");
  *bytes = (int64_t)ctx->ws_bytes;
  return S1IFG_OK;
}

int s1ifg_geom_create(s1ifg_ctx *ctx, const s1ifg_orbit *orbit, const s1ifg_burst *b,
                      s1ifg_geom **out) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!orbit || !b || !out) return set_error(S1IFG_EINVAL, "null argument");
  *out = nullptr;
  const int n = orbit->n;
  if (n < 4 || !orbit->t || !orbit->pos || !orbit->vel)
    return set_error(S1IFG_EINVAL, "orbit needs >= 4 state vectors (got %d)", n);
  const double h = orbit->t[1] - orbit->t[0];
  if (!(h > 0)) return set_error(S1IFG_EINVAL, "orbit times must increase");
  for (int i = 1; i < n; ++i) {
    double d = orbit->t[i] - orbit->t[i - 1];
    if (!(std::fabs(d - h) <= 1e-6 * h))
      return set_error(S1IFG_EINVAL, "orbit state vectors must be uniformly spaced");
  }
  if (b->lines < 2 || b->samples < 2)
    return set_error(S1IFG_EINVAL, "burst %d x %d too small", b->lines, b->samples);
  if (b->first_valid_line < 0 || b->last_valid_line > b->lines ||
      b->first_valid_line >= b->last_valid_line || b->first_valid_sample < 0 ||
      b->last_valid_sample > b->samples || b->first_valid_sample >= b->last_valid_sample)
    return set_error(S1IFG_EINVAL,
                     "valid window lines [%d,%d) samples [%d,%d) outside burst %d x %d",
                     b->first_valid_line, b->last_valid_line, b->first_valid_sample,
                     b->last_valid_sample, b->lines, b->samples);
  if (!(b->prf > 0) || !(b->azimuth_time_interval > 0) || !(b->range_sampling_rate > 0) ||
      !(b->wavelength > 0))
    return set_error(S1IFG_EINVAL, "prf, azimuth interval, sampling rate and wavelength must be > 0");
  if (b->n_fm < 1 || b->n_dc < 1 || !b->fm || !b->dc)
    return set_error(S1IFG_EINVAL, "FM-rate and Doppler records required");

  // orbit polynomials, one per stencil index 1..n-3
  std::vector<double> coef((size_t)(n - 3) * 24);
  for (int k = 1; k <= n - 3; ++k) {
    double s[4], x[4][3], v[4][3];
    for (int i = 0; i < 4; ++i) {
      int j = k - 1 + i;
      s[i] = (orbit->t[j] - orbit->t[k]) / h;
      for (int c = 0; c < 3; ++c) {
        x[i][c] = orbit->pos[j * 3 + c];
        v[i][c] = orbit->vel[j * 3 + c];
      }
    }
    hermite_poly(s, x, v, h, &coef[(size_t)(k - 1) * 24]);
  }

  GeomDev gd;
  std::memset(&gd, 0, sizeof(gd));
  const int L = b->lines, N = b->samples;
  // geo2rdr.cu:477-483
  double tc = b->t_start + L * b->azimuth_time_interval / 2;
  double xp[3], vp[3], ap[3];
  orbit_eval_host(coef, orbit->t, n, tc, xp, vp, ap);
  double vs = std::sqrt(vp[0] * vp[0] + vp[1] * vp[1] + vp[2] * vp[2]);
  double ks = 2.0 * vs * b->radar_frequency / kSol * b->azimuth_steering_rate * kPi / 180.0;
  interp_record(b->fm, b->n_fm, tc, gd.fm);
  interp_record(b->dc, b->n_dc, tc, gd.dc);
  std::vector<double> kt(N), etaref(N);
  double etac0 = 0;
  for (int i = 0; i < N; ++i) {  // geo2rdr.cu:533-542
    double trange = b->slant_range_time + i / b->range_sampling_rate;
    double dt = trange - gd.fm[0];
    double ka = gd.fm[1] + dt * (gd.fm[2] + gd.fm[3] * dt);
    dt = trange - gd.dc[0];
    double fnc = gd.dc[1] + dt * (gd.dc[2] + gd.dc[3] * dt);
    kt[i] = ka * ks / (ka - ks);
    double etac = -fnc / ka;
    if (i == 0) etac0 = etac;
    etaref[i] = etac - etac0;
  }
  // geo2rdr.cu:588-596
  gd.tstart = b->t_start;
  gd.dtaz = 1. / b->prf;
  double tend = gd.tstart + (L - 1) * gd.dtaz;
  gd.tmid = (gd.tstart + tend) * 0.5;
  orbit_eval_host(coef, orbit->t, n, gd.tmid, gd.xmid, gd.vmid, gd.amid);
  gd.rngstart = b->slant_range_time * kSol / 2.0;
  gd.dmrg = kSol / 2.0 / b->range_sampling_rate;
  gd.wvl = b->wavelength;
  gd.ks = ks;
  gd.etac0 = etac0;
  gd.ati = b->azimuth_time_interval;
  gd.lines = L;
  gd.samples = N;
  gd.flv = b->first_valid_line;
  gd.llv = b->last_valid_line;
  gd.fvs = b->first_valid_sample;
  gd.lvs = b->last_valid_sample;

  // device records: [t_k, coef[24], pad] per stencil index (kOrbRec doubles)
  const size_t nrec = (size_t)(n - 3) * 32;
  std::vector<double> recs(nrec, 0.0);
  for (int k = 1; k <= n - 3; ++k) {
    recs[(size_t)(k - 1) * 32] = orbit->t[k];
    for (int m = 0; m < 24; ++m) recs[(size_t)(k - 1) * 32 + 1 + m] = coef[(size_t)(k - 1) * 24 + m];
  }
  // orbit node table: M + 1 nodes per stencil interval
  const int M = orbit_node_count(coef, n - 3, h);
  const size_t nnode = M ? (size_t)(n - 3) * (M + 1) * kNodeRec : 0;
  std::vector<double> nodes(nnode, 0.0);
  for (int k = 1; M && k <= n - 3; ++k)
    for (int m = 0; m <= M; ++m) {
      double *r = &nodes[((size_t)(k - 1) * (M + 1) + m) * kNodeRec];
      r[0] = orbit->t[k] + m * (h / M);
      const double s0 = (r[0] - orbit->t[k]) / h;
      for (int comp = 0; comp < 3; ++comp) {
        double c[8];
        poly_shift(&coef[(size_t)(k - 1) * 24 + comp * 8], s0, c);
        r[1 + comp] = c[0];
        r[4 + comp] = c[1] / h;
        r[7 + comp] = 2 * c[2] / (h * h);
        r[10 + comp] = 6 * c[3] / (h * h * h);
      }
    }
  gd.zd_warm_dt2 = warm_step_limit(coef, orbit->t, n, gd.tstart, tend,
                                    gd.rngstart + N * gd.dmrg);
  gd.inv_dmrg = 1.0 / gd.dmrg;
  gd.k4pi_wvl = 4.0 * kPi / gd.wvl;
  gd.inv_dtaz = b->prf;
  gd.inv_samples = 1.0 / N;
  gd.inv_lines = 1.0 / L;
  // buffer: orbit records | kt | etaref | orbit nodes | GeomDev copy
  // (256-byte aligned)
  const size_t ndev = (sizeof(GeomDev) + sizeof(double) - 1) / sizeof(double);
  const size_t off_node = ((nrec + 2 * (size_t)N) + 31) & ~(size_t)31;
  const size_t off_dev = (off_node + nnode + 31) & ~(size_t)31;
  size_t nbuf = off_dev + ndev;
  double *d = nullptr;
  S1_HIP(hipMalloc(&d, nbuf * sizeof(double)));
  gd.orb.coef = d;
  gd.orb.t0 = orbit->t[0];
  gd.orb.h = h;
  gd.orb.inv_h = 1.0 / h;
  gd.orb.n = n;
  gd.orb.node_m = M;
  gd.orb.node = M ? d + off_node : nullptr;
  gd.kt = d + nrec;
  gd.etaref = d + nrec + N;
  hipError_t e = hipMemcpy(d, recs.data(), nrec * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d + nrec, kt.data(), N * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d + nrec + N, etaref.data(), N * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess && nnode)
    e = hipMemcpy(d + off_node, nodes.data(), nnode * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(d + off_dev, &gd, sizeof(GeomDev), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return hip_error(e, "geometry upload");
  }

  s1ifg_geom *g = new s1ifg_geom();
  g->device = ctx->device;
  g->dev = gd;
  g->d_buf = d;
  g->d_dev = reinterpret_cast<const GeomDev *>(d + off_dev);
  g->n_orb = n;
  g->h_t = new double[n];
  g->h_pos = new double[3 * n];
  g->h_vel = new double[3 * n];
  std::memcpy(g->h_t, orbit->t, n * sizeof(double));
  std::memcpy(g->h_pos, orbit->pos, 3 * n * sizeof(double));
  std::memcpy(g->h_vel, orbit->vel, 3 * n * sizeof(double));
  *out = g;
  return S1IFG_OK;
}

void s1ifg_geom_destroy(s1ifg_geom *g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  (void)hipFree(g->d_buf);
  delete[] g->h_t;
  delete[] g->h_pos;
  delete[] g->h_vel;
  delete g;
}

int s1ifg_geom_query(const s1ifg_geom *g, double *out, int n) {
  if (!g || !out || n < 17) return set_error(S1IFG_EINVAL, "need a geometry and 17 outputs");
  const GeomDev &d = g->dev;
  out[0] = d.ks;
  for (int k = 0; k < 4; ++k) {
    out[1 + k] = d.fm[k];
    out[5 + k] = d.dc[k];
  }
  out[9] = d.etac0;
  out[10] = d.tmid;
  for (int k = 0; k < 3; ++k) {
    out[11 + k] = d.xmid[k];
    out[14 + k] = d.vmid[k];
  }
  return S1IFG_OK;
}

int s1ifg_geom_vectors(const s1ifg_geom *g, double *kt, double *etaref, int n) {
  if (!g || (!kt && !etaref)) return set_error(S1IFG_EINVAL, "need a geometry and an output");
  if (n != g->dev.samples)
    return set_error(S1IFG_EINVAL, "n = %d, the burst has %d samples", n, g->dev.samples);
  int dev = -1;
  S1_HIP(hipGetDevice(&dev));
  if (dev != g->device) S1_HIP(hipSetDevice(g->device));
  // the device copies the kernels read (uploaded once when the geometry was built)
  if (kt) S1_HIP(hipMemcpy(kt, g->dev.kt, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  if (etaref)
    S1_HIP(hipMemcpy(etaref, g->dev.etaref, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  return S1IFG_OK;
}

// ---------------------------------------------------------------- bounds
static double reast(double lat) { return kRa / std::sqrt(1.0 - kRe2 * std::pow(std::sin(lat), 2)); }
static double rnorth(double lat) {
  return kRa * (1.0 - kRe2) / std::pow(1.0 - kRe2 * std::pow(std::sin(lat), 2), 1.5);
}

static double ellipsoidal_hdg(const double *ll1, const double *ll2) {
  // bounds.cu:112-199 (Sodano inverse geodesic, heading at the first point)
  double e = std::sqrt(kRe2), sq = std::sqrt(1.0 - kRe2), f = 1.0 - sq;
  double ep = e * f / (kRe2 - f), nn = f / kRe2;
  double k1 = (16.0 * kRe2 * nn * nn + ep * ep) / (ep * ep);
  double k2 = (16.0 * kRe2 * nn * nn) / (16.0 * kRe2 * nn * nn + ep * ep);
  double k3 = (16.0 * kRe2 * nn * nn) / (ep * ep);
  double k4 = (16.0 * nn - ep * ep) / (16.0 * kRe2 * nn * nn + ep * ep);
  double k5 = 16.0 / (kRe2 * (16.0 * kRe2 * nn * nn + ep * ep));
  double rad = kPi / 180.0;
  double tli = std::tan(ll1[0] * rad), tlf = std::tan(ll2[0] * rad);
  double l = std::fabs(ll2[1] * rad - ll1[1] * rad), lsign = ll2[1] * rad - ll1[1] * rad;
  double sinlon = std::sin(l), coslon = std::cos(l);
  double tbi = sq * tli, tbf = sq * tlf;
  double cbi = 1.0 / std::sqrt(1.0 + tbi * tbi), cbf = 1.0 / std::sqrt(1.0 + tbf * tbf);
  double sbi = tbi * cbi, sbf = tbf * cbf;
  double ac = sbi * sbf, bc = cbi * cbf;
  double cosphi = ac + bc * coslon;
  double sgn = (sinlon > 0) - (sinlon < 0);
  double sinphi = sgn * std::sqrt(1.0 - std::fmin(cosphi * cosphi, 1.0));
  double phi = std::fabs(std::atan2(sinphi, cosphi));
  if (!(kRa * std::fabs(phi) > 1e-6)) return 0.0;
  double ca = (bc * sinlon) / sinphi, cb = ca * ca;
  double cc = (cosphi * (1.0 - cb)) / k1, cd = (-2.0 * ac) / k1, ce = -ac * k2;
  double cf = k3 * cc, cg = phi * phi / sinphi;
  double x = ((phi * (k4 + cb) + sinphi * (cc + cd) + cg * (cf + ce)) * ca) / k5;
  double lam = l + x;
  double t2 = tbf * cbi - std::cos(lam) * sbi;
  double sl = std::sin(lam) * ((lsign > 0) - (lsign < 0));
  return std::atan2(sl, t2);
}

static void cross(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static void normalize(double v[3]) {
  double nrm = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int i = 0; i < 3; ++i) v[i] /= nrm;
}

int s1ifg_geom_bounds(const s1ifg_geom *g, double hmin, double hmax, double ll[4]) {
  if (!g || !ll) return set_error(S1IFG_EINVAL, "null argument");
  // Hermite orbit on the host, stencil as orbit.cu:229-232
  const int n = g->n_orb;
  std::vector<double> coef((size_t)(n - 3) * 24);
  double h = g->h_t[1] - g->h_t[0];
  for (int k = 1; k <= n - 3; ++k) {
    double s[4], x[4][3], v[4][3];
    for (int i = 0; i < 4; ++i) {
      int j = k - 1 + i;
      s[i] = (g->h_t[j] - g->h_t[k]) / h;
      for (int c = 0; c < 3; ++c) {
        x[i][c] = g->h_pos[j * 3 + c];
        v[i][c] = g->h_vel[j * 3 + c];
      }
    }
    hermite_poly(s, x, v, h, &coef[(size_t)(k - 1) * 24]);
  }
  const GeomDev &d = g->dev;
  double t_start = d.tstart, t_end = d.tstart + (d.lines - 1) * d.dtaz;  // geo2rdr.cu:390-397
  double r_start = d.rngstart, r_end = d.rngstart + (d.samples - 1) * d.dmrg;
  double xs[3], vs_[3], xe[3], ve[3], xm[3], vm[3], a[3];
  orbit_eval_host(coef, g->h_t, n, t_start, xs, vs_, a);
  orbit_eval_host(coef, g->h_t, n, t_end, xe, ve, a);
  orbit_eval_host(coef, g->h_t, n, (t_start + t_end) / 2., xm, vm, a);
  double lls[3], lle[3], llm[3];
  xyz2llh(xs, lls);
  xyz2llh(xe, lle);
  double hdg = ellipsoidal_hdg(lls, lle);
  xyz2llh(xm, llm);
  double latm = llm[0] * kPi / 180.0;
  double re = reast(latm), rn = rnorth(latm);
  double rcurv = (re * rn) / (re * std::pow(std::cos(hdg), 2) + rn * std::pow(std::sin(hdg), 2));
  double lat_min = 1e300, lat_max = -1e300, lon_min = 1e300, lon_max = -1e300;
  for (int i = 0; i < 8; ++i) {  // bounds.cu:294-310, rah2ll 260-284
    double r = i < 4 ? r_start : r_end;
    double t = (i % 4 < 2) ? t_start : t_end;
    double hh = (i % 2 == 0) ? hmin : hmax;
    double xyz[3], vel[3];
    orbit_eval_host(coef, g->h_t, n, t, xyz, vel, a);
    double llh[3];
    xyz2llh(xyz, llh);
    double vhat[3] = {vel[0], vel[1], vel[2]};
    normalize(vhat);
    double rlat = llh[0] * kPi / 180.0, rlon = llh[1] * kPi / 180.0;
    double nhat[3] = {-std::cos(rlat) * std::cos(rlon), -std::cos(rlat) * std::sin(rlon),
                      -std::sin(rlat)};
    double chat[3], that[3];
    cross(nhat, vel, chat);  // right-looking (bounds.cu:79-80)
    normalize(chat);
    cross(chat, nhat, that);
    normalize(that);
    double aa = rcurv + llh[2], bb = rcurv + hh;
    double cost = 0.5 * ((aa / r + r / aa - bb / aa * bb / r));
    double sint = std::sqrt(1. - cost * cost);
    double gamm = cost * r;
    double vn = nhat[0] * vhat[0] + nhat[1] * vhat[1] + nhat[2] * vhat[2];
    double vt = vhat[0] * that[0] + vhat[1] * that[1] + vhat[2] * that[2];
    double alpha = (0.0 * r - gamm * vn) / vt;
    double beta = std::sqrt((r * sint) * (r * sint) - alpha * alpha);
    double tar[3], llt[3];
    for (int j = 0; j < 3; ++j) tar[j] = xyz[j] + gamm * nhat[j] + alpha * that[j] + beta * chat[j];
    xyz2llh(tar, llt);
    lat_min = std::fmin(lat_min, llt[0]);
    lat_max = std::fmax(lat_max, llt[0]);
    lon_min = std::fmin(lon_min, llt[1]);
    lon_max = std::fmax(lon_max, llt[1]);
  }
  ll[0] = lat_min;
  ll[1] = lat_max;
  ll[2] = lon_min;
  ll[3] = lon_max;
  (void)llh2xyz;
  return S1IFG_OK;
}

}  // extern "C"
