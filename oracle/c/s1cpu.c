/*
 * s1cpu.c -- C/OpenMP restatement of the reference's interferogram-formation
 * path, for the CPU timing beside the GPU and as a second CPU check of the
 * NumPy oracle.  Test infrastructure: only tests/ and bench.py's CPU
 * baseline load it (oracle/libs1cpu.so through ctypes).
 *
 * Reference code each function follows (file:line in the reference tree):
 *   llh2xyz                  orbit.cu:15-35
 *   hermite / intp_orbit     orbit.cu:170-245 (truncated stencil index)
 *   orbitrangetime           orbit.cu:247-279 (Newton from the burst mid
 *                            time, <= 51 iterations, |dt| < 5e-9 s)
 *   closest_sample, record interpolation, kt / etaref set-up
 *                            geo2rdr.cu:48-59, 465-549
 *   deramp                   geo2rdr.cu:138-158
 *   reproject                geo2rdr.cu:170-248 (bilinear, row blends stored
 *                            as float); an 8-tap Hann-windowed sinc and a
 *                            6-coefficient offset polynomial as the build's
 *                            extensions (no reference counterpart)
 *   crossmul + looks         sario.cu:199-286 (conj(ref) * sec, column then
 *                            row means)
 *   goldstein                goldstein.cu:87-408 (patch grid, 7 x 7 table
 *                            smoothing of |Z|, mean normalisation, S^alpha,
 *                            additive triangular overlap-add weights); the
 *                            FFT is a radix-2 transform written here (cuFFT
 *                            in the reference), in float as cuFFT's
 *   coherence                |sum ifg| / sqrt(sum p1 sum p2) over win x win,
 *                            clipped at the borders (win = 1: the look cell,
 *                            coherence.py:142-144)
 * Geometry is FP64, products are float, as in the reference.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define SOL 299792458.0   /* geo2rdr.cu:26-28 */
#define RA 6378137.0      /* orbit.hpp:6-8 */
#define RE2 0.00669437999015
#define PI 3.14159265358979323846

typedef struct {
  int n;
  const double *t, *pos, *vel; /* [n], [n*3], [n*3] */
} s1c_orbit;

typedef struct {
  int lines, samples, flv, llv, fvs, lvs;
  double t_start, prf, ati, slant_range_time, range_sampling_rate, wavelength,
      radar_frequency, steering_rate;
  int n_fm, n_dc;
  const double *fm, *dc; /* records (t, t0, c0, c1, c2) */
} s1c_burst;

typedef struct {
  int nrow, ncol;
  double lat0, lon0, dlat, dlon;
} s1c_grid;

typedef struct {
  const float *data;
  int nrow, ncol;
  double lat0, lon0, dlat, dlon;
} s1c_dem;

int s1c_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------ geometry */
static void llh2xyz(double lat, double lon, double h, double X[3]) {
  const double rl = lat * PI / 180, ro = lon * PI / 180;
  const double sl = sin(rl), cl = cos(rl);
  const double re = RA / sqrt(1.0 - RE2 * sl * sl);
  X[0] = (re + h) * cl * cos(ro);
  X[1] = (re + h) * cl * sin(ro);
  X[2] = (re * (1 - RE2) + h) * sl;
}

/* 4-node Hermite interpolation of position and velocity: with
 * L_i(t) = prod_{j != i} (t - t_j)/(t_i - t_j) and dl_i = sum 1/(t_i - t_j),
 * a_i = (1 - 2 (t - t_i) dl_i) L_i^2, b_i = (t - t_i) L_i^2,
 * x = sum a_i x_i + b_i v_i and v = sum a_i' x_i + b_i' v_i. */
static void hermite(const double tt[4], const double *xx[4], const double *vv[4], double t,
                    double x[3], double v[3]) {
  double a[4], b[4], a2[4], b2[4];
  for (int i = 0; i < 4; ++i) {
    double li = 1.0, dl = 0.0, hdot = 0.0;
    for (int j = 0; j < 4; ++j) {
      if (j == i) continue;
      dl += 1.0 / (tt[i] - tt[j]);
      li *= (t - tt[j]) / (tt[i] - tt[j]);
      double p = 1.0 / (tt[i] - tt[j]);
      for (int k = 0; k < 4; ++k)
        if (k != i && k != j) p *= (t - tt[k]) / (tt[i] - tt[k]);
      hdot += p;
    }
    const double l2 = li * li;
    a[i] = (1 - 2 * (t - tt[i]) * dl) * l2;
    b[i] = (t - tt[i]) * l2;
    a2[i] = -2 * dl * l2 + (1 - 2 * (t - tt[i]) * dl) * li * 2 * hdot;
    b2[i] = l2 + (t - tt[i]) * li * 2 * hdot;
  }
  for (int c = 0; c < 3; ++c) {
    x[c] = 0;
    v[c] = 0;
    for (int i = 0; i < 4; ++i) {
      x[c] += a[i] * xx[i][c] + b[i] * vv[i][c];
      v[c] += a2[i] * xx[i][c] + b2[i] * vv[i][c];
    }
  }
}

static void intp_orbit(const s1c_orbit *o, double t, double x[3], double v[3]) {
  int k = (int)((t - o->t[0]) / (o->t[1] - o->t[0])); /* truncation */
  if (k < 1) k = 1;
  if (k > o->n - 3) k = o->n - 3;
  double tt[4];
  const double *xx[4], *vv[4];
  for (int i = 0; i < 4; ++i) {
    tt[i] = o->t[k - 1 + i];
    xx[i] = o->pos + 3 * (k - 1 + i);
    vv[i] = o->vel + 3 * (k - 1 + i);
  }
  hermite(tt, xx, vv, t, x, v);
}

/* zero-Doppler Newton; returns t and dr = X - x(t) */
static double orbitrangetime(const s1c_orbit *o, const double X[3], double tmid,
                             const double xmid[3], const double vmid[3], double dr[3]) {
  double t = tmid, x[3] = {xmid[0], xmid[1], xmid[2]}, v[3] = {vmid[0], vmid[1], vmid[2]};
  for (int it = 0; it < 51; ++it) {
    const double d0 = X[0] - x[0], d1 = X[1] - x[1], d2 = X[2] - x[2];
    const double fn = d0 * v[0] + d1 * v[1] + d2 * v[2];
    const double fnp = -v[0] * v[0] - v[1] * v[1] - v[2] * v[2];
    const double tn = t - fn / fnp;
    intp_orbit(o, tn, x, v);
    const double step = fabs(tn - t);
    t = tn;
    if (step < 5.0e-9) break;
  }
  for (int c = 0; c < 3; ++c) dr[c] = X[c] - x[c];
  return t;
}

static int closest_sample(const double *rec, int n, double t0) {
  int idx = 0;
  double mind = 1e10;
  for (int i = 0; i < n; ++i) {
    const double d = fabs(rec[5 * i] - t0);
    if (mind > d) {
      mind = d;
      idx = i;
    }
  }
  return idx;
}

static void interp_record(const double *rec, int n, double tc, double out[4]) {
  const int idx = closest_sample(rec, n, tc);
  const double *r0, *r1;
  if (rec[5 * idx] < tc && idx < n - 1) {
    r0 = rec + 5 * idx;
    r1 = rec + 5 * (idx + 1);
  } else if (rec[5 * idx] > tc && idx > 0) {
    r0 = rec + 5 * (idx - 1);
    r1 = rec + 5 * idx;
  } else {
    for (int k = 0; k < 4; ++k) out[k] = rec[5 * idx + 1 + k];
    return;
  }
  const double frac = (tc - r0[0]) / (r1[0] - r0[0]);
  for (int k = 0; k < 4; ++k) out[k] = r0[1 + k] + frac * (r1[1 + k] - r0[1 + k]);
}

/* scal: [0] ks, [1..4] fm (t0, c0, c1, c2), [5..8] dc, [9] etac0 */
int s1c_deramp_setup(const s1c_orbit *o, const s1c_burst *b, double *kt, double *etaref,
                     double *scal) {
  if (o->n < 4 || b->samples < 1 || b->n_fm < 1 || b->n_dc < 1) return -1;
  const double tc = b->t_start + b->lines * b->ati / 2;
  double x[3], v[3];
  intp_orbit(o, tc, x, v);
  const double vs = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  const double ks = 2.0 * vs * b->radar_frequency / SOL * b->steering_rate * PI / 180.0;
  double fm[4], dc[4];
  interp_record(b->fm, b->n_fm, tc, fm);
  interp_record(b->dc, b->n_dc, tc, dc);
  double etac0 = 0;
  for (int i = 0; i < b->samples; ++i) {
    const double tr = b->slant_range_time + i / b->range_sampling_rate;
    double dt = tr - fm[0];
    const double ka = fm[1] + dt * (fm[2] + fm[3] * dt);
    dt = tr - dc[0];
    const double fnc = dc[1] + dt * (dc[2] + dc[3] * dt);
    kt[i] = ka * ks / (ka - ks);
    const double etac = -fnc / ka;
    if (i == 0) etac0 = etac;
    etaref[i] = etac - etac0;
  }
  scal[0] = ks;
  for (int k = 0; k < 4; ++k) {
    scal[1 + k] = fm[k];
    scal[5 + k] = dc[k];
  }
  scal[9] = etac0;
  return 0;
}

void s1c_deramp(const s1c_burst *b, const double *kt, const double *etaref, const float *in,
                float *out) {
  const int L = b->lines, N = b->samples;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < L; ++i) {
    const double eta = -L * b->ati / 2. + i * b->ati;
    for (int j = 0; j < N; ++j) {
      const double d = eta - etaref[j];
      const double ph = -PI * kt[j] * d * d;
      const double c = cos(ph), s = sin(ph);
      const size_t k = 2 * ((size_t)i * N + j);
      const double ax = in[k], ay = in[k + 1];
      out[k] = (float)(ax * c - ay * s);
      out[k + 1] = (float)(ax * s + ay * c);
    }
  }
}

/* Hann-windowed sinc, taps floor(x) - 3 .. floor(x) + 4, unit sum */
static void sinc_taps(double f, double w[8]) {
  double s = 0;
  for (int k = 0; k < 8; ++k) {
    const double x = f - (k - 3);
    const double sc = fabs(x) < 1e-300 ? 1.0 : sin(PI * x) / (PI * x);
    const double c = cos(PI * x / 8.0);
    w[k] = sc * c * c;
    s += w[k];
  }
  for (int k = 0; k < 8; ++k) w[k] /= s;
}

static double dem_height(const s1c_dem *d, double lat, double lon) {
  double fy = (lat - d->lat0) / d->dlat, fx = (lon - d->lon0) / d->dlon;
  fy = fmin(fmax(fy, 0.0), d->nrow - 1.0);
  fx = fmin(fmax(fx, 0.0), d->ncol - 1.0);
  int iy = (int)floor(fy), ix = (int)floor(fx);
  if (iy > d->nrow - 2) iy = d->nrow - 2;
  if (ix > d->ncol - 2) ix = d->ncol - 2;
  const double wy = fy - iy, wx = fx - ix;
  const float *p = d->data + (size_t)iy * d->ncol + ix;
  const double h00 = p[0], h01 = p[1], h10 = p[d->ncol], h11 = p[d->ncol + 1];
  return (h00 * (1 - wx) + h01 * wx) * (1 - wy) + (h10 * (1 - wx) + h11 * wx) * wy;
}

/* z: the deramped burst; scal from s1c_deramp_setup; out: target grid */
void s1c_reproject(const s1c_orbit *o, const s1c_burst *b, const double *scal,
                   const s1c_grid *g, const s1c_dem *dem, const double *offpoly, const float *z,
                   int interp, float *out) {
  const int L = b->lines, N = b->samples;
  const double dtaz = 1.0 / b->prf;
  const double tmid = (b->t_start + (b->t_start + (L - 1) * dtaz)) * 0.5;
  double xmid[3], vmid[3];
  intp_orbit(o, tmid, xmid, vmid);
  const double rngstart = b->slant_range_time * SOL / 2.0;
  const double dmrg = SOL / 2.0 / b->range_sampling_rate;
  const double ks = scal[0], etac0 = scal[9];
  const double *fm = scal + 1, *dc = scal + 5;
#pragma omp parallel for schedule(dynamic, 4)
  for (int r = 0; r < g->nrow; ++r) {
    const double lat = g->lat0 + r * g->dlat;
    for (int c = 0; c < g->ncol; ++c) {
      const double lon = g->lon0 + c * g->dlon;
      float *po = out + 2 * ((size_t)r * g->ncol + c);
      po[0] = po[1] = 0.f;
      double X[3], dr[3];
      llh2xyz(lat, lon, dem_height(dem, lat, lon), X);
      const double t = orbitrangetime(o, X, tmid, xmid, vmid, dr);
      const double rng = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      const double rgoff0 = (rng - rngstart) / dmrg, azoff0 = (t - b->t_start) / dtaz;
      double drg = 0, daz = 0;
      if (offpoly) {
        const double xs = rgoff0 / N, ys = azoff0 / L;
        drg = offpoly[0] + offpoly[1] * xs + offpoly[2] * ys;
        daz = offpoly[3] + offpoly[4] * xs + offpoly[5] * ys;
      }
      const double rg = rgoff0 + drg, az = azoff0 + daz;
      int valid;
      if (interp == 0)
        valid = rg > b->fvs && rg < b->lvs - 1 && az >= b->flv + 5 && az < b->llv - 5;
      else
        valid = rg >= b->fvs + 3 && rg < b->lvs - 4 && az >= b->flv + 5 && az < b->llv - 5;
      if (!valid) continue;
      const double rp = rng + drg * dmrg;
      double dt = rp * 2 / SOL - fm[0];
      const double ka = fm[1] + dt * (fm[2] + fm[3] * dt);
      dt = rp * 2 / SOL - dc[0];
      const double fnc = dc[1] + dt * (dc[2] + dc[3] * dt);
      const double kt = ka * ks / (ka - ks);
      const double etaref = -fnc / ka - etac0;
      const double eta = -L * dtaz / 2. + t - b->t_start + daz * dtaz;
      const double reramp = -PI * kt * (eta - etaref) * (eta - etaref);
      double rx, ry;
      if (interp == 0) {
        const int ir = (int)rg, ia = (int)az;
        const double fr = rg - ir, fa = az - ia;
        const float *p = z + 2 * ((size_t)ia * N + ir);
        const float *q = p + 2 * (size_t)N;
        const float c1x = (float)(p[0] * (1 - fr) + p[2] * fr);
        const float c1y = (float)(p[1] * (1 - fr) + p[3] * fr);
        const float c2x = (float)(q[0] * (1 - fr) + q[2] * fr);
        const float c2y = (float)(q[1] * (1 - fr) + q[3] * fr);
        rx = c1x * (1 - fa) + c2x * fa;
        ry = c1y * (1 - fa) + c2y * fa;
      } else {
        const int ir = (int)floor(rg), ia = (int)floor(az);
        double wr[8], wa[8];
        sinc_taps(rg - ir, wr);
        sinc_taps(az - ia, wa);
        rx = ry = 0;
        for (int ka_ = 0; ka_ < 8; ++ka_) {
          const float *row = z + 2 * ((size_t)(ia - 3 + ka_) * N + (ir - 3));
          double ax = 0, ay = 0;
          for (int kr = 0; kr < 8; ++kr) {
            ax += wr[kr] * row[2 * kr];
            ay += wr[kr] * row[2 * kr + 1];
          }
          rx += wa[ka_] * ax;
          ry += wa[ka_] * ay;
        }
      }
      const double ph = 4.0 * PI / b->wavelength * rng - reramp;
      const double cp = cos(ph), sp = sin(ph);
      po[0] = (float)(rx * cp - ry * sp);
      po[1] = (float)(rx * sp + ry * cp);
    }
  }
}

/* --------------------------------------------------------- pair products */
void s1c_crossmul_looks(const float *a, const float *bb, int nrow, int ncol, int rl, int cl,
                        float *ifg, float *p1, float *p2) {
  const int nr = nrow / rl, nc = ncol / cl;
#pragma omp parallel for schedule(static)
  for (int rm = 0; rm < nr; ++rm)
    for (int cm = 0; cm < nc; ++cm) {
      double sx = 0, sy = 0, s1 = 0, s2 = 0;
      for (int i = 0; i < rl; ++i) {
        double cx = 0, cy = 0, c1 = 0, c2 = 0;
        for (int j = 0; j < cl; ++j) {
          const size_t k = 2 * ((size_t)(rm * rl + i) * ncol + cm * cl + j);
          const double ar = a[k], ai = a[k + 1], br = bb[k], bi = bb[k + 1];
          cx += ar * br + ai * bi;
          cy += ar * bi - ai * br;
          c1 += ar * ar + ai * ai;
          c2 += br * br + bi * bi;
        }
        sx += cx / cl;
        sy += cy / cl;
        s1 += c1 / cl;
        s2 += c2 / cl;
      }
      const size_t o = (size_t)rm * nc + cm;
      ifg[2 * o] = (float)(sx / rl);
      ifg[2 * o + 1] = (float)(sy / rl);
      if (p1) p1[o] = (float)(s1 / rl);
      if (p2) p2[o] = (float)(s2 / rl);
    }
}

/* ------------------------------------------------------------- Goldstein */
#define GN 32
static const float kGauss7[49] = { /* goldstein.cu:43-52 (project data) */
    0.000004f, 0.000088f, 0.000543f, 0.000948f, 0.000543f, 0.000088f, 0.000004f,
    0.000088f, 0.001915f, 0.011776f, 0.020584f, 0.011776f, 0.001915f, 0.000088f,
    0.000543f, 0.011776f, 0.072410f, 0.126584f, 0.072410f, 0.011776f, 0.000543f,
    0.000948f, 0.020584f, 0.126584f, 0.221141f, 0.126584f, 0.020584f, 0.000948f,
    0.000543f, 0.011776f, 0.072410f, 0.126584f, 0.072410f, 0.011776f, 0.000543f,
    0.000088f, 0.001915f, 0.011776f, 0.020584f, 0.011776f, 0.001915f, 0.000088f,
    0.000004f, 0.000088f, 0.000543f, 0.000948f, 0.000543f, 0.000088f, 0.000004f};

/* In-place 32-point DFT, sign -1 (forward) or +1, unnormalised: radix-2
 * decimation in time over bit-reversed input (Cooley-Tukey). */
static void fft32(float *re, float *im, int stride, int sign, const float *cw, const float *sw) {
  static const int rev[GN] = {0, 16, 8, 24, 4, 20, 12, 28, 2, 18, 10, 26, 6, 22, 14, 30,
                              1, 17, 9, 25, 5, 21, 13, 29, 3, 19, 11, 27, 7, 23, 15, 31};
  float xr[GN], xi[GN];
  for (int k = 0; k < GN; ++k) {
    xr[k] = re[rev[k] * stride];
    xi[k] = im[rev[k] * stride];
  }
  for (int len = 2; len <= GN; len <<= 1) {
    const int half = len >> 1, tstep = GN / len;
    for (int s = 0; s < GN; s += len)
      for (int k = 0; k < half; ++k) {
        const float c = cw[k * tstep], sn = sign * sw[k * tstep]; /* e^{sign i 2 pi k / len} */
        const int a = s + k, bq = a + half;
        const float tr = xr[bq] * c - xi[bq] * sn, ti = xr[bq] * sn + xi[bq] * c;
        xr[bq] = xr[a] - tr;
        xi[bq] = xi[a] - ti;
        xr[a] += tr;
        xi[a] += ti;
      }
  }
  for (int k = 0; k < GN; ++k) {
    re[k * stride] = xr[k];
    im[k * stride] = xi[k];
  }
}

static void fft2(float *re, float *im, int sign, const float *cw, const float *sw) {
  for (int r = 0; r < GN; ++r) fft32(re + r * GN, im + r * GN, 1, sign, cw, sw);
  for (int c = 0; c < GN; ++c) fft32(re + c, im + c, GN, sign, cw, sw);
}

void s1c_goldstein(const float *in, int nrow, int ncol, float alpha, float *out) {
  const int inc = GN / 4;
  int pi_n = (int)ceilf((float)nrow / inc) - 3, pj_n = (int)ceilf((float)ncol / inc) - 3;
  if (pi_n < 1) pi_n = 1;
  if (pj_n < 1) pj_n = 1;
  float cw[GN], sw[GN];
  for (int k = 0; k < GN; ++k) {
    cw[k] = (float)cos(2 * PI * k / GN);
    sw[k] = (float)sin(2 * PI * k / GN);
  }
  double *acc = calloc((size_t)nrow * ncol * 2, sizeof(double));
  double *wsum = calloc((size_t)nrow * ncol, sizeof(double));
  if (!acc || !wsum) {
    free(acc);
    free(wsum);
    return;
  }
  float w1[GN];
  for (int t = 0; t < GN; ++t) w1[t] = (float)(t < GN / 2 ? t : GN - 1 - t);
  /* patch rows p and p + 4 do not overlap, except that the last one may be
   * clamped back onto its predecessors (goldstein.cu:308-313): four phases,
   * each parallel over its unclamped patch rows, then a clamped last row on
   * its own */
  for (int phase = 0; phase < 5; ++phase) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int pi = phase < 4 ? phase : 0; pi < pi_n; pi += phase < 4 ? 4 : 1) {
      const int clamped = pi * inc > nrow - GN;
      if (clamped != (phase == 4)) continue;
      const int oy = pi * inc < nrow - GN ? pi * inc : nrow - GN;
      float re[GN * GN], im[GN * GN], m[GN * GN];
      for (int pj = 0; pj < pj_n; ++pj) {
        const int ox = pj * inc < ncol - GN ? pj * inc : ncol - GN;
        for (int y = 0; y < GN; ++y)
          for (int x = 0; x < GN; ++x) {
            const size_t k = 2 * ((size_t)(oy + y) * ncol + ox + x);
            re[y * GN + x] = in[k];
            im[y * GN + x] = in[k + 1];
          }
        fft2(re, im, -1, cw, sw);
        for (int k = 0; k < GN * GN; ++k) m[k] = sqrtf(re[k] * re[k] + im[k] * im[k]);
        double sum = 0;
        float s[GN * GN];
        for (int y = 0; y < GN; ++y)
          for (int x = 0; x < GN; ++x) {
            float v = 0;
            for (int ky = -3; ky <= 3; ++ky)
              for (int kx = -3; kx <= 3; ++kx)
                v += kGauss7[(ky + 3) * 7 + kx + 3] * m[((y + ky) & 31) * GN + ((x + kx) & 31)];
            s[y * GN + x] = v;
            sum += v;
          }
        const float mean = (float)(sum / (GN * GN));
        for (int k = 0; k < GN * GN; ++k) {
          const float sn = mean > 1e-8f ? s[k] / mean : 0.f;
          const float w = powf(sn, alpha);
          if (w > 0.f) {
            re[k] *= w;
            im[k] *= w;
          }
        }
        fft2(re, im, +1, cw, sw);
        for (int y = 0; y < GN; ++y)
          for (int x = 0; x < GN; ++x) {
            const double w = (double)w1[x] + w1[y];
            const size_t o = (size_t)(oy + y) * ncol + ox + x;
            acc[2 * o] += re[y * GN + x] * w / (GN * GN);
            acc[2 * o + 1] += im[y * GN + x] * w / (GN * GN);
            wsum[o] += w;
          }
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (long o = 0; o < (long)nrow * ncol; ++o) {
    const int ok = wsum[o] > 0;
    out[2 * o] = ok ? (float)(acc[2 * o] / wsum[o]) : 0.f;
    out[2 * o + 1] = ok ? (float)(acc[2 * o + 1] / wsum[o]) : 0.f;
  }
  free(acc);
  free(wsum);
}

/* ------------------------------------------------------------- coherence */
void s1c_coherence(const float *ifg, const float *p1, const float *p2, int nrow, int ncol,
                   int win, float *coh) {
  const int h = win / 2;
#pragma omp parallel for schedule(static)
  for (int r = 0; r < nrow; ++r)
    for (int c = 0; c < ncol; ++c) {
      double sx = 0, sy = 0, a = 0, b = 0;
      for (int y = r - h; y <= r + h; ++y) {
        if (y < 0 || y >= nrow) continue;
        for (int x = c - h; x <= c + h; ++x) {
          if (x < 0 || x >= ncol) continue;
          const size_t k = (size_t)y * ncol + x;
          sx += ifg[2 * k];
          sy += ifg[2 * k + 1];
          a += p1[k];
          b += p2[k];
        }
      }
      const double den = sqrt(a * b);
      coh[(size_t)r * ncol + c] = den > 0 ? (float)(sqrt(sx * sx + sy * sy) / den) : 0.f;
    }
}
