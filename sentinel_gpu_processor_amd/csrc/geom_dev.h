// Device functions of the geometry path (shared by the per-stage and fused
// kernels so that both produce the same bits for the same pixel).
#pragma once

#include "s1ifg_internal.h"
#include "wave_dev.h"

namespace s1ifg {

// Per-stencil record in the orbit buffer: [t_k, coef[24], pad...].
constexpr int kOrbRec = 32;

// Horner evaluation of one stencil record rec = [t_k, coef[24]] (P is a
// pointer to const double in any address space).  p'' and p' start at 0
// and p' picks up c[7] first: their first steps fma(0, s, x) = x are
// taken directly (same values).
template <typename P>
__device__ __forceinline__ void orbit_horner(P rec, double t, double inv_h, double x[3],
                                             double v[3], double a[3]) {
  const double s = (t - rec[0]) * inv_h;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    P c = rec + 1 + comp * 8;
    // m = 6: ddp = 0, dp = c7, p = c7 s + c6;  m = 5: ddp = 2 c7
    double dp = c[7];
    double p = fma(c[7], s, c[6]);
    double ddp = 2.0 * dp;
    dp = fma(dp, s, p);
    p = fma(p, s, c[5]);
#pragma unroll
    for (int m = 4; m >= 0; --m) {
      ddp = fma(ddp, s, 2.0 * dp);
      dp = fma(dp, s, p);
      p = fma(p, s, c[m]);
    }
    x[comp] = p;
    v[comp] = dp * inv_h;
    a[comp] = ddp * inv_h * inv_h;
  }
}

// Position, velocity and acceleration from the polynomial form of the
// 4-node Hermite interpolation (orbit.cu:170-245).  Stencil index: truncated
// (t - t0)/(t1 - t0) clamped to [1, n-3] (orbit.cu:229-232).  When the
// wavefront's active lanes share the stencil (nearly always: a wavefront
// spans milliseconds of azimuth time, the stencils seconds) the record is
// read with scalar loads through the constant address space; otherwise per
// lane.  Both read the same record, so the values are the same.
__device__ __forceinline__ void orbit_eval(const OrbitDev &o, double t, double x[3],
                                           double v[3], double a[3]) {
  int k = (int)((t - o.t0) * o.inv_h);
  k = k > 1 ? k : 1;
  k = k < o.n - 3 ? k : o.n - 3;
  const int k0 = __builtin_amdgcn_readfirstlane(k);
  if (__all(k == k0)) {
    typedef const __attribute__((address_space(4))) double *CP;
    orbit_horner((CP)o.coef + (size_t)(k0 - 1) * kOrbRec, t, o.inv_h, x, v, a);
  } else {
    orbit_horner(o.coef + (size_t)(k - 1) * kOrbRec, t, o.inv_h, x, v, a);
  }
}

// Orbit state at t from the node record rec (any address space): the
// cubic Taylor step of the node's stencil polynomial.
template <typename P>
__device__ __forceinline__ void orbit_node_step(P rec, double t, double x[3], double v[3],
                                                double a[3]) {
  const double e = t - rec[0];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double xc = rec[1 + c], vc = rec[4 + c], ac = rec[7 + c], jc = rec[10 + c];
    x[c] = fma(e, fma(e, fma(e, jc * (1.0 / 6.0), 0.5 * ac), vc), xc);
    v[c] = fma(e, fma(e, 0.5 * jc, ac), vc);
    a[c] = fma(e, jc, ac);
  }
}

// orbit_eval from the orbit node table: the stencil index as orbit_eval,
// then the nearest of the node_m + 1 nodes of that stencil interval and a
// cubic Taylor step, within 1e-10 m and 1e-10 m/s of orbit_eval (the node
// count is chosen on that bound when the geometry is built).  Scalar loads
// when the wavefront shares the node.  Without a node table, or for a time
// outside the interior stencil intervals, orbit_eval.
__device__ __forceinline__ void orbit_eval_fast(const OrbitDev &o, double t, double x[3],
                                                double v[3], double a[3]) {
  const double u = (t - o.t0) * o.inv_h;
  int k = (int)u;
  k = k > 1 ? k : 1;
  k = k < o.n - 3 ? k : o.n - 3;
  const double f = u - k;  // position in stencil interval k
  if (!__all(o.node_m > 0 && f >= 0.0 && f <= 1.0)) {
    orbit_eval(o, t, x, v, a);
    return;
  }
  const int idx = (k - 1) * (o.node_m + 1) + (int)rint(f * o.node_m);
  const int i0 = __builtin_amdgcn_readfirstlane(idx);
  if (__all(idx == i0)) {
    typedef const __attribute__((address_space(4))) double *CP;
    orbit_node_step((CP)o.node + (size_t)i0 * kNodeRec, t, x, v, a);
  } else {
    orbit_node_step(o.node + (size_t)idx * kNodeRec, t, x, v, a);
  }
}

// 1/x in FP64 from the float reciprocal and two Newton steps (relative
// error ~1e-16 for |x| within float range, x != 0).
__device__ __forceinline__ double rcp64(double x) {
  double r = (double)__builtin_amdgcn_rcpf((float)x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// One Newton iteration of zero_doppler from the orbit state (x, v, a):
// d = X - x and the time step dt.
__device__ __forceinline__ void zd_step(const double X[3], const double x[3], const double v[3],
                                        const double a[3], double d[3], double &dt) {
  d[0] = X[0] - x[0];
  d[1] = X[1] - x[1];
  d[2] = X[2] - x[2];
  double fn = d[0] * v[0] + d[1] * v[1] + d[2] * v[2];
  double fp = -(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + (d[0] * a[0] + d[1] * a[1] + d[2] * a[2]);
  dt = -fn * rcp64(fp);
}

// dr = X - x(t + dt) after the converged step, by the second-order Taylor
// step of the orbit.
__device__ __forceinline__ void zd_finish(const double d[3], const double v[3], const double a[3],
                                          double dt, double dr[3]) {
  double h2 = 0.5 * dt * dt;
  dr[0] = d[0] - v[0] * dt - a[0] * h2;
  dr[1] = d[1] - v[1] * dt - a[1] * h2;
  dr[2] = d[2] - v[2] * dt - a[2] * h2;
}

// Zero-Doppler azimuth time of X (orbitrangetime, orbit.cu:247-279).  The
// iteration starts from the burst mid-time state as the reference does
// (geo2rdr.cu:588-600).  The Newton derivative includes the dr.a term, so
// the iteration converges quadratically to the same root; after the step
// that meets the 5e-9 s criterion the orbit is advanced by a second-order
// Taylor step instead of a further interpolation.  Returns dr = X - x(t)
// and, when asked, the orbit velocity and acceleration at t.
__device__ __forceinline__ int zero_doppler(const GeomDev &g, const double X[3], double &t,
                                            double dr[3], double *vout = nullptr,
                                            double *aout = nullptr) {
  double x[3] = {g.xmid[0], g.xmid[1], g.xmid[2]};
  double v[3] = {g.vmid[0], g.vmid[1], g.vmid[2]};
  double a[3] = {g.amid[0], g.amid[1], g.amid[2]};
  t = g.tmid;
  for (int it = 0; it < 51; ++it) {
    double d[3], dt;
    zd_step(X, x, v, a, d, dt);
    t += dt;
    if (fabs(dt) < 5.0e-9) {
      zd_finish(d, v, a, dt, dr);
      if (vout) {
        for (int k = 0; k < 3; ++k) {
          vout[k] = v[k] + a[k] * dt;
          aout[k] = a[k];
        }
      }
      return it + 1;
    }
    orbit_eval_fast(g.orb, t, x, v, a);
  }
  dr[0] = X[0] - x[0];
  dr[1] = X[1] - x[1];
  dr[2] = X[2] - x[2];
  if (vout) {
    for (int k = 0; k < 3; ++k) {
      vout[k] = v[k];
      aout[k] = a[k];
    }
  }
  return 52;
}

// Orbit state at a converged zero-Doppler time: t, x(t), v(t), a(t).
struct ZdState {
  double t, x[3], v[3], a[3];
};

// zero_doppler that also returns the orbit state at the root (for
// zero_doppler_near on nearby points).
__device__ __forceinline__ void zero_doppler_state(const GeomDev &g, const double X[3], double &t,
                                                   double dr[3], ZdState &s) {
  zero_doppler(g, X, t, dr, s.v, s.a);
  s.t = t;
  s.x[0] = X[0] - dr[0];
  s.x[1] = X[1] - dr[1];
  s.x[2] = X[2] - dr[2];
}

// Zero-Doppler time of X from the root state s of a nearby point (a pixel
// of the same look cell): one Newton step from s.t with the exact
// derivative, the orbit advanced by the second-order Taylor step of
// zero_doppler.  Newton's error after the step is f''/(2 f') dt^2; the step
// is accepted when dt^2 <= g.zd_warm_dt2, the bound (set up with the orbit)
// under which that error is below 5e-10 s, a tenth of the reference's
// convergence tolerance.  Longer steps take the full iteration.
__device__ __forceinline__ void zero_doppler_near(const GeomDev &g, const double X[3],
                                                  const ZdState &s, double &t, double dr[3]) {
  double d[3], dt;
  zd_step(X, s.x, s.v, s.a, d, dt);
  if (dt * dt <= g.zd_warm_dt2) {
    t = s.t + dt;
    zd_finish(d, s.v, s.a, dt, dr);
  } else {
    zero_doppler(g, X, t, dr);
  }
}

// Row / column tables of the target grid (FP64): see k_grid_tables.
struct RowTab { double coslat, sinlat, re, fy; };
struct ColTab { double coslon, sinlon, fx; };

// Bilinear DEM height with indices clamped to the DEM.  The sample type is
// tested once, so the four loads are issued together.
__device__ __forceinline__ double dem_height(const DemDev &d, double fy, double fx) {
  fy = fmin(fmax(fy, 0.0), (double)(d.nrow - 1));
  fx = fmin(fmax(fx, 0.0), (double)(d.ncol - 1));
  int iy = min((int)floor(fy), d.nrow - 2);
  int ix = min((int)floor(fx), d.ncol - 2);
  double wy = fy - iy, wx = fx - ix;
  const size_t o0 = (size_t)iy * d.ncol + ix, o1 = o0 + d.ncol;
  double h00, h01, h10, h11;
  if (d.dtype == S1IFG_DEM_I16) {
    const short *p = (const short *)d.data;
    const short a = p[o0], b = p[o0 + 1], c = p[o1], e = p[o1 + 1];
    h00 = a;
    h01 = b;
    h10 = c;
    h11 = e;
  } else {
    const float *p = (const float *)d.data;
    const float a = p[o0], b = p[o0 + 1], c = p[o1], e = p[o1 + 1];
    h00 = a;
    h01 = b;
    h10 = c;
    h11 = e;
  }
  return (h00 * (1 - wx) + h01 * wx) * (1 - wy) + (h10 * (1 - wx) + h11 * wx) * wy;
}

// llh2xyz (orbit.cu:15-35) with the per-row / per-column trigonometry
// precomputed.
__device__ __forceinline__ void target_xyz(const RowTab &r, const ColTab &c, double h,
                                           double X[3]) {
  double rh = (r.re + h) * r.coslat;
  X[0] = rh * c.coslon;
  X[1] = rh * c.sinlon;
  X[2] = (r.re * (1 - kRe2) + h) * r.sinlat;
}

// Reduce a phase to [-pi, pi] in FP64 and return it as float.
__device__ __forceinline__ float reduce_phase(double ph) {
  double k = rint(ph * (1.0 / kTwoPi));
  return (float)fma(-k, kTwoPi, ph);
}

// sin and cos of a phase: reduced in FP64 to [-1/2, 1/2] revolutions
// (|ph| ~ 2e8 rad keeps 2e-8 rad of resolution), then the hardware sine and
// cosine, which take revolutions (absolute error ~1e-6, far below the
// resampling error; the library sincosf would redo the range reduction).
__device__ __forceinline__ void sincos_phase(double ph, float &s, float &c) {
  const double rv = ph * (1.0 / kTwoPi);
  const float r = (float)(rv - rint(rv));
  s = __builtin_amdgcn_sinf(r);
  c = __builtin_amdgcn_cosf(r);
}

// The deramp as the fused kernels apply it to the samples they read
// (deramp_burst's phase, geo2rdr.cu:138-158): z * exp(i phi),
// phi = -pi kt (eta - etaref)^2, evaluated in FP64 in revolutions
// (-kt (eta - etaref)^2 / 2, with kth = -kt / 2 given), reduced to
// [-1/2, 1/2] and rotated with the hardware sine and cosine (as the pixel
// phase, sincos_phase).
__device__ __forceinline__ float2 deramp_rot(double eta, double kth, double etaref, float2 a) {
  const double d = eta - etaref;
  const double rv = kth * d * d;
  const float r = (float)(rv - rint(rv));
  const float s = __builtin_amdgcn_sinf(r), c = __builtin_amdgcn_cosf(r);
  return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
}

// exp(2 pi i rv) of a phase rv in revolutions (FP64), reduced to [-1/2, 1/2]
// before the hardware cosine and sine.
__device__ __forceinline__ float2 phasor_rev(double rv) {
  const float r = (float)(rv - rint(rv));
  return make_float2(__builtin_amdgcn_cosf(r), __builtin_amdgcn_sinf(r));
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// The deramp phasors of one column (kth = -kt / 2, etaref) at the azimuth
// times eta0 + k h, k = 0, 1, ...: z_k = exp(2 pi i kth (eta0 + k h -
// etaref)^2) by a second-order recurrence, z_{k+1} = z_k u_k, u_{k+1} = u_k q,
// with z_0, u_0 = exp(2 pi i kth h (2 d + h)) (d = eta0 - etaref) and q =
// exp(2 pi i 2 kth h^2) from FP64 phases.  Each step rounds two complex
// products in FP32: after k steps the phasor is within ~k 1e-7 of
// deramp_rot's (a few 1e-6 over a staged column), for 2 complex products
// per sample instead of an FP64 phase and a sine and cosine.
struct DerampSeq {
  float2 z, u, q;
};

__device__ __forceinline__ DerampSeq deramp_seq(double eta0, double h, double kth,
                                                double etaref) {
  const double d = eta0 - etaref;
  DerampSeq s;
  s.z = phasor_rev(kth * d * d);
  s.u = phasor_rev(kth * h * (2.0 * d + h));
  s.q = phasor_rev(2.0 * kth * h * h);
  return s;
}

// The next sample of the column deramped: a z_k, then the step to k + 1.
__device__ __forceinline__ float2 deramp_next(DerampSeq &s, float2 a) {
  const float2 o = cmulf(a, s.z);
  s.z = cmulf(s.z, s.u);
  s.u = cmulf(s.u, s.q);
  return o;
}

// Sample (row, col) of a raw burst, deramped (deramp_rot).
__device__ __forceinline__ float2 raw_tap(const GeomDev &g, const float2 *__restrict__ z, int row,
                                          int col) {
  const double eta = -g.lines * g.ati / 2. + row * g.ati;
  return deramp_rot(eta, -0.5 * g.kt[col], g.etaref[col], z[(size_t)row * g.samples + col]);
}

// 8-tap Hann-windowed sinc weights for fractional offset f in [0,1):
// w_k ~ sinc(f - k) * cos^2(pi (f - k)/8), k = -3..4, normalised to sum 1.
// With sin(pi (f - k)) = (-1)^k sin(pi f), sinc(f - k) is
// (-1)^k sin(pi f) / (pi P) * prod_{j != k} (f - j), P = prod_j (f - j); the
// factor sin(pi f) / (pi P) is common to all taps and cancels in the
// normalisation, and so does a constant factor of the window
// (2 cos^2(x/2) = 1 + cos x).  So w_k = q_k / sum_j q_j with
// q_k = (-1)^k (1 + cos(pi (f - k)/4)) prod_{j != k} (f - j) (prefix and
// suffix products; f = 0 gives q = delta_k0 times a constant).  Writes the
// q_k and returns 1 / sum_j q_j: the caller applies the normalisation once
// to its result.
__device__ __forceinline__ float sinc_weights(float f, float q[8]) {
  // hardware sin/cos take revolutions: sin(pi f / 4) = sin(2 pi (f/8))
  const float s4 = __builtin_amdgcn_sinf(0.125f * f);
  const float c4 = __builtin_amdgcn_cosf(0.125f * f);
  // cos(pi (f - k)/4) = c4 cos(pi k/4) + s4 sin(pi k/4), k = -3..4:
  // -u, -s4, v, c4, u, s4, -v, -c4 with u, v = (c4 +- s4) / sqrt(2)
  const float u = 0.70710678f * (c4 + s4), v = 0.70710678f * (c4 - s4);
  // (-1)^k (1 + cos(pi (f - k)/4))
  const float win[8] = {u - 1.0f,  1.0f - s4, -1.0f - v, 1.0f + c4,
                        -1.0f - u, 1.0f + s4, v - 1.0f,  1.0f - c4};
  float d[8], pre[8], suf[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) d[i] = f - (float)(i - 3);
  pre[0] = 1.0f;
#pragma unroll
  for (int i = 1; i < 8; ++i) pre[i] = pre[i - 1] * d[i - 1];
  suf[7] = 1.0f;
#pragma unroll
  for (int i = 6; i >= 0; --i) suf[i] = suf[i + 1] * d[i + 1];
  float sum = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    q[i] = pre[i] * suf[i] * win[i];
    sum += q[i];
  }
  return __builtin_amdgcn_rcpf(sum);
}

// Sampling position in burst g of the target at zero-Doppler time t and
// offset dr from the orbit (reproject, geo2rdr.cu:200-207): the geometric
// range rng and the range / azimuth sampling positions rg / az with the
// offset polynomial shift (drg, daz) applied.
__device__ __forceinline__ void sample_position(const GeomDev &g, const OffPoly &op, double t,
                                                const double dr[3], double &rng, double &rg,
                                                double &az, double &drg, double &daz) {
  rng = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
  double rgoff = (rng - g.rngstart) * g.inv_dmrg;
  double azoff = (t - g.tstart) * g.inv_dtaz;
  drg = 0.0;
  daz = 0.0;
  if (op.on) {
    double xs = rgoff * g.inv_samples, ys = azoff * g.inv_lines;
    drg = op.p[0] + op.p[1] * xs + op.p[2] * ys;
    daz = op.p[3] + op.p[4] * xs + op.p[5] * ys;
  }
  rg = rgoff + drg;
  az = azoff + daz;
}

// Sampling position of X in burst g (reproject, geo2rdr.cu:194-207): the
// zero-Doppler time t, then sample_position.  Returns the Newton iteration
// count (52: not converged).
__device__ __forceinline__ int geo_position(const GeomDev &g, const OffPoly &op,
                                            const double X[3], double &t, double &rng,
                                            double &rg, double &az, double &drg, double &daz) {
  double dr[3];
  const int it = zero_doppler(g, X, t, dr);
  sample_position(g, op, t, dr, rng, rg, az, drg, daz);
  return it;
}

// Validity window of the interpolation (geo2rdr.cu:208-209 for bilinear;
// the 8-tap window for sinc).
template <int INTERP>
__device__ __forceinline__ bool sample_valid(const GeomDev &g, double rg, double az) {
  if (INTERP == S1IFG_INTERP_BILINEAR)
    return rg > g.fvs && rg < g.lvs - 1 && az >= g.flv + 5 && az < g.llv - 5;
  return rg >= g.fvs + 3 && rg < g.lvs - 4 && az >= g.flv + 5 && az < g.llv - 5;
}

// sin / cos of the flattening + reramp phase (geo2rdr.cu:210-219) at the
// sampling position rg (r0 = floor(rg), a valid position: r0 + 1 < samples).
// The reference evaluates kt and etaref at the range time of the shifted
// range rp = rng + drg dmrg, i.e. slant_range_time + rg / fs: exactly the
// range times of the burst's per-sample kt / etaref tables (deramp set-up,
// geo2rdr.cu:533-542) at fractional index rg.  Both are smooth in range
// (polynomials of degree 2 through a ratio), so their linear interpolation
// between samples r0 and r0 + 1 is exact to ~1e-12 relative, far below the
// FP64 phase resolution that matters (1e-8 rad), and replaces two
// polynomials and two FP64 reciprocals per pixel.
__device__ __forceinline__ void reramp_phase(const GeomDev &g, double rng, double rg, int r0,
                                             double daz, double t, float &sp, float &cp) {
  const double f = rg - r0;
  const double k0 = g.kt[r0], k1 = g.kt[r0 + 1], e0 = g.etaref[r0], e1 = g.etaref[r0 + 1];
  const double kt = fma(f, k1 - k0, k0), etaref = fma(f, e1 - e0, e0);
  double eta = -g.lines * g.dtaz / 2. + (t - g.tstart) + daz * g.dtaz;
  double etadiff = eta - etaref;
  double reramp = -kPi * kt * etadiff * etadiff;
  sincos_phase(g.k4pi_wvl * rng - reramp, sp, cp);
}

// Geometry and phase of one target pixel of one burst (reproject,
// geo2rdr.cu:194-219): the sampling position rg/az (offset polynomial
// applied), the geometric range, and sin/cos of the flattening + reramp
// phase.  Returns false outside the validity window (the phase is then not
// computed).
template <int INTERP>
__device__ __forceinline__ bool geocode_geom(const GeomDev &g, const OffPoly &op,
                                             const double X[3], double &rg, double &az,
                                             double &rng, float &sp, float &cp) {
  double t, drg, daz;
  geo_position(g, op, X, t, rng, rg, az, drg, daz);
  if (!sample_valid<INTERP>(g, rg, az)) return false;
  reramp_phase(g, rng, rg, (int)floor(rg), daz, t, sp, cp);
  return true;
}

// 8 x 8 sinc interpolation of the deramped burst at fractional offsets
// (fr, fa) from the window whose top-left tap is base (row stride ld),
// rotated by (cp, sp).  The taps are weighted with the unnormalised q of
// sinc_weights; both normalisations scale the rotation.
__device__ __forceinline__ float2 sinc_interp(const float2 *__restrict__ base, long ld, float fr,
                                              float fa, float sp, float cp) {
  float wr[8], wa[8];
  const float nr = sinc_weights(fr, wr);
  const float na = sinc_weights(fa, wa);
  float rx = 0.f, ry = 0.f;
#pragma unroll
  for (int ia = 0; ia < 8; ++ia) {
    const float2 *row = base + ia * ld;
    float ax = 0.f, ay = 0.f;
#pragma unroll
    for (int ir = 0; ir < 8; ++ir) {
      float2 v = row[ir];
      ax = fmaf(wr[ir], v.x, ax);
      ay = fmaf(wr[ir], v.y, ay);
    }
    rx = fmaf(wa[ia], ax, rx);
    ry = fmaf(wa[ia], ay, ry);
  }
  // normalisation of both weight sets, applied with the rotation
  const float sc = nr * na, c = cp * sc, s = sp * sc;
  return make_float2(rx * c - ry * s, rx * s + ry * c);
}

// sinc_interp on a raw burst: the taps deramped as they are read, column
// by column (a deramp_seq down each of the 8 tap columns), so the azimuth
// sums come first, then the range sum (the same separable weights; the
// summation order differs from sinc_interp's, so the result agrees with it
// to FP32 rounding, not bit for bit).
__device__ __forceinline__ float2 sinc_interp_raw(const GeomDev &g, const float2 *__restrict__ z,
                                                  int ia, int ir, float fr, float fa, float sp,
                                                  float cp) {
  float wr[8], wa[8];
  const float nr = sinc_weights(fr, wr);
  const float na = sinc_weights(fa, wa);
  const double eta0 = -g.lines * g.ati / 2. + (ia - 3) * g.ati;
  const float2 *base = z + (size_t)(ia - 3) * g.samples + (ir - 3);
  float rx = 0.f, ry = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    DerampSeq q = deramp_seq(eta0, g.ati, -0.5 * g.kt[ir - 3 + k], g.etaref[ir - 3 + k]);
    float2 t[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) t[a] = base[(size_t)a * g.samples + k];
    float ax = 0.f, ay = 0.f;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const float2 v = deramp_next(q, t[a]);
      ax = fmaf(wa[a], v.x, ax);
      ay = fmaf(wa[a], v.y, ay);
    }
    rx = fmaf(wr[k], ax, rx);
    ry = fmaf(wr[k], ay, ry);
  }
  const float sc = nr * na, c = cp * sc, s = sp * sc;
  return make_float2(rx * c - ry * s, rx * s + ry * c);
}

// Geocode one target pixel of one burst (reproject, geo2rdr.cu:194-247).
// Returns the complex value (0 outside the validity window).  rgoff/azoff
// are the sampling positions (offset polynomial applied), rng the geometric
// range.
template <int INTERP, bool RAW = false>
__device__ __forceinline__ float2 geocode_pixel(const GeomDev &g, const OffPoly &op,
                                                const double X[3], const float2 *__restrict__ z,
                                                double *rg_out = nullptr,
                                                double *az_out = nullptr,
                                                double *rng_out = nullptr) {
  double rg, az, rng;
  float sp, cp;
  const bool valid = geocode_geom<INTERP>(g, op, X, rg, az, rng, sp, cp);
  if (rg_out) *rg_out = rg;
  if (az_out) *az_out = az;
  if (rng_out) *rng_out = rng;
  if (!valid) return make_float2(0.f, 0.f);
  const int N = g.samples;
  if (INTERP == S1IFG_INTERP_BILINEAR) {
    // geo2rdr.cu:220-237: range blends rounded to float, azimuth blend FP64
    int intr = (int)rg;
    double fr = rg - intr;
    int inta = (int)az;
    double fa = az - inta;
    size_t i1 = (size_t)inta * N + intr;
    float2 b1, b2, b3, b4;
    if (RAW) {  // z is the raw burst: its taps deramped as they are read
      b1 = raw_tap(g, z, inta, intr);
      b2 = raw_tap(g, z, inta, intr + 1);
      b3 = raw_tap(g, z, inta + 1, intr);
      b4 = raw_tap(g, z, inta + 1, intr + 1);
    } else {
      b1 = z[i1];
      b2 = z[i1 + 1];
      b3 = z[i1 + N];
      b4 = z[i1 + N + 1];
    }
    float c1x = (float)(b1.x * (1 - fr) + b2.x * fr);
    float c1y = (float)(b1.y * (1 - fr) + b2.y * fr);
    float c2x = (float)(b3.x * (1 - fr) + b4.x * fr);
    float c2y = (float)(b3.y * (1 - fr) + b4.y * fr);
    float rx = (float)(c1x * (1 - fa) + c2x * fa);
    float ry = (float)(c1y * (1 - fa) + c2y * fa);
    return make_float2(rx * cp - ry * sp, rx * sp + ry * cp);
  }
  int intr = (int)floor(rg);
  int inta = (int)floor(az);
  if (RAW) return sinc_interp_raw(g, z, inta, intr, (float)(rg - intr), (float)(az - inta), sp, cp);
  return sinc_interp(z + (size_t)(inta - 3) * N + (intr - 3), N, (float)(rg - intr),
                     (float)(az - inta), sp, cp);
}

__device__ __forceinline__ void pixel_xyz(const GridDev &gr, const DemDev &dem,
                                          const RowTab *rows, const ColTab *cols,
                                          const float *dh, int r, int c, double X[3]) {
  RowTab rt = rows[r];
  ColTab ct = cols[c];
  double h = dem_height(dem, rt.fy, ct.fx);
  if (dh) h += (double)dh[(size_t)r * gr.ncol + c];
  target_xyz(rt, ct, h, X);
}

// Whether every target pixel of rows [r0, r1] x columns [c0, c1] lies
// outside the validity window of burst g, shown from 8 points: the tile's
// corners at the lowest and the highest height of its pixels, bounded by
// the extreme DEM samples that the tile's bilinear heights blend (a blend
// lies between its samples) plus the extremes of the height correction dh
// (row stride ld) over the tile when there is one.  Over a tile the sampling position is an affine function of
// (row, column, height) up to curvature terms far below a sample: the range
// curvature over the height span is sin^2(look) dh^2 / (8 R), 0.3 m for
// 2500 m of relief at R = 850 km, and over a ground extent of a few hundred
// metres it is ~1e-2 m; the azimuth time's curvature is smaller still.  So
// when all 8 points lie beyond the same edge of the window by kSkipMargin
// samples (or lines), every pixel of the tile lies beyond it, and the
// tile's products are the zeros that the per-pixel test gives.  Called by
// all 64 lanes of a wavefront; the result is uniform.  A DEM box or a tile
// of more than kSkipDemMax samples, or a corner whose Newton iteration does
// not converge, gives false.
//
// skew (optional, uniform): the burst's range samples per azimuth line
// along the tile's first grid column (corners (r0, c0) and (r1, c0) at the
// lowest height), rounded to an integer in [-kSkewMax, kSkewMax]; 0 when
// the corners were not solved or span less than one line.  The pair kernel
// shears a tap window that exceeds its LDS budget by it.
constexpr double kSkipMargin = 16.0;
constexpr int kSkipDemMax = 4096;
constexpr int kSkewMax = 8;

__device__ __forceinline__ int dem_index(double f, int n) {
  f = fmin(fmax(f, 0.0), (double)(n - 1));
  return min((int)floor(f), n - 2);
}

template <int INTERP>
__device__ __forceinline__ bool tile_outside(const GeomDev &g, const OffPoly &op, const DemDev &dem,
                                             const RowTab *rows, const ColTab *cols,
                                             const float *dh, int ld, int r0, int r1, int c0,
                                             int c1, int *skew = nullptr) {
  const int lane = threadIdx.x & 63;
  if (skew) *skew = 0;
  // DEM samples blended by the tile's pixels (the index is monotone in the
  // row / column; one sample of slack on each side)
  const int ya = dem_index(rows[r0].fy, dem.nrow), yb = dem_index(rows[r1].fy, dem.nrow);
  const int xa = dem_index(cols[c0].fx, dem.ncol), xb = dem_index(cols[c1].fx, dem.ncol);
  const int y0 = max(min(ya, yb) - 1, 0), y1 = min(max(ya, yb) + 2, dem.nrow - 1);
  const int x0 = max(min(xa, xb) - 1, 0), x1 = min(max(xa, xb) + 2, dem.ncol - 1);
  const int nx = x1 - x0 + 1, cnt = (y1 - y0 + 1) * nx;
  if (cnt > kSkipDemMax) return false;
  // (8 loads per lane in flight per step)
  float lo = INFINITY, hi = -INFINITY;
  for (int k0 = lane; k0 < cnt; k0 += 8 * 64) {
    float v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int k = k0 + 64 * m;
      v[m] = NAN;
      if (k < cnt) {
        const int y = y0 + k / nx, x = x0 + k - (k / nx) * nx;
        const size_t o = (size_t)y * dem.ncol + x;
        v[m] = dem.dtype == S1IFG_DEM_I16 ? (float)((const short *)dem.data)[o]
                                          : ((const float *)dem.data)[o];
      }
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      lo = fminf(lo, v[m]);
      hi = fmaxf(hi, v[m]);
    }
  }
  float dlo = 0.f, dhi = 0.f;
  if (dh) {
    const int tw = c1 - c0 + 1, tn = (r1 - r0 + 1) * tw;
    if (tn > kSkipDemMax) return false;
    dlo = INFINITY;
    dhi = -INFINITY;
    for (int k0 = lane; k0 < tn; k0 += 8 * 64) {
      float v[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int k = k0 + 64 * m;
        v[m] = NAN;
        if (k < tn) {
          const int y = r0 + k / tw, x = c0 + k - (k / tw) * tw;
          v[m] = dh[(size_t)y * ld + x];
        }
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        dlo = fminf(dlo, v[m]);
        dhi = fmaxf(dhi, v[m]);
      }
    }
  }
  const auto fmin = [](float a, float b) { return fminf(a, b); };
  const auto fmax = [](float a, float b) { return fmaxf(a, b); };
  lo = wave_reduce(lo, fmin);
  hi = wave_reduce(hi, fmax);
  dlo = wave_reduce(dlo, fmin);
  dhi = wave_reduce(dhi, fmax);
  // corner q = lane & 7: bit 0 column, bit 1 row, bit 2 height
  const int q = lane & 7;
  double X[3];
  target_xyz(rows[(q & 2) ? r1 : r0], cols[(q & 1) ? c1 : c0],
             (q & 4) ? (double)hi + (double)dhi : (double)lo + (double)dlo, X);
  double t, rng, rg, az, drg, daz;
  const bool ok = geo_position(g, op, X, t, rng, rg, az, drg, daz) < 52;
  if (skew) {
    // corners q = 0 (r0, c0) and q = 2 (r1, c0), lowest height
    const double ra = wave_read(rg, 0), aa = wave_read(az, 0);
    const double rb = wave_read(rg, 2), ab = wave_read(az, 2);
    const bool oka = __builtin_amdgcn_readlane((int)ok, 0), okb = __builtin_amdgcn_readlane((int)ok, 2);
    if (oka && okb && fabs(ab - aa) >= 1.0) {
      const double k = rint((rb - ra) / (ab - aa));
      *skew = (int)fmin(fmax(k, (double)-kSkewMax), (double)kSkewMax);
    }
  }
  const double lr = INTERP == S1IFG_INTERP_BILINEAR ? g.fvs : g.fvs + 3;
  const double ur = INTERP == S1IFG_INTERP_BILINEAR ? g.lvs - 1 : g.lvs - 4;
  const double la = g.flv + 5, ua = g.llv - 5, M = kSkipMargin;
  return __all(ok && rg < lr - M) || __all(ok && rg >= ur + M) || __all(ok && az < la - M) ||
         __all(ok && az >= ua + M);
}

// deramp_burst element (geo2rdr.cu:138-158): z * exp(i phi),
// phi = -pi kt (eta - etaref)^2 in FP64, reduced before a float sincos.
__device__ __forceinline__ float2 deramp_elem(double eta, double kt, double etaref, float2 a) {
  double d = eta - etaref;
  float p = reduce_phase(-kPi * kt * d * d);
  float s, c;
  sincosf(p, &s, &c);
  return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
}

// Deramp of rows [row0, row0 + kDerampRows) x sample pairs
// [pair0, pair0 + blockDim.x) of one burst, one pair per thread: the
// thread loads kt / etaref of its two samples once and issues the loads of
// all its rows before computing (rows per thread give loads in flight).
// in and out may alias (in-place deramp): each thread reads its elements
// before writing them.  vec: even row length and 16-byte aligned rows
// (float4 access); otherwise float2.
constexpr int kDerampRows = 4;

__device__ __forceinline__ void deramp_tile(const GeomDev &g, const float2 *in, float2 *out,
                                            int row0, int pair0, bool vec) {
  const int N = g.samples;
  const int jj = pair0 + (int)threadIdx.x;
  if (jj >= (N + 1) / 2) return;
  const int j = 2 * jj;
  const int nr = min(kDerampRows, g.lines - row0);
  const double eta0 = -g.lines * g.ati / 2.;
  if (vec) {
    const float4 *in4 = (const float4 *)in;
    float4 *out4 = (float4 *)out;
    const double kt0 = g.kt[j], kt1 = g.kt[j + 1];
    const double er0 = g.etaref[j], er1 = g.etaref[j + 1];
    float4 a[kDerampRows];
#pragma unroll
    for (int r = 0; r < kDerampRows; ++r)
      if (r < nr) a[r] = in4[(((size_t)(row0 + r) * N) >> 1) + jj];
#pragma unroll
    for (int r = 0; r < kDerampRows; ++r) {
      if (r >= nr) break;
      const double eta = eta0 + (row0 + r) * g.ati;
      const float2 o0 = deramp_elem(eta, kt0, er0, make_float2(a[r].x, a[r].y));
      const float2 o1 = deramp_elem(eta, kt1, er1, make_float2(a[r].z, a[r].w));
      out4[(((size_t)(row0 + r) * N) >> 1) + jj] = make_float4(o0.x, o0.y, o1.x, o1.y);
    }
  } else {
    for (int r = 0; r < nr; ++r) {
      const double eta = eta0 + (row0 + r) * g.ati;
      const size_t rowoff = (size_t)(row0 + r) * N;
      for (int k = j; k < j + 2 && k < N; ++k)
        out[rowoff + k] = deramp_elem(eta, g.kt[k], g.etaref[k], in[rowoff + k]);
    }
  }
}

// The same deramp of a burst stored as interleaved CInt16 (re, im int16
// per sample, the Sentinel-1 SLC measurement format), written as complex
// float: the int16 -> float conversion is exact, so the result equals
// deramp_tile of the converted burst bit for bit.  vec: 8-byte loads of two
// samples and 16-byte stores (even row length, aligned rows); otherwise one
// sample at a time.  in and out never alias (different element sizes).
__device__ __forceinline__ void deramp_tile_ci16(const GeomDev &g, const short *__restrict__ in,
                                                 float2 *__restrict__ out, int row0, int pair0,
                                                 bool vec) {
  const int N = g.samples;
  const int jj = pair0 + (int)threadIdx.x;
  if (jj >= (N + 1) / 2) return;
  const int j = 2 * jj;
  const int nr = min(kDerampRows, g.lines - row0);
  const double eta0 = -g.lines * g.ati / 2.;
  if (vec) {
    const short4 *in4 = (const short4 *)in;
    float4 *out4 = (float4 *)out;
    const double kt0 = g.kt[j], kt1 = g.kt[j + 1];
    const double er0 = g.etaref[j], er1 = g.etaref[j + 1];
    short4 a[kDerampRows];
#pragma unroll
    for (int r = 0; r < kDerampRows; ++r)
      if (r < nr) a[r] = in4[(((size_t)(row0 + r) * N) >> 1) + jj];
#pragma unroll
    for (int r = 0; r < kDerampRows; ++r) {
      if (r >= nr) break;
      const double eta = eta0 + (row0 + r) * g.ati;
      const float2 o0 = deramp_elem(eta, kt0, er0, make_float2(a[r].x, a[r].y));
      const float2 o1 = deramp_elem(eta, kt1, er1, make_float2(a[r].z, a[r].w));
      out4[(((size_t)(row0 + r) * N) >> 1) + jj] = make_float4(o0.x, o0.y, o1.x, o1.y);
    }
  } else {
    for (int r = 0; r < nr; ++r) {
      const double eta = eta0 + (row0 + r) * g.ati;
      const size_t rowoff = (size_t)(row0 + r) * N;
      for (int k = j; k < j + 2 && k < N; ++k)
        out[rowoff + k] = deramp_elem(eta, g.kt[k], g.etaref[k],
                                      make_float2(in[2 * (rowoff + k)], in[2 * (rowoff + k) + 1]));
    }
  }
}

// Grid tables entry i (llh2xyz trigonometry, orbit.cu:15-35, and DEM
// sampling coordinates).
__device__ __forceinline__ void grid_table_entry(const GridDev &gr, const DemDev &dem, int i,
                                                 RowTab *rows, ColTab *cols) {
  if (i < gr.nrow) {
    double lat = gr.lat0 + i * gr.dlat;
    double sl, cl;
    sincos(lat * kPi / 180, &sl, &cl);
    RowTab r;
    r.coslat = cl;
    r.sinlat = sl;
    r.re = kRa / sqrt(1. - kRe2 * sl * sl);
    r.fy = (lat - dem.lat0) / dem.dlat;
    rows[i] = r;
  }
  if (i < gr.ncol) {
    double lon = gr.lon0 + i * gr.dlon;
    double sn, cs;
    sincos(lon * kPi / 180, &sn, &cs);
    ColTab c;
    c.coslon = cs;
    c.sinlon = sn;
    c.fx = (lon - dem.lon0) / dem.dlon;
    cols[i] = c;
  }
}

// One multilooked pixel of the fused burst pair: geocode both bursts at
// every target pixel of the rl x cl cell, conj(ref)*sec, |ref|^2, |sec|^2,
// accumulated as column means then a row mean (the order of cpx_col_look /
// cpx_row_look, sario.cu:241-286).
template <int INTERP, bool RAW = false>
__device__ __forceinline__ void pair_ml_pixel(const GeomDev &gr_, const GeomDev &gs,
                                              const OffPoly &ops, const GridDev &gr,
                                              const DemDev &dem, const RowTab *rows,
                                              const ColTab *cols, const float *dh,
                                              const float2 *__restrict__ zr,
                                              const float2 *__restrict__ zs, int rl, int cl,
                                              int rm, int cm, int nc_ml, float2 *ifg, float *pr,
                                              float *ps) {
  const OffPoly none = {{0, 0, 0, 0, 0, 0}, 0};
  float ax = 0.f, ay = 0.f, apr = 0.f, aps = 0.f;
  const float inv_cl = 1.0f / cl, inv_rl = 1.0f / rl;
  for (int i = 0; i < rl; ++i) {
    int r = rm * rl + i;
    float sx = 0.f, sy = 0.f, spr = 0.f, sps = 0.f;
    for (int j = 0; j < cl; ++j) {
      int c = cm * cl + j;
      double X[3];
      pixel_xyz(gr, dem, rows, cols, dh, r, c, X);
      float2 a = geocode_pixel<INTERP, RAW>(gr_, none, X, zr);
      float2 b = geocode_pixel<INTERP, RAW>(gs, ops, X, zs);
      sx += a.x * b.x + a.y * b.y;
      sy += a.x * b.y - a.y * b.x;
      spr += a.x * a.x + a.y * a.y;
      sps += b.x * b.x + b.y * b.y;
    }
    ax += sx * inv_cl;
    ay += sy * inv_cl;
    apr += spr * inv_cl;
    aps += sps * inv_cl;
  }
  size_t o = (size_t)rm * nc_ml + cm;
  ifg[o] = make_float2(ax * inv_rl, ay * inv_rl);
  if (pr) pr[o] = apr * inv_rl;
  if (ps) ps[o] = aps * inv_rl;
}

}  // namespace s1ifg
