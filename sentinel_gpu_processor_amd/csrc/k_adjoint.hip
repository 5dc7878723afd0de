// Adjoint (reverse-mode derivative) of the fused burst pair and of the
// coherence estimators.  No reference counterpart: the reference has no
// gradient path.
//
// Pair: for a real loss L of (ifg_ml, pow_ref, pow_sec), given
// gI = dL/dRe ifg + i dL/dIm ifg and the power gradients on the multilooked
// grid, each target pixel of a look cell (n = rl * cl pixels) gets
//   G_r = (conj(gI) z_s + 2 g_pr z_r) / n,   G_s = (gI z_r + 2 g_ps z_s) / n
// (dL/dz in the same convention).  The forward of each burst is recomputed
// (geometry, interpolation taps and their derivatives; nothing is stored)
// and the chain runs through
//   out = I(az, rg) e^{i phi},  phi = 4 pi R / lambda - reramp(rp, eta),
//   rg = rgoff + drg,  az = azoff + daz,  rp = R + drg dmrg,
//   eta = eta0 + (t - tstart) + daz dtaz,
//   drg = p0 + p1 rgoff/N + p2 azoff/L,  daz = p3 + p4 rgoff/N + p5 azoff/L,
// and for the height h: dt/dh = v.n / (|v|^2 - dr.a), dR/dh = dr.n / R with
// n = dX/dh the ellipsoid normal.  Offset-polynomial gradients are reduced
// per block and then across blocks in a fixed order (deterministic).
//
// Coherence: gamma = |A| / sqrt(B C) over the window (A, B, C window sums);
// dgamma/dA = A / (|A| sqrt(B C)), dgamma/dB = -gamma / (2 B), and the
// window sums are transposed by the same (clipped, symmetric) window sums.
// Pixels whose coherence was filled (gamma > 1) or is 0 get no gradient.
#include <climits>

#include "geom_dev.h"
#include "ifg_dev.h"

using namespace s1ifg;

namespace {

// Sinc weights and their derivatives with respect to the fractional
// offset f (the forward's 8-tap Hann-windowed sinc, in float).
__device__ __forceinline__ void sinc_weights_d(float f, float w[8], float dw[8]) {
  const float s = __builtin_amdgcn_sinf(0.5f * f), c = __builtin_amdgcn_cosf(0.5f * f);
  const float s4 = __builtin_amdgcn_sinf(0.125f * f), c4 = __builtin_amdgcn_cosf(0.125f * f);
  const float ck[8] = {-0.70710678f, 0.0f, 0.70710678f, 1.0f, 0.70710678f, 0.0f, -0.70710678f, -1.0f};
  const float sk[8] = {-0.70710678f, -1.0f, -0.70710678f, 0.0f, 0.70710678f, 1.0f, 0.70710678f, 0.0f};
  const float kPiF = 3.14159265358979f;
  float h[8], dh[8], S = 0.f, dS = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    int k = i - 3;
    float x = f - (float)k;
    float sn = (k & 1) ? -s : s;   // sin(pi x)
    float cs = (k & 1) ? -c : c;   // cos(pi x)
    float sc, dsc;
    if (x == 0.0f) {
      sc = 1.0f;
      dsc = 0.0f;
    } else {
      float ix = __builtin_amdgcn_rcpf(x);
      sc = sn * ix * (1.0f / kPiF);
      dsc = (cs - sc) * ix;
    }
    float cw = c4 * ck[i] + s4 * sk[i];    // cos(pi x / 4)
    float sw = s4 * ck[i] - c4 * sk[i];    // sin(pi x / 4)
    float win = 0.5f * (1.0f + cw);
    float dwin = -0.125f * kPiF * sw;
    h[i] = sc * win;
    dh[i] = dsc * win + sc * dwin;
    S += h[i];
    dS += dh[i];
  }
  float inv = 1.0f / S;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    w[i] = h[i] * inv;
    dw[i] = (dh[i] - w[i] * dS) * inv;
  }
}

// One burst at one target pixel: value and the partials the adjoint needs.
struct BurstAdj {
  float2 z, dzr, dza;       // out, d out / d rg, d out / d az (rotation included)
  double R_rp, R_eta;       // d reramp / d rp, d reramp / d eta
  double dt_dh, drng_dh;
  double xs, ys;
  bool valid;
};

// Geometry part of one burst at one target pixel: sampling position
// (rg, az), the phase rotation (sp, cp) and the chain partials of o (all
// but the interpolated values).  Returns validity.
template <int INTERP>
__device__ __forceinline__ bool burst_geom_adj(const GeomDev &g, const OffPoly &op,
                                               const double X[3], const double nrm[3],
                                               double &rg, double &az, float &sp, float &cp,
                                               BurstAdj &o) {
  double t, dr[3], v[3], a[3];
  zero_doppler(g, X, t, dr, v, a);
  double rng = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
  double rgoff = (rng - g.rngstart) * g.inv_dmrg;
  double azoff = (t - g.tstart) * g.inv_dtaz;
  double xs = rgoff * g.inv_samples, ys = azoff * g.inv_lines;
  double drg = 0.0, daz = 0.0;
  if (op.on) {
    drg = op.p[0] + op.p[1] * xs + op.p[2] * ys;
    daz = op.p[3] + op.p[4] * xs + op.p[5] * ys;
  }
  rg = rgoff + drg;
  az = azoff + daz;
  o.xs = xs;
  o.ys = ys;
  if (INTERP == S1IFG_INTERP_BILINEAR)
    o.valid = rg > g.fvs && rg < g.lvs - 1 && az >= g.flv + 5 && az < g.llv - 5;
  else
    o.valid = rg >= g.fvs + 3 && rg < g.lvs - 4 && az >= g.flv + 5 && az < g.llv - 5;
  o.z = o.dzr = o.dza = make_float2(0.f, 0.f);
  if (!o.valid) {
    o.R_rp = o.R_eta = o.dt_dh = o.drng_dh = 0.0;
    sp = 0.f;
    cp = 1.f;
    return false;
  }
  // reramp and its partials (FP64)
  const double c2 = 2.0 / kSol;
  double rp = rng + drg * g.dmrg;
  double tf = rp * c2 - g.fm[0];
  double ka = g.fm[1] + tf * (g.fm[2] + g.fm[3] * tf);
  double dka = c2 * (g.fm[2] + 2.0 * g.fm[3] * tf);
  double td = rp * c2 - g.dc[0];
  double fnc = g.dc[1] + td * (g.dc[2] + g.dc[3] * td);
  double dfnc = c2 * (g.dc[2] + 2.0 * g.dc[3] * td);
  double ikk = rcp64(ka - g.ks), ika = rcp64(ka);
  double kt = ka * g.ks * ikk;
  double dkt = -g.ks * g.ks * ikk * ikk * dka;
  double etaref = -fnc * ika - g.etac0;
  double detaref = -(dfnc * ka - fnc * dka) * ika * ika;
  double eta = -g.lines * g.dtaz / 2. + (t - g.tstart) + daz * g.dtaz;
  double ed = eta - etaref;
  o.R_rp = -kPi * (dkt * ed * ed - 2.0 * kt * ed * detaref);
  o.R_eta = -2.0 * kPi * kt * ed;
  sincos_phase(g.k4pi_wvl * rng + kPi * kt * ed * ed, sp, cp);
  // height partials
  double vn = v[0] * nrm[0] + v[1] * nrm[1] + v[2] * nrm[2];
  double den = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] - (dr[0] * a[0] + dr[1] * a[1] + dr[2] * a[2]);
  o.dt_dh = vn * rcp64(den);
  o.drng_dh = (dr[0] * nrm[0] + dr[1] * nrm[1] + dr[2] * nrm[2]) * rcp64(rng);
  return true;
}

// Rotate the interpolated value and its derivatives by (cp, sp) into o.
__device__ __forceinline__ void set_rotated(float vx, float vy, float rx, float ry, float ax,
                                            float ay, float sp, float cp, BurstAdj &o) {
  o.z = make_float2(vx * cp - vy * sp, vx * sp + vy * cp);
  o.dzr = make_float2(rx * cp - ry * sp, rx * sp + ry * cp);
  o.dza = make_float2(ax * cp - ay * sp, ax * sp + ay * cp);
}

// 8 x 8 sinc value and its range / azimuth derivatives at fractional
// offsets (fr, fa) from the window whose top-left tap is base (row stride
// ld), rotated into o.
__device__ __forceinline__ void sinc_adj(const float2 *__restrict__ base, long ld, float fr,
                                         float fa, float sp, float cp, BurstAdj &o) {
  float wr[8], dwr[8], wa[8], dwa[8];
  sinc_weights_d(fr, wr, dwr);
  sinc_weights_d(fa, wa, dwa);
  float vx = 0.f, vy = 0.f, rx = 0.f, ry = 0.f, ax = 0.f, ay = 0.f;
#pragma unroll 2
  for (int ia = 0; ia < 8; ++ia) {
    const float2 *row = base + ia * ld;
    float sx = 0.f, sy = 0.f, dx = 0.f, dy = 0.f;
#pragma unroll
    for (int ir = 0; ir < 8; ++ir) {
      float2 q = row[ir];
      sx = fmaf(wr[ir], q.x, sx);
      sy = fmaf(wr[ir], q.y, sy);
      dx = fmaf(dwr[ir], q.x, dx);
      dy = fmaf(dwr[ir], q.y, dy);
    }
    vx = fmaf(wa[ia], sx, vx);
    vy = fmaf(wa[ia], sy, vy);
    rx = fmaf(wa[ia], dx, rx);
    ry = fmaf(wa[ia], dy, ry);
    ax = fmaf(dwa[ia], sx, ax);
    ay = fmaf(dwa[ia], sy, ay);
  }
  set_rotated(vx, vy, rx, ry, ax, ay, sp, cp, o);
}

template <int INTERP>
__device__ __forceinline__ void burst_adj(const GeomDev &g, const OffPoly &op, const double X[3],
                                          const double nrm[3], const float2 *__restrict__ z,
                                          BurstAdj &o) {
  double rg, az;
  float sp, cp;
  if (!burst_geom_adj<INTERP>(g, op, X, nrm, rg, az, sp, cp, o)) return;
  const int N = g.samples;
  if (INTERP == S1IFG_INTERP_BILINEAR) {
    int intr = (int)rg, inta = (int)az;
    float fr = (float)(rg - intr), fa = (float)(az - inta);
    size_t i1 = (size_t)inta * N + intr;
    float2 b1 = z[i1], b2 = z[i1 + 1], b3 = z[i1 + N], b4 = z[i1 + N + 1];
    float c1x = b1.x + (b2.x - b1.x) * fr, c1y = b1.y + (b2.y - b1.y) * fr;
    float c2x = b3.x + (b4.x - b3.x) * fr, c2y = b3.y + (b4.y - b3.y) * fr;
    set_rotated(c1x + (c2x - c1x) * fa, c1y + (c2y - c1y) * fa,
                (b2.x - b1.x) * (1.f - fa) + (b4.x - b3.x) * fa,
                (b2.y - b1.y) * (1.f - fa) + (b4.y - b3.y) * fa, c2x - c1x, c2y - c1y, sp, cp, o);
  } else {
    int intr = (int)floor(rg), inta = (int)floor(az);
    sinc_adj(z + (size_t)(inta - 3) * N + (intr - 3), N, (float)(rg - intr), (float)(az - inta),
             sp, cp, o);
  }
}

// Chain from dL/dout of one burst to the offset-polynomial partials (added to
// gp when given: the secondary) and the height gradient (returned).
__device__ __forceinline__ double burst_chain(const GeomDev &g, const OffPoly &op,
                                              const BurstAdj &b, float2 G, double gp[6]) {
  if (!b.valid) return 0.0;
  double gRG = (double)G.x * b.dzr.x + (double)G.y * b.dzr.y;
  double gAZ = (double)G.x * b.dza.x + (double)G.y * b.dza.y;
  // Re(conj(G) * i * z) = -Im(conj(G) z) = -(G.x z.y - G.y z.x)
  double gPHI = -((double)G.x * b.z.y - (double)G.y * b.z.x);
  double gRP = -gPHI * b.R_rp;
  double gETA = -gPHI * b.R_eta;
  double gDRG = gRG + gRP * g.dmrg;
  double gDAZ = gAZ + gETA * g.dtaz;
  double g_rgoff = gRG, g_azoff = gAZ;
  if (gp) {
    gp[0] += gDRG;
    gp[1] += gDRG * b.xs;
    gp[2] += gDRG * b.ys;
    gp[3] += gDAZ;
    gp[4] += gDAZ * b.xs;
    gp[5] += gDAZ * b.ys;
  }
  if (op.on) {
    g_rgoff += (gDRG * op.p[1] + gDAZ * op.p[4]) * g.inv_samples;
    g_azoff += (gDRG * op.p[2] + gDAZ * op.p[5]) * g.inv_lines;
  }
  double g_rng = gPHI * g.k4pi_wvl + g_rgoff * g.inv_dmrg + gRP;
  double g_t = g_azoff * g.inv_dtaz + gETA;
  return g_t * b.dt_dh + g_rng * b.drng_dh;
}

constexpr int kAdjBlock = 128;

__global__ void k_fill(float *p, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

template <int INTERP>
__global__ __launch_bounds__(kAdjBlock) void k_pair_bwd(
    const GeomDev *__restrict__ gpr, const GeomDev *__restrict__ gps, OffPoly ops, GridDev gr,
    DemDev dem, const RowTab *rows, const ColTab *cols, const float *dh,
    const float2 *__restrict__ zr, const float2 *__restrict__ zs, int rl, int cl, int nr_ml,
    int nc_ml, const float2 *__restrict__ g_ifg, const float *__restrict__ g_pr,
    const float *__restrict__ g_ps, double *__restrict__ partial, float *__restrict__ g_dh) {
  const GeomDev &Gr = const_ref(gpr), &Gs = const_ref(gps);
  const OffPoly none = {{0, 0, 0, 0, 0, 0}, 0};
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double gp[6] = {0, 0, 0, 0, 0, 0};
  if (idx < (long)nr_ml * nc_ml) {
    const int rm = (int)(idx / nc_ml), cm = (int)(idx - (long)rm * nc_ml);
    const float inv_n = 1.0f / (float)(rl * cl);
    float2 gI = g_ifg ? g_ifg[idx] : make_float2(0.f, 0.f);
    gI.x *= inv_n;
    gI.y *= inv_n;
    const float g1 = g_pr ? 2.0f * g_pr[idx] * inv_n : 0.f;
    const float g2 = g_ps ? 2.0f * g_ps[idx] * inv_n : 0.f;
    for (int i = 0; i < rl; ++i) {
      const int r = rm * rl + i;
      const RowTab rt = rows[r];
      for (int j = 0; j < cl; ++j) {
        const int c = cm * cl + j;
        const ColTab ct = cols[c];
        double X[3];
        pixel_xyz(gr, dem, rows, cols, dh, r, c, X);
        const double nrm[3] = {rt.coslat * ct.coslon, rt.coslat * ct.sinlon, rt.sinlat};
        BurstAdj br, bs;
        burst_adj<INTERP>(Gr, none, X, nrm, zr, br);
        burst_adj<INTERP>(Gs, ops, X, nrm, zs, bs);
        // G_r = conj(gI) z_s + g1 z_r ; G_s = gI z_r + g2 z_s
        float2 Gref = make_float2(gI.x * bs.z.x + gI.y * bs.z.y + g1 * br.z.x,
                                  gI.x * bs.z.y - gI.y * bs.z.x + g1 * br.z.y);
        float2 Gsec = make_float2(gI.x * br.z.x - gI.y * br.z.y + g2 * bs.z.x,
                                  gI.x * br.z.y + gI.y * br.z.x + g2 * bs.z.y);
        double gh = burst_chain(Gr, none, br, Gref, nullptr) + burst_chain(Gs, ops, bs, Gsec, gp);
        if (g_dh) g_dh[(size_t)r * gr.ncol + c] = (float)gh;
      }
    }
  }
  // block reduction of the 6 partials (fixed tree order)
  __shared__ double red[6][kAdjBlock];
  for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = gp[k];
  __syncthreads();
  for (int s = kAdjBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[(size_t)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Sinc adjoint on tiles of kBwdRows x kBwdCols target pixels (one per
// thread), the burst windows staged in LDS as in the forward tile kernel:
// the geometry and chain partials of a pixel stay in registers while the
// block stages the bounding box of its taps; a window larger than kBwdWin
// samples is read from global memory instead (same arithmetic).
// 8 x 32: the target grid is rotated by the heading against the burst and
// its rows span about two azimuth lines, so this shape needs ~1,200 samples
// per burst (worst seen ~1,900) where 16 x 16 needed ~1,900 (~3,500).
constexpr int kBwdRows = 8, kBwdCols = 32, kBwdThreads = kBwdRows * kBwdCols;
constexpr int kBwdWin = 2048;
constexpr int kBwdStage = 8;  // window loads in flight per thread while staging
// Skip regions: 2 x 2 tiles (16 x 64 target pixels) share one skip flag per
// burst, tested by k_bwd_flags before the tiles run.
constexpr int kBwdRegion = 2;

struct BwdLds {
  float2 win[2][kBwdWin];
  int box[2][4];
  double red[kBwdThreads / 64][6];
};

// FP64 wavefront sum with the result in every lane (fixed order).
__device__ __forceinline__ double wave_sum_d(double v) {
  return wave_reduce(v, [](double a, double b) { return a + b; });
}

__device__ __forceinline__ int wmin_i(int v) {
  return wave_reduce(v, [](int a, int b) { return min(a, b); });
}
__device__ __forceinline__ int wmax_i(int v) {
  return wave_reduce(v, [](int a, int b) { return max(a, b); });
}

// Skip flags of the backward's regions (2 x 2 tiles): wavefront w of the
// grid tests region w / 2 (row-major, rx per row) against burst w % 2
// (tile_outside), over the region's pixels inside the look cells.
__global__ __launch_bounds__(256) void k_bwd_flags(
    const GeomDev *__restrict__ gpr, const GeomDev *__restrict__ gps, OffPoly ops, GridDev gr,
    DemDev dem, const RowTab *rows, const ColTab *cols, const float *dh, int rl, int cl,
    int nr_ml, int nc_ml, int rx, long nflags, unsigned char *__restrict__ flags) {
  const OffPoly none = {{0, 0, 0, 0, 0, 0}, 0};
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= nflags) return;
  const long t = w >> 1;
  const int b = (int)(w & 1);
  const int ry = (int)(t / rx), cx = (int)(t - (long)ry * rx);
  const int r0 = ry * kBwdRegion * kBwdRows, c0 = cx * kBwdRegion * kBwdCols;
  const int r1 = min(r0 + kBwdRegion * kBwdRows, nr_ml * rl) - 1;
  const int c1 = min(c0 + kBwdRegion * kBwdCols, nc_ml * cl) - 1;
  const bool o = tile_outside<S1IFG_INTERP_SINC>(b ? const_ref(gps) : const_ref(gpr), b ? ops : none, dem, rows, cols,
                                                 dh, gr.ncol, r0, r1, c0, c1);
  if ((threadIdx.x & 63) == 0) flags[w] = o;
}

__global__ __launch_bounds__(kBwdThreads) void k_pair_bwd_tiles(
    const GeomDev *__restrict__ gpr, const GeomDev *__restrict__ gps, OffPoly ops, GridDev gr,
    DemDev dem, const RowTab *rows, const ColTab *cols, const float *dh,
    const float2 *__restrict__ zr, const float2 *__restrict__ zs, int rl, int cl, int nr_ml,
    int nc_ml, const float2 *__restrict__ g_ifg, const float *__restrict__ g_pr,
    const float *__restrict__ g_ps, double *__restrict__ partial, float *__restrict__ g_dh,
    const unsigned char *__restrict__ flags) {
  __shared__ BwdLds L;
  const GeomDev *G[2] = {gpr, gps};
  const float2 *Z[2] = {zr, zs};
  const OffPoly none = {{0, 0, 0, 0, 0, 0}, 0};
  const int tid = threadIdx.x;
  const int r = blockIdx.y * kBwdRows + tid / kBwdCols;
  const int c = blockIdx.x * kBwdCols + (tid & (kBwdCols - 1));
  const bool active = r < nr_ml * rl && c < nc_ml * cl;
  if (tid < 8) L.box[tid >> 2][tid & 3] = (tid & 3) < 2 ? INT_MAX : INT_MIN;
  // the skip flags of the tile's region (k_bwd_flags)
  const long reg = (long)(blockIdx.y / kBwdRegion) * ((gridDim.x + kBwdRegion - 1) / kBwdRegion) +
                   blockIdx.x / kBwdRegion;
  const bool out[2] = {flags[2 * reg] != 0, flags[2 * reg + 1] != 0};
  if (out[0] && out[1]) {
    // no pixel of the tile in either burst: zero gradients
    if (active && g_dh) g_dh[(size_t)r * gr.ncol + c] = 0.f;
    if (tid < 6) partial[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 6 + tid] = 0.0;
    return;
  }
  BurstAdj b[2];
  int ir[2] = {0, 0}, ia[2] = {0, 0};
  float fr[2] = {0.f, 0.f}, fa[2] = {0.f, 0.f}, sp[2] = {0.f, 0.f}, cp[2] = {1.f, 1.f};
  int lo_r[2] = {INT_MAX, INT_MAX}, lo_a[2] = {INT_MAX, INT_MAX};
  int hi_r[2] = {INT_MIN, INT_MIN}, hi_a[2] = {INT_MIN, INT_MIN};
  b[0].valid = b[1].valid = false;
  b[0].z = b[1].z = make_float2(0.f, 0.f);
  if (active) {
    double X[3];
    pixel_xyz(gr, dem, rows, cols, dh, r, c, X);
    const RowTab rt = rows[r];
    const ColTab ct = cols[c];
    const double nrm[3] = {rt.coslat * ct.coslon, rt.coslat * ct.sinlon, rt.sinlat};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      double rg, az;
      if (!out[k] && burst_geom_adj<S1IFG_INTERP_SINC>(*G[k], k ? ops : none, X, nrm, rg, az,
                                                       sp[k], cp[k], b[k])) {
        ir[k] = (int)floor(rg);
        ia[k] = (int)floor(az);
        fr[k] = (float)(rg - ir[k]);
        fa[k] = (float)(az - ia[k]);
        lo_r[k] = hi_r[k] = ir[k];
        lo_a[k] = hi_a[k] = ia[k];
      }
    }
  }
  __syncthreads();  // box initialised
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int v0 = wmin_i(lo_r[k]), v1 = wmin_i(lo_a[k]);
    const int v2 = wmax_i(hi_r[k]), v3 = wmax_i(hi_a[k]);
    if ((tid & 63) == 0) {
      atomicMin(&L.box[k][0], v0);
      atomicMin(&L.box[k][1], v1);
      atomicMax(&L.box[k][2], v2);
      atomicMax(&L.box[k][3], v3);
    }
  }
  __syncthreads();
  int wc[2], wr0[2], wa0[2];
  bool staged[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int mr = L.box[k][0], ma = L.box[k][1], xr = L.box[k][2], xa = L.box[k][3];
    const bool any = mr <= xr;
    const long W = any ? (long)xr - mr + 8 : 0, H = any ? (long)xa - ma + 8 : 0;
    staged[k] = any && W * H <= kBwdWin;
    wc[k] = (int)W;
    wr0[k] = mr - 3;
    wa0[k] = ma - 3;
    if (staged[k]) {
      // element i = row * W + col, kBwdThreads apart per thread; kBwdStage
      // loads are issued before their LDS stores
      const int ns = G[k]->samples, w = wc[k];
      const int n = (int)(W * H), dq = kBwdThreads / w, dr = kBwdThreads - dq * w;
      int row = tid / w, col = tid - row * w;
      for (int i0 = tid; i0 < n; i0 += kBwdStage * kBwdThreads) {
        float2 v[kBwdStage];
#pragma unroll
        for (int m = 0; m < kBwdStage; ++m) {
          if (i0 + m * kBwdThreads < n) v[m] = Z[k][(size_t)(wa0[k] + row) * ns + wr0[k] + col];
          row += dq;
          col += dr;
          if (col >= w) {
            col -= w;
            ++row;
          }
        }
#pragma unroll
        for (int m = 0; m < kBwdStage; ++m)
          if (i0 + m * kBwdThreads < n) L.win[k][i0 + m * kBwdThreads] = v[m];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (!b[k].valid) continue;
    if (staged[k])
      sinc_adj(&L.win[k][(ia[k] - 3 - wa0[k]) * wc[k] + (ir[k] - 3 - wr0[k])], wc[k], fr[k], fa[k],
               sp[k], cp[k], b[k]);
    else
      sinc_adj(Z[k] + (size_t)(ia[k] - 3) * G[k]->samples + (ir[k] - 3), G[k]->samples, fr[k],
               fa[k], sp[k], cp[k], b[k]);
  }
  double gp[6] = {0, 0, 0, 0, 0, 0};
  if (active) {
    const long idx = (long)(r / rl) * nc_ml + c / cl;
    const float inv_n = 1.0f / (float)(rl * cl);
    float2 gI = g_ifg ? g_ifg[idx] : make_float2(0.f, 0.f);
    gI.x *= inv_n;
    gI.y *= inv_n;
    const float g1 = g_pr ? 2.0f * g_pr[idx] * inv_n : 0.f;
    const float g2 = g_ps ? 2.0f * g_ps[idx] * inv_n : 0.f;
    const float2 z0 = b[0].z, z1 = b[1].z;
    // G_r = conj(gI) z_s + g1 z_r ; G_s = gI z_r + g2 z_s
    float2 Gref = make_float2(gI.x * z1.x + gI.y * z1.y + g1 * z0.x,
                              gI.x * z1.y - gI.y * z1.x + g1 * z0.y);
    float2 Gsec = make_float2(gI.x * z0.x - gI.y * z0.y + g2 * z1.x,
                              gI.x * z0.y + gI.y * z0.x + g2 * z1.y);
    double gh = burst_chain(const_ref(gpr), none, b[0], Gref, nullptr) +
                burst_chain(const_ref(gps), ops, b[1], Gsec, gp);
    if (g_dh) g_dh[(size_t)r * gr.ncol + c] = (float)gh;
  }
  // block sum of the 6 partials: wavefront reductions (fixed order), then
  // the wavefronts in order
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double v = wave_sum_d(gp[k]);
    if ((tid & 63) == 0) L.red[tid >> 6][k] = v;
  }
  __syncthreads();
  if (tid < 6) {
    double v = 0.0;
    for (int w = 0; w < kBwdThreads / 64; ++w) v += L.red[w][tid];
    partial[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 6 + tid] = v;
  }
}

// Sum of n records of 6 partials in two fixed-order levels: block b of
// the first level sums records [b chunk, (b + 1) chunk) (thread t adds
// records t, t + 256, ... of the chunk, then a fixed tree over the
// threads); one block then sums the block results the same way.
constexpr int kReduceThreads = 256;
constexpr int kReduceBlocks = 256;

__global__ __launch_bounds__(kReduceThreads) void k_reduce6(const double *__restrict__ in, long n,
                                                            long chunk, double *out) {
  __shared__ double red[6][kReduceThreads];
  double s[6] = {0, 0, 0, 0, 0, 0};
  const long b0 = (long)blockIdx.x * chunk;
  const long b1 = b0 + chunk < n ? b0 + chunk : n;
  for (long b = b0 + threadIdx.x; b < b1; b += kReduceThreads)
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] += in[(size_t)b * 6 + k];
#pragma unroll
  for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = s[k];
  __syncthreads();
  for (int w = kReduceThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 6) out[(size_t)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Coherence adjoint, step 1: per output pixel q the coefficients of the
// window sums (cA complex, cB, cC); win = 1 gives the final gradients.
__global__ void k_coh_bwd_coef(const float2 *__restrict__ ifg, const float *__restrict__ p1,
                               const float *__restrict__ p2, int nr, int nc, int win,
                               const float *__restrict__ g, float2 *__restrict__ cA,
                               float *__restrict__ cB, float *__restrict__ cC) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= nr || c >= nc) return;
  size_t o = (size_t)r * nc + c;
  float gq = g[o];
  float2 ra = make_float2(0.f, 0.f);
  float rb = 0.f, rc = 0.f;
  if (win == 1) {
    float2 v = ifg[o];
    float a1 = sqrtf(p1[o]) + 1e-12f, a2 = sqrtf(p2[o]) + 1e-12f;
    float mag = sqrtf(v.x * v.x + v.y * v.y);
    float gam = mag / a1 / a2;
    if (gam <= 1.0f && mag > 0.f && p1[o] > 0.f && p2[o] > 0.f) {
      float s = gq / (a1 * a2 * mag);
      ra = make_float2(s * v.x, s * v.y);
      rb = -gq * gam / a1 * 0.5f / sqrtf(p1[o]);
      rc = -gq * gam / a2 * 0.5f / sqrtf(p2[o]);
    }
  } else {
    int h = win / 2;
    int r0 = max(r - h, 0), r1 = min(r + h, nr - 1);
    int c0 = max(c - h, 0), c1 = min(c + h, nc - 1);
    float sx = 0.f, sy = 0.f, s1 = 0.f, s2 = 0.f;
    for (int i = r0; i <= r1; ++i)
      for (int j = c0; j <= c1; ++j) {
        size_t b = (size_t)i * nc + j;
        float2 v = ifg[b];
        sx += v.x;
        sy += v.y;
        s1 += p1[b];
        s2 += p2[b];
      }
    float den = sqrtf(s1 * s2);
    float mag = sqrtf(sx * sx + sy * sy);
    if (den > 0.f && mag > 0.f) {
      float gam = mag / den;
      if (gam <= 1.0f) {
        float s = gq / (mag * den);
        ra = make_float2(s * sx, s * sy);
        rb = -0.5f * gq * gam / s1;
        rc = -0.5f * gq * gam / s2;
      }
    }
  }
  cA[o] = ra;
  cB[o] = rb;
  cC[o] = rc;
}

// Step 2 (win >= 3): transpose of the window sums = the same window sums.
__global__ void k_coh_bwd_box(const float2 *__restrict__ cA, const float *__restrict__ cB,
                              const float *__restrict__ cC, int nr, int nc, int win,
                              float2 *__restrict__ gI, float *__restrict__ g1,
                              float *__restrict__ g2) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= nr || c >= nc) return;
  int h = win / 2;
  int r0 = max(r - h, 0), r1 = min(r + h, nr - 1);
  int c0 = max(c - h, 0), c1 = min(c + h, nc - 1);
  float sx = 0.f, sy = 0.f, s1 = 0.f, s2 = 0.f;
  for (int i = r0; i <= r1; ++i)
    for (int j = c0; j <= c1; ++j) {
      size_t b = (size_t)i * nc + j;
      sx += cA[b].x;
      sy += cA[b].y;
      s1 += cB[b];
      s2 += cC[b];
    }
  size_t o = (size_t)r * nc + c;
  if (gI) gI[o] = make_float2(sx, sy);
  if (g1) g1[o] = s1;
  if (g2) g2[o] = s2;
}

// Tiled forms of the two steps (win / 2 <= kCohMaxHalo): the window sums
// come from an LDS tile with halo (the forward's staging); same sums, same
// order.
__global__ __launch_bounds__(kCohTX * kCohTY) void k_coh_bwd_coef_tiled(
    const float2 *__restrict__ ifg, const float *__restrict__ p1, const float *__restrict__ p2,
    int nr, int nc, int win, const float *__restrict__ g, float2 *__restrict__ cA,
    float *__restrict__ cB, float *__restrict__ cC) {
  __shared__ CohTile L;
  int r, c;
  float sx, sy, s1, s2;
  if (!window_sums_tile(ifg, p1, p2, nr, nc, win, blockIdx.x, blockIdx.y, L, r, c, sx, sy, s1,
                        s2))
    return;
  const size_t o = (size_t)r * nc + c;
  const float gq = g[o];
  float2 ra = make_float2(0.f, 0.f);
  float rb = 0.f, rc = 0.f;
  float den = sqrtf(s1 * s2);
  float mag = sqrtf(sx * sx + sy * sy);
  if (den > 0.f && mag > 0.f) {
    float gam = mag / den;
    if (gam <= 1.0f) {
      float s = gq / (mag * den);
      ra = make_float2(s * sx, s * sy);
      rb = -0.5f * gq * gam / s1;
      rc = -0.5f * gq * gam / s2;
    }
  }
  cA[o] = ra;
  cB[o] = rb;
  cC[o] = rc;
}

__global__ __launch_bounds__(kCohTX * kCohTY) void k_coh_bwd_box_tiled(
    const float2 *__restrict__ cA, const float *__restrict__ cB, const float *__restrict__ cC,
    int nr, int nc, int win, float2 *__restrict__ gI, float *__restrict__ g1,
    float *__restrict__ g2) {
  __shared__ CohTile L;
  int r, c;
  float sx, sy, s1, s2;
  if (!window_sums_tile(cA, cB, cC, nr, nc, win, blockIdx.x, blockIdx.y, L, r, c, sx, sy, s1, s2))
    return;
  const size_t o = (size_t)r * nc + c;
  if (gI) gI[o] = make_float2(sx, sy);
  if (g1) g1[o] = s1;
  if (g2) g2[o] = s2;
}

// The same two passes over 64 x kStripRows strips (window_sums_strip).
__global__ __launch_bounds__(kCohTX * kCohTY) void k_coh_bwd_coef_strip(
    const float2 *__restrict__ ifg, const float *__restrict__ p1, const float *__restrict__ p2,
    int nr, int nc, int win, const float *__restrict__ g, float2 *__restrict__ cA,
    float *__restrict__ cB, float *__restrict__ cC) {
  __shared__ CohStrip L;
  window_sums_strip(ifg, p1, p2, nr, nc, win, blockIdx.x, blockIdx.y, L,
                    [&](int r, int c, float sx, float sy, float s1, float s2) {
                      const size_t o = (size_t)r * nc + c;
                      const float gq = g[o];
                      float2 ra = make_float2(0.f, 0.f);
                      float rb = 0.f, rc = 0.f;
                      const float den = sqrtf(s1 * s2);
                      const float mag = sqrtf(sx * sx + sy * sy);
                      if (den > 0.f && mag > 0.f) {
                        const float gam = mag / den;
                        if (gam <= 1.0f) {
                          const float s = gq / (mag * den);
                          ra = make_float2(s * sx, s * sy);
                          rb = -0.5f * gq * gam / s1;
                          rc = -0.5f * gq * gam / s2;
                        }
                      }
                      cA[o] = ra;
                      cB[o] = rb;
                      cC[o] = rc;
                    });
}

__global__ __launch_bounds__(kCohTX * kCohTY) void k_coh_bwd_box_strip(
    const float2 *__restrict__ cA, const float *__restrict__ cB, const float *__restrict__ cC,
    int nr, int nc, int win, float2 *__restrict__ gI, float *__restrict__ g1,
    float *__restrict__ g2) {
  __shared__ CohStrip L;
  window_sums_strip(cA, cB, cC, nr, nc, win, blockIdx.x, blockIdx.y, L,
                    [&](int r, int c, float sx, float sy, float s1, float s2) {
                      const size_t o = (size_t)r * nc + c;
                      if (gI) gI[o] = make_float2(sx, sy);
                      if (g1) g1[o] = s1;
                      if (g2) g2[o] = s2;
                    });
}

}  // namespace

namespace s1ifg {
int grid_tables(s1ifg_ctx *ctx, const GridDev &gr, const DemDev &dem, RowTab **rows,
                ColTab **cols, size_t extra, void **extra_ptr);
}

extern "C" {

int s1ifg_pair_backward(s1ifg_ctx *ctx, const s1ifg_geom *g_ref, const s1ifg_geom *g_sec,
                        const s1ifg_grid *grid, const s1ifg_dem *dem, const float *dh,
                        const double *offpoly_sec, const s1ifg_cf32 *ref_deramped,
                        const s1ifg_cf32 *sec_deramped, int interp, int rl, int cl,
                        const s1ifg_cf32 *g_ifg, const float *g_pow_ref,
                        const float *g_pow_sec, double *g_offpoly, float *g_dh) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if ((rc = check_geom(ctx, g_ref)) || (rc = check_geom(ctx, g_sec))) return rc;
  GridDev gr;
  DemDev dd;
  OffPoly op;
  if ((rc = make_grid(grid, &gr)) || (rc = make_dem(dem, &dd)) ||
      (rc = make_offpoly(offpoly_sec, &op)))
    return rc;
  if (!ref_deramped || !sec_deramped) return set_error(S1IFG_EINVAL, "null burst pointer");
  if (rl < 1 || cl < 1) return set_error(S1IFG_EINVAL, "looks %d x %d", rl, cl);
  if (interp != S1IFG_INTERP_BILINEAR && interp != S1IFG_INTERP_SINC)
    return set_error(S1IFG_EINVAL, "interp %d", interp);
  int nr = gr.nrow / rl, nc = gr.ncol / cl;
  if (nr < 1 || nc < 1) return set_error(S1IFG_EINVAL, "grid smaller than one look cell");
  long nml = (long)nr * nc;
  const bool tiles = interp == S1IFG_INTERP_SINC;
  const dim3 tgrid((unsigned)((nc * cl + kBwdCols - 1) / kBwdCols),
                   (unsigned)((nr * rl + kBwdRows - 1) / kBwdRows));
  unsigned blocks = tiles ? tgrid.x * tgrid.y : (unsigned)((nml + kAdjBlock - 1) / kAdjBlock);
  RowTab *rows;
  ColTab *cols;
  void *part = nullptr;
  // per-block partials, then kReduceBlocks first-level sums, then the skip
  // flags of the tile regions
  const int rx = (int)((tgrid.x + kBwdRegion - 1) / kBwdRegion);
  const long nflags = tiles ? 2L * rx * ((tgrid.y + kBwdRegion - 1) / kBwdRegion) : 0;
  const size_t part_bytes = sizeof(double) * 6 * ((size_t)blocks + kReduceBlocks);
  if ((rc = grid_tables(ctx, gr, dd, &rows, &cols, part_bytes + (size_t)nflags, &part)))
    return rc;
  unsigned char *flags = (unsigned char *)part + part_bytes;
  if (g_dh) {
    long n = (long)gr.nrow * gr.ncol;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       g_dh, n);
    S1_LAUNCH_CHECK("k_fill");
  }
  if (interp == S1IFG_INTERP_BILINEAR)
    hipLaunchKernelGGL(k_pair_bwd<S1IFG_INTERP_BILINEAR>, dim3(blocks), dim3(kAdjBlock), 0,
                       ctx->stream, g_ref->d_dev, g_sec->d_dev, op, gr, dd, rows, cols, dh,
                       (const float2 *)ref_deramped, (const float2 *)sec_deramped, rl, cl, nr, nc,
                       (const float2 *)g_ifg, g_pow_ref, g_pow_sec, (double *)part, g_dh);
  else {
    hipLaunchKernelGGL(k_bwd_flags, dim3((unsigned)((nflags + 3) / 4)), dim3(256), 0,
                       ctx->stream, g_ref->d_dev, g_sec->d_dev, op, gr, dd, rows, cols, dh, rl,
                       cl, nr, nc, rx, nflags, flags);
    S1_LAUNCH_CHECK("k_bwd_flags");
    hipLaunchKernelGGL(k_pair_bwd_tiles, tgrid, dim3(kBwdThreads), 0, ctx->stream,
                       g_ref->d_dev, g_sec->d_dev, op, gr, dd, rows, cols, dh,
                       (const float2 *)ref_deramped, (const float2 *)sec_deramped, rl, cl, nr, nc,
                       (const float2 *)g_ifg, g_pow_ref, g_pow_sec, (double *)part, g_dh,
                       (const unsigned char *)flags);
  }
  S1_LAUNCH_CHECK("k_pair_bwd");
  if (g_offpoly) {
    double *lvl = (double *)part + 6 * (size_t)blocks;
    const long chunk = ((long)blocks + kReduceBlocks - 1) / kReduceBlocks;
    hipLaunchKernelGGL(k_reduce6, dim3(kReduceBlocks), dim3(kReduceThreads), 0, ctx->stream,
                       (const double *)part, (long)blocks, chunk, lvl);
    hipLaunchKernelGGL(k_reduce6, dim3(1), dim3(kReduceThreads), 0, ctx->stream,
                       (const double *)lvl, (long)kReduceBlocks, (long)kReduceBlocks, g_offpoly);
    S1_LAUNCH_CHECK("k_reduce6");
  }
  return S1IFG_OK;
}

int s1ifg_coherence_backward(s1ifg_ctx *ctx, const s1ifg_cf32 *ifg_ml, const float *pow_ref,
                             const float *pow_sec, int nrow, int ncol, int win,
                             const float *g_coh, s1ifg_cf32 *g_ifg, float *g_pow_ref,
                             float *g_pow_sec) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!ifg_ml || !pow_ref || !pow_sec || !g_coh) return set_error(S1IFG_EINVAL, "null pointer");
  if (nrow < 1 || ncol < 1) return set_error(S1IFG_EINVAL, "bad size");
  if (win < 1 || win % 2 == 0) return set_error(S1IFG_EINVAL, "window %d (odd, >= 1)", win);
  size_t n = (size_t)nrow * ncol;
  dim3 blk(64, 4), grd((ncol + 63) / 64, (nrow + 3) / 4);
  if (win == 1) {
    if (!g_ifg || !g_pow_ref || !g_pow_sec)
      return set_error(S1IFG_EINVAL, "win = 1 needs all three gradient outputs");
    hipLaunchKernelGGL(k_coh_bwd_coef, grd, blk, 0, ctx->stream, (const float2 *)ifg_ml, pow_ref,
                       pow_sec, nrow, ncol, 1, g_coh, (float2 *)g_ifg, g_pow_ref, g_pow_sec);
    S1_LAUNCH_CHECK("k_coh_bwd_coef");
    return S1IFG_OK;
  }
  void *ws = nullptr;
  if ((rc = ctx_workspace(ctx, n * 16 + 512, &ws))) return rc;
  float2 *cA = (float2 *)ws;
  float *cB = (float *)((char *)ws + ((n * 8 + 255) & ~(size_t)255));
  float *cC = (float *)((char *)cB + ((n * 4 + 255) & ~(size_t)255));
  if (win / 2 <= kStripMaxHalo) {
    dim3 sgrd((ncol + kCohTX - 1) / kCohTX, (nrow + kStripRows - 1) / kStripRows);
    hipLaunchKernelGGL(k_coh_bwd_coef_strip, sgrd, blk, 0, ctx->stream, (const float2 *)ifg_ml,
                       pow_ref, pow_sec, nrow, ncol, win, g_coh, cA, cB, cC);
    S1_LAUNCH_CHECK("k_coh_bwd_coef_strip");
    hipLaunchKernelGGL(k_coh_bwd_box_strip, sgrd, blk, 0, ctx->stream, cA, cB, cC, nrow, ncol,
                       win, (float2 *)g_ifg, g_pow_ref, g_pow_sec);
    S1_LAUNCH_CHECK("k_coh_bwd_box_strip");
    return S1IFG_OK;
  }
  if (win / 2 <= kCohMaxHalo) {
    hipLaunchKernelGGL(k_coh_bwd_coef_tiled, grd, blk, 0, ctx->stream, (const float2 *)ifg_ml,
                       pow_ref, pow_sec, nrow, ncol, win, g_coh, cA, cB, cC);
    S1_LAUNCH_CHECK("k_coh_bwd_coef_tiled");
    hipLaunchKernelGGL(k_coh_bwd_box_tiled, grd, blk, 0, ctx->stream, cA, cB, cC, nrow, ncol,
                       win, (float2 *)g_ifg, g_pow_ref, g_pow_sec);
    S1_LAUNCH_CHECK("k_coh_bwd_box_tiled");
    return S1IFG_OK;
  }
  hipLaunchKernelGGL(k_coh_bwd_coef, grd, blk, 0, ctx->stream, (const float2 *)ifg_ml, pow_ref,
                     pow_sec, nrow, ncol, win, g_coh, cA, cB, cC);
  S1_LAUNCH_CHECK("k_coh_bwd_coef");
  hipLaunchKernelGGL(k_coh_bwd_box, grd, blk, 0, ctx->stream, cA, cB, cC, nrow, ncol, win,
                     (float2 *)g_ifg, g_pow_ref, g_pow_sec);
  S1_LAUNCH_CHECK("k_coh_bwd_box");
  return S1IFG_OK;
}

}  // extern "C"
