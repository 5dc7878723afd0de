// Fused burst pair with sinc interpolation, the bursts' tap windows staged
// in LDS (shared by the per-pair and batched kernels).
//
// A 256-thread block owns a tile of 16 x 16 look cells.  Two wavefronts
// first test whether the whole tile lies outside either burst
// (tile_outside: about half of a burst's target grid, the corners of its
// latitude / longitude box, does); a burst that the tile misses is not
// geocoded, and a tile that misses both writes zeros.  Each thread owns
// one cell (rl * cl <= kPairMaxPix target pixels).  Then, for burst 0 and
// then burst 1 through the same LDS window: the thread computes the FP64
// geometry and phase of its pixels and keeps the integer / fractional
// sampling positions and sin/cos of the phase in registers; the block
// reduces the bounding box of the 8 x 8 tap windows, loads that window of
// the deramped burst into LDS (stage_window: many loads in flight per
// thread), and every pixel is interpolated from LDS.  A window larger than
// kPairWin samples is read from global memory instead; on raw bursts it is
// first sheared by the tile's skew (row a of the window starts sk columns
// after row a - 1: a grid column runs diagonally through the burst, so its
// tap rows shift in range) and staged if that fits, since a raw burst's
// taps read from global memory are deramped one by one.  All ways run the
// same arithmetic in the same order as pair_ml_pixel, so the products are
// bit-identical to it.
#pragma once

#include <climits>

#include "geom_dev.h"
#include "wave_dev.h"

namespace s1ifg {

constexpr int kPairTile = 16;        // look cells per tile side
constexpr int kPairWin = 4864;       // staged samples of burst 0 (38 KB; 4 blocks per CU)
constexpr int kPairMaxPix = 4;       // target pixels per look cell on this path
constexpr int kStage = 16;           // window loads in flight per thread while staging

struct PairLds {
  float2 win[kPairWin];  // one burst's tap window at a time
  int box[2][4];         // per burst: min intr, min inta, max intr, max inta
  int sbox[2][2];        // per burst: first and last column of the sheared window
  int out[2];            // per burst: every pixel of the tile outside the burst
  int skew[2];           // per burst: shear of the window (tile_outside)
};

// Stage rows [wa0, wa0 + H) x columns [wr0, wr0 + W) of burst z (row
// stride ns) into win (row-major, W per row).  All 256 threads call it;
// element i = row * W + col is 256 apart per thread and kStage loads are
// issued before their LDS stores (row and column advance by the quotient
// and remainder of 256 / W).
__device__ __forceinline__ void stage_window(const float2 *__restrict__ z, int ns, int wa0,
                                             int wr0, int W, int H, float2 *win) {
  const int tid = threadIdx.x;
  const int n = W * H, dq = 256 / W, dr = 256 - dq * W;
  int row = tid / W, col = tid - row * W;
  for (int i0 = tid; i0 < n; i0 += kStage * 256) {
    float2 v[kStage];
#pragma unroll
    for (int k = 0; k < kStage; ++k) {
      if (i0 + k * 256 < n) v[k] = z[(size_t)(wa0 + row) * ns + wr0 + col];
      row += dq;
      col += dr;
      if (col >= W) {
        col -= W;
        ++row;
      }
    }
#pragma unroll
    for (int k = 0; k < kStage; ++k)
      if (i0 + k * 256 < n) win[i0 + k * 256] = v[k];
  }
}

// stage_window from a raw burst, deramped as it is staged: wavefront w
// stages rows w, w + 4, ...; its lanes take consecutive columns (coalesced
// rows), a lane's rotations follow its column's deramp_seq with a step of
// 4 rows, and kStage rows' loads are issued before their rotations.
__device__ __forceinline__ void stage_window_deramp(const float2 *__restrict__ z,
                                                    const GeomDev &g, int wa0, int wr0, int W,
                                                    int H, float2 *win) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ns = g.samples;
  const double eta0 = -g.lines * g.ati / 2.;
  for (int c = lane; c < W; c += 64) {
    const int col = wr0 + c;
    DerampSeq q =
        deramp_seq(eta0 + (wa0 + wave) * g.ati, 4 * g.ati, -0.5 * g.kt[col], g.etaref[col]);
    for (int r0 = wave; r0 < H; r0 += 4 * kStage) {
      float2 v[kStage];
#pragma unroll
      for (int k = 0; k < kStage; ++k) {
        const int r = r0 + 4 * k;
        if (r < H) v[k] = z[(size_t)(wa0 + r) * ns + col];
      }
#pragma unroll
      for (int k = 0; k < kStage; ++k) {
        const int r = r0 + 4 * k;
        if (r < H) win[r * W + c] = deramp_next(q, v[k]);
      }
    }
  }
}

// Floor and ceiling of a / b for b > 0.
__device__ __forceinline__ int floor_div(int a, int b) { return a >= 0 ? a / b : -((b - 1 - a) / b); }
__device__ __forceinline__ int ceil_div(int a, int b) { return -floor_div(-a, b); }

// stage_window_deramp for a sheared window (row r holds burst columns
// [wr0 + r dk, wr0 + r dk + W), dk != 0): the lanes walk burst columns, so
// that each keeps the deramp_seq of one column; a column meets the rows r
// with 0 <= col - wr0 - r dk < W, of which wavefront w takes r = w mod 4.
// Columns outside the burst (never tapped) are not staged.
__device__ __forceinline__ void stage_window_deramp_skew(const float2 *__restrict__ z,
                                                         const GeomDev &g, int wa0, int wr0,
                                                         int W, int H, int dk, float2 *win) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ns = g.samples;
  const double eta0 = -g.lines * g.ati / 2.;
  const int gc0 = wr0 + min(0, dk * (H - 1));
  const int span = W + abs(dk) * (H - 1);
  for (int c = lane; c < span; c += 64) {
    const int col = gc0 + c;
    if ((unsigned)col >= (unsigned)ns) continue;
    const int e = col - wr0;  // the column sits at e - r dk in row r
    int rlo, rhi;
    if (dk > 0) {
      rlo = floor_div(e - W, dk) + 1;
      rhi = floor_div(e, dk);
    } else {
      rlo = ceil_div(-e, -dk);
      rhi = floor_div(W - 1 - e, -dk);
    }
    rlo = max(rlo, 0);
    rhi = min(rhi, H - 1);
    const int r0 = rlo + ((wave - rlo) & 3);  // this wavefront's first row
    if (r0 > rhi) continue;
    DerampSeq q =
        deramp_seq(eta0 + (wa0 + r0) * g.ati, 4 * g.ati, -0.5 * g.kt[col], g.etaref[col]);
    for (int rb = r0; rb <= rhi; rb += 4 * kStage) {
      float2 v[kStage];
#pragma unroll
      for (int k = 0; k < kStage; ++k) {
        const int r = rb + 4 * k;
        if (r <= rhi) v[k] = z[(size_t)(wa0 + r) * ns + col];
      }
#pragma unroll
      for (int k = 0; k < kStage; ++k) {
        const int r = rb + 4 * k;
        if (r <= rhi) win[r * W + (e - r * dk)] = deramp_next(q, v[k]);
      }
    }
  }
}

// sinc_interp_raw out of line, for the rare tiles whose window exceeds the
// LDS budget (one copy shared by every call site).
__device__ __attribute__((noinline)) float2 sinc_interp_raw_global(const GeomDev &g,
                                                                   const float2 *__restrict__ z,
                                                                   int ia, int ir, float fr,
                                                                   float fa, float sp, float cp) {
  return sinc_interp_raw(g, z, ia, ir, fr, fa, sp, cp);
}

// Integer wavefront min / max with the result in every lane.
__device__ __forceinline__ int wave_min_i(int v) {
  return wave_reduce(v, [](int a, int b) { return min(a, b); });
}
__device__ __forceinline__ int wave_max_i(int v) {
  return wave_reduce(v, [](int a, int b) { return max(a, b); });
}

// Tile classes counted into the optional stats array (validation only):
// [0] tiles, [1] burst-tiles skipped by the corner test, [2] burst-tiles
// interpolated from the LDS window, [3] burst-tiles read from global memory
// (window larger than kPairWin), [4] burst-tiles not skipped whose pixels
// all lie outside the validity window, [5] raw burst-tiles interpolated
// from a sheared LDS window.
enum { kStatTiles, kStatSkipped, kStatStaged, kStatGlobal, kStatEmpty, kStatSheared, kStatCount };
static_assert(kStatCount == S1IFG_TILE_STATS, "tile classes of s1ifg_ctx_set_tile_stats");

// One tile (tile_y, tile_x) of look cells.  All 256 threads of the block
// call it (block barriers); rl * cl <= kPairMaxPix.  A block that runs
// another tile afterwards must pass a block barrier first (L is reused).
template <int NP, int CL, bool RAW = false>
__device__ __forceinline__ void pair_tile_sinc(
    const GeomDev &G0, const GeomDev &G1, const OffPoly &op1, const GridDev &gr,
    const DemDev &dem, const RowTab *rows, const ColTab *cols, const float *dh,
    const float2 *__restrict__ z0, const float2 *__restrict__ z1, int rl_, int cl_, int tile_y,
    int tile_x, int nr_ml, int nc_ml, float2 *ifg, float *pr, float *ps, PairLds &L,
    unsigned long long *stats = nullptr) {
  const int cl = CL ? CL : cl_;
  const int rl = NP ? NP / CL : rl_;
  const OffPoly none = {{0, 0, 0, 0, 0, 0}, 0};
  const int tid = threadIdx.x;
  const int cm = tile_x * kPairTile + (tid & (kPairTile - 1));
  const int rm = tile_y * kPairTile + (tid / kPairTile);
  const bool active = cm < nc_ml && rm < nr_ml;
  const int np = NP ? NP : rl * cl;
  if (tid < 8) L.box[tid >> 2][tid & 3] = (tid & 3) < 2 ? INT_MAX : INT_MIN;
  if (tid < 4) L.sbox[tid >> 1][tid & 1] = (tid & 1) ? INT_MIN : INT_MAX;
  // wavefront b < 2: is the whole tile outside burst b?
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (wave < 2) {
    const int nrt = min(nr_ml - tile_y * kPairTile, kPairTile);
    const int nct = min(nc_ml - tile_x * kPairTile, kPairTile);
    const int r0 = tile_y * kPairTile * rl, c0 = tile_x * kPairTile * cl;
    int sk = 0;
    const bool o = tile_outside<S1IFG_INTERP_SINC>(wave ? G1 : G0, wave ? op1 : none, dem, rows,
                                                   cols, dh, gr.ncol, r0, r0 + nrt * rl - 1, c0,
                                                   c0 + nct * cl - 1, RAW ? &sk : nullptr);
    if ((tid & 63) == 0) {
      L.out[wave] = o;
      L.skew[wave] = sk;
    }
  }
  __syncthreads();  // boxes initialised, L.out and L.skew set
  const bool out0 = L.out[0], out1 = L.out[1];
  const int sk0 = __builtin_amdgcn_readfirstlane(L.skew[0]);
  const int sk1 = __builtin_amdgcn_readfirstlane(L.skew[1]);
  if (stats && tid == 0) {
    atomicAdd(&stats[kStatTiles], 1ull);
    atomicAdd(&stats[kStatSkipped], (unsigned long long)(out0 + out1));
  }
  if (out0 && out1) {
    // no pixel of the tile in either burst: the products are zero
    if (active) {
      size_t o = (size_t)rm * nc_ml + cm;
      ifg[o] = make_float2(0.f, 0.f);
      if (pr) pr[o] = 0.f;
      if (ps) ps[o] = 0.f;
    }
    return;
  }
  // Burst 0: geometry of the cell's pixels (rows outer, columns inner),
  // their tap window staged in LDS, interpolation.  The pixel heights are
  // kept for burst 1.
  double hp[kPairMaxPix];
  float2 v0[kPairMaxPix];
  {
    int ir[kPairMaxPix], ia[kPairMaxPix];
    float fr[kPairMaxPix], fa[kPairMaxPix], sn[kPairMaxPix], cs[kPairMaxPix];
    unsigned vmask = 0;
    int lo_r = INT_MAX, lo_a = INT_MAX, hi_r = INT_MIN, hi_a = INT_MIN;
    ZdState zs;  // pixel 0's root: the cell's other pixels start from it
#pragma unroll
    for (int p = 0; p < kPairMaxPix; ++p) {
      ir[p] = ia[p] = 0;
      fr[p] = fa[p] = sn[p] = cs[p] = 0.f;
      hp[p] = 0.0;
      if (p < np && active) {
        const int i = p / cl, j = p - i * cl;
        const int r = rm * rl + i, c = cm * cl + j;
        const RowTab rt = rows[r];
        const ColTab ct = cols[c];
        hp[p] = dem_height(dem, rt.fy, ct.fx);
        if (dh) hp[p] += (double)dh[(size_t)r * gr.ncol + c];
        double X[3];
        target_xyz(rt, ct, hp[p], X);
        double t, dr[3], rng, rg, az, drg, daz;
        if (p == 0)
          zero_doppler_state(G0, X, t, dr, zs);
        else
          zero_doppler_near(G0, X, zs, t, dr);
        sample_position(G0, none, t, dr, rng, rg, az, drg, daz);
        float s, cc;
        if (!out0 && sample_valid<S1IFG_INTERP_SINC>(G0, rg, az)) {
          const int r0 = (int)floor(rg), a0 = (int)floor(az);
          reramp_phase(G0, rng, rg, r0, daz, t, s, cc);
          ir[p] = r0;
          ia[p] = a0;
          fr[p] = (float)(rg - r0);
          fa[p] = (float)(az - a0);
          sn[p] = s;
          cs[p] = cc;
          vmask |= 1u << p;
          lo_r = min(lo_r, r0);
          hi_r = max(hi_r, r0);
          lo_a = min(lo_a, a0);
          hi_a = max(hi_a, a0);
        }
      }
    }
    {
      const int w0 = wave_min_i(lo_r), w1 = wave_min_i(lo_a);
      const int w2 = wave_max_i(hi_r), w3 = wave_max_i(hi_a);
      if ((tid & 63) == 0) {
        atomicMin(&L.box[0][0], w0);
        atomicMin(&L.box[0][1], w1);
        atomicMax(&L.box[0][2], w2);
        atomicMax(&L.box[0][3], w3);
      }
    }
    __syncthreads();
    // stage the tap window (columns min-3 .. max+4, rows likewise)
    const int mr = L.box[0][0], ma = L.box[0][1], xr = L.box[0][2], xa = L.box[0][3];
    const bool any = mr <= xr;
    const long W = any ? (long)xr - mr + 8 : 0, H = any ? (long)xa - ma + 8 : 0;
    const int wa0 = ma - 3;
    // the rectangular window; on raw bursts, when it exceeds kPairWin, the
    // window sheared by the tile's skew (row a holds columns lo + a sk .. hi
    // + a sk, lo and hi the extremes of the taps' columns less their rows'
    // shear)
    bool staged = any && W * H <= kPairWin;
    int wc = (int)W, wr0 = mr - 3, dk = 0;
    if (RAW && any && !staged && sk0 != 0) {  // block-uniform
      int lo_s = INT_MAX, hi_s = INT_MIN;
#pragma unroll
      for (int p = 0; p < kPairMaxPix; ++p) {
        if (vmask & (1u << p)) {
          const int ta = (ia[p] - 3) * sk0, tb = (ia[p] + 4) * sk0;
          lo_s = min(lo_s, ir[p] - 3 - max(ta, tb));
          hi_s = max(hi_s, ir[p] + 4 - min(ta, tb));
        }
      }
      const int w4 = wave_min_i(lo_s), w5 = wave_max_i(hi_s);
      if ((tid & 63) == 0) {
        atomicMin(&L.sbox[0][0], w4);
        atomicMax(&L.sbox[0][1], w5);
      }
      __syncthreads();
      const long S = (long)L.sbox[0][1] - L.sbox[0][0] + 1;
      if (S * H <= kPairWin) {
        staged = true;
        dk = sk0;
        wc = (int)S;
        wr0 = L.sbox[0][0] + wa0 * dk;
      }
    }
    const int ld = wc - dk;  // LDS index step of a tap row
    if (stats && tid == 0 && !out0)
      atomicAdd(&stats[staged ? (dk ? kStatSheared : kStatStaged) : any ? kStatGlobal : kStatEmpty],
                1ull);
    if (staged && RAW) {
      if (dk)
        stage_window_deramp_skew(z0, G0, wa0, wr0, wc, (int)H, dk, L.win);
      else
        stage_window_deramp(z0, G0, wa0, wr0, wc, (int)H, L.win);
    }
    if (staged && !RAW) stage_window(z0, G0.samples, wa0, wr0, wc, (int)H, L.win);
    __syncthreads();
#pragma unroll
    for (int p = 0; p < kPairMaxPix; ++p) {
      v0[p] = make_float2(0.f, 0.f);
      if (p < np && (vmask & (1u << p))) {
        if (staged)
          v0[p] = sinc_interp(&L.win[(ia[p] - 3 - wa0) * ld + (ir[p] - 3 - wr0)], ld, fr[p],
                              fa[p], sn[p], cs[p]);
        else if (RAW)
          v0[p] = sinc_interp_raw_global(G0, z0, ia[p], ir[p], fr[p], fa[p], sn[p], cs[p]);
        else
          v0[p] = sinc_interp(z0 + (size_t)(ia[p] - 3) * G0.samples + (ir[p] - 3), G0.samples,
                              fr[p], fa[p], sn[p], cs[p]);
      }
    }
  }
  // Burst 1: geometry, its tap window in LDS once every thread is done
  // with burst 0's, interpolation
  float2 v1[kPairMaxPix];
  {
    int ir[kPairMaxPix], ia[kPairMaxPix];
    float fr[kPairMaxPix], fa[kPairMaxPix], sn[kPairMaxPix], cs[kPairMaxPix];
    unsigned vmask = 0;
    int lo_r = INT_MAX, lo_a = INT_MAX, hi_r = INT_MIN, hi_a = INT_MIN;
    ZdState zs;  // pixel 0's root: the cell's other pixels start from it
#pragma unroll
    for (int p = 0; p < kPairMaxPix; ++p) {
      ir[p] = ia[p] = 0;
      fr[p] = fa[p] = sn[p] = cs[p] = 0.f;
      if (p < np && active && !out1) {
        const int i = p / cl, j = p - i * cl;
        double X[3];
        target_xyz(rows[rm * rl + i], cols[cm * cl + j], hp[p], X);
        double t, dr[3], rng, rg, az, drg, daz;
        if (p == 0)
          zero_doppler_state(G1, X, t, dr, zs);
        else
          zero_doppler_near(G1, X, zs, t, dr);
        sample_position(G1, op1, t, dr, rng, rg, az, drg, daz);
        float s, cc;
        if (sample_valid<S1IFG_INTERP_SINC>(G1, rg, az)) {
          const int r0 = (int)floor(rg), a0 = (int)floor(az);
          reramp_phase(G1, rng, rg, r0, daz, t, s, cc);
          ir[p] = r0;
          ia[p] = a0;
          fr[p] = (float)(rg - r0);
          fa[p] = (float)(az - a0);
          sn[p] = s;
          cs[p] = cc;
          vmask |= 1u << p;
          lo_r = min(lo_r, r0);
          hi_r = max(hi_r, r0);
          lo_a = min(lo_a, a0);
          hi_a = max(hi_a, a0);
        }
      }
    }
    {
      const int w0 = wave_min_i(lo_r), w1 = wave_min_i(lo_a);
      const int w2 = wave_max_i(hi_r), w3 = wave_max_i(hi_a);
      if ((tid & 63) == 0) {
        atomicMin(&L.box[1][0], w0);
        atomicMin(&L.box[1][1], w1);
        atomicMax(&L.box[1][2], w2);
        atomicMax(&L.box[1][3], w3);
      }
    }
    __syncthreads();  // burst 0's window read by every thread, burst 1's box complete
    // stage the tap window (columns min-3 .. max+4, rows likewise)
    const int mr = L.box[1][0], ma = L.box[1][1], xr = L.box[1][2], xa = L.box[1][3];
    const bool any = mr <= xr;
    const long W = any ? (long)xr - mr + 8 : 0, H = any ? (long)xa - ma + 8 : 0;
    const int wa0 = ma - 3;
    // the rectangular window; on raw bursts, when it exceeds kPairWin, the
    // window sheared by the tile's skew (row a holds columns lo + a sk .. hi
    // + a sk, lo and hi the extremes of the taps' columns less their rows'
    // shear)
    bool staged = any && W * H <= kPairWin;
    int wc = (int)W, wr0 = mr - 3, dk = 0;
    if (RAW && any && !staged && sk1 != 0) {  // block-uniform
      int lo_s = INT_MAX, hi_s = INT_MIN;
#pragma unroll
      for (int p = 0; p < kPairMaxPix; ++p) {
        if (vmask & (1u << p)) {
          const int ta = (ia[p] - 3) * sk1, tb = (ia[p] + 4) * sk1;
          lo_s = min(lo_s, ir[p] - 3 - max(ta, tb));
          hi_s = max(hi_s, ir[p] + 4 - min(ta, tb));
        }
      }
      const int w4 = wave_min_i(lo_s), w5 = wave_max_i(hi_s);
      if ((tid & 63) == 0) {
        atomicMin(&L.sbox[1][0], w4);
        atomicMax(&L.sbox[1][1], w5);
      }
      __syncthreads();
      const long S = (long)L.sbox[1][1] - L.sbox[1][0] + 1;
      if (S * H <= kPairWin) {
        staged = true;
        dk = sk1;
        wc = (int)S;
        wr0 = L.sbox[1][0] + wa0 * dk;
      }
    }
    const int ld = wc - dk;  // LDS index step of a tap row
    if (stats && tid == 0 && !out1)
      atomicAdd(&stats[staged ? (dk ? kStatSheared : kStatStaged) : any ? kStatGlobal : kStatEmpty],
                1ull);
    if (staged && RAW) {
      if (dk)
        stage_window_deramp_skew(z1, G1, wa0, wr0, wc, (int)H, dk, L.win);
      else
        stage_window_deramp(z1, G1, wa0, wr0, wc, (int)H, L.win);
    }
    if (staged && !RAW) stage_window(z1, G1.samples, wa0, wr0, wc, (int)H, L.win);
    __syncthreads();
#pragma unroll
    for (int p = 0; p < kPairMaxPix; ++p) {
      v1[p] = make_float2(0.f, 0.f);
      if (p < np && (vmask & (1u << p))) {
        if (staged)
          v1[p] = sinc_interp(&L.win[(ia[p] - 3 - wa0) * ld + (ir[p] - 3 - wr0)], ld, fr[p],
                              fa[p], sn[p], cs[p]);
        else if (RAW)
          v1[p] = sinc_interp_raw_global(G1, z1, ia[p], ir[p], fr[p], fa[p], sn[p], cs[p]);
        else
          v1[p] = sinc_interp(z1 + (size_t)(ia[p] - 3) * G1.samples + (ir[p] - 3), G1.samples,
                              fr[p], fa[p], sn[p], cs[p]);
      }
    }
  }
  // the look means in the order of pair_ml_pixel
  float ax = 0.f, ay = 0.f, apr = 0.f, aps = 0.f;
  float sx = 0.f, sy = 0.f, spr = 0.f, sps = 0.f;
  const float inv_cl = 1.0f / cl, inv_rl = 1.0f / rl;
#pragma unroll
  for (int p = 0; p < kPairMaxPix; ++p) {
    if (p < np) {
      const float2 a = v0[p], c = v1[p];
      sx += a.x * c.x + a.y * c.y;
      sy += a.x * c.y - a.y * c.x;
      spr += a.x * a.x + a.y * a.y;
      sps += c.x * c.x + c.y * c.y;
      const int j = p - (p / cl) * cl;
      if (j == cl - 1) {
        ax += sx * inv_cl;
        ay += sy * inv_cl;
        apr += spr * inv_cl;
        aps += sps * inv_cl;
        sx = sy = spr = sps = 0.f;
      }
    }
  }
  if (active) {
    size_t o = (size_t)rm * nc_ml + cm;
    ifg[o] = make_float2(ax * inv_rl, ay * inv_rl);
    if (pr) pr[o] = apr * inv_rl;
    if (ps) ps[o] = aps * inv_rl;
  }
}

}  // namespace s1ifg
