// Fused burst pair with sinc interpolation, burst windows staged in LDS
// (shared by the per-pair and batched kernels).
//
// A 256-thread block owns a tile of 16 x 16 look cells.  Two wavefronts
// first test whether the whole tile lies outside either burst
// (tile_outside: about half of a burst's target grid, the corners of its
// latitude / longitude box, does); a burst that the tile misses is not
// geocoded, and a tile that misses both writes zeros.  Each thread owns
// one cell (rl * cl <= kPairMaxPix target pixels): it first computes the
// FP64 geometry and phase of its pixels in both bursts and keeps the
// integer / fractional sampling positions and sin/cos of the phase in
// registers.  The block then reduces the bounding box of the 8 x 8 tap
// windows of each burst, loads that window of the deramped burst into LDS
// with row-contiguous (coalesced) loads, and every pixel is interpolated
// from LDS.  A window larger than kPairWin samples (steep terrain spreads
// the range positions) is read from global memory instead.  Both ways run
// the same arithmetic in the same order as pair_ml_pixel, so the products
// are bit-identical to it: the staging only moves data (each burst sample
// is fetched about 3 times per tile instead of 64 tap gathers per pixel).
#pragma once

#include <climits>

#include "geom_dev.h"

namespace s1ifg {

constexpr int kPairTile = 16;        // look cells per tile side
constexpr int kPairWin = 4864;       // staged samples of burst 0 (38 KB; 4 blocks per CU)
constexpr int kPairMaxPix = 4;       // target pixels per look cell on this path

struct PairLds {
  float2 win[1][kPairWin];
  int box[2][4];  // per burst: min intr, min inta, max intr, max inta
  int out[2];     // per burst: every pixel of the tile outside the burst
};

__device__ __forceinline__ int wave_min_i(int v) {
  for (int o = 32; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// One tile (tile_y, tile_x) of look cells.  All 256 threads of the block
// call it (block barriers); rl * cl <= kPairMaxPix.  A block that runs
// another tile afterwards must pass a block barrier first (L is reused).
__device__ __forceinline__ void pair_tile_sinc(
    const GeomDev &G0, const GeomDev &G1, const OffPoly &op1, const GridDev &gr,
    const DemDev &dem, const RowTab *rows, const ColTab *cols, const float *dh,
    const float2 *__restrict__ z0, const float2 *__restrict__ z1, int rl, int cl, int tile_y,
    int tile_x, int nr_ml, int nc_ml, float2 *ifg, float *pr, float *ps, PairLds &L) {
  const OffPoly none = {{0, 0, 0, 0, 0, 0}, 0};
  const int tid = threadIdx.x;
  const int cm = tile_x * kPairTile + (tid & (kPairTile - 1));
  const int rm = tile_y * kPairTile + (tid / kPairTile);
  const bool active = cm < nc_ml && rm < nr_ml;
  const int np = rl * cl;
  int ir[2][kPairMaxPix], ia[2][kPairMaxPix];
  float fr[2][kPairMaxPix], fa[2][kPairMaxPix], sn[2][kPairMaxPix], cs[2][kPairMaxPix];
  unsigned vmask = 0;
  int lo_r[2] = {INT_MAX, INT_MAX}, lo_a[2] = {INT_MAX, INT_MAX};
  int hi_r[2] = {INT_MIN, INT_MIN}, hi_a[2] = {INT_MIN, INT_MIN};
  if (tid < 8) L.box[tid >> 2][tid & 3] = (tid & 3) < 2 ? INT_MAX : INT_MIN;
  // wavefront b < 2: is the whole tile outside burst b?
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (wave < 2) {
    const int nrt = min(nr_ml - tile_y * kPairTile, kPairTile);
    const int nct = min(nc_ml - tile_x * kPairTile, kPairTile);
    const int r0 = tile_y * kPairTile * rl, c0 = tile_x * kPairTile * cl;
    const bool o = tile_outside<S1IFG_INTERP_SINC>(wave ? G1 : G0, wave ? op1 : none, dem, rows,
                                                   cols, dh, gr.ncol, r0, r0 + nrt * rl - 1, c0,
                                                   c0 + nct * cl - 1);
    if ((tid & 63) == 0) L.out[wave] = o;
  }
  __syncthreads();  // box initialised, L.out set
  const bool out0 = L.out[0], out1 = L.out[1];
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
  // geometry of the cell's pixels (rows outer, columns inner)
#pragma unroll
  for (int p = 0; p < kPairMaxPix; ++p) {
    ir[0][p] = ir[1][p] = ia[0][p] = ia[1][p] = 0;
    fr[0][p] = fr[1][p] = fa[0][p] = fa[1][p] = 0.f;
    sn[0][p] = sn[1][p] = cs[0][p] = cs[1][p] = 0.f;
    if (p < np && active) {
      const int i = p / cl, j = p - i * cl;
      double X[3];
      pixel_xyz(gr, dem, rows, cols, dh, rm * rl + i, cm * cl + j, X);
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        double rg, az, rng;
        float s, c;
        if (!(b ? out1 : out0) &&
            geocode_geom<S1IFG_INTERP_SINC>(b ? G1 : G0, b ? op1 : none, X, rg, az, rng, s, c)) {
          const int r0 = (int)floor(rg), a0 = (int)floor(az);
          ir[b][p] = r0;
          ia[b][p] = a0;
          fr[b][p] = (float)(rg - r0);
          fa[b][p] = (float)(az - a0);
          sn[b][p] = s;
          cs[b][p] = c;
          vmask |= 1u << (2 * p + b);
          lo_r[b] = min(lo_r[b], r0);
          hi_r[b] = max(hi_r[b], r0);
          lo_a[b] = min(lo_a[b], a0);
          hi_a[b] = max(hi_a[b], a0);
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int v0 = wave_min_i(lo_r[b]), v1 = wave_min_i(lo_a[b]);
    const int v2 = wave_max_i(hi_r[b]), v3 = wave_max_i(hi_a[b]);
    if ((tid & 63) == 0) {
      atomicMin(&L.box[b][0], v0);
      atomicMin(&L.box[b][1], v1);
      atomicMax(&L.box[b][2], v2);
      atomicMax(&L.box[b][3], v3);
    }
  }
  __syncthreads();
  // stage the tap windows (columns min-3 .. max+4, rows likewise)
  int wc[2], wr0[2], wa0[2];
  bool staged[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int mr = L.box[b][0], ma = L.box[b][1], xr = L.box[b][2], xa = L.box[b][3];
    const bool any = mr <= xr;
    const long W = any ? (long)xr - mr + 8 : 0, H = any ? (long)xa - ma + 8 : 0;
    // burst 0 from LDS, burst 1 by gathers through the vector L1: the
    // two memory paths run side by side
    staged[b] = b == 0 && any && W * H <= kPairWin;
    wc[b] = (int)W;
    wr0[b] = mr - 3;
    wa0[b] = ma - 3;
    if (staged[b]) {
      const GeomDev &g = b ? G1 : G0;
      const float2 *z = b ? z1 : z0;
      for (int row = tid >> 6; row < H; row += 4)
        for (int col = tid & 63; col < W; col += 64)
          L.win[0][row * W + col] = z[(size_t)(wa0[b] + row) * g.samples + wr0[b] + col];
    }
  }
  __syncthreads();
  // interpolation and look means, in the order of pair_ml_pixel
  float ax = 0.f, ay = 0.f, apr = 0.f, aps = 0.f;
  float sx = 0.f, sy = 0.f, spr = 0.f, sps = 0.f;
  const float inv_cl = 1.0f / cl, inv_rl = 1.0f / rl;
#pragma unroll
  for (int p = 0; p < kPairMaxPix; ++p) {
    if (p < np) {
      float2 v[2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        v[b] = make_float2(0.f, 0.f);
        if (vmask & (1u << (2 * p + b))) {
          if (staged[b]) {
            v[b] = sinc_interp(&L.win[0][(ia[b][p] - 3 - wa0[b]) * wc[b] + (ir[b][p] - 3 - wr0[b])],
                               wc[b], fr[b][p], fa[b][p], sn[b][p], cs[b][p]);
          } else {
            const GeomDev &g = b ? G1 : G0;
            const float2 *z = b ? z1 : z0;
            v[b] = sinc_interp(z + (size_t)(ia[b][p] - 3) * g.samples + (ir[b][p] - 3), g.samples,
                               fr[b][p], fa[b][p], sn[b][p], cs[b][p]);
          }
        }
      }
      const float2 a = v[0], c = v[1];
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
