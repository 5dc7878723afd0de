// Goldstein overlap-add without a patch buffer (shared by the per-image and
// batched kernels).
//
// Work item = one patch row pi (a band of 32 image rows) x one chunk of
// kBandCW consecutive patch columns, processed by one wavefront: it filters
// the chunk's patches in column order on MFMA (gold_mfma_dev.h) and adds
// each weighted patch straight from registers into its own 32 x 64 circular
// LDS buffer holding the band's running overlap-add.  After patch j, the
// columns left of patch j+1's origin are final for the (band, chunk) and are
// written to a partial buffer of 32 rows x kBandCols per item.  The output
// pixel (y, x) is the sum of the partials of the (at most 4) bands and (at
// most 2) chunks that cover it, in increasing band then chunk order, over
// the triangular weight sum (goldstein.cu:257-292).  Every sum has a fixed
// order, so results are bit-reproducible; the partials cost 4 x (1 + 24 /
// (8 kBandCW)) times the output in traffic instead of the 16 x of a patch
// buffer.
#pragma once

#include "gold_dev.h"
#include "gold_mfma_dev.h"

namespace s1ifg {

constexpr int kBandCW = 16;                    // patches per chunk
constexpr int kBandCols = 8 * kBandCW + 24;    // partial width (columns)
constexpr int kBandBuf = 64;                   // circular buffer width
constexpr int kBandLd = kBandBuf + 1;          // padded row stride
constexpr int kBandWaves = 4;                  // wavefronts per block

struct BandLds {
  float2 buf[kBandWaves][kGoldN * kBandLd];
};

__device__ __forceinline__ int gold_origin(int j, int n) { return min(j * (kGoldN / 4), n - kGoldN); }

__host__ __device__ __forceinline__ int band_nchunk(int pj_n) { return (pj_n + kBandCW - 1) / kBandCW; }

// One (band, chunk) item on one wavefront (no block barriers).  buf: the
// wave's LDS buffer (kGoldN x kBandLd); partial: this image's partial buffer
// [pi_n][nchunk][32][kBandCols].
__device__ __forceinline__ void band_item(const mf::Consts &K, const float2 *__restrict__ img,
                                          int nrow, int ncol, int pi_n, int pj_n, int pi, int c,
                                          float alpha, float2 *__restrict__ partial,
                                          float2 *buf) {
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5, n = lane & 31;
  const int nchunk = band_nchunk(pj_n);
  const int j0 = c * kBandCW;
  const int j1 = min(j0 + kBandCW, pj_n);
  const int ox0 = gold_origin(j0, ncol);
  float2 *dst = partial + (size_t)(pi * nchunk + c) * kGoldN * kBandCols;
  for (int i = lane; i < kGoldN * kBandLd; i += 64) buf[i] = make_float2(0.f, 0.f);
  __builtin_amdgcn_wave_barrier();
  int lo = ox0;
  const long rowbase = (long)pi * pj_n;
  for (int j = j0; j < j1; ++j) {
    float2 x[16];
    mf::load_patch(img, nrow, ncol, pj_n, rowbase + j, h, n, x);
    float2 out[16];
    if (alpha == 0.5f)
      mf::filter_patch_regs<true>(K, x, h, n, alpha, out);
    else
      mf::filter_patch_regs<false>(K, x, h, n, alpha, out);
    // add the patch (each lane owns distinct (row, column) entries)
    const int ox = gold_origin(j, ncol);
    float2 *row = buf + n * kBandLd;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        float2 *e = row + ((ox + 4 * h + 8 * q + t) & (kBandBuf - 1));
        float2 v = *e;
        v.x += out[4 * q + t].x;
        v.y += out[4 * q + t].y;
        *e = v;
      }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // columns no later patch of the chunk reaches are final
    const int done = (j + 1 < j1) ? gold_origin(j + 1, ncol) : ox + kGoldN;
    const int w = done - lo;
    for (int i = lane; i < w * kGoldN; i += 64) {
      const int r = i / w, col = lo + (i - r * w);
      float2 *e = buf + r * kBandLd + (col & (kBandBuf - 1));
      dst[(size_t)r * kBandCols + (col - ox0)] = *e;
      *e = make_float2(0.f, 0.f);
    }
    lo = done;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Output pixel (y, x): partial sums of the bands and chunks covering it
// over the triangular weight sum; 0 where the weight sum is 0.
__device__ __forceinline__ float2 band_sum_pixel(const float2 *__restrict__ partial, int nrow,
                                                 int ncol, int pi_n, int pj_n, int y, int x) {
  const int step = kGoldN / 4;
  const int nchunk = band_nchunk(pj_n);
  // weight sum: sum over covering patches (pi, pj) of wx + wy
  float cy = 0.f, sy = 0.f, cx = 0.f, sx = 0.f;
  const int pi_lo = y - (kGoldN - 1) > 0 ? (y - (kGoldN - 1) + step - 1) / step : 0;
  const int pj_lo = x - (kGoldN - 1) > 0 ? (x - (kGoldN - 1) + step - 1) / step : 0;
  for (int pi = pi_lo; pi < pi_n; ++pi) {
    const int o = gold_origin(pi, nrow);
    if (o > y) break;
    if (y >= o + kGoldN) continue;
    const int ty = y - o;
    cy += 1.f;
    sy += (float)(ty < kGoldN / 2 ? ty : kGoldN - 1 - ty);
  }
  for (int pj = pj_lo; pj < pj_n; ++pj) {
    const int o = gold_origin(pj, ncol);
    if (o > x) break;
    if (x >= o + kGoldN) continue;
    const int tx = x - o;
    cx += 1.f;
    sx += (float)(tx < kGoldN / 2 ? tx : kGoldN - 1 - tx);
  }
  const float ws = cx * sy + cy * sx;
  if (!(ws > 0.f)) return make_float2(0.f, 0.f);
  // last patch column with origin <= x, and the chunks that can cover x
  const int jx = x >= gold_origin(pj_n - 1, ncol) ? pj_n - 1 : min(x / step, pj_n - 1);
  const int c_hi = jx / kBandCW;
  const int c_lo = c_hi > 0 ? c_hi - 1 : 0;
  float ax = 0.f, ay = 0.f;
  for (int pi = pi_lo; pi < pi_n; ++pi) {
    const int o = gold_origin(pi, nrow);
    if (o > y) break;
    if (y >= o + kGoldN) continue;
    const int ty = y - o;
    for (int c = c_lo; c <= c_hi && c < nchunk; ++c) {
      const int a = gold_origin(c * kBandCW, ncol);
      const int b = gold_origin(min(c * kBandCW + kBandCW, pj_n) - 1, ncol) + kGoldN;
      if (x < a || x >= b) continue;
      const float2 v =
          partial[((size_t)(pi * nchunk + c) * kGoldN + ty) * kBandCols + (x - a)];
      ax += v.x;
      ay += v.y;
    }
  }
  return make_float2(ax / ws, ay / ws);
}

// band_sum_pixel over a tile of kSumTX x kSumTY output pixels (block
// kSumTX x kSumTY): the factors that depend only on the column (the column
// weight terms and the <= 2 chunks covering it) or only on the row (the row
// weight terms and the <= 5 bands covering it: 4, plus the last band whose
// origin is clamped to nrow - 32) are computed once per tile
// column / row into LDS; every pixel then adds its partials in the same
// order as band_sum_pixel, so the result is the same.  Contains block
// barriers (the last one lets the caller reuse L).
constexpr int kSumTX = 64, kSumTY = 4;

struct BandSumLds {
  float cx[kSumTX], sx[kSumTX];
  int nch[kSumTX], ch[kSumTX][2], xo[kSumTX][2];
  float cy[kSumTY], sy[kSumTY];
  int nb[kSumTY], pi[kSumTY][5], ty[kSumTY][5];
};

__device__ __forceinline__ void band_sum_tile(const float2 *__restrict__ partial, int nrow,
                                              int ncol, int pi_n, int pj_n, int tile_x,
                                              int tile_y, float2 *__restrict__ out,
                                              BandSumLds &L) {
  const int step = kGoldN / 4;
  const int nchunk = band_nchunk(pj_n);
  const int tx = threadIdx.x, tyi = threadIdx.y;
  const int x = tile_x * kSumTX + tx, y = tile_y * kSumTY + tyi;
  if (tyi == 0) {
    // column terms of x
    float cx = 0.f, sx = 0.f;
    int n = 0;
    if (x < ncol) {
      const int pj_lo = x - (kGoldN - 1) > 0 ? (x - (kGoldN - 1) + step - 1) / step : 0;
      for (int pj = pj_lo; pj < pj_n; ++pj) {
        const int o = gold_origin(pj, ncol);
        if (o > x) break;
        if (x >= o + kGoldN) continue;
        const int t = x - o;
        cx += 1.f;
        sx += (float)(t < kGoldN / 2 ? t : kGoldN - 1 - t);
      }
      const int jx = x >= gold_origin(pj_n - 1, ncol) ? pj_n - 1 : min(x / step, pj_n - 1);
      const int c_hi = jx / kBandCW;
      const int c_lo = c_hi > 0 ? c_hi - 1 : 0;
      for (int c = c_lo; c <= c_hi && c < nchunk; ++c) {
        const int a = gold_origin(c * kBandCW, ncol);
        const int b = gold_origin(min(c * kBandCW + kBandCW, pj_n) - 1, ncol) + kGoldN;
        if (x < a || x >= b) continue;
        L.ch[tx][n] = c;
        L.xo[tx][n] = x - a;
        ++n;
      }
    }
    L.cx[tx] = cx;
    L.sx[tx] = sx;
    L.nch[tx] = n;
  } else if (tyi == 1 && tx < kSumTY) {
    // row terms of row tile_y * kSumTY + tx
    const int yy = tile_y * kSumTY + tx;
    float cy = 0.f, sy = 0.f;
    int n = 0;
    if (yy < nrow) {
      const int pi_lo = yy - (kGoldN - 1) > 0 ? (yy - (kGoldN - 1) + step - 1) / step : 0;
      for (int p = pi_lo; p < pi_n; ++p) {
        const int o = gold_origin(p, nrow);
        if (o > yy) break;
        if (yy >= o + kGoldN) continue;
        const int t = yy - o;
        cy += 1.f;
        sy += (float)(t < kGoldN / 2 ? t : kGoldN - 1 - t);
        if (n < 5) {  // always: origins step by 8, only the last one is clamped
          L.pi[tx][n] = p;
          L.ty[tx][n] = t;
          ++n;
        }
      }
    }
    L.cy[tx] = cy;
    L.sy[tx] = sy;
    L.nb[tx] = n;
  }
  __syncthreads();
  if (x < ncol && y < nrow) {
    const float ws = L.cx[tx] * L.sy[tyi] + L.cy[tyi] * L.sx[tx];
    float2 r = make_float2(0.f, 0.f);
    if (ws > 0.f) {
      float ax = 0.f, ay = 0.f;
      const int nb = L.nb[tyi], nch = L.nch[tx];
      for (int i = 0; i < nb; ++i) {
        const size_t rowb = (size_t)L.pi[tyi][i] * nchunk;
        const int t = L.ty[tyi][i];
        for (int k = 0; k < nch; ++k) {
          const float2 v =
              partial[((rowb + L.ch[tx][k]) * kGoldN + t) * kBandCols + L.xo[tx][k]];
          ax += v.x;
          ay += v.y;
        }
      }
      r = make_float2(ax / ws, ay / ws);
    }
    out[(size_t)y * ncol + x] = r;
  }
  __syncthreads();
}

}  // namespace s1ifg
