// Device code of the coherence estimators shared by the per-image and
// batched kernels.
#pragma once

#include "s1ifg_internal.h"

namespace s1ifg {

__device__ __forceinline__ float hash_uniform(uint32_t seed, uint32_t i) {
  uint32_t x = i * 0x9E3779B9u ^ seed;
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return (x >> 8) * (1.0f / 16777216.0f);
}

// Values > 1 (coherence.py:143-144 fills them with unseeded random numbers).
__device__ __forceinline__ float coh_fill(float g, int fill, uint32_t seed, size_t o) {
  if (g > 1.0f) {
    if (fill == S1IFG_COH_FILL_NAN) return __builtin_nanf("");
    if (fill == S1IFG_COH_FILL_HASH) return hash_uniform(seed, (uint32_t)o);
    return 1.0f;
  }
  return g;
}

constexpr int kCohTX = 64, kCohTY = 4, kCohMaxHalo = 7;
constexpr int kCohW = kCohTX + 2 * kCohMaxHalo, kCohH = kCohTY + 2 * kCohMaxHalo;

struct CohTile {
  float sx[kCohH][kCohW], sy[kCohH][kCohW], s1[kCohH][kCohW], s2[kCohH][kCohW];
};

// Window sums of four channels (a complex, b, c) over a win x win window
// centred on each pixel of one 64 x 4 output tile (block 64 x 4): the tile
// and its halo are staged in LDS; sums over the window rows outer, columns
// inner, zero outside the image.  Returns whether the thread's pixel
// (r, c) is inside the image.  Contains a block barrier; callers pass
// another before L is reused.
__device__ __forceinline__ bool window_sums_tile(const float2 *__restrict__ za,
                                                 const float *__restrict__ zb,
                                                 const float *__restrict__ zc, int nr, int nc,
                                                 int win, int tile_x, int tile_y, CohTile &L,
                                                 int &r, int &c, float &ax, float &ay, float &a1,
                                                 float &a2) {
  const int h = win / 2;
  const int tw = kCohTX + 2 * h, th = kCohTY + 2 * h;
  const int c0 = tile_x * kCohTX - h, r0 = tile_y * kCohTY - h;
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  for (int i = tid; i < tw * th; i += blockDim.x * blockDim.y) {
    int ty = i / tw, tx = i - ty * tw;
    int rr = r0 + ty, cc = c0 + tx;
    float a = 0.f, b = 0.f, u = 0.f, v = 0.f;
    if (rr >= 0 && rr < nr && cc >= 0 && cc < nc) {
      size_t o = (size_t)rr * nc + cc;
      float2 z = za[o];
      a = z.x;
      b = z.y;
      u = zb[o];
      v = zc[o];
    }
    L.sx[ty][tx] = a;
    L.sy[ty][tx] = b;
    L.s1[ty][tx] = u;
    L.s2[ty][tx] = v;
  }
  __syncthreads();
  c = tile_x * kCohTX + threadIdx.x;
  r = tile_y * kCohTY + threadIdx.y;
  ax = ay = a1 = a2 = 0.f;
  if (r >= nr || c >= nc) return false;
  for (int i = 0; i < win; ++i)
    for (int j = 0; j < win; ++j) {
      int ty = threadIdx.y + i, tx = threadIdx.x + j;
      ax += L.sx[ty][tx];
      ay += L.sy[ty][tx];
      a1 += L.s1[ty][tx];
      a2 += L.s2[ty][tx];
    }
  return true;
}

// Windowed coherence of one 64 x 4 output tile (block 64 x 4).  Contains
// block barriers.
__device__ __forceinline__ void coherence_tile(const float2 *__restrict__ ifg,
                                               const float *__restrict__ p1,
                                               const float *__restrict__ p2, int nr, int nc,
                                               int win, int fill, uint32_t seed,
                                               float *__restrict__ coh, int tile_x, int tile_y,
                                               CohTile &L) {
  int r, c;
  float ax, ay, a1, a2;
  if (window_sums_tile(ifg, p1, p2, nr, nc, win, tile_x, tile_y, L, r, c, ax, ay, a1, a2)) {
    float den = sqrtf(a1 * a2);
    float g = den > 0.f ? sqrtf(ax * ax + ay * ay) / den : 0.f;
    size_t o = (size_t)r * nc + c;
    coh[o] = coh_fill(g, fill, seed, o);
  }
  __syncthreads();
}

// Window sums of four channels (a complex, b, c) over a win x win window
// centred on each pixel of a strip of 64 columns x kStripRows rows (block
// 64 x 4), for win <= 2 kStripMaxHalo + 1, zero outside the image.  The
// sums are separable: the strip's input rows are read once, 4 at a time,
// into LDS; each row's horizontal win-sums go into a ring of rows; each
// output pixel adds the win row sums above and below it, and f(r, c, sa,
// sb, s1, s2) gets the pixel's sums.  The loads of the next 4 input rows
// are issued before the current rows' sums (they are in flight while the
// block sums).  Contains block barriers (uniform); the last one lets the
// caller reuse L.
constexpr int kStripRows = 32, kStripMaxHalo = 6, kStripRing = 16;
static_assert(kStripRing >= kCohTY + 2 * kStripMaxHalo, "ring holds a window of rows");

struct CohStrip {
  float in[4][kCohTY][kCohTX + 2 * kStripMaxHalo];  // channels x 4 input rows x columns
  float hs[4][kStripRing][kCohTX];                  // horizontal sums, ring of rows
};

template <typename F>
__device__ __forceinline__ void window_sums_strip(const float2 *__restrict__ za,
                                                  const float *__restrict__ zb,
                                                  const float *__restrict__ zc, int nr, int nc,
                                                  int win, int strip_x, int strip_y, CohStrip &L,
                                                  F f) {
  const int h = win / 2, w2 = 2 * h;
  const int x = threadIdx.x, y = threadIdx.y;
  const int c0 = strip_x * kCohTX, rb = strip_y * kStripRows;
  const int re = min(rb + kStripRows, nr);
  // this thread's input columns i = x and x + kCohTX (the latter when
  // below kCohTX + w2), of input row base + y, into registers
  static_assert(kCohTX >= 2 * kStripMaxHalo, "two columns per thread cover the halo");
  float va[2], vb[2], vu[2], vv[2];
  const auto load_rows = [&](int base) {
    const int rr = base + y;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int i = x + k * kCohTX, cc = c0 - h + i;
      va[k] = vb[k] = vu[k] = vv[k] = 0.f;
      if (i < kCohTX + w2 && rr >= 0 && rr < nr && cc >= 0 && cc < nc) {
        const size_t o = (size_t)rr * nc + cc;
        const float2 z = za[o];
        va[k] = z.x;
        vb[k] = z.y;
        vu[k] = zb[o];
        vv[k] = zc[o];
      }
    }
  };
  int nl = rb - h;  // next input row to stage
  // the last input row staged: that of the last output row group
  const int last = rb + (re - 1 - rb) / kCohTY * kCohTY + kCohTY - 1 + h;
  load_rows(nl);
  for (int r0 = rb; r0 < re; r0 += kCohTY) {
    // horizontal sums of the input rows up to r0 + kCohTY - 1 + h
    while (nl <= r0 + kCohTY - 1 + h) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int i = x + k * kCohTX;
        if (i < kCohTX + w2) {
          L.in[0][y][i] = va[k];
          L.in[1][y][i] = vb[k];
          L.in[2][y][i] = vu[k];
          L.in[3][y][i] = vv[k];
        }
      }
      __syncthreads();
      if (nl + kCohTY <= last) load_rows(nl + kCohTY);  // in flight during the sums
      float sa = 0.f, sb = 0.f, su = 0.f, sv = 0.f;
      for (int j = 0; j <= w2; ++j) {
        sa += L.in[0][y][x + j];
        sb += L.in[1][y][x + j];
        su += L.in[2][y][x + j];
        sv += L.in[3][y][x + j];
      }
      const int slot = (nl + y) & (kStripRing - 1);
      L.hs[0][slot][x] = sa;
      L.hs[1][slot][x] = sb;
      L.hs[2][slot][x] = su;
      L.hs[3][slot][x] = sv;
      __syncthreads();
      nl += kCohTY;
    }
    const int r = r0 + y, c = c0 + x;
    if (r < re && c < nc) {
      float ax = 0.f, ay = 0.f, a1 = 0.f, a2 = 0.f;
      for (int i = 0; i <= w2; ++i) {
        const int slot = (r - h + i) & (kStripRing - 1);
        ax += L.hs[0][slot][x];
        ay += L.hs[1][slot][x];
        a1 += L.hs[2][slot][x];
        a2 += L.hs[3][slot][x];
      }
      f(r, c, ax, ay, a1, a2);
    }
  }
  __syncthreads();
}

// Windowed coherence of one strip (window_sums_strip).
__device__ __forceinline__ void coherence_strip(const float2 *__restrict__ ifg,
                                                const float *__restrict__ p1,
                                                const float *__restrict__ p2, int nr, int nc,
                                                int win, int fill, uint32_t seed,
                                                float *__restrict__ coh, int strip_x, int strip_y,
                                                CohStrip &L) {
  window_sums_strip(ifg, p1, p2, nr, nc, win, strip_x, strip_y, L,
                    [&](int r, int c, float ax, float ay, float a1, float a2) {
                      const float den = sqrtf(a1 * a2);
                      const float g = den > 0.f ? sqrtf(ax * ax + ay * ay) / den : 0.f;
                      const size_t o = (size_t)r * nc + c;
                      coh[o] = coh_fill(g, fill, seed, o);
                    });
}

// Look-cell estimator (coherence.py:142): |ifg| / (sqrt(p1)+eps) / (sqrt(p2)+eps).
__device__ __forceinline__ float coherence_cell(float2 v, float p1, float p2) {
  float a1 = sqrtf(p1) + 1e-12f, a2 = sqrtf(p2) + 1e-12f;
  return sqrtf(v.x * v.x + v.y * v.y) / a1 / a2;
}

}  // namespace s1ifg
