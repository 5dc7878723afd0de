// Goldstein overlap-add without a patch buffer (shared by the per-image and
// batched kernels).
//
// Work item = a group of kBandRows consecutive patch rows (a band of up to
// kBandH image rows) x one chunk of kBandCW consecutive patch columns,
// processed by one wavefront: for each patch column of the chunk in order it
// filters the group's patches on MFMA (gold_mfma_dev.h) and adds each
// weighted patch straight from registers into its own kBandH x 32 circular
// LDS buffer holding the band's running overlap-add (a patch column's live
// span is 32 columns: everything left of the current origin is final).
// After patch column j, the columns left of column j+1's origin are final
// for the (group, chunk) and are written to a partial buffer of kBandH rows
// x kBandCols per item.  The output pixel (y, x) is the sum of the partials
// of the (at most 2) groups and (at most 2) chunks that cover it, in
// increasing group then chunk order, over the triangular weight sum
// (goldstein.cu:257-292).  Every sum has a fixed order, so results are
// bit-reproducible.  The partials cost (kBandH / (8 kBandRows)) x (1 + 24 /
// (8 kBandCW)) = 3.4 x the output in traffic (a patch buffer: 16 x; one
// patch row per item: 4.75 x).  Larger groups and chunks cut that further
// but leave too few items (C2: 93 x 98 = 9,114 for 2,048 wavefront slots)
// to balance the chip.  Patches of zeros (no-data regions) are passed over;
// an item writes its partial columns only from its first non-zero patch on
// and records where (its zero mark), so no-data regions cost neither the
// filter nor partial traffic.
#pragma once

#include <climits>

#include "gold_dev.h"
#include "gold_mfma_dev.h"

namespace s1ifg {

constexpr int kBandCW = 8;                           // patch columns per chunk
constexpr int kBandCols = 8 * kBandCW + 24;          // partial width (columns)
constexpr int kBandRows = 2;                         // patch rows per group
constexpr int kBandH = kGoldN + 8 * (kBandRows - 1); // partial height (rows)
constexpr int kBandBuf = 32;                         // circular buffer width
constexpr int kBandLd = kBandBuf + 1;                // padded row stride
constexpr int kBandWaves = 4;                        // wavefronts per block

struct BandLds {
  float2 buf[kBandWaves][kBandH * kBandLd];
  mf::SmoothLds u[kBandWaves];
};

__device__ __forceinline__ int gold_origin(int j, int n) { return min(j * (kGoldN / 4), n - kGoldN); }

__host__ __device__ __forceinline__ int band_nchunk(int pj_n) { return (pj_n + kBandCW - 1) / kBandCW; }
__host__ __device__ __forceinline__ int band_ngroup(int pi_n) {
  return (pi_n + kBandRows - 1) / kBandRows;
}
// Work items and partial-buffer elements of an image's patch grid.
__host__ __device__ __forceinline__ long band_items(int pi_n, int pj_n) {
  return (long)band_ngroup(pi_n) * band_nchunk(pj_n);
}
__host__ __device__ __forceinline__ size_t band_partial_elems(int pi_n, int pj_n) {
  return (size_t)band_items(pi_n, pj_n) * kBandH * kBandCols;
}
// The band sum addresses the partial sums with 32-bit offsets (about 3.4
// elements per image pixel: images up to ~630 million pixels).
__host__ __device__ __forceinline__ bool band_offsets_fit(int pi_n, int pj_n) {
  return band_partial_elems(pi_n, pj_n) <= (size_t)INT_MAX;
}
// A pair's Goldstein buffer: the partial sums, then (256-byte aligned) one
// int per item, its zero mark: the item's partial columns below it (from
// the item's first column) are zeros and were not written (kBandCols: the
// whole item is zeros).
__host__ __device__ __forceinline__ size_t band_zb_offset(int pi_n, int pj_n) {
  return (band_partial_elems(pi_n, pj_n) * 8 + 255) / 256 * 256;
}
__host__ __device__ __forceinline__ size_t band_buffer_bytes(int pi_n, int pj_n) {
  return band_zb_offset(pi_n, pj_n) + ((size_t)band_items(pi_n, pj_n) * 4 + 255) / 256 * 256;
}
__host__ __device__ __forceinline__ int *band_zb(float2 *partial, int pi_n, int pj_n) {
  return (int *)((char *)partial + band_zb_offset(pi_n, pj_n));
}
__host__ __device__ __forceinline__ const int *band_zb(const float2 *partial, int pi_n, int pj_n) {
  return (const int *)((const char *)partial + band_zb_offset(pi_n, pj_n));
}
// Per-image workspace: the buffer above, then (256-byte aligned) the
// work-queue counter of the persistent wavefronts.
__host__ __device__ __forceinline__ size_t band_counter_offset(int pi_n, int pj_n) {
  return band_buffer_bytes(pi_n, pj_n);
}

// Next work item of a persistent wavefront: the first nw items go one to
// each of the nw wavefronts; after that a wavefront takes the next item
// from the queue counter (zeroed before the launch), so that wavefronts
// whose items end early (patches of zeros) take more.
template <typename C>
__device__ __forceinline__ long band_next(C *ctr, long nw) {
  unsigned k = 0;
  if ((threadIdx.x & 63) == 0) k = (unsigned)atomicAdd(ctr, (C)1);
  return nw + (long)(unsigned)__builtin_amdgcn_readfirstlane((int)k);
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One (group, chunk) item on one wavefront (no block barriers).  buf: the
// wave's LDS buffer (kBandH x kBandLd); partial: this image's buffer,
// partial sums [ngroup][nchunk][kBandH][kBandCols] then the items' zero
// marks.  Until the item's first non-zero patch the overlap-add holds zeros:
// the columns that become final before it are not written, and the zero
// mark records where the written columns start.
__device__ __forceinline__ void band_item(const mf::Consts &K, const mf::SmoothLds &U,
                                          const float2 *__restrict__ img, int nrow, int ncol,
                                          int pi_n, int pj_n, int g, int c, float alpha,
                                          float2 *__restrict__ partial, float2 *buf) {
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5, n = lane & 31;
  const int nchunk = band_nchunk(pj_n);
  const int j0 = c * kBandCW;
  const int j1 = min(j0 + kBandCW, pj_n);
  const int p0 = g * kBandRows;
  const int p1 = min(p0 + kBandRows, pi_n);
  const int ox0 = gold_origin(j0, ncol);
  const int oy0 = gold_origin(p0, nrow);
  const int hrows = gold_origin(p1 - 1, nrow) - oy0 + kGoldN;  // rows the group covers
  float2 *dst = partial + ((size_t)g * nchunk + c) * kBandH * kBandCols;
  for (int i = lane; i < kBandH * kBandLd; i += 64) buf[i] = make_float2(0.f, 0.f);
  wave_sync_lds();
  int lo = ox0;
  int zb = -1;  // zero mark, set at the first non-zero patch (wavefront-uniform)
  for (int j = j0; j < j1; ++j) {
    const int ox = gold_origin(j, ncol);
    for (int p = p0; p < p1; ++p) {
      float2 x[16];
      mf::load_patch(img, nrow, ncol, pj_n, (long)p * pj_n + j, h, n, x);
      // A patch of zeros (no data: the corners of a burst's latitude /
      // longitude box, tiles outside the bursts) filters to zeros and adds
      // nothing to the overlap-add: it is passed over (wavefront-uniform
      // branch).
      if (mf::patch_is_zero(x)) continue;
      if (zb < 0) zb = lo - ox0;  // the columns flushed so far hold zeros
      float2 out[16];
      float norm;
      if (alpha == 0.5f)
        mf::filter_patch_regs<true>(K, U, x, h, n, alpha, out, norm);
      else
        mf::filter_patch_regs<false>(K, U, x, h, n, alpha, out, norm);
      // add the weighted patch (within a patch each lane owns distinct
      // entries): per group of 4 columns, the 4 reads are issued before
      // the updates, so the read-modify-write round trips to LDS overlap
      float2 *row = buf + (gold_origin(p, nrow) - oy0 + n) * kBandLd;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float2 *e[4];
        float2 v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          e[t] = row + ((ox + 4 * h + 8 * q + t) & (kBandBuf - 1));
          v[t] = *e[t];
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const float w = mf::patch_weight(h, n, q, t, norm);
          v[t].x = fmaf(out[4 * q + t].x, w, v[t].x);
          v[t].y = fmaf(out[4 * q + t].y, w, v[t].y);
          *e[t] = v[t];
        }
      }
      wave_sync_lds();  // the next patch's lanes add to these entries
    }
    // columns no later patch column of the chunk reaches are final (while
    // every patch so far was zeros they are zeros, not written)
    const int done = (j + 1 < j1) ? gold_origin(j + 1, ncol) : ox + kGoldN;
    const int w = done - lo;
    if (zb >= 0) {
      for (int i = lane; i < w * hrows; i += 64) {
        const int r = i / w, col = lo + (i - r * w);
        float2 *e = buf + r * kBandLd + (col & (kBandBuf - 1));
        dst[(size_t)r * kBandCols + (col - ox0)] = *e;
        *e = make_float2(0.f, 0.f);
      }
    }
    lo = done;
    wave_sync_lds();
  }
  if (lane == 0) band_zb(partial, pi_n, pj_n)[(size_t)g * nchunk + c] = zb < 0 ? kBandCols : zb;
}

// Output pixels of a tile of kSumTX x kSumTY (block kSumTX x kSumTY): the
// partial sums of the groups and chunks covering each pixel (zeros below an
// item's zero mark, not read), in increasing group then chunk order, over
// the triangular weight sum; 0 where the weight sum is 0.  The factors that depend only on the column (the column
// weight terms and the <= 2 chunks covering it) or only on the row (the row
// weight terms and the <= kSumGroups groups covering it) are computed once per tile
// column / row into LDS.  Contains block barriers (the last one lets the
// caller reuse L).
constexpr int kSumTX = 64, kSumTY = 4;
// A pixel is covered by at most 5 consecutive patch rows / columns (4, plus
// a last one whose origin is clamped): they span at most kSumGroups groups
// and, with kBandCW >= 5, at most 2 chunks.
constexpr int kSumGroups = (4 + kBandRows - 1) / kBandRows + 1;
static_assert(kBandCW >= 5, "a pixel's patch columns must span at most 2 chunks");

struct BandSumLds {
  float cx[kSumTX], sx[kSumTX];
  int nch[kSumTX], ch[kSumTX][2], xo[kSumTX][2];
  int cb[kSumTX][2];  // column part of a partial's offset: chunk * kBandH * kBandCols + xo
  float cy[kSumTY], sy[kSumTY];
  int nb[kSumTY], grp[kSumTY][kSumGroups], ty[kSumTY][kSumGroups];
  int rb[kSumTY][kSumGroups];  // row part: (group * nchunk * kBandH + ty) * kBandCols
  // zero marks of the items the tile meets: groups from the first row's
  // first, chunks from the first column's first (kSumTY rows advance the
  // groups by at most one; kSumTX columns meet at most 3 chunks)
  int zb[kSumGroups + 1][3];
};
static_assert(kSumTY <= kGoldN / 4 && kSumTX <= 8 * kBandCW, "zero-mark table bounds");

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
        L.cb[tx][n] = c * (kBandH * kBandCols) + (x - a);
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
        // the covering patches are consecutive: the groups come in order
        const int gp = p / kBandRows;
        if ((n == 0 || L.grp[tx][n - 1] != gp) && n < kSumGroups) {
          L.grp[tx][n] = gp;
          L.ty[tx][n] = yy - gold_origin(gp * kBandRows, nrow);
          L.rb[tx][n] = (gp * nchunk * kBandH + L.ty[tx][n]) * kBandCols;
          ++n;
        }
      }
    }
    L.cy[tx] = cy;
    L.sy[tx] = sy;
    L.nb[tx] = n;
  }
  __syncthreads();
  // the zero marks of the items the tile meets (the tile's first row and
  // column lie in the image, so both have covering patches)
  const int g0 = L.grp[0][0], c0 = L.ch[0][0];
  {
    const int t = tyi * kSumTX + tx;
    if (t < (kSumGroups + 1) * 3) {
      const int g = g0 + t / 3, c = c0 + t % 3;
      L.zb[t / 3][t % 3] = g < band_ngroup(pi_n) && c < nchunk
                               ? band_zb(partial, pi_n, pj_n)[(size_t)g * nchunk + c]
                               : kBandCols;
    }
  }
  __syncthreads();
  if (x < ncol && y < nrow) {
    const float ws = L.cx[tx] * L.sy[tyi] + L.cy[tyi] * L.sx[tx];
    float2 r = make_float2(0.f, 0.f);
    if (ws > 0.f) {
      // all covering partials loaded first, then added in (group, chunk)
      // order; a wavefront is one tile row, so the row terms are uniform
      const int nb = __builtin_amdgcn_readfirstlane(L.nb[tyi]), nch = L.nch[tx];
      float2 v[kSumGroups][2];
#pragma unroll
      for (int i = 0; i < kSumGroups; ++i) {
        const int rbi = __builtin_amdgcn_readfirstlane(L.rb[tyi][i]);
        const int gi = min(max(__builtin_amdgcn_readfirstlane(L.grp[tyi][i]) - g0, 0), kSumGroups);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          v[i][k] = make_float2(0.f, 0.f);
          if (i < nb && k < nch) {
            // a column below the item's zero mark holds zeros, not written
            const int ci = min(max(L.ch[tx][k] - c0, 0), 2);
            if (L.xo[tx][k] >= L.zb[gi][ci]) v[i][k] = partial[rbi + L.cb[tx][k]];
          }
        }
      }
      float ax = 0.f, ay = 0.f;
#pragma unroll
      for (int i = 0; i < kSumGroups; ++i)
#pragma unroll
        for (int k = 0; k < 2; ++k)
          if (i < nb && k < nch) {
            ax += v[i][k].x;
            ay += v[i][k].y;
          }
      const float inv = 1.0f / ws;
      r = make_float2(ax * inv, ay * inv);
    }
    out[(size_t)y * ncol + x] = r;
  }
  __syncthreads();
}

}  // namespace s1ifg
