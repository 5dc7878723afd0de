// Subswath stitching on the GPU (interfere.py:81-150) and the merge of
// subswath mosaics into the full image (interfere.py:153-180).
//
// The bursts of a subswath are processed in order on the context stream; the
// mosaic is the only state between them.  Per burst i >= 1 whose ids differ
// from burst i-1's:
//   1. k_ov_partial: over the burst rows that earlier bursts' boxes cover
//      (elsewhere the mosaic holds no data), pixels where both the mosaic
//      and the burst hold data (the overlap mask) add their phase difference
//      d = conj(old) * new (complex) or exp(i (new - old)) (phase images) to
//      FP64 block partials, with the overlap count and its row range;
//   2. three passes of k_ov_hist: radix selection (11, 11 and 10 bits) of
//      the ranks (n-1)/2 and n/2 among the float32 keys angle(d exp(-i m)),
//      m the angle of the mean difference, over the overlap rows.  Each
//      block derives the state it filters with itself (pass 0 sums the
//      partials in block order, pass p > 0 selects the digits of pass p-1's
//      histogram), so no one-block kernel sits between the passes.  Pass
//      0 stores the keys (one u32 per cell of the overlap rows) in the
//      workspace; passes 1 and 2 read them instead of recomputing;
//   3. k_ov_median (one block): the last digits; the running offset becomes
//      m + median (float32 arithmetic: np.median of float32 is the middle
//      value or the float32 mean of the two middle values);
// then, for every burst i >= 1, k_apply replaces mosaic pixels holding no
// data by the burst's pixels rotated (or shifted and wrapped) by the running
// offset, which carries over to bursts whose ids match their predecessor's,
// or whose overlap is empty, as in the reference.
//
// When no burst's rows reach past its predecessor's into an earlier burst's
// (the usual layout of a subswath: each burst overlaps the next one only),
// the pixels a burst meets in the mosaic are its predecessor's, rotated by
// the predecessor's offset.  Then every boundary's estimate is taken
// against the predecessor's own pixels, all boundaries in the same launches
// (grid y = boundary): the angle of the mean difference shifts by the
// predecessor's offset, the residual angles do not, so k_ov_chain later
// chains the offsets in one thread (the carried offsets included) and the
// bursts are applied in order with them.  Otherwise the boundaries are
// estimated one after the other against the mosaic.
//
// Everything stays on the device (no host round trip per burst).  Partials
// are summed in a fixed order and histograms are integer counts, so results
// do not depend on timing.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <vector>

#include "s1ifg_internal.h"

using namespace s1ifg;

namespace {

constexpr int kOvBlocks = 1024;    // fixed grid of the partial sums
constexpr int kBins = 2048;        // radix digit of the first two passes (11 bits)
constexpr float kPiF = 3.14159265358979f;

constexpr int kHistBlocks = 256;   // grid of the histogram passes
constexpr int kChunk = 1024;       // columns per work item (4 per thread)

struct OvPartial {
  double re, im;
  long long n;
  int rmin, rmax;
};

// Selection state of one radix pass: the key bits selected so far and the
// remaining rank, for the (n-1)/2 and n/2 order statistics.
struct Sel {
  unsigned prefix[2];
  long long k[2];
};

struct StitchState {
  float off;            // running mean_phase_diff (0 before the first estimate)
  float m;              // angle of the mean phase difference
  float mr, mi;         // the mean phase difference (complex64)
  float med;            // median of the residual angles
  long long n;          // overlap count of the current burst
  int rmin, rmax;       // overlap rows (burst window rows)
  Sel sel[3];           // sel[p]: the state histogram pass p filters with
};

struct Workspace {
  StitchState st;
  unsigned hist[3][2][kBins];  // hist[p]: counts of pass p
  OvPartial part[kOvBlocks];
};

struct Win {
  int h, w;          // burst window rows, columns
  int r0, c0;        // window origin in the mosaic
  long ld;           // mosaic row stride (elements)
  int o0, o1;        // burst rows [o0, o1) inside earlier bursts' boxes
};

// Where a boundary estimate reads the "old" pixels: the mosaic, or the
// predecessor burst (origin r0, c0 in the mosaic, h x w, row stride ld).
template <typename T>
struct Old {
  const T *base;
  long ld;
  int r0, c0, h, w;
};

// One boundary of a launch over several (blockIdx.y).
template <typename T>
struct Est {
  const T *b;        // the new burst
  Win w;
  Old<T> old;
  Workspace *ws;
  unsigned *keys;    // (o1 - o0) x w keys
};
constexpr int kMaxEst = 24;  // boundaries per launch (kernel arguments <= 4 KB)
template <typename T>
struct EstBatch {
  Est<T> e[kMaxEst];
};

// Row r of the new burst's window against the old pixels: pointer to the old
// row at the window's first column, and the window columns [lo, hi) the old
// source covers (everything for the mosaic).
template <typename T>
__device__ __forceinline__ const T *old_row(const Est<T> &e, int r, int &lo, int &hi) {
  const int R = e.w.r0 + r - e.old.r0;
  if (R < 0 || R >= e.old.h) {
    lo = hi = 0;
    return e.old.base;
  }
  lo = max(0, e.old.c0 - e.w.c0);
  hi = min(e.w.w, e.old.c0 + e.old.w - e.w.c0);
  return e.old.base + (long)R * e.old.ld + (e.w.c0 - e.old.c0);
}

__device__ __forceinline__ bool has_data(float2 v) { return v.x != 0.f; }
__device__ __forceinline__ bool has_data(float v) { return v != 0.f; }
// np.ndarray != 0 on complex64 (either part non-zero), on float32
__device__ __forceinline__ bool nonzero(float2 v) { return v.x != 0.f || v.y != 0.f; }
__device__ __forceinline__ bool nonzero(float v) { return v != 0.f; }

// complex64 product without contraction (numpy rounds each product)
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(__fsub_rn(__fmul_rn(a.x, b.x), __fmul_rn(a.y, b.y)),
                     __fadd_rn(__fmul_rn(a.x, b.y), __fmul_rn(a.y, b.x)));
}

// phase difference of an overlap pixel (interfere.py:121-126)
__device__ __forceinline__ float2 phase_diff(float2 o, float2 n) {
  return cmul(make_float2(o.x, -o.y), n);
}
__device__ __forceinline__ float2 phase_diff(float o, float n) {
  float s, c;
  sincosf(__fsub_rn(n, o), &s, &c);
  return make_float2(c, s);
}

// float32 bits -> unsigned key of the same order (finite values)
__device__ __forceinline__ unsigned order_key(float f) {
  const unsigned b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float key_value(unsigned u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// rotation exp(-i m) in complex64 (cos m, -sin m)
__device__ __forceinline__ float2 unrotate(float m) {
  float s, c;
  sincosf(m, &s, &c);
  return make_float2(c, -s);
}

template <typename T>
__device__ __forceinline__ float median_key(T o, T n, float2 rot) {
  const float2 d = cmul(phase_diff(o, n), rot);
  return atan2f(d.y, d.x);
}

// Burst 0 resampled into the subswath bounds (CroppedImage.resample: zeros
// outside the burst).  One row of the mosaic per block.
template <typename T>
__global__ __launch_bounds__(256) void k_place_first(const T *__restrict__ b, int bh, int bw,
                                                     int br0, int bc0, T *__restrict__ mosaic,
                                                     int mh, int mw) {
  const int r = blockIdx.x;
  const int rb = r - br0;
  for (int c = threadIdx.x; c < mw; c += 256) {
    const int cb = c - bc0;
    T v{};
    if (rb >= 0 && rb < bh && cb >= 0 && cb < bw) v = b[(long)rb * bw + cb];
    mosaic[(long)r * mw + c] = v;
  }
}

// k_place_first for complex bursts with even widths and column offset
// (16-byte aligned rows): two cells per 16-byte access.
__global__ __launch_bounds__(256) void k_place_first_c2(const float4 *__restrict__ b, int bh,
                                                        int bw2, int br0, int bc02,
                                                        float4 *__restrict__ mosaic, int mh,
                                                        int mw2) {
  const int r = blockIdx.x;
  const int rb = r - br0;
  const bool in_row = rb >= 0 && rb < bh;
  for (int c = threadIdx.x; c < mw2; c += 256) {
    const int cb = c - bc02;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (in_row && cb >= 0 && cb < bw2) v = b[(long)rb * bw2 + cb];
    mosaic[(long)r * mw2 + c] = v;
  }
}

// Fixed-order block sums: s_re / s_im / s_n added, s_lo / s_hi min / max.
struct BlockSums {
  double re[256], im[256];
  long long n[256];
  int lo[256], hi[256];
};

__device__ __forceinline__ void block_sums(BlockSums &S) {
  const int t = threadIdx.x;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      S.re[t] += S.re[t + h];
      S.im[t] += S.im[t + h];
      S.n[t] += S.n[t + h];
      S.lo[t] = min(S.lo[t], S.lo[t + h]);
      S.hi[t] = max(S.hi[t], S.hi[t + h]);
    }
    __syncthreads();
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_ov_partial(const EstBatch<T> B) {
  __shared__ BlockSums S;
  const Est<T> &e = B.e[blockIdx.y];
  const T *__restrict__ b = e.b;
  const Win w = e.w;
  Workspace *__restrict__ ws = e.ws;
  const int t = threadIdx.x;
  double sr = 0.0, si = 0.0;
  long long n = 0;
  int rmin = INT_MAX, rmax = -1;
  const int nchunk = (w.w + kChunk - 1) / kChunk;
  const long items = (long)(w.o1 - w.o0) * nchunk;
  for (long it = blockIdx.x; it < items; it += kOvBlocks) {
    const int rr = (int)(it / nchunk);
    const int r = w.o0 + rr;
    const int cb = (int)(it - (long)rr * nchunk) * kChunk, ce = min(cb + kChunk, w.w);
    const T *brow = b + (long)r * w.w;
    int lo, hi;
    const T *mrow = old_row(e, r, lo, hi);
    bool any = false;
    T o[4], v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int c = cb + t + 256 * k;
      o[k] = c < ce && c >= lo && c < hi ? mrow[c] : T{};
      v[k] = c < ce ? brow[c] : T{};
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (has_data(o[k]) && nonzero(v[k])) {
        const float2 d = phase_diff(o[k], v[k]);
        sr += (double)d.x;
        si += (double)d.y;
        ++n;
        any = true;
      }
    }
    if (any) {
      rmin = min(rmin, r);
      rmax = max(rmax, r);
    }
  }
  S.re[t] = sr;
  S.im[t] = si;
  S.n[t] = n;
  S.lo[t] = rmin;
  S.hi[t] = rmax;
  block_sums(S);
  if (t == 0) ws->part[blockIdx.x] = OvPartial{S.re[0], S.im[0], S.n[0], S.lo[0], S.hi[0]};
}

// The partials in block order (4 per thread, then the tree), by every
// thread of a block: overlap count, angle m of the mean (float32, as
// np.angle of the complex64 mean) and the overlap rows.
struct Overlap {
  long long n;
  float m, mr, mi;
  int rmin, rmax;
};

__device__ __forceinline__ Overlap sum_partials(const Workspace *ws, BlockSums &S) {
  const int t = threadIdx.x;
  constexpr int kPer = kOvBlocks / 256;
  double sr = 0.0, si = 0.0;
  long long n = 0;
  int rmin = INT_MAX, rmax = -1;
  for (int k = 0; k < kPer; ++k) {
    const OvPartial p = ws->part[t * kPer + k];
    sr += p.re;
    si += p.im;
    n += p.n;
    rmin = min(rmin, p.rmin);
    rmax = max(rmax, p.rmax);
  }
  S.re[t] = sr;
  S.im[t] = si;
  S.n[t] = n;
  S.lo[t] = rmin;
  S.hi[t] = rmax;
  block_sums(S);
  Overlap o;
  o.n = S.n[0];
  o.m = o.mr = o.mi = 0.f;
  if (o.n > 0) {
    o.mr = (float)(S.re[0] / (double)o.n);
    o.mi = (float)(S.im[0] / (double)o.n);
    o.m = atan2f(o.mi, o.mr);
  }
  o.rmin = S.lo[0];
  o.rmax = S.hi[0];
  __syncthreads();  // S reusable
  return o;
}

// One radix digit selection by every thread of a block: for each target, the
// digit of histogram H (nb bins) whose cumulative count passes the remaining
// rank.  Returns the next selection state (same value in every thread).
struct SelScratch {
  unsigned long long c[256];
  unsigned pick[2];
  long long rest[2];
};

__device__ __forceinline__ Sel select_digit(const unsigned (*H)[kBins], int nb, int shift,
                                            const Sel &in, SelScratch &X) {
  const int t = threadIdx.x;
  const int per = nb / 256;
  const bool same = in.prefix[0] == in.prefix[1];
  for (int tg = 0; tg < 2; ++tg) {
    const unsigned *hs = H[same ? 0 : tg];
    unsigned x[8];
    unsigned long long c = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      x[k] = k < per ? hs[t * per + k] : 0u;
      c += x[k];
    }
    // inclusive scan of the per-thread counts (Hillis-Steele)
    X.c[t] = c;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      const unsigned long long y = t >= o ? X.c[t - o] : 0ull;
      __syncthreads();
      X.c[t] += y;
      __syncthreads();
    }
    const long long want = in.k[tg];
    const unsigned long long hi = X.c[t], lo = hi - c;
    if ((unsigned long long)want >= lo && (unsigned long long)want < hi) {
      unsigned long long run = lo;
      for (int k = 0; k < per; ++k) {
        if ((unsigned long long)want < run + x[k]) {
          X.pick[tg] = (unsigned)(t * per + k);
          X.rest[tg] = want - (long long)run;
          break;
        }
        run += x[k];
      }
    }
    __syncthreads();
  }
  Sel out;
  for (int tg = 0; tg < 2; ++tg) {
    out.prefix[tg] = in.prefix[tg] | (X.pick[tg] << shift);
    out.k[tg] = X.rest[tg];
  }
  __syncthreads();  // X reusable
  return out;
}

constexpr int kShift[3] = {21, 10, 0}, kWidth[3] = {11, 11, 10};

// Radix pass p over the overlap rows: histogram of bits [shift, shift +
// width) of the keys angle(d exp(-i m)) whose higher bits equal each
// target's selected prefix.  Every block first derives the selection state
// it filters with, identically: pass 0 from the overlap partials, pass p > 0
// by the digit selection over pass p-1's histogram; block 0 stores it (and
// pass 0 the overlap figures) for the following kernels.
template <typename T>
__global__ __launch_bounds__(256) void k_ov_hist(const EstBatch<T> B, int pass) {
  __shared__ unsigned h[2][kBins];
  const Est<T> &e = B.e[blockIdx.y];
  const T *__restrict__ b = e.b;
  const Win w = e.w;
  Workspace *__restrict__ ws = e.ws;
  unsigned *__restrict__ keys = e.keys;
  __shared__ union {
    BlockSums S;
    SelScratch X;
  } U;
  StitchState &gst = ws->st;
  const int t = threadIdx.x;
  long long cnt;
  float m;
  int rmin, rmax;
  Sel sel;
  if (pass == 0) {
    const Overlap o = sum_partials(ws, U.S);
    cnt = o.n;
    m = o.m;
    rmin = o.rmin;
    rmax = o.rmax;
    sel.prefix[0] = sel.prefix[1] = 0u;
    sel.k[0] = (cnt - 1) / 2;
    sel.k[1] = cnt / 2;
    if (blockIdx.x == 0 && t == 0) {
      gst.n = cnt;
      if (cnt > 0) {
        gst.m = m;
        gst.mr = o.mr;
        gst.mi = o.mi;
        gst.rmin = rmin;
        gst.rmax = rmax;
        gst.sel[0] = sel;
      }
    }
    if (cnt == 0) return;
  } else {
    cnt = gst.n;
    if (cnt == 0) return;
    m = gst.m;
    rmin = gst.rmin;
    rmax = gst.rmax;
    sel = select_digit(ws->hist[pass - 1], 1 << kWidth[pass - 1], kShift[pass - 1],
                       gst.sel[pass - 1], U.X);
    if (blockIdx.x == 0 && t == 0) gst.sel[pass] = sel;
  }
  const int shift = kShift[pass], width = kWidth[pass];
  const int nb = 1 << width;
  for (int k = t; k < 2 * kBins; k += 256) (&h[0][0])[k] = 0u;
  __syncthreads();
  const unsigned hmask = (shift + width >= 32) ? 0u : (0xffffffffu << (shift + width));
  const unsigned dmask = (unsigned)nb - 1u;
  const float2 rot = unrotate(m);
  const unsigned p0 = sel.prefix[0], p1 = sel.prefix[1];
  const bool same = p0 == p1;
  const int nchunk = (w.w + kChunk - 1) / kChunk;
  const long items = (long)(rmax - rmin + 1) * nchunk;
  for (long it = blockIdx.x; it < items; it += kHistBlocks) {
    const int rr = (int)(it / nchunk);
    const int r = rmin + rr;
    const int cb = (int)(it - (long)rr * nchunk) * kChunk, ce = min(cb + kChunk, w.w);
    unsigned *krow = keys + (long)(r - w.o0) * w.w;
    unsigned u4[4];
    if (pass == 0) {
      // keys of the overlap cells (0 elsewhere: no finite angle maps to 0),
      // kept for the later passes
      const T *brow = b + (long)r * w.w;
      int lo, hi;
      const T *mrow = old_row(e, r, lo, hi);
      T o[4], v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int c = cb + t + 256 * k;
        o[k] = c < ce && c >= lo && c < hi ? mrow[c] : T{};
        v[k] = c < ce ? brow[c] : T{};
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int c = cb + t + 256 * k;
        u4[k] = has_data(o[k]) && nonzero(v[k]) ? order_key(median_key(o[k], v[k], rot)) : 0u;
        if (c < ce) krow[c] = u4[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int c = cb + t + 256 * k;
        u4[k] = c < ce ? krow[c] : 0u;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const unsigned u = u4[k];
      if (u == 0u) continue;
      const unsigned d = (u >> shift) & dmask;
      if ((u & hmask) == p0) atomicAdd(&h[0][d], 1u);
      if (!same && (u & hmask) == p1) atomicAdd(&h[1][d], 1u);
    }
  }
  __syncthreads();
  unsigned(*gh)[kBins] = ws->hist[pass];
  for (int k = t; k < nb; k += 256) {
    if (h[0][k]) atomicAdd(&gh[0][k], h[0][k]);
    if (!same && h[1][k]) atomicAdd(&gh[1][k], h[1][k]);
  }
}

// One block: the digit selection over the last pass's histogram gives both
// order statistics; the running offset becomes m + median.  All histograms
// are cleared for the next boundary.
template <typename T>
__global__ __launch_bounds__(256) void k_ov_median(const EstBatch<T> B, int chained) {
  __shared__ SelScratch X;
  Workspace *__restrict__ ws = B.e[blockIdx.x].ws;
  StitchState &gst = ws->st;
  const long long cnt = gst.n;
  if (cnt == 0) return;
  const Sel sel = select_digit(ws->hist[2], 1 << kWidth[2], kShift[2], gst.sel[2], X);
  for (int k = threadIdx.x; k < 3 * 2 * kBins; k += 256) (&ws->hist[0][0][0])[k] = 0u;
  if (threadIdx.x == 0) {
    const float a = key_value(sel.prefix[0]), bb = key_value(sel.prefix[1]);
    const float med = (cnt & 1) ? a : __fmul_rn(__fadd_rn(a, bb), 0.5f);
    gst.med = med;
    if (!chained) gst.off = __fadd_rn(gst.m, med);
  }
}

// Replace the mosaic pixels without data by the burst's (interfere.py:
// 135-143): rotated by exp(-i off) (complex64), or shifted by -off and
// wrapped once into [-pi, pi] (float32) when off != 0.  One burst row per
// block; rows outside earlier bursts' boxes hold no data in the mosaic, so
// there the mosaic is not read.  Block 0 records the offset used.
template <typename T>
__device__ __forceinline__ T corrected(T v, float off);
template <>
__device__ __forceinline__ float2 corrected(float2 v, float off) {
  return cmul(v, unrotate(off));
}
template <>
__device__ __forceinline__ float corrected(float v, float off) {
  float x = __fsub_rn(v, off);
  if (x > kPiF) x = __fsub_rn(x, 2.f * kPiF);
  if (x < -kPiF) x = __fadd_rn(x, 2.f * kPiF);
  return x;
}

template <typename T>
__global__ __launch_bounds__(256) void k_apply(const T *__restrict__ b, Win w,
                                               T *__restrict__ mosaic, const float *offp,
                                               double *__restrict__ offsets, int burst) {
  const float off = *offp;
  const int r = blockIdx.x;
  const T *brow = b + (long)r * w.w;
  T *mrow = mosaic + (long)(w.r0 + r) * w.ld + w.c0;
  const bool covered = r >= w.o0 && r < w.o1;
  constexpr int kU = 8;  // loads in flight per thread
  for (int c0 = threadIdx.x; c0 < w.w; c0 += 256 * kU) {
    T v[kU];
    bool put[kU];
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int c = c0 + 256 * k;
      v[k] = c < w.w ? brow[c] : T{};
    }
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int c = c0 + 256 * k;
      put[k] = has_data(v[k]) && !(covered && has_data(mrow[c]));
    }
#pragma unroll
    for (int k = 0; k < kU; ++k)
      if (put[k]) mrow[c0 + 256 * k] = off != 0.f ? corrected(v[k], off) : v[k];
  }
  if (offsets && r == 0 && threadIdx.x == 0) offsets[burst] = (double)off;
}

// k_apply for complex bursts with 16-byte alignment (even window width,
// column offset and row stride): two cells per 16-byte access.
__global__ __launch_bounds__(256) void k_apply_c2(const float4 *__restrict__ b, Win w,
                                                  float2 *__restrict__ mosaic, const float *offp,
                                                  double *__restrict__ offsets, int burst) {
  const float off = *offp;
  const int r = blockIdx.x;
  const int w2 = w.w >> 1;
  const float4 *brow = b + (long)r * w2;
  float2 *mrow = mosaic + (long)(w.r0 + r) * w.ld + w.c0;
  float4 *mrow4 = reinterpret_cast<float4 *>(mrow);
  const bool covered = r >= w.o0 && r < w.o1;
  const float2 rot = unrotate(off);
  constexpr int kU = 4;
  for (int c0 = threadIdx.x; c0 < w2; c0 += 256 * kU) {
    float4 v[kU];
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int c = c0 + 256 * k;
      v[k] = c < w2 ? brow[c] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    float4 o[kU];
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int c = c0 + 256 * k;
      o[k] = covered && c < w2 && (v[k].x != 0.f || v[k].z != 0.f) ? mrow4[c]
                                                                  : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int c = c0 + 256 * k;
      const bool p0 = v[k].x != 0.f && o[k].x == 0.f;
      const bool p1 = v[k].z != 0.f && o[k].z == 0.f;
      float2 a = make_float2(v[k].x, v[k].y), bb = make_float2(v[k].z, v[k].w);
      if (off != 0.f) {
        a = cmul(a, rot);
        bb = cmul(bb, rot);
      }
      if (p0 && p1) {
        mrow4[c] = make_float4(a.x, a.y, bb.x, bb.y);
      } else {
        if (p0) mrow[2 * c] = a;
        if (p1) mrow[2 * c + 1] = bb;
      }
    }
  }
  if (offsets && r == 0 && threadIdx.x == 0) offsets[burst] = (double)off;
}

// The whole mosaic in one pass once the offsets are known (the chained
// form), with the result of placing burst 0 and applying the bursts in
// order (interfere.py:97-143): a pixel starts as burst 0's value where its
// box covers it, else 0; each later burst, in order, replaces it by its own
// value rotated / shifted by its offset where that value has data and the
// pixel has none.  One mosaic row per block.  In the chained form at most
// two bursts cover a row (bursts that share rows are consecutive); a row
// covered by more takes the general loop.
constexpr int kMaxCompose = 96;
template <typename T>
struct ComposeArgs {
  struct B {
    const T *d;
    int top, bottom, left, right;
  } b[kMaxCompose];
};

template <typename T>
__device__ __forceinline__ void compose_step(T &cur, T x, bool first, float off) {
  if (first)
    cur = x;
  else if (!has_data(cur) && has_data(x))
    cur = off != 0.f ? corrected(x, off) : x;
}

template <typename T>
__global__ __launch_bounds__(256) void k_compose(const ComposeArgs<T> A, int n, int top0, int left0,
                                                 int mw, const float *__restrict__ offs,
                                                 T *__restrict__ mosaic,
                                                 double *__restrict__ offsets) {
  const int R = blockIdx.x, gr = top0 + R;
  int cv[2] = {-1, -1}, nc = 0;  // the bursts covering this row, in order
  for (int j = 0; j < n && nc <= 2; ++j)
    if (A.b[j].top <= gr && gr < A.b[j].bottom) {
      if (nc < 2) cv[nc] = j;
      ++nc;
    }
  T *mrow = mosaic + (long)R * mw;
  if (nc > 2) {
    for (int c = threadIdx.x; c < mw; c += 256) {
      const int gc = left0 + c;
      T cur{};
      for (int j = 0; j < n; ++j) {
        const auto &b = A.b[j];
        if (b.top <= gr && gr < b.bottom && b.left <= gc && gc < b.right)
          compose_step(cur, b.d[(long)(gr - b.top) * (b.right - b.left) + (gc - b.left)], j == 0,
                       j ? offs[j] : 0.f);
      }
      mrow[c] = cur;
    }
  } else {
    const T *row[2] = {nullptr, nullptr};
    int l[2] = {0, 0}, r[2] = {0, 0};
    float off[2] = {0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 2; ++k)
      if (cv[k] >= 0) {
        const auto &b = A.b[cv[k]];
        row[k] = b.d + (long)(gr - b.top) * (b.right - b.left) - b.left;
        l[k] = b.left - left0;
        r[k] = b.right - left0;
        off[k] = cv[k] ? offs[cv[k]] : 0.f;
      }
    constexpr int kU = 4;  // cells in flight per thread and source
    for (int c0 = threadIdx.x; c0 < mw; c0 += 256 * kU) {
      T x[2][kU];
#pragma unroll
      for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int c = c0 + 256 * u;
          x[k][u] = c >= l[k] && c < r[k] && c < mw ? row[k][left0 + c] : T{};
        }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int c = c0 + 256 * u;
        T cur{};
#pragma unroll
        for (int k = 0; k < 2; ++k)
          if (c >= l[k] && c < r[k]) compose_step(cur, x[k][u], cv[k] == 0, off[k]);
        if (c < mw) mrow[c] = cur;
      }
    }
  }
  if (offsets && R == 0)
    for (int j = 1 + threadIdx.x; j < n; j += 256) offsets[j] = (double)offs[j];
}

template <typename T>
__global__ __launch_bounds__(256) void k_merge(const T *__restrict__ src, int h, int w,
                                               T *__restrict__ dst, long ld) {
  const int r = blockIdx.x;
  for (int c = threadIdx.x; c < w; c += 256) {
    const T v = src[(long)r * w + c];
    if (has_data(v)) dst[(long)r * ld + c] = v;
  }
}

// Clears the state and histograms of workspaces ws[0 .. gridDim.x).
__global__ void k_init_state(Workspace *__restrict__ ws, double *__restrict__ offsets) {
  Workspace *w = ws + blockIdx.x;
  if (threadIdx.x == 0) {
    w->st.off = 0.f;
    w->st.n = 0;
    if (offsets && blockIdx.x == 0) offsets[0] = 0.0;
  }
  for (int k = threadIdx.x; k < 3 * 2 * kBins; k += blockDim.x) (&w->hist[0][0][0])[k] = 0u;
}

// The running offsets of the bursts from the boundary estimates taken
// against each predecessor's own pixels (one thread): at an estimated
// boundary with overlap, the mean difference is rotated by the
// predecessor's offset (the mosaic holds the predecessor rotated by it)
// and the residual median added (interfere.py:121-133); elsewhere the
// offset carries over.  est[i]: the workspace of burst i's boundary, or -1.
__device__ void ov_chain(const Workspace *__restrict__ ws, const int *est, int n,
                         float *__restrict__ offs) {
  float off = 0.f;
  offs[0] = 0.f;
  for (int i = 1; i < n; ++i) {
    const int e = est[i];
    if (e >= 0 && ws[e].st.n > 0) {
      const StitchState &st = ws[e].st;
      float sn, cs;
      sincosf(off, &sn, &cs);
      const float2 mean = cmul(make_float2(st.mr, st.mi), make_float2(cs, sn));
      off = __fadd_rn(atan2f(mean.y, mean.x), st.med);
    }
    offs[i] = off;
  }
}

constexpr int kMaxChain = 256;  // bursts of the chained form (k_ov_chain's argument)
struct ChainArgs {
  int est[kMaxChain];
};

__global__ void k_ov_chain(const Workspace *__restrict__ ws, const ChainArgs A, int n,
                           float *__restrict__ offs) {
  if (threadIdx.x == 0) ov_chain(ws, A.est, n, offs);
}

template <typename T>
int launch_estimates(hipStream_t s, const std::vector<Est<T>> &es, int chained) {
  for (size_t i0 = 0; i0 < es.size(); i0 += kMaxEst) {
    const int nb = (int)std::min(es.size() - i0, (size_t)kMaxEst);
    EstBatch<T> B{};
    for (int k = 0; k < nb; ++k) B.e[k] = es[i0 + k];
    hipLaunchKernelGGL(k_ov_partial<T>, dim3(kOvBlocks, nb), dim3(256), 0, s, B);
    S1_LAUNCH_CHECK("k_ov_partial");
    for (int pss = 0; pss < 3; ++pss) {
      hipLaunchKernelGGL(k_ov_hist<T>, dim3(kHistBlocks, nb), dim3(256), 0, s, B, pss);
      S1_LAUNCH_CHECK("k_ov_hist");
    }
    hipLaunchKernelGGL(k_ov_median<T>, dim3(nb), dim3(256), 0, s, B, chained);
    S1_LAUNCH_CHECK("k_ov_median");
  }
  return S1IFG_OK;
}

template <typename T>
int stitch_impl(s1ifg_ctx *ctx, const s1ifg_stitch_burst *bursts, int n, T *mosaic, int left,
                int top, int right, int bottom, double *offsets) {
  const int mh = bottom - top, mw = right - left;
  const s1ifg_stitch_burst &b0 = bursts[0];
  // rows [o0, o1) of each burst inside earlier bursts' boxes (burst 0's box
  // clipped to the bounds), and whether every burst meets only its
  // predecessor among them (the chained form)
  std::vector<int> o0(n, 0), o1(n, 0), est(n, -1);
  bool chained = n <= kMaxChain;
  int nest = 0;
  {
    int c0 = max(b0.top, top), c1 = min(b0.bottom, bottom);
    for (int i = 1; i < n; ++i) {
      const s1ifg_stitch_burst &bi = bursts[i];
      const int h = bi.bottom - bi.top;
      o0[i] = min(max(c0 - bi.top, 0), h);
      o1[i] = max(min(c1 - bi.top, h), o0[i]);
      for (int j = 0; j + 1 < i && chained; ++j) {
        const int tj = j == 0 ? max(b0.top, top) : bursts[j].top;
        const int bj = j == 0 ? min(b0.bottom, bottom) : bursts[j].bottom;
        if (tj < bi.bottom && bi.top < bj) chained = false;
      }
      if (bi.estimate) est[i] = nest++;
      c0 = min(c0, bi.top);
      c1 = max(c1, bi.bottom);
    }
  }
  // workspace: states and histograms (one per estimated boundary when
  // chained, else one), the offsets, the keys (per boundary when chained,
  // else the largest band)
  const int nws = chained ? std::max(nest, 1) : 1;
  std::vector<size_t> koff(n, 0);
  size_t nkeys = 0;
  for (int i = 1; i < n; ++i) {
    if (est[i] < 0) continue;
    const size_t band = (size_t)(o1[i] - o0[i]) * (size_t)(bursts[i].right - bursts[i].left);
    if (chained) {
      koff[i] = nkeys;
      nkeys += band;
    } else {
      nkeys = std::max(nkeys, band);
    }
  }
  const size_t ws_bytes = ((size_t)nws * sizeof(Workspace) + 255) & ~(size_t)255;
  const size_t off_bytes = ((size_t)n * sizeof(float) + 255) & ~(size_t)255;
  void *p = nullptr;
  int rc = ctx_workspace(ctx, ws_bytes + off_bytes + nkeys * sizeof(unsigned), &p);
  if (rc) return rc;
  Workspace *ws = (Workspace *)p;
  float *offs = (float *)((char *)p + ws_bytes);
  unsigned *keys = (unsigned *)((char *)p + ws_bytes + off_bytes);
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_init_state, dim3(nws), dim3(256), 0, s, ws, offsets);
  S1_LAUNCH_CHECK("k_init_state");
  auto window = [&](int i) {
    const s1ifg_stitch_burst &bi = bursts[i];
    return Win{bi.bottom - bi.top, bi.right - bi.left, bi.top - top, bi.left - left, (long)mw,
               o0[i], o1[i]};
  };
  if (chained && n <= kMaxCompose) {
    // every boundary against its predecessor's pixels, in one set of
    // launches; the offsets in one thread; the mosaic in one pass
    if (nest > 0) {
      std::vector<Est<T>> es;
      for (int i = 1; i < n; ++i) {
        if (est[i] < 0) continue;
        const s1ifg_stitch_burst &bp = bursts[i - 1];
        Old<T> old{(const T *)bp.data, (long)(bp.right - bp.left), bp.top - top, bp.left - left,
                   bp.bottom - bp.top, bp.right - bp.left};
        es.push_back(
            Est<T>{(const T *)bursts[i].data, window(i), old, ws + est[i], keys + koff[i]});
      }
      if ((rc = launch_estimates<T>(s, es, 1))) return rc;
    }
    ChainArgs A{};
    for (int i = 0; i < n; ++i) A.est[i] = est[i];
    hipLaunchKernelGGL(k_ov_chain, dim3(1), dim3(64), 0, s, (const Workspace *)ws, A, n, offs);
    S1_LAUNCH_CHECK("k_ov_chain");
    ComposeArgs<T> C{};
    for (int i = 0; i < n; ++i)
      C.b[i] = {(const T *)bursts[i].data, bursts[i].top, bursts[i].bottom, bursts[i].left,
                bursts[i].right};
    hipLaunchKernelGGL(k_compose<T>, dim3(mh), dim3(256), 0, s, C, n, top, left, mw,
                       (const float *)offs, mosaic, offsets);
    S1_LAUNCH_CHECK("k_compose");
    return S1IFG_OK;
  }
  const int bw0 = b0.right - b0.left, bc0 = b0.left - left;
  if (sizeof(T) == 8 && mw % 2 == 0 && bw0 % 2 == 0 && bc0 % 2 == 0 &&
      (uintptr_t)b0.data % 16 == 0 && (uintptr_t)mosaic % 16 == 0) {
    hipLaunchKernelGGL(k_place_first_c2, dim3(mh), dim3(256), 0, s, (const float4 *)b0.data,
                       b0.bottom - b0.top, bw0 / 2, b0.top - top, bc0 / 2, (float4 *)mosaic, mh,
                       mw / 2);
    S1_LAUNCH_CHECK("k_place_first_c2");
  } else {
    hipLaunchKernelGGL(k_place_first<T>, dim3(mh), dim3(256), 0, s, (const T *)b0.data,
                       b0.bottom - b0.top, bw0, b0.top - top, bc0, mosaic, mh, mw);
    S1_LAUNCH_CHECK("k_place_first");
  }
  for (int i = 1; i < n; ++i) {
    const s1ifg_stitch_burst &bi = bursts[i];
    const T *bd = (const T *)bi.data;
    const Win w = window(i);
    // the boundaries one after the other against the mosaic
    if (bi.estimate) {
      Old<T> old{(const T *)mosaic, (long)mw, 0, 0, mh, mw};
      std::vector<Est<T>> es{Est<T>{bd, w, old, ws, keys}};
      if ((rc = launch_estimates<T>(s, es, 0))) return rc;
    }
    const float *offp = &ws->st.off;
    const bool pairs = sizeof(T) == 8 && w.w % 2 == 0 && w.c0 % 2 == 0 && w.ld % 2 == 0 &&
                       (uintptr_t)bd % 16 == 0 && (uintptr_t)mosaic % 16 == 0;
    if (pairs) {
      hipLaunchKernelGGL(k_apply_c2, dim3(w.h), dim3(256), 0, s, (const float4 *)bd, w,
                         (float2 *)mosaic, offp, offsets, i);
      S1_LAUNCH_CHECK("k_apply_c2");
    } else {
      hipLaunchKernelGGL(k_apply<T>, dim3(w.h), dim3(256), 0, s, bd, w, mosaic, offp, offsets, i);
      S1_LAUNCH_CHECK("k_apply");
    }
  }
  return S1IFG_OK;
}

}  // namespace

extern "C" {

int s1ifg_stitch_subswath(s1ifg_ctx *ctx, const s1ifg_stitch_burst *bursts, int n, int out_float,
                          void *mosaic, int left, int top, int right, int bottom,
                          double *offsets) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!bursts || n < 1) return set_error(S1IFG_EINVAL, "no bursts");
  if (!mosaic) return set_error(S1IFG_EINVAL, "null mosaic");
  if (right <= left || bottom <= top)
    return set_error(S1IFG_EINVAL, "empty subswath bounds [%d, %d) x [%d, %d)", top, bottom,
                     left, right);
  for (int i = 0; i < n; ++i) {
    const s1ifg_stitch_burst &b = bursts[i];
    if (!b.data) return set_error(S1IFG_EINVAL, "burst %d: null data", i);
    if (b.right <= b.left || b.bottom <= b.top)
      return set_error(S1IFG_EINVAL, "burst %d: empty bounds", i);
    // burst 0 is resampled into the bounds; the others must lie inside
    if (i > 0 && (b.left < left || b.top < top || b.right > right || b.bottom > bottom))
      return set_error(S1IFG_EINVAL,
                       "burst %d: bounds [%d, %d) x [%d, %d) outside the subswath's", i, b.top,
                       b.bottom, b.left, b.right);
  }
  if (out_float)
    return stitch_impl<float>(ctx, bursts, n, (float *)mosaic, left, top, right, bottom,
                              offsets);
  return stitch_impl<float2>(ctx, bursts, n, (float2 *)mosaic, left, top, right, bottom,
                             offsets);
}

int s1ifg_merge_subswath(s1ifg_ctx *ctx, const void *src, int nrow, int ncol, int out_float,
                         void *dst, int64_t ld_dst) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!src || !dst) return set_error(S1IFG_EINVAL, "null data pointer");
  if (nrow < 1 || ncol < 1 || ld_dst < ncol)
    return set_error(S1IFG_EINVAL, "image %d x %d, ld %lld", nrow, ncol, (long long)ld_dst);
  if (out_float)
    hipLaunchKernelGGL(k_merge<float>, dim3(nrow), dim3(256), 0, ctx->stream, (const float *)src,
                       nrow, ncol, (float *)dst, (long)ld_dst);
  else
    hipLaunchKernelGGL(k_merge<float2>, dim3(nrow), dim3(256), 0, ctx->stream,
                       (const float2 *)src, nrow, ncol, (float2 *)dst, (long)ld_dst);
  S1_LAUNCH_CHECK("k_merge");
  return S1IFG_OK;
}

}  // extern "C"
