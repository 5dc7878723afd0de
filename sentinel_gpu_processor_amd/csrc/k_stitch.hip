// Subswath stitching on the GPU (interfere.py:81-150) and the merge of
// subswath mosaics into the full image (interfere.py:153-180).
//
// The bursts of a subswath are processed in order on the context stream; the
// mosaic is the only state between them.  Per burst i >= 1 whose ids differ
// from burst i-1's:
//   1. k_ov_partial: over the burst's window, pixels where both the mosaic
//      and the burst hold data (the overlap mask) add their phase difference
//      d = conj(old) * new (complex) or exp(i (new - old)) (phase images) to
//      FP64 block partials, with the overlap count and its row range;
//   2. k_ov_finalize (one block): partials summed in block order, m = angle
//      of the mean (float32, as np.angle of the complex64 mean), the ranks
//      (n-1)/2 and n/2 of the median, histograms cleared;
//   3. three passes of k_ov_hist + k_ov_select: radix selection (11, 11 and
//      10 bits) of both ranks among the float32 keys angle(d * exp(-i m)),
//      over the overlap rows only; the last pass sets the running offset
//      m + median (float32 arithmetic: np.median of float32 is the middle
//      value or the float32 mean of the two middle values);
// then, for every burst i >= 1, k_apply replaces mosaic pixels holding no
// data by the burst's pixels rotated (or shifted and wrapped) by the running
// offset, which carries over to bursts whose ids match their predecessor's,
// or whose overlap is empty, as in the reference.
//
// Everything stays on the device (no host round trip per burst); the
// partials are summed in a fixed order, so results do not depend on timing.
#include <climits>

#include "s1ifg_internal.h"

using namespace s1ifg;

namespace {

constexpr int kOvBlocks = 1024;    // fixed grid of the partial sums
constexpr int kBins = 2048;        // radix digit of the first two passes (11 bits)
constexpr float kPiF = 3.14159265358979f;

struct OvPartial {
  double re, im;
  long long n;
  int rmin, rmax;
};

struct StitchState {
  float off;            // running mean_phase_diff (0 before the first estimate)
  float m;              // angle of the mean phase difference
  long long n;          // overlap count of the current burst
  long long k[2];       // remaining rank of the (n-1)/2 and n/2 order statistics
  unsigned prefix[2];   // key bits selected so far
  int rmin, rmax;       // overlap rows (burst window rows)
};

struct Workspace {
  StitchState st;
  unsigned hist[2][kBins];
  OvPartial part[kOvBlocks];
};

struct Win {
  int h, w;          // burst window rows, columns
  int r0, c0;        // window origin in the mosaic
  long ld;           // mosaic row stride (elements)
};

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

template <typename T>
__global__ __launch_bounds__(256) void k_ov_partial(const T *__restrict__ b, Win w,
                                                    const T *__restrict__ mosaic,
                                                    Workspace *__restrict__ ws) {
  double sr = 0.0, si = 0.0;
  long long n = 0;
  int rmin = INT_MAX, rmax = -1;
  for (int r = blockIdx.x; r < w.h; r += kOvBlocks) {
    const T *brow = b + (long)r * w.w;
    const T *mrow = mosaic + (long)(w.r0 + r) * w.ld + w.c0;
    bool any = false;
    for (int c = threadIdx.x; c < w.w; c += 256) {
      const T o = mrow[c], v = brow[c];
      if (has_data(o) && nonzero(v)) {
        const float2 d = phase_diff(o, v);
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
  __shared__ double s_re[256], s_im[256];
  __shared__ long long s_n[256];
  __shared__ int s_lo[256], s_hi[256];
  const int t = threadIdx.x;
  s_re[t] = sr;
  s_im[t] = si;
  s_n[t] = n;
  s_lo[t] = rmin;
  s_hi[t] = rmax;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      s_re[t] += s_re[t + h];
      s_im[t] += s_im[t + h];
      s_n[t] += s_n[t + h];
      s_lo[t] = min(s_lo[t], s_lo[t + h]);
      s_hi[t] = max(s_hi[t], s_hi[t + h]);
    }
    __syncthreads();
  }
  if (t == 0) ws->part[blockIdx.x] = OvPartial{s_re[0], s_im[0], s_n[0], s_lo[0], s_hi[0]};
}

// One block of 256 threads: the block partials in a fixed order.
__global__ __launch_bounds__(256) void k_ov_finalize(Workspace *__restrict__ ws) {
  const int t = threadIdx.x;
  constexpr int kPer = kOvBlocks / 256;
  double sr = 0.0, si = 0.0;
  long long n = 0;
  int lo = INT_MAX, hi = -1;
  for (int k = 0; k < kPer; ++k) {
    const OvPartial p = ws->part[t * kPer + k];
    sr += p.re;
    si += p.im;
    n += p.n;
    lo = min(lo, p.rmin);
    hi = max(hi, p.rmax);
  }
  __shared__ double s_re[256], s_im[256];
  __shared__ long long s_n[256];
  __shared__ int s_lo[256], s_hi[256];
  s_re[t] = sr;
  s_im[t] = si;
  s_n[t] = n;
  s_lo[t] = lo;
  s_hi[t] = hi;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      s_re[t] += s_re[t + h];
      s_im[t] += s_im[t + h];
      s_n[t] += s_n[t + h];
      s_lo[t] = min(s_lo[t], s_lo[t + h]);
      s_hi[t] = max(s_hi[t], s_hi[t + h]);
    }
    __syncthreads();
  }
  for (int k = t; k < 2 * kBins; k += 256) (&ws->hist[0][0])[k] = 0u;
  if (t == 0) {
    StitchState &st = ws->st;
    const long long cnt = s_n[0];
    st.n = cnt;
    if (cnt > 0) {
      // np.mean of the complex64 differences, then np.angle in float32
      const float mr = (float)(s_re[0] / (double)cnt), mi = (float)(s_im[0] / (double)cnt);
      st.m = atan2f(mi, mr);
      st.k[0] = (cnt - 1) / 2;
      st.k[1] = cnt / 2;
      st.prefix[0] = st.prefix[1] = 0u;
      st.rmin = s_lo[0];
      st.rmax = s_hi[0];
    }
  }
}

// rotation exp(-i m) in complex64 (cos m, -sin m)
__device__ __forceinline__ float2 unrotate(float m) {
  float s, c;
  sincosf(m, &s, &c);
  return make_float2(c, -s);
}

// Histogram of one radix digit (pass p: bits [shift, shift + width)) of the
// keys whose higher bits equal each target's selected prefix; overlap rows
// only.
template <typename T>
__global__ __launch_bounds__(256) void k_ov_hist(const T *__restrict__ b, Win w,
                                                 const T *__restrict__ mosaic,
                                                 Workspace *__restrict__ ws, int shift, int width) {
  __shared__ unsigned h[2][kBins];
  const StitchState st = ws->st;
  if (st.n == 0) return;
  const int t = threadIdx.x;
  const int nb = 1 << width;
  for (int k = t; k < 2 * kBins; k += 256) (&h[0][0])[k] = 0u;
  __syncthreads();
  const unsigned hmask = (shift + width >= 32) ? 0u : (0xffffffffu << (shift + width));
  const unsigned dmask = (unsigned)nb - 1u;
  const float2 rot = unrotate(st.m);
  const bool same = st.prefix[0] == st.prefix[1];
  for (int r = st.rmin + blockIdx.x; r <= st.rmax; r += gridDim.x) {
    const T *brow = b + (long)r * w.w;
    const T *mrow = mosaic + (long)(w.r0 + r) * w.ld + w.c0;
    for (int c = t; c < w.w; c += 256) {
      const T o = mrow[c], v = brow[c];
      if (!(has_data(o) && nonzero(v))) continue;
      const unsigned u = order_key(median_key(o, v, rot));
      const unsigned d = (u >> shift) & dmask;
      if ((u & hmask) == st.prefix[0]) atomicAdd(&h[0][d], 1u);
      if (!same && (u & hmask) == st.prefix[1]) atomicAdd(&h[1][d], 1u);
    }
  }
  __syncthreads();
  for (int k = t; k < nb; k += 256) {
    if (h[0][k]) atomicAdd(&ws->hist[0][k], h[0][k]);
    if (!same && h[1][k]) atomicAdd(&ws->hist[1][k], h[1][k]);
  }
}

// One block: for each target, the digit whose cumulative count passes the
// remaining rank; the histogram is cleared for the next pass.  After the last
// pass (shift 0) the running offset becomes m + median.
__global__ __launch_bounds__(256) void k_ov_select(Workspace *__restrict__ ws, int shift,
                                                   int width) {
  StitchState &st = ws->st;
  if (st.n == 0) return;
  const int t = threadIdx.x;
  const int nb = 1 << width;
  const int per = nb / 256;
  const bool same = st.prefix[0] == st.prefix[1];
  __shared__ unsigned long long s_c[256];
  __shared__ unsigned s_pick[2];
  __shared__ long long s_rest[2];
  for (int tg = 0; tg < 2; ++tg) {
    const unsigned *hs = ws->hist[same ? 0 : tg];
    unsigned long long c = 0;
    for (int k = 0; k < per; ++k) c += hs[t * per + k];
    s_c[t] = c;
    __syncthreads();
    if (t == 0) {
      unsigned long long run = 0;
      for (int k = 0; k < 256; ++k) {
        const unsigned long long x = s_c[k];
        s_c[k] = run;
        run += x;
      }
    }
    __syncthreads();
    const long long want = st.k[tg];
    const unsigned long long lo = s_c[t], hi = lo + c;
    if ((unsigned long long)want >= lo && (unsigned long long)want < hi) {
      unsigned long long run = lo;
      for (int k = 0; k < per; ++k) {
        const unsigned x = hs[t * per + k];
        if ((unsigned long long)want < run + x) {
          s_pick[tg] = (unsigned)(t * per + k);
          s_rest[tg] = want - (long long)run;
          break;
        }
        run += x;
      }
    }
    __syncthreads();
  }
  for (int k = t; k < 2 * kBins; k += 256) (&ws->hist[0][0])[k] = 0u;
  if (t == 0) {
    for (int tg = 0; tg < 2; ++tg) {
      st.prefix[tg] |= s_pick[tg] << shift;
      st.k[tg] = s_rest[tg];
    }
    if (shift == 0) {
      const float a = key_value(st.prefix[0]), b = key_value(st.prefix[1]);
      const float med = (st.n & 1) ? a : __fmul_rn(__fadd_rn(a, b), 0.5f);
      st.off = __fadd_rn(st.m, med);
    }
  }
}

// Replace the mosaic pixels without data by the burst's (interfere.py:
// 135-143): rotated by exp(-i off) (complex64), or shifted by -off and
// wrapped once into [-pi, pi] (float32) when off != 0.  One burst row per
// block.  Block 0 records the offset used.
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
                                               T *__restrict__ mosaic,
                                               const Workspace *__restrict__ ws,
                                               double *__restrict__ offsets, int burst) {
  const float off = ws->st.off;
  const int r = blockIdx.x;
  const T *brow = b + (long)r * w.w;
  T *mrow = mosaic + (long)(w.r0 + r) * w.ld + w.c0;
  for (int c = threadIdx.x; c < w.w; c += 256) {
    const T o = mrow[c], v = brow[c];
    if (!has_data(o) && has_data(v)) mrow[c] = off != 0.f ? corrected(v, off) : v;
  }
  if (offsets && r == 0 && threadIdx.x == 0) offsets[burst] = (double)off;
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

__global__ void k_init_state(Workspace *__restrict__ ws, double *__restrict__ offsets) {
  if (threadIdx.x == 0) {
    ws->st.off = 0.f;
    ws->st.n = 0;
    if (offsets) offsets[0] = 0.0;
  }
}

template <typename T>
int stitch_impl(s1ifg_ctx *ctx, const s1ifg_stitch_burst *bursts, int n, T *mosaic, int left,
                int top, int right, int bottom, double *offsets) {
  const int mh = bottom - top, mw = right - left;
  void *p = nullptr;
  int rc = ctx_workspace(ctx, sizeof(Workspace), &p);
  if (rc) return rc;
  Workspace *ws = (Workspace *)p;
  hipStream_t s = ctx->stream;
  hipLaunchKernelGGL(k_init_state, dim3(1), dim3(64), 0, s, ws, offsets);
  S1_LAUNCH_CHECK("k_init_state");
  const s1ifg_stitch_burst &b0 = bursts[0];
  hipLaunchKernelGGL(k_place_first<T>, dim3(mh), dim3(256), 0, s, (const T *)b0.data,
                     b0.bottom - b0.top, b0.right - b0.left, b0.top - top, b0.left - left, mosaic,
                     mh, mw);
  S1_LAUNCH_CHECK("k_place_first");
  for (int i = 1; i < n; ++i) {
    const s1ifg_stitch_burst &bi = bursts[i];
    const T *bd = (const T *)bi.data;
    Win w{bi.bottom - bi.top, bi.right - bi.left, bi.top - top, bi.left - left, (long)mw};
    if (bi.estimate) {
      hipLaunchKernelGGL(k_ov_partial<T>, dim3(kOvBlocks), dim3(256), 0, s, bd, w,
                         (const T *)mosaic, ws);
      S1_LAUNCH_CHECK("k_ov_partial");
      hipLaunchKernelGGL(k_ov_finalize, dim3(1), dim3(256), 0, s, ws);
      S1_LAUNCH_CHECK("k_ov_finalize");
      const int shifts[3] = {21, 10, 0}, widths[3] = {11, 11, 10};
      for (int pss = 0; pss < 3; ++pss) {
        hipLaunchKernelGGL(k_ov_hist<T>, dim3(min(w.h, 512)), dim3(256), 0, s, bd, w,
                           (const T *)mosaic, ws, shifts[pss], widths[pss]);
        S1_LAUNCH_CHECK("k_ov_hist");
        hipLaunchKernelGGL(k_ov_select, dim3(1), dim3(256), 0, s, ws, shifts[pss],
                           widths[pss]);
        S1_LAUNCH_CHECK("k_ov_select");
      }
    }
    hipLaunchKernelGGL(k_apply<T>, dim3(w.h), dim3(256), 0, s, bd, w, mosaic,
                       (const Workspace *)ws, offsets, i);
    S1_LAUNCH_CHECK("k_apply");
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
