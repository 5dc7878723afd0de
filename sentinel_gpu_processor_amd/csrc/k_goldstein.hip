// Goldstein adaptive filter (goldstein.cu:87-408), deterministic form.
//
// Pass 1 (one 32-lane group per patch): forward 2-D transform, 7x7
// circular smoothing of |Z| (goldstein.cu:87-123), patch-mean normalisation
// (142-225), Z *= S^alpha where positive (125-140), inverse transform, and
// the triangular weight w = wx + wy over n^2 (227-274).  The weighted patch
// goes to a workspace buffer.
// Pass 2 (one thread per output pixel): sums the contributions of every
// patch covering the pixel in a fixed order (patch row, then patch column)
// and divides by the weight sum (280-292); pixels with weight sum 0 are 0.
//
// The reference accumulates with float atomicAdd in arbitrary order; this
// gather form gives the same sums in a fixed order, so results are
// bit-reproducible run to run.
#include "s1ifg_internal.h"
#include "gold_dev.h"

using namespace s1ifg;

namespace {

constexpr int kN = 32;
constexpr int kPad = 33;
constexpr int kPatchesPerBlock = 4;

__device__ __forceinline__ void fft32_fwd(float2 v[32]) { fft32<false>(v); }
__device__ __forceinline__ void fft32_inv(float2 v[32]) { fft32<true>(v); }

__global__ __launch_bounds__(128) void k_gold_fp32_patches(
    const float2 *__restrict__ img, int nrow, int ncol, int pi_n, int pj_n, float alpha,
    float2 *__restrict__ patches) {
  __shared__ float2 T[kPatchesPerBlock][kN * kPad];
  __shared__ float M[kPatchesPerBlock][kN * kPad];
  const int lane = threadIdx.x & 31;
  const int sub = threadIdx.x >> 5;
  const long patch = (long)blockIdx.x * kPatchesPerBlock + sub;
  const bool active = patch < (long)pi_n * pj_n;
  const int pi = active ? (int)(patch / pj_n) : 0;
  const int pj = active ? (int)(patch % pj_n) : 0;
  const int oy = min(pi * (kN / 4), nrow - kN);  // goldstein.cu:308-313
  const int ox = min(pj * (kN / 4), ncol - kN);
  float2 *Ts = T[sub];
  float *Ms = M[sub];

  float2 v[32];
  // rows: lane = row index
  const float2 *src = img + (size_t)(oy + lane) * ncol + ox;
#pragma unroll
  for (int k = 0; k < 32; ++k) v[k] = active ? src[k] : make_float2(0.f, 0.f);
  fft32_fwd(v);
#pragma unroll
  for (int k = 0; k < 32; ++k) Ts[lane * kPad + k] = v[k];
  __syncthreads();
  // columns: lane = column (frequency kx)
#pragma unroll
  for (int k = 0; k < 32; ++k) v[k] = Ts[k * kPad + lane];
  fft32_fwd(v);  // v[ky] = Z[ky][lane]
#pragma unroll
  for (int k = 0; k < 32; ++k) Ms[k * kPad + lane] = sqrtf(v[k].x * v[k].x + v[k].y * v[k].y);
  __syncthreads();

  float S[32];
  smooth_column(Ms, kPad, lane, S);
  float sum = 0.f;
#pragma unroll
  for (int k = 0; k < 32; ++k) sum += S[k];
#pragma unroll
  for (int off = 16; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 32);
  const float mean = sum / (kN * kN);
  weight_spectrum(v, S, mean, alpha);

  fft32_inv(v);  // v[r] = column-inverse at spatial row r, frequency kx = lane
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; ++k) Ts[k * kPad + lane] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; ++k) v[k] = Ts[lane * kPad + k];
  fft32_inv(v);  // v[x] = n^2 * filtered spatial value at (row lane, col x)
  if (!active) return;
  float2 *dst = patches + (size_t)patch * (kN * kN) + lane * kN;
  const float wy = (float)(lane < kN / 2 ? lane : kN - 1 - lane);
  const float norm = 1.0f / (kN * kN);
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    float wx = (float)(k < kN / 2 ? k : kN - 1 - k);
    float w = (wx + wy) * norm;
    dst[k] = make_float2(v[k].x * w, v[k].y * w);
  }
}

}  // namespace

namespace s1ifg {

// Pass 2: fixed-order gather of the weighted patches.
__global__ void k_gold_gather(const float2 *__restrict__ patches, int nrow, int ncol, int pi_n,
                              int pj_n, float2 *__restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= ncol || y >= nrow) return;
  const int step = kN / 4;
  int pi_lo = y - (kN - 1) > 0 ? (y - (kN - 1) + step - 1) / step : 0;
  int pj_lo = x - (kN - 1) > 0 ? (x - (kN - 1) + step - 1) / step : 0;
  float sx = 0.f, sy = 0.f, ws = 0.f;
  for (int pi = pi_lo; pi < pi_n; ++pi) {
    int oy = min(pi * step, nrow - kN);
    if (oy > y) break;
    if (y >= oy + kN) continue;
    int ty = y - oy;
    float wy = (float)(ty < kN / 2 ? ty : kN - 1 - ty);
    for (int pj = pj_lo; pj < pj_n; ++pj) {
      int ox = min(pj * step, ncol - kN);
      if (ox > x) break;
      if (x >= ox + kN) continue;
      int tx = x - ox;
      float wx = (float)(tx < kN / 2 ? tx : kN - 1 - tx);
      float2 v = patches[((size_t)pi * pj_n + pj) * (kN * kN) + ty * kN + tx];
      sx += v.x;
      sy += v.y;
      ws += wx + wy;
    }
  }
  float2 o = make_float2(0.f, 0.f);
  if (ws > 0.f) o = make_float2(sx / ws, sy / ws);
  out[(size_t)y * ncol + x] = o;
}

int goldstein_grid(int nrow, int ncol, int *pi_n, int *pj_n) {
  // goldstein.cu:334-340
  int inc = kN / 4;
  *pi_n = (int)ceilf((float)nrow / inc) - 3;
  *pj_n = (int)ceilf((float)ncol / inc) - 3;
  if (*pi_n <= 0) *pi_n = 1;
  if (*pj_n <= 0) *pj_n = 1;
  return 0;
}

int goldstein_mfma(s1ifg_ctx *ctx, const float2 *in, int nrow, int ncol, float alpha,
                   int pi_n, int pj_n, float2 *patches);

}  // namespace s1ifg

extern "C" int s1ifg_goldstein(s1ifg_ctx *ctx, const s1ifg_cf32 *in, int nrow, int ncol,
                               int n_win, float alpha, int mode, s1ifg_cf32 *out) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!in || !out) return set_error(S1IFG_EINVAL, "null data pointer");
  if (n_win != kN) return set_error(S1IFG_ERANGE, "only n_win = 32 is supported (got %d)", n_win);
  if (nrow < kN || ncol < kN)
    return set_error(S1IFG_EINVAL, "image %d x %d smaller than the %d-pixel window", nrow, ncol, kN);
  if (!(alpha >= 0.f) || !std::isfinite(alpha)) return set_error(S1IFG_EINVAL, "alpha %g", alpha);
  if (mode != S1IFG_GOLD_FP32 && mode != S1IFG_GOLD_FP16_SPLIT)
    return set_error(S1IFG_EINVAL, "mode %d", mode);
  if (in == out) return set_error(S1IFG_EINVAL, "in-place filtering is not supported");
  int pi_n, pj_n;
  goldstein_grid(nrow, ncol, &pi_n, &pj_n);
  size_t npatch = (size_t)pi_n * pj_n;
  void *ws = nullptr;
  if ((rc = ctx_workspace(ctx, npatch * kN * kN * sizeof(float2), &ws))) return rc;
  float2 *patches = (float2 *)ws;
  if (mode == S1IFG_GOLD_FP32) {
    unsigned blocks = (unsigned)((npatch + kPatchesPerBlock - 1) / kPatchesPerBlock);
    hipLaunchKernelGGL(k_gold_fp32_patches, dim3(blocks), dim3(128), 0, ctx->stream,
                       (const float2 *)in, nrow, ncol, pi_n, pj_n, alpha, patches);
    S1_LAUNCH_CHECK("k_gold_fp32_patches");
  } else {
    if ((rc = goldstein_mfma(ctx, (const float2 *)in, nrow, ncol, alpha, pi_n, pj_n, patches)))
      return rc;
  }
  dim3 blk(64, 4), grd((ncol + 63) / 64, (nrow + 3) / 4);
  hipLaunchKernelGGL(k_gold_gather, grd, blk, 0, ctx->stream, patches, nrow, ncol, pi_n, pj_n,
                     (float2 *)out);
  S1_LAUNCH_CHECK("k_gold_gather");
  return S1IFG_OK;
}
