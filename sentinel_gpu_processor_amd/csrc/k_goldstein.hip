// Goldstein adaptive filter (goldstein.cu:87-408), deterministic form.
// FP16 split mode: the fused band kernels of k_gold_mfma.hip.  FP32 mode:
//
// Pass 1: every 32x32 patch (origin min(8 p, n - 32), goldstein.cu:334-342)
// is filtered -- forward 2-D transform, 7x7 circular smoothing of |Z|
// (goldstein.cu:87-123), patch-mean normalisation (142-225), Z *= S^alpha
// where positive (125-140), inverse transform -- and stored multiplied by
// the triangular weight w = wx + wy over n^2 (227-274) in a workspace
// buffer, with FP32 radix-2 FFTs (gold_dev.h).
// Pass 2 (one thread per output pixel): sums the contributions of every
// patch covering the pixel in a fixed order and divides by the weight sum
// (280-292); pixels with weight sum 0 are 0.
//
// The reference accumulates with float atomicAdd in arbitrary order; this
// gather form gives the same sums in a fixed order, so results are
// bit-reproducible run to run.
#include "gold_dev.h"
#include "s1ifg_internal.h"

using namespace s1ifg;

namespace {

constexpr int kPatchesPerBlock = 4;

__global__ __launch_bounds__(128) void k_gold_fp32_patches(
    const float2 *__restrict__ img, int nrow, int ncol, int pi_n, int pj_n, float alpha,
    float2 *__restrict__ patches) {
  __shared__ float2 T[kPatchesPerBlock][kGoldN * kGoldPad];
  __shared__ float M[kPatchesPerBlock][kGoldN * kGoldPad];
  const int lane = threadIdx.x & 31;
  const int sub = threadIdx.x >> 5;
  const long patch = (long)blockIdx.x * kPatchesPerBlock + sub;
  const bool active = patch < (long)pi_n * pj_n;
  fp32_patch_group(img, nrow, ncol, pj_n, patch, active, alpha,
                   patches + (size_t)(active ? patch : 0) * (kGoldN * kGoldN), T[sub], M[sub],
                   lane);
}

__global__ void k_gold_gather(const float2 *__restrict__ patches, int nrow, int ncol, int pi_n,
                              int pj_n, float2 *__restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= ncol || y >= nrow) return;
  out[(size_t)y * ncol + x] = gold_gather_pixel(patches, nrow, ncol, pi_n, pj_n, y, x);
}

}  // namespace

namespace s1ifg {

int goldstein_grid(int nrow, int ncol, int *pi_n, int *pj_n) {
  // goldstein.cu:334-340
  int inc = kGoldN / 4;
  *pi_n = (int)ceilf((float)nrow / inc) - 3;
  *pj_n = (int)ceilf((float)ncol / inc) - 3;
  if (*pi_n <= 0) *pi_n = 1;
  if (*pj_n <= 0) *pj_n = 1;
  return 0;
}

size_t band_workspace_bytes(int pi_n, int pj_n);
bool band_fits(int pi_n, int pj_n);
int goldstein_band(s1ifg_ctx *ctx, const float2 *in, int nrow, int ncol, float alpha, int pi_n,
                   int pj_n, void *ws, float2 *out);

}  // namespace s1ifg

extern "C" int s1ifg_goldstein(s1ifg_ctx *ctx, const s1ifg_cf32 *in, int nrow, int ncol,
                               int n_win, float alpha, int mode, s1ifg_cf32 *out) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!in || !out) return set_error(S1IFG_EINVAL, "null data pointer");
  if (n_win != kGoldN) return set_error(S1IFG_ERANGE, "only n_win = 32 is supported (got %d)", n_win);
  if (nrow < kGoldN || ncol < kGoldN)
    return set_error(S1IFG_EINVAL, "image %d x %d smaller than the %d-pixel window", nrow, ncol,
                     kGoldN);
  if (!(alpha >= 0.f) || !std::isfinite(alpha)) return set_error(S1IFG_EINVAL, "alpha %g", alpha);
  if (mode != S1IFG_GOLD_FP32 && mode != S1IFG_GOLD_FP16_SPLIT)
    return set_error(S1IFG_EINVAL, "mode %d", mode);
  if (in == out) return set_error(S1IFG_EINVAL, "in-place filtering is not supported");
  int pi_n, pj_n;
  goldstein_grid(nrow, ncol, &pi_n, &pj_n);
  size_t npatch = (size_t)pi_n * pj_n;
  void *ws = nullptr;
  if (mode == S1IFG_GOLD_FP16_SPLIT) {
    if (!band_fits(pi_n, pj_n))
      return set_error(S1IFG_ERANGE, "image %d x %d too large for the fp16_split band sums", nrow,
                       ncol);
    if ((rc = ctx_workspace(ctx, band_workspace_bytes(pi_n, pj_n), &ws))) return rc;
    return goldstein_band(ctx, (const float2 *)in, nrow, ncol, alpha, pi_n, pj_n, ws,
                          (float2 *)out);
  }
  if ((rc = ctx_workspace(ctx, npatch * kGoldN * kGoldN * sizeof(float2), &ws))) return rc;
  float2 *patches = (float2 *)ws;
  unsigned blocks = (unsigned)((npatch + kPatchesPerBlock - 1) / kPatchesPerBlock);
  hipLaunchKernelGGL(k_gold_fp32_patches, dim3(blocks), dim3(128), 0, ctx->stream,
                     (const float2 *)in, nrow, ncol, pi_n, pj_n, alpha, patches);
  S1_LAUNCH_CHECK("k_gold_fp32_patches");
  dim3 blk(64, 4), grd((ncol + 63) / 64, (nrow + 3) / 4);
  hipLaunchKernelGGL(k_gold_gather, grd, blk, 0, ctx->stream, patches, nrow, ncol, pi_n, pj_n,
                     (float2 *)out);
  S1_LAUNCH_CHECK("k_gold_gather");
  return S1IFG_OK;
}
