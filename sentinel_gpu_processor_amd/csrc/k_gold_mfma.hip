// Goldstein patch pass on MFMA for one image (the filter itself is in
// gold_mfma_dev.h).
#include "gold_mfma_dev.h"

namespace s1ifg {
namespace {

constexpr int kWaves = 4;

// Persistent wavefronts: each wave builds the constant operands once and
// filters patches wave_id, wave_id + nwaves, ... with the next patch
// prefetched (the loop bound is uniform per wavefront; no block barriers).
__global__ __launch_bounds__(256) void k_gold_mfma_patches(const float2 *__restrict__ img,
                                                           int nrow, int ncol, int pi_n,
                                                           int pj_n, float alpha,
                                                           float2 *__restrict__ patches) {
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5, n = lane & 31;
  const long total = (long)pi_n * pj_n;
  const long nwaves = (long)gridDim.x * kWaves;
  long patch = (long)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (patch >= total) return;
  mf::Consts K;
  mf::build_consts(h, n, K);
  float2 nxt[16];
  mf::load_patch(img, nrow, ncol, pj_n, patch, h, n, nxt);
  for (; patch < total; patch += nwaves) {
    float2 x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) x[r] = nxt[r];
    if (patch + nwaves < total) mf::load_patch(img, nrow, ncol, pj_n, patch + nwaves, h, n, nxt);
    mf::filter_patch(K, x, h, n, alpha, patches + (size_t)patch * (mf::kN * mf::kN));
  }
}

}  // namespace

int device_cu_count(int device) {
  static int ncu[64] = {0};
  int dev = (device >= 0 && device < 64) ? device : 0;
  if (!ncu[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        v <= 0)
      v = 256;
    ncu[dev] = v;
  }
  return ncu[dev];
}

int goldstein_mfma(s1ifg_ctx *ctx, const float2 *in, int nrow, int ncol, float alpha, int pi_n,
                   int pj_n, float2 *patches) {
  long total = (long)pi_n * pj_n;
  long want = (total + kWaves - 1) / kWaves;
  long cap = 2L * device_cu_count(ctx->device);
  unsigned blocks = (unsigned)(want < cap ? want : cap);
  hipLaunchKernelGGL(k_gold_mfma_patches, dim3(blocks), dim3(64 * kWaves), 0, ctx->stream, in,
                     nrow, ncol, pi_n, pj_n, alpha, patches);
  S1_LAUNCH_CHECK("k_gold_mfma_patches");
  return S1IFG_OK;
}

}  // namespace s1ifg
