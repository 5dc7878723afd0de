// Goldstein filter in FP16 hi/lo split on MFMA with the overlap-add fused
// into the patch pass (gold_band_dev.h): the filtered patches never leave
// the chip; per-band partial sums do.
#include "gold_band_dev.h"

namespace s1ifg {
namespace {

// Persistent wavefronts over the (band, chunk) items (no block barriers),
// the items after the first one per wavefront from the queue counter ctr.
__global__ __launch_bounds__(64 * kBandWaves) __attribute__((amdgpu_waves_per_eu(2))) void k_gold_band(
    const float2 *__restrict__ img, int nrow, int ncol, int pi_n, int pj_n, float alpha,
    float2 *__restrict__ partial, unsigned *ctr) {
  __shared__ BandLds L;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nchunk = band_nchunk(pj_n);
  const long total = band_items(pi_n, pj_n);
  const long nw = (long)gridDim.x * kBandWaves;
  long t = (long)blockIdx.x * kBandWaves + wave;
  if (t >= total) return;
  mf::Consts K;
  mf::build_consts(lane >> 5, lane & 31, K, L.u[wave]);
  wave_sync_lds();  // L.u[wave] written
  for (; t < total; t = band_next(ctr, nw))
    band_item(K, L.u[wave], img, nrow, ncol, pi_n, pj_n, (int)(t / nchunk), (int)(t % nchunk), alpha,
              partial, L.buf[wave]);
}

// The queue counter, zeroed by a kernel (not a memset: graph replays).
__global__ void k_band_reset(unsigned *ctr) {
  if (threadIdx.x == 0) *ctr = 0u;
}

__global__ __launch_bounds__(kSumTX * kSumTY) void k_gold_band_sum(
    const float2 *__restrict__ partial, int nrow, int ncol, int pi_n, int pj_n,
    float2 *__restrict__ out) {
  __shared__ BandSumLds L;
  band_sum_tile(partial, nrow, ncol, pi_n, pj_n, blockIdx.x, blockIdx.y, out, L);
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

size_t band_workspace_bytes(int pi_n, int pj_n) {
  return band_counter_offset(pi_n, pj_n) + 256;
}

bool band_fits(int pi_n, int pj_n) { return band_offsets_fit(pi_n, pj_n); }

int goldstein_band(s1ifg_ctx *ctx, const float2 *in, int nrow, int ncol, float alpha, int pi_n,
                   int pj_n, void *ws, float2 *out) {
  float2 *partial = (float2 *)ws;
  unsigned *ctr = (unsigned *)((char *)ws + band_counter_offset(pi_n, pj_n));
  const long items = band_items(pi_n, pj_n);
  const long want = (items + kBandWaves - 1) / kBandWaves;
  const long cap = 2L * device_cu_count(ctx->device);
  const unsigned blocks = (unsigned)(want < cap ? want : cap);
  hipLaunchKernelGGL(k_band_reset, dim3(1), dim3(64), 0, ctx->stream, ctr);
  S1_LAUNCH_CHECK("k_band_reset");
  hipLaunchKernelGGL(k_gold_band, dim3(blocks), dim3(64 * kBandWaves), 0, ctx->stream, in, nrow, ncol, pi_n,
                     pj_n, alpha, partial, ctr);
  S1_LAUNCH_CHECK("k_gold_band");
  dim3 blk(kSumTX, kSumTY), grd((ncol + kSumTX - 1) / kSumTX, (nrow + kSumTY - 1) / kSumTY);
  hipLaunchKernelGGL(k_gold_band_sum, grd, blk, 0, ctx->stream, partial, nrow, ncol, pi_n, pj_n,
                     out);
  S1_LAUNCH_CHECK("k_gold_band_sum");
  return S1IFG_OK;
}

}  // namespace s1ifg
