// Pair products on the target / multilooked grid: fused crossmul + looks,
// amplitude looks, coherence.
#include "ifg_dev.h"
#include "s1ifg_internal.h"

using namespace s1ifg;

namespace {

// conj_mul + cpx_col_look + cpx_row_look (sario.cu:199-286) fused: one thread
// per multilooked pixel, column means first, then the row mean.  Power looks
// of both images on the same cells ride along.
__global__ void k_crossmul_looks(const float2 *__restrict__ ref, long ld_ref,
                                 const float2 *__restrict__ sec, long ld_sec, int rl, int cl,
                                 int nr, int nc, float2 *__restrict__ ifg, float *__restrict__ pr,
                                 float *__restrict__ ps) {
  int cm = blockIdx.x * blockDim.x + threadIdx.x;
  int rm = blockIdx.y * blockDim.y + threadIdx.y;
  if (rm >= nr || cm >= nc) return;
  float ax = 0.f, ay = 0.f, apr = 0.f, aps = 0.f;
  const float inv_cl = 1.0f / cl, inv_rl = 1.0f / rl;
  for (int i = 0; i < rl; ++i) {
    const float2 *a = ref + (size_t)(rm * rl + i) * ld_ref + (size_t)cm * cl;
    const float2 *b = sec + (size_t)(rm * rl + i) * ld_sec + (size_t)cm * cl;
    float sx = 0.f, sy = 0.f, spr = 0.f, sps = 0.f;
    for (int j = 0; j < cl; ++j) {
      float2 u = a[j], v = b[j];
      sx += u.x * v.x + u.y * v.y;
      sy += u.x * v.y - u.y * v.x;
      spr += u.x * u.x + u.y * u.y;
      sps += v.x * v.x + v.y * v.y;
    }
    ax += sx * inv_cl;
    ay += sy * inv_cl;
    apr += spr * inv_cl;
    aps += sps * inv_cl;
  }
  size_t o = (size_t)rm * nc + cm;
  ifg[o] = make_float2(ax * inv_rl, ay * inv_rl);
  if (pr) pr[o] = apr * inv_rl;
  if (ps) ps[o] = aps * inv_rl;
}

// cpx_col_look + cpx_row_look (sario.cu:241-286) on one image.
__global__ void k_looks(const float2 *__restrict__ s, long ld, int rl, int cl, int nr, int nc,
                        float2 *__restrict__ out) {
  int cm = blockIdx.x * blockDim.x + threadIdx.x;
  int rm = blockIdx.y * blockDim.y + threadIdx.y;
  if (rm >= nr || cm >= nc) return;
  const float inv_cl = 1.0f / cl, inv_rl = 1.0f / rl;
  float ax = 0.f, ay = 0.f;
  for (int i = 0; i < rl; ++i) {
    const float2 *a = s + (size_t)(rm * rl + i) * ld + (size_t)cm * cl;
    float sx = 0.f, sy = 0.f;
    for (int j = 0; j < cl; ++j) {
      sx += a[j].x;
      sy += a[j].y;
    }
    ax += sx * inv_cl;
    ay += sy * inv_cl;
  }
  out[(size_t)rm * nc + cm] = make_float2(ax * inv_rl, ay * inv_rl);
}

// point_angle (sario.cu:231-239).
__global__ void k_phase(const float2 *__restrict__ z, long n, float *__restrict__ out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = atan2f(z[i].y, z[i].x);
}

// multilook/powlooks (coherence.py:46-48, sario.py:416-420).
__global__ void k_amplitude(const float2 *__restrict__ s, long ld, int rl, int cl, int nr,
                            int nc, float *__restrict__ amp) {
  int cm = blockIdx.x * blockDim.x + threadIdx.x;
  int rm = blockIdx.y * blockDim.y + threadIdx.y;
  if (rm >= nr || cm >= nc) return;
  size_t o = (size_t)rm * nc + cm;
  if (rl == 1 && cl == 1) {  // sario.py:417-418: |s|, then sqrt
    float2 v = s[(size_t)rm * ld + cm];
    amp[o] = sqrtf(sqrtf(v.x * v.x + v.y * v.y));
    return;
  }
  const float inv_cl = 1.0f / cl, inv_rl = 1.0f / rl;
  float acc = 0.f;
  for (int i = 0; i < rl; ++i) {
    const float2 *a = s + (size_t)(rm * rl + i) * ld + (size_t)cm * cl;
    float sp = 0.f;
    for (int j = 0; j < cl; ++j) sp += a[j].x * a[j].x + a[j].y * a[j].y;
    acc += sp * inv_cl;
  }
  amp[o] = sqrtf(acc * inv_rl);
}

// coherence.py:142-144 (win = 1) and the windowed estimator (win >= 3).
__global__ void k_coherence(const float2 *__restrict__ ifg, const float *__restrict__ p1,
                            const float *__restrict__ p2, int nr, int nc, int win, int fill,
                            uint32_t seed, float *__restrict__ coh) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= nr || c >= nc) return;
  size_t o = (size_t)r * nc + c;
  float g;
  if (win == 1) {
    g = coherence_cell(ifg[o], p1[o], p2[o]);
  } else {
    int h = win / 2;
    int r0 = max(r - h, 0), r1 = min(r + h, nr - 1);
    int c0 = max(c - h, 0), c1 = min(c + h, nc - 1);
    float sx = 0.f, sy = 0.f, s1 = 0.f, s2 = 0.f;
    for (int i = r0; i <= r1; ++i) {
      size_t b = (size_t)i * nc;
      for (int j = c0; j <= c1; ++j) {
        float2 v = ifg[b + j];
        sx += v.x;
        sy += v.y;
        s1 += p1[b + j];
        s2 += p2[b + j];
      }
    }
    float den = sqrtf(s1 * s2);
    g = den > 0.f ? sqrtf(sx * sx + sy * sy) / den : 0.f;
  }
  coh[o] = coh_fill(g, fill, seed, o);
}

// Windowed coherence with the 64 x 4 output tile and its halo staged in LDS
// (each input element is read from HBM about once); same sums, same order
// (rows outer, columns inner) as k_coherence.
__global__ __launch_bounds__(256) void k_coherence_tiled(
    const float2 *__restrict__ ifg, const float *__restrict__ p1, const float *__restrict__ p2,
    int nr, int nc, int win, int fill, uint32_t seed, float *__restrict__ coh) {
  __shared__ CohTile L;
  coherence_tile(ifg, p1, p2, nr, nc, win, fill, seed, coh, blockIdx.x, blockIdx.y, L);
}

// Windowed coherence over 64 x kStripRows strips with separable window
// sums (coherence_strip).
__global__ __launch_bounds__(256) void k_coherence_strip(
    const float2 *__restrict__ ifg, const float *__restrict__ p1, const float *__restrict__ p2,
    int nr, int nc, int win, int fill, uint32_t seed, float *__restrict__ coh) {
  __shared__ CohStrip L;
  coherence_strip(ifg, p1, p2, nr, nc, win, fill, seed, coh, blockIdx.x, blockIdx.y, L);
}

}  // namespace

extern "C" {

int s1ifg_crossmul_looks(s1ifg_ctx *ctx, const s1ifg_cf32 *ref, int64_t ld_ref,
                         const s1ifg_cf32 *sec, int64_t ld_sec, int nrow, int ncol, int rl,
                         int cl, s1ifg_cf32 *ifg_ml, float *pow_ref, float *pow_sec) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!ref || !sec || !ifg_ml) return set_error(S1IFG_EINVAL, "null data pointer");
  if (rl < 1 || cl < 1) return set_error(S1IFG_EINVAL, "looks %d x %d", rl, cl);
  if (nrow < rl || ncol < cl) return set_error(S1IFG_EINVAL, "image %d x %d smaller than one look", nrow, ncol);
  if (ld_ref < ncol || ld_sec < ncol) return set_error(S1IFG_EINVAL, "row stride < ncol");
  int nr = nrow / rl, nc = ncol / cl;
  dim3 blk(64, 4), grd((nc + 63) / 64, (nr + 3) / 4);
  hipLaunchKernelGGL(k_crossmul_looks, grd, blk, 0, ctx->stream, (const float2 *)ref,
                     (long)ld_ref, (const float2 *)sec, (long)ld_sec, rl, cl, nr, nc,
                     (float2 *)ifg_ml, pow_ref, pow_sec);
  S1_LAUNCH_CHECK("k_crossmul_looks");
  return S1IFG_OK;
}

int s1ifg_looks(s1ifg_ctx *ctx, const s1ifg_cf32 *in, int64_t ld, int nrow, int ncol, int rl,
                int cl, s1ifg_cf32 *out) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!in || !out) return set_error(S1IFG_EINVAL, "null data pointer");
  if (rl < 1 || cl < 1) return set_error(S1IFG_EINVAL, "looks %d x %d", rl, cl);
  if (nrow < rl || ncol < cl || ld < ncol) return set_error(S1IFG_EINVAL, "bad image size / stride");
  int nr = nrow / rl, nc = ncol / cl;
  dim3 blk(64, 4), grd((nc + 63) / 64, (nr + 3) / 4);
  hipLaunchKernelGGL(k_looks, grd, blk, 0, ctx->stream, (const float2 *)in, (long)ld, rl, cl, nr,
                     nc, (float2 *)out);
  S1_LAUNCH_CHECK("k_looks");
  return S1IFG_OK;
}

int s1ifg_phase(s1ifg_ctx *ctx, const s1ifg_cf32 *in, int64_t n, float *out) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (n < 0 || (n > 0 && (!in || !out))) return set_error(S1IFG_EINVAL, "bad arguments");
  if (n == 0) return S1IFG_OK;
  hipLaunchKernelGGL(k_phase, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     (const float2 *)in, (long)n, out);
  S1_LAUNCH_CHECK("k_phase");
  return S1IFG_OK;
}

int s1ifg_amplitude(s1ifg_ctx *ctx, const s1ifg_cf32 *slc, int64_t ld, int nrow, int ncol,
                    int rl, int cl, float *amp) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!slc || !amp) return set_error(S1IFG_EINVAL, "null data pointer");
  if (rl < 1 || cl < 1) return set_error(S1IFG_EINVAL, "looks %d x %d", rl, cl);
  if (nrow < rl || ncol < cl || ld < ncol) return set_error(S1IFG_EINVAL, "bad image size / stride");
  int nr = nrow / rl, nc = ncol / cl;
  dim3 blk(64, 4), grd((nc + 63) / 64, (nr + 3) / 4);
  hipLaunchKernelGGL(k_amplitude, grd, blk, 0, ctx->stream, (const float2 *)slc, (long)ld, rl, cl,
                     nr, nc, amp);
  S1_LAUNCH_CHECK("k_amplitude");
  return S1IFG_OK;
}

int s1ifg_coherence(s1ifg_ctx *ctx, const s1ifg_cf32 *ifg_ml, const float *pow_ref,
                    const float *pow_sec, int nrow, int ncol, int win, int fill, uint32_t seed,
                    float *coh) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!ifg_ml || !pow_ref || !pow_sec || !coh) return set_error(S1IFG_EINVAL, "null data pointer");
  if (nrow < 1 || ncol < 1) return set_error(S1IFG_EINVAL, "image %d x %d", nrow, ncol);
  if (win < 1 || win % 2 == 0 || win > 63) return set_error(S1IFG_EINVAL, "window %d (odd, 1..63)", win);
  if (fill < S1IFG_COH_FILL_CLAMP || fill > S1IFG_COH_FILL_HASH)
    return set_error(S1IFG_EINVAL, "fill rule %d", fill);
  dim3 blk(64, 4), grd((ncol + 63) / 64, (nrow + 3) / 4);
  if (win > 1 && win / 2 <= kStripMaxHalo) {
    dim3 sgrd((ncol + kCohTX - 1) / kCohTX, (nrow + kStripRows - 1) / kStripRows);
    hipLaunchKernelGGL(k_coherence_strip, sgrd, blk, 0, ctx->stream, (const float2 *)ifg_ml,
                       pow_ref, pow_sec, nrow, ncol, win, fill, seed, coh);
    S1_LAUNCH_CHECK("k_coherence_strip");
    return S1IFG_OK;
  }
  if (win > 1 && win / 2 <= kCohMaxHalo) {
    hipLaunchKernelGGL(k_coherence_tiled, grd, blk, 0, ctx->stream, (const float2 *)ifg_ml,
                       pow_ref, pow_sec, nrow, ncol, win, fill, seed, coh);
    S1_LAUNCH_CHECK("k_coherence_tiled");
    return S1IFG_OK;
  }
  hipLaunchKernelGGL(k_coherence, grd, blk, 0, ctx->stream, (const float2 *)ifg_ml, pow_ref,
                     pow_sec, nrow, ncol, win, fill, seed, coh);
  S1_LAUNCH_CHECK("k_coherence");
  return S1IFG_OK;
}

}  // extern "C"
