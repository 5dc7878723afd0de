// Device helpers of the Goldstein filter shared by the FP32 and MFMA paths.
#pragma once

#include "s1ifg_internal.h"

namespace s1ifg {

// goldstein.cu:43-52: the reference's 7x7 smoothing table (row = dy + 3,
// column = dx + 3), kept as project data.
__device__ constexpr float kGauss7[49] = {
    0.000004f, 0.000088f, 0.000543f, 0.000948f, 0.000543f, 0.000088f, 0.000004f,
    0.000088f, 0.001915f, 0.011776f, 0.020584f, 0.011776f, 0.001915f, 0.000088f,
    0.000543f, 0.011776f, 0.072410f, 0.126584f, 0.072410f, 0.011776f, 0.000543f,
    0.000948f, 0.020584f, 0.126584f, 0.221141f, 0.126584f, 0.020584f, 0.000948f,
    0.000543f, 0.011776f, 0.072410f, 0.126584f, 0.072410f, 0.011776f, 0.000543f,
    0.000088f, 0.001915f, 0.011776f, 0.020584f, 0.011776f, 0.001915f, 0.000088f,
    0.000004f, 0.000088f, 0.000543f, 0.000948f, 0.000543f, 0.000088f, 0.000004f};

// cos / sin of 2 pi m / 32, m = 0..15
__device__ constexpr float kCos32[16] = {
    1.000000000e+00f, 9.807852804e-01f, 9.238795325e-01f, 8.314696123e-01f,
    7.071067812e-01f, 5.555702330e-01f, 3.826834324e-01f, 1.950903220e-01f,
    0.0f,             -1.950903220e-01f, -3.826834324e-01f, -5.555702330e-01f,
    -7.071067812e-01f, -8.314696123e-01f, -9.238795325e-01f, -9.807852804e-01f};
__device__ constexpr float kSin32[16] = {
    0.000000000e+00f, 1.950903220e-01f, 3.826834324e-01f, 5.555702330e-01f,
    7.071067812e-01f, 8.314696123e-01f, 9.238795325e-01f, 9.807852804e-01f,
    1.000000000e+00f, 9.807852804e-01f, 9.238795325e-01f, 8.314696123e-01f,
    7.071067812e-01f, 5.555702330e-01f, 3.826834324e-01f, 1.950903220e-01f};

__device__ constexpr int brev5(int i) {
  return ((i & 1) << 4) | ((i & 2) << 2) | (i & 4) | ((i & 8) >> 2) | ((i & 16) >> 4);
}

// In-register radix-2 decimation-in-time FFT of length 32, unnormalised.
// INV = false: exp(-i ...) (forward); INV = true: exp(+i ...).
template <bool INV>
__device__ __forceinline__ void fft32(float2 v[32]) {
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    int j = brev5(i);
    if (j > i) {
      float2 t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
#pragma unroll
  for (int len = 2; len <= 32; len <<= 1) {
    const int half = len >> 1;
    const int tstep = 32 / len;
#pragma unroll
    for (int i = 0; i < 32; i += len) {
#pragma unroll
      for (int k = 0; k < half; ++k) {
        const float wr = kCos32[k * tstep];
        const float wi = INV ? kSin32[k * tstep] : -kSin32[k * tstep];
        float2 a = v[i + k];
        float2 b = v[i + k + half];
        float2 bw = make_float2(b.x * wr - b.y * wi, b.x * wi + b.y * wr);
        v[i + k] = make_float2(a.x + bw.x, a.y + bw.y);
        v[i + k + half] = make_float2(a.x - bw.x, a.y - bw.y);
      }
    }
  }
}

// S[ky] for one spectral column kx = c: circular 7x7 correlation of the
// magnitude plane M (row stride `ld`) with kGauss7, summed in the
// reference's order (dy outer, dx inner; goldstein.cu:105-119), using a
// register window that slides down the column.
__device__ __forceinline__ void smooth_column(const float *M, int ld, int c, float S[32]) {
  int cx[7];
#pragma unroll
  for (int d = 0; d < 7; ++d) cx[d] = (c + d - 3) & 31;
  float w[7][7];
#pragma unroll
  for (int d = 0; d < 7; ++d) {
    const int row = (d - 3) & 31;
#pragma unroll
    for (int e = 0; e < 7; ++e) w[d][e] = M[row * ld + cx[e]];
  }
#pragma unroll
  for (int ky = 0; ky < 32; ++ky) {
    float s = 0.f;
#pragma unroll
    for (int d = 0; d < 7; ++d)
#pragma unroll
      for (int e = 0; e < 7; ++e) s = fmaf(w[d][e], kGauss7[d * 7 + e], s);
    S[ky] = s;
    if (ky < 31) {
#pragma unroll
      for (int d = 0; d < 6; ++d)
#pragma unroll
        for (int e = 0; e < 7; ++e) w[d][e] = w[d + 1][e];
      const int row = (ky + 4) & 31;
#pragma unroll
      for (int e = 0; e < 7; ++e) w[6][e] = M[row * ld + cx[e]];
    }
  }
}

// goldstein.cu:125-140, 166-175: S <- S/mean (0 when mean <= 1e-8),
// Z <- Z * S^alpha where S^alpha > 0.
__device__ __forceinline__ void weight_spectrum(float2 v[32], const float S[32], float mean,
                                                float alpha) {
  const bool ok = mean > 1e-8f;
  const float inv = ok ? 1.0f / mean : 0.f;
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    float sn = ok ? S[k] * inv : 0.f;
    float w = powf(sn, alpha);
    if (w > 0.0f) {
      v[k].x *= w;
      v[k].y *= w;
    }
  }
}

constexpr int kGoldN = 32;
constexpr int kGoldPad = 33;

// FP32 patch filter for one 32-lane group (row/column radix-2 FFTs in
// registers, transposes through the group's LDS planes T and M).  Every
// thread of the block must call it (it contains block barriers); `active`
// masks groups without a patch.
__device__ __forceinline__ void fp32_patch_group(const float2 *__restrict__ img, int nrow,
                                                 int ncol, int pj_n, long patch, bool active,
                                                 float alpha, float2 *__restrict__ dst,
                                                 float2 *Ts, float *Ms, int lane) {
  const int pi = active ? (int)(patch / pj_n) : 0;
  const int pj = active ? (int)(patch % pj_n) : 0;
  const int oy = active ? min(pi * (kGoldN / 4), nrow - kGoldN) : 0;  // goldstein.cu:308-313
  const int ox = active ? min(pj * (kGoldN / 4), ncol - kGoldN) : 0;
  float2 v[32];
  const float2 *src = img + (size_t)(oy + lane) * ncol + ox;
#pragma unroll
  for (int k = 0; k < 32; ++k) v[k] = active ? src[k] : make_float2(0.f, 0.f);
  fft32<false>(v);
#pragma unroll
  for (int k = 0; k < 32; ++k) Ts[lane * kGoldPad + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; ++k) v[k] = Ts[k * kGoldPad + lane];
  fft32<false>(v);  // v[ky] = Z[ky][lane]
#pragma unroll
  for (int k = 0; k < 32; ++k) Ms[k * kGoldPad + lane] = sqrtf(v[k].x * v[k].x + v[k].y * v[k].y);
  __syncthreads();
  float S[32];
  smooth_column(Ms, kGoldPad, lane, S);
  float sum = 0.f;
#pragma unroll
  for (int k = 0; k < 32; ++k) sum += S[k];
#pragma unroll
  for (int off = 16; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 32);
  const float mean = sum / (kGoldN * kGoldN);
  weight_spectrum(v, S, mean, alpha);
  fft32<true>(v);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; ++k) Ts[k * kGoldPad + lane] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; ++k) v[k] = Ts[lane * kGoldPad + k];
  fft32<true>(v);  // v[x] = n^2 * filtered value at (row lane, col x)
  if (!active) return;
  float2 *d = dst + lane * kGoldN;
  const float wy = (float)(lane < kGoldN / 2 ? lane : kGoldN - 1 - lane);
  const float norm = 1.0f / (kGoldN * kGoldN);
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    float wx = (float)(k < kGoldN / 2 ? k : kGoldN - 1 - k);
    float w = (wx + wy) * norm;
    d[k] = make_float2(v[k].x * w, v[k].y * w);
  }
}

// Output pixel (y, x): sum of the weighted patches covering it in a fixed
// order (patch row, then patch column) over the weight sum
// (goldstein.cu:257-292); 0 where the weight sum is 0.
__device__ __forceinline__ float2 gold_gather_pixel(const float2 *__restrict__ patches, int nrow,
                                                    int ncol, int pi_n, int pj_n, int y, int x) {
  const int step = kGoldN / 4;
  int pi_lo = y - (kGoldN - 1) > 0 ? (y - (kGoldN - 1) + step - 1) / step : 0;
  int pj_lo = x - (kGoldN - 1) > 0 ? (x - (kGoldN - 1) + step - 1) / step : 0;
  float sx = 0.f, sy = 0.f, ws = 0.f;
  for (int pi = pi_lo; pi < pi_n; ++pi) {
    int oy = min(pi * step, nrow - kGoldN);
    if (oy > y) break;
    if (y >= oy + kGoldN) continue;
    int ty = y - oy;
    float wy = (float)(ty < kGoldN / 2 ? ty : kGoldN - 1 - ty);
    for (int pj = pj_lo; pj < pj_n; ++pj) {
      int ox = min(pj * step, ncol - kGoldN);
      if (ox > x) break;
      if (x >= ox + kGoldN) continue;
      int tx = x - ox;
      float wx = (float)(tx < kGoldN / 2 ? tx : kGoldN - 1 - tx);
      float2 v = patches[((size_t)pi * pj_n + pj) * (kGoldN * kGoldN) + ty * kGoldN + tx];
      sx += v.x;
      sy += v.y;
      ws += wx + wy;
    }
  }
  return ws > 0.f ? make_float2(sx / ws, sy / ws) : make_float2(0.f, 0.f);
}

}  // namespace s1ifg
