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

}  // namespace s1ifg
