// Goldstein patch filter on MFMA (gfx950): DFT-by-GEMM with hi/lo split FP16
// operands and FP32 accumulation (v_mfma_f32_32x32x16_f16).  Device code
// shared by the per-image and batched kernels.
//
// One wavefront filters one 32x32 patch entirely in registers:
//   Y  = X F          (row transform)       Z^T = Y^T F   (column transform)
//   S  = l1 B1^T M B1 + l2 B2^T M B2       (7x7 circular smoothing of M=|Z|)
//   Zw = Z * (S/mean(S))^alpha where positive
//   Y' = Zw conj(F)   z'^T = Y'^T conj(F)  (unnormalised inverse)
// F is the 32-point DFT matrix; B_r are banded circulant matrices built from
// the two leading eigenvectors of the reference's 7x7 smoothing table
// (goldstein.cu:43-52); the omitted eigen-terms are below 1.3e-6 of the
// table maximum.
//
// Lane layout.  Lane l (h = l/32, n = l%32) holds 16 values of a matrix row
// n at columns col(r) = 4h + 8(r/4) + (r%4), r = 0..15.  This is exactly the
// row set of the 32x32 MFMA result held by a lane, so every product's
// result feeds the next product's A operand without data movement; the
// K index of MFMA k-slot (8h + j) in chunk c is
// kappa(c,h,j) = 16c + 8(j/4) + 4h + (j%4), and the constant B operands are
// generated in that order.
//
// Split.  x = hi + lo/2^11 with hi = fp16(x), lo = fp16((x - hi) 2^11);
// products keep hi*hi + (hi*lo + lo*hi)/2^11 (lo*lo dropped), accumulated
// in FP32.  Each patch is scaled by a power of two to max |x| < 1 first.
#pragma once

#include "gold_dev.h"
#include "s1ifg_internal.h"

namespace s1ifg {
namespace mf {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kN = 32;
constexpr float kLoScale = 2048.0f;
constexpr float kLoInv = 1.0f / 2048.0f;
// leading eigen-pairs of the 7x7 smoothing table (even eigenvectors)
constexpr float kLam1 = 3.698581918e-01f;
constexpr float kLam2 = -5.889286314e-05f;
__device__ constexpr float kU1[7] = {3.315716208e-03f, 7.196210134e-02f, 4.425208916e-01f,
                                     7.732861339e-01f, 4.425208916e-01f, 7.196210134e-02f,
                                     3.315716208e-03f};
__device__ constexpr float kU2[7] = {8.714356827e-03f, -6.391971615e-02f, -5.432523281e-01f,
                                     6.335854132e-01f, -5.432523281e-01f, -6.391971615e-02f,
                                     8.714356827e-03f};

__device__ __forceinline__ f32x16 mfma(half8 a, half8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int kappa(int c, int h, int j) {
  return 16 * c + 8 * (j >> 2) + 4 * h + (j & 3);
}

__device__ __forceinline__ void split16(const float v[16], half8 hi[2], half8 lo[2]) {
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      float x = v[8 * c + j];
      _Float16 xh = (_Float16)x;
      hi[c][j] = xh;
      lo[c][j] = (_Float16)((x - (float)xh) * kLoScale);
    }
}

__device__ __forceinline__ void to_half16(const float v[16], half8 hi[2]) {
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int j = 0; j < 8; ++j) hi[c][j] = (_Float16)v[8 * c + j];
}

struct CplxOp {  // constant complex B operand, split, per chunk
  half8 rh[2], rl[2], ih[2], il[2];
};

struct Consts {
  CplxOp F;
  half8 u1h[2], u1l[2], u2h[2];
};

__device__ __forceinline__ void split1(float x, _Float16 &h, _Float16 &l) {
  h = (_Float16)x;
  l = (_Float16)((x - (float)h) * kLoScale);
}

// Constant B operands of lane (h, n) in kappa order.
__device__ __forceinline__ void build_consts(int h, int n, Consts &K) {
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = kappa(c, h, j);
      const int m = (k * n) & 31;
      const float cs = m < 16 ? kCos32[m] : -kCos32[m - 16];
      const float sn = m < 16 ? kSin32[m] : -kSin32[m - 16];
      _Float16 a, b;
      split1(cs, a, b);
      K.F.rh[c][j] = a;
      K.F.rl[c][j] = b;
      split1(-sn, a, b);  // Im F = -sin
      K.F.ih[c][j] = a;
      K.F.il[c][j] = b;
      const int d = ((k - n + 16) & 31) - 16;  // circular offset
      const float b1 = (d >= -3 && d <= 3) ? kU1[d + 3] : 0.f;
      const float b2 = (d >= -3 && d <= 3) ? kU2[d + 3] : 0.f;
      split1(b1, a, b);
      K.u1h[c][j] = a;
      K.u1l[c][j] = b;
      K.u2h[c][j] = (_Float16)b2;
    }
}

// C = A * B (B = F) or A * conj(B) (CONJ), A given by split real/imag
// parts, B constant.  The minus signs of the complex product go on a
// negated copy of A's imaginary (B = F) or real (CONJ) part.  Four
// accumulator chains: the hi*hi products of the real and imaginary parts,
// and the hi*lo + lo*hi products of each (both carry the 2^11 scale of the
// lo parts), interleaved so that consecutive MFMAs do not wait on each
// other.
template <bool CONJ>
__device__ __forceinline__ void cgemm(const half8 arh[2], const half8 arl[2], const half8 aih[2],
                                      const half8 ail[2], const CplxOp &b, float cr[16],
                                      float ci[16]) {
  f32x16 rm = {}, im = {}, rc = {}, ic = {};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (!CONJ) {
      // Cr = Ar Fr - Ai Fi, Ci = Ar Fi + Ai Fr
      const half8 nih = -aih[c], nil = -ail[c];
      rm = mfma(arh[c], b.rh[c], rm);
      im = mfma(arh[c], b.ih[c], im);
      rc = mfma(arh[c], b.rl[c], rc);
      ic = mfma(arh[c], b.il[c], ic);
      rm = mfma(nih, b.ih[c], rm);
      im = mfma(aih[c], b.rh[c], im);
      rc = mfma(arl[c], b.rh[c], rc);
      ic = mfma(arl[c], b.ih[c], ic);
      rc = mfma(nih, b.il[c], rc);
      ic = mfma(aih[c], b.rl[c], ic);
      rc = mfma(nil, b.ih[c], rc);
      ic = mfma(ail[c], b.rh[c], ic);
    } else {
      // Cr = Ar Fr + Ai Fi, Ci = Ai Fr - Ar Fi
      const half8 nrh = -arh[c], nrl = -arl[c];
      rm = mfma(arh[c], b.rh[c], rm);
      im = mfma(nrh, b.ih[c], im);
      rc = mfma(arh[c], b.rl[c], rc);
      ic = mfma(nrh, b.il[c], ic);
      rm = mfma(aih[c], b.ih[c], rm);
      im = mfma(aih[c], b.rh[c], im);
      rc = mfma(arl[c], b.rh[c], rc);
      ic = mfma(nrl, b.ih[c], ic);
      rc = mfma(aih[c], b.il[c], rc);
      ic = mfma(aih[c], b.rl[c], ic);
      rc = mfma(ail[c], b.ih[c], rc);
      ic = mfma(ail[c], b.rh[c], ic);
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    cr[r] = fmaf(rc[r], kLoInv, rm[r]);
    ci[r] = fmaf(ic[r], kLoInv, im[r]);
  }
}

template <bool CONJ>
__device__ __forceinline__ void cgemm_v(const float xr[16], const float xi[16], const CplxOp &b,
                                        float yr[16], float yi[16]) {
  half8 arh[2], arl[2], aih[2], ail[2];
  split16(xr, arh, arl);
  split16(xi, aih, ail);
  cgemm<CONJ>(arh, arl, aih, ail, b, yr, yi);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Patch p of an image: row n, columns col(r) (goldstein.cu:298-323 gather).
__device__ __forceinline__ void load_patch(const float2 *__restrict__ img, int nrow, int ncol,
                                           int pj_n, long p, int h, int n, float2 v[16]) {
  const int pi = (int)(p / pj_n), pj = (int)(p % pj_n);
  const int oy = min(pi * (kN / 4), nrow - kN);  // goldstein.cu:308-313
  const int ox = min(pj * (kN / 4), ncol - kN);
  const float2 *src = img + (size_t)(oy + n) * ncol + ox + 4 * h;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int t = 0; t < 4; ++t) v[4 * q + t] = src[8 * q + t];
}

// Filter one patch held in x (lane layout) into its weighted result
// w(row, col) z'/n^2 in the same layout: lane (h, n) holds row n, columns
// 4h + 8q + t (q, t = 0..3) in out[4q + t].
__device__ __forceinline__ void filter_patch_regs(const Consts &K, const float2 x[16], int h,
                                                  int n, float alpha, float2 out[16]) {
  float xr[16], xi[16];
  float mx = 0.f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    xr[r] = x[r].x;
    xi[r] = x[r].y;
    mx = fmaxf(mx, fmaxf(fabsf(xr[r]), fabsf(xi[r])));
  }
  mx = wave_max(mx);
  int e = 0;
  if (mx > 0.f) frexpf(mx, &e);
  const float scale = ldexpf(1.0f, -e);  // max |x| * scale in [0.5, 1)
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    xr[r] *= scale;
    xi[r] *= scale;
  }
  // forward transform
  float yr[16], yi[16], zr[16], zi[16];
  cgemm_v<false>(xr, xi, K.F, yr, yi);
  cgemm_v<false>(yr, yi, K.F, zr, zi);
  // smoothing of |Z|
  float m[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) m[r] = sqrtf(zr[r] * zr[r] + zi[r] * zi[r]);
  half8 mh[2], ml[2];
  split16(m, mh, ml);
  f32x16 t1m = {}, t1c = {}, t2 = {};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    t1m = mfma(mh[c], K.u1h[c], t1m);
    t1c = mfma(mh[c], K.u1l[c], t1c);
    t2 = mfma(mh[c], K.u2h[c], t2);
    t1c = mfma(ml[c], K.u1h[c], t1c);
  }
  // S = l1 (B1^T M B1 + (l2/l1) B2^T M B2): the second pass runs on B_r,
  // l2/l1 scales the rank-2 term's data and l1 the sum
  float t1[16], t2s[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    t1[r] = fmaf(t1c[r], kLoInv, t1m[r]);
    t2s[r] = t2[r] * (kLam2 / kLam1);
  }
  half8 t1h[2], t1l[2], t2h[2];
  split16(t1, t1h, t1l);
  to_half16(t2s, t2h);
  f32x16 sm = {}, sc = {}, sm2 = {};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    sm = mfma(t1h[c], K.u1h[c], sm);
    sc = mfma(t1h[c], K.u1l[c], sc);
    sm2 = mfma(t2h[c], K.u2h[c], sm2);
    sc = mfma(t1l[c], K.u1h[c], sc);
  }
  float s[16];
  float lsum = 0.f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    s[r] = fmaf(sc[r], kLoInv, sm[r] + sm2[r]) * kLam1;
    lsum += s[r];
  }
  const float mean = wave_sum(lsum) * (1.0f / (kN * kN));
  // goldstein.cu:170-175 on the unscaled data: mean > 1e-8
  const bool ok = mean > 1e-8f * scale;
  const float inv = ok ? 1.0f / mean : 0.f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    float sn = s[r] * inv;
    if (sn > 0.f) {  // powf(sn, alpha) > 0 only here
      float w = __builtin_amdgcn_exp2f(alpha * __builtin_amdgcn_logf(sn));
      zr[r] *= w;
      zi[r] *= w;
    }
  }
  // inverse transform
  cgemm_v<true>(zr, zi, K.F, yr, yi);
  cgemm_v<true>(yr, yi, K.F, zr, zi);
  // weighted output (goldstein.cu:257-267), undo the patch scale
  const float wy = (float)(n < kN / 2 ? n : kN - 1 - n);
  const float norm = 1.0f / (kN * kN * scale);
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int col = 4 * h + 8 * q + t;
      const float wx = (float)(col < kN / 2 ? col : kN - 1 - col);
      const float w = (wx + wy) * norm;
      out[4 * q + t] = make_float2(zr[4 * q + t] * w, zi[4 * q + t] * w);
    }
}

}  // namespace mf
}  // namespace s1ifg
