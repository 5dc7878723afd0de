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
// Split.  Every MFMA A operand v is split v = hi + lo with hi = fp16(v),
// lo = fp16(v - hi); the B constants are split the same way.  A product
// keeps hi*hi + hi*lo + lo*hi (lo*lo, ~2^-22 relative, dropped); all three
// carry the same scale and go into one FP32 accumulator.  Operand ranges:
// the patch is scaled by a power of two to max component in [4, 8); a
// 32-point transform grows the maximum by at most 32 sqrt(2), so
// |Y| < 2^8.5, |Z| < 2^14 and the smoothing passes stay below 1.9 * 2^14.5
// (all within FP16); the weighted spectrum is rescaled like the patch
// before the inverse transform, and both scales are undone exactly at the
// end.  Values below 2^-11 of their operand's maximum have subnormal lo
// parts: their absolute error stays below 2^-27 of that maximum.
#pragma once

#include "gold_dev.h"
#include "s1ifg_internal.h"

namespace s1ifg {
namespace mf {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kN = 32;
constexpr int kNormExp = 3;  // max component of a normalised operand in [4, 8)
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

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

typedef _Float16 half2v __attribute__((ext_vector_type(2)));

// hi + lo split of v (|v| < 2^16).  hi rounds toward zero, two values per
// packing conversion; lo = v - hi is exact in FP32 (|lo| < ulp(hi)) and is
// rounded to fp16 by one mixed-precision FMA v * 1 - hi per value that
// reads hi from its packed half (the 1 is hidden from constant folding so
// that the FMA is not rewritten as a subtraction with separate
// conversions).
__device__ __forceinline__ void split16(const float v[16], half8 hi[2], half8 lo[2]) {
  float one = 1.0f;
  asm("" : "+v"(one));
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
      const half2v h =
          __builtin_bit_cast(half2v, __builtin_amdgcn_cvt_pkrtz(v[8 * c + j], v[8 * c + j + 1]));
      hi[c][j] = h[0];
      hi[c][j + 1] = h[1];
      lo[c][j] = (_Float16)fmaf(v[8 * c + j], one, -(float)h[0]);
      lo[c][j + 1] = (_Float16)fmaf(v[8 * c + j + 1], one, -(float)h[1]);
    }
}

__device__ __forceinline__ void to_half16(const float v[16], float s, half8 hi[2]) {
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int j = 0; j < 8; ++j) hi[c][j] = (_Float16)(v[8 * c + j] * s);
}

struct CplxOp {  // constant complex B operand, split, per chunk; -Im too
  half8 rh[2], rl[2], ih[2], il[2], nih[2], nil[2];
};

struct Consts {
  CplxOp F;
};

// The smoothing operands, in LDS between uses (they are needed once per
// patch): per wavefront [6][64] half8, entry [k][lane] with k = u1h[0],
// u1h[1], u1l[0], u1l[1], u2h[0], u2h[1].
typedef half8 SmoothLds[6][64];

__device__ __forceinline__ void split1(float x, _Float16 &h, _Float16 &l) {
  h = (_Float16)x;
  l = (_Float16)(x - (float)h);
}

// Constant B operands of lane (h, n) in kappa order; the smoothing
// operands go to U (this wavefront's; visible to it after a wavefront
// barrier).
__device__ __forceinline__ void build_consts(int h, int n, Consts &K, SmoothLds &U) {
  half8 u1h[2], u1l[2], u2h[2];
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
      K.F.nih[c][j] = -a;
      K.F.nil[c][j] = -b;
      const int d = ((k - n + 16) & 31) - 16;  // circular offset
      const float b1 = (d >= -3 && d <= 3) ? kU1[d + 3] : 0.f;
      const float b2 = (d >= -3 && d <= 3) ? kU2[d + 3] : 0.f;
      split1(b1, a, b);
      u1h[c][j] = a;
      u1l[c][j] = b;
      u2h[c][j] = (_Float16)b2;
    }
  const int lane = 32 * h + n;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    U[c][lane] = u1h[c];
    U[2 + c][lane] = u1l[c];
    U[4 + c][lane] = u2h[c];
  }
}

// C = A * B (B = F) or A * conj(B) (CONJ), A given by split real/imag
// parts, B constant.  The minus signs of the complex product come with the
// negated copy of B's imaginary part (fp16 negation is exact).  One
// accumulator for the real and one for the imaginary part, alternating, the
// hi*lo terms of a K chunk before its hi*hi terms.
template <bool CONJ>
__device__ __forceinline__ void cgemm(const half8 arh[2], const half8 arl[2], const half8 aih[2],
                                      const half8 ail[2], const CplxOp &b, float cr[16],
                                      float ci[16]) {
  f32x16 re = {}, im = {};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (!CONJ) {
      // Cr = Ar Fr - Ai Fi, Ci = Ar Fi + Ai Fr
      re = mfma(arl[c], b.rh[c], re);
      im = mfma(arl[c], b.ih[c], im);
      re = mfma(arh[c], b.rl[c], re);
      im = mfma(arh[c], b.il[c], im);
      re = mfma(ail[c], b.nih[c], re);
      im = mfma(ail[c], b.rh[c], im);
      re = mfma(aih[c], b.nil[c], re);
      im = mfma(aih[c], b.rl[c], im);
      re = mfma(arh[c], b.rh[c], re);
      im = mfma(arh[c], b.ih[c], im);
      re = mfma(aih[c], b.nih[c], re);
      im = mfma(aih[c], b.rh[c], im);
    } else {
      // Cr = Ar Fr + Ai Fi, Ci = Ai Fr - Ar Fi
      re = mfma(arl[c], b.rh[c], re);
      im = mfma(ail[c], b.rh[c], im);
      re = mfma(arh[c], b.rl[c], re);
      im = mfma(aih[c], b.rl[c], im);
      re = mfma(ail[c], b.ih[c], re);
      im = mfma(arl[c], b.nih[c], im);
      re = mfma(aih[c], b.il[c], re);
      im = mfma(arh[c], b.nil[c], im);
      re = mfma(arh[c], b.rh[c], re);
      im = mfma(aih[c], b.rh[c], im);
      re = mfma(aih[c], b.ih[c], re);
      im = mfma(arh[c], b.nih[c], im);
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    cr[r] = re[r];
    ci[r] = im[r];
  }
}

// Y = X B or X conj(B) for X = xr + i xi.
template <bool CONJ>
__device__ __forceinline__ void cgemm_v(const float xr[16], const float xi[16], const CplxOp &b,
                                        float yr[16], float yi[16]) {
  half8 arh[2], arl[2], aih[2], ail[2];
  split16(xr, arh, arl);
  split16(xi, aih, ail);
  cgemm<CONJ>(arh, arl, aih, ail, b, yr, yi);
}

// Wavefront reductions with the result in every lane, on the VALU:
// DPP within rows of 16 lanes (quad permutations, then the half-row and row
// mirrors, which pair quads and half rows once their lanes agree), then
// the gfx950 lane swaps between rows 0/1 and 2/3 (permlane16_swap) and
// between the two halves (permlane32_swap).  Called by full wavefronts.
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float x) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xF, 0xF,
                                                        false));
}

template <typename Op>
__device__ __forceinline__ float wave_reduce(float v, Op op) {
  v = op(v, dpp_mov<0xB1>(v));   // quad_perm [1, 0, 3, 2]
  v = op(v, dpp_mov<0x4E>(v));   // quad_perm [2, 3, 0, 1]
  v = op(v, dpp_mov<0x141>(v));  // row_half_mirror
  v = op(v, dpp_mov<0x140>(v));  // row_mirror
  const unsigned u = __builtin_bit_cast(unsigned, v);
  const auto r16 = __builtin_amdgcn_permlane16_swap(u, u, false, false);
  v = op(__builtin_bit_cast(float, (unsigned)r16[0]), __builtin_bit_cast(float, (unsigned)r16[1]));
  const unsigned w = __builtin_bit_cast(unsigned, v);
  const auto r32 = __builtin_amdgcn_permlane32_swap(w, w, false, false);
  return op(__builtin_bit_cast(float, (unsigned)r32[0]), __builtin_bit_cast(float, (unsigned)r32[1]));
}

__device__ __forceinline__ float wave_sum(float v) {
  return wave_reduce(v, [](float a, float b) { return a + b; });
}

__device__ __forceinline__ float wave_max(float v) {
  return wave_reduce(v, [](float a, float b) { return fmaxf(a, b); });
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

// Whether the patch held in x (lane layout) is all zeros, +0 or -0: the OR
// of the component bit patterns without the sign bit is 0 in every lane (a
// NaN or an infinity has exponent bits set).  Wavefront-uniform.
__device__ __forceinline__ bool patch_is_zero(const float2 x[16]) {
  unsigned b = 0;
#pragma unroll
  for (int r = 0; r < 16; ++r)
    b |= __builtin_bit_cast(unsigned, x[r].x) | __builtin_bit_cast(unsigned, x[r].y);
  return !__any((b & 0x7fffffffu) != 0);
}

// Filter one patch held in x (lane layout) into z' in the same layout:
// lane (h, n) holds row n, columns 4h + 8q + t (q, t = 0..3) in out[4q + t];
// z' norm is the filtered patch over n^2 (patch_weight adds the triangular
// weight).
// SQRT: alpha = 0.5 (the usual Goldstein exponent), the weight is one
// square root.
template <bool SQRT>
__device__ __forceinline__ void filter_patch_regs(const Consts &K, const SmoothLds &U,
                                                  const float2 x[16], int h, int n, float alpha,
                                                  float2 out[16], float &norm) {
  const int lane = 32 * h + n;
  float xr[16], xi[16];
  float mx = 0.f;
#pragma unroll
  for (int r = 0; r < 16; ++r) mx = fmaxf(mx, fmaxf(fabsf(x[r].x), fabsf(x[r].y)));
  mx = wave_max(mx);
  int e = 0;
  if (mx > 0.f) frexpf(mx, &e);  // max |x| in [2^(e-1), 2^e)
  e = max(e, -123);               // keeps s0 = 2^(kNormExp - e) finite
  const float s0 = ldexpf(1.0f, kNormExp - e);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    xr[r] = x[r].x * s0;
    xi[r] = x[r].y * s0;
  }
  // forward transform: Z s0
  float yr[16], yi[16], zr[16], zi[16];
  cgemm_v<false>(xr, xi, K.F, yr, yi);
  cgemm_v<false>(yr, yi, K.F, zr, zi);
  // smoothing of |Z|
  float m[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) m[r] = __builtin_amdgcn_sqrtf(zr[r] * zr[r] + zi[r] * zi[r]);
  half8 mh[2], ml[2];
  split16(m, mh, ml);
  half8 u1h[2], u1l[2], u2h[2];
  asm volatile("" ::: "memory");  // read here, not held in registers across patches
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    u1h[c] = U[c][lane];
    u1l[c] = U[2 + c][lane];
    u2h[c] = U[4 + c][lane];
  }
  f32x16 t1 = {}, t2 = {};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    t1 = mfma(mh[c], u1l[c], t1);
    t2 = mfma(mh[c], u2h[c], t2);
    t1 = mfma(ml[c], u1h[c], t1);
    t1 = mfma(mh[c], u1h[c], t1);
  }
  // S = l1 (B1^T M B1 + (l2/l1) B2^T M B2): the second pass runs on B_r,
  // l2/l1 scales the rank-2 term's data and l1 the sum
  float t1v[16], t2v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    t1v[r] = t1[r];
    t2v[r] = t2[r];
  }
  half8 t1h[2], t1l[2], t2h[2];
  split16(t1v, t1h, t1l);
  to_half16(t2v, kLam2 / kLam1, t2h);
  f32x16 sm = {};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    sm = mfma(t2h[c], u2h[c], sm);
    sm = mfma(t1h[c], u1l[c], sm);
    sm = mfma(t1l[c], u1h[c], sm);
    sm = mfma(t1h[c], u1h[c], sm);
  }
  // S = l1 sm; the mean of S and S / mean with l1 applied once
  float lsum = 0.f;
#pragma unroll
  for (int r = 0; r < 16; ++r) lsum += sm[r];
  const float mean = wave_sum(lsum) * (kLam1 / (kN * kN));
  // goldstein.cu:170-175 on the unscaled data: mean > 1e-8 (S carries s0)
  const bool ok = mean > 1e-8f * s0;
  const float inv = ok ? kLam1 / mean : 0.f;
  float mw = 0.f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    float sn = sm[r] * inv;
    if (sn > 0.f) {  // powf(sn, alpha) > 0 only here; Z is kept elsewhere
      float w = SQRT ? __builtin_amdgcn_sqrtf(sn)
                     : __builtin_amdgcn_exp2f(alpha * __builtin_amdgcn_logf(sn));
      zr[r] *= w;
      zi[r] *= w;
    }
    mw = fmaxf(mw, fmaxf(fabsf(zr[r]), fabsf(zi[r])));
  }
  mw = wave_max(mw);
  int e2 = 0;
  if (mw > 0.f) frexpf(mw, &e2);
  e2 = max(e2, -123);
  const float s2 = ldexpf(1.0f, kNormExp - e2);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    zr[r] *= s2;
    zi[r] *= s2;
  }
  // inverse transform of the rescaled spectrum: z' n^2 s0 s2
  cgemm_v<true>(zr, zi, K.F, yr, yi);
  cgemm_v<true>(yr, yi, K.F, zr, zi);
  // z' / n^2 = result / (s0 s2 n^2) = result 2^(e + e2 - 2 kNormExp - 10)
  norm = ldexpf(1.0f, e + e2 - 2 * kNormExp - 10);
#pragma unroll
  for (int r = 0; r < 16; ++r) out[r] = make_float2(zr[r], zi[r]);
}

// Triangular output weight (goldstein.cu:257-267) of lane (h, n)'s entry
// 4q + t times norm.
__device__ __forceinline__ float patch_weight(int h, int n, int q, int t, float norm) {
  const float wy = (float)(n < kN / 2 ? n : kN - 1 - n);
  const int col = 4 * h + 8 * q + t;
  const float wx = (float)(col < kN / 2 ? col : kN - 1 - col);
  return fmaf(wx, norm, wy * norm);
}

}  // namespace mf
}  // namespace s1ifg
