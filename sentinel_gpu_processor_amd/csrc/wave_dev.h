// Wavefront reductions on the VALU, result in every lane: DPP within rows
// of 16 lanes (quad permutations [1,0,3,2] and [2,3,0,1], then the
// half-row and row mirrors, which pair quads and half rows once their
// lanes agree), then the gfx950 lane swaps between rows 0/1 and 2/3
// (permlane16_swap) and between the two halves (permlane32_swap).  Fixed
// order; called with all 64 lanes active.
#pragma once

#include <hip/hip_runtime.h>

namespace s1ifg {

template <int CTRL>
__device__ __forceinline__ unsigned dpp32(unsigned x) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}

// T: a 32-bit or 64-bit trivially copyable type; op: associative combine.
template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
  constexpr int kWords = sizeof(T) / 4;
  struct Words {
    unsigned w[kWords];
  };
  const auto dpp = [](T x, auto ctrl) {
    Words u = __builtin_bit_cast(Words, x);
#pragma unroll
    for (int k = 0; k < kWords; ++k) u.w[k] = dpp32<decltype(ctrl)::value>(u.w[k]);
    return __builtin_bit_cast(T, u);
  };
  v = op(v, dpp(v, std::integral_constant<int, 0xB1>()));
  v = op(v, dpp(v, std::integral_constant<int, 0x4E>()));
  v = op(v, dpp(v, std::integral_constant<int, 0x141>()));
  v = op(v, dpp(v, std::integral_constant<int, 0x140>()));
  {
    const Words u = __builtin_bit_cast(Words, v);
    Words a, b;
#pragma unroll
    for (int k = 0; k < kWords; ++k) {
      const auto r = __builtin_amdgcn_permlane16_swap(u.w[k], u.w[k], false, false);
      a.w[k] = (unsigned)r[0];
      b.w[k] = (unsigned)r[1];
    }
    v = op(__builtin_bit_cast(T, a), __builtin_bit_cast(T, b));
  }
  {
    const Words u = __builtin_bit_cast(Words, v);
    Words a, b;
#pragma unroll
    for (int k = 0; k < kWords; ++k) {
      const auto r = __builtin_amdgcn_permlane32_swap(u.w[k], u.w[k], false, false);
      a.w[k] = (unsigned)r[0];
      b.w[k] = (unsigned)r[1];
    }
    v = op(__builtin_bit_cast(T, a), __builtin_bit_cast(T, b));
  }
  return v;
}

// Lane l's value of a double (wavefront-uniform result; l a constant lane
// index).
__device__ __forceinline__ double wave_read(double v, int l) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

}  // namespace s1ifg
