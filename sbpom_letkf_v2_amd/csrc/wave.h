// wave.h -- wavefront-wide FP64 reductions on DPP moves (no LDS traffic),
// shared by the eigensolve, adjoint and transform kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace letkf {

// Wavefront sum of a double with DPP moves (no LDS traffic): butterfly
// within quads, half-row and row mirrors, then the row broadcasts 15 and 31
// fold the four rows into lane 63, which readlane hands to every lane.
// With every row enabled (the permutations within a row: every lane reads
// a valid lane) the move's "old" operand is never used and is left
// undefined, so no zero is materialised first; the row broadcasts (rows
// masked off) keep 0 in the lanes they do not write.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double v) {
  const long long x = __double_as_longlong(v);
  int lo, hi;
  if constexpr (ROWMASK == 0xF) {
    lo = __builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp((int)(x >> 32), CTRL, 0xF, 0xF, false);
  } else {
    lo = __builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWMASK, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), CTRL, ROWMASK, 0xF, false);
  }
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_d<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  v += dpp_d<0x141, 0xF>(v);  // row_half_mirror
  v += dpp_d<0x140, 0xF>(v);  // row_mirror
  v += dpp_d<0x142, 0xA>(v);  // row_bcast15 -> rows 1, 3
  v += dpp_d<0x143, 0xC>(v);  // row_bcast31 -> rows 2, 3
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)x, 63);
  const int hi = __builtin_amdgcn_readlane((int)(x >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Two wavefront sums interleaved (one reduction latency for both).
__device__ __forceinline__ void wave_sum2(double& a, double& b) {
  a += dpp_d<0xB1, 0xF>(a);
  b += dpp_d<0xB1, 0xF>(b);
  a += dpp_d<0x4E, 0xF>(a);
  b += dpp_d<0x4E, 0xF>(b);
  a += dpp_d<0x141, 0xF>(a);
  b += dpp_d<0x141, 0xF>(b);
  a += dpp_d<0x140, 0xF>(a);
  b += dpp_d<0x140, 0xF>(b);
  a += dpp_d<0x142, 0xA>(a);
  b += dpp_d<0x142, 0xA>(b);
  a += dpp_d<0x143, 0xC>(a);
  b += dpp_d<0x143, 0xC>(b);
  const long long x = __double_as_longlong(a), y = __double_as_longlong(b);
  const int xl = __builtin_amdgcn_readlane((int)x, 63), xh = __builtin_amdgcn_readlane((int)(x >> 32), 63);
  const int yl = __builtin_amdgcn_readlane((int)y, 63), yh = __builtin_amdgcn_readlane((int)(y >> 32), 63);
  a = __longlong_as_double(((long long)xh << 32) | (unsigned)xl);
  b = __longlong_as_double(((long long)yh << 32) | (unsigned)yl);
}

// The same sums for floats (one DPP move per step).
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_f(float v) {
  if constexpr (ROWMASK == 0xF)
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
  else
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xF, false));
}

__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_f<0xB1, 0xF>(v);
  v += dpp_f<0x4E, 0xF>(v);
  v += dpp_f<0x141, 0xF>(v);
  v += dpp_f<0x140, 0xF>(v);
  v += dpp_f<0x142, 0xA>(v);
  v += dpp_f<0x143, 0xC>(v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ void wave_sum2(float& a, float& b) {
  a += dpp_f<0xB1, 0xF>(a);
  b += dpp_f<0xB1, 0xF>(b);
  a += dpp_f<0x4E, 0xF>(a);
  b += dpp_f<0x4E, 0xF>(b);
  a += dpp_f<0x141, 0xF>(a);
  b += dpp_f<0x141, 0xF>(b);
  a += dpp_f<0x140, 0xF>(a);
  b += dpp_f<0x140, 0xF>(b);
  a += dpp_f<0x142, 0xA>(a);
  b += dpp_f<0x142, 0xA>(b);
  a += dpp_f<0x143, 0xC>(a);
  b += dpp_f<0x143, 0xC>(b);
  a = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a), 63));
  b = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b), 63));
}

}  // namespace letkf
