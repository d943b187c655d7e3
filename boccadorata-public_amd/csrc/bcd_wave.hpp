// bcd_wave.hpp -- sums and maxima across one 64-lane wave (gfx950 device code).
#pragma once

#include <hip/hip_runtime.h>

namespace bcd {

// x from another lane of the same 16-lane row, by a DPP control (quad_perm,
// row_mirror, row_half_mirror): a VALU operand modifier, with no LDS round
// trip as a shuffle takes.  Every source lane lies inside the row.
template <int CTRL>
__device__ __forceinline__ double dpp_row(double x) {
  const long long v = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(v & 0xffffffffll), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(v >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// x of lane l (the same value in every lane).
__device__ __forceinline__ double lane_get(double x, int l) {
  const long long v = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the 64 lanes (all active); every lane gets the same total.  Each
// 16-lane row sums pairwise (lanes i ^ 1, i ^ 2, then the mirrored half and
// the mirrored row; a pair's two lanes add the same two values, so the row
// agrees), then the four row sums are added in a fixed order.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_row<0xB1>(x);    // quad_perm [1,0,3,2]
  x += dpp_row<0x4E>(x);    // quad_perm [2,3,0,1]
  x += dpp_row<0x141>(x);   // row_half_mirror
  x += dpp_row<0x140>(x);   // row_mirror
  return (lane_get(x, 0) + lane_get(x, 16)) + (lane_get(x, 32) + lane_get(x, 48));
}

// Maximum over the 64 lanes (all active), the same way.
__device__ __forceinline__ double wave_max(double x) {
  x = fmax(x, dpp_row<0xB1>(x));
  x = fmax(x, dpp_row<0x4E>(x));
  x = fmax(x, dpp_row<0x141>(x));
  x = fmax(x, dpp_row<0x140>(x));
  return fmax(fmax(lane_get(x, 0), lane_get(x, 16)), fmax(lane_get(x, 32), lane_get(x, 48)));
}

}  // namespace bcd
