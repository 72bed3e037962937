// XCD-aware block order: the hardware deals consecutive workgroups to the 8
// XCDs of the chip in turn; xcd_tile gives each XCD a contiguous range of
// tiles so that tiles sharing cache lines meet in one L2.
#pragma once

#include <hip/hip_runtime.h>

namespace fc {

__device__ __forceinline__ int xcd_tile(int bid, int nb) {
  const int q8 = nb >> 3, r8 = nb & 7;
  const int xcd = bid & 7, slot = bid >> 3;
  return (xcd < r8 ? xcd * (q8 + 1) : r8 * (q8 + 1) + (xcd - r8) * q8) + slot;
}

}  // namespace fc
