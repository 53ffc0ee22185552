// wave.h -- wavefront-wide scans for gfx950 wave64 kernels (internal).
#pragma once
#include <hip/hip_runtime.h>

namespace zw {

// Wave-wide inclusive scans on DPP (data-parallel primitives: a VALU operand
// read from another lane of the same row, no LDS round trip): within each
// 16-lane row by row shifts of 1, 2, 4, 8, then across rows by broadcasting
// lane 15 into row 1 and 3 and lane 31 into rows 2 and 3.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp0(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, ROWS, 0xF, true);
}
__device__ __forceinline__ int wave_scan_add(int x) {
  x += dpp0<0x111, 0xF>(x);  // row_shr:1
  x += dpp0<0x112, 0xF>(x);  // row_shr:2
  x += dpp0<0x114, 0xF>(x);  // row_shr:4
  x += dpp0<0x118, 0xF>(x);  // row_shr:8
  x += dpp0<0x142, 0xA>(x);  // row_bcast:15 into rows 1, 3
  x += dpp0<0x143, 0xC>(x);  // row_bcast:31 into rows 2, 3
  return x;
}
__device__ __forceinline__ unsigned umax_(unsigned a, unsigned b) { return a > b ? a : b; }
__device__ __forceinline__ unsigned wave_scan_max(unsigned x) {
  x = umax_(x, (unsigned)dpp0<0x111, 0xF>((int)x));
  x = umax_(x, (unsigned)dpp0<0x112, 0xF>((int)x));
  x = umax_(x, (unsigned)dpp0<0x114, 0xF>((int)x));
  x = umax_(x, (unsigned)dpp0<0x118, 0xF>((int)x));
  x = umax_(x, (unsigned)dpp0<0x142, 0xA>((int)x));
  x = umax_(x, (unsigned)dpp0<0x143, 0xC>((int)x));
  return x;
}
}  // namespace zw
