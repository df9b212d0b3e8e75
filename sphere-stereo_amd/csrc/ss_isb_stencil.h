// Stencil tables of the ISB filter's up pass (isb_filter.cu:173-208), shared
// by the forward passes (isb.hip) and their adjoints (grad.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ss {

// (row, col) of the 25 up-sampling taps in the 3x3 coarse neighbourhood,
// 4 bits each (row*4 + col), in the order of isb_filter.cu:173-208:
// stencil 1 = taps 0-8, stencil 2 = 9-14, stencil 3 = 15-20, stencil 4 = 21-24.
constexpr uint64_t kUpTapsLo = 0x5A68495A98654210ull;
constexpr uint64_t kUpTapsHi = 0xA965A9216ull;

__host__ __device__ constexpr int up_tap_code(int k) {
  return k < 16 ? (int)((kUpTapsLo >> (4 * k)) & 15) : (int)((kUpTapsHi >> (4 * (k - 16))) & 15);
}
__host__ __device__ constexpr int up_tap_flat(int k) { return (up_tap_code(k) >> 2) * 3 + (up_tap_code(k) & 3); }
__host__ __device__ constexpr int up_stencil(int k) { return k < 9 ? 0 : (k < 15 ? 1 : (k < 21 ? 2 : 3)); }
// spatial base weights {1,1,1,1,8,1,1,1,1}, {8,8,1,1,1,1} x2, {1,1,1,1}
__host__ __device__ constexpr float up_base(int k) {
  return (k == 4 || k == 9 || k == 10 || k == 15 || k == 16) ? 8.0f : 1.0f;
}
constexpr int kStencilBegin[5] = {0, 9, 15, 21, 25};

// Value selects (by-value parameters keep the choice in registers; selecting
// between lvalues lets the compiler pick between stack addresses instead).
__device__ __forceinline__ int pick3(int i, int a, int b, int c) { return i == 0 ? a : (i == 1 ? b : c); }
__device__ __forceinline__ uint32_t pick4(int i, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
  return i == 0 ? a : (i == 1 ? b : (i == 2 ? c : d));
}

}  // namespace ss
