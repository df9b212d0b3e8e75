// Issue-rate probes for gfx950: how many SIMD cycles one wave64 instruction of
// a class occupies, measured on long streams of independent instructions (8
// chains per lane, 8 waves per SIMD), and the throughput of LDS atomics.
// These are the rates that the kernels' roofline notes price instruction
// mixes with.  ALU-only inline asm (no memory operands in any asm block);
// every kernel writes one value per thread into a buffer sized for its grid.
// Built by _build.py into sphere_stereo/issue_rates; run by
// `python bench.py --issue-rates`.  With `--check` it only runs the
// exhaustive equality checks of the exact fast forms that the kernels use
// (exit status 1 on any mismatch).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#define CHECK(x)                                                             \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                              \
    }                                                                        \
  } while (0)

namespace {

constexpr int kBlock = 256;
constexpr int kIters = 2048;
constexpr int kUnroll = 4;
constexpr double kClockHz = 2.4e9;  // MI355X peak engine clock

__global__ void __launch_bounds__(kBlock) k_valu0(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_add_f32 %0, %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32 %0, %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32 %0, %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32 %0, %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32 %0, %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32 %0, %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32 %0, %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32 %0, %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu1(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_sub_f32 %0, %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_f32 %0, %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_f32 %0, %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_f32 %0, %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_f32 %0, %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_f32 %0, %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_f32 %0, %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_f32 %0, %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu2(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_mul_f32 %0, %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_f32 %0, %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_f32 %0, %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_f32 %0, %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_f32 %0, %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_f32 %0, %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_f32 %0, %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_f32 %0, %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu3(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fmac_f32 %0, %1, %2" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu4(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_fma_f32 %0, -%0, %1, %2" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fma_f32 %0, -%0, %1, %2" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fma_f32 %0, -%0, %1, %2" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fma_f32 %0, -%0, %1, %2" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fma_f32 %0, -%0, %1, %2" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fma_f32 %0, -%0, %1, %2" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fma_f32 %0, -%0, %1, %2" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_fma_f32 %0, -%0, %1, %2" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu5(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_add_f32_e64 %0, |%0|, |%1|" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_e64 %0, |%0|, |%1|" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_e64 %0, |%0|, |%1|" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_e64 %0, |%0|, |%1|" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_e64 %0, |%0|, |%1|" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_e64 %0, |%0|, |%1|" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_e64 %0, |%0|, |%1|" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_e64 %0, |%0|, |%1|" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu6(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_add_u32 %0, %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_u32 %0, %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_u32 %0, %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_u32 %0, %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_u32 %0, %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_u32 %0, %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_u32 %0, %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_u32 %0, %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu7(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_sub_u32 %0, %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_u32 %0, %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_u32 %0, %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_u32 %0, %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_u32 %0, %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_u32 %0, %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_u32 %0, %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sub_u32 %0, %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu8(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_and_b32 %0, %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_and_b32 %0, %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_and_b32 %0, %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_and_b32 %0, %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_and_b32 %0, %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_and_b32 %0, %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_and_b32 %0, %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_and_b32 %0, %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu9(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_or_b32 %0, %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_or_b32 %0, %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_or_b32 %0, %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_or_b32 %0, %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_or_b32 %0, %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_or_b32 %0, %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_or_b32 %0, %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_or_b32 %0, %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu10(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_mov_b32 %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32 %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32 %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32 %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32 %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32 %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32 %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32 %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu11(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_min_f32 %0, %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_min_f32 %0, %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_min_f32 %0, %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_min_f32 %0, %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_min_f32 %0, %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_min_f32 %0, %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_min_f32 %0, %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_min_f32 %0, %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu12(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_max_f32 %0, %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_max_f32 %0, %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_max_f32 %0, %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_max_f32 %0, %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_max_f32 %0, %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_max_f32 %0, %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_max_f32 %0, %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_max_f32 %0, %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu13(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_med3_f32 %0, %0, %1, %2" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_med3_f32 %0, %0, %1, %2" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_med3_f32 %0, %0, %1, %2" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_med3_f32 %0, %0, %1, %2" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_med3_f32 %0, %0, %1, %2" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_med3_f32 %0, %0, %1, %2" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_med3_f32 %0, %0, %1, %2" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_med3_f32 %0, %0, %1, %2" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu14(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_floor_f32 %0, %0" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_floor_f32 %0, %0" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_floor_f32 %0, %0" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_floor_f32 %0, %0" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_floor_f32 %0, %0" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_floor_f32 %0, %0" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_floor_f32 %0, %0" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_floor_f32 %0, %0" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu15(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_cvt_i32_f32 %0, %0" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_i32_f32 %0, %0" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_i32_f32 %0, %0" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_i32_f32 %0, %0" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_i32_f32 %0, %0" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_i32_f32 %0, %0" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_i32_f32 %0, %0" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_i32_f32 %0, %0" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu16(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_cvt_f32_i32 %0, %0" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f32_i32 %0, %0" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f32_i32 %0, %0" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f32_i32 %0, %0" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f32_i32 %0, %0" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f32_i32 %0, %0" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f32_i32 %0, %0" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f32_i32 %0, %0" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu17(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_cvt_f16_f32 %0, %0" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f16_f32 %0, %0" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f16_f32 %0, %0" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f16_f32 %0, %0" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f16_f32 %0, %0" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f16_f32 %0, %0" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f16_f32 %0, %0" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cvt_f16_f32 %0, %0" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu18(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_ldexp_f32 %0, %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_ldexp_f32 %0, %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_ldexp_f32 %0, %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_ldexp_f32 %0, %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_ldexp_f32 %0, %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_ldexp_f32 %0, %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_ldexp_f32 %0, %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_ldexp_f32 %0, %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu19(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshlrev_b32 %0, 1, %0" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu20(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_bfe_u32 %0, %0, 8, 8" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_bfe_u32 %0, %0, 8, 8" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_bfe_u32 %0, %0, 8, 8" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_bfe_u32 %0, %0, 8, 8" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_bfe_u32 %0, %0, 8, 8" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_bfe_u32 %0, %0, 8, 8" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_bfe_u32 %0, %0, 8, 8" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_bfe_u32 %0, %0, 8, 8" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu21(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu22(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu23(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu24(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_sad_u8 %0, %0, %1, %2" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sad_u8 %0, %0, %1, %2" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sad_u8 %0, %0, %1, %2" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sad_u8 %0, %0, %1, %2" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sad_u8 %0, %0, %1, %2" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sad_u8 %0, %0, %1, %2" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sad_u8 %0, %0, %1, %2" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sad_u8 %0, %0, %1, %2" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu25(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu26(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_cndmask_b32_e64 %0, %0, %1, %3" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cndmask_b32_e64 %0, %0, %1, %3" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cndmask_b32_e64 %0, %0, %1, %3" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cndmask_b32_e64 %0, %0, %1, %3" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cndmask_b32_e64 %0, %0, %1, %3" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cndmask_b32_e64 %0, %0, %1, %3" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cndmask_b32_e64 %0, %0, %1, %3" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_cndmask_b32_e64 %0, %0, %1, %3" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu27(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_mov_b32_dpp %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32_dpp %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32_dpp %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32_dpp %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32_dpp %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32_dpp %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32_dpp %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_mov_b32_dpp %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu28(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_add_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_add_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu29(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_sqrt_f32 %0, %0" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sqrt_f32 %0, %0" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sqrt_f32 %0, %0" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sqrt_f32 %0, %0" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sqrt_f32 %0, %0" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sqrt_f32 %0, %0" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sqrt_f32 %0, %0" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_sqrt_f32 %0, %0" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu30(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_rcp_f32 %0, %0" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_rcp_f32 %0, %0" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_rcp_f32 %0, %0" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_rcp_f32 %0, %0" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_rcp_f32 %0, %0" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_rcp_f32 %0, %0" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_rcp_f32 %0, %0" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_rcp_f32 %0, %0" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}
__global__ void __launch_bounds__(kBlock) k_valu31(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float c = 0.999f, d = 1e-7f;
  const uint64_t mask = 0x5555555555555555ull;
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      asm volatile("v_exp_f32 %0, %0" : "+v"(r0) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_exp_f32 %0, %0" : "+v"(r1) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_exp_f32 %0, %0" : "+v"(r2) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_exp_f32 %0, %0" : "+v"(r3) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_exp_f32 %0, %0" : "+v"(r4) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_exp_f32 %0, %0" : "+v"(r5) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_exp_f32 %0, %0" : "+v"(r6) : "v"(c), "v"(d), "s"(mask)); asm volatile("v_exp_f32 %0, %0" : "+v"(r7) : "v"(c), "v"(d), "s"(mask));
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}

// Composite sequences: s_nop beside fp32 adds, compare -> s_nop -> select
// triples, and two forms of the correctly rounded square root.
template <int MODE>
__global__ void __launch_bounds__(kBlock) k_seq(float* out, float seed) {
  float r0 = seed + threadIdx.x, r1 = r0 + 1, r2 = r0 + 2, r3 = r0 + 3, r4 = r0 + 4, r5 = r0 + 5, r6 = r0 + 6, r7 = r0 + 7;
  const float d = 1e-7f;
  for (int it = 0; it < kIters; ++it) {
    if constexpr (MODE == 0) {
      asm volatile("v_add_f32 %0, %0, %8\n s_nop 0\n v_add_f32 %1, %1, %8\n s_nop 0\n v_add_f32 %2, %2, %8\n s_nop 0\n"
                   " v_add_f32 %3, %3, %8\n s_nop 0\n v_add_f32 %4, %4, %8\n s_nop 0\n v_add_f32 %5, %5, %8\n s_nop 0\n"
                   " v_add_f32 %6, %6, %8\n s_nop 0\n v_add_f32 %7, %7, %8\n s_nop 0"
                   : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7) : "v"(d));
    } else if constexpr (MODE == 1) {
      asm volatile("v_add_f32 %0, %0, %8\n s_nop 1\n v_add_f32 %1, %1, %8\n s_nop 1\n v_add_f32 %2, %2, %8\n s_nop 1\n"
                   " v_add_f32 %3, %3, %8\n s_nop 1\n v_add_f32 %4, %4, %8\n s_nop 1\n v_add_f32 %5, %5, %8\n s_nop 1\n"
                   " v_add_f32 %6, %6, %8\n s_nop 1\n v_add_f32 %7, %7, %8\n s_nop 1"
                   : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7) : "v"(d));
    } else {
      asm volatile("v_cmp_lt_f32 vcc, %0, %8\n s_nop 1\n v_cndmask_b32 %0, %0, %8, vcc\n"
                   " v_cmp_lt_f32 vcc, %1, %8\n s_nop 1\n v_cndmask_b32 %1, %1, %8, vcc\n"
                   " v_cmp_lt_f32 vcc, %2, %8\n s_nop 1\n v_cndmask_b32 %2, %2, %8, vcc\n"
                   " v_cmp_lt_f32 vcc, %3, %8\n s_nop 1\n v_cndmask_b32 %3, %3, %8, vcc\n"
                   " v_cmp_lt_f32 vcc, %4, %8\n s_nop 1\n v_cndmask_b32 %4, %4, %8, vcc\n"
                   " v_cmp_lt_f32 vcc, %5, %8\n s_nop 1\n v_cndmask_b32 %5, %5, %8, vcc\n"
                   " v_cmp_lt_f32 vcc, %6, %8\n s_nop 1\n v_cndmask_b32 %6, %6, %8, vcc\n"
                   " v_cmp_lt_f32 vcc, %7, %8\n s_nop 1\n v_cndmask_b32 %7, %7, %8, vcc"
                   : "+v"(r0), "+v"(r1), "+v"(r2), "+v"(r3), "+v"(r4), "+v"(r5), "+v"(r6), "+v"(r7) : "v"(d) : "vcc");
    }
  }
  out[blockIdx.x * kBlock + threadIdx.x] = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}

// The correctly rounded square root for normal inputs as the compiler emits
// it: v_sqrt_f32, then s - ulp and s + ulp tested with fused residuals.
__device__ __forceinline__ float sqrt_select(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u);
  const float sp = __uint_as_float(__float_as_uint(s) + 1u);
  float t = (0.0f >= __builtin_fmaf(-sm, s, x)) ? sm : s;
  t = (0.0f < __builtin_fmaf(-sp, s, x)) ? sp : t;
  return t;
}
// The same rounding decision in integer arithmetic: the residuals are never
// -0 or NaN here, so e <= 0 iff (int)bits(e) - 1 < 0 and e > 0 iff
// -(int)bits(e) < 0.
__device__ __forceinline__ float sqrt_integer(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const uint32_t sb = __float_as_uint(s);
  const float sm = __uint_as_float(sb - 1u), sp = __uint_as_float(sb + 1u);
  const uint32_t em = __float_as_uint(__builtin_fmaf(-sm, s, x));
  const uint32_t ep = __float_as_uint(__builtin_fmaf(-sp, s, x));
  return __uint_as_float(sb - ((em - 1u) >> 31) + ((0u - ep) >> 31));
}
template <int WHICH>
__global__ void __launch_bounds__(kBlock) k_sqrt(float* out, float seed) {
  float x[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) x[i] = seed + (float)(threadIdx.x * 8 + i);
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = (WHICH == 0 ? sqrt_select(x[i]) : sqrt_integer(x[i])) + 1.5f;
  }
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < 8; ++i) s += x[i];
  out[blockIdx.x * kBlock + threadIdx.x] = s;
}
// The refined reciprocal fma(fma(-d, r0, 1), r0, r0), r0 = v_rcp_f32(d),
// against the correctly rounded 1.0f / d (this file is compiled with
// correctly rounded division) on every float in [lo, hi) and its negation;
// counts the inputs where they differ.
__global__ void __launch_bounds__(kBlock) k_rcp_compare(unsigned long long* bad, uint32_t lo, uint32_t hi,
                                                        uint32_t per) {
  const uint64_t b0 = (uint64_t)lo + ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * per;
  unsigned long long n = 0;
  for (uint32_t i = 0; i < per; ++i) {
    const uint64_t b = b0 + i;
    if (b >= hi) break;
#pragma unroll
    for (int sgn = 0; sgn < 2; ++sgn) {
      const float d = __uint_as_float((uint32_t)b | (sgn ? 0x80000000u : 0u));
      const float r0 = __builtin_amdgcn_rcpf(d);
      const float r = __builtin_fmaf(__builtin_fmaf(-d, r0, 1.0f), r0, r0);
      n += __float_as_uint(r) != __float_as_uint(1.0f / d);
    }
  }
  if (n) atomicAdd(bad, n);
}

// Both forms on every float in [lo, hi): each thread a run of `per` bit
// patterns; counts the inputs where they differ.
__global__ void __launch_bounds__(kBlock) k_sqrt_compare(unsigned long long* bad, uint32_t lo, uint32_t hi,
                                                         uint32_t per) {
  const uint64_t b0 = (uint64_t)lo + ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * per;
  unsigned long long n = 0;
  for (uint32_t i = 0; i < per; ++i) {
    const uint64_t b = b0 + i;
    if (b >= hi) break;
    const float x = __uint_as_float((uint32_t)b);
    n += __float_as_uint(sqrt_select(x)) != __float_as_uint(sqrt_integer(x));
  }
  if (n) atomicAdd(bad, n);
}

// LDS atomics: 8 per iteration at offsets kStep dwords apart, base address by
// pattern; every address stays inside the 8192-dword array.
constexpr int kWords = 8192, kStep = 1024, kLdsIters = 4096;
static_assert(64 * 3 * 2 + 8 + 7 * kStep + 2 <= kWords, "LDS probe addresses inside the array");
// MODE 0 ds_add_u32, 1 ds_add_f32, 2 ds_add_u64; PAT 0 stride 1, 1 stride 3, 2 one address per wave
template <int MODE, int PAT>
__global__ void __launch_bounds__(kBlock) k_lds(float* out, int seed) {
  __shared__ __align__(16) uint32_t s[kWords];
  for (int i = threadIdx.x; i < kWords; i += kBlock) s[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int base = PAT == 0 ? lane : (PAT == 1 ? lane * 3 : 0);
  if constexpr (MODE == 2) base *= 2;
  base += wave * 2;
  const uint32_t v = (uint32_t)(seed + lane);
  for (int it = 0; it < kLdsIters; ++it) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if constexpr (MODE == 0) atomicAdd(&s[base + j * kStep], v);
      else if constexpr (MODE == 1) atomicAdd(reinterpret_cast<float*>(&s[base + j * kStep]), 1.0f);
      else atomicAdd(reinterpret_cast<unsigned long long*>(&s[base + j * kStep]), (unsigned long long)v);
    }
  }
  __syncthreads();
  out[blockIdx.x * kBlock + threadIdx.x] = (float)s[threadIdx.x];
}

// HBM streams: read, write and copy of n float4 per buffer, grid-stride
// with 16-byte accesses (the read kernel folds what it reads into one value
// per thread, so the loads stay).
__global__ void __launch_bounds__(kBlock) k_stream_read(const float4* __restrict__ in, size_t n, float* out) {
  float acc = 0.0f;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
    const float4 v = in[i];
    acc += (v.x + v.y) + (v.z + v.w);
  }
  out[(size_t)blockIdx.x * kBlock + threadIdx.x] = acc;
}
__global__ void __launch_bounds__(kBlock) k_stream_write(float4* __restrict__ o, size_t n, float v) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
    o[i] = make_float4(v, v, v, v);
}
__global__ void __launch_bounds__(kBlock) k_stream_copy(const float4* __restrict__ in, float4* __restrict__ o,
                                                        size_t n) {
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) o[i] = in[i];
}

typedef void (*ValuFn)(float*, float);

int time_kernel(ValuFn k, int blocks, float* out, float* best_ms) {
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  hipLaunchKernelGGL(k, dim3(blocks), dim3(kBlock), 0, 0, out, 1.0f);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  float best = 1e30f;
  for (int rep = 0; rep < 5; ++rep) {
    CHECK(hipEventRecord(a, 0));
    hipLaunchKernelGGL(k, dim3(blocks), dim3(kBlock), 0, 0, out, 1.0f + rep);
    CHECK(hipEventRecord(b, 0));
    CHECK(hipEventSynchronize(b));
    float ms;
    CHECK(hipEventElapsedTime(&ms, a, b));
    if (ms < best) best = ms;
  }
  CHECK(hipEventDestroy(a));
  CHECK(hipEventDestroy(b));
  *best_ms = best;
  return 0;
}

template <int MODE, int PAT>
int time_lds(int blocks, float* out, float* best_ms) {
  hipEvent_t a, b;
  CHECK(hipEventCreate(&a));
  CHECK(hipEventCreate(&b));
  hipLaunchKernelGGL((k_lds<MODE, PAT>), dim3(blocks), dim3(kBlock), 0, 0, out, 1);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  float best = 1e30f;
  for (int rep = 0; rep < 5; ++rep) {
    CHECK(hipEventRecord(a, 0));
    hipLaunchKernelGGL((k_lds<MODE, PAT>), dim3(blocks), dim3(kBlock), 0, 0, out, rep);
    CHECK(hipEventRecord(b, 0));
    CHECK(hipEventSynchronize(b));
    float ms;
    CHECK(hipEventElapsedTime(&ms, a, b));
    if (ms < best) best = ms;
  }
  CHECK(hipEventDestroy(a));
  CHECK(hipEventDestroy(b));
  *best_ms = best;
  return 0;
}

}  // namespace

// Exhaustive equality checks of the exact fast forms the kernels rely on;
// returns the number of mismatching inputs, or -1 on a HIP error.
long long run_checks() {
  unsigned long long* bad = nullptr;
  if (hipMalloc(&bad, sizeof *bad) != hipSuccess) return -1;
  long long total = 0;
  const uint32_t per = 4096;
  struct C { const char* what; uint32_t lo, hi; int rcp; };
  const C cs[] = {{"refined reciprocal vs 1.0f / d, normal |d| < 2^126, both signs", 0x00800000u, 0x7E800000u, 1},
                  {"square root, select form vs integer form, [2^-96, inf)", 0x0F800000u, 0x7F800000u, 0}};
  for (const C& c : cs) {
    if (hipMemset(bad, 0, sizeof *bad) != hipSuccess) return -1;
    const uint64_t threads = ((uint64_t)(c.hi - c.lo) + per - 1) / per;
    const dim3 grid((unsigned)((threads + kBlock - 1) / kBlock));
    if (c.rcp)
      hipLaunchKernelGGL(k_rcp_compare, grid, dim3(kBlock), 0, 0, bad, c.lo, c.hi, per);
    else
      hipLaunchKernelGGL(k_sqrt_compare, grid, dim3(kBlock), 0, 0, bad, c.lo, c.hi, per);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return -1;
    unsigned long long nbad = 0;
    if (hipMemcpy(&nbad, bad, sizeof nbad, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    printf("%s: %llu of %llu inputs differ\n", c.what, nbad,
           (unsigned long long)(c.hi - c.lo) * (c.rcp ? 2ull : 1ull));
    total += (long long)nbad;
  }
  hipFree(bad);
  return total;
}

int main(int argc, char** argv) {
  if (argc > 1 && strcmp(argv[1], "--check") == 0) {
    const long long n = run_checks();
    printf("status %s\n", n == 0 ? "ok" : "mismatch");
    return n == 0 ? 0 : 1;
  }
  hipDeviceProp_t p;
  CHECK(hipGetDeviceProperties(&p, 0));
  const int cus = p.multiProcessorCount;
  const double simds = cus * 4.0;
  printf("# %s, %d CUs; cycles at %.1f GHz\n", p.gcnArchName, cus, kClockHz * 1e-9);
  const int blocks = cus * 8;  // 8 waves per SIMD
  float* out = nullptr;
  CHECK(hipMalloc(&out, (size_t)blocks * kBlock * sizeof(float)));
  struct Row { const char* name; const char* cls; ValuFn k; };
  const Row rows[] = {
    {"v_add_f32", "fp32 add/sub/mul/fma", k_valu0},
    {"v_sub_f32", "fp32 add/sub/mul/fma", k_valu1},
    {"v_mul_f32", "fp32 add/sub/mul/fma", k_valu2},
    {"v_fmac_f32", "fp32 add/sub/mul/fma", k_valu3},
    {"v_fma_f32 (VOP3)", "fp32 add/sub/mul/fma", k_valu4},
    {"v_add_f32_e64 |a|+|b|", "fp32 add/sub/mul/fma", k_valu5},
    {"v_add_u32", "int add/sub, bitwise, mov", k_valu6},
    {"v_sub_u32", "int add/sub, bitwise, mov", k_valu7},
    {"v_and_b32", "int add/sub, bitwise, mov", k_valu8},
    {"v_or_b32", "int add/sub, bitwise, mov", k_valu9},
    {"v_mov_b32", "int add/sub, bitwise, mov", k_valu10},
    {"v_min_f32", "min/max, compare, select, shift, cvt", k_valu11},
    {"v_max_f32", "min/max, compare, select, shift, cvt", k_valu12},
    {"v_med3_f32", "min/max, compare, select, shift, cvt", k_valu13},
    {"v_floor_f32", "min/max, compare, select, shift, cvt", k_valu14},
    {"v_cvt_i32_f32", "min/max, compare, select, shift, cvt", k_valu15},
    {"v_cvt_f32_i32", "min/max, compare, select, shift, cvt", k_valu16},
    {"v_cvt_f16_f32", "min/max, compare, select, shift, cvt", k_valu17},
    {"v_ldexp_f32", "min/max, compare, select, shift, cvt", k_valu18},
    {"v_lshlrev_b32", "min/max, compare, select, shift, cvt", k_valu19},
    {"v_bfe_u32", "min/max, compare, select, shift, cvt", k_valu20},
    {"v_lshl_add_u32", "min/max, compare, select, shift, cvt", k_valu21},
    {"v_mul_lo_u32", "min/max, compare, select, shift, cvt", k_valu22},
    {"v_mad_u32_u24", "min/max, compare, select, shift, cvt", k_valu23},
    {"v_sad_u8", "min/max, compare, select, shift, cvt", k_valu24},
    {"v_perm_b32", "min/max, compare, select, shift, cvt", k_valu25},
    {"v_cndmask_b32_e64", "min/max, compare, select, shift, cvt", k_valu26},
    {"v_mov_b32_dpp", "DPP", k_valu27},
    {"v_add_f32_dpp", "DPP", k_valu28},
    {"v_sqrt_f32", "transcendental", k_valu29},
    {"v_rcp_f32", "transcendental", k_valu30},
    {"v_exp_f32", "transcendental", k_valu31},
  };
  printf("| instruction | class | SIMD cycles per wave64 instruction |\n|---|---|---|\n");
  for (const Row& r : rows) {
    float ms;
    if (time_kernel(r.k, blocks, out, &ms)) return 1;
    const double instr = (double)blocks * (kBlock / 64) * kIters * kUnroll * 8;
    printf("| %s | %s | %.2f |\n", r.name, r.cls, simds * kClockHz * ms * 1e-3 / instr);
    fflush(stdout);
  }
  // the same streams with fewer waves per SIMD: 8 independent chains per lane
  // then leave part of each instruction's latency exposed
  const int picks[] = {3, 0, 4, 26, 29};  // v_fmac, v_add, v_fma (VOP3), v_cndmask_e64, v_sqrt
  printf("\n| instruction | 1 wave per SIMD | 2 | 4 | 8 |\n|---|---|---|---|---|\n");
  for (int pi : picks) {
    const Row& r = rows[pi];
    printf("| %s |", r.name);
    for (int w = 1; w <= 8; w *= 2) {
      float ms;
      if (time_kernel(r.k, cus * w, out, &ms)) return 1;
      const double instr = (double)cus * w * (kBlock / 64) * kIters * kUnroll * 8;
      printf(" %.2f |", simds * kClockHz * ms * 1e-3 / instr);
    }
    printf("\n");
    fflush(stdout);
  }
  struct SRow { const char* name; ValuFn k; double units; };
  const SRow srows[] = {
      {"v_add_f32 + s_nop 0 (per pair)", k_seq<0>, 8},
      {"v_add_f32 + s_nop 1 (per pair)", k_seq<1>, 8},
      {"v_cmp vcc + s_nop 1 + v_cndmask vcc (per triple)", k_seq<2>, 8},
      {"correctly rounded sqrt, compiler's select form (+ 1 add)", k_sqrt<0>, 8},
      {"correctly rounded sqrt, integer form (+ 1 add)", k_sqrt<1>, 8},
  };
  printf("\n| sequence | SIMD cycles per unit |\n|---|---|\n");
  for (const SRow& r : srows) {
    float ms;
    if (time_kernel(r.k, blocks, out, &ms)) return 1;
    const double units = (double)blocks * (kBlock / 64) * kIters * r.units;
    printf("| %s | %.2f |\n", r.name, simds * kClockHz * ms * 1e-3 / units);
    fflush(stdout);
  }
  {
    unsigned long long* bad = nullptr;
    CHECK(hipMalloc(&bad, sizeof *bad));
    CHECK(hipMemset(bad, 0, sizeof *bad));
    const uint32_t lo = 0x0F800000u, hi = 0x7F800000u, per = 4096;  // [2^-96, +inf)
    const uint64_t threads = ((uint64_t)(hi - lo) + per - 1) / per;
    hipLaunchKernelGGL(k_sqrt_compare, dim3((unsigned)((threads + kBlock - 1) / kBlock)), dim3(kBlock), 0, 0, bad, lo,
                       hi, per);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    unsigned long long nbad = 0;
    CHECK(hipMemcpy(&nbad, bad, sizeof nbad, hipMemcpyDeviceToHost));
    CHECK(hipFree(bad));
    printf("\nthe two square-root forms differ on %llu of the %u floats in [2^-96, inf)\n", nbad, hi - lo);
  }
  const int lblocks = cus * 4;  // 4 blocks (16 waves, 128 KiB) per CU
  struct LRow { const char* name; int (*f)(int, float*, float*); };
  const LRow lrows[] = {
      {"ds_add_u32, 64 distinct dwords", time_lds<0, 0>},
      {"ds_add_u32, stride 3 dwords", time_lds<0, 1>},
      {"ds_add_u32, one address per wave", time_lds<0, 2>},
      {"ds_add_f32, 64 distinct dwords", time_lds<1, 0>},
      {"ds_add_f32, stride 3 dwords", time_lds<1, 1>},
      {"ds_add_u64, 64 distinct qwords", time_lds<2, 0>},
  };
  printf("\n| LDS atomic | CU cycles per wave64 instruction | lane-operations per CU cycle |\n|---|---|---|\n");
  for (const LRow& r : lrows) {
    float ms;
    if (r.f(lblocks, out, &ms)) return 1;
    const double lane_ops = (double)lblocks * kBlock * kLdsIters * 8;
    const double cu_cycles = cus * kClockHz * ms * 1e-3;
    printf("| %s | %.2f | %.2f |\n", r.name, cu_cycles / (lane_ops / 64), lane_ops / cu_cycles);
    fflush(stdout);
  }
  {
    // 2 GiB per buffer: far beyond the 256 MB Infinity Cache
    const size_t n = ((size_t)2 << 30) / sizeof(float4);
    float4 *a = nullptr, *b = nullptr;
    float* acc = nullptr;
    const int sblocks = cus * 16;
    CHECK(hipMalloc(&a, n * sizeof(float4)));
    CHECK(hipMalloc(&b, n * sizeof(float4)));
    CHECK(hipMalloc(&acc, (size_t)sblocks * kBlock * sizeof(float)));
    hipLaunchKernelGGL(k_stream_write, dim3(sblocks), dim3(kBlock), 0, 0, a, n, 1.0f);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    printf("\n| HBM stream (2 GiB buffers, 16-byte accesses) | GB/s | fraction of 8 TB/s |\n|---|---|---|\n");
    for (int which = 0; which < 3; ++which) {
      float best = 1e30f;
      for (int rep = 0; rep < 5; ++rep) {
        CHECK(hipEventRecord(e0, 0));
        if (which == 0)
          hipLaunchKernelGGL(k_stream_read, dim3(sblocks), dim3(kBlock), 0, 0, a, n, acc);
        else if (which == 1)
          hipLaunchKernelGGL(k_stream_write, dim3(sblocks), dim3(kBlock), 0, 0, b, n, (float)rep);
        else
          hipLaunchKernelGGL(k_stream_copy, dim3(sblocks), dim3(kBlock), 0, 0, a, b, n);
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) best = ms;
      }
      const double bytes = (double)n * sizeof(float4) * (which == 2 ? 2 : 1);
      const double gbs = bytes / (best * 1e-3) * 1e-9;
      printf("| %s | %.0f | %.2f |\n", which == 0 ? "read" : (which == 1 ? "write" : "copy (read + write)"), gbs,
             gbs / 8000.0);
      fflush(stdout);
    }
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    CHECK(hipFree(acc));
  }
  CHECK(hipFree(out));
  printf("status ok\n");
  return 0;
}
