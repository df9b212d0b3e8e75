// Cost-volume element types and vector loads/stores of V consecutive
// candidates.  The volume may be stored as fp32 (the reference's type), fp16
// or bf16; every kernel widens to fp32 on load and does all of its arithmetic
// in fp32, so only the stored values are rounded (round to nearest even).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ss {

// bf16 as its bit pattern: the upper half of an fp32
struct bf16_t {
  uint16_t u;
};

__device__ __forceinline__ float to_f32(float v) { return v; }
__device__ __forceinline__ float to_f32(_Float16 v) { return (float)v; }
__device__ __forceinline__ float to_f32(bf16_t v) { return __uint_as_float((uint32_t)v.u << 16); }

// fp32 -> 16-bit patterns, round to nearest even.  fp16 uses the hardware
// conversion (v_cvt_f16_f32 rounds to nearest even and keeps fp16
// subnormals); bf16 rounds the fp32 bit pattern (NaN stays a quiet NaN).
// The stored value is the fp32 result rounded to fp16: the empty asm pins
// that fp32 value in a register, because otherwise the backend folds the
// product or sum that made it into one v_fma_mix with a single rounding
// straight to fp16, which differs from the two roundings at exact fp16 ties.
__device__ __forceinline__ uint32_t f16_bits(float v) {
  asm("" : "+v"(v));
  return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)v);
}
__device__ __forceinline__ uint32_t bf16_bits(float v) {
  const uint32_t u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

template <typename T>
__device__ __forceinline__ uint32_t half_bits(float v);
template <>
__device__ __forceinline__ uint32_t half_bits<_Float16>(float v) {
  return f16_bits(v);
}
template <>
__device__ __forceinline__ uint32_t half_bits<bf16_t>(float v) {
  return bf16_bits(v);
}

template <typename T>
__device__ __forceinline__ float lo_f32(uint32_t w);
template <typename T>
__device__ __forceinline__ float hi_f32(uint32_t w);
template <>
__device__ __forceinline__ float lo_f32<_Float16>(uint32_t w) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(w & 0xffffu));
}
template <>
__device__ __forceinline__ float hi_f32<_Float16>(uint32_t w) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)(w >> 16));
}
template <>
__device__ __forceinline__ float lo_f32<bf16_t>(uint32_t w) {
  return __uint_as_float(w << 16);
}
template <>
__device__ __forceinline__ float hi_f32<bf16_t>(uint32_t w) {
  return __uint_as_float(w & 0xffff0000u);
}

// n consecutive elements at p (p aligned to the vector used) widened to fp32
template <int n, typename T>
__device__ __forceinline__ void load_n(const T* __restrict__ p, float* o) {
  if constexpr (sizeof(T) == 4) {
    if constexpr (n % 4 == 0) {
#pragma unroll
      for (int i = 0; i < n / 4; ++i) {
        const float4 t = reinterpret_cast<const float4*>(p)[i];
        o[4 * i] = t.x;
        o[4 * i + 1] = t.y;
        o[4 * i + 2] = t.z;
        o[4 * i + 3] = t.w;
      }
    } else if constexpr (n % 2 == 0) {
#pragma unroll
      for (int i = 0; i < n / 2; ++i) {
        const float2 t = reinterpret_cast<const float2*>(p)[i];
        o[2 * i] = t.x;
        o[2 * i + 1] = t.y;
      }
    } else {
#pragma unroll
      for (int i = 0; i < n; ++i) o[i] = p[i];
    }
  } else {
    if constexpr (n % 8 == 0) {
#pragma unroll
      for (int i = 0; i < n / 8; ++i) {
        const uint4 t = reinterpret_cast<const uint4*>(p)[i];
        const uint32_t w[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          o[8 * i + 2 * j] = lo_f32<T>(w[j]);
          o[8 * i + 2 * j + 1] = hi_f32<T>(w[j]);
        }
      }
    } else if constexpr (n % 4 == 0) {
#pragma unroll
      for (int i = 0; i < n / 4; ++i) {
        const uint2 t = reinterpret_cast<const uint2*>(p)[i];
        o[4 * i] = lo_f32<T>(t.x);
        o[4 * i + 1] = hi_f32<T>(t.x);
        o[4 * i + 2] = lo_f32<T>(t.y);
        o[4 * i + 3] = hi_f32<T>(t.y);
      }
    } else if constexpr (n % 2 == 0) {
#pragma unroll
      for (int i = 0; i < n / 2; ++i) {
        const uint32_t t = reinterpret_cast<const uint32_t*>(p)[i];
        o[2 * i] = lo_f32<T>(t);
        o[2 * i + 1] = hi_f32<T>(t);
      }
    } else {
#pragma unroll
      for (int i = 0; i < n; ++i) o[i] = to_f32(p[i]);
    }
  }
}

// n fp32 values rounded to T and stored at p
template <int n, typename T>
__device__ __forceinline__ void store_n(T* __restrict__ p, const float* v) {
  if constexpr (sizeof(T) == 4) {
    if constexpr (n % 4 == 0) {
#pragma unroll
      for (int i = 0; i < n / 4; ++i)
        reinterpret_cast<float4*>(p)[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    } else if constexpr (n % 2 == 0) {
#pragma unroll
      for (int i = 0; i < n / 2; ++i) reinterpret_cast<float2*>(p)[i] = make_float2(v[2 * i], v[2 * i + 1]);
    } else {
#pragma unroll
      for (int i = 0; i < n; ++i) p[i] = v[i];
    }
  } else {
    auto pack = [&](int i) __attribute__((always_inline)) {
      return half_bits<T>(v[i]) | (half_bits<T>(v[i + 1]) << 16);
    };
    if constexpr (n % 8 == 0) {
#pragma unroll
      for (int i = 0; i < n / 8; ++i)
        reinterpret_cast<uint4*>(p)[i] = make_uint4(pack(8 * i), pack(8 * i + 2), pack(8 * i + 4), pack(8 * i + 6));
    } else if constexpr (n % 4 == 0) {
#pragma unroll
      for (int i = 0; i < n / 4; ++i) reinterpret_cast<uint2*>(p)[i] = make_uint2(pack(4 * i), pack(4 * i + 2));
    } else if constexpr (n % 2 == 0) {
#pragma unroll
      for (int i = 0; i < n / 2; ++i) reinterpret_cast<uint32_t*>(p)[i] = pack(2 * i);
    } else {
#pragma unroll
      for (int i = 0; i < n; ++i) reinterpret_cast<uint16_t*>(p)[i] = (uint16_t)half_bits<T>(v[i]);
    }
  }
}

// Candidates handled per inner step of the filter: one 16-byte vector (4
// fp32 or 8 16-bit values) when V allows it.
template <int V, typename T>
constexpr int chunk_of() {
  if constexpr (sizeof(T) == 2 && V % 8 == 0) return 8;
  return V % 4 == 0 ? 4 : V;
}

}  // namespace ss

// Dispatch of a volume type code (ss_volume_type) onto the storage type T:
// CALL is a macro taking T.
#define SS_DISPATCH_T(vt, CALL)                 \
  switch (vt) {                                 \
    case SS_VOLUME_F32: return CALL(float);     \
    case SS_VOLUME_F16: return CALL(_Float16);  \
    case SS_VOLUME_BF16: return CALL(ss::bf16_t); \
    default: return hipErrorInvalidValue;       \
  }
