// Input stage on gfx950: the per-frame image preparation of the reference's
// read_input_images (utils.py:157-204): the decoded fisheye converted to
// float32 in [0, 255] (uint8 as is, uint16 / 255, utils.py:174-177), then
// cv2.resize(image, (cols, rows), cv2.INTER_AREA) to the matching and stitch
// resolutions (utils.py:198-201).  INTER_AREA sits in cv2.resize's dst slot
// there, so the interpolation is cv2's default INTER_LINEAR: half-pixel
// centres, the source coordinate (d + 0.5) * (S / D) - 0.5 evaluated in
// double and rounded to float, clamped to the first / last sample, then a
// horizontal pass and a vertical pass, each a * w0 + b * w1.
#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kBlock = 256;

// cv2 INTER_LINEAR coefficients of one destination coordinate.
__device__ __forceinline__ void resize_axis(int d, int S, double scale, int& s0, int& s1, float& w0, float& w1) {
  float f = (float)(((double)d + 0.5) * scale - 0.5);
  int s = (int)floorf(f);
  f = f - (float)s;
  if (s < 0) {
    f = 0.0f;
    s = 0;
  }
  if (s >= S - 1) {
    f = 0.0f;
    s = S - 1;
  }
  s0 = s;
  s1 = s + 1 < S ? s + 1 : S - 1;
  w0 = 1.0f - f;
  w1 = f;
}

template <typename T>
__device__ __forceinline__ float input_value(T v) {
  if constexpr (sizeof(T) == 2) return (float)v / 255.0f;  // uint16 (utils.py:176)
  else return (float)v;                                    // uint8, or float32 already in [0, 255]
}

// One lane per destination pixel, all its channels (HWC in and out).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_resize_linear(const T* __restrict__ src, int Hs, int Ws, int C,
                                                          float* __restrict__ dst, int Hd, int Wd, double scale_x,
                                                          double scale_y) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= Hd * Wd) return;
  const int dx = i % Wd, dy = i / Wd;
  int x0, x1, y0, y1;
  float a0, a1, b0, b1;
  resize_axis(dx, Ws, scale_x, x0, x1, a0, a1);
  resize_axis(dy, Hs, scale_y, y0, y1, b0, b1);
  const T* r0 = src + (size_t)y0 * Ws * C;
  const T* r1 = src + (size_t)y1 * Ws * C;
  for (int c = 0; c < C; ++c) {
    const float h0 = input_value(r0[x0 * C + c]) * a0 + input_value(r0[x1 * C + c]) * a1;
    const float h1 = input_value(r1[x0 * C + c]) * a0 + input_value(r1[x1 * C + c]) * a1;
    dst[(size_t)i * C + c] = h0 * b0 + h1 * b1;
  }
}

inline int blocks_for(long long threads) { return (int)((threads + kBlock - 1) / kBlock); }

}  // namespace

// cv2's scale: 1 / (D / S), both in double
inline double resize_scale(int S, int D) { return 1.0 / ((double)D / (double)S); }

hipError_t launch_resize_linear(const void* src, int src_type, int Hs, int Ws, int C, float* dst, int Hd, int Wd,
                                hipStream_t s) {
  const dim3 grid(blocks_for((long long)Hd * Wd));
  const double sx = resize_scale(Ws, Wd), sy = resize_scale(Hs, Hd);
  switch (src_type) {
    case SS_INPUT_U8:
      hipLaunchKernelGGL(k_resize_linear<uint8_t>, grid, dim3(kBlock), 0, s, static_cast<const uint8_t*>(src), Hs,
                         Ws, C, dst, Hd, Wd, sx, sy);
      break;
    case SS_INPUT_U16:
      hipLaunchKernelGGL(k_resize_linear<uint16_t>, grid, dim3(kBlock), 0, s, static_cast<const uint16_t*>(src), Hs,
                         Ws, C, dst, Hd, Wd, sx, sy);
      break;
    case SS_INPUT_F32:
      hipLaunchKernelGGL(k_resize_linear<float>, grid, dim3(kBlock), 0, s, static_cast<const float*>(src), Hs, Ws,
                         C, dst, Hd, Wd, sx, sy);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ss
