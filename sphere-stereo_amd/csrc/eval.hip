// Output conversion and evaluation on gfx950: inverse distance
// (main.py:114) and the error statistics of evaluate_rgbd_panorama
// (utils.py:233-241): mean absolute error, root mean square error and bad
// pixel ratios of two maps.
//
// The error of a pixel is the fp32 |a - b| the reference forms with numpy;
// the sums are accumulated in fp64 (numpy sums fp32 pairwise), in a fixed
// order: per-block partials, then one block adds them up, so the result does
// not depend on scheduling.
#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxThr = 4;
constexpr int kStats = 3 + kMaxThr;  // sum |e|, sum e^2, counted pixels, pixels above each threshold
constexpr int kReduceBlocks = 1024;

__global__ void __launch_bounds__(kBlock) k_inverse(const float* __restrict__ d, float* __restrict__ inv, size_t n) {
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) inv[i] = 1.0f / d[i];
}

struct StatArgs {
  const float* a;
  const float* b;
  size_t n;
  float thr[kMaxThr];
  int nthr, finite_only;
  double* partial;  // gridDim.x x kStats
};

__global__ void __launch_bounds__(kBlock) k_error_partial(StatArgs s) {
  double acc[kStats];
#pragma unroll
  for (int k = 0; k < kStats; ++k) acc[k] = 0.0;
  const size_t stride = (size_t)gridDim.x * kBlock;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < s.n; i += stride) {
    const float va = s.a[i], vb = s.b[i];
    if (s.finite_only && !(isfinite(va) && isfinite(vb))) continue;
    const float e = fabsf(va - vb);
    acc[0] += (double)e;
    acc[1] += (double)(e * e);  // err2 = err * err in fp32 (utils.py:236)
    acc[2] += 1.0;
#pragma unroll
    for (int k = 0; k < kMaxThr; ++k)
      if (k < s.nthr && e > s.thr[k]) acc[3 + k] += 1.0;
  }
  // block reduction in a fixed tree order
  __shared__ double red[kStats][kBlock];
#pragma unroll
  for (int k = 0; k < kStats; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
#pragma unroll
      for (int k = 0; k < kStats; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < kStats) s.partial[(size_t)blockIdx.x * kStats + threadIdx.x] = red[threadIdx.x][0];
}

// one block: the partials added in block order, one stat per lane
__global__ void __launch_bounds__(64) k_error_final(const double* __restrict__ partial, int nblocks,
                                                     double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= kStats) return;
  double sum = 0.0;
  for (int b = 0; b < nblocks; ++b) sum += partial[(size_t)b * kStats + k];
  out[k] = sum;
}

}  // namespace

hipError_t launch_inverse(const float* d, float* inv, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_inverse, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, d, inv, n);
  return hipGetLastError();
}

int error_stats_blocks(size_t n) {
  const size_t b = (n + kBlock - 1) / kBlock;
  return (int)(b < (size_t)kReduceBlocks ? (b < 1 ? 1 : b) : kReduceBlocks);
}

size_t error_stats_partial_bytes(size_t n) { return (size_t)error_stats_blocks(n) * kStats * sizeof(double); }

int error_stats_max_thresholds() { return kMaxThr; }

hipError_t launch_error_stats(const float* a, const float* b, size_t n, const float* thr, int nthr, int finite_only,
                              double* partial, double* out, hipStream_t s) {
  StatArgs st{};
  st.a = a;
  st.b = b;
  st.n = n;
  for (int k = 0; k < nthr && k < kMaxThr; ++k) st.thr[k] = thr[k];
  st.nthr = nthr;
  st.finite_only = finite_only;
  st.partial = partial;
  const int blocks = error_stats_blocks(n);
  hipLaunchKernelGGL(k_error_partial, dim3(blocks), dim3(kBlock), 0, s, st);
  hipLaunchKernelGGL(k_error_final, dim3(1), dim3(64), 0, s, partial, blocks, out);
  return hipGetLastError();
}

}  // namespace ss
