// Sphere sweep, camera selection and guide construction on gfx950.
//
// Behaviour: depth_estimation.py:83-175 and utils.py:57-94, 144-155 of the
// reference.  The reference evaluates the sweep as a chain of full-volume
// torch ops per camera (~10 GB of HBM traffic per camera per reference at
// 1024^2, 32 candidates); here one pass writes the candidate-contiguous cost
// volume directly.  A group of 8 lanes owns a reference pixel; each lane
// sweeps V consecutive candidates and stores them as one 16-byte vector, so
// the volume is written with full-line coalescing.  Only the selected camera
// of each pixel is sampled (the reference samples every camera and keeps the
// selected one), which leaves the result unchanged.
#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kBlock = 256;

inline int blocks_for(long long threads) { return (int)((threads + kBlock - 1) / kBlock); }

// grid_coord with the division by a power-of-two size done as an exact
// multiply by its reciprocal (a compile-time choice, so that the other
// form is not evaluated as well).
template <bool POW2>
__device__ __forceinline__ float grid_coord2(float u, float size, float inv) {
  const float t = u + 0.5f;
  const float q = POW2 ? t * inv : t / size;
  const float g = q * 2.0f - 1.0f;
  return ((g + 1.0f) * size - 1.0f) * 0.5f;
}

// Zero-padded bilinear sample of an HWC image with nch channels, tap order
// and accumulation of torch's grid_sample (nw, ne, sw, se from 0).  Taps
// outside the image read a clamped texel with weight 0, which adds exactly 0.
template <int NCH>
__device__ __forceinline__ void sample_zero(const float* __restrict__ img, int H, int W, float ix, float iy,
                                            float* out) {
#pragma unroll
  for (int c = 0; c < NCH; ++c) out[c] = 0.0f;
  const float x0f = floorf(ix), y0f = floorf(iy);
  if (!(x0f >= -1.0f && x0f <= (float)(W - 1) && y0f >= -1.0f && y0f <= (float)(H - 1))) return;
  const int x0 = (int)x0f, y0 = (int)y0f;
  const float x1f = x0f + 1.0f, y1f = y0f + 1.0f;
  float wnw = (x1f - ix) * (y1f - iy), wne = (ix - x0f) * (y1f - iy);
  float wsw = (x1f - ix) * (iy - y0f), wse = (ix - x0f) * (iy - y0f);
  const bool bx0 = x0 >= 0, bx1 = x0 + 1 < W, by0 = y0 >= 0, by1 = y0 + 1 < H;
  wnw = (bx0 && by0) ? wnw : 0.0f;
  wne = (bx1 && by0) ? wne : 0.0f;
  wsw = (bx0 && by1) ? wsw : 0.0f;
  wse = (bx1 && by1) ? wse : 0.0f;
  const int xa = bx0 ? x0 : 0, xb = bx1 ? x0 + 1 : W - 1;
  const int ya = by0 ? y0 : 0, yb = by1 ? y0 + 1 : H - 1;
  // 32-bit offsets: an image plane holds fewer than 2^31 / NCH texels
  const float* pnw = img + (ya * W + xa) * NCH;
  const float* pne = img + (ya * W + xb) * NCH;
  const float* psw = img + (yb * W + xa) * NCH;
  const float* pse = img + (yb * W + xb) * NCH;
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    float s = pnw[c] * wnw;
    s = s + pne[c] * wne;
    s = s + psw[c] * wsw;
    s = s + pse[c] * wse;
    out[c] = s;
  }
}

// rgb2yCbCr (utils.py:144-155) + truncating uint8 cast, packed
__global__ void __launch_bounds__(kBlock) k_guide(const float* __restrict__ img, int HW, uint32_t* __restrict__ g) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= HW) return;
  const float c0 = img[3 * p], c1 = img[3 * p + 1], c2 = img[3 * p + 2];
  float y = ((16.0f + 0.1826f * c0) + 0.6142f * c1) + 0.062f * c2;
  float cb = ((128.0f - 0.1006f * c0) - 0.3386f * c1) + 0.4392f * c2;
  float cr = ((128.0f + 0.4392f * c0) - 0.3989f * c1) - 0.0403f * c2;
  y = y < 16.0f ? 16.0f : (y > 235.0f ? 235.0f : y);
  cb = cb < 16.0f ? 16.0f : (cb > 240.0f ? 240.0f : cb);
  cr = cr < 16.0f ? 16.0f : (cr > 240.0f ? 240.0f : cr);
  g[p] = pack_guide(y, cb, cr);
}

// unit rays of the reference (depth_estimation.py:141-143) + selection index
__global__ void __launch_bounds__(kBlock) k_ray_lut(TorchCam ref, const int8_t* __restrict__ sel, int H, int W,
                                                    float4* __restrict__ lut) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= H * W) return;
  float px, py, pz;
  bool valid;
  unproject_t(ref, (float)(p % W), (float)(p / W), px, py, pz, valid);
  lut[p] = make_float4(px, py, pz, (float)sel[p]);
}

// RGBD_Estimator.select_camera (depth_estimation.py:83-135), one pixel per lane
__global__ void __launch_bounds__(kBlock) k_select_camera(SelectArgs a) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  const int HW = a.H * a.W;
  if (p >= HW) return;
  float rx, ry, rz;
  bool ref_valid;
  unproject_t(a.ref, (float)(p % a.W), (float)(p / a.W), rx, ry, rz, ref_valid);
  const float mref = a.masks[(size_t)a.refidx * HW + p];
  int best = -1;
  float maxd = 1.0f;
  for (int k = 0; k < a.C; ++k) {
    const float* rt = a.rt[k];
    float nx, ny, nz, fx, fy, fz;
    rigid(rt, rx * a.mn, ry * a.mn, rz * a.mn, nx, ny, nz);
    rigid(rt, rx * a.mx, ry * a.mx, rz * a.mx, fx, fy, fz);
    const float nn = sqrtf((nx * nx + ny * ny) + nz * nz);
    const float nf = sqrtf((fx * fx + fy * fy) + fz * fz);
    nx = nx / nn;
    ny = ny / nn;
    nz = nz / nn;
    fx = fx / nf;
    fy = fy / nf;
    fz = fz / nf;
    float un, vn, uf, vf;
    bool okn, okf;
    project_t(a.cam[k], nx, ny, nz, un, vn, okn);
    project_t(a.cam[k], fx, fy, fz, uf, vf, okf);
    const float dx = un - uf, dy = vn - vf;
    const float disp = sqrtf(dx * dx + dy * dy);
    const float* mk = a.masks + (size_t)k * HW;
    float mnear, mfar;
    const float gun = a.pow2W ? grid_coord2<true>(un, a.Wf, a.invW) : grid_coord2<false>(un, a.Wf, a.invW);
    const float gvn = a.pow2H ? grid_coord2<true>(vn, a.Hf, a.invH) : grid_coord2<false>(vn, a.Hf, a.invH);
    const float guf = a.pow2W ? grid_coord2<true>(uf, a.Wf, a.invW) : grid_coord2<false>(uf, a.Wf, a.invW);
    const float gvf = a.pow2H ? grid_coord2<true>(vf, a.Hf, a.invH) : grid_coord2<false>(vf, a.Hf, a.invH);
    sample_zero<1>(mk, a.H, a.W, gun, gvn, &mnear);
    sample_zero<1>(mk, a.H, a.W, guf, gvf, &mfar);
    if (disp > maxd && ref_valid && okn && okf && mref >= 0.9f && mnear >= 0.9f && mfar >= 0.9f) {
      maxd = disp;
      best = k;
    }
  }
  a.sel[p] = (int8_t)best;
}

// Sweep + raw cost (depth_estimation.py:141-175)
template <int V, bool POW2W, bool POW2H>
__global__ void __launch_bounds__(kBlock) k_sweep(SweepArgs a) {
  constexpr int G = 8, NP = G * V;
  const int lane = threadIdx.x & (G - 1);
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int pix = (tile * kBlock + (int)threadIdx.x) / G;
  if (pix >= a.H * a.W) return;
  const float4 rl = a.raylut[pix];
  const int k = (int)rl.w;
  const float r0 = a.ref_img[3 * pix], r1 = a.ref_img[3 * pix + 1], r2 = a.ref_img[3 * pix + 2];
  const int n0 = lane * V;
  float out[V];
  if (k < 0) {
    // nothing sampled: the volume holds zeros there (depth_estimation.py:150-152)
    float c = (r0 + r1) + r2;
    c = c > 500.0f ? 500.0f : c;
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = (n0 + v < a.N) ? c : 0.0f;
  } else {
    const TorchCam& cam = a.cam[k];
    const float* rt = a.rt[k];
    const float* img = a.img[k];
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int n = n0 + v;
      float c = 0.0f;
      if (n < a.N) {
        const float d = a.cands[n];
        float qx, qy, qz, u, w;
        bool valid;
        rigid(rt, d * rl.x, d * rl.y, d * rl.z, qx, qy, qz);
        project_t(cam, qx, qy, qz, u, w, valid);
        float s[3];
        sample_zero<3>(img, a.H, a.W, grid_coord2<POW2W>(u, a.Wf, a.invW), grid_coord2<POW2H>(w, a.Hf, a.invH), s);
        c = (fabsf(s[0] - r0) + fabsf(s[1] - r1)) + fabsf(s[2] - r2);
        c = c > 500.0f ? 500.0f : c;
      }
      out[v] = c;
    }
  }
  float* dst = a.cost + (size_t)pix * NP + n0;
  if constexpr (V % 4 == 0) {
#pragma unroll
    for (int i = 0; i < V / 4; ++i)
      reinterpret_cast<float4*>(dst)[i] = make_float4(out[4 * i], out[4 * i + 1], out[4 * i + 2], out[4 * i + 3]);
  } else if constexpr (V % 2 == 0) {
#pragma unroll
    for (int i = 0; i < V / 2; ++i) reinterpret_cast<float2*>(dst)[i] = make_float2(out[2 * i], out[2 * i + 1]);
  } else {
#pragma unroll
    for (int v = 0; v < V; ++v) dst[v] = out[v];
  }
}

template <int V>
hipError_t launch_sweep_v(const SweepArgs& a, hipStream_t s) {
  const dim3 grid(blocks_for((long long)a.H * a.W * 8));
  if (a.pow2W && a.pow2H)
    hipLaunchKernelGGL((k_sweep<V, true, true>), grid, dim3(kBlock), 0, s, a);
  else if (a.pow2W)
    hipLaunchKernelGGL((k_sweep<V, true, false>), grid, dim3(kBlock), 0, s, a);
  else if (a.pow2H)
    hipLaunchKernelGGL((k_sweep<V, false, true>), grid, dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_sweep<V, false, false>), grid, dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_guide(const float* img, int HW, uint32_t* guide, hipStream_t s) {
  hipLaunchKernelGGL(k_guide, dim3(blocks_for(HW)), dim3(kBlock), 0, s, img, HW, guide);
  return hipGetLastError();
}

hipError_t launch_ray_lut(const TorchCam& ref, const int8_t* sel, int H, int W, float4* lut, hipStream_t s) {
  hipLaunchKernelGGL(k_ray_lut, dim3(blocks_for((long long)H * W)), dim3(kBlock), 0, s, ref, sel, H, W, lut);
  return hipGetLastError();
}

hipError_t launch_select_camera(const SelectArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_select_camera, dim3(blocks_for((long long)a.H * a.W)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_sweep(const SweepArgs& a, int V, hipStream_t s) {
  switch (V) {
    case 1: return launch_sweep_v<1>(a, s);
    case 2: return launch_sweep_v<2>(a, s);
    case 3: return launch_sweep_v<3>(a, s);
    case 4: return launch_sweep_v<4>(a, s);
    case 6: return launch_sweep_v<6>(a, s);
    case 8: return launch_sweep_v<8>(a, s);
    case 12: return launch_sweep_v<12>(a, s);
    case 16: return launch_sweep_v<16>(a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ss
