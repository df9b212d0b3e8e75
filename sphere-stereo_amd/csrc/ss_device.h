// Device-side building blocks shared by the sphere-sweeping stereo kernels.
//
// Every floating-point expression here is written in the operation order of
// the reference (KAIST-VCLAB/sphere-stereo) and the library is compiled with
// -ffp-contract=off and IEEE division/sqrt, so each operation is one
// round-to-nearest fp32 operation.  That is what makes argmin indices, guide
// bytes and z-test winners reproducible bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ss_types.h"

namespace ss {

// ---------------------------------------------------------------------------
// exp: Cody-Waite reduction (ln2 = C1 + C2) and the degree-6 polynomial of the
// Cephes library's expf (S. L. Moshier).  Used wherever the reference calls
// __expf (isb_filter.cu:87,229,236,241,248) or expf (stitcher.cu:389).
// The scaling by 2^n is one correctly rounded v_ldexp_f32, so subnormal
// results keep their value (fp32 denormals are not flushed in this build).
// ---------------------------------------------------------------------------
// Branch-free: the polynomial runs on the argument clamped to the finite
// range (unchanged inside it) and the out-of-range results are selected at
// the end, so the lanes of a wavefront never diverge here.
__device__ __forceinline__ float expf_cw(float x) {
  const float xc = fminf(fmaxf(x, -104.0f), 88.7228394f);
  float n = floorf(xc * 1.44269504088896341f + 0.5f);
  float r = xc - n * 0.693359375f;
  r = r - n * -2.12194440e-4f;
  float z = r * r;
  float p = 1.9875691500e-4f;
  p = p * r + 1.3981999507e-3f;
  p = p * r + 8.3334519073e-3f;
  p = p * r + 4.1665795894e-2f;
  p = p * r + 1.6666665459e-1f;
  p = p * r + 5.0000001201e-1f;
  p = p * z + r + 1.0f;
  float e = ldexpf(p, (int)n);
  e = x > 88.7228394f ? __builtin_huge_valf() : e;
  return x >= -104.0f ? e : ((x != x) ? x : 0.0f);
}

// ---------------------------------------------------------------------------
// Fast forms of the correctly rounded fp32 square root and division, for
// operands in the normal range.  They are the sequences the compiler emits
// for sqrtf and '/' in this build (-fhip-fp32-correctly-rounded-divide-
// sqrt, read from its gfx950 assembly) without the guards that only act on
// extreme operands, so they return the same bits wherever those guards are
// idle.
// ---------------------------------------------------------------------------
// v_sqrt_f32, then s - ulp and s + ulp tested with fused residuals; the full
// form also scales inputs below 2^-96 by 2^32 and passes +-0 and +inf
// through.  Equal to sqrtf(x) for x in [2^-96, +inf).
__device__ __forceinline__ float sqrt_rn_normal(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __uint_as_float(__float_as_uint(s) - 1u);
  const float sp = __uint_as_float(__float_as_uint(s) + 1u);
  float t = (0.0f >= __builtin_fmaf(-sm, s, x)) ? sm : s;
  t = (0.0f < __builtin_fmaf(-sp, s, x)) ? sp : t;
  return t;
}

// Division: a Newton-refined reciprocal of d, then two fused residual
// corrections of n * r.  The full form first scales the operands by 2^+-64
// (v_div_scale: only for denormal or near-overflow operands, an exponent
// gap of 96 or more, or a numerator below 2^-103) and ends with v_div_fixup
// (special values; for finite non-zero results it only restates the sign).
// The reciprocal refinement depends on d alone, so divisions by one d share
// it.
struct RcpRN {
  float d, r;
};
__device__ __forceinline__ RcpRN rcp_refined(float d) {
  const float r0 = __builtin_amdgcn_rcpf(d);
  const float e = __builtin_fmaf(-d, r0, 1.0f);
  return RcpRN{d, __builtin_fmaf(e, r0, r0)};
}
// 1.0f / d correctly rounded, for normal d with |d| < 2^126: the refined
// reciprocal equals RN(1/d) on every such input of either sign (checked
// exhaustively on the GPU: csrc/probes/issue_rates.hip, `--check`); three
// instructions instead of the guarded division sequence.
__device__ __forceinline__ float rcp_rn(float d) { return rcp_refined(d).r; }

__device__ __forceinline__ float div_rn_normal(float n, const RcpRN& q) {
  const float q0 = n * q.r;
  const float e1 = __builtin_fmaf(-q.d, q0, n);
  const float q1 = __builtin_fmaf(e1, q.r, q0);
  const float e2 = __builtin_fmaf(-q.d, q1, n);
  return __builtin_fmaf(e2, q.r, q1);
}

// ---------------------------------------------------------------------------
// Double-sphere camera, torch form (utils.py:57-94): fp32 tensor arithmetic
// whose Python-double scalars were rounded to fp32 once on the host.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void unproject_t(const TorchCam& c, float u, float v, float& px, float& py,
                                            float& pz, bool& valid) {
  float mx = (u - c.cs0) / c.fs0;
  float my = (v - c.cs1) / c.fs1;
  float r2 = mx * mx + my * my;
  float t = 1.0f - c.am2_1 * r2;
  float tc = t > 0.0f ? t : 0.0f;
  float mz = (1.0f - c.a2 * r2) / ((c.a * sqrtf(tc) + 1.0f) - c.a);
  float mz2 = mz * mz;
  float k = (mz * c.xi + sqrtf(mz2 + c.one_m_xi2 * r2)) / (mz2 + r2);
  px = k * mx;
  py = k * my;
  pz = k * mz - c.xi;
  valid = t >= 0.0f;
}

__device__ __forceinline__ void project_t(const TorchCam& c, float qx, float qy, float qz, float& u, float& v,
                                          bool& valid) {
  float xy2 = qx * qx + qy * qy;
  float d1 = sqrtf(xy2 + qz * qz);
  float cc = c.xi * d1 + qz;
  float d2 = sqrtf(xy2 + cc * cc);
  float nrm = c.a * d2 + c.one_m_a * cc;
  valid = qz > c.mw2 * d1;
  u = (c.fs0 * qx) / nrm + c.cs0;
  v = (c.fs1 * qy) / nrm + c.cs1;
}

// project_t (without the validity flag) with the fast square roots and one
// shared reciprocal for its two divisions.  Returns whether the result is
// certainly project_t's: |nrm| >= 2^-30, given camera parameters, transforms
// and candidate distances within sweep_fast_bounds (api.hip) and unit rays.
// Under those bounds the numerators stay below 2^62 and both square-root
// inputs below 2^96, so only a tiny nrm can wake a guard: it needs a point
// within 2^-48 of the camera centre (a square-root input below 2^-96) or
// makes the exponent gap reach 96.  A numerator so small that the division
// would scale it gives a quotient below 2^-73, which leaves u = q + cs0 equal
// to cs0 either way because |cs| >= 2^-16; a zero numerator gives a zero
// quotient whose sign cs0 absorbs.
__device__ __forceinline__ bool project_t_fast(const TorchCam& c, float qx, float qy, float qz, float& u, float& v) {
  float xy2 = qx * qx + qy * qy;
  float d1 = sqrt_rn_normal(xy2 + qz * qz);
  float cc = c.xi * d1 + qz;
  float d2 = sqrt_rn_normal(xy2 + cc * cc);
  float nrm = c.a * d2 + c.one_m_a * cc;
  const RcpRN q = rcp_refined(nrm);
  u = div_rn_normal(c.fs0 * qx, q) + c.cs0;
  v = div_rn_normal(c.fs1 * qy, q) + c.cs1;
  return fabsf(nrm) >= 0x1p-30f;
}

// project_t with its intermediate values (d1, cc, d2, nrm) as well, for a
// chain rule through it: project_t_fast's operations where `fast` allows
// them and they are certainly exact, else project_t's.  u, v and the
// intermediates are project_t's bits either way.
__device__ __forceinline__ void project_t_parts(const TorchCam& c, float qx, float qy, float qz, bool fast, float& u,
                                                float& v, float& d1, float& cc, float& d2, float& nrm) {
  const float xy2 = qx * qx + qy * qy;
  if (fast) {
    d1 = sqrt_rn_normal(xy2 + qz * qz);
    cc = c.xi * d1 + qz;
    d2 = sqrt_rn_normal(xy2 + cc * cc);
    nrm = c.a * d2 + c.one_m_a * cc;
    if (fabsf(nrm) >= 0x1p-30f) {
      const RcpRN q = rcp_refined(nrm);
      u = div_rn_normal(c.fs0 * qx, q) + c.cs0;
      v = div_rn_normal(c.fs1 * qy, q) + c.cs1;
      return;
    }
  }
  d1 = sqrtf(xy2 + qz * qz);
  cc = c.xi * d1 + qz;
  d2 = sqrtf(xy2 + cc * cc);
  nrm = c.a * d2 + c.one_m_a * cc;
  u = (c.fs0 * qx) / nrm + c.cs0;
  v = (c.fs1 * qy) / nrm + c.cs1;
}

// torch.matmul(cat([p, 1]), rt.T) for one 3x4 row-major transform, as the
// k-ordered multiply-add chain of a GEMM micro-kernel.
__device__ __forceinline__ void rigid(const float* rt, float px, float py, float pz, float& qx, float& qy,
                                      float& qz) {
  qx = __builtin_fmaf(pz, rt[2], __builtin_fmaf(py, rt[1], px * rt[0])) + rt[3];
  qy = __builtin_fmaf(pz, rt[6], __builtin_fmaf(py, rt[5], px * rt[4])) + rt[7];
  qz = __builtin_fmaf(pz, rt[10], __builtin_fmaf(py, rt[9], px * rt[8])) + rt[11];
}

// grid_sample's align_corners=False round trip: normalise (u+0.5)/size*2-1
// as depth_estimation.py:160-161 does, then torch's ((g+1)*size-1)/2.
__device__ __forceinline__ float grid_coord(float u, float size) {
  float g = ((u + 0.5f) / size) * 2.0f - 1.0f;
  return ((g + 1.0f) * size - 1.0f) / 2.0f;
}

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

// grid_coord2 in four operations for a power-of-two size, with inv2 = 2/size
// and half = size/2 (both exact).  Each step equals the corresponding steps
// of the long form exactly: t*(2/size) is the exact product 2*(t/size), so
// fma(t, inv2, -1) = RN(2*(t/size) - 1); (g+1)*size is exact, so
// fma(s, half, -0.5) = RN((s*size - 1)/2).  (Where t/size would be a
// subnormal, both forms give g = -1.  Where s*size would overflow, both
// coordinates fall outside the image and sample zero.)
__device__ __forceinline__ float grid_coord_pow2(float u, float inv2, float half) {
  const float t = u + 0.5f;
  const float g = __builtin_fmaf(t, inv2, -1.0f);
  const float s = g + 1.0f;
  return __builtin_fmaf(s, half, -0.5f);
}

// ---------------------------------------------------------------------------
// Double-sphere camera, stitcher form (stitcher.cu:121-168): fp32 throughout,
// no clamp in unproject; float2/float divides by reciprocal multiply.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void unproject_c(const ss_intrinsics& c, float u, float v, float& px, float& py,
                                            float& pz) {
  float mx = (u - c.cx) / c.fx;
  float my = (v - c.cy) / c.fy;
  float r2 = mx * mx + my * my;
  float mz = (1.0f - c.alpha * c.alpha * r2) /
             (c.alpha * sqrtf(1.0f - (2.0f * c.alpha - 1.0f) * r2) + 1.0f - c.alpha);
  float mz2 = mz * mz;
  float k = (mz * c.xi + sqrtf(mz2 + (1.0f - c.xi * c.xi) * r2)) / (mz2 + r2);
  px = mx * k;
  py = my * k;
  pz = mz * k - c.xi;
}

template <bool WITH_VALID>
__device__ __forceinline__ void project_c(const ss_intrinsics& c, float px, float py, float pz, float& u, float& v,
                                          bool& valid) {
  float d1 = sqrtf((px * px + py * py) + pz * pz);
  float cc = c.xi * d1 + pz;
  float d2 = sqrtf((px * px + py * py) + cc * cc);
  float nrm = c.alpha * d2 + (1.0f - c.alpha) * cc;
  if (WITH_VALID) {
    float w1 = (double)c.alpha > 0.5 ? (1.0f - c.alpha) / c.alpha : c.alpha / (1.0f - c.alpha);
    float w2 = (w1 + c.xi) / sqrtf(2.0f * w1 * c.xi + c.xi * c.xi + 1.0f);
    valid = valid && (pz > -w2 * d1);
  }
  float inv = 1.0f / nrm;
  u = (c.fx * px) * inv + c.cx;
  v = (c.fy * py) * inv + c.cy;
}

// __float2int_rn: round half to even, NaN -> 0, saturating
__device__ __forceinline__ int cvt_rni(float v) {
  if (v != v) return 0;
  const float r = rintf(v);
  if (r >= 2147483647.0f) return 2147483647;
  if (r <= -2147483648.0f) return -2147483647 - 1;
  return (int)r;
}

// reprojectDistanceKernel (stitcher.cu:181-204) for pixel i of an H x W map
// holding distance d, as an exact atomic z-minimum into dout (H x W, filled
// with 1e8 beforehand): distances are >= 0 (or NaN, which never wins), so the
// uint order of the bit patterns is the float order.
__device__ __forceinline__ void reproject_splat(const ss_intrinsics& c, const float* t, float* dout, int H, int W,
                                                int i, float d) {
  float px, py, pz, u, v;
  bool valid = true;
  unproject_c(c, (float)(i % W), (float)(i / W), px, py, pz);
  px = px * d - t[0];
  py = py * d - t[1];
  pz = pz * d - t[2];
  project_c<false>(c, px, py, pz, u, v, valid);
  const float dist = sqrtf((px * px + py * py) + pz * pz);
  const long long o = (long long)cvt_rni(v) * W + cvt_rni(u);
  if (o >= 0 && o < (long long)H * W && dist < 1e8f)
    atomicMin(reinterpret_cast<unsigned int*>(dout) + o, __float_as_uint(dist));
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
// Workgroups are dealt round-robin to the 8 XCDs; remapping gives each XCD a
// contiguous run of tiles so that stencil halos are shared inside one L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  int q = n >> 3;
  if (b >= (q << 3)) return b;
  return (b & 7) * q + (b >> 3);
}

// Sum of absolute byte differences over the three guide channels (byte 3 is
// zero in both words); the exact integer the reference's float absSum gives.
__device__ __forceinline__ float guide_sad(uint32_t a, uint32_t b) {
  return (float)__builtin_amdgcn_sad_u8(a, b, 0u);
}

// The same distance as an integer, the index of a filter's weight table.
__device__ __forceinline__ uint32_t guide_sad_u(uint32_t a, uint32_t b) { return __builtin_amdgcn_sad_u8(a, b, 0u); }

__device__ __forceinline__ float gch(uint32_t g, int c) { return (float)((g >> (8 * c)) & 0xffu); }

// float -> uint8 by truncation (vec_utils.cuh:111-114); inputs are convex
// combinations of bytes, inside [0, 256).
__device__ __forceinline__ uint32_t pack_guide(float a, float b, float c) {
  return (uint32_t)(uint8_t)a | ((uint32_t)(uint8_t)b << 8) | ((uint32_t)(uint8_t)c << 16);
}

}  // namespace ss
