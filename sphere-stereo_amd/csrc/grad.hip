// Differentiation form on gfx950: a soft-argmin readout over the filtered
// cost volume, hand-written adjoints of the ISB filter and of the sweep, and
// the Gauss-Newton normal equations of a photometric calibration refinement.
//
// The forward is the plain path's own sweep (sweep.hip) and ISB filter
// (isb.hip) on an fp32 candidate-contiguous volume; only the readout differs:
// depth_estimation.py:180-198 takes a hard argmin with parabola refinement,
// which has no useful derivative, so this form reads the inverse distance as
//   iota = sum_n n * softmax_n(-beta * F_n),  1/d = 1/d_0 + (1/d_{N-1} - 1/d_0) * iota / (N - 1)
// (the candidates are uniform in inverse distance, depth_estimation.py:145-146).
//
// Adjoints.  For a fixed guide the ISB filter is linear in the cost, so its
// adjoint is the transposed stencil chain; nothing of the forward volume
// pyramid is kept: the guide pyramid (3 bytes a pixel) is recomputed and the
// normalised stencil weights of every level are tabulated from it, then the
// transposed passes run as gathers (each output element owned by one lane,
// no atomics, deterministic).  The sweep adjoint recomputes every sample
// (projection, bilinear taps) from the rays, and returns the gradient with
// respect to the images (summed per block in an LDS window of the sampled
// image, then added with float atomics), to each camera's double-sphere parameters (fs, cs, xi, alpha) and
// to each (camera <- reference) 3 x 4 transform; the transform gradients
// chain to 6-DoF extrinsics on the host.  Per-block partial sums are reduced
// in a fixed order, so the parameter gradients are deterministic too.
// Piecewise-constant parts have zero derivative: the camera selection, the
// uint8 guide and the floor of the bilinear cell.
#include <stdio.h>
#include <string.h>

#include <cmath>

#include "ss_isb_stencil.h"
#include "ss_kernels.h"

namespace ss {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ void st4(float* p, float4 v) { *reinterpret_cast<float4*>(p) = v; }
__device__ __forceinline__ float4 fma4(float w, float4 x, float4 acc) {
  return make_float4(acc.x + w * x.x, acc.y + w * x.y, acc.z + w * x.z, acc.w + w * x.w);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---------------------------------------------------------------------------
// Soft-argmin readout and its adjoint (one lane per pixel)
// ---------------------------------------------------------------------------
struct SoftArgs {
  const float* F;  // filtered volume [P][NP]
  int P, N, NP;
  float beta, inv0, dinv;  // dinv = 1/d_{N-1} - 1/d_0
  const float* ginvd;      // backward: dL/d(1/d) [P]
  float* out;              // forward: 1/d [P]; backward: dL/dF [P][NP]
};

// A group of 8 lanes per pixel, lane j holding candidates [j*V, j*V + V) of
// the NP = 8 * V stride, so that a wavefront reads 8 whole pixel rows.
constexpr int kSoftGroup = 8;

__device__ __forceinline__ float group_min8(float v) {
#pragma unroll
  for (int off = 4; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, kSoftGroup));
  return v;
}
__device__ __forceinline__ float group_sum8(float v) {
#pragma unroll
  for (int off = 4; off > 0; off >>= 1) v += __shfl_xor(v, off, kSoftGroup);
  return v;
}

// The softmax statistics of one pixel: minimum m, z = sum_n e_n and
// s = sum_n n e_n with e_n = exp(-beta (F_n - m)), in every lane of the group.
struct SoftStats {
  float m, z, s;
};
__device__ __forceinline__ SoftStats soft_stats(const float* f, int n0, int V, int N, float beta) {
  float m = __builtin_huge_valf();
  for (int v = 0; v < V; ++v)
    if (n0 + v < N) m = fminf(m, f[v]);
  m = group_min8(m);
  float z = 0.0f, s = 0.0f;
  for (int v = 0; v < V; ++v)
    if (n0 + v < N) {
      const float e = expf(-beta * (f[v] - m));
      z += e;
      s += (float)(n0 + v) * e;
    }
  return SoftStats{m, group_sum8(z), group_sum8(s)};
}

__global__ void __launch_bounds__(kBlock) k_soft_argmin(SoftArgs a) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const int lane = (int)(threadIdx.x & (kSoftGroup - 1));
  long long p = t / kSoftGroup;
  const bool active = p < a.P;
  if (!active) p = a.P - 1;  // the whole group takes part in the shuffles
  const int V = a.NP / kSoftGroup, n0 = lane * V;
  const SoftStats st = soft_stats(a.F + (size_t)p * a.NP + n0, n0, V, a.N, a.beta);
  if (active && lane == 0) a.out[p] = a.inv0 + a.dinv * ((st.s / st.z) / (float)(a.N - 1));
}

// dL/dF_n = dL/diota * softmax_n * (n - iota) * (-beta)
__global__ void __launch_bounds__(kBlock) k_soft_argmin_bwd(SoftArgs a) {
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const int lane = (int)(threadIdx.x & (kSoftGroup - 1));
  long long p = t / kSoftGroup;
  const bool active = p < a.P;
  if (!active) p = a.P - 1;
  const int V = a.NP / kSoftGroup, n0 = lane * V;
  const float* f = a.F + (size_t)p * a.NP + n0;
  const SoftStats st = soft_stats(f, n0, V, a.N, a.beta);
  if (!active) return;
  const float iota = st.s / st.z;
  const float gi = a.ginvd[p] * a.dinv / (float)(a.N - 1);
  const float c = -a.beta * gi / st.z;
  float* g = a.out + (size_t)p * a.NP + n0;
  for (int v = 0; v < V; ++v) {
    const int n = n0 + v;
    g[v] = n < a.N ? c * expf(-a.beta * (f[v] - st.m)) * ((float)n - iota) : 0.0f;
  }
}

// ---------------------------------------------------------------------------
// Guide pyramid and normalised stencil weights of every level
// ---------------------------------------------------------------------------
// Down pass into level s (isb_filter.cu:44-119): the coarse guide, as the
// forward computes it, and the weights with which the coarse cost takes its
// 9 fine inputs: wdn[o][0] = 1/ws for the centre, wdn[o][1 + j] = w_j/ws for
// neighbour j in the order of isb_filter.cu:68-77.
constexpr int kWdn = 12;
struct GDownArgs {
  const uint32_t* gin;
  int hin, win;
  uint32_t* gout;
  int hout, wout;
  float var_inv_i;
  float* wdn;
};

__global__ void __launch_bounds__(kBlock) k_gdown(GDownArgs a) {
  const int pix = blockIdx.x * kBlock + threadIdx.x;
  if (pix >= a.hout * a.wout) return;
  const int x = pix % a.wout, y = pix / a.wout;
  const int x2 = 2 * x, y2 = 2 * y;
  const int xp = min(x2 + 1, a.win - 1), yp = min(y2 + 1, a.hin - 1);
  const int xm = max(x2 - 1, 0), ym = max(y2 - 1, 0);
  const int nb[8] = {ym * a.win + xm, y2 * a.win + xm, yp * a.win + xm, yp * a.win + x2,
                     yp * a.win + xp, y2 * a.win + xp, ym * a.win + xp, ym * a.win + x2};
  const uint32_t gc = a.gin[y2 * a.win + x2];
  float w[8], ws = 1.0f;
  float s0 = gch(gc, 0), s1 = gch(gc, 1), s2 = gch(gc, 2);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t gn = a.gin[nb[j]];
    const float d = guide_sad(gc, gn);
    w[j] = ((j & 1) ? 0.5f : 0.25f) * (expf_cw(-d * d * a.var_inv_i) + 0.01f);
    ws = ws + w[j];
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t gn = a.gin[nb[j]];
    s0 = s0 + gch(gn, 0) * w[j];
    s1 = s1 + gch(gn, 1) * w[j];
    s2 = s2 + gch(gn, 2) * w[j];
  }
  const float inv = 1.0f / ws;
  a.gout[pix] = pack_guide(s0 * inv, s1 * inv, s2 * inv);
  float* o = a.wdn + (size_t)pix * kWdn;
  o[0] = 1.0f / ws;
#pragma unroll
  for (int j = 0; j < 8; ++j) o[1 + j] = w[j] / ws;
  o[9] = o[10] = o[11] = 0.0f;
}

// Up pass out of owner level s+1 into level s (isb_filter.cu:132-330): the
// weights with which fine pixel f (stencil t of owner o) takes its taps and
// its own down-pass value, wup[o][k] = w_k / ws_t for the 25 taps and
// wup[o][25 + t] = weight_up / ws_t, and the updated fine guide, which is the
// coarse guide of the next finer up pass (not written for level 0).
constexpr int kWup = 32;
struct GUpArgs {
  const uint32_t* gin;  // updated guide of level s+1
  int hin, win;
  uint32_t* gio;        // guide of level s (down-pass values, updated in place)
  int hio, wio;
  float wu, wd, var_inv_i;
  float* wup;
  int write_guide;
};

__global__ void __launch_bounds__(kBlock) k_gup(GUpArgs a) {
  const int pix = blockIdx.x * kBlock + threadIdx.x;
  if (pix >= a.hin * a.win) return;
  const int x = pix % a.win, y = pix / a.win;
  const int r0 = max(y - 1, 0), r2 = min(y + 1, a.hin - 1);
  const int c0 = max(x - 1, 0), c2 = min(x + 1, a.win - 1);
  const int x2 = 2 * x, y2 = 2 * y;
  const int xpu = min(x2 + 1, a.wio - 1), ypu = min(y2 + 1, a.hio - 1);
  const bool setY = y2 != ypu, setX = x2 != xpu;
  const bool doit[4] = {true, setY, setX, setY && setX};
  const int fine[4] = {y2 * a.wio + x2, ypu * a.wio + x2, y2 * a.wio + xpu, ypu * a.wio + xpu};
  uint32_t gF[4], gD[9], gout[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (int s = 0; s < 4; ++s) gF[s] = a.gio[fine[s]];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) gD[r * 3 + c] = a.gin[pick3(r, r0, y, r2) * a.win + pick3(c, c0, x, c2)];
  float w[25];
#pragma unroll
  for (int k = 0; k < 25; ++k) {
    const float d = guide_sad(gD[up_tap_flat(k)], gF[up_stencil(k)]);
    w[k] = up_base(k) * (a.wd * (expf_cw(-d * d * a.var_inv_i) + 0.01f));
  }
  float* o = a.wup + (size_t)pix * kWup;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    float ws = a.wu;
#pragma unroll
    for (int k = kStencilBegin[s]; k < kStencilBegin[s + 1]; ++k) ws = ws + w[k];
    const float inv = 1.0f / ws;
#pragma unroll
    for (int k = kStencilBegin[s]; k < kStencilBegin[s + 1]; ++k) o[k] = w[k] / ws;
    o[25 + s] = a.wu / ws;
    if (a.write_guide) {
      float s0 = gch(gF[s], 0) * a.wu, s1 = gch(gF[s], 1) * a.wu, s2 = gch(gF[s], 2) * a.wu;
#pragma unroll
      for (int k = kStencilBegin[s]; k < kStencilBegin[s + 1]; ++k) {
        const uint32_t g = gD[up_tap_flat(k)];
        s0 = s0 + gch(g, 0) * w[k];
        s1 = s1 + gch(g, 1) * w[k];
        s2 = s2 + gch(g, 2) * w[k];
      }
      gout[s] = pack_guide(s0 * inv, s1 * inv, s2 * inv);
    }
  }
  o[29] = o[30] = o[31] = 0.0f;
  if (a.write_guide) {
#pragma unroll
    for (int s = 0; s < 4; ++s)
      if (doit[s]) a.gio[fine[s]] = gout[s];
  }
}

// ---------------------------------------------------------------------------
// Transposed filter passes, one lane per (pixel, 4 candidates)
// ---------------------------------------------------------------------------
// Up-pass adjoint between fine level s and coarse level s+1.  Lane (c, q):
//  (a) for the fine pixels c owns: gC_s[f] = wup[c][25 + t] * gF_s[f]
//      (the part of the fine level's gradient that reaches its down-pass
//      values directly);
//  (b) gF_{s+1}[c] = sum over the owners o of c's 3x3 neighbourhood, their
//      fine pixels f and the taps k of f that read c, of wup[o][k] * gF_s[f].
struct UpAdjArgs {
  const float* gF;  // dL/d(filtered level s) [Pf][NP]
  int hio, wio;
  int hin, win;
  const float* wup;
  float* gC;   // out: direct part of dL/d(down-pass level s) [Pf][NP]
  float* gFc;  // out: dL/d(filtered level s+1) [Pc][NP]
  int NP;
};

__global__ void __launch_bounds__(kBlock) k_up_adj(UpAdjArgs a) {
  const int nq = a.NP >> 2;
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const long long pc_ = t / nq;
  if (pc_ >= (long long)a.hin * a.win) return;
  const int pc = (int)pc_, q4 = 4 * (int)(t - pc_ * nq);
  const int x = pc % a.win, y = pc / a.win;
  {
    const int x2 = 2 * x, y2 = 2 * y;
    const int xpu = min(x2 + 1, a.wio - 1), ypu = min(y2 + 1, a.hio - 1);
    const bool doit[4] = {true, y2 != ypu, x2 != xpu, y2 != ypu && x2 != xpu};
    const int fine[4] = {y2 * a.wio + x2, ypu * a.wio + x2, y2 * a.wio + xpu, ypu * a.wio + xpu};
    const float* W = a.wup + (size_t)pc * kWup;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      if (!doit[s]) continue;
      const float4 g = ld4(a.gF + (size_t)fine[s] * a.NP + q4);
      const float w = W[25 + s];
      st4(a.gC + (size_t)fine[s] * a.NP + q4, make_float4(w * g.x, w * g.y, w * g.z, w * g.w));
    }
  }
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int dy = -1; dy <= 1; ++dy) {
    const int oy = y + dy;
    if (oy < 0 || oy >= a.hin) continue;
    const bool mr[3] = {max(oy - 1, 0) == y, oy == y, min(oy + 1, a.hin - 1) == y};
    const int oy2 = 2 * oy, ypu = min(oy2 + 1, a.hio - 1);
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx) {
      const int ox = x + dx;
      if (ox < 0 || ox >= a.win) continue;
      const bool mc[3] = {max(ox - 1, 0) == x, ox == x, min(ox + 1, a.win - 1) == x};
      const int ox2 = 2 * ox, xpu = min(ox2 + 1, a.wio - 1);
      const bool doit[4] = {true, oy2 != ypu, ox2 != xpu, oy2 != ypu && ox2 != xpu};
      const int fine[4] = {oy2 * a.wio + ox2, ypu * a.wio + ox2, oy2 * a.wio + xpu, ypu * a.wio + xpu};
      const float* W = a.wup + (size_t)(oy * a.win + ox) * kWup;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        float weff = 0.0f;
#pragma unroll
        for (int k = kStencilBegin[s]; k < kStencilBegin[s + 1]; ++k) {
          const int code = up_tap_code(k);
          if (mr[code >> 2] && mc[code & 3]) weff += W[k];
        }
        if (doit[s] && weff != 0.0f) acc = fma4(weff, ld4(a.gF + (size_t)fine[s] * a.NP + q4), acc);
      }
    }
  }
  st4(a.gFc + (size_t)pc * a.NP + q4, acc);
}

// Down-pass adjoint from coarse level s into fine level s-1 (in place on the
// fine level, which already holds its direct part): gC_{s-1}[f] +=
// sum over the coarse pixels o whose 3x3 stencil reads f of wdn[o][.] * gC_s[o].
struct DownAdjArgs {
  const float* gCc;  // dL/d(down-pass level s), complete [Pc][NP]
  int hout, wout;
  const float* wdn;
  float* gCf;  // dL/d(down-pass level s-1), in/out [Pf][NP]
  int hin, win;
  int NP;
};

__global__ void __launch_bounds__(kBlock) k_down_adj(DownAdjArgs a) {
  const int nq = a.NP >> 2;
  const long long t = (long long)blockIdx.x * kBlock + threadIdx.x;
  const long long pf_ = t / nq;
  if (pf_ >= (long long)a.hin * a.win) return;
  const int pf = (int)pf_, q4 = 4 * (int)(t - pf_ * nq);
  const int fx = pf % a.win, fy = pf / a.win;
  float4 acc = ld4(a.gCf + (size_t)pf * a.NP + q4);
  const int oy0 = max((fy - 1) >> 1, 0), oy1 = min((fy + 1) >> 1, a.hout - 1);
  const int ox0 = max((fx - 1) >> 1, 0), ox1 = min((fx + 1) >> 1, a.wout - 1);
  for (int oy = oy0; oy <= oy1; ++oy) {
    const int y2 = 2 * oy;
    const bool mym = max(y2 - 1, 0) == fy, my2 = y2 == fy, myp = min(y2 + 1, a.hin - 1) == fy;
    for (int ox = ox0; ox <= ox1; ++ox) {
      const int x2 = 2 * ox;
      const bool mxm = max(x2 - 1, 0) == fx, mx2 = x2 == fx, mxp = min(x2 + 1, a.win - 1) == fx;
      const float* W = a.wdn + (size_t)(oy * a.wout + ox) * kWdn;
      float weff = 0.0f;
      if (my2 && mx2) weff += W[0];
      if (mym && mxm) weff += W[1];
      if (my2 && mxm) weff += W[2];
      if (myp && mxm) weff += W[3];
      if (myp && mx2) weff += W[4];
      if (myp && mxp) weff += W[5];
      if (my2 && mxp) weff += W[6];
      if (mym && mxp) weff += W[7];
      if (mym && mx2) weff += W[8];
      if (weff != 0.0f) acc = fma4(weff, ld4(a.gCc + (size_t)(oy * a.wout + ox) * a.NP + q4), acc);
    }
  }
  st4(a.gCf + (size_t)pf * a.NP + q4, acc);
}

// ---------------------------------------------------------------------------
// Sweep adjoint
// ---------------------------------------------------------------------------
// Per camera, the gradient record is 18 floats: fs0, fs1, cs0, cs1, xi,
// alpha, then the 3 x 4 (camera <- reference) transform row-major.  Record C
// holds the reference camera's parameters through its rays (first 6).
constexpr int kRec = 18;

struct SweepBwdArgs {
  const float4* raylut;
  const float* img[SS_MAX_CAMERAS];
  float* gimg[SS_MAX_CAMERAS];  // nullable entries: dL/dimage, accumulated
  TorchCam cam[SS_MAX_CAMERAS];
  float rt[SS_MAX_CAMERAS][12];
  float cands[SS_MAX_CANDIDATES];
  const float* gcost;  // dL/d(clamped cost) [P][NP]
  float* partial;      // [gridDim.x][C + 1][kRec]
  int C, ref, H, W, N, NP;
  float Wf, Hf, invW, invH, inv2W, inv2H, halfW, halfH;
  int pow2, fastproj;
};

// first-order dual number over the 6 parameters of one camera
struct D6 {
  float v, d[6];
};
__device__ __forceinline__ D6 dconst(float v) {
  D6 r;
  r.v = v;
#pragma unroll
  for (int i = 0; i < 6; ++i) r.d[i] = 0.0f;
  return r;
}
__device__ __forceinline__ D6 dvar(float v, int i) {
  D6 r = dconst(v);
  r.d[i] = 1.0f;
  return r;
}
__device__ __forceinline__ D6 operator+(const D6& a, const D6& b) {
  D6 r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < 6; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
__device__ __forceinline__ D6 operator-(const D6& a, const D6& b) {
  D6 r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < 6; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
__device__ __forceinline__ D6 operator*(const D6& a, const D6& b) {
  D6 r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < 6; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
__device__ __forceinline__ D6 operator/(const D6& a, const D6& b) {
  D6 r;
  r.v = a.v / b.v;
#pragma unroll
  for (int i = 0; i < 6; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
  return r;
}
__device__ __forceinline__ D6 dsqrt(const D6& a) {
  D6 r;
  r.v = sqrtf(a.v);
  const float h = r.v > 0.0f ? 0.5f / r.v : 0.0f;
#pragma unroll
  for (int i = 0; i < 6; ++i) r.d[i] = a.d[i] * h;
  return r;
}

// utils.py:57-73 over duals of (fs0, fs1, cs0, cs1, xi, alpha): the unit ray
// of pixel (u, v) and its derivatives.  The clamp at 0 inside the square root
// has zero derivative where it acts.
__device__ void unproject_dual(const TorchCam& c, float u, float v, D6 p[3]) {
  const D6 fs0 = dvar(c.fs0, 0), fs1 = dvar(c.fs1, 1), cs0 = dvar(c.cs0, 2), cs1 = dvar(c.cs1, 3);
  const D6 xi = dvar(c.xi, 4), al = dvar(c.a, 5), one = dconst(1.0f);
  const D6 mx = (dconst(u) - cs0) / fs0, my = (dconst(v) - cs1) / fs1;
  const D6 r2 = mx * mx + my * my;
  D6 t = one - (dconst(2.0f) * al - one) * r2;
  if (!(t.v > 0.0f)) t = dconst(0.0f);
  const D6 mz = (one - al * al * r2) / ((al * dsqrt(t) + one) - al);
  const D6 mz2 = mz * mz;
  const D6 k = (mz * xi + dsqrt(mz2 + (one - xi * xi) * r2)) / (mz2 + r2);
  p[0] = k * mx;
  p[1] = k * my;
  p[2] = k * mz - xi;
}

// Image gradients of a block's samples are first summed in an LDS window of
// the sampled image: per chunk of kBwdChunk candidates, a box of the
// sampled image around the taps of the block's lanes that sample its window
// camera (the smallest camera index among them) is found from three of the
// chunk's candidates (first, middle, last; the samples of one pixel move
// along a smooth epipolar curve), widened by kBwdMargin texels.  When it
// holds at most kWinTexels texels, taps inside it are added there (LDS
// atomics) and the window is then added to the image with one global atomic
// per touched texel and channel.  The samples of 16 consecutive candidates
// of a 16 x 16 pixel tile sweep a patch of ~1000 texels (median box ~1700
// texels at 768^2, 48 candidates), against ~16000 tap updates: ~16x fewer
// global atomics.  (Tiles of 256 pixels of one row map to tilted curves
// whose boxes mostly do not fit.)  A chunk whose box exceeds the window is
// halved (16 -> 8 -> 4 candidates) until it fits: at 768^2 and 48
// candidates, 10 % of the 16-candidate chunks did not fit, and their 13.6 %
// of the taps, sent to global atomics, took ~60 % of the kernel's time.
// Other lanes' taps, taps outside the box and chunks whose box does not fit
// even at 4 candidates use global atomics directly, so the box only decides
// speed, never the sums.
//
// The window sums in fixed point: per chunk, a power-of-two scale 2^(18-e)
// from the block's largest |dL/dcost| over the chunk, gmax < 2^e, bounds
// every contribution by 2^18 and a texel's sum of at most 4096 of them
// (kBlock pixels x kBwdChunk candidates) by 2^30, so the sums are exact
// int32 LDS atomics (order-independent, and measured ~30 % faster than
// float LDS atomics); each contribution is rounded to 2^(e-18), a relative
// 2^-19 of gmax.
constexpr int kBwdChunk = 16;
constexpr int kBwdMinChunk = 4;  // chunks stay multiples of 4 candidates (the 16-byte dL/dcost loads)
constexpr int kBwdTile = 16;      // kBwdTile^2 == kBlock pixels per workgroup
constexpr int kWinTexels = 3072;  // 36 KiB of LDS: four workgroups per CU
constexpr int kBwdMargin = 2;
static_assert(kBwdTile * kBwdTile == kBlock, "one lane per pixel of a tile");
static_assert(kBwdChunk % 4 == 0 && kBwdMinChunk % 4 == 0 && kBwdChunk % kBwdMinChunk == 0 &&
                  (kBwdChunk / kBwdMinChunk & (kBwdChunk / kBwdMinChunk - 1)) == 0,
              "chunks start at a multiple of 4 candidates");

// The sample of candidate n of a pixel with ray rl in camera (cam, rt): the
// point in the camera, project_t's intermediates, the grid_sample
// coordinates (the forward's bits), which of its four taps lie in the image
// (bits 0-3: nw, ne, sw, se; bit 4: any tap can touch the image) and the raw
// texels of its taps, loaded from clamped addresses so that no load waits
// behind a branch (the caller zeroes the taps outside the image).
struct BwdSample {
  float qx, qy, qz, d1, cc, d2, nrm, ix, iy;
  int x0, y0;
  unsigned taps;
  float vv[4][3];
};

__device__ __forceinline__ void bwd_fetch(const SweepBwdArgs& a, const TorchCam& cam, const float* rt,
                                          const float4& rl, const float* img, float d, BwdSample& s) {
  rigid(rt, d * rl.x, d * rl.y, d * rl.z, s.qx, s.qy, s.qz);
  float u, v;
  project_t_parts(cam, s.qx, s.qy, s.qz, a.fastproj != 0, u, v, s.d1, s.cc, s.d2, s.nrm);
  if (a.pow2) {
    s.ix = grid_coord_pow2(u, a.inv2W, a.halfW);
    s.iy = grid_coord_pow2(v, a.inv2H, a.halfH);
  } else {
    s.ix = grid_coord2<false>(u, a.Wf, a.invW);
    s.iy = grid_coord2<false>(v, a.Hf, a.invH);
  }
  const float x0f = floorf(s.ix), y0f = floorf(s.iy);
  const bool touch = x0f >= -1.0f && x0f <= (float)(a.W - 1) && y0f >= -1.0f && y0f <= (float)(a.H - 1);
  s.x0 = (int)(touch ? x0f : 0.0f);
  s.y0 = (int)(touch ? y0f : 0.0f);
  const bool bx0 = s.x0 >= 0, bx1 = s.x0 + 1 < a.W, by0 = s.y0 >= 0, by1 = s.y0 + 1 < a.H;
  s.taps = touch ? ((bx0 && by0 ? 1u : 0u) | (bx1 && by0 ? 2u : 0u) | (bx0 && by1 ? 4u : 0u) |
                    (bx1 && by1 ? 8u : 0u) | 16u)
                 : 0u;
  const int xa = bx0 ? s.x0 : 0, xb = bx1 ? s.x0 + 1 : a.W - 1;
  const int ya = by0 ? s.y0 : 0, yb = by1 ? s.y0 + 1 : a.H - 1;
  const float* const p[4] = {img + ((size_t)ya * a.W + xa) * 3, img + ((size_t)ya * a.W + xb) * 3,
                             img + ((size_t)yb * a.W + xa) * 3, img + ((size_t)yb * a.W + xb) * 3};
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int c = 0; c < 3; ++c) s.vv[t][c] = p[t][c];
}

// The grid_sample coordinates of candidate n (the box pass).
__device__ __forceinline__ bool bwd_sample_coords(const SweepBwdArgs& a, const TorchCam& cam, const float* rt,
                                                  const float4& rl, int n, float& ix, float& iy) {
  const float d = a.cands[n];
  float qx, qy, qz, u, v;
  bool valid;
  rigid(rt, d * rl.x, d * rl.y, d * rl.z, qx, qy, qz);
  if (!a.fastproj || !project_t_fast(cam, qx, qy, qz, u, v)) project_t(cam, qx, qy, qz, u, v, valid);
  if (a.pow2) {
    ix = grid_coord_pow2(u, a.inv2W, a.halfW);
    iy = grid_coord_pow2(v, a.inv2H, a.halfH);
  } else {
    ix = grid_coord2<false>(u, a.Wf, a.invW);
    iy = grid_coord2<false>(v, a.Hf, a.invH);
  }
  const float x0f = floorf(ix), y0f = floorf(iy);
  return x0f >= -1.0f && x0f <= (float)(a.W - 1) && y0f >= -1.0f && y0f <= (float)(a.H - 1);
}

__device__ __forceinline__ int block_min(int v, int* slot) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMin(slot, v);
  return v;
}
__device__ __forceinline__ int block_max(int v, int* slot) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(slot, v);
  return v;
}

__global__ void __launch_bounds__(kBlock) k_sweep_bwd(SweepBwdArgs a) {
  __shared__ float win[kWinTexels * 3];
  __shared__ int wbox[6];  // window camera, min x, max x, min y, max y, gmax bits
  // a workgroup owns a kBwdTile x kBwdTile tile of reference pixels: its
  // samples of one candidate fall on a compact patch of the sampled image
  const int tiles_x = (a.W + kBwdTile - 1) / kBwdTile;
  const int ty = (int)blockIdx.x / tiles_x, tx = (int)blockIdx.x - ty * tiles_x;
  const int px = tx * kBwdTile + (int)(threadIdx.x % kBwdTile), py = ty * kBwdTile + (int)(threadIdx.x / kBwdTile);
  const bool active = px < a.W && py < a.H;
  const int p = active ? py * a.W + px : 0;
  const float4 rl = a.raylut[p];
  const int k = active && (int)rl.w < a.C ? (int)rl.w : -1;
  const float* rimg = a.img[a.ref] + 3 * (size_t)p;
  const float r[3] = {rimg[0], rimg[1], rimg[2]};
  float acc[kRec];
#pragma unroll
  for (int i = 0; i < kRec; ++i) acc[i] = 0.0f;
  float gray[3] = {0.0f, 0.0f, 0.0f}, gref[3] = {0.0f, 0.0f, 0.0f};
  const int kk = k < 0 ? 0 : k;
  const TorchCam cam = a.cam[kk];
  float rt[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) rt[i] = a.rt[kk][i];
  const float* img = a.img[kk];
  float* gimg = k >= 0 ? a.gimg[kk] : nullptr;
  const float* gc = a.gcost + (size_t)p * a.NP;

  // the block's window camera (block-uniform; -1: none)
  if (threadIdx.x == 0) wbox[0] = 0x7fffffff;
  __syncthreads();
  block_min(k >= 0 && gimg ? k : 0x7fffffff, &wbox[0]);
  __syncthreads();
  const int kw = wbox[0] == 0x7fffffff ? -1 : wbox[0];
  const bool inwin = kw >= 0 && k == kw;  // this lane's taps may go to the window
  float* const gwin = kw >= 0 ? a.gimg[kw] : nullptr;

  for (int n0 = 0, len = kBwdChunk; n0 < a.N;) {
    // the chunk [n0, n1): len candidates, halved down to kBwdMinChunk while
    // the window box of its samples exceeds kWinTexels (the next chunk starts
    // at twice the length that fitted)
    int n1, wx0, wy0, ww, wh;
    float qscale, qinv;
    bool use_win;
    for (;;) {
      n1 = min(n0 + len, a.N);
      // the window box from the chunk's first, middle and last candidates
      int bx0 = 0x7fffffff, bx1 = -1, by0 = 0x7fffffff, by1 = -1;
      if (inwin) {
        const int probe[3] = {n0, (n0 + n1 - 1) / 2, n1 - 1};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          float ix, iy;
          if (!bwd_sample_coords(a, cam, rt, rl, probe[i], ix, iy)) continue;
          const int x0 = (int)floorf(ix), y0 = (int)floorf(iy);
          bx0 = min(bx0, max(x0 - kBwdMargin, 0));
          bx1 = max(bx1, min(x0 + 1 + kBwdMargin, a.W - 1));
          by0 = min(by0, max(y0 - kBwdMargin, 0));
          by1 = max(by1, min(y0 + 1 + kBwdMargin, a.H - 1));
        }
      }
      // the largest |dL/dcost| of the chunk over the lanes that add to the window
      float gl = 0.0f;
      if (inwin && active) {
        for (int n = n0; n < n1; n += 4) {
          const float4 g4 = ld4(gc + n);
          const float v[4] = {g4.x, g4.y, g4.z, g4.w};
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (n + i < n1) gl = fmaxf(gl, fabsf(v[i]));
        }
      }
      if (threadIdx.x == 0) {
        wbox[1] = 0x7fffffff;
        wbox[2] = -1;
        wbox[3] = 0x7fffffff;
        wbox[4] = -1;
        wbox[5] = 0;
      }
      __syncthreads();
      block_min(bx0, &wbox[1]);
      block_max(bx1, &wbox[2]);
      block_min(by0, &wbox[3]);
      block_max(by1, &wbox[4]);
      block_max(__float_as_int(gl), &wbox[5]);  // non-negative floats order as their bits
      __syncthreads();
      int gexp = 0;
      (void)frexpf(__int_as_float(wbox[5]), &gexp);
      gexp = max(gexp, -100);
      qscale = ldexpf(1.0f, 18 - gexp);
      qinv = ldexpf(1.0f, gexp - 18);
      wx0 = wbox[1];
      wy0 = wbox[3];
      ww = wbox[2] - wx0 + 1;
      wh = wbox[4] - wy0 + 1;
      const bool boxed = kw >= 0 && wbox[2] >= 0 && wbox[4] >= 0;
      use_win = boxed && ww * wh <= kWinTexels;  // block-uniform, as is everything below
      if (use_win || !boxed || len <= kBwdMinChunk) break;
      len /= 2;
      __syncthreads();  // every lane has read the box before it is reset
    }
    if (use_win) {
      for (int i = (int)threadIdx.x; i < ww * wh * 3; i += kBlock) win[i] = 0.0f;
      __syncthreads();
    }
    // the chunk's candidates in a register (lane i: candidate n0 + i mod 16),
    // read back with readlane: a scalar load in the loop would wait for the
    // LDS atomics before it (they share one counter).  The empty asm keeps
    // the load here, where every lane runs it: sunk into the branch below,
    // lanes that skip the branch would leave their copy unset.
    float cand_v = a.cands[min(n0 + (int)(threadIdx.x & (kBwdChunk - 1)), a.N - 1)];
    asm volatile("" : "+v"(cand_v));
    auto cand = [&](int n) __attribute__((always_inline)) {
      return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cand_v), n - n0));
    };

    if (active && k < 0) {
      // no camera sampled: the cost is |0 - r| summed (depth_estimation.py:150-152)
      for (int n = n0; n < n1; ++n) {
        const float g = gc[n];
        const float S = (fabsf(-r[0]) + fabsf(-r[1])) + fabsf(-r[2]);
        const float gS = S <= 500.0f ? g : 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) gref[c] -= (-r[c]) > 0.0f ? gS : ((-r[c]) < 0.0f ? -gS : 0.0f);
      }
    } else if (active) {
      // candidate n + 1's projection and texel loads are issued before
      // candidate n is used, so their latency overlaps n's arithmetic
      BwdSample cur, nxt;
      bwd_fetch(a, cam, rt, rl, img, cand(n0), cur);
      float4 g4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      for (int n = n0; n < n1; ++n) {
        if (n + 1 < n1) bwd_fetch(a, cam, rt, rl, img, cand(n + 1), nxt);
        // dL/dcost four candidates at a time (rows are NP >= 8 floats, 32-B aligned)
        const int nq = n & 3;
        if (nq == 0) g4 = ld4(gc + n);
        const float g = nq == 0 ? g4.x : (nq == 1 ? g4.y : (nq == 2 ? g4.z : g4.w));
        // zero-padded bilinear sample (the forward's taps and order) and its
        // derivatives in ix, iy
        float sv[3] = {0.0f, 0.0f, 0.0f}, sx[3] = {0.0f, 0.0f, 0.0f}, sy[3] = {0.0f, 0.0f, 0.0f};
        float wx0f = 0.0f, wx1f = 0.0f, wy0f = 0.0f, wy1f = 0.0f;
        const bool touch = (cur.taps & 16u) != 0u;
        if (touch) {
          const float x0f = (float)cur.x0, y0f = (float)cur.y0;
          wx1f = cur.ix - x0f;
          wx0f = (x0f + 1.0f) - cur.ix;
          wy1f = cur.iy - y0f;
          wy0f = (y0f + 1.0f) - cur.iy;
          float vv[4][3];
#pragma unroll
          for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int c = 0; c < 3; ++c) vv[t][c] = (cur.taps >> t) & 1u ? cur.vv[t][c] : 0.0f;
          const float wnw = wx0f * wy0f, wne = wx1f * wy0f, wsw = wx0f * wy1f, wse = wx1f * wy1f;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            float tt = vv[0][c] * wnw;
            tt = tt + vv[1][c] * wne;
            tt = tt + vv[2][c] * wsw;
            tt = tt + vv[3][c] * wse;
            sv[c] = tt;
            sx[c] = wy0f * (vv[1][c] - vv[0][c]) + wy1f * (vv[3][c] - vv[2][c]);
            sy[c] = wx0f * (vv[2][c] - vv[0][c]) + wx1f * (vv[3][c] - vv[1][c]);
          }
        }
        const float S = (fabsf(sv[0] - r[0]) + fabsf(sv[1] - r[1])) + fabsf(sv[2] - r[2]);
        // torch.clamp(max=500) passes the gradient where S <= 500
        const float gS = S <= 500.0f ? g : 0.0f;
        float gs[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float dlt = sv[c] - r[c];
          gs[c] = dlt > 0.0f ? gS : (dlt < 0.0f ? -gS : 0.0f);
          gref[c] -= gs[c];
        }
        // a sample with no tap in the image has no derivative in the
        // sampled image nor in its coordinates
        if (touch && gS != 0.0f) {
          if (gimg) {
            const float wt[4] = {wx0f * wy0f, wx1f * wy0f, wx0f * wy1f, wx1f * wy1f};
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              if (!((cur.taps >> t) & 1u)) continue;
              const int tx = cur.x0 + (t & 1), ty = cur.y0 + (t >> 1);
              const int lx = tx - wx0, ly = ty - wy0;
              if (use_win && inwin && lx >= 0 && lx < ww && ly >= 0 && ly < wh) {
                int* dst = reinterpret_cast<int*>(win) + (ly * ww + lx) * 3;
#pragma unroll
                for (int c = 0; c < 3; ++c) atomicAdd(dst + c, __float2int_rn(wt[t] * gs[c] * qscale));
              } else {
                float* dst = gimg + ((size_t)ty * a.W + tx) * 3;
#pragma unroll
                for (int c = 0; c < 3; ++c) unsafeAtomicAdd(dst + c, wt[t] * gs[c]);
              }
            }
          }
          // the grid_sample coordinate round trip is the identity in u: dix/du = 1
          const float gu = (gs[0] * sx[0] + gs[1] * sx[1]) + gs[2] * sx[2];
          const float gv = (gs[0] * sy[0] + gs[1] * sy[1]) + gs[2] * sy[2];
          // utils.py:75-94 backwards, on the forward's intermediates; the
          // reciprocals of the derivative need no correct rounding
          const float qx = cur.qx, qy = cur.qy, qz = cur.qz, d1 = cur.d1, cc = cur.cc, d2 = cur.d2;
          const float inv = __builtin_amdgcn_rcpf(cur.nrm);
          acc[0] += gu * qx * inv;
          acc[1] += gv * qy * inv;
          acc[2] += gu;
          acc[3] += gv;
          float gqx = gu * cam.fs0 * inv, gqy = gv * cam.fs1 * inv, gqz = 0.0f;
          const float gnrm = -(gu * cam.fs0 * qx + gv * cam.fs1 * qy) * inv * inv;
          acc[5] += gnrm * (d2 - cc);
          const float gd2 = gnrm * cam.a;
          float gcc = gnrm * cam.one_m_a;
          float gxy2 = 0.0f;
          if (d2 > 0.0f) {
            const float r2 = __builtin_amdgcn_rcpf(d2);
            gxy2 += gd2 * 0.5f * r2;
            gcc += gd2 * cc * r2;
          }
          acc[4] += gcc * d1;
          const float gd1 = gcc * cam.xi;
          gqz += gcc;
          if (d1 > 0.0f) {
            const float r1 = __builtin_amdgcn_rcpf(d1);
            gxy2 += gd1 * 0.5f * r1;
            gqz += gd1 * qz * r1;
          }
          gqx += 2.0f * qx * gxy2;
          gqy += 2.0f * qy * gxy2;
          const float d = cand(n);
          const float gq[3] = {gqx, gqy, gqz}, Pv[3] = {d * rl.x, d * rl.y, d * rl.z};
#pragma unroll
          for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int i = 0; i < 3; ++i) acc[6 + 4 * j + i] += gq[j] * Pv[i];
            acc[6 + 4 * j + 3] += gq[j];
          }
#pragma unroll
          for (int i = 0; i < 3; ++i) gray[i] += d * ((rt[i] * gq[0] + rt[4 + i] * gq[1]) + rt[8 + i] * gq[2]);
        }
        cur = nxt;
      }
    }

    if (use_win) {
      // the window into the image: one global atomic per touched texel and channel
      __syncthreads();
      for (int i = (int)threadIdx.x; i < ww * wh * 3; i += kBlock) {
        const int q = reinterpret_cast<const int*>(win)[i];
        if (q != 0) {
          const int t = i / 3, c = i - 3 * t;
          const int ly = t / ww, lx = t - ly * ww;
          unsafeAtomicAdd(gwin + ((size_t)(wy0 + ly) * a.W + wx0 + lx) * 3 + c, (float)q * qinv);
        }
      }
    }
    __syncthreads();  // the window and its box are reused by the next chunk
    n0 = n1;
    len = min(2 * len, kBwdChunk);
  }
  if (active && a.gimg[a.ref]) {
    float* dst = a.gimg[a.ref] + 3 * (size_t)p;
#pragma unroll
    for (int c = 0; c < 3; ++c) unsafeAtomicAdd(dst + c, gref[c]);
  }
  // the reference camera's parameters through its rays
  float gint[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (active && k >= 0) {
    D6 ray[3];
    unproject_dual(a.cam[a.ref], (float)(p % a.W), (float)(p / a.W), ray);
#pragma unroll
    for (int i = 0; i < 6; ++i) gint[i] = (gray[0] * ray[0].d[i] + gray[1] * ray[1].d[i]) + gray[2] * ray[2].d[i];
  }
  // per-block sums, one record per camera (fixed order: deterministic)
  __shared__ float red[kWaves][kRec];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float* out = a.partial + (size_t)blockIdx.x * (a.C + 1) * kRec;
  for (int kc = 0; kc <= a.C; ++kc) {
    const bool mine = kc < a.C ? k == kc : true;
    const bool any = __any(mine);  // wavefront-uniform
#pragma unroll
    for (int i = 0; i < kRec; ++i) {
      float v = 0.0f;
      if (any) {
        v = kc < a.C ? (mine ? acc[i] : 0.0f) : (i < 6 ? gint[i] : 0.0f);
        v = wave_sum(v);
      }
      if (lane == 0) red[wv][i] = v;
    }
    __syncthreads();
    if (threadIdx.x < kRec) {
      float s = 0.0f;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) s += red[w][threadIdx.x];
      out[kc * kRec + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

// Column sums of nb block records of `width` values: one workgroup per
// column, each lane a fixed stride of records, then a fixed-order tree, so
// the result does not depend on scheduling.  T is the accumulation type.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sum_records(const float* partial, int nb, int width, T* out) {
  __shared__ T red[kBlock];
  const int col = blockIdx.x, t = threadIdx.x;
  T s = 0;
  for (int b = t; b < nb; b += kBlock) s += (T)partial[(size_t)b * width + col];
  red[t] = s;
  __syncthreads();
#pragma unroll
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  if (t == 0) out[col] = red[0];
}

// ---------------------------------------------------------------------------
// Photometric Gauss-Newton: residuals of reference pixels reprojected into a
// target camera at their estimated inverse distance, linearised in 8
// parameters of the target: a right perturbation exp(delta) of its pose
// rt_k (rotation omega, translation v) and relative focal scales
// (fs0 *= 1 + dfx, fs1 *= 1 + dfy).  With inv(rt_k exp(delta)) =
// exp(-delta) inv(rt_k), a point q in the target frame moves by
// dq = q x omega - v.
// ---------------------------------------------------------------------------
constexpr int kGN = 8;
constexpr int kNE = kGN * (kGN + 1) / 2 + kGN + 2;  // J^T J (upper), J^T r, loss, count
struct PhotoArgs {
  const float4* raylut;
  const float* invd;
  const float* ref_img;
  const float* img;
  const float* mask;  // nullable: target validity mask (>= 0.9 kept)
  TorchCam cam;
  float rt[12];
  int H, W, target, skip_selected;
  float huber;
  float* partial;  // [gridDim.x][kNE]
};

__global__ void __launch_bounds__(kBlock) k_photo_normal(PhotoArgs a) {
  const int P = a.H * a.W;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  float e[kNE];
#pragma unroll
  for (int i = 0; i < kNE; ++i) e[i] = 0.0f;
  bool use = p < P;
  float4 rl = make_float4(0.0f, 0.0f, 1.0f, -1.0f);
  float id = 0.0f;
  if (use) {
    rl = a.raylut[p];
    id = a.invd[p];
    const int sel = (int)rl.w;
    use = id > 0.0f && sel >= 0 && !(a.skip_selected && sel == a.target);
  }
  if (use) {
    const float X[3] = {rl.x / id, rl.y / id, rl.z / id};
    float q[3];
    rigid(a.rt, X[0], X[1], X[2], q[0], q[1], q[2]);
    const TorchCam& c = a.cam;
    const float xy2 = q[0] * q[0] + q[1] * q[1];
    const float d1 = sqrtf(xy2 + q[2] * q[2]);
    const float cc = c.xi * d1 + q[2];
    const float d2 = sqrtf(xy2 + cc * cc);
    const float nrm = c.a * d2 + c.one_m_a * cc;
    const bool front = q[2] > c.mw2 * d1 && nrm > 0.0f && d1 > 0.0f && d2 > 0.0f;
    const float inv = 1.0f / nrm;
    const float u = c.fs0 * q[0] * inv + c.cs0, v = c.fs1 * q[1] * inv + c.cs1;
    // interior cells only, so that every tap and its derivative are defined
    const float x0f = floorf(u), y0f = floorf(v);
    use = front && x0f >= 0.0f && y0f >= 0.0f && x0f + 1.0f <= (float)(a.W - 1) && y0f + 1.0f <= (float)(a.H - 1);
    if (use) {
      const int x0 = (int)x0f, y0 = (int)y0f;
      const float fx = u - x0f, fy = v - y0f;
      const size_t i00 = (size_t)y0 * a.W + x0;
      if (a.mask) {
        const float m = (1.0f - fy) * ((1.0f - fx) * a.mask[i00] + fx * a.mask[i00 + 1]) +
                        fy * ((1.0f - fx) * a.mask[i00 + a.W] + fx * a.mask[i00 + a.W + 1]);
        use = m >= 0.9f;
      }
      if (use) {
        // d(u, v)/dq of the double-sphere projection
        const float dd1[3] = {q[0] / d1, q[1] / d1, q[2] / d1};
        float dcc[3] = {c.xi * dd1[0], c.xi * dd1[1], c.xi * dd1[2] + 1.0f};
        float dn[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const float dd2 = ((j < 2 ? q[j] : 0.0f) + cc * dcc[j]) / d2;
          dn[j] = c.a * dd2 + c.one_m_a * dcc[j];
        }
        float Ju[kGN], Jv[kGN];
        float du[3], dv[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          du[j] = c.fs0 * ((j == 0 ? 1.0f : 0.0f) - q[0] * inv * dn[j]) * inv;
          dv[j] = c.fs1 * ((j == 1 ? 1.0f : 0.0f) - q[1] * inv * dn[j]) * inv;
        }
        // dq/domega = [q]_x (dq = q x omega), dq/dv = -I
        const float qx[3][3] = {{0.0f, -q[2], q[1]}, {q[2], 0.0f, -q[0]}, {-q[1], q[0], 0.0f}};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          Ju[i] = du[0] * qx[0][i] + du[1] * qx[1][i] + du[2] * qx[2][i];
          Jv[i] = dv[0] * qx[0][i] + dv[1] * qx[1][i] + dv[2] * qx[2][i];
          Ju[3 + i] = -du[i];
          Jv[3 + i] = -dv[i];
        }
        Ju[6] = u - c.cs0;
        Jv[6] = 0.0f;
        Ju[7] = 0.0f;
        Jv[7] = v - c.cs1;
        const float* t00 = a.img + i00 * 3;
        const float* t01 = t00 + 3;
        const float* t10 = t00 + (size_t)a.W * 3;
        const float* t11 = t10 + 3;
        const float* rp = a.ref_img + (size_t)p * 3;
        float loss = 0.0f;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          const float I = (1.0f - fy) * ((1.0f - fx) * t00[ch] + fx * t01[ch]) + fy * ((1.0f - fx) * t10[ch] + fx * t11[ch]);
          const float Iu = (1.0f - fy) * (t01[ch] - t00[ch]) + fy * (t11[ch] - t10[ch]);
          const float Iv = (1.0f - fx) * (t10[ch] - t00[ch]) + fx * (t11[ch] - t01[ch]);
          const float res = I - rp[ch];
          const float ar = fabsf(res);
          const float w = ar <= a.huber ? 1.0f : a.huber / ar;
          loss += ar <= a.huber ? 0.5f * res * res : a.huber * (ar - 0.5f * a.huber);
          float J[kGN];
#pragma unroll
          for (int i = 0; i < kGN; ++i) J[i] = Iu * Ju[i] + Iv * Jv[i];
          int o = 0;
#pragma unroll
          for (int i = 0; i < kGN; ++i)
#pragma unroll
            for (int j = i; j < kGN; ++j) e[o++] += w * J[i] * J[j];
#pragma unroll
          for (int i = 0; i < kGN; ++i) e[o + i] += w * J[i] * res;
        }
        e[kNE - 2] = loss;
        e[kNE - 1] = 1.0f;
      }
    }
  }
  __shared__ float red[kWaves][kNE];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < kNE; ++i) {
    const float v = wave_sum(e[i]);
    if (lane == 0) red[wv][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < kNE) {
    float s = 0.0f;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s += red[w][threadIdx.x];
    a.partial[(size_t)blockIdx.x * kNE + threadIdx.x] = s;
  }
}

// The normal-equation solve, in fp64 (one lane): the summed records (fp64),
// a damped diagonal (Levenberg-Marquardt, H + lambda * diag(H)), Cholesky,
// delta = -H^-1 b, then the loss and the residual count.
__global__ void __launch_bounds__(64) k_gn_solve(const double* sum, double lambda, double* out) {
  if (threadIdx.x != 0) return;
  double H[kGN][kGN], b[kGN];
  int o = 0;
  for (int i = 0; i < kGN; ++i)
    for (int j = i; j < kGN; ++j) {
      H[i][j] = sum[o];
      H[j][i] = sum[o];
      ++o;
    }
  for (int i = 0; i < kGN; ++i) b[i] = sum[o + i];
  for (int i = 0; i < kGN; ++i) H[i][i] = H[i][i] * (1.0 + lambda) + 1e-9;
  // Cholesky H = L L^T in place (lower triangle)
  bool ok = true;
  for (int j = 0; j < kGN && ok; ++j) {
    double d = H[j][j];
    for (int k = 0; k < j; ++k) d -= H[j][k] * H[j][k];
    if (!(d > 0.0)) {
      ok = false;
      break;
    }
    const double l = sqrt(d);
    H[j][j] = l;
    for (int i = j + 1; i < kGN; ++i) {
      double s = H[i][j];
      for (int k = 0; k < j; ++k) s -= H[i][k] * H[j][k];
      H[i][j] = s / l;
    }
  }
  double x[kGN];
  if (ok) {
    for (int i = 0; i < kGN; ++i) {
      double s = -b[i];
      for (int k = 0; k < i; ++k) s -= H[i][k] * x[k];
      x[i] = s / H[i][i];
    }
    for (int i = kGN - 1; i >= 0; --i) {
      double s = x[i];
      for (int k = i + 1; k < kGN; ++k) s -= H[k][i] * x[k];
      x[i] = s / H[i][i];
    }
  } else {
    for (int i = 0; i < kGN; ++i) x[i] = 0.0;
  }
  for (int i = 0; i < kGN; ++i) out[i] = x[i];
  out[kGN] = sum[kNE - 2];
  out[kGN + 1] = sum[kNE - 1];
  out[kGN + 2] = ok ? 1.0 : 0.0;
}

inline int blocks_of(long long n) { return (int)((n + kBlock - 1) / kBlock); }
// k_sweep_bwd's grid: one workgroup per kBwdTile x kBwdTile tile
inline int sweep_bwd_blocks(int H, int W) {
  return ((H + kBwdTile - 1) / kBwdTile) * ((W + kBwdTile - 1) / kBwdTile);
}

int levels_of(int H, int W) {
  if (H < 1 || W < 1) return -1;
  const double l = fmin(log2((double)W), log2((double)H));
  return (int)(l - 1);
}

struct GradLevel {
  int h, w;
  uint32_t* g;
  float *wdn, *wup, *gF, *gC;
};

// 256-byte aligned carving of the backward workspace (nullptr base: sizing)
struct Carve {
  char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t n) {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += n * sizeof(T);
    return p;
  }
};

size_t carve_grad_levels(void* ws, int H, int W, int NP, int L, GradLevel* lv) {
  Carve cv{static_cast<char*>(ws)};
  lv[0].h = H;
  lv[0].w = W;
  lv[0].g = nullptr;
  lv[0].wdn = lv[0].gF = lv[0].gC = nullptr;
  lv[0].wup = nullptr;
  for (int s = 1; s < L; ++s) {
    lv[s].h = (int)ceil(H / pow(2.0, s));
    lv[s].w = (int)ceil(W / pow(2.0, s));
    const size_t P = (size_t)lv[s].h * lv[s].w;
    lv[s].g = cv.take<uint32_t>(P);
    lv[s].wdn = cv.take<float>(P * kWdn);
    lv[s].wup = cv.take<float>(P * kWup);
    lv[s].gF = cv.take<float>(P * NP);
    lv[s].gC = cv.take<float>(P * NP);
  }
  return cv.off;
}

int padded_np(int N) {
  const int v = lanes_v(N);
  return v < 0 ? -1 : 8 * v;
}

int fail_arg(const char* m) { return report_error(SS_ERR_ARG, m); }

#define GR_HIP(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) {                                                           \
      char m_[256];                                                                   \
      snprintf(m_, sizeof m_, "%s: %s", #expr, hipGetErrorString(e_));                \
      return report_error(SS_ERR_HIP, m_);                                            \
    }                                                                                 \
  } while (0)

inline hipStream_t as_stream(ss_stream s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

}  // namespace ss

using namespace ss;

extern "C" {

int ss_soft_argmin(const float* d_cost, int H, int W, int N, float beta, float inv0, float inv1, float* d_invd,
                   ss_stream stream) {
  if (!d_cost || !d_invd) return fail_arg("null pointer argument");
  if (H < 1 || W < 1 || N < 2 || N > SS_MAX_CANDIDATES) return fail_arg("sizes out of range (N >= 2)");
  const long long P = (long long)H * W;
  if (P >= (1ll << 31)) return fail_arg("image too large");
  SoftArgs a{d_cost, (int)P, N, padded_np(N), beta, inv0, inv1 - inv0, nullptr, d_invd};
  hipLaunchKernelGGL(k_soft_argmin, dim3(blocks_of(P * kSoftGroup)), dim3(kBlock), 0, as_stream(stream), a);
  GR_HIP(hipGetLastError());
  return SS_OK;
}

int ss_soft_argmin_backward(const float* d_cost, const float* d_grad_invd, int H, int W, int N, float beta,
                            float inv0, float inv1, float* d_grad_cost, ss_stream stream) {
  if (!d_cost || !d_grad_invd || !d_grad_cost) return fail_arg("null pointer argument");
  if (H < 1 || W < 1 || N < 2 || N > SS_MAX_CANDIDATES) return fail_arg("sizes out of range (N >= 2)");
  const long long P = (long long)H * W;
  if (P >= (1ll << 31)) return fail_arg("image too large");
  SoftArgs a{d_cost, (int)P, N, padded_np(N), beta, inv0, inv1 - inv0, d_grad_invd, d_grad_cost};
  hipLaunchKernelGGL(k_soft_argmin_bwd, dim3(blocks_of(P * kSoftGroup)), dim3(kBlock), 0, as_stream(stream), a);
  GR_HIP(hipGetLastError());
  return SS_OK;
}

size_t ss_isb_backward_workspace_bytes(int H, int W, int N) {
  const int L = levels_of(H, W), NP = padded_np(N);
  if (L < 1 || NP < 8 || (long long)H * W >= (1ll << 31)) return 0;
  GradLevel lv[32];
  if (L > 32) return 0;
  return carve_grad_levels(nullptr, H, W, NP, L, lv) + 256;
}

int ss_isb_filter_backward(const uint32_t* d_guide, const float* d_grad_out, float* d_grad_in, int H, int W, int N,
                           double sigma_i, double sigma_s, void* d_workspace, size_t workspace_bytes,
                           ss_stream stream) {
  if (!d_guide || !d_grad_out || !d_grad_in || !d_workspace) return fail_arg("null pointer argument");
  if (d_grad_out == d_grad_in) return fail_arg("d_grad_out and d_grad_in must not alias");
  const int L = levels_of(H, W), NP = padded_np(N);
  if (L < 1 || L > 32 || N < 2 || NP < 8 || (long long)H * W >= (1ll << 31))
    return fail_arg("sizes out of range (N >= 2, at least one pyramid level)");
  const size_t need = ss_isb_backward_workspace_bytes(H, W, N);
  if (workspace_bytes < need) return report_error(SS_ERR_WORKSPACE, "workspace smaller than required");
  hipStream_t s = as_stream(stream);
  GradLevel lv[32];
  carve_grad_levels(d_workspace, H, W, NP, L, lv);
  const size_t P0 = (size_t)H * W;
  if (L == 1) {
    GR_HIP(hipMemcpyAsync(d_grad_in, d_grad_out, P0 * NP * sizeof(float), hipMemcpyDeviceToDevice, s));
    return SS_OK;
  }
  const double var_inv_s = 1 / (2 * sigma_s * sigma_s);
  const float var_inv_i = (float)(1 / (2 * sigma_i * sigma_i));
  auto gl = [&](int lvl) -> const uint32_t* { return lvl == 0 ? d_guide : lv[lvl].g; };
  // guide pyramid and down-pass weights (isb_filter.py:87-104)
  for (int l = 1; l < L; ++l) {
    GDownArgs a{gl(l - 1), lv[l - 1].h, lv[l - 1].w, lv[l].g, lv[l].h, lv[l].w, var_inv_i, lv[l].wdn};
    hipLaunchKernelGGL(k_gdown, dim3(blocks_of((long long)lv[l].h * lv[l].w)), dim3(kBlock), 0, s, a);
  }
  // guide updates and up-pass weights (isb_filter.py:106-134), coarse to fine
  for (int l = L - 2; l >= 0; --l) {
    const double dist = pow(2.0, l) - 0.5;
    const double wdn = exp(-(dist * dist) * var_inv_s);
    GUpArgs a{lv[l + 1].g, lv[l + 1].h, lv[l + 1].w, l == 0 ? nullptr : lv[l].g, lv[l].h, lv[l].w,
              (float)(1 - wdn), (float)wdn, var_inv_i, lv[l + 1].wup, l != 0};
    if (l == 0) a.gio = const_cast<uint32_t*>(d_guide);  // read only: write_guide = 0
    hipLaunchKernelGGL(k_gup, dim3(blocks_of((long long)lv[l + 1].h * lv[l + 1].w)), dim3(kBlock), 0, s, a);
  }
  // transposed up passes, fine to coarse
  for (int l = 0; l <= L - 2; ++l) {
    const float* gF = l == 0 ? d_grad_out : lv[l].gF;
    float* gC = l == 0 ? d_grad_in : lv[l].gC;
    UpAdjArgs a{gF, lv[l].h, lv[l].w, lv[l + 1].h, lv[l + 1].w, lv[l + 1].wup, gC, lv[l + 1].gF, NP};
    const long long n = (long long)lv[l + 1].h * lv[l + 1].w * (NP / 4);
    hipLaunchKernelGGL(k_up_adj, dim3(blocks_of(n)), dim3(kBlock), 0, s, a);
  }
  // transposed down passes, coarse to fine; the top level's cost gradient is
  // its filtered-level gradient
  for (int l = L - 1; l >= 1; --l) {
    const float* gCc = l == L - 1 ? lv[l].gF : lv[l].gC;
    float* gCf = l - 1 == 0 ? d_grad_in : lv[l - 1].gC;
    DownAdjArgs a{gCc, lv[l].h, lv[l].w, lv[l].wdn, gCf, lv[l - 1].h, lv[l - 1].w, NP};
    const long long n = (long long)lv[l - 1].h * lv[l - 1].w * (NP / 4);
    hipLaunchKernelGGL(k_down_adj, dim3(blocks_of(n)), dim3(kBlock), 0, s, a);
  }
  GR_HIP(hipGetLastError());
  return SS_OK;
}

size_t ss_sweep_backward_workspace_bytes(int H, int W, int C) {
  if (H < 1 || W < 1 || C < 1 || C > SS_MAX_CAMERAS || (long long)H * W >= (1ll << 31)) return 0;
  return (size_t)sweep_bwd_blocks(H, W) * (C + 1) * kRec * sizeof(float);
}

int ss_sweep_backward(const float* d_raylut, const float* const* d_images, int C, const ss_camera* cams,
                      const float* rts, int ref, const float* cands, int N, int H, int W, const float* d_grad_cost,
                      float* const* d_grad_images, float* d_grad_params, void* d_workspace, size_t workspace_bytes,
                      ss_stream stream) {
  if (!d_raylut || !d_images || !cams || !rts || !cands || !d_grad_cost || !d_grad_params || !d_workspace)
    return fail_arg("null pointer argument");
  if (C < 1 || C > SS_MAX_CAMERAS || ref < 0 || ref >= C) return fail_arg("camera count or reference out of range");
  if (N < 2 || N > SS_MAX_CANDIDATES || H < 1 || W < 1 || (long long)H * W >= (1ll << 31))
    return fail_arg("sizes out of range (N >= 2)");
  if (workspace_bytes < ss_sweep_backward_workspace_bytes(H, W, C))
    return report_error(SS_ERR_WORKSPACE, "workspace smaller than required");
  SweepBwdArgs a;
  memset(&a, 0, sizeof a);
  a.raylut = reinterpret_cast<const float4*>(d_raylut);
  for (int k = 0; k < C; ++k) {
    if (!d_images[k]) return fail_arg("null image pointer");
    a.img[k] = d_images[k];
    a.gimg[k] = d_grad_images ? d_grad_images[k] : nullptr;
    a.cam[k] = make_torch_cam(cams[k]);
    memcpy(a.rt[k], rts + 12 * k, sizeof(float) * 12);
  }
  memcpy(a.cands, cands, sizeof(float) * N);
  a.gcost = d_grad_cost;
  a.partial = static_cast<float*>(d_workspace);
  a.C = C;
  a.ref = ref;
  a.H = H;
  a.W = W;
  a.N = N;
  a.NP = padded_np(N);
  a.Wf = (float)W;
  a.Hf = (float)H;
  a.invW = 1.0f / (float)W;
  a.invH = 1.0f / (float)H;
  a.inv2W = 2.0f / (float)W;
  a.inv2H = 2.0f / (float)H;
  a.halfW = (float)W * 0.5f;
  a.halfH = (float)H * 0.5f;
  a.pow2 = (W & (W - 1)) == 0 && (H & (H - 1)) == 0;
  // the forward's choice of projection form (the same bounds as the sweep)
  bool fast = true;
  auto in = [](float v, float lo, float hi) { return std::isfinite(v) && fabsf(v) >= lo && fabsf(v) <= hi; };
  const float big = 0x1p20f;
  for (int k = 0; k < C; ++k) {
    const TorchCam& c = a.cam[k];
    if (!in(c.fs0, 0.0f, big) || !in(c.fs1, 0.0f, big) || !in(c.cs0, 0x1p-16f, big) || !in(c.cs1, 0x1p-16f, big) ||
        !in(c.xi, 0.0f, 16.0f) || !in(c.a, 0.0f, 16.0f) || !in(c.one_m_a, 0.0f, 16.0f))
      fast = false;
    for (int i = 0; i < 12; ++i)
      if (!in(a.rt[k][i], 0.0f, big)) fast = false;
  }
  for (int n = 0; n < N; ++n)
    if (!in(a.cands[n], 0.0f, big)) fast = false;
  a.fastproj = fast;
  hipStream_t s = as_stream(stream);
  const int nb = sweep_bwd_blocks(H, W);
  hipLaunchKernelGGL(k_sweep_bwd, dim3(nb), dim3(kBlock), 0, s, a);
  const int width = (C + 1) * kRec;
  hipLaunchKernelGGL(k_sum_records<float>, dim3(width), dim3(kBlock), 0, s, a.partial, nb, width, d_grad_params);
  GR_HIP(hipGetLastError());
  return SS_OK;
}

int ss_photometric_blocks(int H, int W) {
  if (H < 1 || W < 1 || (long long)H * W >= (1ll << 31)) return -1;
  return blocks_of((long long)H * W);
}

int ss_photometric_normal_equations(const float* d_raylut, const float* d_invd, const float* d_ref_image,
                                    const float* d_image, const float* d_mask, const ss_camera* cam, const float* rt,
                                    int target, int skip_selected, int H, int W, float huber, float* d_partial,
                                    ss_stream stream) {
  if (!d_raylut || !d_invd || !d_ref_image || !d_image || !cam || !rt || !d_partial)
    return fail_arg("null pointer argument");
  if (H < 2 || W < 2 || (long long)H * W >= (1ll << 31)) return fail_arg("sizes out of range");
  if (!(huber > 0.0f)) return fail_arg("huber threshold must be positive");
  PhotoArgs a;
  memset(&a, 0, sizeof a);
  a.raylut = reinterpret_cast<const float4*>(d_raylut);
  a.invd = d_invd;
  a.ref_img = d_ref_image;
  a.img = d_image;
  a.mask = d_mask;
  a.cam = make_torch_cam(*cam);
  memcpy(a.rt, rt, sizeof(float) * 12);
  a.H = H;
  a.W = W;
  a.target = target;
  a.skip_selected = skip_selected;
  a.huber = huber;
  a.partial = d_partial;
  hipLaunchKernelGGL(k_photo_normal, dim3(blocks_of((long long)H * W)), dim3(kBlock), 0, as_stream(stream), a);
  GR_HIP(hipGetLastError());
  return SS_OK;
}

int ss_photometric_record_floats(void) { return kNE; }

int ss_gauss_newton_solve(const float* d_partial, int nblocks, double lambda, double* d_out, ss_stream stream) {
  if (!d_partial || !d_out || nblocks < 1) return fail_arg("null pointer or no records");
  // the sums go to d_out[16..16+kNE) (the caller's buffer holds 64 doubles)
  double* sums = d_out + 16;
  hipLaunchKernelGGL(k_sum_records<double>, dim3(kNE), dim3(kBlock), 0, as_stream(stream), d_partial, nblocks, kNE,
                     sums);
  hipLaunchKernelGGL(k_gn_solve, dim3(1), dim3(64), 0, as_stream(stream), sums, lambda, d_out);
  GR_HIP(hipGetLastError());
  return SS_OK;
}

}  // extern "C"
