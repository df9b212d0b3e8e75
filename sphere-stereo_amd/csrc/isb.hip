// Multiscale edge-preserving ("ISB") cost-volume filter on gfx950.
//
// Behaviour: isb_filter.cu:44-330 and isb_filter.py:72-136 of the reference.
// Design: the volume is candidate-contiguous ([h][w][NP], NP = G*V).  A group
// of G lanes owns one coarse pixel; lane j holds candidates [j*V, j*V+V) and
// moves them with 16-byte loads, so a wavefront streams 8 whole pixels per
// instruction.  The guide weights of a pixel are computed once, spread over
// the G lanes (lane j evaluates weight j) and broadcast with wavefront
// shuffles.  G = 1 is the scalar form used for the N = 1 distance filter.
// The last up-pass can be fused with the argmin/refinement so that the
// filtered level-0 volume is never written back to HBM.
//
// Element types: the level-0 volume (the sweep output, most of the traffic)
// may be fp32, fp16 or bf16 (T below); the coarser pyramid levels are always
// fp32.  With sigma_s = 25 the level-0 up pass takes ~99.98 % of its weight
// from level 1, so rounding the coarse levels to 16 bits moved the argmin of
// several per cent of the pixels, while rounding level 0 changed next to
// nothing.
#include <string.h>

#include <type_traits>

#include "ss_isb_stencil.h"
#include "ss_kernels.h"
#include "ss_ragged.h"

namespace ss {

namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Down-sampling: guideDownsample2xKernel (isb_filter.cu:44-119)
// ---------------------------------------------------------------------------
// One coarse pixel of the down pass, handled by the G lanes of a group.
// Inactive groups (past the end of the level) compute on a valid pixel so
// that the shuffles see every lane, and store nothing.  T is the storage
// type of the input level (level 0 may be 16-bit); the output level is fp32,
// as are all sums and weights.
template <int G, int V, typename T>
__device__ __forceinline__ void down_pixel(const IsbDownArgs& a, int pix, bool active, int lane) {
  constexpr int NP = G * V;
  constexpr int CH = chunk_of<V, T>();
  const T* cin = static_cast<const T*>(a.cin);
  float* cout = static_cast<float*>(a.cout);

  const int x = pix % a.wout, y = pix / a.wout;
  const int x2 = 2 * x, y2 = 2 * y;
  const int xp = min(x2 + 1, a.win - 1), yp = min(y2 + 1, a.hin - 1);
  const int xm = max(x2 - 1, 0), ym = max(y2 - 1, 0);
  // neighbour order of isb_filter.cu:68-77
  const int nb[8] = {ym * a.win + xm, y2 * a.win + xm, yp * a.win + xm, yp * a.win + x2,
                     yp * a.win + xp, y2 * a.win + xp, ym * a.win + xp, ym * a.win + x2};
  const int ctr = y2 * a.win + x2;
  const uint32_t gc = a.gin[ctr];

  // neighbour j's guide and weight factor (isb_filter.cu:80-93)
  auto nguide = [&](int j) __attribute__((always_inline)) -> uint32_t {
    const int row = pick3((j == 0 || j == 6 || j == 7) ? 0 : ((j == 1 || j == 5) ? 1 : 2), ym, y2, yp);
    const int col = pick3((j <= 2) ? 0 : ((j == 3 || j == 7) ? 1 : 2), xm, x2, xp);
    return a.gin[row * a.win + col];
  };
  auto weight = [&](int j, uint32_t gn) __attribute__((always_inline)) -> float {
    const float base = (j & 1) ? 0.5f : 0.25f;
    return base * (a.wtab[guide_sad_u(gc, gn)] + 0.01f);
  };

  // G > 1: lane l evaluates neighbours l, l + G, ... and the group shares
  // the weights; glv[i] keeps neighbour l + i G's guide for the guide sums
  constexpr int PER = G == 1 ? 1 : 8 / G;
  float w[8];
  uint32_t glv[PER];
  if constexpr (G == 1) {
#pragma unroll
    for (int j = 0; j < 8; ++j) w[j] = weight(j, nguide(j));
  } else {
    static_assert(G == 4 || G == 8, "groups of 4 or 8 lanes");
    float wl[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      glv[i] = nguide(lane + i * G);
      wl[i] = weight(lane + i * G, glv[i]);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) w[j] = __shfl(wl[j / G], j % G, G);
  }
  float ws = 1.0f;
#pragma unroll
  for (int j = 0; j < 8; ++j) ws = ws + w[j];

  // the cost pass first: its loads do not wait behind the guide store
  const size_t base_c = (size_t)lane * V;
#pragma unroll
  for (int c0 = 0; c0 < V; c0 += CH) {
    float acc[CH], t[CH];
    load_n<CH>(cin + (size_t)ctr * NP + base_c + c0, acc);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      load_n<CH>(cin + (size_t)nb[j] * NP + base_c + c0, t);
#pragma unroll
      for (int v = 0; v < CH; ++v) acc[v] = acc[v] + t[v] * w[j];
    }
#pragma unroll
    for (int v = 0; v < CH; ++v) acc[v] = acc[v] / ws;
    if (active) store_n<CH>(cout + (size_t)pix * NP + base_c + c0, acc);
  }

  // guide: each channel's weighted sum in the reference's tap order
  // 1 <= ws <= 1 + 8 * 0.5 * 1.01: rcp_rn is exactly 1.0f / ws
  const float inv = rcp_rn(ws);
  if constexpr (G == 1) {
    if (active) {
      float s0 = gch(gc, 0), s1 = gch(gc, 1), s2 = gch(gc, 2);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t gn = a.gin[nb[j]];
        s0 = s0 + gch(gn, 0) * w[j];
        s1 = s1 + gch(gn, 1) * w[j];
        s2 = s2 + gch(gn, 2) * w[j];
      }
      a.gout[pix] = pack_guide(s0 * inv, s1 * inv, s2 * inv);
    }
  } else {
    // lane c of the group sums channel c (lanes 3..7 repeat channel 0); the
    // neighbour guides come from the lanes that loaded them for the weights
    const int ch = lane < 3 ? lane : 0;
    float sc = gch(gc, ch);
#pragma unroll
    for (int j = 0; j < 8; ++j) sc = sc + gch(__shfl(glv[j / G], j % G, G), ch) * w[j];
    const uint32_t byte = (uint32_t)(uint8_t)(sc * inv);
    const uint32_t b1 = __shfl(byte, 1, G), b2 = __shfl(byte, 2, G);
    if (lane == 0 && active) a.gout[pix] = byte | (b1 << 8) | (b2 << 16);
  }
}

// xcd: the XCD-aware tile order (xcd_tile), else tiles in launch order
template <int G, int V, typename T>
__global__ void __launch_bounds__(kBlock) k_isb_down(Batch<IsbDownArgs> b, int xcd) {
  const IsbDownArgs a = b.a[blockIdx.y];
  const int lane = (G == 1) ? 0 : (int)(threadIdx.x & (G - 1));
  const int tile = xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  int pix = (tile * kBlock + (int)threadIdx.x) / G;
  const int npix = a.hout * a.wout;
  const bool active = pix < npix;
  if (G == 1 && !active) return;
  if (!active) pix = npix - 1;
  down_pixel<G, V, T>(a, pix, active, lane);
}

// ---------------------------------------------------------------------------
// Argmin + parabola refinement + index->distance for one pixel whose NP
// candidates are spread over a G-lane group (depth_estimation.py:180-198).
// ---------------------------------------------------------------------------
// The group part: first minimum (index, value), maximum, and the values at
// the clamped neighbours of the minimum, in every lane of the group.
struct ArgminGroup {
  float m, M, L, R;
  int im;
};

template <int G, int V>
__device__ __forceinline__ ArgminGroup argmin_reduce_group(const float* vals, int lane, const ArgminParams& p) {
  const int n0 = lane * V;
  float m = __builtin_huge_valf(), M = -__builtin_huge_valf();
  int im = 0x7fffffff;
#pragma unroll
  for (int v = 0; v < V; ++v) {
    if (n0 + v < p.N) {
      if (vals[v] < m) {
        m = vals[v];
        im = n0 + v;
      }
      M = fmaxf(M, vals[v]);
    }
  }
  if constexpr (G > 1) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
      const float om = __shfl_xor(m, off, G);
      const int oi = __shfl_xor(im, off, G);
      const float oM = __shfl_xor(M, off, G);
      if (om < m || (om == m && oi < im)) {
        m = om;
        im = oi;
      }
      M = fmaxf(M, oM);
    }
  }
  const int il = max(im - 1, 0), ir = min(im + 1, p.N - 1);
  float L, R;
  {
    // pick element (il % V) / (ir % V) from the owning lane
    float sl = vals[0], sr = vals[0];
#pragma unroll
    for (int v = 1; v < V; ++v) {
      sl = (il % V == v) ? vals[v] : sl;
      sr = (ir % V == v) ? vals[v] : sr;
    }
    if constexpr (G > 1) {
      L = __shfl(sl, il / V, G);
      R = __shfl(sr, ir / V, G);
    } else {
      L = sl;
      R = sr;
    }
  }
  return ArgminGroup{m, M, L, R, im};
}

// The per-pixel part: parabola refinement and index -> distance.
__device__ __forceinline__ float argmin_finish(const ArgminGroup& g, const ArgminParams& p) {
  float var = (0.5f * (g.L - g.R)) / (((g.L + g.R) - 2.0f * g.m) + 1e-8f);
  var = var < -0.5f ? -0.5f : (var > 0.5f ? 0.5f : var);
  if (g.im == p.N - 1 || g.im == 0) var = 0.0f;
  float idx = (float)g.im + var;
  if (g.M == g.m) idx = (float)(p.N - 1);
  float d = p.d0 / (((p.A * idx) / (float)(p.N - 1)) + 1.0f);
  if (fabsf(g.M - g.m) < 1e-8f) d = p.dN1;
  return d;
}

template <int G, int V>
__device__ __forceinline__ float argmin_refine_group(const float* vals, int lane, const ArgminParams& p) {
  return argmin_finish(argmin_reduce_group<G, V>(vals, lane, p), p);
}

// ---------------------------------------------------------------------------
// The neighbour lane's value (0 <-> 1, 2 <-> 3 of each quad of four lanes):
// a DPP move, one VALU instruction and no LDS round trip.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int quad_swap(int v) {
  constexpr int kCtrl = 1 | (0 << 2) | (3 << 4) | (2 << 6);
  return __builtin_amdgcn_update_dpp(0, v, kCtrl, 0xF, 0xF, false);
}
__device__ __forceinline__ float quad_swap(float v) { return __int_as_float(quad_swap(__float_as_int(v))); }

// The fused argmin's form: through LDS with groups of 4 or 8 lanes and at
// most 8 candidates per lane (16 or 32 KB of LDS per workgroup), in
// registers otherwise.  Each fine pixel's filtered values go to LDS as soon
// as they are computed, so the pass holds fewer registers than the register
// form (78 against 90 VGPRs at 8 lanes x 4 candidates, fp32) and still
// issues every volume load before the weights are computed; it also takes
// fewer VALU instructions (807 against 937 per wave).
template <int G, int V, typename T>
constexpr bool lds_argmin() {
  return G >= 4 && V <= 8;
}

// ---------------------------------------------------------------------------
// Up-sampling: guideUpsample2xKernel (isb_filter.cu:132-330).  FUSE_ARGMIN
// replaces the level-0 write-back with the argmin of argmin_refine_group.
// ---------------------------------------------------------------------------
// One coarse pixel of the up pass (its four fine pixels), handled by the G
// lanes of a group; inactive groups as in down_pixel.
template <int G, int V, typename T, bool FUSE_ARGMIN, bool REPROJ = false>
__device__ __forceinline__ void up_pixel(const IsbUpArgs& a, int pix, bool active, int lane) {
  static_assert(!REPROJ || (G == 1 && V == 1 && !FUSE_ARGMIN), "reprojection: the one-lane distance filter");
  constexpr int NP = G * V;
  constexpr int CH = chunk_of<V, T>();
  const float* cin = static_cast<const float*>(a.cin);  // coarse level: fp32
  T* cio = static_cast<T*>(a.cio);                      // fine level: T

  const int x = pix % a.win, y = pix / a.win;
  const int r0 = max(y - 1, 0), r2 = min(y + 1, a.hin - 1);
  const int c0 = max(x - 1, 0), c2 = min(x + 1, a.win - 1);
  const int x2 = 2 * x, y2 = 2 * y;
  const int xpu = min(x2 + 1, a.wio - 1), ypu = min(y2 + 1, a.hio - 1);
  const bool setY = y2 != ypu, setX = x2 != xpu;
  const bool doit[4] = {true, setY, setX, setY && setX};
  const int fine[4] = {y2 * a.wio + x2, ypu * a.wio + x2, y2 * a.wio + xpu, ypu * a.wio + xpu};
  const int rows[3] = {r0, y, r2}, cols[3] = {c0, x, c2};
  int tap[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) tap[r * 3 + c] = rows[r] * a.win + cols[c];

  uint32_t gF[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) gF[s] = a.gio[fine[s]];

  auto weight = [&](int k) __attribute__((always_inline)) -> float {
    const int code = up_tap_code(k);
    const int rr = code >> 2, cc = code & 3;
    const int row = pick3(rr, r0, y, r2);
    const int col = pick3(cc, c0, x, c2);
    const int s = up_stencil(k);
    const uint32_t gf = pick4(s, gF[0], gF[1], gF[2], gF[3]);
    return up_base(k) * (a.weight_down * (a.wtab[guide_sad_u(a.gin[row * a.win + col], gf)] + 0.01f));
  };

  float w[25];
  if constexpr (G == 1) {
#pragma unroll
    for (int k = 0; k < 25; ++k) w[k] = weight(k);
  } else {
    // lane l evaluates taps l, l + G, ... of the first 24 and the group
    // shares them; every lane evaluates the last
    static_assert(G == 4 || G == 8, "groups of 4 or 8 lanes");
    float wv[24 / G];
#pragma unroll
    for (int i = 0; i < 24 / G; ++i) wv[i] = weight(lane + i * G);
#pragma unroll
    for (int k = 0; k < 24; ++k) w[k] = __shfl(wv[k / G], k % G, G);
    w[24] = weight(24);
  }

  float inv[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    float ws = a.weight_up;
#pragma unroll
    for (int k = kStencilBegin[s]; k < kStencilBegin[s + 1]; ++k) ws = ws + w[k];
    // ws = weight_up + weight_down * (sum of base * (table + 0.01)) with
    // weight_up + weight_down = 1 and at least 4 base weights >= 1, so
    // 0.04 <= ws <= 1 + 16 * 1.01: rcp_rn is exactly 1.0f / ws
    inv[s] = rcp_rn(ws);
  }

  const size_t base_c = (size_t)lane * V;
  constexpr bool kLdsArgmin = FUSE_ARGMIN && lds_argmin<G, V, T>();
  // the fused argmin's inputs: fine pixel s's candidates of this lane, in
  // registers, or written to the group's LDS rows as soon as they are done
  float res[FUSE_ARGMIN && !kLdsArgmin ? 4 : 1][FUSE_ARGMIN && !kLdsArgmin ? V : 1];
  constexpr int kAbufGroups = kLdsArgmin ? kBlock / G : 1, kAbufNP = kLdsArgmin ? NP : 1;
  __shared__ __align__(16) float abuf[kAbufGroups][4][kAbufNP];
  float* const buf = &abuf[kLdsArgmin ? threadIdx.x / G : 0][0][0];
#pragma unroll
  for (int cb = 0; cb < V; cb += CH) {
    float cD[9][CH];
#pragma unroll
    for (int t = 0; t < 9; ++t) load_n<CH>(cin + (size_t)tap[t] * NP + base_c + cb, cD[t]);
    // every fine pixel is computed (at the last row or column two of them
    // coincide, as the clamped indices make them) and only stores are
    // guarded: straight-line code, no branch per fine pixel
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      float o[CH];
      load_n<CH>(cio + (size_t)fine[s] * NP + base_c + cb, o);
#pragma unroll
      for (int v = 0; v < CH; ++v) {
        float acc = a.weight_up * o[v];
#pragma unroll
        for (int k = kStencilBegin[s]; k < kStencilBegin[s + 1]; ++k) acc = acc + w[k] * cD[up_tap_flat(k)][v];
        o[v] = acc * inv[s];
      }
      if constexpr (kLdsArgmin) {
        store_n<CH>(buf + s * NP + lane * V + cb, o);
      } else if constexpr (FUSE_ARGMIN) {
#pragma unroll
        for (int v = 0; v < CH; ++v) res[s][cb + v] = o[v];
      } else {
        if (active && doit[s]) {
          store_n<CH>(cio + (size_t)fine[s] * NP + base_c + cb, o);
          if constexpr (REPROJ) reproject_splat(a.rp_c, a.rp_t, a.rp_out, a.hio, a.wio, fine[s], o[0]);
        }
      }
    }
  }

  // guide (not needed after the level-0 pass, whose guide is a copy at the
  // reference's call sites, depth_estimation.py:177): each fine pixel's
  // channel sums in the reference's tap order, after the cost pass so that
  // its loads do not wait behind these stores
  if constexpr (!FUSE_ARGMIN) {
    if (a.skip_guide) {
      // launch-uniform: nothing reads this fine guide afterwards
    } else if constexpr (G == 1) {
      if (active) {
        uint32_t gD[9];
#pragma unroll
        for (int t = 0; t < 9; ++t) gD[t] = a.gin[tap[t]];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          float s0 = gch(gF[s], 0) * a.weight_up, s1 = gch(gF[s], 1) * a.weight_up,
                s2 = gch(gF[s], 2) * a.weight_up;
#pragma unroll
          for (int k = kStencilBegin[s]; k < kStencilBegin[s + 1]; ++k) {
            const uint32_t g = gD[up_tap_flat(k)];
            s0 = s0 + gch(g, 0) * w[k];
            s1 = s1 + gch(g, 1) * w[k];
            s2 = s2 + gch(g, 2) * w[k];
          }
          if (doit[s]) a.gio[fine[s]] = pack_guide(s0 * inv[s], s1 * inv[s], s2 * inv[s]);
        }
      }
    } else {
      // lane c of the group sums channel c of the four fine pixels (lanes
      // 3..7 repeat channel 0); byte s of `mine` is fine pixel s's channel
      const int ch = lane < 3 ? lane : 0;
      float gD[9];
#pragma unroll
      for (int t = 0; t < 9; ++t) gD[t] = gch(a.gin[tap[t]], ch);
      uint32_t mine = 0;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        float acc = gch(gF[s], ch) * a.weight_up;
#pragma unroll
        for (int k = kStencilBegin[s]; k < kStencilBegin[s + 1]; ++k) acc = acc + gD[up_tap_flat(k)] * w[k];
        mine |= (uint32_t)(uint8_t)(acc * inv[s]) << (8 * s);
      }
      const uint32_t m1 = __shfl(mine, 1, G), m2 = __shfl(mine, 2, G);
      if (lane == 0 && active) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
          if (doit[s])
            a.gio[fine[s]] = ((mine >> (8 * s)) & 0xffu) | (((m1 >> (8 * s)) & 0xffu) << 8) |
                             (((m2 >> (8 * s)) & 0xffu) << 16);
      }
    }
  }

  if constexpr (FUSE_ARGMIN) {
    // (fine[] are the fine pixels' indices in the level-0 distance map too)
    if constexpr (lds_argmin<G, V, T>()) {
      // The group's filtered values go to LDS (fine pixel major, candidate
      // contiguous); G / 4 lanes then scan each fine pixel's candidates, a
      // contiguous part each, merge their partial results with a DPP move,
      // and read the refinement's neighbours back from LDS.  First minimum,
      // maximum and neighbours are those of argmin_reduce_group.
      constexpr int LPP = G / 4, SPAN = NP / LPP;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int s = lane / LPP, n0 = (lane % LPP) * SPAN, N = a.am.N;
      const float* const pv = buf + s * NP;
      float vals[SPAN];
      load_n<SPAN>(pv + n0, vals);
      float m = __builtin_huge_valf(), M = -__builtin_huge_valf();
      int im = 0x7fffffff;
      if (N == NP) {
#pragma unroll
        for (int v = 0; v < SPAN; ++v) {
          if (vals[v] < m) {
            m = vals[v];
            im = n0 + v;
          }
          M = fmaxf(M, vals[v]);
        }
      } else {
#pragma unroll
        for (int v = 0; v < SPAN; ++v) {
          if (n0 + v < N) {
            if (vals[v] < m) {
              m = vals[v];
              im = n0 + v;
            }
            M = fmaxf(M, vals[v]);
          }
        }
      }
      if constexpr (LPP == 2) {
        const float om = quad_swap(m), oM = quad_swap(M);
        const int oi = quad_swap(im);
        if (om < m || (om == m && oi < im)) {
          m = om;
          im = oi;
        }
        M = fmaxf(M, oM);
      }
      const ArgminGroup g{m, M, pv[max(im - 1, 0)], pv[min(im + 1, N - 1)], im};
      const float d = argmin_finish(g, a.am);
      // the next call of this wavefront rewrites the buffer
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const bool mine_s = s == 0 ? doit[0] : (s == 1 ? doit[1] : (s == 2 ? doit[2] : doit[3]));
      const int at = s == 0 ? fine[0] : (s == 1 ? fine[1] : (s == 2 ? fine[2] : fine[3]));
      if (active && lane % LPP == 0 && mine_s) {
        a.dist[at] = d;
        if (a.fill) a.fill[at] = 1e8f;
      }
    } else {
      // the group reductions need every lane; the refinement of fine pixel s
      // (three divisions) then runs once, in lane s, instead of in all 8
      ArgminGroup mine = {};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const ArgminGroup g = argmin_reduce_group<G, V>(res[s], lane, a.am);
        if (lane == s) mine = g;
      }
      if (active && lane < 4 && doit[lane]) {
        a.dist[fine[lane]] = argmin_finish(mine, a.am);
        if (a.fill) a.fill[fine[lane]] = 1e8f;
      }
    }
  }
}

// The fused level-0 pass with the LDS argmin of up_pixel and up to 4
// candidates per lane is held to 96 VGPRs (at least 5 waves per SIMD).
template <int V, typename T, bool FUSE_ARGMIN>
constexpr int up_plain_waves() {
  return FUSE_ARGMIN && V <= 4 && lds_argmin<8, V, T>() ? 5 : 1;
}

template <int G, int V, typename T, bool FUSE_ARGMIN, bool REPROJ = false>
__global__ void __launch_bounds__(kBlock, (up_plain_waves<V, T, FUSE_ARGMIN>())) k_isb_up(Batch<IsbUpArgs> b, int xcd) {
  const IsbUpArgs a = b.a[blockIdx.y];
  const int lane = (G == 1) ? 0 : (int)(threadIdx.x & (G - 1));
  const int tile = xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  int pix = (tile * kBlock + (int)threadIdx.x) / G;
  const int npix = a.hin * a.win;
  const bool active = pix < npix;
  if (G == 1 && !active) return;
  if (!active) pix = npix - 1;
  up_pixel<G, V, T, FUSE_ARGMIN, REPROJ>(a, pix, active, lane);
}

// The coarse end of the pyramid in one workgroup per batch entry: the down passes into
// levels 1..n-1 of the table, then the up passes back into level 0, each
// separated by a workgroup barrier.  Levels 1..n-1 live only in LDS (they
// are not needed once the pass is over); level 0 is in global memory.  The
// per-pixel functions work on generic pointers, so the same code serves both.
// The tail starts at pyramid level 1 or deeper, so all its levels are fp32.
constexpr int kTailBlock = 1024;

template <int G, int V>
__device__ __forceinline__ void tail_body(const TailLevels& t, uint32_t* const g0, void* const c0, char* smem) {
  const int lane = (G == 1) ? 0 : (int)(threadIdx.x & (G - 1));
  const int slot = threadIdx.x / G, slots = kTailBlock / G;
  auto gp = [&](int k) __attribute__((always_inline)) -> uint32_t* {
    return k == 0 ? g0 : reinterpret_cast<uint32_t*>(smem + t.goff[k]);
  };
  auto cp = [&](int k) __attribute__((always_inline)) -> void* {
    return k == 0 ? c0 : static_cast<void*>(smem + t.coff[k]);
  };
  // the weight table in LDS: every pass of the tail looks weights up there
  float* const wtab = reinterpret_cast<float*>(smem + t.wtab_off);
  for (int i = (int)threadIdx.x; i < kWeightTable; i += kTailBlock) wtab[i] = t.wtab[i];
  __syncthreads();
  for (int k = 1; k < t.n; ++k) {
    const IsbDownArgs a{gp(k - 1), cp(k - 1), t.h[k - 1], t.w[k - 1], gp(k), cp(k), t.h[k], t.w[k], wtab};
    const int npix = t.h[k] * t.w[k];
    for (int base = 0; base < npix; base += slots)
      if (base + slot < npix) down_pixel<G, V, float>(a, base + slot, true, lane);
    __syncthreads();
  }
  for (int k = t.n - 2; k >= 0; --k) {
    const IsbUpArgs a{gp(k + 1), cp(k + 1), t.h[k + 1], t.w[k + 1], gp(k), cp(k), t.h[k], t.w[k],
                      t.wu[k],   t.wd[k],   wtab,   nullptr,  ArgminParams{}, nullptr};
    const int npix = t.h[k + 1] * t.w[k + 1];
    for (int base = 0; base < npix; base += slots)
      if (base + slot < npix) up_pixel<G, V, float, false>(a, base + slot, true, lane);
    __syncthreads();
  }
}

template <int G, int V>
__global__ void __launch_bounds__(kTailBlock) k_isb_tail(IsbTailArgs t) {
  extern __shared__ __align__(16) char smem[];
  tail_body<G, V>(t.t, t.g0[blockIdx.x], t.c0[blockIdx.x], smem);
}

// ---------------------------------------------------------------------------
// Ragged forms: a wavefront's item is kIsbItemPix8 (G = 8) or kIsbItemPix1
// (G = 1) consecutive coarse pixels of one entry, one step; the entry's V
// picks the instantiation.
// ---------------------------------------------------------------------------
template <int G>
constexpr int item_pix() {
  return G == 1 ? kIsbItemPix1 : kIsbItemPix8;
}

template <int G, int V, typename T>
__device__ __forceinline__ void down_item(const IsbDownArgs& a, int item, int lane, int slot) {
  const int npix = a.hout * a.wout;
  int pix = item * item_pix<G>() + slot;
  const bool active = pix < npix;
  if (!active) pix = npix - 1;
  down_pixel<G, V, T>(a, pix, active, lane);
}

template <int G, int V, typename T, bool FUSE>
__device__ __forceinline__ void up_item(const IsbUpArgs& a, int item, int lane, int slot) {
  const int npix = a.hin * a.win;
  int pix = item * item_pix<G>() + slot;
  const bool active = pix < npix;
  if (!active) pix = npix - 1;
  up_pixel<G, V, T, FUSE>(a, pix, active, lane);
}

// Calls f(integral_constant<V>) for the entry's V.  A launch takes the
// entries of one V (VMAX = V) and instantiates that V only, so that its
// register budget, hence its occupancy, is that of its own candidate count.
template <int VMAX, typename F>
__device__ __forceinline__ void with_v(int V, F&& f) {
  if (V == VMAX) f(std::integral_constant<int, VMAX>());
}

// Occupancy targets (waves per SIMD, the second __launch_bounds__ argument)
// of the ragged passes per V, each the most that V reaches without register
// spills.  The entry lookup makes the compiler's own choice lower than the
// plain kernels' for the same V.
template <int G, int V>
constexpr int down_waves() {
  return G == 1 ? 7 : (V <= 2 ? 6 : V <= 4 ? 5 : 4);
}
template <int G, int V, bool FUSE>
constexpr int up_waves() {
  return G == 1 ? 4 : (FUSE ? (V == 1 ? 5 : V == 2 ? 4 : V <= 4 ? 3 : V <= 8 ? 2 : 1) : (V <= 8 ? 3 : 2));
}

template <int G, typename T, int VMAX, bool LOOP>
__global__ void __launch_bounds__(kBlock, (down_waves<G, VMAX>())) k_isb_down_ragged(RaggedSpan<IsbDownEntry> sp) {
  const int l64 = (int)(threadIdx.x & 63);
  const int lane = G == 1 ? 0 : (l64 & (G - 1)), slot = l64 / G;
  for_each_wave_item<LOOP>(sp, [&](int e, int t) {
    const IsbDownEntry en = load_const(sp.e + e);
    const IsbDownArgs& a = en.a;
    with_v<VMAX>(en.V, [&](auto v) { down_item<G, decltype(v)::value, T>(a, t, lane, slot); });
  });
}

template <int G, typename T, bool FUSE, int VMAX, bool LOOP>
__global__ void __launch_bounds__(kBlock, (up_waves<G, VMAX, FUSE>())) k_isb_up_ragged(RaggedSpan<IsbUpEntry> sp) {
  const int l64 = (int)(threadIdx.x & 63);
  const int lane = G == 1 ? 0 : (l64 & (G - 1)), slot = l64 / G;
  for_each_wave_item<LOOP>(sp, [&](int e, int t) {
    const IsbUpEntry en = load_const(sp.e + e);
    const IsbUpArgs& a = en.a;
    with_v<VMAX>(en.V, [&](auto v) { up_item<G, decltype(v)::value, T, FUSE>(a, t, lane, slot); });
  });
}

// one workgroup per entry, as k_isb_tail
template <int G>
__global__ void __launch_bounds__(kTailBlock) k_isb_tail_ragged(const TailEntry* __restrict__ te) {
  extern __shared__ __align__(16) char smem[];
  const TailEntry& e = te[blockIdx.x];
  auto body = [&](auto v) { tail_body<G, decltype(v)::value>(e.t, e.g0, e.c0, smem); };
  if constexpr (G == 1) {
    with_v<1>(e.V, body);
  } else {
    with_v<1>(e.V, body);
    with_v<2>(e.V, body);
    with_v<3>(e.V, body);
    with_v<4>(e.V, body);
    with_v<6>(e.V, body);
    with_v<8>(e.V, body);
    with_v<12>(e.V, body);
    with_v<16>(e.V, body);
  }
}

// Weight tables: entry d of table i is the guide weight factor
// exp(-d * d * vi[i]) with the operations and exp of the passes' own
// expression (isb_filter.cu:87 and 229-248: __expf(-d*d*var_inv)), so a
// lookup returns the bits the expression would.
struct WeightTableArgs {
  float vi[kMaxBatch];
  float* tab[kMaxBatch];
};
__global__ void __launch_bounds__(kBlock) k_weight_tables(WeightTableArgs t) {
  const int d = blockIdx.x * kBlock + threadIdx.x;
  if (d >= kWeightTable) return;
  const float df = (float)d;
  t.tab[blockIdx.y][d] = expf_cw(-df * df * t.vi[blockIdx.y]);
}

// Argmin over an already filtered level-0 volume (unfused form).
template <int G, int V, typename T>
__global__ void __launch_bounds__(kBlock) k_argmin(const T* __restrict__ cost, int npix, ArgminParams p,
                                                   float* __restrict__ dist) {
  constexpr int NP = G * V;
  const int lane = (G == 1) ? 0 : (int)(threadIdx.x & (G - 1));
  int pix = (int)((blockIdx.x * kBlock + threadIdx.x) / G);
  const bool active = pix < npix;
  if (G == 1 && !active) return;
  if (!active) pix = npix - 1;
  float vals[V];
  load_n<V>(cost + (size_t)pix * NP + (size_t)lane * V, vals);
  const float d = argmin_refine_group<G, V>(vals, lane, p);
  if (active && lane == 0) dist[pix] = d;
}

// Layout changes for the planar drop-in (ISB_Filter.apply's [N][H][W]).
__global__ void k_planar_to_cc(const float* __restrict__ src, float* __restrict__ dst, int HW, int N, int NP) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= HW) return;
  for (int n = 0; n < NP; ++n) dst[(size_t)p * NP + n] = n < N ? src[(size_t)n * HW + p] : 0.0f;
}

__global__ void k_cc_to_planar(const float* __restrict__ src, float* __restrict__ dst, int HW, int N, int NP) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= HW) return;
  for (int n = 0; n < N; ++n) dst[(size_t)n * HW + p] = src[(size_t)p * NP + n];
}

__global__ void k_guide3_to_packed(const uint8_t* __restrict__ g3, uint32_t* __restrict__ g, int HW) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= HW) return;
  g[p] = (uint32_t)g3[3 * p] | ((uint32_t)g3[3 * p + 1] << 8) | ((uint32_t)g3[3 * p + 2] << 16);
}

__global__ void k_packed_to_guide3(const uint32_t* __restrict__ g, uint8_t* __restrict__ g3, int HW) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= HW) return;
  const uint32_t v = g[p];
  g3[3 * p] = (uint8_t)v;
  g3[3 * p + 1] = (uint8_t)(v >> 8);
  g3[3 * p + 2] = (uint8_t)(v >> 16);
}

inline int blocks_for(long long threads) { return (int)((threads + kBlock - 1) / kBlock); }

template <typename A>
Batch<A> batch_of(const A* a, int nb) {
  Batch<A> b;
  memset(&b, 0, sizeof b);
  for (int i = 0; i < nb; ++i) b.a[i] = a[i];
  return b;
}

// Every entry of a batch has the shape of entry 0.
template <int G, int V, typename T>
hipError_t launch_down_gv(const IsbDownArgs* a, int nb, int xcd, hipStream_t s) {
  const dim3 grid(blocks_for((long long)a[0].hout * a[0].wout * G), nb);
  hipLaunchKernelGGL((k_isb_down<G, V, T>), grid, dim3(kBlock), 0, s, batch_of(a, nb), xcd);
  return hipGetLastError();
}

template <int G, int V>
hipError_t launch_tail_gv(const IsbTailArgs& t, int nb, hipStream_t s) {
  hipLaunchKernelGGL((k_isb_tail<G, V>), dim3(nb), dim3(kTailBlock), t.lds_bytes, s, t);
  return hipGetLastError();
}

template <int G, int V, typename T>
hipError_t launch_up_gv(const IsbUpArgs* a, int nb, bool fuse, int xcd, hipStream_t s, bool reproj = false) {
  const dim3 grid(blocks_for((long long)a[0].hin * a[0].win * G), nb);
  if constexpr (G == 1 && V == 1 && std::is_same<T, float>::value) {
    if (reproj) {
      hipLaunchKernelGGL((k_isb_up<1, 1, float, false, true>), grid, dim3(kBlock), 0, s, batch_of(a, nb), xcd);
      return hipGetLastError();
    }
  }
  if (reproj) return hipErrorInvalidValue;
  if (fuse)
    hipLaunchKernelGGL((k_isb_up<G, V, T, true>), grid, dim3(kBlock), 0, s, batch_of(a, nb), xcd);
  else
    hipLaunchKernelGGL((k_isb_up<G, V, T, false>), grid, dim3(kBlock), 0, s, batch_of(a, nb), xcd);
  return hipGetLastError();
}

template <int V, typename T>
hipError_t launch_argmin_vt(const void* cost, int npix, const ArgminParams& p, float* dist, hipStream_t s) {
  hipLaunchKernelGGL((k_argmin<8, V, T>), dim3(blocks_for((long long)npix * 8)), dim3(kBlock), 0, s,
                     static_cast<const T*>(cost), npix, p, dist);
  return hipGetLastError();
}

#define SS_DISPATCH_V(V, CALL)                       \
  switch (V) {                                       \
    case 1: return CALL(1);                          \
    case 2: return CALL(2);                          \
    case 3: return CALL(3);                          \
    case 4: return CALL(4);                          \
    case 6: return CALL(6);                          \
    case 8: return CALL(8);                          \
    case 12: return CALL(12);                        \
    case 16: return CALL(16);                        \
    default: return hipErrorInvalidValue;            \
  }

// G = 4 holds twice the candidates per lane of G = 8 for the same padded
// count (NP = 8 V8 = 4 V4, V4 = 2 V8); it is built for NP <= 64.
#define SS_DISPATCH_V4(V, CALL)                      \
  switch (V) {                                       \
    case 2: return CALL(2);                          \
    case 4: return CALL(4);                          \
    case 6: return CALL(6);                          \
    case 8: return CALL(8);                          \
    case 12: return CALL(12);                        \
    case 16: return CALL(16);                        \
    default: return hipErrorInvalidValue;            \
  }

template <typename T>
hipError_t down_t(const IsbDownArgs* a, int nb, int G, int V, int xcd, hipStream_t s) {
  if (G == 4) {
#define SS_CALL(v) launch_down_gv<4, v, T>(a, nb, xcd, s)
    SS_DISPATCH_V4(V, SS_CALL)
#undef SS_CALL
  }
  if (G != 8) return hipErrorInvalidValue;
#define SS_CALL(v) launch_down_gv<8, v, T>(a, nb, xcd, s)
  SS_DISPATCH_V(V, SS_CALL)
#undef SS_CALL
}

template <typename T>
hipError_t up_t(const IsbUpArgs* a, int nb, int G, int V, bool fuse, int xcd, hipStream_t s) {
  if (G == 4) {
#define SS_CALL(v) launch_up_gv<4, v, T>(a, nb, fuse, xcd, s)
    SS_DISPATCH_V4(V, SS_CALL)
#undef SS_CALL
  }
  if (G != 8) return hipErrorInvalidValue;
#define SS_CALL(v) launch_up_gv<8, v, T>(a, nb, fuse, xcd, s)
  SS_DISPATCH_V(V, SS_CALL)
#undef SS_CALL
}

template <typename T>
hipError_t argmin_t(const void* cost, int npix, int V, const ArgminParams& p, float* dist, hipStream_t s) {
#define SS_CALL(v) launch_argmin_vt<v, T>(cost, npix, p, dist, s)
  SS_DISPATCH_V(V, SS_CALL)
#undef SS_CALL
}

}  // namespace

// vt: the element type of the level-0 side of the pass (the input of a down
// pass, the fine level of an up pass); the other side is fp32.  G = 1 (the
// distance filter) exists for fp32 only.
hipError_t launch_isb_down(const IsbDownArgs* a, int nb, int G, int V, int vt, int xcd, hipStream_t s) {
  if (nb < 1 || nb > kMaxBatch) return hipErrorInvalidValue;
  if (G == 1)
    return (V == 1 && vt == SS_VOLUME_F32) ? launch_down_gv<1, 1, float>(a, nb, xcd, s) : hipErrorInvalidValue;
#define SS_CALL(T) down_t<T>(a, nb, G, V, xcd, s)
  SS_DISPATCH_T(vt, SS_CALL)
#undef SS_CALL
}

hipError_t launch_weight_tables(const float* vi, float* const* tab, int n, hipStream_t s) {
  if (n < 1 || n > kMaxBatch) return hipErrorInvalidValue;
  WeightTableArgs t;
  memset(&t, 0, sizeof t);
  for (int i = 0; i < n; ++i) {
    t.vi[i] = vi[i];
    t.tab[i] = tab[i];
  }
  hipLaunchKernelGGL(k_weight_tables, dim3((kWeightTable + kBlock - 1) / kBlock, n), dim3(kBlock), 0, s, t);
  return hipGetLastError();
}

// The tail's levels are in LDS; it keeps groups of 8 lanes whatever the
// passes before and after it use (its levels never leave the tail).
hipError_t launch_isb_tail(const IsbTailArgs& t, int nb, int G, int V, hipStream_t s) {
  if (nb < 1 || nb > kMaxBatch) return hipErrorInvalidValue;
  if (G == 1) return V == 1 ? launch_tail_gv<1, 1>(t, nb, s) : hipErrorInvalidValue;
  if (G == 4) {
    if (V % 2) return hipErrorInvalidValue;
    V /= 2;
  }
#define SS_CALL(v) launch_tail_gv<8, v>(t, nb, s)
  SS_DISPATCH_V(V, SS_CALL)
#undef SS_CALL
}

hipError_t launch_isb_up(const IsbUpArgs* a, int nb, int G, int V, int vt, bool fuse_argmin, int xcd, hipStream_t s,
                         bool reproj) {
  if (nb < 1 || nb > kMaxBatch) return hipErrorInvalidValue;
  if (G == 1)
    return (V == 1 && vt == SS_VOLUME_F32 && !fuse_argmin) ? launch_up_gv<1, 1, float>(a, nb, false, xcd, s, reproj)
                                                           : hipErrorInvalidValue;
  if (reproj) return hipErrorInvalidValue;
#define SS_CALL(T) up_t<T>(a, nb, G, V, fuse_argmin, xcd, s)
  SS_DISPATCH_T(vt, SS_CALL)
#undef SS_CALL
}

hipError_t launch_argmin(const void* cost, int npix, int V, int vt, const ArgminParams& p, float* dist,
                         hipStream_t s) {
#define SS_CALL(T) argmin_t<T>(cost, npix, V, p, dist, s)
  SS_DISPATCH_T(vt, SS_CALL)
#undef SS_CALL
}

hipError_t launch_planar_to_cc(const float* src, float* dst, int HW, int N, int NP, hipStream_t s) {
  hipLaunchKernelGGL(k_planar_to_cc, dim3(blocks_for(HW)), dim3(kBlock), 0, s, src, dst, HW, N, NP);
  return hipGetLastError();
}

hipError_t launch_cc_to_planar(const float* src, float* dst, int HW, int N, int NP, hipStream_t s) {
  hipLaunchKernelGGL(k_cc_to_planar, dim3(blocks_for(HW)), dim3(kBlock), 0, s, src, dst, HW, N, NP);
  return hipGetLastError();
}

hipError_t launch_guide3_to_packed(const uint8_t* g3, uint32_t* g, int HW, hipStream_t s) {
  hipLaunchKernelGGL(k_guide3_to_packed, dim3(blocks_for(HW)), dim3(kBlock), 0, s, g3, g, HW);
  return hipGetLastError();
}

hipError_t launch_packed_to_guide3(const uint32_t* g, uint8_t* g3, int HW, hipStream_t s) {
  hipLaunchKernelGGL(k_packed_to_guide3, dim3(blocks_for(HW)), dim3(kBlock), 0, s, g, g3, HW);
  return hipGetLastError();
}

template <int G, typename T, int VMAX, bool LOOP>
hipError_t down_ragged_t(const RaggedSpan<IsbDownEntry>& sp, hipStream_t s) {
  const auto k = &k_isb_down_ragged<G, T, VMAX, LOOP>;
  const int grid = ragged_grid(reinterpret_cast<const void*>(k), kBlock, 0, sp, kBlock / 64, LOOP);
  hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, sp);
  return hipGetLastError();
}

template <int G, typename T, bool FUSE, int VMAX, bool LOOP>
hipError_t up_ragged_t(const RaggedSpan<IsbUpEntry>& sp, hipStream_t s) {
  const auto k = &k_isb_up_ragged<G, T, FUSE, VMAX, LOOP>;
  const int grid = ragged_grid(reinterpret_cast<const void*>(k), kBlock, 0, sp, kBlock / 64, LOOP);
  hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, sp);
  return hipGetLastError();
}

#define SS_DISPATCH_VMAX(vmax, CALL)            \
  switch (vmax) {                               \
    case 1: return CALL(1);                     \
    case 2: return CALL(2);                     \
    case 3: return CALL(3);                     \
    case 4: return CALL(4);                     \
    case 6: return CALL(6);                     \
    case 8: return CALL(8);                     \
    case 12: return CALL(12);                   \
    case 16: return CALL(16);                   \
    default: return hipErrorInvalidValue;       \
  }

template <typename T>
hipError_t down_ragged_vt(const RaggedSpan<IsbDownEntry>& sp, int vmax, bool loop, hipStream_t s) {
#define SS_CALL(m) (loop ? down_ragged_t<8, T, m, true>(sp, s) : down_ragged_t<8, T, m, false>(sp, s))
  SS_DISPATCH_VMAX(vmax, SS_CALL)
#undef SS_CALL
}

template <typename T, bool FUSE>
hipError_t up_ragged_vt(const RaggedSpan<IsbUpEntry>& sp, int vmax, bool loop, hipStream_t s) {
#define SS_CALL(m) (loop ? up_ragged_t<8, T, FUSE, m, true>(sp, s) : up_ragged_t<8, T, FUSE, m, false>(sp, s))
  SS_DISPATCH_VMAX(vmax, SS_CALL)
#undef SS_CALL
}

int ragged_vmax(int V) { return V; }

hipError_t launch_isb_down_ragged(const RaggedSpan<IsbDownEntry>& sp, int G, int vt, int vmax, bool loop,
                                  hipStream_t s) {
  if (sp.n < 1 || sp.total < 1) return hipErrorInvalidValue;
  if (G == 1) {
    if (vt != SS_VOLUME_F32 || vmax != 1) return hipErrorInvalidValue;
    return loop ? down_ragged_t<1, float, 1, true>(sp, s) : down_ragged_t<1, float, 1, false>(sp, s);
  }
#define SS_CALL(T) down_ragged_vt<T>(sp, vmax, loop, s)
  SS_DISPATCH_T(vt, SS_CALL)
#undef SS_CALL
}

hipError_t launch_isb_up_ragged(const RaggedSpan<IsbUpEntry>& sp, int G, int vt, int vmax, bool fuse_argmin, bool loop,
                                hipStream_t s) {
  if (sp.n < 1 || sp.total < 1) return hipErrorInvalidValue;
  if (G == 1) {
    if (vt != SS_VOLUME_F32 || vmax != 1 || fuse_argmin) return hipErrorInvalidValue;
    return loop ? up_ragged_t<1, float, false, 1, true>(sp, s) : up_ragged_t<1, float, false, 1, false>(sp, s);
  }
  if (fuse_argmin) {
#define SS_CALL(T) up_ragged_vt<T, true>(sp, vmax, loop, s)
    SS_DISPATCH_T(vt, SS_CALL)
#undef SS_CALL
  }
#define SS_CALL(T) up_ragged_vt<T, false>(sp, vmax, loop, s)
  SS_DISPATCH_T(vt, SS_CALL)
#undef SS_CALL
}

hipError_t launch_isb_tail_ragged(const TailEntry* t, int n, int G, unsigned lds_bytes, hipStream_t s) {
  if (n < 1) return hipErrorInvalidValue;
  if (G == 1)
    hipLaunchKernelGGL(k_isb_tail_ragged<1>, dim3(n), dim3(kTailBlock), lds_bytes, s, t);
  else
    hipLaunchKernelGGL(k_isb_tail_ragged<8>, dim3(n), dim3(kTailBlock), lds_bytes, s, t);
  return hipGetLastError();
}

}  // namespace ss
