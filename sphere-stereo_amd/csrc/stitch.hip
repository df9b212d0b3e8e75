// Panorama stitching on gfx950: z-tested reprojection, inpainting, blending
// tables and the RGB-D merge.
//
// Behaviour: stitcher.cu:181-474 and stitcher.py:49-220 of the reference, with
// two intentional changes that remove the reference's data races:
//   * reprojection takes an exact z-minimum with an integer atomicMin on the
//     bit pattern of the (non-negative) distance instead of a non-atomic
//     read-compare-write run twice (stitcher.cu:200-202, stitcher.py:184-192);
//   * each inpainting pass reads the previous pass's map and writes a second
//     buffer (Jacobi) instead of updating the map in place while other threads
//     read it (stitcher.cu:310, 319).
#include <string.h>

#include "ss_kernels.h"
#include "ss_ragged.h"

namespace ss {

namespace {

constexpr int kBlock = 256;
inline int blocks_for(long long threads) { return (int)((threads + kBlock - 1) / kBlock); }

// bilinear read with floor and no bounds check (stitcher.cu:89-116)
__device__ __forceinline__ float interp_f(const float* __restrict__ m, int cols, float u, float v) {
  const int u1 = (int)floorf(u), v1 = (int)floorf(v), u2 = u1 + 1, v2 = v1 + 1;
  const float u1f = (float)u1, u2f = (float)u2, v1f = (float)v1, v2f = (float)v2;
  const float w1 = (u2f - u) * (v2f - v), w2 = (u2f - u) * (v - v1f), w3 = (u - u1f) * (v2f - v),
              w4 = (u - u1f) * (v - v1f);
  const float p1 = m[(size_t)v1 * cols + u1], p2 = m[(size_t)v2 * cols + u1], p3 = m[(size_t)v1 * cols + u2],
              p4 = m[(size_t)v2 * cols + u2];
  return ((w1 * p1 + w2 * p2) + w3 * p3) + w4 * p4;
}

__global__ void __launch_bounds__(kBlock) k_fill(float* __restrict__ p, float v, size_t n) {
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

// reprojectDistanceKernel (stitcher.cu:181-204) as an exact atomic z-minimum
// (reproject_splat); map blockIdx.y of the batch
__device__ __forceinline__ void reproject_pixel(const ReprojectArgs& a, int H, int W, int i) {
  reproject_splat(a.c, a.t, a.dout, H, W, i, a.din[i]);
}

__global__ void __launch_bounds__(kBlock) k_reproject(Batch<ReprojectArgs> b, int H, int W) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= H * W) return;
  reproject_pixel(b.a[blockIdx.y], H, W, i);
}

// ragged: a wavefront's item is kPixItem consecutive pixels of one map, a lane each
template <bool LOOP>
__global__ void __launch_bounds__(kBlock) k_reproject_ragged(RaggedSpan<ReprojEntry> sp) {
  const int lane = (int)(threadIdx.x & 63);
  for_each_wave_item<LOOP>(sp, [&](int e, int t) {
    const ReprojEntry r = load_const(sp.e + e);
    const int n = r.H * r.W;
    const int i = t * kPixItem + lane;
    if (i < n) reproject_pixel(r.a, r.H, r.W, i);
  });
}

// one Jacobi pass of inpaintKernel (stitcher.cu:283-323)
__global__ void __launch_bounds__(kBlock) k_inpaint_step(const float* __restrict__ din, float* __restrict__ dout,
                                                         const uchar2* __restrict__ lut, int H, int W,
                                                         double thr_empty, double thr_src) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= H * W) return;
  const float d = din[i];
  float r = d;
  const int x = i % W, y = i / W;
  if ((double)d >= thr_empty && x < W - 1 && y < H - 1 && y >= 1 && x >= 1) {
    const uchar2 b = lut[i];
    const int bs[2] = {b.x, b.y};
    float val = 0.0f, wsum = 0.0f;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int nx = x + (bs[s] & 3) - 1, ny = y + ((bs[s] & 12) >> 2) - 1;
      const float w = (float)(bs[s] & 240);
      const float sd = din[ny * W + nx];
      if ((double)sd <= thr_src && w > 0.0f) {
        val = val + w * sd;
        wsum = wsum + w;
      }
    }
    if (wsum > 0.0f) r = val / wsum;
  }
  dout[i] = r;
}

// Several Jacobi passes of inpaintKernel in one launch.  A workgroup owns a
// kTile x kTile block of the map; it stages the block with a halo of kSteps
// pixels (map values and table bytes) in LDS, twice, and lists the pixels
// that can change.  Each pass recomputes only the listed pixels, reading one
// LDS copy and writing the other; unlisted pixels never change, so both
// copies keep them.  A pass moves information by at most one pixel, so the
// block's own pixels, kSteps inside the staged region, are exact after
// kSteps passes.
//
// Listed: empty, inside both the map and the staged region, and within
// `steps` pixels (Chebyshev distance) of a pixel that can act as a source
// (value <= thr_src).  A pixel changes only when one of its table neighbours,
// one pixel away, holds a source value; values become sources only by being
// filled from one; so a pixel farther than `steps` from every source of the
// staged region keeps its value for the whole launch (the interior of big
// holes, the area outside a fisheye's field of view).  The launch also stops
// after a pass that changes nothing, since every later pass would repeat it.
// The tile side TILE is a template parameter (the frame path's inpainting
// autotuner picks it per map size; the ragged executor uses kTile): 48 +
// 2 * 16 = 80 is 79 KB of LDS, 2 workgroups per CU; 32 (50 KB, 3 per CU) and
// 16 (19 KB) stage more halo per tile.  The result does not depend on it.
// blockIdx.z selects the map of the batch.
constexpr int kTile = 48;
constexpr int kSteps = 16;
template <int TILE>
struct InpaintGeom {
  static constexpr int kRegion = TILE + 2 * kSteps;
  static constexpr int kRowWords = (kRegion + 31) / 32;  // a row of the region as a bit mask
  static_assert(kRegion * kRegion % kBlock == 0 && kRegion * kRegion <= 65536 && kRegion <= 96 &&
                    kRegion * kRowWords <= kBlock,
                "region fits the staging loop, the list and the row masks");
};

// A region row's bit mask (bits 0..95 in three words) shifted by k (0..31)
// towards higher or lower x.
__device__ __forceinline__ void row_shl(const uint32_t (&m)[3], int k, uint32_t (&o)[3]) {
  o[0] = m[0] << k;
  o[1] = (m[1] << k) | (k ? m[0] >> (32 - k) : 0u);
  o[2] = (m[2] << k) | (k ? m[1] >> (32 - k) : 0u);
}
__device__ __forceinline__ void row_shr(const uint32_t (&m)[3], int k, uint32_t (&o)[3]) {
  o[2] = m[2] >> k;
  o[1] = (m[1] >> k) | (k ? m[2] << (32 - k) : 0u);
  o[0] = (m[0] >> k) | (k ? m[1] << (32 - k) : 0u);
}

// The passes of one tile (bx, by) of a map: a workgroup's whole body (it
// returns with every thread past its last barrier).
template <int TILE>
__device__ __forceinline__ void inpaint_tile(const float* __restrict__ din, float* __restrict__ dout,
                                             const uchar2* __restrict__ lut, int H, int W, int steps,
                                             double thr_empty, double thr_src, int bx, int by) {
  constexpr int kRegion = InpaintGeom<TILE>::kRegion, kRowWords = InpaintGeom<TILE>::kRowWords;
  __shared__ float bufa[kRegion * kRegion];
  __shared__ float bufb[kRegion * kRegion];
  __shared__ uchar2 tab[kRegion * kRegion];
  __shared__ uint16_t list[kRegion * kRegion];
  __shared__ uint32_t srcm[kRegion][kRowWords];  // sources, then sources dilated along rows
  __shared__ uint32_t near_[kRegion][kRowWords]; // within `steps` of a source
  __shared__ unsigned count, changed[2];  // changed[s & 1] == s + 1: pass s changed a pixel
  const int tid = (int)threadIdx.x;
  const int ox = bx * TILE - kSteps, oy = by * TILE - kSteps;
  const int reach = min(steps, kSteps);
  // A workgroup of the persistent ragged form runs several tiles in a row:
  // every thread must be done with the previous tile's read-out (bufa/bufb)
  // and its last test of `changed` before this tile resets and restages them.
  __syncthreads();
  if (tid == 0) {
    count = 0;
    changed[0] = 0;
    changed[1] = 0;
  }
  // stage: all loads of a thread are issued before any is used
  constexpr int kPer = kRegion * kRegion / kBlock;
  float dv[kPer];
  uchar2 tv[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = tid + k * kBlock;
    const int gx = ox + i % kRegion, gy = oy + i / kRegion;
    const bool in = gx >= 0 && gx < W && gy >= 0 && gy < H;
    const int g = in ? gy * W + gx : 0;
    dv[k] = in ? din[g] : 0.0f;
    tv[k] = in ? lut[g] : make_uchar2(0, 0);
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = tid + k * kBlock;
    bufa[i] = dv[k];
    bufb[i] = dv[k];
    tab[i] = tv[k];
  }
  __syncthreads();
  // source bits, one word per thread (pixels outside the map, staged as 0,
  // count as sources: no listed pixel reads them, so they can only list a
  // pixel that would not change, never drop one that would)
  if (tid < kRegion * kRowWords) {
    const int y = tid / kRowWords, w = tid % kRowWords;
    uint32_t b = 0;
    for (int x = 32 * w; x < min(32 * w + 32, kRegion); ++x)
      if ((double)bufa[y * kRegion + x] <= thr_src) b |= 1u << (x - 32 * w);
    srcm[y][w] = b;
  }
  __syncthreads();
  // dilate along the row (one thread per row), then across rows
  uint32_t hrow[3] = {0u, 0u, 0u};
  if (tid < kRegion) {
    const uint32_t m[3] = {srcm[tid][0], srcm[tid][1], kRowWords > 2 ? srcm[tid][2] : 0u};
    hrow[0] = m[0];
    hrow[1] = m[1];
    hrow[2] = m[2];
    for (int k = 1; k <= reach; ++k) {
      uint32_t l[3], r[3];
      row_shl(m, k, l);
      row_shr(m, k, r);
      hrow[0] |= l[0] | r[0];
      hrow[1] |= l[1] | r[1];
      hrow[2] |= l[2] | r[2];
    }
  }
  __syncthreads();  // every row read
  if (tid < kRegion)
    for (int w = 0; w < kRowWords; ++w) srcm[tid][w] = hrow[w];
  __syncthreads();
  if (tid < kRegion * kRowWords) {
    const int y = tid / kRowWords, w = tid % kRowWords;
    uint32_t b = 0;
    for (int yy = max(y - reach, 0); yy <= min(y + reach, kRegion - 1); ++yy) b |= srcm[yy][w];
    near_[y][w] = b;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = tid + k * kBlock;
    const int lx = i % kRegion, ly = i / kRegion;
    const int gx = ox + lx, gy = oy + ly;
    if ((double)dv[k] >= thr_empty && gx >= 1 && gy >= 1 && gx < W - 1 && gy < H - 1 && lx >= 1 && ly >= 1 &&
        lx < kRegion - 1 && ly < kRegion - 1 && ((near_[ly][lx >> 5] >> (lx & 31)) & 1u))
      list[atomicAdd(&count, 1u)] = (uint16_t)i;
  }
  __syncthreads();
  const int n = (int)count;
  float* a = bufa;
  float* b = bufb;
  for (int s = 0; s < steps && n > 0; ++s) {
    bool moved = false;
    for (int j = threadIdx.x; j < n; j += kBlock) {
      const int i = list[j];
      const float d = a[i];
      float r = d;
      if ((double)d >= thr_empty) {
        const int lx = i % kRegion, ly = i / kRegion;
        const uchar2 t = tab[i];
        const int bs[2] = {t.x, t.y};
        float val = 0.0f, wsum = 0.0f;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          // table offsets are 0..2 (createInpaintingWeightsKernel), so the
          // neighbour of a listed pixel lies inside the staged region
          const int nx = lx + (bs[k] & 3) - 1, ny = ly + ((bs[k] & 12) >> 2) - 1;
          const float w = (float)(bs[k] & 240);
          const float sd = a[min(ny, kRegion - 1) * kRegion + min(nx, kRegion - 1)];
          if ((double)sd <= thr_src && w > 0.0f) {
            val = val + w * sd;
            wsum = wsum + w;
          }
        }
        if (wsum > 0.0f) r = val / wsum;
      }
      b[i] = r;
      moved = moved || __float_as_uint(r) != __float_as_uint(d);
    }
    if (moved) changed[s & 1] = s + 1;  // every writer stores the same value
    __syncthreads();
    float* t = a;
    a = b;
    b = t;
    // uniform: pass s changed nothing, nor would the rest.  (The flag of
    // pass s is next written in pass s + 2, after every lane has read it.)
    if (changed[s & 1] != (unsigned)(s + 1)) break;
  }
  for (int i = threadIdx.x; i < TILE * TILE; i += kBlock) {
    const int lx = kSteps + i % TILE, ly = kSteps + i / TILE;
    const int gx = ox + lx, gy = oy + ly;
    if (gx < W && gy < H) dout[gy * W + gx] = a[ly * kRegion + lx];
  }
}

template <int TILE>
__global__ void __launch_bounds__(kBlock) k_inpaint_multi(Batch<InpaintArgs> bt, int H, int W, int steps,
                                                          double thr_empty, double thr_src) {
  const InpaintArgs& a = bt.a[blockIdx.z];
  inpaint_tile<TILE>(a.din, a.dout, reinterpret_cast<const uchar2*>(a.lut), H, W, steps, thr_empty, thr_src,
                     blockIdx.x, blockIdx.y);
}

// ragged: a workgroup's item is one tile of one map, with the map's own
// step count (0: the tile is copied)
template <bool LOOP>
__global__ void __launch_bounds__(kBlock) k_inpaint_ragged(RaggedSpan<InpaintEntry> sp) {
  for_each_block_item<LOOP>(sp, [&](int e, int t) {
    const InpaintEntry m = load_const(sp.e + e);
    const int by = t / m.tiles_x, bx = t - by * m.tiles_x;
    inpaint_tile<kTile>(m.a.din, m.a.dout, reinterpret_cast<const uchar2*>(m.a.lut), m.H, m.W, m.steps, m.thr_empty,
                        m.thr_src, bx, by);
  });
}

// createInpaintingWeightsKernel (stitcher.cu:212-272)
__global__ void __launch_bounds__(kBlock) k_inpaint_lut(ss_intrinsics c, float tx, float ty, float tz, int H, int W,
                                                        float mn, float mx, uchar2* __restrict__ lut) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= H * W) return;
  float ux, uy, uz, pcx, pcy, pfx, pfy;
  bool valid = true;
  unproject_c(c, (float)(i % W), (float)(i / W), ux, uy, uz);
  project_c<false>(c, ux * mn - tx, uy * mn - ty, uz * mn - tz, pcx, pcy, valid);
  project_c<false>(c, ux * mx - tx, uy * mx - ty, uz * mx - tz, pfx, pfy, valid);
  float dx = pfx - pcx, dy = pfy - pcy;
  const float il = 1.0f / sqrtf(dx * dx + dy * dy);
  dx = dx * il;
  dy = dy * il;
  float wt0 = 0.0f, wt1 = 0.0f;
  int nx0 = 0, ny0 = 0, nx1 = 0, ny1 = 0;
#pragma unroll
  for (int n = -1; n <= 1; ++n)
#pragma unroll
    for (int m = -1; m <= 1; ++m) {
      if (!(n || m)) continue;
      const float ip = 1.0f / sqrtf((float)n * (float)n + (float)m * (float)m);
      const float w = ((float)n * ip) * dx + ((float)m * ip) * dy;
      if (w > wt1) {
        wt0 = wt1;
        nx0 = nx1;
        ny0 = ny1;
        wt1 = w;
        nx1 = n;
        ny1 = m;
      } else if (w > wt0) {
        wt0 = w;
        nx0 = n;
        ny0 = m;
      }
    }
  const unsigned b0 = ((unsigned)(uint8_t)(wt0 * 255.0f) & 240u) + 4u * (unsigned)(1 + ny0) + (unsigned)(1 + nx0);
  const unsigned b1 = ((unsigned)(uint8_t)(wt1 * 255.0f) & 240u) + 4u * (unsigned)(1 + ny1) + (unsigned)(1 + nx1);
  lut[i] = make_uchar2((uint8_t)b0, (uint8_t)b1);
}

// Separable box filter: row sums over the padded rows, then column sums.
__global__ void __launch_bounds__(kBlock) k_box_rows(const float* __restrict__ in, float* __restrict__ rows, int H,
                                                     int W, int r, float pad) {
  const int Hp = H + 2 * r;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= Hp * W) return;
  const int yy = i / W, x = i % W, y = yy - r;
  float s = 0.0f;
  for (int j = -r; j <= r; ++j) {
    const int xx = x + j;
    const float v = (y < 0 || y >= H || xx < 0 || xx >= W) ? pad : in[(size_t)y * W + xx];
    s = s + v;
  }
  rows[i] = s;
}

__global__ void __launch_bounds__(kBlock) k_box_cols(const float* __restrict__ rows, float* __restrict__ out, int H,
                                                     int W, int r, float k) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= H * W) return;
  const int y = i / W, x = i % W;
  float s = 0.0f;
  for (int t = 0; t <= 2 * r; ++t) s = s + rows[(size_t)(y + t) * W + x];
  out[i] = s * k;
}

// createBlendingLutKernel (stitcher.cu:338-401)
__global__ void __launch_bounds__(kBlock) k_blend_lut(BlendArgs a) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  const int P = a.PH * a.PW;
  if (idx >= P) return;
  const double kPi = 3.14159265;  // stitcher.cu:325
  const int x = idx % a.PW, y = idx / a.PW;
  const float phi = (float)(((double)((float)y + 0.5f) * kPi) / a.PH - kPi / 2.0);
  const float theta = (float)(((double)(((float)x + 0.5f) * 2.0f) * kPi) / a.PW + kPi);
  const float cp = (float)cos((double)phi), sp = (float)sin((double)phi);
  const float ct = (float)cos((double)theta), st = (float)sin((double)theta);
  const float ux = cp * st, uy = sp, uz = cp * ct;
  float bw[SS_MAX_REFERENCES];
  float bsum = 0.0f;
  const float Wm = (float)a.W - 1.1f, Hm = (float)a.H - 1.1f;
#pragma unroll
  for (int r = 0; r < SS_MAX_REFERENCES; ++r) {
    if (r >= a.R) break;
    const float* rm = a.rot[r];
    const float fx = (rm[0] * ux + rm[1] * uy) + rm[2] * uz;
    const float fy = (rm[3] * ux + rm[4] * uy) + rm[5] * uz;
    const float fz = (rm[6] * ux + rm[7] * uy) + rm[8] * uz;
    bool valid = true;
    float u, v;
    project_c<true>(a.intr[r], fx, fy, fz, u, v, valid);
    u = fminf(fmaxf(u, 0.1f), Wm);
    v = fminf(fmaxf(v, 0.1f), Hm);
    a.sampling[(size_t)r * P + idx] = make_float2(u, v);
    bw[r] = 1e-8f;
    const float tx = a.trans[r][0], ty = a.trans[r][1], tz = a.trans[r][2];
    float nu, nv, fu, fv;
    project_c<true>(a.intr[r], fx * a.mn - tx, fy * a.mn - ty, fz * a.mn - tz, nu, nv, valid);
    nu = fminf(fmaxf(nu, 0.1f), Wm);
    nv = (float)(r * a.H) + fminf(fmaxf(nv, 0.1f), Hm);
    project_c<true>(a.intr[r], fx * a.mx - tx, fy * a.mx - ty, fz * a.mx - tz, fu, fv, valid);
    fu = fminf(fmaxf(fu, 0.1f), Wm);
    fv = (float)(r * a.H) + fminf(fmaxf(fv, 0.1f), Hm);
    if (valid && (double)interp_f(a.masks, a.W, nu, nv) > 0.99 && (double)interp_f(a.masks, a.W, fu, fv) > 0.99) {
      const float dx = fu - nu, dy = fv - nv;
      const float ds = sqrtf(dx * dx + dy * dy);
      bw[r] = expf_cw((float)((double)(-ds * ds) / (1e-4 * a.H * a.W)));
    }
    bsum = bsum + bw[r];
  }
#pragma unroll
  for (int r = 0; r < SS_MAX_REFERENCES; ++r) {
    if (r >= a.R) break;
    a.weights[(size_t)r * P + idx] = bw[r] / bsum;
  }
}

__global__ void __launch_bounds__(kBlock) k_normalize(float* __restrict__ w, int R, int P) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= P) return;
  float s = 0.0f;
  for (int r = 0; r < R; ++r) s = s + w[(size_t)r * P + p];
  for (int r = 0; r < R; ++r) w[(size_t)r * P + p] = w[(size_t)r * P + p] / s;
}

// mergeRGBDPanoramaKernel (stitcher.cu:423-474), one lane per panorama pixel.
// Each reference's work is a chain of three dependent gathers (blend tables,
// then the two distance maps at the table's position, then the stitch image
// at the reprojected position).  With RN references known at compile time
// the chains of all references are issued side by side, stage by stage, so
// their memory latencies overlap; the sums still run in reference order, as
// the reference kernel accumulates them.  RN = 0 is the generic loop.
template <int RN>
__device__ __forceinline__ void merge_pixel(const MergeArgs& a, int idx) {
  constexpr int RM = RN > 0 ? RN : 1;
  const int P = a.PH * a.PW;
  const float rx = (float)a.Ws / (float)a.W, ry = (float)a.Hs / (float)a.H;
  const float Wsm = (float)a.Ws - 1.1f, Hsm = (float)a.Hs - 1.1f;
  float inv_sum = 0.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
  const int R = RN > 0 ? RN : a.R;
  for (int r0 = 0; r0 < R; r0 += RM) {
    float bw[RM], rd[RM], dc[RM], u[RM], v[RM];
    float2 uv[RM];
#pragma unroll
    for (int j = 0; j < RM; ++j) {
      const int r = r0 + j;
      bw[j] = a.weights[(size_t)r * P + idx];
      uv[j] = a.sampling[(size_t)r * P + idx];
    }
#pragma unroll
    for (int j = 0; j < RM; ++j) {
      const int r = r0 + j;
      // the reference offsets v into the stacked maps (an fp32 add), then samples
      const float vo = uv[j].y + (float)(r * a.H);
      rd[j] = interp_f(a.reproj[r] - (size_t)r * a.H * a.W, a.W, uv[j].x, vo);
      dc[j] = interp_f(a.dists[r] - (size_t)r * a.H * a.W, a.W, uv[j].x, vo);
    }
#pragma unroll
    for (int j = 0; j < RM; ++j) {
      const int r = r0 + j;
      inv_sum = inv_sum + (bw[j] * 1.0f) / rd[j];
      const float d = fminf(rd[j], dc[j]);
      float px, py, pz;
      bool valid = true;
      unproject_c(a.intr[r], uv[j].x, uv[j].y, px, py, pz);
      px = px * d + a.trans[r][0];
      py = py * d + a.trans[r][1];
      pz = pz * d + a.trans[r][2];
      project_c<false>(a.intr[r], px, py, pz, u[j], v[j], valid);
      u[j] = fminf(fmaxf(u[j] * rx, 0.1f), Wsm);
      v[j] = fminf(fmaxf(v[j] * ry, 0.1f), Hsm) + (float)(r * a.Hs);
    }
    float t[RM][4][3];
#pragma unroll
    for (int j = 0; j < RM; ++j) {
      // bilinear uint8 read (stitcher.cu:56-83) on the stacked images
      const int r = r0 + j;
      const uint8_t* img = a.imgs[r] - (size_t)r * a.Hs * a.Ws * 3;
      const int u1 = (int)floorf(u[j]), v1 = (int)floorf(v[j]);
      const uint8_t* p1 = img + ((size_t)v1 * a.Ws + u1) * 3;
      const uint8_t* p2 = p1 + (size_t)a.Ws * 3;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        t[j][0][c] = (float)p1[c];
        t[j][1][c] = (float)p2[c];
        t[j][2][c] = (float)p1[3 + c];
        t[j][3][c] = (float)p2[3 + c];
      }
    }
#pragma unroll
    for (int j = 0; j < RM; ++j) {
      const float u1f = floorf(u[j]), v1f = floorf(v[j]), u2f = u1f + 1.0f, v2f = v1f + 1.0f;
      const float w1 = (u2f - u[j]) * (v2f - v[j]), w2 = (u2f - u[j]) * (v[j] - v1f),
                  w3 = (u[j] - u1f) * (v2f - v[j]), w4 = (u[j] - u1f) * (v[j] - v1f);
      float sc[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) sc[c] = ((t[j][0][c] * w1 + t[j][1][c] * w2) + t[j][2][c] * w3) + t[j][3][c] * w4;
      c0 = c0 + sc[0] * bw[j];
      c1 = c1 + sc[1] * bw[j];
      c2 = c2 + sc[2] * bw[j];
    }
  }
  a.dist_out[idx] = 1.0f / inv_sum;
  a.rgb_out[3 * idx] = (uint8_t)c0;
  a.rgb_out[3 * idx + 1] = (uint8_t)c1;
  a.rgb_out[3 * idx + 2] = (uint8_t)c2;
}

template <int RN>
__global__ void __launch_bounds__(kBlock) k_merge(MergeArgs a) {
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= a.PH * a.PW) return;
  merge_pixel<RN>(a, idx);
}

// ragged: a wavefront's item is kPixItem consecutive panorama pixels of one
// frame (a lane each); every frame of the launch has RN references (RN = 0:
// any count, the generic loop), as k_merge<RN>
template <bool LOOP, int RN>
__global__ void __launch_bounds__(kBlock) k_merge_ragged(RaggedSpan<MergeArgs> sp) {
  const int lane = (int)(threadIdx.x & 63);
  for_each_wave_item<LOOP>(sp, [&](int e, int t) {
    const int i = t * kPixItem + lane;
    if constexpr (RN > 0) {
      // every index into the record is fixed: read it through the constant
      // address space (scalar loads into SGPRs)
      const MergeArgs a = load_const(sp.e + e);
      if (i < a.PH * a.PW) merge_pixel<RN>(a, i);
    } else {
      const MergeArgs& a = sp.e[e];
      if (i < a.PH * a.PW) merge_pixel<RN>(a, i);
    }
  });
}

// Four elements per lane (16-byte load, 4-byte store) when both buffers are
// aligned for it; blockIdx.y selects the image of the batch.
template <bool VEC4>
__global__ void __launch_bounds__(kBlock) k_to_u8(Batch<ToU8Args> b, size_t n) {
  const float* __restrict__ in = b.a[blockIdx.y].in;
  uint8_t* __restrict__ out = b.a[blockIdx.y].out;
  const size_t i = ((size_t)blockIdx.x * kBlock + threadIdx.x) * 4;
  if (i >= n) return;
  if (VEC4 && i + 4 <= n) {
    const float4 v = *reinterpret_cast<const float4*>(in + i);
    const uint32_t q = (uint32_t)(uint8_t)v.x | ((uint32_t)(uint8_t)v.y << 8) | ((uint32_t)(uint8_t)v.z << 16) |
                       ((uint32_t)(uint8_t)v.w << 24);
    *reinterpret_cast<uint32_t*>(out + i) = q;
    return;
  }
  for (size_t k = i; k < i + 4 && k < n; ++k) out[k] = (uint8_t)in[k];
}

// ragged: a wavefront's item is kCastItem consecutive values of one image,
// kCastRounds rounds of 4 per lane (each round a contiguous 1 KB of the
// image; in a whole item every round's load is issued before the stores,
// so a wavefront keeps kCastRounds loads in flight); src is the run's
// stitch-image table
__device__ __forceinline__ uint32_t pack_u8x4(const float4& v) {
  return (uint32_t)(uint8_t)v.x | ((uint32_t)(uint8_t)v.y << 8) | ((uint32_t)(uint8_t)v.z << 16) |
         ((uint32_t)(uint8_t)v.w << 24);
}
template <bool LOOP>
__global__ void __launch_bounds__(kBlock) k_to_u8_ragged(RaggedSpan<CastEntry> sp, const float* const* src) {
  const int lane = (int)(threadIdx.x & 63);
  for_each_wave_item<LOOP>(sp, [&](int e, int t) {
    const CastEntry c = load_const(sp.e + e);
    const float* __restrict__ in = src[c.src];
    uint8_t* __restrict__ out = c.out;
    const bool vec = (reinterpret_cast<uintptr_t>(in) & 15) == 0 && (reinterpret_cast<uintptr_t>(out) & 3) == 0;
    const long long base = (long long)t * kCastItem, i0 = base + 4 * lane;
    if (vec && base + kCastItem <= c.n) {
      float4 v[kCastRounds];
#pragma unroll
      for (int r = 0; r < kCastRounds; ++r) v[r] = *reinterpret_cast<const float4*>(in + i0 + 256 * r);
#pragma unroll
      for (int r = 0; r < kCastRounds; ++r) *reinterpret_cast<uint32_t*>(out + i0 + 256 * r) = pack_u8x4(v[r]);
      return;
    }
    // the image's last item, or unaligned buffers
    for (int r = 0; r < kCastRounds; ++r) {
      const long long i = i0 + 256 * r;
      if (i >= c.n) break;
      if (vec && i + 4 <= c.n)
        *reinterpret_cast<uint32_t*>(out + i) = pack_u8x4(*reinterpret_cast<const float4*>(in + i));
      else
        for (long long k = i; k < i + 4 && k < c.n; ++k) out[k] = (uint8_t)in[k];
    }
  });
}

__global__ void k_publish(PublishArgs a) {
  if ((int)threadIdx.x < a.n) a.dst[threadIdx.x] = a.src[threadIdx.x];
}

template <typename A>
Batch<A> batch_of(const A* a, int nb) {
  Batch<A> b;
  memset(&b, 0, sizeof b);
  for (int i = 0; i < nb; ++i) b.a[i] = a[i];
  return b;
}

}  // namespace

hipError_t launch_fill(float* p, float v, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for((long long)n)), dim3(kBlock), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_reproject(const ReprojectArgs* a, int nb, int H, int W, hipStream_t s) {
  if (nb < 1 || nb > kMaxBatch) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_reproject, dim3(blocks_for((long long)H * W), nb), dim3(kBlock), 0, s, batch_of(a, nb), H, W);
  return hipGetLastError();
}

hipError_t launch_inpaint_step(const float* din, float* dout, const uint8_t* lut, int H, int W, double thr_empty,
                               double thr_src, hipStream_t s) {
  hipLaunchKernelGGL(k_inpaint_step, dim3(blocks_for((long long)H * W)), dim3(kBlock), 0, s, din, dout,
                     reinterpret_cast<const uchar2*>(lut), H, W, thr_empty, thr_src);
  return hipGetLastError();
}

template <int TILE>
hipError_t launch_inpaint_tile(const InpaintArgs* a, int nb, int H, int W, int steps, double thr_empty,
                               double thr_src, hipStream_t s) {
  const dim3 grid((W + TILE - 1) / TILE, (H + TILE - 1) / TILE, nb);
  hipLaunchKernelGGL(k_inpaint_multi<TILE>, grid, dim3(kBlock), 0, s, batch_of(a, nb), H, W, steps, thr_empty,
                     thr_src);
  return hipGetLastError();
}

// tile: 48, 32 or 16 (inpaint_tiles)
hipError_t launch_inpaint_multi(const InpaintArgs* a, int nb, int H, int W, int steps, double thr_empty,
                                double thr_src, int tile, hipStream_t s) {
  if (nb < 1 || nb > kMaxBatch) return hipErrorInvalidValue;
  switch (tile) {
    case 48: return launch_inpaint_tile<48>(a, nb, H, W, steps, thr_empty, thr_src, s);
    case 32: return launch_inpaint_tile<32>(a, nb, H, W, steps, thr_empty, thr_src, s);
    case 16: return launch_inpaint_tile<16>(a, nb, H, W, steps, thr_empty, thr_src, s);
    default: return hipErrorInvalidValue;
  }
}

int inpaint_tile_count() { return 3; }
int inpaint_tile_of(int variant) {
  static const int kTiles[3] = {48, 32, 16};
  return variant >= 0 && variant < 3 ? kTiles[variant] : kTile;
}

int inpaint_steps_per_launch() { return kSteps; }

hipError_t launch_inpaint_lut(const ss_intrinsics& c, const float* t, int H, int W, float mn, float mx, uint8_t* lut,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_inpaint_lut, dim3(blocks_for((long long)H * W)), dim3(kBlock), 0, s, c, t[0], t[1], t[2], H,
                     W, mn, mx, reinterpret_cast<uchar2*>(lut));
  return hipGetLastError();
}

hipError_t launch_box_filter(const float* in, float* out, float* tmp, int B, int H, int W, int r, float pad,
                             hipStream_t s) {
  const float k = 1.0f / (float)((2 * r + 1) * (2 * r + 1));
  for (int b = 0; b < B; ++b) {
    hipLaunchKernelGGL(k_box_rows, dim3(blocks_for((long long)(H + 2 * r) * W)), dim3(kBlock), 0, s,
                       in + (size_t)b * H * W, tmp, H, W, r, pad);
    hipLaunchKernelGGL(k_box_cols, dim3(blocks_for((long long)H * W)), dim3(kBlock), 0, s, tmp,
                       out + (size_t)b * H * W, H, W, r, k);
  }
  return hipGetLastError();
}

hipError_t launch_blend_lut(const BlendArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_blend_lut, dim3(blocks_for((long long)a.PH * a.PW)), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_normalize(float* w, int R, int P, hipStream_t s) {
  hipLaunchKernelGGL(k_normalize, dim3(blocks_for(P)), dim3(kBlock), 0, s, w, R, P);
  return hipGetLastError();
}

hipError_t launch_merge(const MergeArgs& a, hipStream_t s) {
  const dim3 grid(blocks_for((long long)a.PH * a.PW));
  switch (a.R) {
    case 1: hipLaunchKernelGGL(k_merge<1>, grid, dim3(kBlock), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_merge<2>, grid, dim3(kBlock), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_merge<3>, grid, dim3(kBlock), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_merge<4>, grid, dim3(kBlock), 0, s, a); break;
    default: hipLaunchKernelGGL(k_merge<0>, grid, dim3(kBlock), 0, s, a); break;
  }
  return hipGetLastError();
}

hipError_t launch_to_u8(const ToU8Args* a, int nb, size_t n, hipStream_t s) {
  if (nb < 1 || nb > kMaxBatch) return hipErrorInvalidValue;
  bool vec = true;
  for (int i = 0; i < nb; ++i)
    vec = vec && (reinterpret_cast<uintptr_t>(a[i].in) & 15) == 0 && (reinterpret_cast<uintptr_t>(a[i].out) & 3) == 0;
  const dim3 grid(blocks_for((long long)((n + 3) / 4)), nb);
  if (vec)
    hipLaunchKernelGGL(k_to_u8<true>, grid, dim3(kBlock), 0, s, batch_of(a, nb), n);
  else
    hipLaunchKernelGGL(k_to_u8<false>, grid, dim3(kBlock), 0, s, batch_of(a, nb), n);
  return hipGetLastError();
}

hipError_t launch_publish(const void* const* src, int n, const void** dst, hipStream_t s) {
  for (int i0 = 0; i0 < n; i0 += kPublishMax) {
    PublishArgs a;
    memset(&a, 0, sizeof a);
    a.n = n - i0 < kPublishMax ? n - i0 : kPublishMax;
    a.dst = dst + i0;
    for (int i = 0; i < a.n; ++i) a.src[i] = src[i0 + i];
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(kPublishMax), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_reproject_ragged(const RaggedSpan<ReprojEntry>& sp, bool loop, hipStream_t s) {
  if (sp.n < 1 || sp.total < 1) return hipErrorInvalidValue;
  const auto k = loop ? &k_reproject_ragged<true> : &k_reproject_ragged<false>;
  const int grid = ragged_grid(reinterpret_cast<const void*>(k), kBlock, 0, sp, kBlock / 64, loop);
  hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, sp);
  return hipGetLastError();
}

hipError_t launch_inpaint_ragged(const RaggedSpan<InpaintEntry>& sp, bool loop, hipStream_t s) {
  if (sp.n < 1 || sp.total < 1) return hipErrorInvalidValue;
  const auto k = loop ? &k_inpaint_ragged<true> : &k_inpaint_ragged<false>;
  const int grid = ragged_grid(reinterpret_cast<const void*>(k), kBlock, 0, sp, 1, loop);
  hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, sp);
  return hipGetLastError();
}

int inpaint_tile() { return kTile; }

hipError_t launch_to_u8_ragged(const RaggedSpan<CastEntry>& sp, const float* const* src, bool loop, hipStream_t s) {
  if (sp.n < 1 || sp.total < 1) return hipErrorInvalidValue;
  const auto k = loop ? &k_to_u8_ragged<true> : &k_to_u8_ragged<false>;
  const int grid = ragged_grid(reinterpret_cast<const void*>(k), kBlock, 0, sp, kBlock / 64, loop);
  hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, sp, src);
  return hipGetLastError();
}

template <int RN>
hipError_t merge_ragged_rn(const RaggedSpan<MergeArgs>& sp, bool loop, hipStream_t s) {
  const auto k = loop ? &k_merge_ragged<true, RN> : &k_merge_ragged<false, RN>;
  const int grid = ragged_grid(reinterpret_cast<const void*>(k), kBlock, 0, sp, kBlock / 64, loop);
  hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, sp);
  return hipGetLastError();
}

hipError_t launch_merge_ragged(const RaggedSpan<MergeArgs>& sp, int rn, bool loop, hipStream_t s) {
  if (sp.n < 1 || sp.total < 1) return hipErrorInvalidValue;
  switch (rn) {
    case 1: return merge_ragged_rn<1>(sp, loop, s);
    case 2: return merge_ragged_rn<2>(sp, loop, s);
    case 3: return merge_ragged_rn<3>(sp, loop, s);
    case 4: return merge_ragged_rn<4>(sp, loop, s);
    case 0: return merge_ragged_rn<0>(sp, loop, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ss
