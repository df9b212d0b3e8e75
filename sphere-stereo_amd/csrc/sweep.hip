// Sphere sweep, camera selection and guide construction on gfx950.
//
// Behaviour: depth_estimation.py:83-175 and utils.py:57-94, 144-155 of the
// reference.  The reference evaluates the sweep as a chain of full-volume
// torch ops per camera (~10 GB of HBM traffic per camera per reference at
// 1024^2, 32 candidates); here one pass writes the candidate-contiguous cost
// volume directly, one lane per reference pixel sweeping all candidates.
// Only the selected camera of each pixel is sampled (the reference samples
// every camera and keeps the selected one), which leaves the result
// unchanged.
#include "ss_kernels.h"
#include "ss_ragged.h"

namespace ss {

namespace {

constexpr int kBlock = 256;

inline int blocks_for(long long threads) { return (int)((threads + kBlock - 1) / kBlock); }

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

// Sweep + raw cost (depth_estimation.py:141-175) for every reference of the
// launch: blocks [r*blocks_per_ref, (r+1)*blocks_per_ref) sweep reference r.
// One lane per reference pixel, over a 64 x 4 pixel tile per workgroup (one
// wavefront per row segment).  Every lane of a wavefront samples the same
// candidate at the same time, so the 64 lanes' bilinear taps fall on
// neighbouring texels of the sampled image (coalesced gathers).  Costs go
// to an LDS block of 64 pixels x 32 candidates per wavefront, which the
// wavefront then writes out as whole lines (8 lanes per pixel's 32
// candidates): stored straight from the lanes, 4 candidates at a time, each
// store would touch 64 lines partially, and lines evicted half-written
// multiplied the HBM write traffic about three times.  The lane also writes
// its pixel's guide (utils.py:144-155), which needs the same reference colour.
//
// Two scheduling parameters are template arguments, picked per configuration
// by ss_sweep_autotune: CH candidates are projected before their taps are
// loaded (more loads in flight per lane against more registers), and GRP
// candidates are staged per flush (LDS per workgroup, hence occupancy,
// against flush count).  The arithmetic of a sample does not depend on them.
constexpr int kTileW = 64, kTileH = kBlock / kTileW;

// One wavefront's 64-pixel row segment (y, x0..x0+63) of a reference: P
// holds the sizes, candidates and grid_sample constants (SweepArgs or
// SweepPair), cams / rts the cameras and the reference's transforms into
// them, img_of(k) camera k's image; st is the wavefront's LDS block.
template <typename T, bool POW2, int CH, int GRP, typename P, typename ImgOf>
__device__ __forceinline__ void sweep_row(const P& a, const float4* __restrict__ raylut,
                                          const float* __restrict__ refimg, void* cost, uint32_t* guide,
                                          const TorchCam* cams, const float (*rts)[12], ImgOf img_of, int y, int x0,
                                          int lane, float* st) {
  constexpr int kChunk = CH, kGroup = GRP;
  constexpr int kStride = kGroup + 1;  // LDS words per pixel (conflict-free)
  // lanes past the right edge compute on a valid pixel and store nothing, so
  // that the whole wavefront takes part in the flush
  const bool active = x0 + lane < a.W;
  const int x = active ? x0 + lane : a.W - 1;
  const int pix = y * a.W + x;
  const float4 rl = raylut[pix];
  const int k = (int)rl.w;
  const float r0 = refimg[3 * pix], r1 = refimg[3 * pix + 1], r2 = refimg[3 * pix + 2];
  if (guide && active) {
    float gy = ((16.0f + 0.1826f * r0) + 0.6142f * r1) + 0.062f * r2;
    float cb = ((128.0f - 0.1006f * r0) - 0.3386f * r1) + 0.4392f * r2;
    float cr = ((128.0f + 0.4392f * r0) - 0.3989f * r1) - 0.0403f * r2;
    gy = gy < 16.0f ? 16.0f : (gy > 235.0f ? 235.0f : gy);
    cb = cb < 16.0f ? 16.0f : (cb > 240.0f ? 240.0f : cb);
    cr = cr < 16.0f ? 16.0f : (cr > 240.0f ? 240.0f : cr);
    guide[pix] = pack_guide(gy, cb, cr);
  }
  // register copies: read through pointers, the camera and transform would
  // be reloaded after every store (the compiler cannot rule out aliasing)
  const int kk = k < 0 ? 0 : k;
  const TorchCam cam = cams[kk];
  float rt[12];
#pragma unroll
  for (int i = 0; i < 12; ++i) rt[i] = rts[kk][i];
  const float* img = img_of(kk);
  // nothing sampled: the volume holds zeros there (depth_estimation.py:150-152)
  float cnone = (r0 + r1) + r2;
  cnone = cnone > 500.0f ? 500.0f : cnone;
  const float wlim = (float)(a.W - 1), hlim = (float)(a.H - 1);
  T* const row = static_cast<T*>(cost) + (size_t)(y * a.W + x0) * a.NP;

  for (int g0 = 0; g0 < a.NP; g0 += kGroup) {
    const int cnt = a.NP - g0 < kGroup ? a.NP - g0 : kGroup;  // a multiple of 8
#pragma unroll 1
    for (int j = 0; j < cnt; j += kChunk) {
      const int n0 = g0 + j;
      float out[kChunk];
      if (k < 0) {
#pragma unroll
        for (int v = 0; v < kChunk; ++v) out[v] = (n0 + v < a.N) ? cnone : 0.0f;
      } else {
        // sample coordinates of the chunk's candidates (pads sample a
        // dummy interior point and are stored as 0)
        float ix[kChunk], iy[kChunk];
        bool inner = true;
#pragma unroll
        for (int v = 0; v < kChunk; ++v) {
          ix[v] = 0.5f;
          iy[v] = 0.5f;
          if (n0 + v < a.N) {
            const float d = a.cands[n0 + v];
            float qx, qy, qz, u, w;
            bool valid;
            rigid(rt, d * rl.x, d * rl.y, d * rl.z, qx, qy, qz);
            // the fast form where it is certainly exact, else the full one
            if (!a.fastproj || !project_t_fast(cam, qx, qy, qz, u, w)) project_t(cam, qx, qy, qz, u, w, valid);
            if constexpr (POW2) {
              ix[v] = grid_coord_pow2(u, a.inv2W, a.halfW);
              iy[v] = grid_coord_pow2(w, a.inv2H, a.halfH);
            } else {
              ix[v] = grid_coord2<false>(u, a.Wf, a.invW);
              iy[v] = grid_coord2<false>(w, a.Hf, a.invH);
            }
          }
          const float x0f = floorf(ix[v]), y0f = floorf(iy[v]);
          inner = inner && x0f >= 0.0f && x0f < wlim && y0f >= 0.0f && y0f < hlim;
        }
        if (inner) {
          // all taps of the chunk inside the image: issue every load, then combine
          // with the arithmetic of sample_zero (no tap clamped or masked)
          float tn[kChunk][6], ts[kChunk][6];
#pragma unroll
          for (int v = 0; v < kChunk; ++v) {
            const float* pn = img + (unsigned)((int)floorf(iy[v]) * a.W + (int)floorf(ix[v])) * 3u;
            const float* ps = pn + (unsigned)a.W * 3u;
#pragma unroll
            for (int i = 0; i < 6; ++i) {
              tn[v][i] = pn[i];
              ts[v][i] = ps[i];
            }
          }
#pragma unroll
          for (int v = 0; v < kChunk; ++v) {
            const float x0f = floorf(ix[v]), y0f = floorf(iy[v]);
            const float x1f = x0f + 1.0f, y1f = y0f + 1.0f;
            const float wnw = (x1f - ix[v]) * (y1f - iy[v]), wne = (ix[v] - x0f) * (y1f - iy[v]);
            const float wsw = (x1f - ix[v]) * (iy[v] - y0f), wse = (ix[v] - x0f) * (iy[v] - y0f);
            float sv[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              float tt = tn[v][c] * wnw;
              tt = tt + tn[v][3 + c] * wne;
              tt = tt + ts[v][c] * wsw;
              tt = tt + ts[v][3 + c] * wse;
              sv[c] = tt;
            }
            float c = (fabsf(sv[0] - r0) + fabsf(sv[1] - r1)) + fabsf(sv[2] - r2);
            c = c > 500.0f ? 500.0f : c;
            out[v] = (n0 + v < a.N) ? c : 0.0f;
          }
        } else {
#pragma unroll
          for (int v = 0; v < kChunk; ++v) {
            float c = 0.0f;
            if (n0 + v < a.N) {
              float sv[3];
              sample_zero<3>(img, a.H, a.W, ix[v], iy[v], sv);
              c = (fabsf(sv[0] - r0) + fabsf(sv[1] - r1)) + fabsf(sv[2] - r2);
              c = c > 500.0f ? 500.0f : c;
            }
            out[v] = c;
          }
        }
      }
#pragma unroll
      for (int v = 0; v < kChunk; ++v) st[lane * kStride + j + v] = out[v];
    }
    // the wavefront's block: all its lanes' LDS writes precede the reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // flush: quad q = 4 consecutive candidates of one pixel; lanes take
    // consecutive quads, so a store instruction covers whole lines
    const int qpp = cnt / 4;  // quads per pixel
    for (int q = lane; q < kTileW * qpp; q += kTileW) {
      const int p = q / qpp, c4 = q - p * qpp;
      if (x0 + p < a.W) {
        const float* src = st + p * kStride + 4 * c4;
        const float vals[4] = {src[0], src[1], src[2], src[3]};
        store_n<4>(row + (size_t)p * a.NP + g0 + 4 * c4, vals);
      }
    }
    // the next group overwrites the block
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// The trailing cast workgroups of a sweep launch: block b of `blocks` takes
// every blocks-th group of 256 quads, each lane one quad (a 16-byte load and
// a 4-byte store; the truncating conversion of the stitcher's k_to_u8).
__device__ __forceinline__ void sweep_cast(const SweepCast& c, int b) {
  const unsigned total = c.quads * (unsigned)c.n;
  for (unsigned q = (unsigned)b * kBlock + threadIdx.x; q < total; q += (unsigned)c.blocks * kBlock) {
    const int img = (int)(q / c.quads);
    const size_t i = (size_t)(q - (unsigned)img * c.quads) * 4;
    const float4 v = *reinterpret_cast<const float4*>(c.in[img] + i);
    const uint32_t w = (uint32_t)(uint8_t)v.x | ((uint32_t)(uint8_t)v.y << 8) | ((uint32_t)(uint8_t)v.z << 16) |
                       ((uint32_t)(uint8_t)v.w << 24);
    *reinterpret_cast<uint32_t*>(c.out[img] + i) = w;
  }
}

template <typename T, bool POW2, int CH, int GRP>
__global__ void __launch_bounds__(kBlock) k_sweep(SweepArgs a) {
  __shared__ float stage[kTileH][kTileW * (GRP + 1)];
  // the sweep's tiles first (XCD-aware order over them), then the cast blocks
  const int nsweep = (int)gridDim.x - a.cast.blocks;
  if ((int)blockIdx.x >= nsweep) {
    sweep_cast(a.cast, (int)blockIdx.x - nsweep);
    return;
  }
  const int tile = xcd_tile(blockIdx.x, nsweep);
  const int r = tile / a.blocks_per_ref;
  const int t = tile - r * a.blocks_per_ref;
  const int ty = t / a.tiles_x, tx = t - ty * a.tiles_x;
  const int lane = (int)(threadIdx.x & (kTileW - 1)), wv = (int)(threadIdx.x / kTileW);
  const int y = ty * kTileH + wv;
  if (y >= a.H) return;  // a whole wavefront
  const SweepRef& rf = a.ref[r];
  sweep_row<T, POW2, CH, GRP>(
      a, rf.raylut, rf.img, rf.cost, rf.guide, a.cam, a.rt + rf.rt0, [&](int k) { return a.img[k]; }, y,
      tx * kTileW, lane, stage[wv]);
}

// The uniform part of a SweepPair as sweep_row reads it (in registers; the
// candidates through the constant address space).
#if defined(__HIP_DEVICE_COMPILE__)
using ConstFloatPtr = const SS_CONST_AS float*;
using ConstPairPtr = const SS_CONST_AS SweepPair*;
#else
using ConstFloatPtr = const float*;
using ConstPairPtr = const SweepPair*;
#endif
struct RowGeom {
  int H, W, N, NP, fastproj;
  float Wf, Hf, invW, invH, inv2W, inv2H, halfW, halfH;
  ConstFloatPtr cands;
};

// The ragged form: items are the 64 x 4 tiles of every pair of the span
// (pair-major, then tile row, then tile), a wavefront per tile row as in
// k_sweep; img is the run's image table.
template <typename T, bool LOOP, int CH, int GRP>
__global__ void __launch_bounds__(kBlock) k_sweep_ragged(RaggedSpan<SweepPair> sp, const float* const* img) {
  __shared__ float stage[kTileH][kTileW * (GRP + 1)];
  const int lane = (int)(threadIdx.x & (kTileW - 1)), wv = (int)(threadIdx.x / kTileW);
  for_each_block_item<LOOP>(sp, [&](int e, int t) {
    const SweepPair* pp = sp.e + e;
    const ConstPairPtr pc = (ConstPairPtr)pp;
    const RowGeom g{pc->H,     pc->W,    pc->N,     pc->NP,    pc->fastproj, pc->Wf,    pc->Hf,
                    pc->invW,  pc->invH, pc->inv2W, pc->inv2H, pc->halfW,    pc->halfH, pc->cands};
    const int tiles_x = pc->tiles_x;
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    const int y = ty * kTileH + wv;
    if (y >= g.H) return;  // this wavefront's row of the tile is past the image
    const float* const* im = img + pc->img0;
    const float* refimg = im[pc->refcam];
    const float4* raylut = pc->raylut;
    void* cost = pc->cost;
    uint32_t* guide = pc->guide;
    auto img_of = [&](int k) { return im[k]; };
    if (pc->pow2)
      sweep_row<T, true, CH, GRP>(g, raylut, refimg, cost, guide, pp->cam, pp->rt, img_of, y, tx * kTileW, lane,
                                  stage[wv]);
    else
      sweep_row<T, false, CH, GRP>(g, raylut, refimg, cost, guide, pp->cam, pp->rt, img_of, y, tx * kTileW, lane,
                                   stage[wv]);
  });
}

template <typename T, int CH, int GRP>
void launch_sweep_v(const SweepArgs& a, hipStream_t s) {
  const dim3 grid(a.blocks_per_ref * a.R + a.cast.blocks);
  if (a.pow2)
    hipLaunchKernelGGL((k_sweep<T, true, CH, GRP>), grid, dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_sweep<T, false, CH, GRP>), grid, dim3(kBlock), 0, s, a);
}

// the variants: (candidate chunk, flush group)
template <typename T>
hipError_t launch_sweep_t(const SweepArgs& a, int variant, hipStream_t s) {
  switch (variant) {
    case 1: launch_sweep_v<T, 2, 32>(a, s); break;
    case 2: launch_sweep_v<T, 4, 16>(a, s); break;
    case 3: launch_sweep_v<T, 2, 16>(a, s); break;
    default: launch_sweep_v<T, 4, 32>(a, s); break;
  }
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

int sweep_variant_count() { return 4; }
void sweep_variant_info(int v, int* chunk, int* group) {
  static const int kInfo[4][2] = {{4, 32}, {2, 32}, {4, 16}, {2, 16}};
  *chunk = kInfo[v][0];
  *group = kInfo[v][1];
}

template <typename T, bool LOOP, int CH, int GRP>
hipError_t sweep_ragged_v(const RaggedSpan<SweepPair>& sp, const float* const* img, hipStream_t s) {
  const auto k = &k_sweep_ragged<T, LOOP, CH, GRP>;
  const int grid = ragged_grid(reinterpret_cast<const void*>(k), kBlock, 0, sp, 1, LOOP);
  hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, sp, img);
  return hipGetLastError();
}

// the scheduling variants of k_sweep (launch_sweep_t)
template <typename T, bool LOOP>
hipError_t sweep_ragged_t(const RaggedSpan<SweepPair>& sp, const float* const* img, int variant, hipStream_t s) {
  switch (variant) {
    case 1: return sweep_ragged_v<T, LOOP, 2, 32>(sp, img, s);
    case 2: return sweep_ragged_v<T, LOOP, 4, 16>(sp, img, s);
    case 3: return sweep_ragged_v<T, LOOP, 2, 16>(sp, img, s);
    default: return sweep_ragged_v<T, LOOP, 4, 32>(sp, img, s);
  }
}

hipError_t launch_sweep_ragged(const RaggedSpan<SweepPair>& sp, const float* const* img, int vt, int variant,
                               bool loop, hipStream_t s) {
  if (sp.n < 1 || sp.total < 1) return hipErrorInvalidValue;
#define SS_CALL(T) (loop ? sweep_ragged_t<T, true>(sp, img, variant, s) : sweep_ragged_t<T, false>(sp, img, variant, s))
  SS_DISPATCH_T(vt, SS_CALL)
#undef SS_CALL
}

hipError_t launch_sweep(const SweepArgs& a, int vt, int variant, hipStream_t s) {
#define SS_CALL(T) launch_sweep_t<T>(a, variant, s)
  SS_DISPATCH_T(vt, SS_CALL)
#undef SS_CALL
}

}  // namespace ss
