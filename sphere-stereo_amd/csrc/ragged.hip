// Ragged persistent executor: estimate_RGBD_panorama (depth_estimation.py:
// 205-243, stitcher.py:166-220) for a batch of frames whose rigs differ in
// matching resolution, candidate count, reference count, camera count,
// sigmas, inpainting settings and panorama size.  Every pass of the frame is
// one launch for the whole batch:
//
//   publish (per-run image pointers into the table)
//   sweep + guides          one launch per sweep variant the rigs are tuned
//                           to (ss_sweep_autotune), items = 64 x 4 tiles
//   cost filter             one launch per pyramid level (+ one for the
//                           single-workgroup coarse tails), the level-0 up
//                           pass fused with the argmin
//   distance filter         the same with G = 1
//   reprojection, inpainting (one launch per 16 passes), uint8 casts, merge
//
// A pyramid level k holds the level-k pass of every pair that has one (the
// pyramids of different resolutions have different depths and tail starts),
// one launch per candidate-lane count V present (each kernel has the
// register budget of its own V).  Each XCD takes an eighth of every
// pair, pair after pair (ss_ragged.h), so small and large pairs share
// launches and no pair waits for a launch of its own.  A launch is
// persistent (as many workgroups as the GPU holds resident, each worker
// looping over its items) or dispatched (a worker per item), per pass group
// as the caller's schedule asks.  The per-item arithmetic is the plain
// path's (the same device functions), so the results are the plain path's
// bit for bit.
//
// Everything but the image pointers is fixed at creation: the launch
// sequence, its records (in a device table the caller allocates: the library
// allocates nothing) and the weight tables of every pair's two filters,
// built once into the table.  A run publishes the image pointers into the
// table with one small kernel (captured by value into a graph like every
// other launch) and enqueues the passes.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

#include "ss_kernels.h"

namespace ss {

// Persistent form: the resident workgroups of the kernel (occupancy x CU
// count, computed once per kernel: the process is assumed to use one kind of
// GPU), no more than the items need, and at least one per part.
// Dispatched form: enough workgroups in each part for one worker per item.
int ragged_grid_of(const void* kernel, int block, size_t lds, const int* xtotal, int workers_per_block, bool loop) {
  long long most = 0;
  for (int x = 0; x < kParts; ++x) most = std::max(most, (long long)xtotal[x]);
  const long long per_part = std::max(1LL, (most + workers_per_block - 1) / workers_per_block);
  if (!loop) return (int)(per_part * kParts);
  static std::mutex mu;
  static std::vector<std::pair<const void*, int>> cache;
  int cap = 0;
  {
    std::lock_guard<std::mutex> lk(mu);
    for (const auto& c : cache)
      if (c.first == kernel) cap = c.second;
  }
  if (cap == 0) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, lds) != hipSuccess) per = 1;
    cap = std::max(1, per) * std::max(1, cus);
    std::lock_guard<std::mutex> lk(mu);
    cache.emplace_back(kernel, cap);
  }
  return (int)std::max((long long)kParts, std::min((long long)cap, per_part * kParts));
}

namespace {

// The table image: records are laid out in a host buffer at the offsets
// they take in the device table (base: the device address; nullptr while
// sizing).
struct Image {
  std::vector<char> h;
  char* base = nullptr;
  size_t off = 0;
  template <typename T>
  size_t take(size_t count, size_t align = 64) {
    off = (off + align - 1) & ~(align - 1);
    const size_t o = off;
    off += count * sizeof(T);
    if (h.size() < off) h.resize(off, 0);
    return o;
  }
  template <typename T>
  T* host(size_t o) {
    return reinterpret_cast<T*>(h.data() + o);
  }
  template <typename T>
  T* dev(size_t o) const {
    return base ? reinterpret_cast<T*>(base + o) : nullptr;
  }
};

template <typename E>
struct Pass {
  std::vector<E> e;
  std::vector<int> items;
  void add(const E& x, int n) {
    e.push_back(x);
    items.push_back(n);
  }
};

// Lays out a pass's records, the entries' item ranges and the XCD sequences
// of their chunks.
template <typename E>
RaggedSpan<E> place(Image& im, const Pass<E>& p) {
  RaggedSpan<E> sp{};
  sp.n = (int)p.e.size();
  if (sp.n == 0) return sp;
  const size_t oe = im.take<E>(p.e.size());
  const size_t of = im.take<int>(p.e.size() + 1);
  const size_t ox = im.take<int>((size_t)kParts * (p.e.size() + 1));
  memcpy(im.host<E>(oe), p.e.data(), sizeof(E) * p.e.size());
  int* first = im.host<int>(of);
  first[0] = 0;
  for (size_t k = 0; k < p.items.size(); ++k) first[k + 1] = first[k] + p.items[k];
  sp.total = first[sp.n];
  int* xf = im.host<int>(ox);
  for (int x = 0; x < kParts; ++x) {
    int* f = xf + x * (sp.n + 1);
    f[0] = 0;
    for (int k = 0; k < sp.n; ++k) {
      const long long n = p.items[k];
      f[k + 1] = f[k] + (int)((x + 1) * n / kParts - x * n / kParts);
    }
    sp.xtotal[x] = f[sp.n];
  }
  sp.e = im.dev<const E>(oe);
  sp.first = im.dev<const int>(of);
  sp.xfirst = im.dev<const int>(ox);
  return sp;
}

struct DownLaunch {
  RaggedSpan<IsbDownEntry> sp;
  int G, vt, vmax;
};
struct UpLaunch {
  RaggedSpan<IsbUpEntry> sp;
  int G, vt, vmax;
  bool fuse;
};
struct TailLaunch {
  const TailEntry* t = nullptr;
  int n = 0, G = 0;
  unsigned lds = 0;
};
struct Filter {
  std::vector<DownLaunch> down;
  TailLaunch tail;
  std::vector<UpLaunch> up;
};

}  // namespace

struct RaggedPlan {
  int F = 0, vt = 0;
  // per pass group: the persistent form (true) or the dispatched form
  bool loop_sweep = true, loop_filter = true, loop_stitch = true;
  int n_img = 0, n_stitch = 0;
  std::vector<int> img0, stitch0;  // per frame: first slot in the run tables
  const float** run_img = nullptr;
  const float** run_stitch = nullptr;
  struct SweepLaunch {
    RaggedSpan<SweepPair> sp;
    int variant;
  };
  std::vector<SweepLaunch> sweep;  // one launch per sweep variant the pairs are tuned to
  Filter cost, dist;
  RaggedSpan<ReprojEntry> reproj{};
  std::vector<RaggedSpan<InpaintEntry>> inpaint;
  RaggedSpan<CastEntry> cast{};
  // the merge: one launch per reference count present (rn 1-4; 0: more),
  // each kernel with the register budget of its own count
  struct MergeLaunch {
    RaggedSpan<MergeArgs> sp;
    int rn;
  };
  std::vector<MergeLaunch> merge;
  // creation only: the weight tables to build (vi, destination)
  std::vector<float> wt_vi;
  std::vector<float*> wt_dst;
  size_t bytes = 0;
};

namespace {

// One filter (cost: G = 8, volume type vt; distance: G = 1, fp32) over every
// pair: lv[p] the pair's levels, L[p] its depth, wtab[p] its weight table.
void plan_filter(Image& im, int G, int vt, const std::vector<const PyrLevel*>& lv,
                 const std::vector<int>& L, const std::vector<int>& V, const std::vector<float*>& wtab,
                 const std::vector<double>& sigma_s, const std::vector<float*>& dist,
                 const std::vector<ArgminParams>& am, const std::vector<float*>& fill, Filter* out) {
  const int P = (int)lv.size();
  std::vector<int> s0(P), last_down(P), first_up(P);
  int max_down = 0, max_up = -1;
  for (int p = 0; p < P; ++p) {
    s0[p] = tail_start(lv[p], L[p], G, G * V[p]);
    last_down[p] = s0[p] ? s0[p] : L[p] - 1;
    first_up[p] = last_down[p] - 1;
    max_down = std::max(max_down, last_down[p]);
    max_up = std::max(max_up, first_up[p]);
  }
  // one launch per level and V present at that level
  std::vector<int> classes;
  for (int p = 0; p < P; ++p)
    if (std::find(classes.begin(), classes.end(), ragged_vmax(V[p])) == classes.end())
      classes.push_back(ragged_vmax(V[p]));
  std::sort(classes.begin(), classes.end());
  for (int k = 1; k <= max_down; ++k) {
    for (int vm : classes) {
      Pass<IsbDownEntry> pass;
      for (int p = 0; p < P; ++p) {
        if (k > last_down[p] || ragged_vmax(V[p]) != vm) continue;
        const PyrLevel* v = lv[p];
        IsbDownEntry e{};
        e.a = IsbDownArgs{v[k - 1].g, v[k - 1].c, v[k - 1].h, v[k - 1].w, v[k].g, v[k].c, v[k].h, v[k].w, wtab[p]};
        e.V = V[p];
        pass.add(e, isb_items(v[k].h * v[k].w, G));
      }
      if (!pass.e.empty())
        out->down.push_back(DownLaunch{place(im, pass), G, k == 1 ? vt : SS_VOLUME_F32, vm});
    }
  }
  std::vector<TailEntry> tails;
  unsigned lds = 0;
  for (int p = 0; p < P; ++p) {
    if (!s0[p]) continue;
    TailEntry t{};
    lds = std::max(lds, tail_levels(lv[p], s0[p], L[p], G * V[p], wtab[p], sigma_s[p], &t.t));
    t.g0 = lv[p][s0[p]].g;
    t.c0 = lv[p][s0[p]].c;
    t.V = V[p];
    tails.push_back(t);
  }
  if (!tails.empty()) {
    const size_t o = im.take<TailEntry>(tails.size());
    memcpy(im.host<TailEntry>(o), tails.data(), sizeof(TailEntry) * tails.size());
    out->tail.t = im.dev<const TailEntry>(o);
    out->tail.n = (int)tails.size();
    out->tail.G = G;
    out->tail.lds = lds;
  }
  for (int k = max_up; k >= 0; --k) {
    const bool fuse = !dist.empty() && k == 0;
    for (int vm : classes) {
      Pass<IsbUpEntry> pass;
      for (int p = 0; p < P; ++p) {
        if (k > first_up[p] || ragged_vmax(V[p]) != vm) continue;
        const PyrLevel* v = lv[p];
        double wu, wd;
        up_weights(sigma_s[p], k, &wu, &wd);
        IsbUpEntry e{};
        e.a = IsbUpArgs{v[k + 1].g, v[k + 1].c, v[k + 1].h, v[k + 1].w, v[k].g, v[k].c, v[k].h, v[k].w,
                        (float)wu,  (float)wd,  wtab[p],    fuse ? dist[p] : nullptr, fuse ? am[p] : ArgminParams{},
                        fuse ? fill[p] : nullptr};
        e.V = V[p];
        pass.add(e, isb_items(v[k + 1].h * v[k + 1].w, G));
      }
      if (!pass.e.empty())
        out->up.push_back(UpLaunch{place(im, pass), G, k == 0 ? vt : SS_VOLUME_F32, vm, fuse});
    }
  }
}

// Lays out the whole plan; with im.base == nullptr it only measures.
int build(const ss_frame_desc* d, int F, int schedule, Image& im, RaggedPlan* pl) {
  // SS_RAGGED_AUTO: the form measured faster per pass group on MI355X (the
  // dispatched one for every group, on a 16-rig batch of 512^2..1024^2,
  // 16..64 candidates, 2..4 references)
  pl->loop_sweep = schedule == SS_RAGGED_PERSISTENT;
  pl->loop_filter = schedule == SS_RAGGED_PERSISTENT;
  pl->loop_stitch = schedule == SS_RAGGED_PERSISTENT;
  pl->F = F;
  pl->vt = d[0].volume_type;
  const int per = inpaint_steps_per_launch(), tile = inpaint_tile();
  int max_launches = 0;
  for (int f = 0; f < F; ++f) {
    const int K = (d[f].inpaint_iterations + per - 1) / per;
    max_launches = std::max(max_launches, K + (K & 1));
  }
  for (int f = 0; f < F; ++f) {
    pl->img0.push_back(pl->n_img);
    pl->stitch0.push_back(pl->n_stitch);
    pl->n_img += d[f].C;
    pl->n_stitch += d[f].R;
  }
  const size_t oimg = im.take<const float*>(pl->n_img), ost = im.take<const float*>(pl->n_stitch);
  pl->run_img = im.dev<const float*>(oimg);
  pl->run_stitch = im.dev<const float*>(ost);

  int P = 0;
  for (int f = 0; f < F; ++f) P += d[f].R;
  // the weight tables: pair p's cost filter (sigma_i) and distance filter
  // (sigma_i / 2, depth_estimation.py:201-202)
  const size_t owt = im.take<float>((size_t)2 * P * kWeightTable);

  std::vector<std::vector<PyrLevel>> lvs(P, std::vector<PyrLevel>(32)), dls(P, std::vector<PyrLevel>(32));
  std::vector<const PyrLevel*> lvp(P), dlp(P);
  std::vector<int> L(P), V(P), V1(P, 1);
  std::vector<float*> wc(P), wd(P), dist(P), fill(P);
  std::vector<double> ss_cost(P), ss_dist(P);
  std::vector<ArgminParams> am(P);
  Pass<SweepPair> sweep[4];  // by the configuration's tuned sweep variant (ss_sweep_get_variant)
  Pass<ReprojEntry> reproj;
  Pass<CastEntry> cast;
  Pass<MergeArgs> merge[5];  // by reference count: 1-4, and 0 for more
  int p = 0;
  for (int f = 0; f < F; ++f) {
    const ss_frame_desc& c = d[f];
    TorchCam cams[SS_MAX_CAMERAS];
    for (int k = 0; k < c.C; ++k) cams[k] = make_torch_cam(c.cams[k]);
    for (int r = 0; r < c.R; ++r, ++p) {
      L[p] = estimate_levels(c.d_workspace[r], c.H, c.W, c.N, c.volume_type, c.d_dist[r], lvs[p].data(),
                             dls[p].data());
      lvp[p] = lvs[p].data();
      dlp[p] = dls[p].data();
      V[p] = lanes_v(c.N);
      wc[p] = im.dev<float>(owt + sizeof(float) * (size_t)(2 * p) * kWeightTable);
      wd[p] = im.dev<float>(owt + sizeof(float) * (size_t)(2 * p + 1) * kWeightTable);
      pl->wt_vi.push_back((float)(1 / (2 * c.sigma_i * c.sigma_i)));
      pl->wt_dst.push_back(wc[p]);
      pl->wt_vi.push_back((float)(1 / (2 * (c.sigma_i / 2) * (c.sigma_i / 2))));
      pl->wt_dst.push_back(wd[p]);
      ss_cost[p] = c.sigma_s;
      ss_dist[p] = c.sigma_s / 2;
      dist[p] = c.d_dist[r];
      fill[p] = c.d_reproj[r];
      const float d0 = c.cands[0], dN1 = c.cands[c.N - 1];
      am[p] = ArgminParams{c.N, d0, dN1, (float)(d0 / dN1) - 1.0f};

      SweepPair sp{};
      sp.raylut = reinterpret_cast<const float4*>(c.d_raylut[r]);
      sp.cost = lvs[p][0].c;
      sp.guide = lvs[p][0].g;
      sp.img0 = pl->img0[f];
      sp.refcam = c.refs[r];
      sp.H = c.H;
      sp.W = c.W;
      sp.N = c.N;
      sp.NP = 8 * V[p];
      sp.tiles_x = (c.W + 63) / 64;
      sp.pow2 = (c.W & (c.W - 1)) == 0 && (c.H & (c.H - 1)) == 0;
      sp.Wf = (float)c.W;
      sp.Hf = (float)c.H;
      sp.invW = 1.0f / (float)c.W;
      sp.invH = 1.0f / (float)c.H;
      sp.inv2W = 2.0f / (float)c.W;
      sp.inv2H = 2.0f / (float)c.H;
      sp.halfW = (float)c.W * 0.5f;
      sp.halfH = (float)c.H * 0.5f;
      for (int k = 0; k < c.C; ++k) {
        sp.cam[k] = cams[k];
        memcpy(sp.rt[k], c.rts[r] + 12 * k, sizeof(float) * 12);
      }
      memcpy(sp.cands, c.cands, sizeof(float) * c.N);
      sp.fastproj = sweep_fast_bounds(cams, c.C, c.rts[r], c.C, c.cands, c.N);
      const int var = ss_sweep_get_variant(c.H, c.W, c.N, c.volume_type);
      sweep[var >= 0 && var < 4 ? var : 0].add(sp, sweep_items(c.H, c.W));

      ReprojEntry re{};
      re.a.din = c.d_dist[r];
      re.a.dout = c.d_reproj[r];
      re.a.c = c.intr[r];
      memcpy(re.a.t, c.trans[r], sizeof re.a.t);
      re.H = c.H;
      re.W = c.W;
      reproj.add(re, pix_items((long long)c.H * c.W));

      CastEntry ce{};
      ce.src = pl->stitch0[f] + r;
      ce.out = c.d_stitch_u8[r];
      ce.n = (long long)c.Hs * c.Ws * 3;
      if (ce.n > 0) cast.add(ce, cast_items(ce.n));
    }
    MergeArgs m{};
    m.sampling = reinterpret_cast<const float2*>(c.d_sampling);
    m.weights = c.d_weights;
    for (int r = 0; r < c.R; ++r) {
      m.reproj[r] = c.d_reproj[r];
      m.dists[r] = c.d_dist[r];
      m.imgs[r] = c.d_stitch_u8[r];
      m.intr[r] = c.intr[r];
      memcpy(m.trans[r], c.trans[r], sizeof(float) * 3);
    }
    m.R = c.R;
    m.H = c.H;
    m.W = c.W;
    m.Hs = c.Hs;
    m.Ws = c.Ws;
    m.PH = c.PH;
    m.PW = c.PW;
    m.dist_out = c.d_dist_pano;
    m.rgb_out = c.d_rgb_pano;
    merge[c.R >= 1 && c.R <= 4 ? c.R : 0].add(m, pix_items((long long)c.PH * c.PW));
  }
  for (int v = 0; v < 4; ++v)
    if (!sweep[v].e.empty()) pl->sweep.push_back(RaggedPlan::SweepLaunch{place(im, sweep[v]), v});
  plan_filter(im, 8, pl->vt, lvp, L, V, wc, ss_cost, dist, am, fill, &pl->cost);
  plan_filter(im, 1, SS_VOLUME_F32, dlp, L, V1, wd, ss_dist, {}, {}, {}, &pl->dist);
  pl->reproj = place(im, reproj);
  // inpainting: launch j runs passes 16j.. of every map that has them, from
  // the map into the scratch buffer for even j, back for odd j; a map with
  // an odd launch count gets a copy back (steps = 0) in launch K
  for (int j = 0; j < max_launches; ++j) {
    Pass<InpaintEntry> pass;
    p = 0;
    for (int f = 0; f < F; ++f) {
      const ss_frame_desc& c = d[f];
      const int K = (c.inpaint_iterations + per - 1) / per;
      for (int r = 0; r < c.R; ++r, ++p) {
        const bool work = j < K, copy = (K & 1) && j == K;
        if (!work && !copy) continue;
        InpaintEntry e{};
        e.a.din = (j & 1) ? c.d_scratch[r] : c.d_reproj[r];
        e.a.dout = (j & 1) ? c.d_reproj[r] : c.d_scratch[r];
        e.a.lut = c.d_inpaint_lut[r];
        e.H = c.H;
        e.W = c.W;
        e.steps = work ? std::min(per, c.inpaint_iterations - per * j) : 0;
        e.tiles_x = (c.W + tile - 1) / tile;
        e.thr_empty = c.thr_empty;
        e.thr_src = c.thr_src;
        pass.add(e, e.tiles_x * ((c.H + tile - 1) / tile));
      }
    }
    if (!pass.e.empty()) pl->inpaint.push_back(place(im, pass));
  }
  pl->cast = place(im, cast);
  for (int rn : {1, 2, 3, 4, 0})
    if (!merge[rn].e.empty()) pl->merge.push_back(RaggedPlan::MergeLaunch{place(im, merge[rn]), rn});
  pl->bytes = im.off;
  return SS_OK;
}

#define RG_HIP(expr)                                                            \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      char m_[256];                                                             \
      snprintf(m_, sizeof m_, "%s: %s", #expr, hipGetErrorString(e_));          \
      return report_error(SS_ERR_HIP, m_);                                      \
    }                                                                           \
  } while (0)

int check_ragged(const ss_frame_desc* d, int F) {
  for (int f = 0; f < F; ++f) {
    if (d[f].volume_type != d[0].volume_type)
      return report_error(SS_ERR_ARG, "ragged batch: the frames' volume types differ");
    if (d[f].N < 2) return report_error(SS_ERR_UNSUPPORTED, "ragged batch: every frame needs N >= 2 candidates");
    if (isb_levels(d[f].H, d[f].W) < 2)
      return report_error(SS_ERR_UNSUPPORTED, "ragged batch: matching resolutions below 8 x 8 are not supported");
    if (lanes_v(d[f].N) < 0) return report_error(SS_ERR_ARG, "ragged batch: unsupported candidate count");
  }
  return SS_OK;
}

}  // namespace

size_t ragged_table_bytes(const ss_frame_desc* d, int F) {
  if (check_ragged(d, F) != SS_OK) return 0;
  Image im;
  RaggedPlan pl;
  return build(d, F, SS_RAGGED_PERSISTENT, im, &pl) == SS_OK ? pl.bytes : 0;
}

int ragged_create(const ss_frame_desc* d, int F, int schedule, void* table, size_t bytes, hipStream_t s,
                  RaggedPlan** out) {
  int rc = check_ragged(d, F);
  if (rc) return rc;
  if (schedule != SS_RAGGED_PERSISTENT && schedule != SS_RAGGED_DISPATCHED && schedule != SS_RAGGED_AUTO)
    return report_error(SS_ERR_ARG, "ragged batch: unknown schedule");
  if (!table || (reinterpret_cast<uintptr_t>(table) & 255))
    return report_error(SS_ERR_ARG, "ragged batch: the table must be a 256-byte aligned device buffer");
  RaggedPlan* pl = new (std::nothrow) RaggedPlan();
  if (!pl) return report_error(SS_ERR_ARG, "out of host memory");
  Image im;
  im.base = static_cast<char*>(table);
  auto bail = [&](int code) {
    delete pl;
    return code;
  };
  rc = build(d, F, schedule, im, pl);
  if (rc) return bail(rc);
  if (bytes < pl->bytes) {
    char m[128];
    snprintf(m, sizeof m, "ragged batch: table of %zu bytes needed, %zu given", pl->bytes, bytes);
    return bail(report_error(SS_ERR_WORKSPACE, m));
  }
  // the records (and zeroed pointer slots), then the weight tables on top
  if (hipMemcpyAsync(table, im.h.data(), pl->bytes, hipMemcpyHostToDevice, s) != hipSuccess)
    return bail(report_error(SS_ERR_HIP, "ragged batch: table upload failed"));
  for (size_t i = 0; i < pl->wt_vi.size(); i += kMaxBatch) {
    const int n = (int)std::min(pl->wt_vi.size() - i, (size_t)kMaxBatch);
    if (launch_weight_tables(pl->wt_vi.data() + i, pl->wt_dst.data() + i, n, s) != hipSuccess)
      return bail(report_error(SS_ERR_HIP, "ragged batch: weight tables failed"));
  }
  if (hipStreamSynchronize(s) != hipSuccess)
    return bail(report_error(SS_ERR_HIP, "ragged batch: table set-up failed"));
  *out = pl;
  return SS_OK;
}

int ragged_enqueue(const RaggedPlan* pl, const float* const* match, const float* const* stitch, hipStream_t s) {
  RG_HIP(launch_publish(reinterpret_cast<const void* const*>(match), pl->n_img,
                        reinterpret_cast<const void**>(pl->run_img), s));
  RG_HIP(launch_publish(reinterpret_cast<const void* const*>(stitch), pl->n_stitch,
                        reinterpret_cast<const void**>(pl->run_stitch), s));
  for (const auto& l : pl->sweep) RG_HIP(launch_sweep_ragged(l.sp, pl->run_img, pl->vt, l.variant, pl->loop_sweep, s));
  const bool lf = pl->loop_filter, ls = pl->loop_stitch;
  for (const Filter* fl : {&pl->cost, &pl->dist}) {
    for (const DownLaunch& l : fl->down) RG_HIP(launch_isb_down_ragged(l.sp, l.G, l.vt, l.vmax, lf, s));
    if (fl->tail.n) RG_HIP(launch_isb_tail_ragged(fl->tail.t, fl->tail.n, fl->tail.G, fl->tail.lds, s));
    for (const UpLaunch& l : fl->up) RG_HIP(launch_isb_up_ragged(l.sp, l.G, l.vt, l.vmax, l.fuse, lf, s));
  }
  RG_HIP(launch_reproject_ragged(pl->reproj, ls, s));
  for (const auto& sp : pl->inpaint) RG_HIP(launch_inpaint_ragged(sp, ls, s));
  if (pl->cast.n) RG_HIP(launch_to_u8_ragged(pl->cast, pl->run_stitch, ls, s));
  for (const auto& l : pl->merge) RG_HIP(launch_merge_ragged(l.sp, l.rn, ls, s));
  return SS_OK;
}

int ragged_images(const RaggedPlan* pl, int* n_match, int* n_stitch) {
  *n_match = pl->n_img;
  *n_stitch = pl->n_stitch;
  return SS_OK;
}

void ragged_free(RaggedPlan* pl) { delete pl; }

}  // namespace ss
