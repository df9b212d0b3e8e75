// Per-frame executor: estimate_RGBD_panorama (depth_estimation.py:205-243 and
// stitcher.py:166-220) as one call, for one frame or a batch of frames with
// one configuration.  Every pass runs once for all of the frame's references
// (of all the batch's frames, in chunks of up to 16 pairs): one sweep launch writes every reference's volume and
// guide, and each later pass (cost filter levels, fused argmin, distance
// filter, uint8 cast, reprojection, inpainting) is one launch whose grid
// covers every reference.  The full-size passes then fill the GPU with the
// work of all references at once, and the latency-bound coarse-level passes
// cost one launch per frame instead of one per reference.  (Per-reference
// chains on separate streams were measured to run one after the other: the
// second chain started about 250 us after the sweep, well after its work had
// been queued.  A variant with the references in two groups, each with its
// own sweep on its own stream, ran 3 % faster, but only because the command
// processor drained one queue for a few hundred microseconds before it
// served the other, so the overlap it got was incidental, and timing a
// kernel with events then counted the waits on the other queue.)  The frame
// can be captured once into a HIP graph and replayed, which removes the
// per-launch host cost of its ~30 kernels.  A batch can spread its chunks of
// frames over up to 4 lanes (streams): the sweeps, bound by vector-ALU
// issue, run one chunk after the other, and each chunk's HBM- and
// latency-bound passes run beside the next chunk's sweeps (8 frames of
// 1024^2 in two chunks of 4: about 5 % more panoramas/s than one stream).
// Chunks hold whole frames: cutting one frame's references over two lanes
// ran 4 % slower than one stream, so a one-frame executor has one lane.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

#include "ss_kernels.h"

namespace {

// One capture stream per device, shared by every executor on it, created on
// first use (default priority) and kept until the process ends.  The mutex
// serialises the enqueueing (host work only) of frames on the device, and
// with it the use of the capture stream.
constexpr int kMaxDevices = 64;
struct StreamPool {
  std::once_flag once;
  hipError_t err = hipSuccess;
  hipStream_t capture = nullptr;
  std::mutex mu;
};
StreamPool g_pool[kMaxDevices];

hipError_t pool_init(StreamPool& p) {
  std::call_once(p.once, [&p] { p.err = hipStreamCreateWithFlags(&p.capture, hipStreamNonBlocking); });
  return p.err;
}

// Lane streams of one executor (ss_frame_set_lanes): at most as many as the
// hardware queues HIP gives a process by default.
constexpr int kMaxLanes = 4;

// A batch of F frames (F = 1 for ss_frame_create) with one configuration:
// the rigs, tables, buffers and images of the frames differ.
struct Frame {
  std::vector<ss_frame_desc> d;
  StreamPool* pool = nullptr;
  int dev = 0;
  int lanes = 1;
  hipStream_t lane[kMaxLanes] = {};
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> swept, done;  // per chunk
  // graph capture: one stream per chunk (chunk k on lane k mod L would make
  // the lanes wait on each other in a cycle once a lane takes two chunks,
  // which stream capture cannot record)
  std::vector<hipStream_t> cap;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const float*> cap_match, cap_stitch;
  ss::RaggedPlan* ragged = nullptr;  // the ragged persistent form (ss_frame_ragged_create)
  // entry-0 workspaces of chunks whose guide weight tables (which depend on
  // sigma_i alone) an eager run has computed: later runs skip that launch.
  // The descriptors' workspaces are fixed for the executor's lifetime; a
  // capture computes the tables inside the graph and marks nothing.
  std::vector<const void*> tables_done;
  size_t n_match = 0, n_stitch = 0;  // images per run
};

#define FR_HIP(expr)                                                            \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      char m_[256];                                                             \
      snprintf(m_, sizeof m_, "%s: %s", #expr, hipGetErrorString(e_));          \
      return ss::report_error(SS_ERR_HIP, m_);                                  \
    }                                                                           \
  } while (0)

#define FR_CALL(expr)         \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != SS_OK) return rc_; \
  } while (0)

// Frames per chunk.  With one lane, as many as fit kMaxBatch frame-reference
// pairs.  With L lanes the batch is cut into at least L chunks (or one per
// frame when there are fewer than L frames).  A chunk always holds whole
// frames: a frame's references stay in one chunk, so its merge waits for
// nothing but its own chunk.
int chunk_frames(int F, int R, int L) {
  const int per = ss::kMaxBatch / R;  // frames per launch (R <= kMaxBatch)
  return std::min(per, L <= 1 ? F : (F + L - 1) / L);
}

int ensure_capture_streams(Frame* f, size_t n) {
  while (f->cap.size() < n) {
    hipStream_t st = nullptr;
    FR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    f->cap.push_back(st);
  }
  return SS_OK;
}

int ensure_events(Frame* f, size_t n) {
  while (f->swept.size() < n) {
    hipEvent_t a = nullptr, b = nullptr;
    FR_HIP(hipEventCreateWithFlags(&a, hipEventDisableTiming));
    const hipError_t e = hipEventCreateWithFlags(&b, hipEventDisableTiming);
    if (e != hipSuccess) {
      (void)hipEventDestroy(a);
      FR_HIP(e);
    }
    f->swept.push_back(a);
    f->done.push_back(b);
  }
  return SS_OK;
}

// The events (and, for a capture, the streams) the chunks of a run use,
// created before the run (not while a stream captures).
int prepare_lanes(Frame* f, bool capturing) {
  if (f->ragged || f->lanes <= 1) return SS_OK;
  const int F = (int)f->d.size();
  const int cf = chunk_frames(F, f->d[0].R, f->lanes);
  const size_t K = (size_t)((F + cf - 1) / cf);
  FR_CALL(ensure_events(f, K));
  if (capturing) FR_CALL(ensure_capture_streams(f, K));
  return SS_OK;
}

// The batch's work, enqueued after the work already on `s`, which waits for
// all of it.  Each chunk of frames runs its frames' sweeps (one launch per
// frame for its references), then every later pass once for the whole
// chunk, then each frame's merge.  A chunk's volumes are produced and
// consumed together, so its passes meet them in the last-level cache where
// they fit.  With one lane every chunk runs on `s`.  With L lanes chunk k
// runs on lane k mod L and starts its sweeps when chunk k - 1's sweeps are
// over: the sweeps (bound by vector-ALU issue) run one after the other, and
// each chunk's filters, reprojection, inpainting and merges (bound by HBM
// and by launch latency) run beside the next chunk's sweeps.  Under capture
// (`capturing`) chunk k goes to a stream of its own that also waits for
// chunk k - L, which records the same order.
int enqueue(Frame* f, const float* const* match, const float* const* stitch, hipStream_t s, bool capturing) {
  if (f->ragged) return ss::ragged_enqueue(f->ragged, match, stitch, s);
  const int F = (int)f->d.size();
  const ss_frame_desc& c = f->d[0];  // the shared configuration
  const int R = c.R, C = c.C;
  const int cf = chunk_frames(F, R, f->lanes);
  const int K = (F + cf - 1) / cf;
  const int used = std::min(f->lanes, K);  // lanes that get a chunk
  const bool multi = used > 1;
  if (multi) {
    if (f->swept.size() < (size_t)K || (capturing && f->cap.size() < (size_t)K))
      return ss::report_error(SS_ERR_ARG, "frame executor: lane events not prepared");
    FR_HIP(hipEventRecord(f->fork, s));
  }
  for (int k = 0; k < K; ++k) {
    const int f0 = k * cf, nf = std::min(cf, F - f0), n = nf * R;
    hipStream_t ls = !multi ? s : capturing ? f->cap[k] : f->lane[k % used];
    if (multi) {
      if (k < used || capturing) FR_HIP(hipStreamWaitEvent(ls, f->fork, 0));
      if (capturing && k >= used) FR_HIP(hipStreamWaitEvent(ls, f->done[k - used], 0));
      if (k > 0) FR_HIP(hipStreamWaitEvent(ls, f->swept[k - 1], 0));
    }
    float* dist[ss::kMaxBatch];
    const float* cdist[ss::kMaxBatch];
    void* ws[ss::kMaxBatch];
    float* reproj[ss::kMaxBatch];
    float* scratch[ss::kMaxBatch];
    const uint8_t* lut[ss::kMaxBatch];
    ss_intrinsics intr[ss::kMaxBatch];
    const float* trans[ss::kMaxBatch];
    ss::ToU8Args cast[ss::kMaxBatch];
    bool cast_all = true;
    for (int fi = f0; fi < f0 + nf; ++fi) {
      const ss_frame_desc& d = f->d[fi];
      const float* rts[SS_MAX_REFERENCES];
      for (int r = 0; r < R; ++r) {
        rts[r] = d.rts[r];
        const int j = (fi - f0) * R + r;
        dist[j] = d.d_dist[r];
        cdist[j] = d.d_dist[r];
        ws[j] = d.d_workspace[r];
        reproj[j] = d.d_reproj[r];
        scratch[j] = d.d_scratch[r];
        lut[j] = d.d_inpaint_lut[r];
        intr[j] = d.intr[r];
        trans[j] = d.trans[r];
        cast[j] = ss::ToU8Args{stitch[(size_t)fi * R + r], d.d_stitch_u8[r]};
      }
      // the frame's uint8 stitch-image casts ride on its sweep launch
      bool cast_f = false;
      FR_CALL(ss::sweep_references(d.d_raylut, match + (size_t)fi * C, C, d.cams, rts, d.refs, R, d.cands, d.N, d.H,
                                   d.W, d.volume_type, d.d_workspace, ls, cast + (fi - f0) * R, R,
                                   (size_t)c.Hs * c.Ws * 3, &cast_f));
      cast_all = cast_all && cast_f;
    }
    if (multi) FR_HIP(hipEventRecord(f->swept[k], ls));
    // the fused argmin pass also fills the reprojection targets with 1e8
    const bool ready =
        !capturing && std::find(f->tables_done.begin(), f->tables_done.end(), ws[0]) != f->tables_done.end();
    // ... and the distance filter's last pass reprojects the final distances
    bool reprojected = false;
    FR_CALL(ss::estimate_after_sweep(n, c.H, c.W, c.N, c.volume_type, c.cands, c.sigma_i, c.sigma_s, dist, ws, ls,
                                     reproj, -1, true, ready, intr, trans, &reprojected));
    if (!capturing && !ready) f->tables_done.push_back(ws[0]);
    if (!reprojected) FR_CALL(ss::reproject_into(n, cdist, reproj, c.H, c.W, intr, trans, false, ls));
    FR_CALL(ss::inpaint_maps(n, reproj, scratch, lut, c.H, c.W, c.inpaint_iterations, c.thr_empty, c.thr_src, ls));
    if (!cast_all) FR_CALL(ss::to_u8_maps(cast, n, (size_t)c.Hs * c.Ws * 3, ls));
    for (int fi = f0; fi < f0 + nf; ++fi) {
      const ss_frame_desc& d = f->d[fi];
      const float* rp[SS_MAX_REFERENCES];
      const float* ds[SS_MAX_REFERENCES];
      const uint8_t* imgs[SS_MAX_REFERENCES];
      for (int r = 0; r < R; ++r) {
        rp[r] = d.d_reproj[r];
        ds[r] = d.d_dist[r];
        imgs[r] = d.d_stitch_u8[r];
      }
      FR_CALL(ss_merge(d.d_sampling, d.d_weights, rp, ds, imgs, d.Hs, d.Ws, &d.trans[0][0], d.intr, R, d.H, d.W, d.PH,
                       d.PW, d.d_dist_pano, d.d_rgb_pano, ls));
    }
    if (multi) FR_HIP(hipEventRecord(f->done[k], ls));
  }
  // `s` waits for the last chunk of every lane (under capture: every chunk's
  // stream, each joined back explicitly)
  for (int k = capturing ? 0 : K - used; multi && k < K; ++k) FR_HIP(hipStreamWaitEvent(s, f->done[k], 0));
  return SS_OK;
}

// The checks of one frame descriptor.
int check_desc(const ss_frame_desc& d) {
  if (d.C < 1 || d.C > SS_MAX_CAMERAS || d.R < 1 || d.R > SS_MAX_REFERENCES || d.N < 1 || d.N > SS_MAX_CANDIDATES ||
      d.H < 4 || d.W < 4 || d.Hs < 2 || d.Ws < 2 || d.PH < 1 || d.PW < 1 || d.inpaint_iterations < 0)
    return ss::report_error(SS_ERR_ARG, "frame descriptor sizes out of range");
  if (d.volume_type != SS_VOLUME_F32 && d.volume_type != SS_VOLUME_F16 && d.volume_type != SS_VOLUME_BF16)
    return ss::report_error(SS_ERR_ARG, "frame descriptor: unknown volume type");
  if (d.workspace_bytes < ss_estimate_workspace_bytes(d.H, d.W, d.N, d.volume_type))
    return ss::report_error(SS_ERR_WORKSPACE, "frame descriptor: workspace_bytes below ss_estimate_workspace_bytes");
  for (int r = 0; r < d.R; ++r)
    if (d.refs[r] < 0 || d.refs[r] >= d.C || !d.d_raylut[r] || !d.d_workspace[r] || !d.d_dist[r] ||
        !d.d_inpaint_lut[r] || !d.d_reproj[r] || !d.d_scratch[r] || !d.d_stitch_u8[r])
      return ss::report_error(SS_ERR_ARG, "frame descriptor: bad reference index or null buffer");
  if (!d.d_sampling || !d.d_weights || !d.d_dist_pano || !d.d_rgb_pano)
    return ss::report_error(SS_ERR_ARG, "frame descriptor: null stitching table or output");
  return SS_OK;
}

// Whether b has a's configuration (everything a batch's frames share).
bool same_config(const ss_frame_desc& a, const ss_frame_desc& b) {
  return a.C == b.C && a.R == b.R && a.N == b.N && a.H == b.H && a.W == b.W && a.Hs == b.Hs && a.Ws == b.Ws &&
         a.PH == b.PH && a.PW == b.PW && a.inpaint_iterations == b.inpaint_iterations &&
         a.volume_type == b.volume_type && a.sigma_i == b.sigma_i && a.sigma_s == b.sigma_s &&
         a.thr_empty == b.thr_empty && a.thr_src == b.thr_src && memcmp(a.cands, b.cands, sizeof(float) * a.N) == 0;
}

int drop_graph(Frame* f) {
  if (f->exec) FR_HIP(hipGraphExecDestroy(f->exec));
  if (f->graph) FR_HIP(hipGraphDestroy(f->graph));
  f->exec = nullptr;
  f->graph = nullptr;
  return SS_OK;
}

}  // namespace

extern "C" {

int ss_frame_batch_create(const ss_frame_desc* descs, int nframes, void** frame) {
  if (!descs || !frame) return ss::report_error(SS_ERR_ARG, "null descriptor array or output handle");
  if (nframes < 1 || nframes > SS_MAX_BATCH_FRAMES)
    return ss::report_error(SS_ERR_ARG, "frame count outside [1, SS_MAX_BATCH_FRAMES]");
  for (int i = 0; i < nframes; ++i) {
    const int rc = check_desc(descs[i]);
    if (rc != SS_OK) return rc;
    if (!same_config(descs[0], descs[i]))
      return ss::report_error(SS_ERR_ARG, "frame descriptors of a batch differ in sizes, volume type, candidates, "
                                          "sigmas or inpainting settings");
  }
  int dev = 0;
  FR_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return ss::report_error(SS_ERR_UNSUPPORTED, "device index beyond the pool");
  FR_HIP(pool_init(g_pool[dev]));
  Frame* f = new (std::nothrow) Frame();
  if (!f) return ss::report_error(SS_ERR_ARG, "out of host memory");
  f->d.assign(descs, descs + nframes);
  f->pool = &g_pool[dev];
  f->dev = dev;
  f->n_match = (size_t)nframes * descs[0].C;
  f->n_stitch = (size_t)nframes * descs[0].R;
  *frame = f;
  return SS_OK;
}

size_t ss_frame_ragged_table_bytes(const ss_frame_desc* descs, int nframes) {
  if (!descs || nframes < 1 || nframes > SS_MAX_BATCH_FRAMES) return 0;
  for (int i = 0; i < nframes; ++i)
    if (check_desc(descs[i]) != SS_OK) return 0;
  return ss::ragged_table_bytes(descs, nframes);
}

int ss_frame_ragged_create(const ss_frame_desc* descs, int nframes, int schedule, void* d_table, size_t table_bytes,
                           ss_stream stream, void** frame) {
  if (!descs || !frame) return ss::report_error(SS_ERR_ARG, "null descriptor array or output handle");
  if (nframes < 1 || nframes > SS_MAX_BATCH_FRAMES)
    return ss::report_error(SS_ERR_ARG, "frame count outside [1, SS_MAX_BATCH_FRAMES]");
  for (int i = 0; i < nframes; ++i) {
    const int rc = check_desc(descs[i]);
    if (rc != SS_OK) return rc;
  }
  int dev = 0;
  FR_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return ss::report_error(SS_ERR_UNSUPPORTED, "device index beyond the pool");
  FR_HIP(pool_init(g_pool[dev]));
  ss::RaggedPlan* plan = nullptr;
  FR_CALL(ss::ragged_create(descs, nframes, schedule, d_table, table_bytes, reinterpret_cast<hipStream_t>(stream),
                            &plan));
  Frame* f = new (std::nothrow) Frame();
  if (!f) {
    ss::ragged_free(plan);
    return ss::report_error(SS_ERR_ARG, "out of host memory");
  }
  f->d.assign(descs, descs + nframes);
  f->pool = &g_pool[dev];
  f->dev = dev;
  f->ragged = plan;
  int nm = 0, ns = 0;
  ss::ragged_images(plan, &nm, &ns);
  f->n_match = (size_t)nm;
  f->n_stitch = (size_t)ns;
  *frame = f;
  return SS_OK;
}

int ss_frame_create(const ss_frame_desc* desc, void** frame) { return ss_frame_batch_create(desc, 1, frame); }

int ss_frame_batch_run(void* frame, const float* const* d_match, const float* const* d_stitch, int use_graph,
                       ss_stream stream) {
  Frame* f = static_cast<Frame*>(frame);
  if (!f || !d_match || !d_stitch) return ss::report_error(SS_ERR_ARG, "null frame or image array");
  const size_t nm = f->n_match, ns = f->n_stitch;
  for (size_t k = 0; k < nm; ++k)
    if (!d_match[k]) return ss::report_error(SS_ERR_ARG, "null matching image");
  for (size_t r = 0; r < ns; ++r)
    if (!d_stitch[r]) return ss::report_error(SS_ERR_ARG, "null stitch image");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(f->pool->mu);
  FR_CALL(prepare_lanes(f, use_graph != 0));
  if (!use_graph) return enqueue(f, d_match, d_stitch, s, false);
  if (ss::profiler_active()) return ss::report_error(SS_ERR_UNSUPPORTED, "graph replay while kernel timing is on");

  // captured on the device's capture stream (the caller's may be the null
  // stream, which cannot be captured), replayed on the caller's stream: a
  // replay routed through another stream paid two cross-queue waits per
  // frame, measured at ~40 us
  hipStream_t main = f->pool->capture;
  const bool same = f->exec != nullptr && std::equal(d_match, d_match + nm, f->cap_match.begin()) &&
                    std::equal(d_stitch, d_stitch + ns, f->cap_stitch.begin());
  if (!same) {
    FR_CALL(drop_graph(f));
    FR_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeRelaxed));
    const int rc = enqueue(f, d_match, d_stitch, main, true);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(main, &g);
    if (rc != SS_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    FR_HIP(e);
    f->graph = g;
    FR_HIP(hipGraphInstantiate(&f->exec, f->graph, nullptr, nullptr, 0));
    f->cap_match.assign(d_match, d_match + nm);
    f->cap_stitch.assign(d_stitch, d_stitch + ns);
  }
  FR_HIP(hipGraphLaunch(f->exec, s));
  return SS_OK;
}

int ss_frame_run(void* frame, const float* const* d_match, const float* const* d_stitch, int use_graph,
                 ss_stream stream) {
  return ss_frame_batch_run(frame, d_match, d_stitch, use_graph, stream);
}

int ss_frame_set_lanes(void* frame, int lanes) {
  Frame* f = static_cast<Frame*>(frame);
  if (!f) return ss::report_error(SS_ERR_ARG, "null frame");
  if (lanes < 1 || lanes > kMaxLanes) return ss::report_error(SS_ERR_ARG, "lanes outside [1, 4]");
  if (f->ragged && lanes != 1) return ss::report_error(SS_ERR_UNSUPPORTED, "the ragged executor runs on one lane");
  int dev = 0;
  FR_HIP(hipGetDevice(&dev));
  if (dev != f->dev) return ss::report_error(SS_ERR_ARG, "the frame executor belongs to another device");
  std::lock_guard<std::mutex> lk(f->pool->mu);
  if (lanes == f->lanes) return SS_OK;
  if (lanes > 1 && !f->fork) FR_HIP(hipEventCreateWithFlags(&f->fork, hipEventDisableTiming));
  for (int l = 0; l < lanes; ++l)
    if (!f->lane[l]) FR_HIP(hipStreamCreateWithFlags(&f->lane[l], hipStreamNonBlocking));
  FR_CALL(drop_graph(f));  // a captured graph keeps the old schedule
  f->lanes = lanes;
  return SS_OK;
}

int ss_frame_get_lanes(void* frame) {
  const Frame* f = static_cast<const Frame*>(frame);
  return f ? f->lanes : -1;
}

int ss_frame_destroy(void* frame) {
  Frame* f = static_cast<Frame*>(frame);
  if (!f) return SS_OK;
  drop_graph(f);
  ss::ragged_free(f->ragged);
  for (hipStream_t l : f->lane)
    if (l) (void)hipStreamDestroy(l);
  for (hipStream_t l : f->cap) (void)hipStreamDestroy(l);
  for (hipEvent_t e : f->swept) (void)hipEventDestroy(e);
  for (hipEvent_t e : f->done) (void)hipEventDestroy(e);
  if (f->fork) (void)hipEventDestroy(f->fork);
  delete f;
  return SS_OK;
}

}  // extern "C"
