// Per-frame executor: estimate_RGBD_panorama (depth_estimation.py:205-243 and
// stitcher.py:166-220) as one call.  Every pass runs once for all of the
// frame's references: one sweep launch writes every reference's volume and
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
// per-launch host cost of its ~30 kernels.
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>

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

struct Frame {
  ss_frame_desc d;
  StreamPool* pool = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const float* cap_match[SS_MAX_CAMERAS] = {};
  const float* cap_stitch[SS_MAX_REFERENCES] = {};
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

// The frame's work, in order on `s`.
int enqueue(Frame* f, const float* const* match, const float* const* stitch, hipStream_t s) {
  const ss_frame_desc& d = f->d;
  const int R = d.R;
  const float* rts[SS_MAX_REFERENCES];
  const float* dists[SS_MAX_REFERENCES];
  const float* trans[SS_MAX_REFERENCES];
  ss::ToU8Args cast[SS_MAX_REFERENCES];
  for (int r = 0; r < R; ++r) {
    rts[r] = d.rts[r];
    dists[r] = d.d_dist[r];
    trans[r] = d.trans[r];
    cast[r] = ss::ToU8Args{stitch[r], d.d_stitch_u8[r]};
  }
  FR_CALL(ss::sweep_references(d.d_raylut, match, d.C, d.cams, rts, d.refs, R, d.cands, d.N, d.H, d.W,
                               d.volume_type, d.d_workspace, s));
  // the fused argmin pass also fills the reprojection targets with 1e8
  FR_CALL(ss::estimate_after_sweep(R, d.H, d.W, d.N, d.volume_type, d.cands, d.sigma_i, d.sigma_s, d.d_dist,
                                   d.d_workspace, s, d.d_reproj));
  FR_CALL(ss::reproject_into(R, dists, d.d_reproj, d.H, d.W, d.intr, trans, false, s));
  FR_CALL(ss::inpaint_maps(R, d.d_reproj, d.d_scratch, d.d_inpaint_lut, d.H, d.W, d.inpaint_iterations, d.thr_empty,
                           d.thr_src, s));
  FR_CALL(ss::to_u8_maps(cast, R, (size_t)d.Hs * d.Ws * 3, s));
  const float* reproj[SS_MAX_REFERENCES];
  const uint8_t* imgs[SS_MAX_REFERENCES];
  for (int r = 0; r < R; ++r) {
    reproj[r] = d.d_reproj[r];
    imgs[r] = d.d_stitch_u8[r];
  }
  return ss_merge(d.d_sampling, d.d_weights, reproj, dists, imgs, d.Hs, d.Ws, &d.trans[0][0], d.intr, R, d.H, d.W,
                  d.PH, d.PW, d.d_dist_pano, d.d_rgb_pano, s);
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

int ss_frame_create(const ss_frame_desc* desc, void** frame) {
  if (!desc || !frame) return ss::report_error(SS_ERR_ARG, "null descriptor or output handle");
  const ss_frame_desc& d = *desc;
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
  int dev = 0;
  FR_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return ss::report_error(SS_ERR_UNSUPPORTED, "device index beyond the pool");
  FR_HIP(pool_init(g_pool[dev]));
  Frame* f = new (std::nothrow) Frame();
  if (!f) return ss::report_error(SS_ERR_ARG, "out of host memory");
  f->d = d;
  f->pool = &g_pool[dev];
  *frame = f;
  return SS_OK;
}

int ss_frame_run(void* frame, const float* const* d_match, const float* const* d_stitch, int use_graph,
                 ss_stream stream) {
  Frame* f = static_cast<Frame*>(frame);
  if (!f || !d_match || !d_stitch) return ss::report_error(SS_ERR_ARG, "null frame or image array");
  for (int k = 0; k < f->d.C; ++k)
    if (!d_match[k]) return ss::report_error(SS_ERR_ARG, "null matching image");
  for (int r = 0; r < f->d.R; ++r)
    if (!d_stitch[r]) return ss::report_error(SS_ERR_ARG, "null stitch image");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(f->pool->mu);
  if (!use_graph) return enqueue(f, d_match, d_stitch, s);
  if (ss::profiler_active()) return ss::report_error(SS_ERR_UNSUPPORTED, "graph replay while kernel timing is on");

  // captured on the device's capture stream (the caller's may be the null
  // stream, which cannot be captured), replayed on the caller's stream: a
  // replay routed through another stream paid two cross-queue waits per
  // frame, measured at ~40 us
  hipStream_t main = f->pool->capture;
  bool same = f->exec != nullptr;
  for (int k = 0; k < f->d.C && same; ++k) same = f->cap_match[k] == d_match[k];
  for (int r = 0; r < f->d.R && same; ++r) same = f->cap_stitch[r] == d_stitch[r];
  if (!same) {
    FR_CALL(drop_graph(f));
    FR_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeRelaxed));
    const int rc = enqueue(f, d_match, d_stitch, main);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(main, &g);
    if (rc != SS_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    FR_HIP(e);
    f->graph = g;
    FR_HIP(hipGraphInstantiate(&f->exec, f->graph, nullptr, nullptr, 0));
    for (int k = 0; k < f->d.C; ++k) f->cap_match[k] = d_match[k];
    for (int r = 0; r < f->d.R; ++r) f->cap_stitch[r] = d_stitch[r];
  }
  FR_HIP(hipGraphLaunch(f->exec, s));
  return SS_OK;
}

int ss_frame_destroy(void* frame) {
  Frame* f = static_cast<Frame*>(frame);
  if (!f) return SS_OK;
  drop_graph(f);
  delete f;
  return SS_OK;
}

}  // extern "C"
