// Per-frame executor: estimate_RGBD_panorama (depth_estimation.py:205-243 and
// stitcher.py:166-220) as one call.  Each pass runs once for a whole group of
// references: one launch sweeps every reference of the group (and writes its
// guide), and each later pass (cost filter levels, fused argmin, distance
// filter, reprojection, inpainting, uint8 cast) is one launch whose grid
// covers the group.  The full-size passes then fill the GPU, and the
// latency-bound coarse-level passes cost one launch per group instead of one
// per reference.
//
// With two or more references, the odd ones run as a second group on a
// second stream, which waits only for the work that preceded the frame on
// the caller's stream; the merge waits for both groups.  On the MI355X boxes
// measured, the command processor drains one queue's kernels for a few
// hundred microseconds before it serves the other, so the two groups overlap
// only in part (the second group's sweep overlaps the first group's
// reprojection and inpainting): 744 us per frame against 765 us with one
// group on one stream (2 references, 1024^2, 32 candidates).  (Chains per
// reference that waited for a shared sweep started up to 250 us after their
// inputs were ready.)  The frame can be captured once into a HIP graph and
// replayed, which removes the per-launch host cost of its ~60 kernels.
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>

#include "ss_kernels.h"

namespace {

// The capture stream and the second group's stream, one pair per device,
// shared by every executor on it, created on first use (default priority)
// and kept until the process ends.  The mutex serialises the enqueueing
// (host work only) of frames on the device, and with it the use of both.
constexpr int kMaxDevices = 64;
struct StreamPool {
  std::once_flag once;
  hipError_t err = hipSuccess;
  hipStream_t capture = nullptr, lane = nullptr;
  std::mutex mu;
};
StreamPool g_pool[kMaxDevices];

hipError_t pool_init(StreamPool& p) {
  std::call_once(p.once, [&p] {
    p.err = hipStreamCreateWithFlags(&p.capture, hipStreamNonBlocking);
    if (p.err == hipSuccess) p.err = hipStreamCreateWithFlags(&p.lane, hipStreamNonBlocking);
  });
  return p.err;
}

struct Frame {
  ss_frame_desc d;
  StreamPool* pool = nullptr;
  hipEvent_t start = nullptr, join = nullptr;
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

// One half of the references (r = l, l + 2, ...) from sweep to uint8 cast.
int enqueue_half(const ss_frame_desc& d, int l, int L, const float* const* match, const float* const* stitch,
                 hipStream_t s) {
  const float* raylut[SS_MAX_REFERENCES];
  const float* rts[SS_MAX_REFERENCES];
  int refs[SS_MAX_REFERENCES];
  void* ws[SS_MAX_REFERENCES];
  float* dist[SS_MAX_REFERENCES];
  const float* cdist[SS_MAX_REFERENCES];
  float* reproj[SS_MAX_REFERENCES];
  float* scratch[SS_MAX_REFERENCES];
  const uint8_t* lut[SS_MAX_REFERENCES];
  ss_intrinsics intr[SS_MAX_REFERENCES];
  const float* trans[SS_MAX_REFERENCES];
  ss::ToU8Args cast[SS_MAX_REFERENCES];
  int n = 0;
  for (int r = l; r < d.R; r += L, ++n) {
    raylut[n] = d.d_raylut[r];
    rts[n] = d.rts[r];
    refs[n] = d.refs[r];
    ws[n] = d.d_workspace[r];
    dist[n] = d.d_dist[r];
    cdist[n] = d.d_dist[r];
    reproj[n] = d.d_reproj[r];
    scratch[n] = d.d_scratch[r];
    lut[n] = d.d_inpaint_lut[r];
    intr[n] = d.intr[r];
    trans[n] = d.trans[r];
    cast[n] = ss::ToU8Args{stitch[r], d.d_stitch_u8[r]};
  }
  FR_CALL(ss::sweep_references(raylut, match, d.C, d.cams, rts, refs, n, d.cands, d.N, d.H, d.W, d.volume_type, ws,
                               s));
  // the fused argmin pass also fills the reprojection targets with 1e8
  FR_CALL(ss::estimate_after_sweep(n, d.H, d.W, d.N, d.volume_type, d.cands, d.sigma_i, d.sigma_s, dist, ws, s,
                                   reproj));
  FR_CALL(ss::reproject_into(n, cdist, reproj, d.H, d.W, intr, trans, false, s));
  FR_CALL(ss::inpaint_maps(n, reproj, scratch, lut, d.H, d.W, d.inpaint_iterations, d.thr_empty, d.thr_src, s));
  return ss::to_u8_maps(cast, n, (size_t)d.Hs * d.Ws * 3, s);
}

// The frame's work on `s` (half of the references on the pool's lane).
int enqueue(Frame* f, const float* const* match, const float* const* stitch, hipStream_t s) {
  const ss_frame_desc& d = f->d;
  const int R = d.R;
  const int L = R >= 2 ? 2 : 1;
  if (L == 2) {
    FR_HIP(hipEventRecord(f->start, s));
    FR_HIP(hipStreamWaitEvent(f->pool->lane, f->start, 0));
    FR_CALL(enqueue_half(d, 1, L, match, stitch, f->pool->lane));
    FR_HIP(hipEventRecord(f->join, f->pool->lane));
  }
  FR_CALL(enqueue_half(d, 0, L, match, stitch, s));
  if (L == 2) FR_HIP(hipStreamWaitEvent(s, f->join, 0));
  const float* dists[SS_MAX_REFERENCES];
  const float* reproj[SS_MAX_REFERENCES];
  const uint8_t* imgs[SS_MAX_REFERENCES];
  for (int r = 0; r < R; ++r) {
    dists[r] = d.d_dist[r];
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
  if (hipEventCreateWithFlags(&f->start, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&f->join, hipEventDisableTiming) != hipSuccess) {
    ss_frame_destroy(f);
    return ss::report_error(SS_ERR_HIP, "hipEventCreateWithFlags failed");
  }
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
  if (f->start) (void)hipEventDestroy(f->start);
  if (f->join) (void)hipEventDestroy(f->join);
  delete f;
  return SS_OK;
}

}  // extern "C"
