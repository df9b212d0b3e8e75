// Per-frame executor: estimate_RGBD_panorama (depth_estimation.py:205-243 and
// stitcher.py:166-220) as one call.  One sweep launch covers every reference
// (and writes their guides); after it, each reference's chain runs on a
// stream of its own (the chains are independent until the merge), so one
// reference's latency-bound coarse pyramid passes overlap the other's
// full-size passes.  The frame can be captured once into a HIP graph and
// replayed, which removes the per-launch host cost of its ~70 kernels.
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <new>

#include "ss_kernels.h"

namespace {

// Streams shared by every executor on a device, created on first use and
// kept until the process ends.  A device has a few hardware queues (4 by
// default) and each new stream is bound to one of them; streams made per
// executor end up sharing queues, and two reference chains on one queue run
// one after the other.  One pool per device keeps a frame's concurrent
// chains on distinct queues however many executors exist.  Reference r runs
// on lane r % kLanes.  The mutex serialises the enqueueing (host work only)
// of frames that share the pool, and with it the use of the capture stream.
constexpr int kLanes = 3;
constexpr int kMaxDevices = 64;
struct StreamPool {
  std::once_flag once;
  hipError_t err = hipSuccess;
  hipStream_t lane[kLanes] = {};
  hipStream_t capture = nullptr;
  std::mutex mu;
};
StreamPool g_pool[kMaxDevices];

hipError_t pool_init(StreamPool& p) {
  std::call_once(p.once, [&p] {
    for (int i = 0; i < kLanes && p.err == hipSuccess; ++i)
      p.err = hipStreamCreateWithFlags(&p.lane[i], hipStreamNonBlocking);
    if (p.err == hipSuccess) p.err = hipStreamCreateWithFlags(&p.capture, hipStreamNonBlocking);
  });
  return p.err;
}

struct Frame {
  ss_frame_desc d;
  StreamPool* pool = nullptr;
  hipEvent_t pre = nullptr, done = nullptr, swept = nullptr;
  hipEvent_t join[SS_MAX_REFERENCES] = {};
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

// The frame's work after the events on `s`; `s` waits for all of it.
int enqueue(Frame* f, const float* const* match, const float* const* stitch, hipStream_t s) {
  const ss_frame_desc& d = f->d;
  const float* reproj[SS_MAX_REFERENCES];
  const float* dists[SS_MAX_REFERENCES];
  const uint8_t* imgs[SS_MAX_REFERENCES];
  // One sweep launch for every reference.  (Staggering per-reference sweeps
  // so that one reference's chain overlaps the next sweep measured no faster:
  // the ALU-bound sweep occupies the CUs and the overlapped passes crawl.)
  const float* rts[SS_MAX_REFERENCES];
  for (int r = 0; r < d.R; ++r) rts[r] = d.rts[r];
  FR_CALL(ss::sweep_references(d.d_raylut, match, d.C, d.cams, rts, d.refs, d.R, d.cands, d.N, d.H, d.W,
                               d.volume_type, d.d_workspace, s));
  FR_HIP(hipEventRecord(f->swept, s));
  for (int r = 0; r < d.R; ++r) {
    hipStream_t sr = f->pool->lane[r % kLanes];
    FR_HIP(hipStreamWaitEvent(sr, f->swept, 0));
    // the fused argmin pass also fills the reprojection target with 1e8
    FR_CALL(ss::estimate_after_sweep(d.H, d.W, d.N, d.volume_type, d.cands, d.sigma_i, d.sigma_s, d.d_dist[r],
                                     d.d_workspace[r], sr, d.d_reproj[r]));
    // (issued before the sweeps end, this memory-bound cast would crawl
    // beside the ALU-bound sweep and hold the chain back)
    FR_CALL(ss_to_u8(stitch[r], d.d_stitch_u8[r], (size_t)d.Hs * d.Ws * 3, sr));
    FR_CALL(ss::reproject_into(d.d_dist[r], d.d_reproj[r], d.H, d.W, &d.intr[r], d.trans[r], false, sr));
    FR_CALL(ss_inpaint(d.d_reproj[r], d.d_scratch[r], d.d_inpaint_lut[r], d.H, d.W, d.inpaint_iterations,
                       d.thr_empty, d.thr_src, sr));
    FR_HIP(hipEventRecord(f->join[r], sr));
    reproj[r] = d.d_reproj[r];
    dists[r] = d.d_dist[r];
    imgs[r] = d.d_stitch_u8[r];
  }
  for (int r = 0; r < d.R; ++r) FR_HIP(hipStreamWaitEvent(s, f->join[r], 0));
  return ss_merge(d.d_sampling, d.d_weights, reproj, dists, imgs, d.Hs, d.Ws, &d.trans[0][0], d.intr, d.R, d.H, d.W,
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
  int rc = SS_OK;
  auto check = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == SS_OK) {
      char m[256];
      snprintf(m, sizeof m, "%s: %s", what, hipGetErrorString(e));
      rc = ss::report_error(SS_ERR_HIP, m);
    }
  };
  for (int r = 0; r < d.R; ++r)
    check(hipEventCreateWithFlags(&f->join[r], hipEventDisableTiming), "hipEventCreateWithFlags");
  check(hipEventCreateWithFlags(&f->swept, hipEventDisableTiming), "hipEventCreateWithFlags");
  check(hipEventCreateWithFlags(&f->pre, hipEventDisableTiming), "hipEventCreateWithFlags");
  check(hipEventCreateWithFlags(&f->done, hipEventDisableTiming), "hipEventCreateWithFlags");
  if (rc != SS_OK) {
    ss_frame_destroy(f);
    return rc;
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

  hipStream_t main = f->pool->capture;
  bool same = f->exec != nullptr;
  for (int k = 0; k < f->d.C && same; ++k) same = f->cap_match[k] == d_match[k];
  for (int r = 0; r < f->d.R && same; ++r) same = f->cap_stitch[r] == d_stitch[r];
  // the replay stream starts after the caller's work
  FR_HIP(hipEventRecord(f->pre, s));
  FR_HIP(hipStreamWaitEvent(main, f->pre, 0));
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
  FR_HIP(hipGraphLaunch(f->exec, main));
  FR_HIP(hipEventRecord(f->done, main));
  FR_HIP(hipStreamWaitEvent(s, f->done, 0));
  return SS_OK;
}

int ss_frame_destroy(void* frame) {
  Frame* f = static_cast<Frame*>(frame);
  if (!f) return SS_OK;
  drop_graph(f);
  for (int r = 0; r < SS_MAX_REFERENCES; ++r)
    if (f->join[r]) (void)hipEventDestroy(f->join[r]);
  if (f->swept) (void)hipEventDestroy(f->swept);
  if (f->pre) (void)hipEventDestroy(f->pre);
  if (f->done) (void)hipEventDestroy(f->done);
  delete f;
  return SS_OK;
}

}  // extern "C"
