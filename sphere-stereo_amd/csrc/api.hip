// C ABI of libsphere_stereo.so: argument checks, workspace carving and the
// per-reference / per-frame launch sequences.  Declarations and the reference
// interfaces each entry point replaces are in include/sphere_stereo.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "ss_kernels.h"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

int ok() {
  g_err[0] = '\0';
  return SS_OK;
}

#define SS_HIP(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(SS_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define SS_REQUIRE(cond, ...)                                 \
  do {                                                        \
    if (!(cond)) return fail(SS_ERR_ARG, __VA_ARGS__);        \
  } while (0)

inline hipStream_t as_stream(ss_stream s) { return reinterpret_cast<hipStream_t>(s); }

// Kernel timing: a pool of events created when timing is enabled, and the
// (stage, start, end) records of the bracketed launches.
struct Profiler {
  std::mutex mu;
  bool on = false;
  unsigned mask = 0;
  std::vector<hipEvent_t> pool;
  size_t next = 0;
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
};
Profiler g_prof;
constexpr size_t kEventPool = 1 << 14;

// Brackets the launches made during its lifetime when `stage` is profiled.
struct StageTimer {
  int stage;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  StageTimer(int st, hipStream_t str) : stage(st), s(str) {
    if (!g_prof.on || !(g_prof.mask & (1u << st))) return;
    std::lock_guard<std::mutex> lk(g_prof.mu);
    if (g_prof.next + 2 > g_prof.pool.size()) return;
    a = g_prof.pool[g_prof.next++];
    b = g_prof.pool[g_prof.next++];
    (void)hipEventRecord(a, s);
  }
  ~StageTimer() {
    if (!a) return;
    (void)hipEventRecord(b, s);
    std::lock_guard<std::mutex> lk(g_prof.mu);
    g_prof.recs.push_back({stage, a, b});
  }
};

// Carves 256-byte aligned regions out of one workspace; with base == nullptr
// it only measures.
struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* b) : base(static_cast<char*>(b)) {}
  template <typename T>
  T* take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
};

int levels_of(int H, int W) {
  if (H < 1 || W < 1) return -1;
  const double l = fmin(log2((double)W), log2((double)H));
  return (int)(l - 1);
}

int padded(int N) {
  if (N == 1) return 1;
  const int v = ss::lanes_v(N);
  return v < 0 ? -1 : 8 * v;
}

using Level = ss::PyrLevel;

size_t elem_bytes(int vt) { return vt == SS_VOLUME_F32 ? 4 : 2; }
bool valid_vtype(int vt) { return vt == SS_VOLUME_F32 || vt == SS_VOLUME_F16 || vt == SS_VOLUME_BF16; }

// Level sizes ceil(rows / 2^s) x ceil(cols / 2^s) (isb_filter.py:64-70) and
// their buffers for s >= 1 (fp32 volumes whatever the level-0 type).
void carve_levels(Carver& cv, Level* lv, int L, int H, int W, int NP) {
  for (int s = 1; s < L; ++s) {
    lv[s].h = (int)ceil(H / pow(2.0, s));
    lv[s].w = (int)ceil(W / pow(2.0, s));
    lv[s].g = cv.take<uint32_t>((size_t)lv[s].h * lv[s].w);
    lv[s].c = cv.take<float>((size_t)lv[s].h * lv[s].w * NP);
  }
}

}  // namespace

namespace ss {

// Up-pass weights into level k (isb_filter.py:110-112), in double as the
// reference's Python computes them.
void up_weights(double sigma_s, int k, double* wu, double* wd) {
  const double var_inv_s = 1 / (2 * sigma_s * sigma_s);
  const double d = pow(2.0, k) - 0.5;
  *wd = exp(-(d * d) * var_inv_s);
  *wu = 1 - *wd;
}

// The coarse end of a pyramid runs in one workgroup from level s0 on: every
// level it produces needs at most kTailLanes lanes (pixels x G, the lanes per
// pixel), i.e. at most two pixels per thread of the 1024-thread workgroup, so
// that one CU does each pass in one or two wavefronts' issue time (from four
// pixels per thread on, a pass is VALU-bound on that CU and costs more than
// the launch it replaces), and all of its levels (guide + NP values per
// pixel, and the weight table) fit in the LDS budget.  With 8 lanes per pixel
// the tail starts where the next level holds 256 pixels; the one-lane distance
// filter reaches one level further (2048 pixels).  The tail starts at level
// 1 or deeper (its levels are all fp32, and a fused level-0 pass stays a
// regular launch).  Returns 0 when the pyramid has no such tail.
int tail_start(const PyrLevel* l0, int L, int G, int NP) {
  const long long kTailLanes = 2048;
  auto lds_need = [&](int from) {
    size_t b = 0;
    for (int k = from + 1; k < L; ++k) {
      const size_t px = (size_t)l0[k].h * l0[k].w;
      b += ((px * 4 + 15) & ~size_t(15)) + ((px * NP * 4 + 15) & ~size_t(15));
    }
    return b;
  };
  int s0 = L - 1;
  for (int k = 1; k < L - 1; ++k)
    if ((long long)l0[k + 1].h * l0[k + 1].w * G <= kTailLanes && lds_need(k) <= (size_t)kTailLdsBytes) {
      s0 = k;
      break;
    }
  return (L - 1 - s0 >= 2 && L - s0 <= kMaxTailLevels) ? s0 : 0;
}

// The level table of a tail from s0 (levels s0..L-1 of l0); returns its LDS bytes.
unsigned tail_levels(const PyrLevel* l0, int s0, int L, int NP, const float* wtab, double sigma_s, TailLevels* t) {
  memset(t, 0, sizeof *t);
  t->n = L - s0;
  t->wtab = wtab;
  unsigned off = 0;
  for (int e = 0; e < t->n; ++e) {
    t->h[e] = l0[s0 + e].h;
    t->w[e] = l0[s0 + e].w;
    if (e > 0) {
      const unsigned px = (unsigned)(t->h[e] * t->w[e]);
      t->goff[e] = off;
      off += (px * 4 + 15) & ~15u;
      t->coff[e] = off;
      off += (px * (unsigned)(NP * 4) + 15) & ~15u;
    }
    double wu, wd;
    up_weights(sigma_s, s0 + e, &wu, &wd);
    t->wu[e] = (float)wu;
    t->wd[e] = (float)wd;
  }
  t->wtab_off = off;
  off += (unsigned)(kWeightTable * sizeof(float));
  return off;
}

// The bounds under which the sweep's projections may take project_t_fast:
// finite parameters, |fs|, |transform entries|, |candidates| <= 2^20,
// 2^-16 <= |cs| <= 2^20, |xi|, |alpha|, |1 - alpha| <= 16 (the rays are unit
// vectors by construction).  Real rigs are far inside them.  rt: npairs
// transforms of 12 floats.
bool sweep_fast_bounds(const TorchCam* cams, int C, const float* rt, int npairs, const float* cands, int N) {
  auto in = [](float v, float lo, float hi) { return std::isfinite(v) && fabsf(v) >= lo && fabsf(v) <= hi; };
  const float big = 0x1p20f;
  for (int k = 0; k < C; ++k) {
    const TorchCam& c = cams[k];
    if (!in(c.fs0, 0.0f, big) || !in(c.fs1, 0.0f, big) || !in(c.cs0, 0x1p-16f, big) || !in(c.cs1, 0x1p-16f, big) ||
        !in(c.xi, 0.0f, 16.0f) || !in(c.a, 0.0f, 16.0f) || !in(c.one_m_a, 0.0f, 16.0f))
      return false;
  }
  for (int i = 0; i < 12 * npairs; ++i)
    if (!in(rt[i], 0.0f, big)) return false;
  for (int n = 0; n < N; ++n)
    if (!in(cands[n], 0.0f, big)) return false;
  return true;
}

}  // namespace ss

namespace {

// ISB_Filter.apply's two launch loops (isb_filter.py:87-134) for nb pyramids
// of the same shape in the same launches (lv[b] are the levels of entry b).
// With fuse_last, the level-0 up-pass computes the argmin into dist[b]
// instead of writing the filtered volume.  vt is the element type of level
// 0; levels >= 1 are fp32.
// wtab: the filter's guide weight table (launch_weight_tables with
// var_inv_i = 1 / (2 sigma_i^2), isb_filter.py:88), already enqueued on s.
// rp_out / rp_intr / rp_t (optional, G = V = 1): the level-0 pass also
// reprojects each pixel's final value into rp_out[b] (reproject_splat).
int run_pyramid(Level* const* lv, int nb, int L, int G, int V, int vt, const float* wtab, double sigma_s,
                bool fuse_last, const ss::ArgminParams& am, float* const* dist, hipStream_t s, int stage_down,
                int stage_up, float* const* fill = nullptr, int xcd = 1, bool guide0_unused = false,
                float* const* rp_out = nullptr, const ss_intrinsics* rp_intr = nullptr,
                const float* const* rp_t = nullptr) {
  if (nb < 1 || nb > ss::kMaxBatch) return fail(SS_ERR_ARG, "batch of %d pyramids outside [1, %d]", nb, ss::kMaxBatch);
  const Level* l0 = lv[0];  // the shapes (identical for every entry)
  const int NP = G * V;
  const int s0 = ss::tail_start(l0, L, G, NP);
  const bool tail = s0 > 0;
  const int last_down = tail ? s0 : L - 1;
  ss::IsbDownArgs da[ss::kMaxBatch];
  ss::IsbUpArgs ua[ss::kMaxBatch];
  // The first down pass takes the entries last to first: the sweep wrote
  // them first to last, so the volumes still in the last-level cache (256 MB
  // of Infinity Cache, less than two fp32 1024^2 x 32 volumes) are read
  // before they are evicted; the fused level-0 up pass then takes them first
  // to last, again most recently read first.
  for (int k = 1; k <= last_down; ++k) {
    for (int b = 0; b < nb; ++b) {
      const Level* v = lv[k == 1 ? nb - 1 - b : b];
      da[b] = ss::IsbDownArgs{v[k - 1].g, v[k - 1].c, v[k - 1].h, v[k - 1].w, v[k].g, v[k].c, v[k].h, v[k].w, wtab};
    }
    StageTimer t(stage_down, s);
    SS_HIP(ss::launch_isb_down(da, nb, G, V, k == 1 ? vt : SS_VOLUME_F32, xcd, s));
  }
  if (tail) {
    ss::IsbTailArgs ta;
    memset(&ta, 0, sizeof ta);
    for (int b = 0; b < nb; ++b) {
      ta.g0[b] = lv[b][s0].g;
      ta.c0[b] = lv[b][s0].c;
    }
    ta.lds_bytes = tail_levels(l0, s0, L, NP, wtab, sigma_s, &ta.t);
    StageTimer t(stage_down, s);
    SS_HIP(ss::launch_isb_tail(ta, nb, G, V, s));
  }
  for (int k = (tail ? s0 : L - 1) - 1; k >= 0; --k) {
    double wu, wd;
    ss::up_weights(sigma_s, k, &wu, &wd);
    const bool fuse = fuse_last && k == 0;
    for (int b = 0; b < nb; ++b) {
      const Level* v = lv[b];
      ua[b] = ss::IsbUpArgs{v[k + 1].g, v[k + 1].c, v[k + 1].h, v[k + 1].w, v[k].g, v[k].c, v[k].h, v[k].w,
                            (float)wu,  (float)wd,  wtab,       fuse ? dist[b] : nullptr, am,
                            fuse && fill ? fill[b] : nullptr};
      ua[b].skip_guide = k == 0 && guide0_unused ? 1 : 0;
      if (k == 0 && rp_out) {
        ua[b].rp_out = rp_out[b];
        ua[b].rp_c = rp_intr[b];
        memcpy(ua[b].rp_t, rp_t[b], sizeof ua[b].rp_t);
      }
    }
    StageTimer t(fuse ? SS_STAGE_ISB_UP_ARGMIN : stage_up, s);
    SS_HIP(ss::launch_isb_up(ua, nb, G, V, k == 0 ? vt : SS_VOLUME_F32, fuse, xcd, s, k == 0 && rp_out != nullptr));
  }
  if (fuse_last && L == 1) {
    for (int b = 0; b < nb; ++b) {
      StageTimer t(SS_STAGE_ISB_UP_ARGMIN, s);
      // the argmin kernel works in groups of 8 lanes
      SS_HIP(ss::launch_argmin(lv[b][0].c, l0[0].h * l0[0].w, NP / 8, vt, am, dist[b], s));
    }
  }
  return SS_OK;
}

bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

int check_cams(const ss_camera* cams, const float* rts, int C, int ref) {
  SS_REQUIRE(cams && rts, "null camera or transform array");
  SS_REQUIRE(C >= 1 && C <= SS_MAX_CAMERAS, "camera count %d outside [1, %d]", C, SS_MAX_CAMERAS);
  SS_REQUIRE(ref >= 0 && ref < C, "reference index %d outside [0, %d)", ref, C);
  return SS_OK;
}

int check_sweep(const float* const* d_images, int C, const ss_camera* cams, const float* cands, int N, int H,
                int W, int vt) {
  SS_REQUIRE(d_images && cams && cands, "null pointer argument");
  SS_REQUIRE(C >= 1 && C <= SS_MAX_CAMERAS, "camera count %d outside [1, %d]", C, SS_MAX_CAMERAS);
  SS_REQUIRE(N >= 1 && N <= SS_MAX_CANDIDATES, "candidate count %d outside [1, %d]", N, SS_MAX_CANDIDATES);
  SS_REQUIRE(H >= 1 && W >= 1, "bad matching resolution %dx%d", W, H);
  SS_REQUIRE(valid_vtype(vt), "unknown volume type %d", vt);
  for (int k = 0; k < C; ++k) SS_REQUIRE(d_images[k], "null image pointer for camera %d", k);
  return SS_OK;
}

// ss::sweep_fast_bounds on a launch's argument record
bool sweep_fast_bounds(const ss::SweepArgs& a, int C, int npairs) {
  return ss::sweep_fast_bounds(a.cam, C, &a.rt[0][0], npairs, a.cands, a.N);
}

// The autotuners keep variant 0 (the default) unless another is at least 2 %
// faster: closer timings are within run-to-run noise.
constexpr float kTuneMargin = 0.98f;

// The fastest variant of ms[0..n) (ms < 0: not run) against the default.
int pick_variant(const float* ms, int n) {
  int best = 0;
  for (int v = 1; v < n; ++v)
    if (ms[v] >= 0.0f && ms[v] < ms[best]) best = v;
  return ms[best] < kTuneMargin * ms[0] ? best : 0;
}

// The sweep variant chosen per configuration by ss_sweep_autotune (or
// ss_sweep_set_variant); configurations never tuned use variant 0.
struct SweepTune {
  int H, W, N, vt, variant;
};
std::mutex g_tune_mu;
std::vector<SweepTune> g_tune;

int sweep_variant_for(int H, int W, int N, int vt) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (const SweepTune& t : g_tune)
    if (t.H == H && t.W == W && t.N == N && t.vt == vt) return t.variant;
  return 0;
}

void set_sweep_variant(int H, int W, int N, int vt, int variant) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (size_t i = 0; i < g_tune.size(); ++i)
    if (g_tune[i].H == H && g_tune[i].W == W && g_tune[i].N == N && g_tune[i].vt == vt) {
      if (variant < 0)
        g_tune.erase(g_tune.begin() + (long)i);
      else
        g_tune[i].variant = variant;
      return;
    }
  if (variant >= 0) g_tune.push_back(SweepTune{H, W, N, vt, variant});
}

// The cost filter's scheduling variants: lanes per pixel group (8: V = NP/8
// candidates a lane, or 4: twice as many, which halves the per-pixel weight
// broadcasts, argmin rounds and guide sums per candidate) x tile order
// (XCD-aware or launch order).  Every variant computes each value with the
// same operations in the same order.  Groups of 4 exist for NP <= 64.
constexpr int kIsbVariants = 4;
struct IsbVariant {
  int lanes, xcd;
};
constexpr IsbVariant kIsbVariant[kIsbVariants] = {{8, 1}, {8, 0}, {4, 1}, {4, 0}};

bool isb_variant_ok(int v, int N) {
  if (v < 0 || v >= kIsbVariants || N < 2) return v == 0;
  return kIsbVariant[v].lanes == 8 || padded(N) <= 64;
}

std::vector<SweepTune> g_isb_tune;  // (H, W, N, vt) -> variant, guarded by g_tune_mu

int isb_variant_for(int H, int W, int N, int vt) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (const SweepTune& t : g_isb_tune)
    if (t.H == H && t.W == W && t.N == N && t.vt == vt) return t.variant;
  return 0;
}

void set_isb_variant(int H, int W, int N, int vt, int variant) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (size_t i = 0; i < g_isb_tune.size(); ++i)
    if (g_isb_tune[i].H == H && g_isb_tune[i].W == W && g_isb_tune[i].N == N && g_isb_tune[i].vt == vt) {
      if (variant < 0)
        g_isb_tune.erase(g_isb_tune.begin() + (long)i);
      else
        g_isb_tune[i].variant = variant;
      return;
    }
  if (variant >= 0) g_isb_tune.push_back(SweepTune{H, W, N, vt, variant});
}

// The inpainting tile recorded per map size (H, W); 0 (48 x 48) if none.
std::vector<SweepTune> g_inpaint_tune;  // N = vt = 0; guarded by g_tune_mu

int inpaint_variant_for(int H, int W) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (const SweepTune& t : g_inpaint_tune)
    if (t.H == H && t.W == W) return t.variant;
  return 0;
}

void set_inpaint_variant(int H, int W, int variant) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (size_t i = 0; i < g_inpaint_tune.size(); ++i)
    if (g_inpaint_tune[i].H == H && g_inpaint_tune[i].W == W) {
      if (variant < 0)
        g_inpaint_tune.erase(g_inpaint_tune.begin() + (long)i);
      else
        g_inpaint_tune[i].variant = variant;
      return;
    }
  if (variant >= 0) g_inpaint_tune.push_back(SweepTune{H, W, 0, 0, variant});
}

// (G, V, xcd) of the cost filter of a configuration: its tuned variant, or
// `variant` when >= 0
void isb_schedule(int H, int W, int N, int vt, int variant, int* G, int* V, int* xcd) {
  int v = variant >= 0 ? variant : isb_variant_for(H, W, N, vt);
  if (!isb_variant_ok(v, N)) v = 0;
  const int NP = padded(N);
  *G = N == 1 ? 1 : kIsbVariant[v].lanes;
  *V = N == 1 ? 1 : NP / *G;
  *xcd = kIsbVariant[v].xcd;
}

// One sweep launch over references refs[0..R) (R * C <= kSweepMaxPairs):
// rts[r] holds reference r's C transforms, costs[r] / guides[r] its outputs
// (guides[r] may be null).  Requires N >= 2 (the volume stride is 8 * V).
// variant < 0: the configuration's tuned variant
int launch_sweep_refs(const float* const* d_raylut, const float* const* d_images, int C, const ss_camera* cams,
                      const float* const* rts, const int* refs, int R, const float* cands, int N, int H, int W,
                      int vt, void* const* costs, uint32_t* const* guides, hipStream_t s, int variant = -1,
                      const ss::SweepCast* cast = nullptr) {
  ss::SweepArgs a;
  memset(&a, 0, sizeof a);
  if (cast) a.cast = *cast;
  for (int k = 0; k < C; ++k) {
    a.img[k] = d_images[k];
    a.cam[k] = ss::make_torch_cam(cams[k]);
  }
  for (int r = 0; r < R; ++r) {
    SS_REQUIRE(refs[r] >= 0 && refs[r] < C, "reference index %d outside [0, %d)", refs[r], C);
    SS_REQUIRE(d_raylut[r] && costs[r] && rts[r], "null ray table, cost volume or transforms");
    a.ref[r].raylut = reinterpret_cast<const float4*>(d_raylut[r]);
    a.ref[r].img = d_images[refs[r]];
    a.ref[r].cost = costs[r];
    a.ref[r].guide = guides ? guides[r] : nullptr;
    a.ref[r].rt0 = r * C;
    for (int k = 0; k < C; ++k) memcpy(a.rt[r * C + k], rts[r] + 12 * k, sizeof(float) * 12);
  }
  memcpy(a.cands, cands, sizeof(float) * N);
  const int V = ss::lanes_v(N);
  SS_REQUIRE(V > 0, "candidate count %d unsupported", N);
  a.R = R;
  a.H = H;
  a.W = W;
  a.N = N;
  a.NP = 8 * V;
  a.tiles_x = (W + 63) / 64;
  a.blocks_per_ref = a.tiles_x * ((H + 3) / 4);
  a.Wf = (float)W;
  a.Hf = (float)H;
  a.invW = 1.0f / (float)W;
  a.invH = 1.0f / (float)H;
  a.inv2W = 2.0f / (float)W;
  a.inv2H = 2.0f / (float)H;
  a.halfW = (float)W * 0.5f;
  a.halfH = (float)H * 0.5f;
  a.pow2 = pow2(W) && pow2(H);
  a.fastproj = sweep_fast_bounds(a, C, R * C);
  StageTimer t(SS_STAGE_SWEEP, s);
  SS_HIP(ss::launch_sweep(a, vt, variant < 0 ? sweep_variant_for(H, W, N, vt) : variant, s));
  return SS_OK;
}

}  // namespace

namespace ss {
int report_error(int code, const char* what) { return fail(code, "%s", what); }
bool profiler_active() { return g_prof.on; }
}  // namespace ss

extern "C" {

int ss_abi_version(void) { return SS_ABI_VERSION; }

const char* ss_last_error(void) { return g_err; }

int ss_padded_candidates(int N) { return padded(N); }

int ss_isb_levels(int H, int W) { return levels_of(H, W); }

int ss_select_camera(const ss_camera* cams, const float* rts, int C, int ref, const float* d_masks, int H, int W,
                     double min_dist, double max_dist, int8_t* d_sel, ss_stream stream) {
  int rc = check_cams(cams, rts, C, ref);
  if (rc) return rc;
  SS_REQUIRE(d_masks && d_sel, "null pointer argument");
  SS_REQUIRE(H >= 1 && W >= 1, "bad matching resolution %dx%d", W, H);
  ss::SelectArgs a;
  memset(&a, 0, sizeof a);
  a.ref = ss::make_torch_cam(cams[ref]);
  for (int k = 0; k < C; ++k) {
    a.cam[k] = ss::make_torch_cam(cams[k]);
    memcpy(a.rt[k], rts + 12 * k, sizeof(float) * 12);
  }
  a.masks = d_masks;
  a.sel = d_sel;
  a.C = C;
  a.refidx = ref;
  a.H = H;
  a.W = W;
  a.mn = (float)min_dist;
  a.mx = (float)max_dist;
  a.Wf = (float)W;
  a.Hf = (float)H;
  a.invW = 1.0f / (float)W;
  a.invH = 1.0f / (float)H;
  a.pow2W = pow2(W);
  a.pow2H = pow2(H);
  SS_HIP(ss::launch_select_camera(a, as_stream(stream)));
  return ok();
}

int ss_ray_lut(const ss_camera* ref, const int8_t* d_sel, int H, int W, float* d_raylut, ss_stream stream) {
  SS_REQUIRE(ref && d_sel && d_raylut, "null pointer argument");
  SS_REQUIRE(H >= 1 && W >= 1, "bad matching resolution %dx%d", W, H);
  SS_HIP(ss::launch_ray_lut(ss::make_torch_cam(*ref), d_sel, H, W, reinterpret_cast<float4*>(d_raylut),
                            as_stream(stream)));
  return ok();
}

int ss_inpaint_lut(const ss_intrinsics* intr, const float* t3, int H, int W, double min_dist, double max_dist,
                   uint8_t* d_lut, ss_stream stream) {
  SS_REQUIRE(intr && t3 && d_lut, "null pointer argument");
  SS_REQUIRE(H >= 1 && W >= 1, "bad matching resolution %dx%d", W, H);
  SS_HIP(ss::launch_inpaint_lut(*intr, t3, H, W, (float)min_dist, (float)max_dist, d_lut, as_stream(stream)));
  return ok();
}

int ss_box_filter(const float* d_in, float* d_out, float* d_tmp, int B, int H, int W, int radius, float pad_value,
                  ss_stream stream) {
  SS_REQUIRE(d_in && d_out && d_tmp, "null pointer argument");
  SS_REQUIRE(B >= 1 && H >= 1 && W >= 1 && radius >= 0, "bad box filter shape");
  SS_REQUIRE(d_in != d_out, "box filter cannot run in place");
  SS_HIP(ss::launch_box_filter(d_in, d_out, d_tmp, B, H, W, radius, pad_value, as_stream(stream)));
  return ok();
}

int ss_blend_lut(const ss_intrinsics* intr, const float* rot, const float* trans, int R, const float* d_masks, int H,
                 int W, int PH, int PW, double min_dist, double max_dist, float* d_sampling, float* d_weights,
                 ss_stream stream) {
  SS_REQUIRE(intr && rot && trans && d_masks && d_sampling && d_weights, "null pointer argument");
  SS_REQUIRE(R >= 1 && R <= SS_MAX_REFERENCES, "reference count %d outside [1, %d]", R, SS_MAX_REFERENCES);
  SS_REQUIRE(H >= 2 && W >= 2 && PH >= 1 && PW >= 1, "bad blend table shape");
  ss::BlendArgs a;
  memset(&a, 0, sizeof a);
  for (int r = 0; r < R; ++r) {
    a.intr[r] = intr[r];
    memcpy(a.rot[r], rot + 9 * r, sizeof(float) * 9);
    memcpy(a.trans[r], trans + 3 * r, sizeof(float) * 3);
  }
  a.masks = d_masks;
  a.sampling = reinterpret_cast<float2*>(d_sampling);
  a.weights = d_weights;
  a.R = R;
  a.H = H;
  a.W = W;
  a.PH = PH;
  a.PW = PW;
  a.mn = (float)min_dist;
  a.mx = (float)max_dist;
  SS_HIP(ss::launch_blend_lut(a, as_stream(stream)));
  return ok();
}

int ss_normalize_weights(float* d_w, int R, int P, ss_stream stream) {
  SS_REQUIRE(d_w && R >= 1 && P >= 1, "bad argument");
  SS_HIP(ss::launch_normalize(d_w, R, P, as_stream(stream)));
  return ok();
}

int ss_guide(const float* d_img, int H, int W, uint32_t* d_guide, ss_stream stream) {
  SS_REQUIRE(d_img && d_guide && H >= 1 && W >= 1, "bad argument");
  StageTimer t(SS_STAGE_GUIDE, as_stream(stream));
  SS_HIP(ss::launch_guide(d_img, H * W, d_guide, as_stream(stream)));
  return ok();
}

int ss_sweep(const float* d_raylut, const float* const* d_images, int C, const ss_camera* cams, const float* rts,
             int ref, const float* cands, int N, int H, int W, int vtype, void* d_cost, uint32_t* d_guide,
             ss_stream stream) {
  SS_REQUIRE(d_cost && d_raylut && rts, "null pointer argument");
  int rc = check_sweep(d_images, C, cams, cands, N, H, W, vtype);
  if (rc) return rc;
  SS_REQUIRE(ref >= 0 && ref < C, "reference index %d outside [0, %d)", ref, C);
  SS_REQUIRE(N >= 2, "the sweep needs at least 2 candidates (1 candidate gives d_{N-1} everywhere)");
  rc = launch_sweep_refs(&d_raylut, d_images, C, cams, &rts, &ref, 1, cands, N, H, W, vtype, &d_cost, &d_guide,
                         as_stream(stream));
  return rc ? rc : ok();
}

int ss_sweep_variant_count(void) { return ss::sweep_variant_count(); }

int ss_sweep_variant_info(int variant, int* chunk, int* group) {
  SS_REQUIRE(chunk && group, "null output");
  SS_REQUIRE(variant >= 0 && variant < ss::sweep_variant_count(), "sweep variant %d outside [0, %d)", variant,
             ss::sweep_variant_count());
  ss::sweep_variant_info(variant, chunk, group);
  return ok();
}

int ss_sweep_set_variant(int H, int W, int N, int vtype, int variant) {
  SS_REQUIRE(variant >= -1 && variant < ss::sweep_variant_count(), "sweep variant %d outside [-1, %d)", variant,
             ss::sweep_variant_count());
  set_sweep_variant(H, W, N, vtype, variant);
  return ok();
}

int ss_sweep_get_variant(int H, int W, int N, int vtype) { return sweep_variant_for(H, W, N, vtype); }

int ss_isb_variant_count(void) { return kIsbVariants; }

int ss_isb_variant_info(int variant, int* lanes, int* xcd_order) {
  SS_REQUIRE(lanes && xcd_order, "null output");
  SS_REQUIRE(variant >= 0 && variant < kIsbVariants, "filter variant %d outside [0, %d)", variant, kIsbVariants);
  *lanes = kIsbVariant[variant].lanes;
  *xcd_order = kIsbVariant[variant].xcd;
  return ok();
}

int ss_isb_set_variant(int H, int W, int N, int vtype, int variant) {
  SS_REQUIRE(variant >= -1 && variant < kIsbVariants, "filter variant %d outside [-1, %d)", variant, kIsbVariants);
  SS_REQUIRE(variant < 0 || isb_variant_ok(variant, N), "filter variant %d does not exist for %d candidates", variant,
             N);
  set_isb_variant(H, W, N, vtype, variant);
  return ok();
}

int ss_isb_get_variant(int H, int W, int N, int vtype) { return isb_variant_for(H, W, N, vtype); }

int ss_sweep_autotune(const float* d_raylut, const float* const* d_images, int C, const ss_camera* cams,
                      const float* rts, int ref, const float* cands, int N, int H, int W, int vtype, void* d_cost,
                      int repeats, ss_stream stream, float* ms_out, int* best_out) {
  int rc = check_cams(cams, rts, C, ref);
  if (rc) return rc;
  rc = check_sweep(d_images, C, cams, cands, N, H, W, vtype);
  if (rc) return rc;
  SS_REQUIRE(d_raylut && d_cost, "null pointer argument");
  SS_REQUIRE(N >= 2, "the sweep needs at least two candidates");
  SS_REQUIRE(repeats >= 1 && repeats <= 1000, "repeats %d outside [1, 1000]", repeats);
  hipStream_t s = as_stream(stream);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  SS_HIP(hipStreamIsCapturing(s, &cs));
  SS_REQUIRE(cs == hipStreamCaptureStatusNone, "ss_sweep_autotune times launches: not while capturing a graph");
  hipEvent_t e0, e1;
  SS_HIP(hipEventCreate(&e0));
  SS_HIP(hipEventCreate(&e1));
  const float* lut[1] = {d_raylut};
  const float* rt[1] = {rts};
  void* cost[1] = {d_cost};
  const int nv = ss::sweep_variant_count();
  float msv[16];
  for (int v = 0; v < 16; ++v) msv[v] = -1.0f;
  for (int v = 0; v < nv && rc == SS_OK; ++v) {
    rc = launch_sweep_refs(lut, d_images, C, cams, rt, &ref, 1, cands, N, H, W, vtype, cost, nullptr, s, v);
    if (rc) break;
    if (hipEventRecord(e0, s) != hipSuccess) {
      rc = fail(SS_ERR_HIP, "hipEventRecord failed");
      break;
    }
    for (int i = 0; i < repeats && rc == SS_OK; ++i)
      rc = launch_sweep_refs(lut, d_images, C, cams, rt, &ref, 1, cands, N, H, W, vtype, cost, nullptr, s, v);
    if (rc) break;
    float ms = 0.0f;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = fail(SS_ERR_HIP, "timing the sweep variants failed");
      break;
    }
    ms /= (float)repeats;
    if (ms_out) ms_out[v] = ms;
    msv[v] = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  const int best = pick_variant(msv, nv);
  set_sweep_variant(H, W, N, vtype, best);
  if (best_out) *best_out = best;
  return ok();
}

size_t ss_isb_workspace_bytes(int H, int W, int N, int vtype, int planar) {
  const int L = levels_of(H, W), NP = padded(N);
  if (L < 1 || NP < 0 || !valid_vtype(vtype)) return 0;
  if (planar && vtype != SS_VOLUME_F32) return 0;
  Carver cv(nullptr);  // the pyramid levels are fp32 whatever the level-0 type
  Level lv[32];
  if (planar) {
    cv.take<uint32_t>((size_t)H * W);
    cv.take<float>((size_t)H * W * NP);
  }
  cv.take<float>(ss::kWeightTable);
  carve_levels(cv, lv, L, H, W, NP);
  return cv.off + 256;
}

// The guide weight table of a filter with sigma_i (isb_filter.py:88), built on s.
static int build_weight_table(Carver& cv, double sigma_i, hipStream_t s, float** tab) {
  *tab = cv.take<float>(ss::kWeightTable);
  const float vi = (float)(1 / (2 * sigma_i * sigma_i));
  float* const t[1] = {*tab};
  SS_HIP(ss::launch_weight_tables(&vi, t, 1, s));
  return SS_OK;
}

static int isb_filter_impl(uint32_t* d_guide, void* d_cost, int H, int W, int N, int vt, double sigma_i,
                           double sigma_s, Carver& cv, hipStream_t s) {
  const int L = levels_of(H, W);
  const int NP = padded(N);
  float* wtab = nullptr;
  const int rc = build_weight_table(cv, sigma_i, s, &wtab);
  if (rc) return rc;
  Level lv[32];
  lv[0] = Level{H, W, d_guide, d_cost};
  carve_levels(cv, lv, L, H, W, NP);
  int G, V, xcd;
  isb_schedule(H, W, N, vt, -1, &G, &V, &xcd);
  ss::ArgminParams am{};
  Level* const one[1] = {lv};
  return run_pyramid(one, 1, L, G, V, vt, wtab, sigma_s, false, am, nullptr, s, SS_STAGE_ISB_DOWN,
                     SS_STAGE_ISB_UP, nullptr, xcd);
}

static int check_isb_args(int H, int W, int N, int vt) {
  SS_REQUIRE(H >= 1 && W >= 1, "bad resolution %dx%d", W, H);
  SS_REQUIRE(levels_of(H, W) >= 1, "resolution %dx%d gives no pyramid level (min side must be >= 4)", W, H);
  SS_REQUIRE(N >= 1 && padded(N) > 0, "candidate count %d unsupported (1..%d)", N, SS_MAX_CANDIDATES);
  SS_REQUIRE(valid_vtype(vt), "unknown volume type %d", vt);
  SS_REQUIRE(N > 1 || vt == SS_VOLUME_F32, "a one-candidate volume is filtered in fp32 only");
  return SS_OK;
}

static int check_isb(int H, int W, int N, int vt, void* ws, size_t bytes, int planar) {
  const int rc = check_isb_args(H, W, N, vt);
  if (rc) return rc;
  const size_t need = ss_isb_workspace_bytes(H, W, N, vt, planar);
  if (!ws || bytes < need)
    return fail(SS_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need, ws ? bytes : (size_t)0);
  return SS_OK;
}

int ss_isb_filter(uint32_t* d_guide, void* d_cost, int H, int W, int N, int vtype, double sigma_i, double sigma_s,
                  void* d_workspace, size_t workspace_bytes, ss_stream stream) {
  SS_REQUIRE(d_guide && d_cost, "null pointer argument");
  int rc = check_isb(H, W, N, vtype, d_workspace, workspace_bytes, 0);
  if (rc) return rc;
  Carver cv(d_workspace);
  rc = isb_filter_impl(d_guide, d_cost, H, W, N, vtype, sigma_i, sigma_s, cv, as_stream(stream));
  return rc ? rc : ok();
}

// ---------------------------------------------------------------------------
// The allocation-free ISB filter.  Its scratch (the call's guide-weight table
// and pyramid levels 1..L-1, laid out as ss_isb_filter's workspace) is a
// file-scope device array: every GPU the library runs on holds one copy until
// the process ends, and every call on that GPU shares it, so calls that
// overlap on one GPU race.  No part of it must start at zero: a call writes
// every byte it reads (the weight table first, each level by the down pass
// into it), so the last call's contents are harmless.  On the host the name
// `scratch` is a stand-in, not a device address: its address on the current
// device is looked up once per device (hipGetSymbolAddress) and only that
// address reaches the kernels.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[SS_ISB_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

// The current device's copy of `scratch`, looked up on the first call there.
static int scratch_on_current_device(void** out) {
  int dev = -1;
  SS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SS_ERR_ARG, "device %d outside [0, %d)", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(SS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SS_OK;
}

int ss_isb_filter_noalloc(uint32_t* d_guide, void* d_cost, int H, int W, int N, int vtype, double sigma_i,
                          double sigma_s, ss_stream stream) {
  SS_REQUIRE(d_guide && d_cost, "null pointer argument");
  int rc = check_isb_args(H, W, N, vtype);
  if (rc) return rc;
  const size_t need = ss_isb_workspace_bytes(H, W, N, vtype, 0);
  if (need > (size_t)SS_ISB_NOALLOC_SCRATCH_BYTES)
    return fail(SS_ERR_WORKSPACE,
                "ss_isb_filter_noalloc: %dx%d with %d candidates needs %zu bytes of scratch, more than its static "
                "%zu (use ss_isb_filter with a workspace)",
                W, H, N, need, (size_t)SS_ISB_NOALLOC_SCRATCH_BYTES);
  void* ws = nullptr;
  rc = scratch_on_current_device(&ws);
  if (rc) return rc;
  Carver cv(ws);
  rc = isb_filter_impl(d_guide, d_cost, H, W, N, vtype, sigma_i, sigma_s, cv, as_stream(stream));
  return rc ? rc : ok();
}

int ss_isb_filter_planar(uint8_t* d_guide, float* d_cost, int H, int W, int N, double sigma_i, double sigma_s,
                         void* d_workspace, size_t workspace_bytes, ss_stream stream) {
  SS_REQUIRE(d_guide && d_cost, "null pointer argument");
  int rc = check_isb(H, W, N, SS_VOLUME_F32, d_workspace, workspace_bytes, 1);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  Carver cv(d_workspace);
  const int NP = padded(N);
  uint32_t* g = cv.take<uint32_t>((size_t)H * W);
  float* c = cv.take<float>((size_t)H * W * NP);
  SS_HIP(ss::launch_guide3_to_packed(d_guide, g, H * W, s));
  SS_HIP(ss::launch_planar_to_cc(d_cost, c, H * W, N, NP, s));
  rc = isb_filter_impl(g, c, H, W, N, SS_VOLUME_F32, sigma_i, sigma_s, cv, s);
  if (rc) return rc;
  SS_HIP(ss::launch_packed_to_guide3(g, d_guide, H * W, s));
  SS_HIP(ss::launch_cc_to_planar(c, d_cost, H * W, N, NP, s));
  return ok();
}

int ss_argmin_refine(const void* d_cost, int H, int W, int N, int vtype, float d0, float dN1, float A,
                     float* d_dist, ss_stream stream) {
  SS_REQUIRE(d_cost && d_dist && H >= 1 && W >= 1, "bad argument");
  SS_REQUIRE(N >= 2 && N <= SS_MAX_CANDIDATES, "candidate count %d outside [2, %d]", N, SS_MAX_CANDIDATES);
  SS_REQUIRE(valid_vtype(vtype), "unknown volume type %d", vtype);
  ss::ArgminParams am{N, d0, dN1, A};
  SS_HIP(ss::launch_argmin(d_cost, H * W, ss::lanes_v(N), vtype, am, d_dist, as_stream(stream)));
  return ok();
}

// Workspace of ss_estimate_distance: level-0 guide and volume, the cost
// pyramid, and the distance pyramid (which shares the guide levels).
static size_t estimate_layout(Carver& cv, int H, int W, int N, int vt, uint32_t** g0, void** c0, Level* lv,
                              Level* dl, float** wtab = nullptr) {
  const int L = levels_of(H, W), NP = padded(N < 2 ? 2 : N);
  const size_t eb = elem_bytes(vt);
  *g0 = cv.take<uint32_t>((size_t)H * W);
  *c0 = cv.take<char>((size_t)H * W * NP * eb);
  float* t = cv.take<float>(2 * ss::kWeightTable);  // cost filter, distance filter
  if (wtab) *wtab = t;
  carve_levels(cv, lv, L, H, W, NP);
  for (int s = 1; s < L; ++s) {
    dl[s].h = lv[s].h;
    dl[s].w = lv[s].w;
    dl[s].g = lv[s].g;  // the distance filter reuses the guide pyramid buffers
    dl[s].c = cv.take<float>((size_t)lv[s].h * lv[s].w);
  }
  return cv.off + 256;
}

size_t ss_estimate_workspace_bytes(int H, int W, int N, int vtype) {
  if (levels_of(H, W) < 1 || N < 1 || padded(N) < 0 || !valid_vtype(vtype)) return 0;
  Carver cv(nullptr);
  uint32_t* g0;
  void* c0;
  Level lv[32], dl[32];
  return estimate_layout(cv, H, W, N, vtype, &g0, &c0, lv, dl);
}

}  // extern "C"

namespace ss {

size_t volume_elem_bytes(int vt) { return elem_bytes(vt); }

void estimate_level0(void* ws, int H, int W, int N, int vt, uint32_t** g0, void** c0) {
  Carver cv(ws);
  Level lv[32], dl[32];
  estimate_layout(cv, H, W, N, vt, g0, c0, lv, dl);
}

int isb_levels(int H, int W) { return levels_of(H, W); }

int estimate_levels(void* ws, int H, int W, int N, int vt, float* dist, PyrLevel* lv, PyrLevel* dl) {
  Carver cv(ws);
  uint32_t* g0;
  void* c0;
  estimate_layout(cv, H, W, N, vt, &g0, &c0, lv, dl);
  lv[0] = PyrLevel{H, W, g0, c0};
  dl[0] = PyrLevel{H, W, g0, dist};
  return levels_of(H, W);
}

int estimate_after_sweep(int nb, int H, int W, int N, int vt, const float* cands, double sigma_i, double sigma_s,
                         float* const* d_dist, void* const* ws, hipStream_t s, float* const* fill, int isb_variant,
                         bool distance_filter, bool tables_ready, const ss_intrinsics* rp_intr,
                         const float* const* rp_t, bool* reprojected) {
  if (reprojected) *reprojected = false;
  if (nb < 1 || nb > kMaxBatch) return fail(SS_ERR_ARG, "batch of %d references outside [1, %d]", nb, kMaxBatch);
  const int L = levels_of(H, W);
  Level lv[kMaxBatch][32], dl[kMaxBatch][32];
  Level* lvp[kMaxBatch];
  Level* dlp[kMaxBatch];
  uint32_t* g0[kMaxBatch];
  void* c0[kMaxBatch];
  float* wtab = nullptr;  // entry 0's tables serve the whole batch (one sigma_i)
  for (int b = 0; b < nb; ++b) {
    Carver cv(ws[b]);
    estimate_layout(cv, H, W, N, vt, &g0[b], &c0[b], lv[b], dl[b], b == 0 ? &wtab : nullptr);
    lv[b][0] = Level{H, W, g0[b], c0[b]};
    dl[b][0] = Level{H, W, g0[b], d_dist[b]};
    lvp[b] = lv[b];
    dlp[b] = dl[b];
  }
  const float d0 = cands[0], dN1 = cands[N - 1];
  if (!tables_ready) {
    // the cost filter's table (sigma_i) and the distance filter's (sigma_i / 2,
    // depth_estimation.py:201-202), one launch
    const float vi[2] = {(float)(1 / (2 * sigma_i * sigma_i)), (float)(1 / (2 * (sigma_i / 2) * (sigma_i / 2)))};
    float* const t[2] = {wtab, wtab + ss::kWeightTable};
    SS_HIP(ss::launch_weight_tables(vi, t, 2, s));
  }
  if (N == 1) {
    // one candidate: max == min everywhere, so the reference returns d_{N-1}
    // (depth_estimation.py:198) whatever the cost
    for (int b = 0; b < nb; ++b) {
      SS_HIP(ss::launch_fill(d_dist[b], dN1, (size_t)H * W, s));
      if (fill) SS_HIP(ss::launch_fill(fill[b], 1e8f, (size_t)H * W, s));
    }
  } else {
    ss::ArgminParams am{N, d0, dN1, (float)(d0 / dN1) - 1.0f};
    int G, V, xcd;
    isb_schedule(H, W, N, vt, isb_variant, &G, &V, &xcd);
    const int rc = run_pyramid(lvp, nb, L, G, V, vt, wtab, sigma_s, true, am, d_dist, s, SS_STAGE_ISB_DOWN,
                               SS_STAGE_ISB_UP, L > 1 ? fill : nullptr, xcd);
    if (rc) return rc;
    if (fill && L == 1)
      for (int b = 0; b < nb; ++b) SS_HIP(ss::launch_fill(fill[b], 1e8f, (size_t)H * W, s));
  }
  if (!distance_filter) return SS_OK;
  // the level-0 guide is not read after the distance filter (the reference
  // filters a copy of it, depth_estimation.py:201): its last up pass skips it.
  // With rp_intr, that pass also reprojects the final distances into fill.
  const bool rp = rp_intr && rp_t && fill && L > 1;
  const int rc = run_pyramid(dlp, nb, L, 1, 1, SS_VOLUME_F32, wtab + ss::kWeightTable, sigma_s / 2, false,
                             ss::ArgminParams{}, nullptr, s, SS_STAGE_DIST_DOWN, SS_STAGE_DIST_UP, nullptr, 1, true,
                             rp ? fill : nullptr, rp ? rp_intr : nullptr, rp ? rp_t : nullptr);
  if (rc == SS_OK && reprojected) *reprojected = rp;
  return rc;
}

// The cost filter of the estimate path (its fused argmin included) timed per
// scheduling variant on the level-0 guide and volume in `ws`.
int time_filter_variants(void* ws, int H, int W, int N, int vt, const float* cands, double sigma_i, double sigma_s,
                         float* d_dist, int repeats, hipStream_t s, float* ms_out, int* best_out) {
  hipEvent_t e0, e1;
  SS_HIP(hipEventCreate(&e0));
  SS_HIP(hipEventCreate(&e1));
  void* const wss[1] = {ws};
  float* const dists[1] = {d_dist};
  int rc = SS_OK;
  float msv[kIsbVariants];
  for (int v = 0; v < kIsbVariants && rc == SS_OK; ++v) {
    msv[v] = -1.0f;
    if (ms_out) ms_out[v] = -1.0f;
    if (!isb_variant_ok(v, N)) continue;
    rc = estimate_after_sweep(1, H, W, N, vt, cands, sigma_i, sigma_s, dists, wss, s, nullptr, v, false);
    if (rc) break;
    if (hipEventRecord(e0, s) != hipSuccess) {
      rc = fail(SS_ERR_HIP, "hipEventRecord failed");
      break;
    }
    for (int i = 0; i < repeats && rc == SS_OK; ++i)
      rc = estimate_after_sweep(1, H, W, N, vt, cands, sigma_i, sigma_s, dists, wss, s, nullptr, v, false);
    if (rc) break;
    float ms = 0.0f;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = fail(SS_ERR_HIP, "timing the filter variants failed");
      break;
    }
    ms /= (float)repeats;
    if (ms_out) ms_out[v] = ms;
    msv[v] = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  const int best = pick_variant(msv, kIsbVariants);
  set_isb_variant(H, W, N, vt, best);
  if (best_out) *best_out = best;
  return SS_OK;
}

// Sweeps (and guides) of R references into their estimate workspaces, in as
// few launches as the transform table allows.  With N == 1 there is no sweep
// to run and only the guides are computed.
int sweep_references(const float* const* d_raylut, const float* const* d_images, int C, const ss_camera* cams,
                     const float* const* rts, const int* refs, int R, const float* cands, int N, int H, int W,
                     int vt, void* const* ws, hipStream_t s, const ToU8Args* cast, int ncast, size_t cast_n,
                     bool* cast_done) {
  // the uint8 casts ride on the (last) sweep launch when they fit SweepCast
  SweepCast sc;
  memset(&sc, 0, sizeof sc);
  bool fuse = cast && ncast > 0 && ncast <= SS_MAX_REFERENCES && cast_n % 4 == 0 && cast_n / 4 > 0 &&
              cast_n / 4 * (size_t)ncast < ((size_t)1 << 31) && N > 1;
  for (int i = 0; fuse && i < ncast; ++i)
    fuse = (reinterpret_cast<uintptr_t>(cast[i].in) & 15) == 0 && (reinterpret_cast<uintptr_t>(cast[i].out) & 3) == 0;
  if (fuse) {
    for (int i = 0; i < ncast; ++i) {
      sc.in[i] = cast[i].in;
      sc.out[i] = cast[i].out;
    }
    sc.n = ncast;
    sc.quads = (unsigned)(cast_n / 4);
    // about 8 quads per lane, at most 1024 workgroups
    const size_t total = (size_t)sc.quads * ncast;
    sc.blocks = (int)std::min<size_t>(1024, (total + 256 * 8 - 1) / (256 * 8));
  }
  if (cast_done) *cast_done = fuse;
  void* costs[SS_MAX_REFERENCES];
  uint32_t* guides[SS_MAX_REFERENCES];
  for (int r = 0; r < R; ++r) estimate_level0(ws[r], H, W, N, vt, &guides[r], &costs[r]);
  if (N == 1) {
    for (int r = 0; r < R; ++r) {
      StageTimer t(SS_STAGE_GUIDE, s);
      SS_HIP(ss::launch_guide(d_images[refs[r]], H * W, guides[r], s));
    }
    return SS_OK;
  }
  const int per = kSweepMaxPairs / C;
  for (int r0 = 0; r0 < R; r0 += per) {
    const int n = R - r0 < per ? R - r0 : per;
    const int rc = launch_sweep_refs(d_raylut + r0, d_images, C, cams, rts + r0, refs + r0, n, cands, N, H, W, vt,
                                     costs + r0, guides + r0, s, -1, fuse && r0 + n >= R ? &sc : nullptr);
    if (rc) return rc;
  }
  return SS_OK;
}

int reproject_into(int nb, const float* const* d_in, float* const* d_out, int H, int W, const ss_intrinsics* intr,
                   const float* const* t3, bool fill, hipStream_t s) {
  if (nb < 1 || nb > kMaxBatch) return fail(SS_ERR_ARG, "batch of %d maps outside [1, %d]", nb, kMaxBatch);
  StageTimer t(SS_STAGE_REPROJECT, s);
  ReprojectArgs a[kMaxBatch];
  for (int b = 0; b < nb; ++b) {
    if (fill) SS_HIP(ss::launch_fill(d_out[b], 1e8f, (size_t)H * W, s));
    a[b].din = d_in[b];
    a[b].dout = d_out[b];
    a[b].c = intr[b];
    memcpy(a[b].t, t3[b], sizeof a[b].t);
  }
  SS_HIP(ss::launch_reproject(a, nb, H, W, s));
  return SS_OK;
}

int to_u8_maps(const ToU8Args* a, int nb, size_t n, hipStream_t s) {
  if (nb < 1 || nb > kMaxBatch) return fail(SS_ERR_ARG, "batch of %d images outside [1, %d]", nb, kMaxBatch);
  if (n == 0) return SS_OK;
  StageTimer t(SS_STAGE_TO_U8, s);
  SS_HIP(ss::launch_to_u8(a, nb, n, s));
  return SS_OK;
}

int inpaint_maps(int nb, float* const* map, float* const* tmp, const uint8_t* const* lut, int H, int W,
                 int iterations, double thr_empty, double thr_src, hipStream_t s) {
  return inpaint_maps_v(nb, map, tmp, lut, H, W, iterations, thr_empty, thr_src, -1, s);
}

int inpaint_maps_v(int nb, float* const* map, float* const* tmp, const uint8_t* const* lut, int H, int W,
                   int iterations, double thr_empty, double thr_src, int variant, hipStream_t s) {
  if (nb < 1 || nb > kMaxBatch) return fail(SS_ERR_ARG, "batch of %d maps outside [1, %d]", nb, kMaxBatch);
  InpaintArgs a[kMaxBatch];
  for (int b = 0; b < nb; ++b) a[b] = InpaintArgs{map[b], tmp[b], lut[b]};
  const int per = ss::inpaint_steps_per_launch();
  const int tile = ss::inpaint_tile_of(variant >= 0 ? variant : inpaint_variant_for(H, W));
  bool in_map = true;  // where the current values are
  for (int i = 0; i < iterations; i += per) {
    StageTimer tm(SS_STAGE_INPAINT, s);
    SS_HIP(ss::launch_inpaint_multi(a, nb, H, W, iterations - i < per ? iterations - i : per, thr_empty, thr_src,
                                    tile, s));
    for (int b = 0; b < nb; ++b) {
      const float* t = a[b].din;
      a[b].din = a[b].dout;
      a[b].dout = const_cast<float*>(t);
    }
    in_map = !in_map;
  }
  if (!in_map)
    for (int b = 0; b < nb; ++b)
      SS_HIP(hipMemcpyAsync(map[b], tmp[b], sizeof(float) * H * W, hipMemcpyDeviceToDevice, s));
  return SS_OK;
}

}  // namespace ss

extern "C" {

int ss_estimate_distance(const float* d_raylut, const float* const* d_images, int C, const ss_camera* cams,
                         const float* rts, int ref, const float* cands, int N, int H, int W, double sigma_i,
                         double sigma_s, int vtype, const uint8_t* d_guide3, float* d_dist, void* d_workspace,
                         size_t workspace_bytes, ss_stream stream) {
  hipStream_t s = as_stream(stream);
  SS_REQUIRE(d_dist && d_raylut && rts, "null pointer argument");
  int rc = check_sweep(d_images, C, cams, cands, N, H, W, vtype);
  if (rc) return rc;
  SS_REQUIRE(ref >= 0 && ref < C, "reference index %d outside [0, %d)", ref, C);
  SS_REQUIRE(levels_of(H, W) >= 1, "resolution %dx%d gives no pyramid level (min side must be >= 4)", W, H);
  const size_t need = ss_estimate_workspace_bytes(H, W, N, vtype);
  if (!d_workspace || workspace_bytes < need)
    return fail(SS_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need, workspace_bytes);
  if (d_guide3) {
    // the caller's guide; the sweep then writes only the volume
    uint32_t* g0;
    void* c0;
    ss::estimate_level0(d_workspace, H, W, N, vtype, &g0, &c0);
    {
      StageTimer t(SS_STAGE_GUIDE, s);
      SS_HIP(ss::launch_guide3_to_packed(d_guide3, g0, H * W, s));
    }
    if (N > 1) {
      uint32_t* none = nullptr;
      rc = launch_sweep_refs(&d_raylut, d_images, C, cams, &rts, &ref, 1, cands, N, H, W, vtype, &c0, &none, s);
      if (rc) return rc;
    }
  } else {
    void* const ws[1] = {d_workspace};
    rc = ss::sweep_references(&d_raylut, d_images, C, cams, &rts, &ref, 1, cands, N, H, W, vtype, ws, s);
    if (rc) return rc;
  }
  float* const dists[1] = {d_dist};
  void* const wss[1] = {d_workspace};
  rc = ss::estimate_after_sweep(1, H, W, N, vtype, cands, sigma_i, sigma_s, dists, wss, s, nullptr);
  return rc ? rc : ok();
}

int ss_estimate_filter_autotune(void* d_workspace, size_t workspace_bytes, int H, int W, int N, int vtype,
                                const float* cands, double sigma_i, double sigma_s, float* d_dist, int repeats,
                                ss_stream stream, float* ms_out, int* best_out) {
  SS_REQUIRE(d_workspace && cands && d_dist, "null pointer argument");
  SS_REQUIRE(N >= 2 && N <= SS_MAX_CANDIDATES, "candidate count %d outside [2, %d]", N, SS_MAX_CANDIDATES);
  SS_REQUIRE(H >= 1 && W >= 1 && levels_of(H, W) >= 1, "resolution %dx%d gives no pyramid level", W, H);
  SS_REQUIRE(valid_vtype(vtype), "unknown volume type %d", vtype);
  SS_REQUIRE(repeats >= 1 && repeats <= 1000, "repeats %d outside [1, 1000]", repeats);
  const size_t need = ss_estimate_workspace_bytes(H, W, N, vtype);
  if (workspace_bytes < need)
    return fail(SS_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need, workspace_bytes);
  hipStream_t s = as_stream(stream);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  SS_HIP(hipStreamIsCapturing(s, &cs));
  SS_REQUIRE(cs == hipStreamCaptureStatusNone, "ss_estimate_filter_autotune times launches: not while capturing");
  const int rc = ss::time_filter_variants(d_workspace, H, W, N, vtype, cands, sigma_i, sigma_s, d_dist, repeats, s,
                                          ms_out, best_out);
  return rc ? rc : ok();
}

int ss_reproject(const float* d_in, float* d_out, int H, int W, const ss_intrinsics* intr, const float* t3,
                 ss_stream stream) {
  SS_REQUIRE(d_in && d_out && intr && t3 && H >= 1 && W >= 1, "bad argument");
  SS_REQUIRE(d_in != d_out, "reprojection cannot run in place");
  float* const outs[1] = {d_out};
  const int rc = ss::reproject_into(1, &d_in, outs, H, W, intr, &t3, true, as_stream(stream));
  return rc ? rc : ok();
}

int ss_inpaint(float* d_map, float* d_tmp, const uint8_t* d_lut, int H, int W, int iterations, double thr_empty,
               double thr_src, ss_stream stream) {
  SS_REQUIRE(d_map && d_tmp && d_lut && H >= 1 && W >= 1 && iterations >= 0, "bad argument");
  const int rc = ss::inpaint_maps(1, &d_map, &d_tmp, &d_lut, H, W, iterations, thr_empty, thr_src, as_stream(stream));
  return rc ? rc : ok();
}

int ss_inpaint_variant_count(void) { return ss::inpaint_tile_count(); }

int ss_inpaint_variant_info(int variant, int* tile) {
  SS_REQUIRE(tile, "null output");
  SS_REQUIRE(variant >= 0 && variant < ss::inpaint_tile_count(), "inpainting variant %d outside [0, %d)", variant,
             ss::inpaint_tile_count());
  *tile = ss::inpaint_tile_of(variant);
  return ok();
}

int ss_inpaint_set_variant(int H, int W, int variant) {
  SS_REQUIRE(variant >= -1 && variant < ss::inpaint_tile_count(), "inpainting variant %d outside [-1, %d)", variant,
             ss::inpaint_tile_count());
  set_inpaint_variant(H, W, variant);
  return ok();
}

int ss_inpaint_get_variant(int H, int W) { return inpaint_variant_for(H, W); }

int ss_inpaint_autotune(const float* d_map, int nmaps, float* d_work, float* d_tmp, const uint8_t* d_lut, int H,
                        int W, int iterations, double thr_empty, double thr_src, int repeats, ss_stream stream,
                        float* ms_out, int* best_out) {
  SS_REQUIRE(d_map && d_work && d_tmp && d_lut && H >= 1 && W >= 1 && iterations >= 1, "bad argument");
  SS_REQUIRE(nmaps >= 1 && nmaps <= ss::kMaxBatch, "map count %d outside [1, %d]", nmaps, ss::kMaxBatch);
  SS_REQUIRE(d_map != d_work && d_work != d_tmp, "the map, its working copies and the scratch must differ");
  SS_REQUIRE(repeats >= 1 && repeats <= 1000, "repeats %d outside [1, 1000]", repeats);
  hipStream_t s = as_stream(stream);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  SS_HIP(hipStreamIsCapturing(s, &cs));
  SS_REQUIRE(cs == hipStreamCaptureStatusNone, "ss_inpaint_autotune times launches: not while capturing");
  hipEvent_t e0, e1;
  SS_HIP(hipEventCreate(&e0));
  SS_HIP(hipEventCreate(&e1));
  const size_t bytes = (size_t)H * W * sizeof(float);
  int rc = SS_OK;
  float msv[8];
  for (int v = 0; v < 8; ++v) msv[v] = -1.0f;
  // each run inpaints fresh copies of the map, nmaps in one batch as the frame
  // executors run a frame's references: the copies cost every variant the same
  float* m[ss::kMaxBatch];
  float* t[ss::kMaxBatch];
  const uint8_t* l[ss::kMaxBatch];
  for (int b = 0; b < nmaps; ++b) {
    m[b] = d_work + (size_t)b * H * W;
    t[b] = d_tmp + (size_t)b * H * W;
    l[b] = d_lut;
  }
  auto run = [&](int v) -> int {
    for (int b = 0; b < nmaps; ++b)
      if (hipMemcpyAsync(m[b], d_map, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(SS_ERR_HIP, "copying the map failed");
    return ss::inpaint_maps_v(nmaps, m, t, l, H, W, iterations, thr_empty, thr_src, v, s);
  };
  for (int v = 0; v < ss::inpaint_tile_count() && rc == SS_OK; ++v) {
    rc = run(v);
    if (rc) break;
    if (hipEventRecord(e0, s) != hipSuccess) {
      rc = fail(SS_ERR_HIP, "hipEventRecord failed");
      break;
    }
    for (int i = 0; i < repeats && rc == SS_OK; ++i) rc = run(v);
    if (rc) break;
    float ms = 0.0f;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = fail(SS_ERR_HIP, "timing the inpainting variants failed");
      break;
    }
    ms /= (float)repeats;
    if (ms_out) ms_out[v] = ms;
    msv[v] = ms;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  const int best = pick_variant(msv, ss::inpaint_tile_count());
  set_inpaint_variant(H, W, best);
  if (best_out) *best_out = best;
  return ok();
}

int ss_merge(const float* d_sampling, const float* d_weights, const float* const* d_reproj,
             const float* const* d_dists, const uint8_t* const* d_imgs, int Hs, int Ws, const float* trans,
             const ss_intrinsics* intr, int R, int H, int W, int PH, int PW, float* d_dist_pano, uint8_t* d_rgb_pano,
             ss_stream stream) {
  SS_REQUIRE(d_sampling && d_weights && d_reproj && d_dists && d_imgs && trans && intr && d_dist_pano && d_rgb_pano,
             "null pointer argument");
  SS_REQUIRE(R >= 1 && R <= SS_MAX_REFERENCES, "reference count %d outside [1, %d]", R, SS_MAX_REFERENCES);
  SS_REQUIRE(H >= 2 && W >= 2 && Hs >= 2 && Ws >= 2 && PH >= 1 && PW >= 1, "bad merge shape");
  ss::MergeArgs a;
  memset(&a, 0, sizeof a);
  for (int r = 0; r < R; ++r) {
    SS_REQUIRE(d_reproj[r] && d_dists[r] && d_imgs[r], "null map or image pointer for reference %d", r);
    a.reproj[r] = d_reproj[r];
    a.dists[r] = d_dists[r];
    a.imgs[r] = d_imgs[r];
    a.intr[r] = intr[r];
    memcpy(a.trans[r], trans + 3 * r, sizeof(float) * 3);
  }
  a.sampling = reinterpret_cast<const float2*>(d_sampling);
  a.weights = d_weights;
  a.R = R;
  a.H = H;
  a.W = W;
  a.Hs = Hs;
  a.Ws = Ws;
  a.PH = PH;
  a.PW = PW;
  a.dist_out = d_dist_pano;
  a.rgb_out = d_rgb_pano;
  StageTimer t(SS_STAGE_MERGE, as_stream(stream));
  SS_HIP(ss::launch_merge(a, as_stream(stream)));
  return ok();
}

int ss_to_u8(const float* d_in, uint8_t* d_out, size_t n, ss_stream stream) {
  SS_REQUIRE(d_in && d_out, "null pointer argument");
  if (n == 0) return ok();
  const ss::ToU8Args a{d_in, d_out};
  const int rc = ss::to_u8_maps(&a, 1, n, as_stream(stream));
  return rc ? rc : ok();
}

int ss_resize_linear(const void* d_src, int src_type, int Hs, int Ws, int channels, float* d_dst, int Hd, int Wd,
                     ss_stream stream) {
  SS_REQUIRE(d_src && d_dst, "null pointer argument");
  SS_REQUIRE(src_type == SS_INPUT_U8 || src_type == SS_INPUT_U16 || src_type == SS_INPUT_F32, "unknown input type %d",
             src_type);
  SS_REQUIRE(channels >= 1 && channels <= 4, "channel count %d outside [1, 4]", channels);
  SS_REQUIRE(Hs >= 1 && Ws >= 1 && Hd >= 1 && Wd >= 1, "bad sizes %dx%d -> %dx%d", Ws, Hs, Wd, Hd);
  SS_REQUIRE((long long)Hd * Wd <= 0x7fffffffLL && (long long)Hs * Ws * channels <= 0x7fffffffLL,
             "image too large");
  SS_HIP(ss::launch_resize_linear(d_src, src_type, Hs, Ws, channels, d_dst, Hd, Wd, as_stream(stream)));
  return ok();
}

int ss_inverse_distance(const float* d_dist, float* d_inv, size_t n, ss_stream stream) {
  SS_REQUIRE(d_dist && d_inv, "null pointer argument");
  if (n == 0) return ok();
  SS_HIP(ss::launch_inverse(d_dist, d_inv, n, as_stream(stream)));
  return ok();
}

size_t ss_error_stats_workspace_bytes(size_t n) { return ss::error_stats_partial_bytes(n); }

int ss_error_stats(const float* d_a, const float* d_b, size_t n, const float* thresholds, int nthr, int finite_only,
                   double* d_out, void* d_workspace, size_t workspace_bytes, ss_stream stream) {
  SS_REQUIRE(d_a && d_b && d_out, "null pointer argument");
  SS_REQUIRE(n >= 1, "empty maps");
  SS_REQUIRE(nthr >= 0 && nthr <= ss::error_stats_max_thresholds() && (nthr == 0 || thresholds),
             "threshold count %d outside [0, %d]", nthr, ss::error_stats_max_thresholds());
  const size_t need = ss::error_stats_partial_bytes(n);
  if (!d_workspace || workspace_bytes < need)
    return fail(SS_ERR_WORKSPACE, "workspace of %zu bytes needed, %zu given", need, workspace_bytes);
  SS_HIP(ss::launch_error_stats(d_a, d_b, n, thresholds, nthr, finite_only, static_cast<double*>(d_workspace), d_out,
                                as_stream(stream)));
  return ok();
}

int ss_profile_enable(int on, unsigned stage_mask) {
  std::lock_guard<std::mutex> lk(g_prof.mu);
  if (on && g_prof.pool.empty()) {
    g_prof.pool.resize(kEventPool);
    for (auto& e : g_prof.pool) SS_HIP(hipEventCreate(&e));
  }
  g_prof.on = on != 0;
  g_prof.mask = stage_mask;
  return ok();
}

int ss_profile_collect(double* ms, int* launches) {
  SS_REQUIRE(ms && launches, "null output arrays");
  std::lock_guard<std::mutex> lk(g_prof.mu);
  for (int i = 0; i < SS_STAGE_COUNT; ++i) {
    ms[i] = 0.0;
    launches[i] = 0;
  }
  for (const auto& r : g_prof.recs) {
    SS_HIP(hipEventSynchronize(r.b));
    float t = 0.0f;
    SS_HIP(hipEventElapsedTime(&t, r.a, r.b));
    ms[r.stage] += t;
    launches[r.stage] += 1;
  }
  g_prof.recs.clear();
  g_prof.next = 0;
  return ok();
}

const char* ss_stage_name(int stage) {
  static const char* names[SS_STAGE_COUNT] = {"guide",      "sweep",     "isb_down", "isb_up",
                                              "isb_up_argmin", "dist_down", "dist_up",  "reproject",
                                              "inpaint",    "merge",     "to_u8"};
  return (stage >= 0 && stage < SS_STAGE_COUNT) ? names[stage] : "";
}

}  // extern "C"
