// Kernel argument records and launch functions (host side of the .hip files).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ss_device.h"
#include "ss_types.h"
#include "ss_volume.h"

namespace ss {

// Volumes are passed untyped; the kernels cast them to their storage type.
// The guide weights exp(-d^2 var_inv_i) of a filter depend on the guide
// distance d alone, an integer in [0, 765] (a sum of three byte
// differences): each filter call tabulates them once (kWeightTable floats,
// launch_weight_tables) and its passes look them up.
constexpr int kWeightTable = 768;

struct IsbDownArgs {
  const uint32_t* gin;
  const void* cin;
  int hin, win;
  uint32_t* gout;
  void* cout;
  int hout, wout;
  const float* wtab;
};

struct ArgminParams {
  int N;
  float d0, dN1, A;
};

struct IsbUpArgs {
  const uint32_t* gin;
  const void* cin;
  int hin, win;
  uint32_t* gio;
  void* cio;
  int hio, wio;
  float weight_up, weight_down;
  const float* wtab;
  float* dist;  // fused argmin output (level-0 distance map)
  ArgminParams am;
  float* fill;  // nullable: with the fused argmin, set to 1e8 at every level-0 pixel
  int skip_guide = 0;  // 1: the fine level's guide is not needed afterwards (not written)
  // the one-lane level-0 pass of a distance filter may also reproject every
  // fine pixel's final distance (reproject_splat) into rp_out (1e8-filled)
  float* rp_out = nullptr;
  ss_intrinsics rp_c = {};
  float rp_t[3] = {0.0f, 0.0f, 0.0f};
};

// One launch runs the same pass for up to kMaxBatch references of equal
// shape (a frame's references, or those of a batch of frames): the kernel
// takes entry blockIdx.y of the table (blockIdx.z for grids of 2-D tiles).
// Only the buffers differ between entries, so one launch replaces a chain per
// reference, and the small coarse-level passes, which cannot fill the GPU,
// cost one launch per batch.  (16 entries keep the largest table, 16 up-pass
// records, at 1.5 KB of kernel arguments.)
constexpr int kMaxBatch = 16;
static_assert(kMaxBatch >= SS_MAX_REFERENCES, "a frame's references fit one launch");
template <typename A>
struct Batch {
  A a[kMaxBatch];
};

// Levels s0..s0+n-1 of a pyramid for the single-workgroup coarse tail:
// entry 0 in global memory, entries 1..n-1 at byte offsets in LDS.  Workgroup
// b runs the tail of batch entry b (its own level-s0 buffers g0[b], c0[b]).
constexpr int kMaxTailLevels = 16;
constexpr int kTailLdsBytes = 60 * 1024;
struct TailLevels {
  int n;
  unsigned goff[kMaxTailLevels], coff[kMaxTailLevels];
  int h[kMaxTailLevels], w[kMaxTailLevels];
  float wu[kMaxTailLevels], wd[kMaxTailLevels];  // up pass into entry k
  const float* wtab;
  unsigned wtab_off;  // the tail's LDS copy of wtab
};
static_assert(kTailLdsBytes + kWeightTable * 4 + 16 <= 64 * 1024, "tail levels and weight table in 64 KiB of LDS");
struct IsbTailArgs {
  TailLevels t;
  uint32_t* g0[kMaxBatch];
  void* c0[kMaxBatch];
  unsigned lds_bytes;
};

// One sweep launch covers up to SS_MAX_REFERENCES references of a rig whose
// (reference, camera) transforms fit kSweepMaxPairs; ref[r].rt0 is the index
// of reference r's first transform in rt.
constexpr int kSweepMaxPairs = 32;
struct SweepRef {
  const float4* raylut;
  const float* img;  // the reference's own image
  void* cost;
  uint32_t* guide;   // nullable: packed guide written from the same pass
  int rt0;
};
// Images converted to uint8 (the stitcher's cast, ToU8Args) by `blocks`
// extra workgroups at the end of a sweep launch: they run in the slots the
// sweep's last round of tiles leaves free.  Each image has `quads` groups of 4
// values (16-byte aligned input, 4-byte aligned output); blocks == 0: none.
struct SweepCast {
  const float* in[SS_MAX_REFERENCES];
  uint8_t* out[SS_MAX_REFERENCES];
  int n, blocks;
  unsigned quads;
};
struct SweepArgs {
  SweepRef ref[SS_MAX_REFERENCES];
  const float* img[SS_MAX_CAMERAS];
  TorchCam cam[SS_MAX_CAMERAS];
  float rt[kSweepMaxPairs][12];
  float cands[SS_MAX_CANDIDATES];
  int R, H, W, N, NP;            // NP: candidate stride of the volume (a multiple of 8)
  int tiles_x, blocks_per_ref;   // 64 x 4 pixel tiles per row of tiles / per reference
  // grid_sample's coordinate round trip: size, 1/size, 2/size, size/2
  float Wf, Hf, invW, invH, inv2W, inv2H, halfW, halfH;
  int pow2;      // both sizes are powers of two
  int fastproj;  // the parameters are within sweep_fast_bounds: project_t_fast applies
  SweepCast cast;
};

struct SelectArgs {
  TorchCam ref;
  TorchCam cam[SS_MAX_CAMERAS];
  float rt[SS_MAX_CAMERAS][12];
  const float* masks;
  int8_t* sel;
  int C, refidx, H, W;
  float mn, mx;
  float Wf, Hf, invW, invH;
  int pow2W, pow2H;
};

// per-reference buffers of the stitching passes
struct ReprojectArgs {
  const float* din;
  float* dout;
  ss_intrinsics c;
  float t[3];
};
struct InpaintArgs {
  const float* din;
  float* dout;
  const uint8_t* lut;
};
struct ToU8Args {
  const float* in;
  uint8_t* out;
};

struct MergeArgs {
  const float2* sampling;
  const float* weights;
  const float* reproj[SS_MAX_REFERENCES];
  const float* dists[SS_MAX_REFERENCES];
  const uint8_t* imgs[SS_MAX_REFERENCES];
  ss_intrinsics intr[SS_MAX_REFERENCES];
  float trans[SS_MAX_REFERENCES][3];
  int R, H, W, Hs, Ws, PH, PW;
  float* dist_out;
  uint8_t* rgb_out;
};

struct BlendArgs {
  ss_intrinsics intr[SS_MAX_REFERENCES];
  float rot[SS_MAX_REFERENCES][9];
  float trans[SS_MAX_REFERENCES][3];
  const float* masks;
  float2* sampling;
  float* weights;
  int R, H, W, PH, PW;
  float mn, mx;
};

// ---------------------------------------------------------------------------
// Ragged persistent launches (ragged.hip): one launch per pass covers the
// frame-reference pairs of many frames whatever their shapes.  Each launch is
// sized to what the GPU holds resident and its workers (wavefronts, or
// workgroups for passes with workgroup barriers) loop over the launch's work
// items.  The records below live in a device table built once per batch; a
// launch gets a span of them.
// ---------------------------------------------------------------------------
// Scheduling: every entry's items are split into kParts contiguous chunks,
// chunk x = items [x n / kParts, (x + 1) n / kParts) of an entry of n items;
// the workgroups with blockIdx.x % kParts == x (dispatched round-robin over
// the 8 XCDs, so all on one XCD) take chunk x of every entry, entry after
// entry.  All XCDs thus move through the entries together, as in a launch
// per entry, and the data of the entries in flight is shared in the
// last-level cache; each XCD gets an eighth of every entry, so the XCDs are
// balanced without a cost model.  No atomics: measured on MI355X,
// device-scope atomic adds on shared counters serialise at a few ns each
// across the chip, which made a dynamic work queue slower than the hardware
// dispatcher on every pass shorter than the sweep.
constexpr int kParts = 8;

// Entries e[0..n) of a launch; entry k owns items [first[k], first[k+1]),
// first[n] = total.  xfirst[x * (n + 1) + k]: the position of entry k's chunk
// x in XCD x's sequence (xtotal[x] = xfirst[x * (n + 1) + n]).
template <typename E>
struct RaggedSpan {
  const E* e;
  const int* first;
  const int* xfirst;
  int n, total;
  int xtotal[kParts];
};

// Items of the ragged ISB passes: one wavefront step, 64 / G coarse pixels
// (G = 8: groups of 8 lanes per pixel; G = 1: a lane per pixel).
constexpr int kIsbItemPix8 = 8, kIsbItemPix1 = 64;

// One reference of one frame in the ragged sweep: an item is a 64 x 4 pixel
// tile, swept by one workgroup (a wavefront per row), as in k_sweep.
struct SweepPair {
  const float4* raylut;
  void* cost;
  uint32_t* guide;
  int img0;    // index of the frame's camera 0 in the run's image table
  int refcam;  // camera index of the reference
  int H, W, N, NP, tiles_x, pow2, fastproj;
  float Wf, Hf, invW, invH, inv2W, inv2H, halfW, halfH;
  TorchCam cam[SS_MAX_CAMERAS];
  float rt[SS_MAX_CAMERAS][12];
  float cands[SS_MAX_CANDIDATES];
};

struct IsbDownEntry {
  IsbDownArgs a;
  int V;
};
struct IsbUpEntry {
  IsbUpArgs a;
  int V;
};

// One pyramid's coarse tail in the ragged tail launch (one workgroup each).
struct TailEntry {
  TailLevels t;
  uint32_t* g0;
  void* c0;
  int V;
};

struct ReprojEntry {
  ReprojectArgs a;
  int H, W;
};
// steps == 0: copy din to dout (the last launch of an odd count)
struct InpaintEntry {
  InpaintArgs a;
  int H, W, steps, tiles_x;
  double thr_empty, thr_src;
};
// per-run pointers written into the device table before a ragged run
constexpr int kPublishMax = 64;
struct PublishArgs {
  const void* src[kPublishMax];
  const void** dst;
  int n;
};
struct CastEntry {
  int src;  // index in the run's stitch-image table
  uint8_t* out;
  long long n;
};

// api.hip: error text of the calling thread; whether kernel timing is on
int report_error(int code, const char* what);
bool profiler_active();

// api.hip: pyramid levels (level s: ceil(H / 2^s) x ceil(W / 2^s)) and their buffers
struct PyrLevel {
  int h, w;
  uint32_t* g;
  void* c;
};
int isb_levels(int H, int W);
void up_weights(double sigma_s, int k, double* wu, double* wd);
int tail_start(const PyrLevel* l0, int L, int G, int NP);
unsigned tail_levels(const PyrLevel* l0, int s0, int L, int NP, const float* wtab, double sigma_s, TailLevels* t);
// an estimate workspace's cost pyramid lv[0..L) (level 0 = guide + volume)
// and distance pyramid dl[0..L) (level 0 = guide + dist); returns L
int estimate_levels(void* ws, int H, int W, int N, int vt, float* dist, PyrLevel* lv, PyrLevel* dl);
bool sweep_fast_bounds(const TorchCam* cams, int C, const float* rt, int npairs, const float* cands, int N);

// isb.hip (vt: ss_volume_type of the level-0 side of a pass, i.e. the input
// of a down pass or the fine level of an up pass; pyramid levels >= 1 are
// fp32, and so is everything with G = 1).  a[0..nb) are the batch entries.
// xcd: XCD-aware tile order (1) or launch order (0); G: 1, 4 or 8 lanes per pixel
hipError_t launch_isb_down(const IsbDownArgs* a, int nb, int G, int V, int vt, int xcd, hipStream_t s);
// reproj: G = V = 1 only, the entries' rp_* fields set
hipError_t launch_isb_up(const IsbUpArgs* a, int nb, int G, int V, int vt, bool fuse_argmin, int xcd, hipStream_t s,
                         bool reproj = false);
hipError_t launch_isb_tail(const IsbTailArgs& t, int nb, int G, int V, hipStream_t s);
// n weight tables in one launch: tab[i][d] = exp(-d^2 vi[i]) for d < kWeightTable
hipError_t launch_weight_tables(const float* vi, float* const* tab, int n, hipStream_t s);
hipError_t launch_argmin(const void* cost, int npix, int V, int vt, const ArgminParams& p, float* dist,
                         hipStream_t s);
hipError_t launch_planar_to_cc(const float* src, float* dst, int HW, int N, int NP, hipStream_t s);
hipError_t launch_cc_to_planar(const float* src, float* dst, int HW, int N, int NP, hipStream_t s);
hipError_t launch_guide3_to_packed(const uint8_t* g3, uint32_t* g, int HW, hipStream_t s);
hipError_t launch_packed_to_guide3(const uint32_t* g, uint8_t* g3, int HW, hipStream_t s);

// sweep.hip
hipError_t launch_guide(const float* img, int HW, uint32_t* guide, hipStream_t s);
hipError_t launch_ray_lut(const TorchCam& ref, const int8_t* sel, int H, int W, float4* lut, hipStream_t s);
hipError_t launch_select_camera(const SelectArgs& a, hipStream_t s);
hipError_t launch_sweep(const SweepArgs& a, int vt, int variant, hipStream_t s);
// the sweep's scheduling variants (candidate chunk, flush group); 0 is the default
int sweep_variant_count();
void sweep_variant_info(int v, int* chunk, int* group);

// api.hip: pieces of ss_estimate_distance the frame executor composes
size_t volume_elem_bytes(int vt);
// level-0 guide and volume of an estimate workspace
void estimate_level0(void* ws, int H, int W, int N, int vt, uint32_t** g0, void** c0);
// everything after the sweep, for nb references in the same launches: cost
// filter + argmin, then the distance filter; fill (nullable, else nb maps of
// H x W) is set to 1e8 on the way (the reprojection targets)
// the cost filter (fused argmin included) timed per scheduling variant on the
// level-0 guide and volume of an estimate workspace; records the fastest
int time_filter_variants(void* ws, int H, int W, int N, int vt, const float* cands, double sigma_i, double sigma_s,
                         float* d_dist, int repeats, hipStream_t s, float* ms_out, int* best_out);
// tables_ready: entry 0's workspace already holds the guide weight tables of
// sigma_i (the frame executor computes them on its first eager run)
// rp_intr / rp_t (optional, with fill): the distance filter's last pass also
// reprojects the final distances into fill[b] (reproject_into's work);
// *reprojected tells whether it did (it cannot without an up pass, L == 1)
int estimate_after_sweep(int nb, int H, int W, int N, int vt, const float* cands, double sigma_i, double sigma_s,
                         float* const* d_dist, void* const* ws, hipStream_t s, float* const* fill, int isb_variant = -1,
                         bool distance_filter = true, bool tables_ready = false,
                         const ss_intrinsics* rp_intr = nullptr, const float* const* rp_t = nullptr,
                         bool* reprojected = nullptr);
// the z-tested reprojection of nb maps, with or without the 1e8 fill of the
// targets first
int reproject_into(int nb, const float* const* d_in, float* const* d_out, int H, int W, const ss_intrinsics* intr,
                   const float* const* t3, bool fill, hipStream_t s);
// uint8 casts of nb images of n values each, in one launch
int to_u8_maps(const ToU8Args* a, int nb, size_t n, hipStream_t s);
// `iterations` inpainting passes over nb maps (the result ends in map[b])
int inpaint_maps(int nb, float* const* map, float* const* tmp, const uint8_t* const* lut, int H, int W,
                 int iterations, double thr_empty, double thr_src, hipStream_t s);
// one sweep launch over R references (each with its own workspace)
// cast (optional): ncast images of cast_n values converted to uint8 by the
// last sweep launch's trailing workgroups when sizes and alignment allow
// (*cast_done tells whether they were; otherwise the caller casts them)
int sweep_references(const float* const* d_raylut, const float* const* d_images, int C, const ss_camera* cams,
                     const float* const* rts, const int* refs, int R, const float* cands, int N, int H, int W,
                     int vt, void* const* ws, hipStream_t s, const ToU8Args* cast = nullptr, int ncast = 0,
                     size_t cast_n = 0, bool* cast_done = nullptr);

// stitch.hip
hipError_t launch_fill(float* p, float v, size_t n, hipStream_t s);
hipError_t launch_reproject(const ReprojectArgs* a, int nb, int H, int W, hipStream_t s);
hipError_t launch_inpaint_step(const float* din, float* dout, const uint8_t* lut, int H, int W, double thr_empty,
                               double thr_src, hipStream_t s);
// tile: the side of a workgroup's tile, inpaint_tile_of(variant) for the
// inpaint_tile_count() variants (48, 32, 16); every tile gives the same map
hipError_t launch_inpaint_multi(const InpaintArgs* a, int nb, int H, int W, int steps, double thr_empty,
                                double thr_src, int tile, hipStream_t s);
int inpaint_steps_per_launch();
int inpaint_tile_count();
int inpaint_tile_of(int variant);
// inpaint_maps with the tile of the variant recorded for (H, W), or of
// `variant` when >= 0
int inpaint_maps_v(int nb, float* const* map, float* const* tmp, const uint8_t* const* lut, int H, int W,
                   int iterations, double thr_empty, double thr_src, int variant, hipStream_t s);
hipError_t launch_inpaint_lut(const ss_intrinsics& c, const float* t, int H, int W, float mn, float mx, uint8_t* lut,
                              hipStream_t s);
hipError_t launch_box_filter(const float* in, float* out, float* tmp, int B, int H, int W, int r, float pad,
                             hipStream_t s);
hipError_t launch_blend_lut(const BlendArgs& a, hipStream_t s);
hipError_t launch_normalize(float* w, int R, int P, hipStream_t s);
hipError_t launch_merge(const MergeArgs& a, hipStream_t s);
hipError_t launch_to_u8(const ToU8Args* a, int nb, size_t n, hipStream_t s);

// input.hip
hipError_t launch_resize_linear(const void* src, int src_type, int Hs, int Ws, int C, float* dst, int Hd, int Wd,
                                hipStream_t s);

// eval.hip
hipError_t launch_inverse(const float* d, float* inv, size_t n, hipStream_t s);
int error_stats_blocks(size_t n);
size_t error_stats_partial_bytes(size_t n);
int error_stats_max_thresholds();
hipError_t launch_error_stats(const float* a, const float* b, size_t n, const float* thr, int nthr, int finite_only,
                              double* partial, double* out, hipStream_t s);

// Ragged launches.  loop = true: the persistent form (ragged_grid sizes the
// grid to the resident workgroups, and every worker loops over its items);
// loop = false: the dispatched form (a worker per item, at most one each).
int ragged_grid_of(const void* kernel, int block, size_t lds, const int* xtotal, int workers_per_block, bool loop);
template <typename E>
int ragged_grid(const void* kernel, int block, size_t lds, const RaggedSpan<E>& sp, int workers_per_block, bool loop) {
  return ragged_grid_of(kernel, block, lds, sp.xtotal, workers_per_block, loop);
}
hipError_t launch_publish(const void* const* src, int n, const void** dst, hipStream_t s);
// variant: one of the sweep's scheduling variants (sweep_variant_info)
hipError_t launch_sweep_ragged(const RaggedSpan<SweepPair>& sp, const float* const* img, int vt, int variant,
                               bool loop, hipStream_t s);
// vmax: the V of the span's entries (ragged_vmax: one V per launch)
int ragged_vmax(int V);
hipError_t launch_isb_down_ragged(const RaggedSpan<IsbDownEntry>& sp, int G, int vt, int vmax, bool loop,
                                  hipStream_t s);
hipError_t launch_isb_up_ragged(const RaggedSpan<IsbUpEntry>& sp, int G, int vt, int vmax, bool fuse_argmin, bool loop,
                                hipStream_t s);
hipError_t launch_isb_tail_ragged(const TailEntry* t, int n, int G, unsigned lds_bytes, hipStream_t s);
hipError_t launch_reproject_ragged(const RaggedSpan<ReprojEntry>& sp, bool loop, hipStream_t s);
hipError_t launch_inpaint_ragged(const RaggedSpan<InpaintEntry>& sp, bool loop, hipStream_t s);
hipError_t launch_to_u8_ragged(const RaggedSpan<CastEntry>& sp, const float* const* src, bool loop, hipStream_t s);
// rn: the reference count of every entry of sp (1-4), or 0 (any count, a loop)
hipError_t launch_merge_ragged(const RaggedSpan<MergeArgs>& sp, int rn, bool loop, hipStream_t s);
// items of one entry of each ragged pass
inline int sweep_items(int H, int W) { return ((H + 3) / 4) * ((W + 63) / 64); }
inline int isb_items(int npix, int G) {
  const int per = G == 1 ? kIsbItemPix1 : kIsbItemPix8;
  return (npix + per - 1) / per;
}
constexpr int kPixItem = 64;    // reprojection and merge: pixels per item (a lane each)
constexpr int kCastRounds = 8;                // uint8 cast: rounds of 4 values per lane in an item
constexpr int kCastItem = 256 * kCastRounds;  // uint8 cast: values per item (a wavefront)
inline int pix_items(long long n) { return (int)((n + kPixItem - 1) / kPixItem); }
inline int cast_items(long long n) { return (int)((n + kCastItem - 1) / kCastItem); }
int inpaint_tile();  // the side of an inpainting item (a tile of the map)

// ragged.hip: the ragged persistent executor of a batch of frames
struct RaggedPlan;
size_t ragged_table_bytes(const ss_frame_desc* d, int F);
int ragged_create(const ss_frame_desc* d, int F, int schedule, void* table, size_t bytes, hipStream_t s,
                  RaggedPlan** out);
int ragged_enqueue(const RaggedPlan* pl, const float* const* match, const float* const* stitch, hipStream_t s);
int ragged_images(const RaggedPlan* pl, int* n_match, int* n_stitch);
void ragged_free(RaggedPlan* pl);

}  // namespace ss
