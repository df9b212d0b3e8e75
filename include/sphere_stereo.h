/*
 * sphere_stereo.h - C ABI of the MI355X (gfx950) sphere-sweeping stereo library
 * (libsphere_stereo.so).  It replaces the native and tensor-op stages of the
 * per-frame path of KAIST-VCLAB/sphere-stereo:
 *
 *   python/depth_estimation.py  RGBD_Estimator (sweep, cost, argmin, camera selection)
 *   python/isb_filter.py + .cu  ISB_Filter.apply, guideDownsample2xKernel, guideUpsample2xKernel
 *   python/stitcher.py + .cu    Stitcher, reprojectDistanceKernel, inpaintKernel,
 *                               createInpaintingWeightsKernel, createBlendingLutKernel,
 *                               mergeRGBDPanoramaKernel
 *
 * Conventions
 *   - Every pointer named d_* is device memory; everything else is host memory.
 *   - Every call is asynchronous on `stream` (a hipStream_t passed as void*;
 *     NULL = the null stream) and returns SS_OK or an ss_status code.  The text
 *     of the last error of the calling thread is returned by ss_last_error().
 *   - The library allocates no device memory in its entry points; callers pass
 *     the workspaces whose sizes ss_*_workspace_bytes() report.
 *   - Images are HWC float32 in the channel order of the caller (BGR, as cv2
 *     reads them).  The cost volume inside the library is candidate-contiguous:
 *     [H][W][Np] with Np = ss_padded_candidates(N) >= N, pads never read as
 *     results.  The reference's planar [N][H][W] layout is accepted by
 *     ss_isb_filter_planar (the drop-in for ISB_Filter.apply).
 */
#ifndef SPHERE_STEREO_H
#define SPHERE_STEREO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SS_ABI_VERSION 2
#define SS_MAX_CAMERAS 8
#define SS_MAX_CANDIDATES 128
#define SS_MAX_REFERENCES 8
#define SS_MAX_BATCH_FRAMES 64   /* frames of one ss_frame_batch_create */

typedef void *ss_stream; /* hipStream_t */

typedef enum {
    SS_OK = 0,
    SS_ERR_ARG = 1,         /* invalid argument (sizes, pointers, counts) */
    SS_ERR_HIP = 2,         /* a HIP runtime call failed */
    SS_ERR_UNSUPPORTED = 3, /* valid for the reference but outside this build's limits */
    SS_ERR_WORKSPACE = 4    /* workspace smaller than required */
} ss_status;

/* Element type of the stored cost volume.  All arithmetic is fp32; with F16 /
 * BF16 only the values written to the volume (sweep output, pyramid levels)
 * are rounded to nearest even.  F32 is the reference's type. */
typedef enum {
    SS_VOLUME_F32 = 0,
    SS_VOLUME_F16 = 1,
    SS_VOLUME_BF16 = 2
} ss_volume_type;

/* Double-sphere camera as the reference's torch code uses it (utils.py:40-94):
 * fs = fl * matching_scale and cs = principal * matching_scale are the fp32
 * products the reference forms; xi and alpha stay Python doubles. */
typedef struct {
    float fs[2];
    float cs[2];
    double xi;
    double alpha;
} ss_camera;

/* Stitcher intrinsics: the 24-byte Intrinsics record of stitcher.cu:31-35,
 * filled as vectorize_calibration (stitcher.py:35-46) fills it. */
typedef struct {
    float fx, fy, cx, cy, xi, alpha;
} ss_intrinsics;

/* Library ----------------------------------------------------------------- */

/* Has no counterpart in the reference (CuPy compiled the kernels per object). */
int ss_abi_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char *ss_last_error(void);
/* Padded candidate stride Np used by the candidate-contiguous volume. */
int ss_padded_candidates(int N);
/* Pyramid level count int(min(log2 W, log2 H) - 1) of isb_filter.py:62. */
int ss_isb_levels(int H, int W);

/* Init-time builders -------------------------------------------------------- */

/* RGBD_Estimator.select_camera for one reference (depth_estimation.py:83-135).
 * rts: C x 12 host floats, the 3x4 rows of inv(rt_k) @ rt_ref; d_masks: C x H x W;
 * d_sel: H x W int8, -1 where no camera qualifies. */
int ss_select_camera(const ss_camera *cams, const float *rts, int C, int ref, const float *d_masks, int H, int W,
                     double min_dist, double max_dist, int8_t *d_sel, ss_stream stream);

/* Per-pixel unit ray of the reference camera (utils.py:57-73, called at
 * depth_estimation.py:141-143) with the selected camera index stored in the
 * 4th lane as a float: d_raylut is H x W x 4 float32. */
int ss_ray_lut(const ss_camera *ref, const int8_t *d_sel, int H, int W, float *d_raylut, ss_stream stream);

/* createInpaintingWeightsKernel (stitcher.cu:212-272): d_lut is H x W x 2 uint8. */
int ss_inpaint_lut(const ss_intrinsics *intr, const float *t3, int H, int W, double min_dist, double max_dist,
                   uint8_t *d_lut, ss_stream stream);

/* Normalised (2r+1)^2 box filter of stitcher.py:145-148 / 162-164 over B planes
 * of H x W; outside reads pad_value.  d_tmp holds (H + 2r) x W floats. */
int ss_box_filter(const float *d_in, float *d_out, float *d_tmp, int B, int H, int W, int radius, float pad_value,
                  ss_stream stream);

/* createBlendingLutKernel (stitcher.cu:338-401).  rot: R x 9 host floats
 * (inverse rotations), trans: R x 3 host floats; d_masks: R x H x W smoothed
 * masks; outputs d_sampling R x PH x PW x 2 and d_weights R x PH x PW. */
int ss_blend_lut(const ss_intrinsics *intr, const float *rot, const float *trans, int R, const float *d_masks, int H,
                 int W, int PH, int PW, double min_dist, double max_dist, float *d_sampling, float *d_weights,
                 ss_stream stream);

/* d_w[r][p] /= sum_r d_w[r][p] (stitcher.py:164). */
int ss_normalize_weights(float *d_w, int R, int P, ss_stream stream);

/* Per-frame stages ---------------------------------------------------------- */

/* rgb2yCbCr + uint8 cast (utils.py:144-155, depth_estimation.py:227), packed
 * as one uint32 per pixel (bytes: ch0, ch1, ch2, 0). */
int ss_guide(const float *d_img, int H, int W, uint32_t *d_guide, ss_stream stream);

/* Sweep + raw cost (depth_estimation.py:141-175) into the candidate-contiguous
 * volume d_cost (H x W x Np elements of vtype).  d_images: C host-side device
 * pointers (H x W x 3 float32); rts as ss_select_camera; cands: N host floats
 * (1/linspace(1/min, 1/max, N)).  d_guide (nullable) also receives the packed
 * guide of d_images[ref], as ss_guide computes it, from the same pass. */
int ss_sweep(const float *d_raylut, const float *const *d_images, int C, const ss_camera *cams, const float *rts,
             int ref, const float *cands, int N, int H, int W, int vtype, void *d_cost, uint32_t *d_guide,
             ss_stream stream);

/* Sweep autotuning (no counterpart in the reference).  The sweep has
 * ss_sweep_variant_count() scheduling variants that compute identical bits:
 * variant v projects `chunk` candidates before loading their taps and stages
 * `group` candidates in LDS per flush (ss_sweep_variant_info).  Every sweep
 * of a configuration (H, W, N, vtype) uses the variant recorded for it, 0 if
 * none.  ss_sweep_autotune times every variant on the given inputs (one
 * warm-up and `repeats` timed launches each, synchronising `stream`; not
 * while it is being captured), records the fastest for the configuration
 * and returns it; ms_out (nullable) receives the per-launch milliseconds of
 * each variant.  d_cost is scratch of ss_sweep's size.  The record is per
 * process; graphs captured before it keep their variant. */
int ss_sweep_variant_count(void);
int ss_sweep_variant_info(int variant, int *chunk, int *group);
int ss_sweep_set_variant(int H, int W, int N, int vtype, int variant); /* -1 forgets the record */
int ss_sweep_get_variant(int H, int W, int N, int vtype);
int ss_sweep_autotune(const float *d_raylut, const float *const *d_images, int C, const ss_camera *cams,
                      const float *rts, int ref, const float *cands, int N, int H, int W, int vtype, void *d_cost,
                      int repeats, ss_stream stream, float *ms_out, int *best_out);

/* The cost filter's scheduling variants (no counterpart in the reference,
 * whose filter has one launch shape): lanes per pixel group (8, or 4 for
 * padded candidate counts up to 64) x tile order (XCD-aware or launch order).
 * Every variant gives the same bits.  The variant recorded for a
 * configuration (set, or picked by ss_estimate_filter_autotune; otherwise 0)
 * is used by ss_isb_filter, ss_isb_filter_planar, ss_isb_filter_noalloc,
 * ss_estimate_distance and the frame executors of ss_frame_create /
 * ss_frame_batch_create (not by the single ragged executor). */
int ss_isb_variant_count(void);
int ss_isb_variant_info(int variant, int *lanes, int *xcd_order);
int ss_isb_set_variant(int H, int W, int N, int vtype, int variant); /* -1 forgets the record */
int ss_isb_get_variant(int H, int W, int N, int vtype);

/* Bytes of workspace for ss_isb_filter (vtype volume) / ss_isb_filter_planar
 * (planar = 1, vtype must be SS_VOLUME_F32): pyramid levels 1..L-1 of guides
 * and volume, the guide-weight table of the call (768 floats), plus the
 * level-0 repack for the planar form. */
size_t ss_isb_workspace_bytes(int H, int W, int N, int vtype, int planar);

/* ISB_Filter.apply (isb_filter.py:72-136) on the library layout: d_guide
 * (H x W packed guide) and d_cost (H x W x Np of vtype) are filtered in place. */
int ss_isb_filter(uint32_t *d_guide, void *d_cost, int H, int W, int N, int vtype, double sigma_i, double sigma_s,
                  void *d_workspace, size_t workspace_bytes, ss_stream stream);

/* ss_isb_filter without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * scratch (what ss_isb_workspace_bytes(H, W, N, vtype, 0) reports) comes from
 * a static device array of SS_ISB_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends (e.g. 1024 x 1024 with 32 candidates needs 44 MiB, 768 x
 * 768 with 48 needs 37 MiB); larger calls return SS_ERR_WORKSPACE.  The call
 * allocates nothing.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (different streams, other threads) race.  Same
 * results as ss_isb_filter. */
#define SS_ISB_NOALLOC_SCRATCH_BYTES (48u << 20)
int ss_isb_filter_noalloc(uint32_t *d_guide, void *d_cost, int H, int W, int N, int vtype, double sigma_i,
                          double sigma_s, ss_stream stream);

/* Drop-in ISB_Filter.apply on the reference layout: d_guide H x W x 3 uint8,
 * d_cost N x H x W float32, both filtered in place. */
int ss_isb_filter_planar(uint8_t *d_guide, float *d_cost, int H, int W, int N, double sigma_i, double sigma_s,
                         void *d_workspace, size_t workspace_bytes, ss_stream stream);

/* Argmin, parabola refinement and index->distance (depth_estimation.py:180-198)
 * on a filtered candidate-contiguous volume.  d0, dN1: first/last candidate;
 * A = d0/dN1 - 1 in fp32.  d_dist: H x W. */
int ss_argmin_refine(const void *d_cost, int H, int W, int N, int vtype, float d0, float dN1, float A,
                     float *d_dist, ss_stream stream);

/* Bytes of workspace for ss_estimate_distance with a vtype volume. */
size_t ss_estimate_workspace_bytes(int H, int W, int N, int vtype);

/* estimate_fisheye_distance (depth_estimation.py:137-203) for one reference:
 * guide, sweep, cost filter with the final up-pass fused with the argmin,
 * then the distance filter at (sigma_i/2, sigma_s/2).  d_guide3 (H x W x 3
 * uint8) is the guide argument of the reference; NULL computes it from
 * d_images[ref] as estimate_RGBD_panorama does (depth_estimation.py:227).
 * The cost volume is stored as vtype; the distance map is always fp32.
 * d_dist receives the filtered H x W distance map. */
int ss_estimate_distance(const float *d_raylut, const float *const *d_images, int C, const ss_camera *cams,
                         const float *rts, int ref, const float *cands, int N, int H, int W, double sigma_i,
                         double sigma_s, int vtype, const uint8_t *d_guide3, float *d_dist, void *d_workspace,
                         size_t workspace_bytes, ss_stream stream);

/* Times the cost filter with its fused argmin (depth_estimation.py:177-198)
 * in every valid scheduling variant on the level-0 guide and volume that
 * d_workspace (an ss_estimate_distance workspace) holds from a previous
 * sweep, `repeats` runs each, and records the fastest for (H, W, N, vtype).
 * d_dist (H x W) receives the distance map; ms_out[ss_isb_variant_count()]
 * the per-run ms (-1 for variants that do not exist for N).  Not while the
 * stream is being captured. */
int ss_estimate_filter_autotune(void *d_workspace, size_t workspace_bytes, int H, int W, int N, int vtype,
                                const float *cands, double sigma_i, double sigma_s, float *d_dist, int repeats,
                                ss_stream stream, float *ms_out, int *best_out);

/* reprojectDistanceKernel (stitcher.cu:181-204) with stitcher.py:184-192's
 * fill: d_out is set to 1e8, then every source pixel takes part in an exact
 * atomic z-minimum at rint(y)*W + rint(x) (indices outside the map dropped). */
int ss_reproject(const float *d_in, float *d_out, int H, int W, const ss_intrinsics *intr, const float *t3,
                 ss_stream stream);

/* iterations passes of inpaintKernel (stitcher.cu:283-323) in Jacobi form:
 * d_map is updated, d_tmp (H x W) is scratch.  thr_empty = max_dist + 0.1;
 * thr_src = the reference's "MAX_DIST + 0.1f" as a double. */
int ss_inpaint(float *d_map, float *d_tmp, const uint8_t *d_lut, int H, int W, int iterations, double thr_empty,
               double thr_src, ss_stream stream);

/* Inpainting tile (no counterpart in the reference, whose inpaintKernel runs
 * one pass per launch): the multi-pass launches of ss_inpaint and of the
 * frame executors (ss_frame_create / ss_frame_batch_create) stage tiles of
 * ss_inpaint_variant_info(v) pixels a side (48, 32, 16) with a 16-pixel halo;
 * every tile gives the same map.  The variant recorded for the map size is
 * used, 0 if none.  ss_inpaint_autotune times every variant on nmaps copies
 * of d_map inpainted in one batch, as a frame's references are (into d_work,
 * with d_tmp as scratch, nmaps x H x W floats each; `repeats` runs each, not
 * while the stream is being captured), records the fastest for (H, W) and
 * returns it; ms_out (nullable) receives the per-run ms of each. */
int ss_inpaint_variant_count(void);
int ss_inpaint_variant_info(int variant, int *tile);
int ss_inpaint_set_variant(int H, int W, int variant); /* -1 forgets the record */
int ss_inpaint_get_variant(int H, int W);
int ss_inpaint_autotune(const float *d_map, int nmaps, float *d_work, float *d_tmp, const uint8_t *d_lut, int H,
                        int W, int iterations, double thr_empty, double thr_src, int repeats, ss_stream stream,
                        float *ms_out, int *best_out);

/* mergeRGBDPanoramaKernel (stitcher.cu:423-474).  d_reproj, d_dists, d_imgs:
 * R host-side arrays of device pointers (H x W float32 each; Hs x Ws x 3 uint8
 * each); intr: R records; trans: R x 3 host floats.  Outputs d_dist_pano
 * PH x PW float32 and d_rgb_pano PH x PW x 3 uint8. */
int ss_merge(const float *d_sampling, const float *d_weights, const float *const *d_reproj,
             const float *const *d_dists, const uint8_t *const *d_imgs, int Hs, int Ws, const float *trans,
             const ss_intrinsics *intr, int R, int H, int W, int PH, int PW, float *d_dist_pano, uint8_t *d_rgb_pano,
             ss_stream stream);

/* Truncating float32 -> uint8 cast of n values (depth_estimation.py:239-240). */
int ss_to_u8(const float *d_in, uint8_t *d_out, size_t n, ss_stream stream);

/* Input stage ----------------------------------------------------------------- */

/* Element types of a decoded input image (ss_resize_linear). */
typedef enum { SS_INPUT_U8 = 0, SS_INPUT_U16 = 1, SS_INPUT_F32 = 2 } ss_input_type;

/* Replaces read_input_images' per-image preparation (utils.py:157-204): the
 * HWC image (channels 1..4) converted to float32 in [0, 255] (uint8 as is,
 * uint16 / 255 as utils.py:176; a float32 image is taken as already in
 * [0, 255]) and resized to Hd x Wd as its cv2.resize(image, (cols, rows),
 * cv2.INTER_AREA) calls do (utils.py:198-201; that flag lands in the dst
 * slot, so the interpolation is INTER_LINEAR: half-pixel centres, source
 * coordinate (d + 0.5) * (S / D) - 0.5 in double rounded to float, clamped
 * to the first / last sample, a horizontal then a vertical a*w0 + b*w1).
 * d_dst: Hd x Wd x channels float32. */
int ss_resize_linear(const void *d_src, int src_type, int Hs, int Ws, int channels, float *d_dst, int Hd, int Wd,
                     ss_stream stream);

/* Output conversion and evaluation -------------------------------------------- */

/* 1 / distance (main.py:114): the inverse-distance map the reference saves and
 * evaluates, one IEEE fp32 division per element. */
int ss_inverse_distance(const float *d_dist, float *d_inv, size_t n, ss_stream stream);

/* The error statistics of evaluate_rgbd_panorama (utils.py:233-241) between
 * d_a (evaluated) and d_b (reference), n elements each.  d_out (device, 7
 * doubles) receives: sum |a-b|, sum (a-b)^2 (each error formed in fp32 as the
 * reference does, summed in fp64 in a fixed order), the number of pixels
 * counted, then for each of the nthr (<= 4) thresholds the number of pixels
 * with |a-b| > threshold.  MAE = [0]/[2], RMSE = sqrt([1]/[2]), bad-pixel
 * ratio = 100 * [3+k]/[2].  With finite_only, pixels where either map is not
 * finite are skipped (the reference counts every pixel, so a NaN propagates
 * into its sums). */
size_t ss_error_stats_workspace_bytes(size_t n);
int ss_error_stats(const float *d_a, const float *d_b, size_t n, const float *thresholds, int nthr, int finite_only,
                   double *d_out, void *d_workspace, size_t workspace_bytes, ss_stream stream);

/* Per-frame executor ---------------------------------------------------------- */

/* Everything estimate_RGBD_panorama (depth_estimation.py:205-243) needs besides
 * the frame's images: the rig, the per-reference tables built at set-up, the
 * device buffers the frame writes, and the panorama outputs. */
typedef struct {
    int C, R, N, H, W;                /* cameras, references, candidates, matching rows and cols */
    int Hs, Ws, PH, PW;               /* stitch image rows/cols, panorama rows/cols */
    int inpaint_iterations;
    int volume_type;                  /* ss_volume_type of the cost volumes */
    int refs[SS_MAX_REFERENCES];      /* camera index of each reference */
    ss_camera cams[SS_MAX_CAMERAS];
    float rts[SS_MAX_REFERENCES][SS_MAX_CAMERAS * 12]; /* per reference, as ss_sweep */
    float cands[SS_MAX_CANDIDATES];
    double sigma_i, sigma_s;
    double thr_empty, thr_src;        /* as ss_inpaint */
    const float *d_raylut[SS_MAX_REFERENCES];
    void *d_workspace[SS_MAX_REFERENCES]; /* ss_estimate_workspace_bytes(H, W, N, volume_type) each */
    size_t workspace_bytes;
    float *d_dist[SS_MAX_REFERENCES]; /* per-reference distance maps, H x W */
    ss_intrinsics intr[SS_MAX_REFERENCES];
    float trans[SS_MAX_REFERENCES][3];
    const uint8_t *d_inpaint_lut[SS_MAX_REFERENCES];
    float *d_reproj[SS_MAX_REFERENCES];
    float *d_scratch[SS_MAX_REFERENCES]; /* H x W each */
    uint8_t *d_stitch_u8[SS_MAX_REFERENCES]; /* Hs x Ws x 3 each */
    const float *d_sampling, *d_weights;
    float *d_dist_pano;
    uint8_t *d_rgb_pano;
} ss_frame_desc;

/* No counterpart in the reference (its per-frame loop is Python).  Creates an
 * executor for desc (copied) on the current device.  Graphs are captured on
 * one stream per device that every executor shares (created on first use, at
 * default priority, kept until the process ends).  Calls of ss_frame_run on
 * one device are serialised while they enqueue work. */
int ss_frame_create(const ss_frame_desc *desc, void **frame);
/* Enqueues one frame on `stream`, after the work already there.  Every pass
 * is one launch for all references: the sweep (which also writes the
 * guides), each level of the cost filter, the fused argmin, each level of
 * the distance filter, the reprojection, the inpainting passes, the uint8
 * casts of the stitch images, then the merge.  d_match: C float32 H x W x 3
 * images; d_stitch: R float32 Hs x Ws x 3 images.  With use_graph, the frame
 * is captured into a HIP graph on the first call (and again whenever the
 * image pointers change) and replayed on `stream` afterwards. */
int ss_frame_run(void *frame, const float *const *d_match, const float *const *d_stitch, int use_graph,
                 ss_stream stream);
/* No counterpart in the reference (the batching form; the reference runs one
 * frame per call).  Creates an executor for nframes frames (descs copied)
 * that share one configuration: C, R, N, H, W, Hs, Ws, PH, PW, the volume
 * type, the candidates, sigma_i, sigma_s and the inpainting settings must be
 * equal; the rigs (cameras, transforms, viewpoint tables), the buffers and the
 * images are each frame's own.  Every pass after the sweep runs once for a
 * chunk of frames (as many as fit 16 frame-reference pairs), so the
 * latency-bound coarse pyramid levels cost one launch per chunk.  Free with
 * ss_frame_destroy. */
int ss_frame_batch_create(const ss_frame_desc *descs, int nframes, void **frame);
/* Enqueues the batch's frames on `stream`, after the work already there.
 * d_match: nframes * C images (frame-major), d_stitch: nframes * R images,
 * as ss_frame_run's for each frame; use_graph as ss_frame_run. */
int ss_frame_batch_run(void *frame, const float *const *d_match, const float *const *d_stitch, int use_graph,
                       ss_stream stream);
/* No counterpart in the reference (its frames run one after the other on one
 * stream).  Sets the lanes (1 to 4, default 1) of an executor of
 * ss_frame_batch_create: with L > 1 a run cuts its frames into L or more
 * chunks of whole frames (at most one chunk per frame) that alternate over L
 * streams of the executor (created here, default priority): chunk k's sweeps
 * start when chunk k - 1's are over, so each chunk's filters, reprojection,
 * inpainting and merges run beside the next chunk's sweeps.  The stream
 * given to the run still waits for all of it, and every result is the same,
 * bit for bit.  An executor of one frame runs on one stream whatever its
 * lanes.  Drops a captured graph.  The ragged executor has one lane. */
int ss_frame_set_lanes(void *frame, int lanes);
/* The executor's lanes (-1 for a null handle). */
int ss_frame_get_lanes(void *frame);
/* No counterpart in the reference (the batching form).  The ragged persistent
 * form of ss_frame_batch_create: the nframes frames may differ in every size
 * and setting (C, R, N, H, W, Hs, Ws, PH, PW, candidates, sigmas, inpainting
 * settings); they must share the volume type and have N >= 2 and a matching
 * resolution of at least 8 x 8.  Each pass of the frame (the sweep, each
 * level of the two filters, the coarse tails, the reprojection, each group of
 * 16 inpainting passes, the uint8 casts, the merge) is one launch for all the
 * frames (the filter levels one per padded candidate count present; the
 * sweep one per tuned sweep variant).  In
 * the persistent form a launch is sized to the workgroups the GPU holds
 * resident and each worker loops over work items (row segments, pixel
 * blocks, tiles); in the dispatched form each worker takes one item.  Each
 * XCD takes an eighth of every frame-reference pair, pair after pair.  The launch records, the filters' weight tables and
 * the per-run image pointer slots live in d_table, a device buffer of
 * ss_frame_ragged_table_bytes(descs, nframes) bytes (256-byte aligned) that
 * the caller keeps until ss_frame_destroy; it is written on `stream` during
 * the call, which returns once it is ready.  Runs go through
 * ss_frame_batch_run (d_match: every frame's C images, frame after frame;
 * d_stitch: every frame's R images), with or without a graph.  Runs of one
 * executor share the pointer slots: they must not overlap on different
 * streams.  Returns 0 from ss_frame_ragged_table_bytes for descriptors it
 * rejects. */
size_t ss_frame_ragged_table_bytes(const ss_frame_desc *descs, int nframes);
/* schedule: SS_RAGGED_PERSISTENT (every pass in the persistent form),
 * SS_RAGGED_DISPATCHED (every pass with a worker per item, balanced by the
 * hardware dispatcher), SS_RAGGED_AUTO (per pass, the form measured faster). */
#define SS_RAGGED_PERSISTENT 0
#define SS_RAGGED_DISPATCHED 1
#define SS_RAGGED_AUTO 2
int ss_frame_ragged_create(const ss_frame_desc *descs, int nframes, int schedule, void *d_table, size_t table_bytes,
                           ss_stream stream, void **frame);
int ss_frame_destroy(void *frame);

/* Differentiation form --------------------------------------------------------
 * No counterpart in the reference (it has no backward).  The forward is the
 * plain path's ss_ray_lut + ss_sweep + ss_isb_filter on an fp32 volume, read
 * out by ss_soft_argmin instead of the hard argmin + parabola of
 * depth_estimation.py:180-198; the calls below are the adjoints of those
 * stages, in reverse order, and the Gauss-Newton pieces of a photometric
 * calibration refinement.  All require N >= 2. */

/* Soft-argmin readout: iota = sum_n n softmax_n(-beta * F_n) over the filtered
 * fp32 volume d_cost (H x W x Np), d_invd[p] = inv0 + (inv1 - inv0) * iota /
 * (N - 1), where inv0 = 1/d_0 and inv1 = 1/d_{N-1} (the candidates are uniform
 * in inverse distance, depth_estimation.py:145-146). */
int ss_soft_argmin(const float *d_cost, int H, int W, int N, float beta, float inv0, float inv1, float *d_invd,
                   ss_stream stream);

/* Its adjoint: d_grad_cost (H x W x Np, pads written 0) = dL/dF given
 * d_grad_invd = dL/d(invd) (H x W) and the same filtered volume. */
int ss_soft_argmin_backward(const float *d_cost, const float *d_grad_invd, int H, int W, int N, float beta,
                            float inv0, float inv1, float *d_grad_cost, ss_stream stream);

/* Adjoint of ss_isb_filter (ISB_Filter.apply, isb_filter.py:72-136) on an fp32
 * volume, which for a fixed guide is linear in the cost.  d_guide: the level-0
 * packed guide the forward filter was given (before the filter updated it);
 * d_grad_out: dL/d(filtered volume); d_grad_in (not aliasing it): dL/d(input
 * volume).  The guide pyramid is recomputed, no forward level is needed. */
size_t ss_isb_backward_workspace_bytes(int H, int W, int N);
int ss_isb_filter_backward(const uint32_t *d_guide, const float *d_grad_out, float *d_grad_in, int H, int W, int N,
                           double sigma_i, double sigma_s, void *d_workspace, size_t workspace_bytes,
                           ss_stream stream);

/* Adjoint of ss_sweep (depth_estimation.py:141-175, with the clamp at 500)
 * for reference `ref`, arguments as ss_sweep plus d_grad_cost = dL/d(clamped
 * cost volume).  d_grad_images (nullable; entries nullable): C HWC float32
 * buffers to which dL/d(image) is ADDED.  d_grad_params receives (C + 1) x 18
 * floats: record k < C is camera k's (fs0, fs1, cs0, cs1, xi, alpha) followed
 * by its 3 x 4 (camera <- reference) transform row-major, as sampled from this
 * reference; record C holds the reference camera's own 6 parameters through
 * its rays.  The selection and the bilinear cell are held fixed. */
size_t ss_sweep_backward_workspace_bytes(int H, int W, int C);
int ss_sweep_backward(const float *d_raylut, const float *const *d_images, int C, const ss_camera *cams,
                      const float *rts, int ref, const float *cands, int N, int H, int W, const float *d_grad_cost,
                      float *const *d_grad_images, float *d_grad_params, void *d_workspace, size_t workspace_bytes,
                      ss_stream stream);

/* Photometric Gauss-Newton normal equations of one (reference, target camera)
 * pair: each reference pixel with d_invd > 0 and a selected camera (d_raylut
 * as ss_ray_lut writes it; with skip_selected, one other than `target`) is placed at ray / invd, moved into the
 * target by rt (3 x 4, target <- reference), projected with *cam and compared
 * with the target image (residual per channel, Huber weight at `huber`);
 * pixels outside the image interior or below 0.9 of d_mask (nullable) are
 * skipped.  Parameters: a right perturbation exp(delta) of the target's pose
 * (rotation 3, translation 3) and relative scales of fs0 and fs1.  Writes one
 * record of ss_photometric_record_floats() floats per block of
 * ss_photometric_blocks(H, W) blocks at d_partial. */
int ss_photometric_blocks(int H, int W);
int ss_photometric_record_floats(void);
int ss_photometric_normal_equations(const float *d_raylut, const float *d_invd, const float *d_ref_image,
                                    const float *d_image, const float *d_mask, const ss_camera *cam, const float *rt,
                                    int target, int skip_selected, int H, int W, float huber, float *d_partial,
                                    ss_stream stream);
/* Sums nblocks records (in fp64) and solves (H + lambda diag H) delta = -b in
 * fp64 on the GPU (Cholesky).  d_out: 64 doubles; d_out[0..11) = delta[8],
 * loss, residual count, and 1.0 if the factorisation succeeded (else delta =
 * 0); the rest is scratch. */
int ss_gauss_newton_solve(const float *d_partial, int nblocks, double lambda, double *d_out, ss_stream stream);

/* Kernel timing ------------------------------------------------------------- */
/* No counterpart in the reference (it has no timers).  While enabled, every
 * launch of a stage whose bit is set in stage_mask is bracketed by a pair of
 * hipEvents recorded on the launch's stream.  ss_profile_enable(0, 0) stops
 * recording; ss_profile_collect waits for the recorded events and returns,
 * per stage, the summed milliseconds and the launch count (arrays of
 * SS_STAGE_COUNT), then clears the records. */
enum {
    SS_STAGE_GUIDE = 0,
    SS_STAGE_SWEEP,
    SS_STAGE_ISB_DOWN,
    SS_STAGE_ISB_UP,
    SS_STAGE_ISB_UP_ARGMIN,
    SS_STAGE_DIST_DOWN,
    SS_STAGE_DIST_UP,
    SS_STAGE_REPROJECT,
    SS_STAGE_INPAINT,
    SS_STAGE_MERGE,
    SS_STAGE_TO_U8,
    SS_STAGE_COUNT
};
int ss_profile_enable(int on, unsigned stage_mask);
int ss_profile_collect(double *ms, int *launches);
const char *ss_stage_name(int stage);

#ifdef __cplusplus
}
#endif
#endif /* SPHERE_STEREO_H */
