/*
 * tdvh.h — C ABI of the MI355X (gfx950) stereo-matching hot path.
 *
 * Plain C: pointers, sizes and status codes. Device pointers are HIP device
 * addresses; `stream` is a hipStream_t passed as void* (NULL = the null
 * stream). All calls are asynchronous on `stream` unless stated otherwise.
 * Every call returns TDVH_OK (0) or a negative tdvh_status; the message of
 * the last failure on the calling thread is available from
 * tdvh_last_error().
 *
 * Each entry point names the reference function it replaces (paths relative
 * to LibThunderVision/src/tdvision/ of otaviog/ThunderVision).
 */
#ifndef TDVH_H
#define TDVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ----------------------------------------------------------- */
typedef enum {
    TDVH_OK = 0,
    TDVH_ERR_INVALID_ARG = -1,
    TDVH_ERR_UNSUPPORTED = -2,
    TDVH_ERR_DEVICE = -3,      /* a HIP runtime call failed */
    TDVH_ERR_OUT_OF_MEMORY = -4,
    TDVH_ERR_BUSY = -5         /* admission refused: budget exhausted */
} tdvh_status;

/* Message of the last failed call on this thread ("" if none).
 * Replaces the CUException text of cuerr.hpp:73-78. */
const char *tdvh_last_error(void);

/* ---- context ---------------------------------------------------------- */
/* Per-device context owning scratch memory (the LocalDSIMem / TmpCudaMem
 * role of dsimem.cpp:22-42 and tmpcudamem.cpp) and cached path tables. A
 * context is used by one host thread at a time; several contexts may run
 * concurrently. */
typedef struct tdvh_ctx tdvh_ctx;

int tdvh_ctx_create(int device, tdvh_ctx **out);
int tdvh_ctx_destroy(tdvh_ctx *ctx);
/* Releases cached scratch (SemiGlobalDev::finished, semiglobaldev.cpp:30-35) */
int tdvh_ctx_trim(tdvh_ctx *ctx);
/* Bytes of device scratch currently held by the context. */
size_t tdvh_ctx_scratch_bytes(const tdvh_ctx *ctx);

/* ---- stage timing (CudaBenchmarker, benchmark.cpp:7-31) ---------------- */
/* While enabled, every kernel this context launches is bracketed by a pair of
 * hipEvents recorded on the launch stream; durations accumulate per kernel
 * name. Reading synchronises on the recorded events. */
int tdvh_profile_enable(tdvh_ctx *ctx, int on);
int tdvh_profile_reset(tdvh_ctx *ctx);
/* Entry `index` of the per-kernel table: name, total milliseconds and launch
 * count. Returns TDVH_ERR_INVALID_ARG past the last entry. */
int tdvh_profile_read(tdvh_ctx *ctx, int index, char *name, size_t name_len,
                      double *total_ms, int64_t *launches);

/* ---- disparity-space image (DSIMem, dsimem.hpp:12-136) ---------------- */
typedef enum {
    TDVH_U16 = 1, /* BT2 costs, SGM volumes; invalid = 0xFFFF */
    TDVH_U32 = 2, /* SSD costs; invalid = 0xFFFFFFFF */
    TDVH_F32 = 3, /* NCC costs 1 - rho; invalid = +inf */
    TDVH_I32 = 4  /* reference-geometry SGM volume; invalid = 0x3FFFFFFF */
} tdvh_dtype;

typedef enum {
    /* d contiguous; element (x, y, d) at x*stride_x + y*stride_y + d.
     * The reference layout [x][y][d] (dsimem.cpp:34-35, dsimemutil.h:6) is
     * stride_x = H*D, stride_y = D; row-major is stride_x = D,
     * stride_y = W*D. */
    TDVH_LAYOUT_STRIDED = 0,
    /* Device-native uint16 layout read and written by the SGM kernels:
     * pixels row-major; each pixel is Dp = 32*ceil(D/32) uint16 stored as
     * 16*m uint32 words (m = Dp/32); word j*m + r holds d = j*m + r in its
     * low half and d = Dp/2 + j*m + r in its high half; d >= D is invalid.
     * stride_x/stride_y are ignored. */
    TDVH_LAYOUT_SPLIT16 = 1
} tdvh_layout;

typedef struct {
    void *ptr;
    int32_t width, height, depth;
    int32_t dtype;  /* tdvh_dtype */
    int32_t layout; /* tdvh_layout */
    int64_t stride_x, stride_y; /* elements; TDVH_LAYOUT_STRIDED only */
} tdvh_dsi;

/* Bytes needed by a DSI of this shape/dtype/layout (strides row-major). */
size_t tdvh_dsi_bytes(int width, int height, int depth, int dtype, int layout);
/* Fills `out` as a row-major strided or split DSI over `ptr`. */
int tdvh_dsi_init(tdvh_dsi *out, void *ptr, int width, int height, int depth,
                  int dtype, int layout);
/* Copies/reorders a DSI between layouts (same dtype, same shape). */
int tdvh_dsi_convert(tdvh_ctx *ctx, const tdvh_dsi *src, tdvh_dsi *dst,
                     void *stream);

/* ---- matching costs (MatchingCost stages) ----------------------------- */
/* Images are uint8, `pitch` bytes per row, W x H, device memory.
 * Cells with x - d < 0 are invalid (birchfieldcost.cu:52). Out-of-image
 * window reads clamp to the edge (texture semantics, birchfieldcost.cu:160). */

/* Birchfield-Tomasi in 2x units (0..510), uint16 output, STRIDED or SPLIT16.
 * Replaces BirchfieldCostRun (birchfieldcost.cu:194-212). */
int tdvh_cost_bt(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right,
                 int width, int height, int pitch, tdvh_dsi *out, void *stream);

/* (2*radius+1)^2 sum of squared differences, uint32 STRIDED output.
 * Replaces SSDDevRun (ssd.cu:79-114; radius 3 there). */
int tdvh_cost_ssd(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right,
                  int width, int height, int pitch, int radius, tdvh_dsi *out,
                  void *stream);

/* Normalised cross-correlation cost 1 - rho in FP32, STRIDED output;
 * zero_mean = 0 is the reference's plain NCC (crosscorrelation.cu:11-28),
 * 1 the zero-mean variant. A zero denominator gives +inf (invalid).
 * Replaces CrossCorrelationDevRun (crosscorrelation.cu:53-76). */
int tdvh_cost_ncc(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right,
                  int width, int height, int pitch, int radius, int zero_mean,
                  tdvh_dsi *out, void *stream);

/* ---- optimisers (Optimizer stages) ------------------------------------ */

typedef enum {
    /* reference-geometry quirks (semiglobalcpu.cpp, semiglobal.cu) */
    TDVH_SGM_REF_GEOMETRY = 1,  /* path table of semiglobal.cu:57-230 */
    TDVH_SGM_START_ASSIGN = 2,  /* semiglobalcpu.cpp:81 */
    TDVH_SGM_MIN_SKIP_LAST = 4  /* semiglobalcpu.cpp:95-97 */
} tdvh_sgm_flags;

typedef struct {
    int32_t p1, p2;   /* constant penalties of integer cost volumes: [0, 1024]
                         for uint16, [0, 2^30] for uint32 */
    int32_t dir_mask; /* bit i: direction i, order (1,0) (-1,0) (0,1) (0,-1)
                         (1,1) (-1,-1) (-1,1) (1,-1); 0xFF = all 8 */
    int32_t flags;    /* tdvh_sgm_flags; 0 = standard geometry */
    float fp1, fp2;   /* penalties of float cost volumes (finite, >= 0); the
                         reference's are 30/255 and 150/255 (semiglobal.cu:254-255) */
} tdvh_sgm_params;

/* Semi-global aggregation + lowest-index WTA. C is uint16 (STRIDED or
 * SPLIT16; BT costs), uint32 (STRIDED; SSD costs) or float (STRIDED; NCC
 * costs). `S_out` (optional, may be NULL) receives the aggregated volume:
 * uint16 in standard geometry, int32 with TDVH_SGM_REF_GEOMETRY (uint16
 * input only), uint32 or float (strided) for uint32 and float input.
 * uint32 arithmetic saturates at 0xFFFFFFFF (the invalid value); float
 * arithmetic is IEEE with the directions summed in increasing index. A step
 * whose predecessor vector is entirely invalid restarts the path (L = C).
 * `disp` receives int16 disparities (row pitch in elements).
 * Replaces SemiGlobalDevRun (semiglobal.cu:474-521). */
int tdvh_sgm(tdvh_ctx *ctx, const tdvh_dsi *C, const tdvh_sgm_params *params,
             tdvh_dsi *S_out, int16_t *disp, int disp_pitch, void *stream);

typedef enum {
    TDVH_WTA_LOWEST = 0, /* lowest d among equal minima (wtacpu.cpp:28) */
    TDVH_WTA_TREE = 1    /* pairwise tree, ties to the right operand
                            (wta.cu:62-78); D must be a power of two */
} tdvh_wta_mode;

/* Winner-take-all over any dtype/layout. For F32 the minimum cost wins.
 * Replaces WTARunDev (wta.cu:90-100). */
int tdvh_wta(tdvh_ctx *ctx, const tdvh_dsi *S, int mode, int16_t *disp,
             int disp_pitch, void *stream);

typedef enum {
    TDVH_DP_START_ZERO = 1,    /* dynamicprogcpu.cpp:67-79 */
    TDVH_DP_REF_BACKTRACK = 2, /* dynamicprogcpu.cpp:81-91 */
    /* uint32 volumes whose costs the caller expects below 2^22 and INVALID
     * (0xFFFFFFFF) exactly on the cells with d > x (true of tdvh_cost_ssd
     * volumes with radius <= 3: at most 49 * 255^2), with p_occ below 2^20:
     * the sums run in 32 bits with per-column renormalisation. The kernel
     * checks that contract on every cost and recomputes each row that breaks
     * it with 64-bit sums, so the results are the same either way. No
     * counterpart in the reference (an execution choice). */
    TDVH_DP_COST_BOUNDED = 4
} tdvh_dp_flags;

/* Scanline dynamic programming over a STRIDED cost volume with occlusion
 * penalty `p_occ` on +-1 transitions and the tie rule of
 * dynamicprogcpu.cpp:42-56; D <= 512. uint16 (BT) and uint32 (SSD) costs
 * accumulate exactly (64-bit sums, or 32-bit sums with per-column
 * renormalisation for uint16 volumes and under TDVH_DP_COST_BOUNDED, with
 * every row that breaks the bound recomputed in 64 bits); float (NCC) costs
 * accumulate in float with the penalty (float)p_occ.
 * Replaces DynamicProgDevRun (dynamicprog.cu:106-122). */
int tdvh_dp(tdvh_ctx *ctx, const tdvh_dsi *C, uint32_t p_occ, int flags,
            int16_t *disp, int disp_pitch, void *stream);

/* tdvh_dp over a float volume with a float occlusion penalty (finite,
 * >= 0; the reference's DP has none, dynamicprogcpu.cpp:29-59, i.e. 0).
 * Replaces DynamicProgDevRun on CrossCorrelationDev volumes. */
int tdvh_dp_f32(tdvh_ctx *ctx, const tdvh_dsi *C, float p_occ, int flags,
                int16_t *disp, int disp_pitch, void *stream);

/* ---- fused matchers --------------------------------------------------- */

/* SSD window + lowest-index WTA without a cost volume.
 * Replaces FastWTADevRun (fastwta.cu:57-83). */
int tdvh_fast_wta_ssd(tdvh_ctx *ctx, const uint8_t *left,
                      const uint8_t *right, int width, int height, int pitch,
                      int depth, int radius, int16_t *disp, int disp_pitch,
                      void *stream);

/* SSD window (radius 1..3, as tdvh_cost_ssd) + scanline DP (as tdvh_dp with
 * occlusion penalty `p_occ` < 2^20 and the TDVH_DP_START_ZERO /
 * TDVH_DP_REF_BACKTRACK flags) without a cost volume: each row's costs are
 * computed by the wave that scans the row. The disparities of tdvh_cost_ssd
 * followed by tdvh_dp. D <= 256, width <= 10236 (the staged rows, 16 bytes
 * per column, within a workgroup's 160 KiB of LDS). No counterpart in the
 * reference, whose SSDDev + DynamicProgDev chain computes the same map
 * through a volume (ssd.cu:79-114, dynamicprog.cu:106-122). */
int tdvh_stereo_ssd_dp(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right,
                       int width, int height, int pitch, int depth, int radius,
                       uint32_t p_occ, int flags, int16_t *disp, int disp_pitch,
                       void *stream);

/* NCC window (zero-mean or plain, as tdvh_cost_ncc) + lowest-index WTA
 * without a cost volume: the disparities of tdvh_cost_ncc followed by
 * tdvh_wta (TDVH_WTA_LOWEST). No counterpart in the reference (it fuses only
 * SSD, fastwta.cu); the CrossCorrelationDev + WTADev chain computes the same
 * map through a volume (crosscorrelation.cu:53-76, wta.cu:90-100). */
int tdvh_fast_wta_ncc(tdvh_ctx *ctx, const uint8_t *left,
                      const uint8_t *right, int width, int height, int pitch,
                      int depth, int radius, int zero_mean, int16_t *disp,
                      int disp_pitch, void *stream);

/* tdvh_fast_wta_ncc without any device allocation and without a context:
 * the box sums live in a 64 MiB array of device memory that each GPU holds
 * from the library's load until the process ends, so images up to 4,194,304
 * pixels (W*H*16 bytes) are accepted; larger ones return
 * TDVH_ERR_UNSUPPORTED. Runs on the calling thread's current device and
 * `stream`. Calls on one GPU share the array: calls that overlap on one GPU
 * (other streams, other threads) race and must be serialised by the caller.
 * No counterpart in the reference; for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
int tdvh_fast_wta_ncc_noalloc(const uint8_t *left, const uint8_t *right, int width,
                              int height, int pitch, int depth, int radius,
                              int zero_mean, int16_t *disp, int disp_pitch,
                              void *stream);

typedef enum {
    TDVH_FUSE_STAGED = 0, /* BT volume in HBM, then SGM, then WTA */
    TDVH_FUSE_COST = 1    /* BT recomputed inside the SGM scan; no C volume */
} tdvh_fusion;

/* BT2 + 8-direction SGM + WTA for one pair, from images to disparities.
 * Replaces SemiGlobalLabDevRun (semiglobal_lab.cu:232-282). */
int tdvh_stereo_bt_sgm(tdvh_ctx *ctx, const uint8_t *left,
                       const uint8_t *right, int width, int height, int pitch,
                       int depth, const tdvh_sgm_params *params, int fusion,
                       int16_t *disp, int disp_pitch, void *stream);

/* int16 disparity -> float d/D (the value convention of wtacpu.cpp:35-36 and
 * wta.cu:141 that downstream stages such as RGBConv expect). */
int tdvh_disp_to_unit_float(tdvh_ctx *ctx, const int16_t *disp, int disp_pitch,
                            int width, int height, int depth, float *out,
                            int out_pitch, void *stream);

/* uint8 gray from float [0,1] (round(v*255), clamped): the integer profile's
 * view of the reference's FloatImage inputs (misc.cpp:18-49). */
int tdvh_unit_float_to_u8(tdvh_ctx *ctx, const float *src, int src_pitch,
                          int width, int height, uint8_t *dst, int dst_pitch,
                          void *stream);

/* ---- reference path table (SGPaths::getDescCPU, semiglobal.cu:57-81) -- */
typedef struct {
    uint16_t sx, sy, ex, ey;
    int16_t dx, dy;
    uint16_t size;
    uint16_t pad;
} tdvh_sgpath;

/* Host function. Returns the path count W + H + 2(W + H - 1); fills `paths`
 * when not NULL. */
size_t tdvh_sgm_paths(int width, int height, tdvh_sgpath *paths);

/* ---- batched execution ------------------------------------------------- */
/* Many pairs of uneven size and disparity range planned once into device
 * work-item lists (BT rows, line groups of one SGM direction ordered longest
 * first, runs of pixels for the plane sums) and run as one captured
 * hipGraph whose depth classes are parallel branches, each a persistent
 * launch fed by a device work queue or one launch per phase
 * (tdvh_batch_config). There is no counterpart in the reference (it runs
 * one launch per pair, semiglobal.cu:507-516); every pair gets the
 * arithmetic of tdvh_stereo_bt_sgm. */
typedef struct {
    const uint8_t *left, *right; /* device */
    int32_t width, height, pitch, depth;
    int16_t *disp;               /* device, row pitch = width */
} tdvh_pair;

/* ---- around the path: filters, formats, transfers, rectification, 3D ------- */

/* Gaussian 5x5 prefilter of a uint8 image: binomial kernel [1 4 6 4 1]^2/256,
 * edge-replicated border, rounded half up. Replaces MedianFilterCPU::updateImpl
 * (medianfiltercpu.cpp:7-19, cvSmooth CV_GAUSSIAN 5), the "median filter" stage
 * DevStereoMatcher puts in front of the matching cost
 * (devstereomatcher.cpp:22-29). dst must not alias src. */
int tdvh_prefilter_gauss5(tdvh_ctx *ctx, const uint8_t *src, int src_pitch, int width,
                          int height, uint8_t *dst, int dst_pitch, void *stream);

/* Opening of int16 disparities: `iterations` 3x3 erosions, then as many 3x3
 * dilations (a (2k+1)^2 square each way; outside pixels ignored). Replaces
 * Dilate::updateImpl (dilate.cpp:12-22, cvErode/cvDilate with 3 iterations),
 * which follows the matcher in Reconstruction (reconstruction.cpp:30-31).
 * May run in place. */
int tdvh_postfilter_open(tdvh_ctx *ctx, const int16_t *src, int src_pitch, int width,
                         int height, int iterations, int16_t *dst, int dst_pitch, void *stream);

enum { TDVH_RGB = 0, TDVH_BGR = 1 };

/* 3- or 4-channel uint8 colour to gray, (4899 R + 9617 G + 1868 B + 8192) >> 14.
 * Replaces misc::create8UGray (misc.cpp:10-16, cvCvtColor CV_RGB2GRAY). */
int tdvh_rgb_to_gray(tdvh_ctx *ctx, const uint8_t *src, int src_pitch, int width, int height,
                     int channels, int order, uint8_t *dst, int dst_pitch, void *stream);

/* Rectification remap of a uint8 image: output pixel (x, y) samples src at
 * (map_x, map_y)[y][x], bilinear with coordinates rounded to 1/32 pixel and
 * integer weights summing to 1024 ((sum + 512) >> 10); neighbours outside the
 * source read `border`. Replaces cvRemap in RectificationCV::update
 * (rectificationcv.cpp:292-293). Maps are float, pitch in elements. */
int tdvh_remap_u8(tdvh_ctx *ctx, const uint8_t *src, int src_pitch, int width, int height,
                  const float *map_x, const float *map_y, int map_pitch, int out_width,
                  int out_height, int border, uint8_t *dst, int dst_pitch, void *stream);

/* Disparity to 3D: for pixel (x, y), [X Y Z w] = Q [x, H - y, disp * disp_scale, 1]
 * (Q row-major 4x4, evaluated in double), xyz = (X, Y, Z) / w (0 where
 * |w| <= FLT_EPSILON), Z *= z_scale. Output float triples, pitch in floats.
 * Replaces CVReprojector::reproject (cvreprojector.cpp:11-26; the reference
 * passes the d/D float disparity, i.e. disp_scale = 1/D, and z_scale = 24). */
int tdvh_reproject(tdvh_ctx *ctx, const int16_t *disp, int disp_pitch, int width, int height,
                   const float q[16], float disp_scale, float z_scale, float *xyz,
                   int xyz_pitch, void *stream);

/* Page-locked host memory for asynchronous transfers (hipHostMalloc). */
int tdvh_host_alloc(size_t bytes, void **out);
int tdvh_host_free(void *p);

/* One asynchronous 2D copy (host<->device or device<->device, pitches in
 * bytes). Replaces the per-row cudaMemcpy loops of FloatImageImpl::devMem and
 * ::cpuMem (floatimage.cpp:84-91, 118-124). */
int tdvh_copy_2d(tdvh_ctx *ctx, void *dst, size_t dst_pitch, const void *src,
                 size_t src_pitch, size_t width_bytes, int rows, void *stream);

/* Copies a densely packed cost volume to host memory in one transfer, in its
 * own layout. Replaces LocalDSIMem::toCpuMem (dsimem.cpp:55-92), which issues
 * one cudaMemcpy per (x, y). */
int tdvh_dsi_download(tdvh_ctx *ctx, const tdvh_dsi *dsi, void *host, void *stream);

/* ---- stream/graph executor for a fixed frame shape ----------------------------- */

/* A device-resident stereo pipeline for one frame shape, replayed as one
 * hipGraph per frame: [Gaussian prefilter of both inputs] -> fused BT + SGM-8
 * + WTA -> [erode/dilate postfilter, 3 iterations]. Replaces the reference's
 * execution of DevStereoMatcher's WorkUnits on one CUDAProcess thread with a
 * blocking kernel sequence per stage (cudaprocess.cpp:9-30,
 * devstereomatcher.cpp:14-49) and the Dilate stage of Reconstruction
 * (reconstruction.cpp:30-31). tdvh_pipeline_run enqueues one frame
 * asynchronously on `stream`: when the images and the disparity buffer are
 * device memory the graph reads and writes them in place (a graph is captured
 * for each new set of buffers and the last set is kept, so a serving loop over
 * fixed buffers replays without copies); host buffers are copied into the
 * pipeline's own device buffers and back around a second graph. The first run
 * executes eagerly. The pipeline owns a private context; runs on one pipeline
 * must not overlap (they share its buffers). */
typedef struct tdvh_pipeline tdvh_pipeline;
enum { TDVH_PIPE_PREFILTER = 1, TDVH_PIPE_POSTFILTER = 2 };
int tdvh_pipeline_create(tdvh_ctx *ctx, int width, int height, int depth,
                         const tdvh_sgm_params *params, int flags, tdvh_pipeline **out);
int tdvh_pipeline_run(tdvh_pipeline *p, const uint8_t *left, const uint8_t *right, int pitch,
                      int16_t *disp, int disp_pitch, void *stream);
int tdvh_pipeline_destroy(tdvh_pipeline *p);

typedef struct tdvh_batch tdvh_batch;

/* Plans a batch for `n` pairs (item lists, scratch). The pair descriptors
 * are copied; their device buffers must stay valid while the batch runs. */
int tdvh_batch_create(tdvh_ctx *ctx, const tdvh_pair *pairs, int n,
                      const tdvh_sgm_params *params, tdvh_batch **out);
int tdvh_batch_destroy(tdvh_batch *b);
/* Execution knobs (autotuner), 0 keeps the current value:
 *  lines_per_wave (tile height of the SGM line scans): 4 (16 lanes per
 *    line) or 8 (8 lanes per line, used for pairs with D <= 64);
 *  waves_per_group: 1, 2 or 4 waves per workgroup;
 *  prefetch: 4 or 8 steps of the line scans' load ring (8 means the
 *    deepest ring that keeps two waves per SIMD);
 *  fusion_depth: 1 = one launch per phase (BT, line scans, plane sums);
 *    2 = BT, then per depth class one persistent launch that takes line
 *    groups and plane sums from a device work queue, each pair's sums
 *    released by a completion counter once its 8 directions are done;
 *    3 = per class one persistent launch for all three phases;
 *  sum_lag (fusion depth 2 and 3): a pair's plane sums are queued after the
 *    line groups of the next sum_lag pairs of its class, so that they
 *    overlap later pairs' scans; -1 queues them after all line groups.
 * Defaults: 4, 1, 8, 1, -1. */
int tdvh_batch_config(tdvh_batch *b, int lines_per_wave, int waves_per_group,
                      int prefetch, int fusion_depth, int sum_lag);
/* Enqueues the whole batch on `stream`. With use_graph != 0 the launch
 * sequence is captured once into a hipGraph and replayed. */
int tdvh_batch_run(tdvh_batch *b, int use_graph, void *stream);
/* Synchronises `stream` (the batch's results are then complete); reports a
 * work queue that stalled. */
int tdvh_batch_wait(tdvh_batch *b, void *stream);
/* Total W*H*D cells of the batch. */
int64_t tdvh_batch_cells(const tdvh_batch *b);
/* Device bytes held by the batch. */
size_t tdvh_batch_scratch_bytes(const tdvh_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* TDVH_H */
