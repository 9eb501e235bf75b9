/*
 * tdvref — CPU oracle for the stereo-matching hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP
 * implementation in thundervision_amd/ is compared against. Only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product path never calls into it and has no CPU fallback.
 *
 * It restates ThunderVision's CPU algorithms in the integer profile used by
 * this project (uint8 images, uint16 BT costs, uint32 SSD costs, FP64 NCC,
 * constant SGM penalties, lowest-index winner-take-all). Each function cites
 * the reference file:line it follows. Nothing here is copied from the
 * reference; the loops are written from the algorithm descriptions.
 *
 * Conventions (all functions):
 *   - images are W x H uint8, row-major, pitch = W bytes;
 *   - cost volumes are row-major [y][x][d] (offset (y*W + x)*D + d);
 *   - a cell is valid iff x - d >= 0 (reference birchfieldcost.cu:52,
 *     ssd.cu:43, crosscorrelation.cu:153, fastwta.cu:233);
 *   - out-of-image reads clamp to the edge (reference texture fetches with
 *     unnormalised coordinates, birchfieldcost.cu:160-163);
 *   - invalid cells hold TDVREF_INV16 / TDVREF_INV32 / +inf.
 */
#ifndef TDVREF_H
#define TDVREF_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDVREF_INV16 0xFFFFu
#define TDVREF_INV32 0xFFFFFFFFu
/* invalid value of the signed reference-geometry SGM volume */
#define TDVREF_INF_I32 0x3FFFFFFF

/* ---- matching costs -------------------------------------------------- */

/* Birchfield-Tomasi, single pixel, in 2x units (0..510), uint16.
 * Follows birchfieldcost.cu:14-59. */
void tdvref_cost_bt2(const uint8_t *L, const uint8_t *R, int W, int H, int D,
                     uint16_t *C);

/* Windowed sum of squared differences, (2r+1)^2 window, uint32.
 * Follows ssd.cu:15-50 (r = 3 there). Brute force. */
void tdvref_cost_ssd(const uint8_t *L, const uint8_t *R, int W, int H, int D,
                     int radius, uint32_t *C);

/* Same values as tdvref_cost_ssd, computed with per-disparity box sums
 * (exact integer arithmetic, O(1) per cell). */
void tdvref_cost_ssd_box(const uint8_t *L, const uint8_t *R, int W, int H,
                         int D, int radius, uint32_t *C);

/* Normalised cross-correlation, FP64: cost = 1 - rho, window (2r+1)^2,
 * zero_mean = 0 gives the reference's plain NCC (crosscorrelation.cu:11-28),
 * zero_mean = 1 the zero-mean variant. Zero denominator -> +inf (invalid).
 * If rho_out is not NULL it receives rho (NaN where invalid). */
void tdvref_cost_ncc(const uint8_t *L, const uint8_t *R, int W, int H, int D,
                     int radius, int zero_mean, double *C, double *rho_out);

/* ---- semi-global matching --------------------------------------------- */

/* The 8 directions, index -> (dx, dy): 0 (1,0) 1 (-1,0) 2 (0,1) 3 (0,-1)
 * 4 (1,1) 5 (-1,-1) 6 (-1,1) 7 (1,-1). */
extern const int tdvref_dir_dx[8];
extern const int tdvref_dir_dy[8];

/* Path cost L_r of one direction in standard geometry (every pixel visited
 * exactly once per direction). Follows the recurrence of
 * semiglobalcpu.cpp:87-125 with constant P1/P2 (no intensity adjustment).
 * Lr: uint16 [y][x][d]. */
void tdvref_sgm_lr(const uint16_t *C, int W, int H, int D, int P1, int P2,
                   int dir, uint16_t *Lr);

/* Aggregated volume S = sum of L_r over the directions in dir_mask
 * (bit i = direction i). Standard geometry. uint16 (S <= 8*(510+P2)). */
void tdvref_sgm_std(const uint16_t *C, int W, int H, int D, int P1, int P2,
                    int dir_mask, uint16_t *S);

/* SGM over uint32 (SSD) costs: the recurrence of tdvref_sgm_lr evaluated as
 *   t = min(L'(d), min(L'(d-1), L'(d+1)) + P1), E = min(t - m, P2), L = C + E
 * with every addition saturating at TDVREF_INV32 (the invalid value), L = C
 * at a path start and after a predecessor with no valid cell, and
 * S = L_first + L_second + ... over the directions of dir_mask in
 * increasing index. */
void tdvref_sgm_u32(const uint32_t *C, int W, int H, int D, uint32_t P1, uint32_t P2,
                    int dir_mask, uint32_t *S);

/* The same over float (NCC) costs in IEEE single precision with that
 * operation order; +inf (and NaN, -inf costs) invalid. */
void tdvref_sgm_f32(const float *C, int W, int H, int D, float P1, float P2, int dir_mask,
                    float *S);

/* Reference-geometry flags. */
#define TDVREF_SGM_START_ASSIGN 1  /* semiglobalcpu.cpp:81 */
#define TDVREF_SGM_MIN_SKIP_LAST 2 /* semiglobalcpu.cpp:95-97 */

typedef struct {
    uint16_t sx, sy, ex, ey; /* start and end pixel */
    int16_t dx, dy;          /* direction start -> end */
    uint16_t size;           /* pixels on the path */
    uint16_t pad;
} tdvref_sgpath;

/* Number of paths of the reference path table: W + H + 2(W + H - 1). */
size_t tdvref_sgm_path_count(int W, int H);

/* Fills the reference path table (semiglobal.cu:57-230), including its
 * end-point arithmetic for W != H. Returns the number of paths. */
size_t tdvref_sgm_paths(int W, int H, tdvref_sgpath *paths);

/* Reference-geometry SGM: the path table traversed sequentially, each path
 * forward from start then backward from end (semiglobalcpu.cpp:196-222).
 * int32 S with TDVREF_INF_I32 for invalid cells. */
void tdvref_sgm_ref(const uint16_t *C, int W, int H, int D, int P1, int P2,
                    int flags, int32_t *S);

/* Per-direction visit counts of the reference path table: visits[dir*W*H +
 * y*W + x] counts how many traversals in direction dir touch (x, y). */
void tdvref_sgm_ref_visits(int W, int H, int32_t *visits);

/* ---- winner-take-all ------------------------------------------------------- */

/* Lowest d minimising the cost (strict <), wtacpu.cpp:14-39. */
void tdvref_wta_u16(const uint16_t *S, int W, int H, int D, int16_t *disp);
void tdvref_wta_u32(const uint32_t *S, int W, int H, int D, int16_t *disp);
void tdvref_wta_i32(const int32_t *S, int W, int H, int D, int16_t *disp);
void tdvref_wta_f64(const double *S, int W, int H, int D, int16_t *disp);
void tdvref_wta_f32(const float *S, int W, int H, int D, int16_t *disp);

/* GPU tree tie-break of wta.cu:62-78 (ties go to the right operand of each
 * pairwise comparison); D must be a power of two. */
void tdvref_wta_tree_u16(const uint16_t *S, int W, int H, int D, int16_t *disp);

/* ---- scanline dynamic programming ----------------------------------------- */

#define TDVREF_DP_START_ZERO 1     /* dynamicprogcpu.cpp:67-79 starts at d=0 */
#define TDVREF_DP_REF_BACKTRACK 2  /* dynamicprogcpu.cpp:81-91 output shift */

/* Per-row DP over uint32 costs with occlusion penalty p_occ on +-1
 * transitions and the tie rule of dynamicprogcpu.cpp:42-56. Accumulates in
 * uint64. If acc_out is not NULL it receives A(W-1, y, d) as uint64
 * [y][d]. */
void tdvref_dp(const uint32_t *C, int W, int H, int D, uint32_t p_occ,
               int flags, int16_t *disp, uint64_t *acc_out);

/* Per-row DP over float costs in IEEE single precision (+inf invalid):
 * c1 = A(x-1, d-1) + p_occ, c3 = A(x-1, d+1) + p_occ, A = C + chosen, with
 * the tie rule, start and backtrack of tdvref_dp. */
void tdvref_dp_f32(const float *C, int W, int H, int D, float p_occ, int flags, int16_t *disp);

/* ---- fused SSD + WTA (fastwta.cu:32-53) ------------------------------------ */
void tdvref_fast_wta_ssd(const uint8_t *L, const uint8_t *R, int W, int H,
                         int D, int radius, int16_t *disp);

/* ---- around the path: filters, formats, rectification, reprojection ---------
 * (tdvref_io.c). These restate the OpenCV calls the reference makes at the
 * matcher boundary in the integer profile; OpenCV itself is not available, so
 * they are pinned by closed forms, known answers and independent restatements
 * (scipy) in the tests, not against OpenCV's outputs. */

/* 5x5 binomial Gaussian ([1 4 6 4 1] x [1 4 6 4 1] / 256), edge-replicated
 * border, rounded half up: cvSmooth(CV_GAUSSIAN, 5) of
 * medianfiltercpu.cpp:7-19 on uint8 images. */
void tdvref_gauss5_u8(const uint8_t *src, int W, int H, uint8_t *dst);

/* `iterations` 3x3 erosions then `iterations` 3x3 dilations (neighbours
 * outside the image ignored): cvErode / cvDilate with 3 iterations of
 * dilate.cpp:12-22, on int16 disparities. */
void tdvref_open_i16(const int16_t *src, int W, int H, int iterations, int16_t *dst);

/* RGB (order 0) or BGR (order 1) uint8 pixels with `channels` 3 or 4 to
 * gray: (4899 R + 9617 G + 1868 B + 8192) >> 14, the 14-bit fixed-point
 * luma of cvCvtColor(CV_RGB2GRAY) on 8-bit data (misc.cpp:10-16). */
void tdvref_rgb_to_gray_u8(const uint8_t *src, int W, int H, int channels, int order,
                           uint8_t *dst);

/* Bilinear remap (cvRemap, rectificationcv.cpp:292-293): output pixel i
 * samples src at (mapx[i], mapy[i]); coordinates are rounded to 1/32 pixel,
 * the four weights (32-ax)(32-ay), ax(32-ay), (32-ax)ay, ax*ay sum to 1024,
 * result (sum + 512) >> 10; neighbours outside the image read `border`. */
void tdvref_remap_u8(const uint8_t *src, int W, int H, const float *mapx, const float *mapy,
                     int Wo, int Ho, int border, uint8_t *dst);

/* Disparity to 3D (cvreprojector.cpp:11-26): pixel (x, y) with disparity
 * value v = disp * disp_scale (float) maps to [X Y Z w] = Q [x, H - y, v, 1]
 * evaluated in double in the order ((q0 x + q1 y') + q2 v) + q3 per row,
 * then xyz = (X, Y, Z) * (1 / w) (all 0 when |w| <= FLT_EPSILON) stored as
 * float, and Z *= z_scale (the reference uses 24); out is float [y][x][3]. */
void tdvref_reproject(const int16_t *disp, int W, int H, const float Q[16], float disp_scale,
                      float z_scale, float *out);

#ifdef __cplusplus
}
#endif

#endif /* TDVREF_H */
