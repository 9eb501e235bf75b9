// The steps on either side of the stereo path, device-resident:
//   - Gaussian 5x5 prefilter of the input images (the "median filter" stage
//     of DevStereoMatcher, medianfiltercpu.cpp:7-19),
//   - erode/dilate postfilter of the disparities (dilate.cpp:12-22),
//   - colour to gray conversion (misc.cpp:10-16),
//   - rectification remap (rectificationcv.cpp:292-293),
//   - disparity to 3D reprojection (cvreprojector.cpp:11-26),
//   - pinned host buffers and single-call 2D / volume transfers in place of
//     the reference's per-row and per-(x, y) cudaMemcpy loops
//     (floatimage.cpp:84-91, 118-124; dsimem.cpp:55-92).
// Images are row-major with explicit pitches; all kernels are HBM-bound
// stencils or per-pixel maps, tiled through LDS where a pixel is read by
// several outputs.
#include <cfloat>

#include "kernels.hpp"

namespace tdvh {

constexpr int TILE_W = 64, TILE_H = 32; // output tile of the stencil kernels

// 5x5 binomial Gaussian, edge-replicated border, (S + 128) >> 8.
__global__ __launch_bounds__(256) void gauss5_kernel(const uint8_t *__restrict__ src, int sp,
                                                     int W, int H, uint8_t *__restrict__ dst,
                                                     int dp)
{
    __shared__ uint8_t tin[TILE_H + 4][TILE_W + 4];
    __shared__ uint16_t th[TILE_H + 4][TILE_W];
    const int x0 = blockIdx.x * TILE_W, y0 = blockIdx.y * TILE_H;
    const int t = threadIdx.x;
    for (int i = t; i < (TILE_H + 4) * (TILE_W + 4); i += 256) {
        const int r = i / (TILE_W + 4), c = i % (TILE_W + 4);
        const int y = min(max(y0 + r - 2, 0), H - 1), x = min(max(x0 + c - 2, 0), W - 1);
        tin[r][c] = src[(size_t)y * sp + x];
    }
    __syncthreads();
    const int c = t % TILE_W, r0 = t / TILE_W;
    for (int r = r0; r < TILE_H + 4; r += 4)
        th[r][c] = (uint16_t)(tin[r][c] + 4 * tin[r][c + 1] + 6 * tin[r][c + 2] +
                              4 * tin[r][c + 3] + tin[r][c + 4]);
    __syncthreads();
    const int x = x0 + c;
    if (x >= W) return;
    for (int r = r0; r < TILE_H; r += 4) {
        const int y = y0 + r;
        if (y >= H) break;
        const int s = th[r][c] + 4 * th[r + 1][c] + 6 * th[r + 2][c] + 4 * th[r + 3][c] +
                      th[r + 4][c];
        dst[(size_t)y * dp + x] = (uint8_t)((s + 128) >> 8);
    }
}

// Min (MAX_ = false) or max over the (2R+1)^2 window clipped to the image
// (clamped reads repeat edge pixels, which are inside the clipped window).
template <bool MAX_>
__global__ __launch_bounds__(256) void morph_kernel(const int16_t *__restrict__ src, int sp,
                                                    int W, int H, int R,
                                                    int16_t *__restrict__ dst, int dp)
{
    extern __shared__ int16_t lds[];
    const int TW = TILE_W + 2 * R, TH = TILE_H + 2 * R;
    int16_t *tin = lds;            // TH x TW
    int16_t *th = lds + TH * TW;   // TH x TILE_W
    const int x0 = blockIdx.x * TILE_W, y0 = blockIdx.y * TILE_H;
    const int t = threadIdx.x;
    for (int i = t; i < TH * TW; i += 256) {
        const int r = i / TW, c = i % TW;
        const int y = min(max(y0 + r - R, 0), H - 1), x = min(max(x0 + c - R, 0), W - 1);
        tin[i] = src[(size_t)y * sp + x];
    }
    __syncthreads();
    const int c = t % TILE_W, r0 = t / TILE_W;
    for (int r = r0; r < TH; r += 4) {
        int v = tin[r * TW + c];
        for (int k = 1; k <= 2 * R; k++) {
            const int u = tin[r * TW + c + k];
            v = MAX_ ? max(v, u) : min(v, u);
        }
        th[r * TILE_W + c] = (int16_t)v;
    }
    __syncthreads();
    const int x = x0 + c;
    if (x >= W) return;
    for (int r = r0; r < TILE_H; r += 4) {
        const int y = y0 + r;
        if (y >= H) break;
        int v = th[r * TILE_W + c];
        for (int k = 1; k <= 2 * R; k++) {
            const int u = th[(r + k) * TILE_W + c];
            v = MAX_ ? max(v, u) : min(v, u);
        }
        dst[(size_t)y * dp + x] = (int16_t)v;
    }
}

// 14-bit fixed-point luma of 3- or 4-channel RGB (order 0) or BGR (1).
__global__ __launch_bounds__(256) void gray_kernel(const uint8_t *__restrict__ src, int sp,
                                                   int W, int H, int ch, int order,
                                                   uint8_t *__restrict__ dst, int dp)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    const uint8_t *p = src + (size_t)y * sp + (size_t)x * ch;
    const int r = order ? p[2] : p[0], g = p[1], b = order ? p[0] : p[2];
    dst[(size_t)y * dp + x] = (uint8_t)((4899 * r + 9617 * g + 1868 * b + 8192) >> 14);
}

// Bilinear remap with 1/32-pixel coordinates and integer weights (sum 1024).
__global__ __launch_bounds__(256) void remap_kernel(const uint8_t *__restrict__ src, int sp,
                                                    int W, int H,
                                                    const float *__restrict__ mapx,
                                                    const float *__restrict__ mapy, int mp,
                                                    int Wo, int Ho, int border,
                                                    uint8_t *__restrict__ dst, int dp)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= Wo) return;
    const size_t m = (size_t)y * mp + x;
    const long long fx = llround((double)mapx[m] * 32.0), fy = llround((double)mapy[m] * 32.0);
    const long long xa = fx >> 5, ya = fy >> 5;
    const int ax = (int)(fx & 31), ay = (int)(fy & 31);
    auto at = [&](long long xx, long long yy) -> int {
        return (xx < 0 || yy < 0 || xx >= W || yy >= H) ? border
                                                         : src[(size_t)yy * sp + (size_t)xx];
    };
    const int s = (32 - ax) * (32 - ay) * at(xa, ya) + ax * (32 - ay) * at(xa + 1, ya) +
                  (32 - ax) * ay * at(xa, ya + 1) + ax * ay * at(xa + 1, ya + 1);
    dst[(size_t)y * dp + x] = (uint8_t)((s + 512) >> 10);
}

struct Q16 {
    double q[16];
};

// xyz = (Q [x, H - y, disp * ds, 1]).xyz / w in double with explicit
// rounding (no contraction), z scaled by zs; zeros where |w| <= FLT_EPSILON.
__global__ __launch_bounds__(256) void reproject_kernel(const int16_t *__restrict__ disp, int sp,
                                                        int W, int H, Q16 Q, float ds, float zs,
                                                        float *__restrict__ out, int op)
{
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    const double vx = (double)(float)x, vy = (double)(float)(H - y);
    const double vd = (double)__fmul_rn((float)disp[(size_t)y * sp + x], ds);
    double r[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
        r[i] = __dadd_rn(__dadd_rn(__dadd_rn(__dmul_rn(Q.q[4 * i], vx),
                                             __dmul_rn(Q.q[4 * i + 1], vy)),
                                   __dmul_rn(Q.q[4 * i + 2], vd)),
                         Q.q[4 * i + 3]);
    float *o = out + (size_t)y * op + (size_t)x * 3;
    if (fabs(r[3]) > (double)FLT_EPSILON) {
        const double iw = __ddiv_rn(1.0, r[3]);
        o[0] = __double2float_rn(__dmul_rn(r[0], iw));
        o[1] = __double2float_rn(__dmul_rn(r[1], iw));
        o[2] = __fmul_rn(__double2float_rn(__dmul_rn(r[2], iw)), zs);
    } else {
        o[0] = o[1] = o[2] = 0.0f;
    }
}

static dim3 tiles(int W, int H) { return dim3((W + TILE_W - 1) / TILE_W, (H + TILE_H - 1) / TILE_H); }

} // namespace tdvh

using namespace tdvh;

extern "C" int tdvh_prefilter_gauss5(tdvh_ctx *ctx, const uint8_t *src, int src_pitch, int width,
                                     int height, uint8_t *dst, int dst_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(src && dst, "null pointer argument");
    TDVH_REQUIRE(src != dst, "the prefilter cannot run in place");
    TDVH_REQUIRE(width > 0 && height > 0 && src_pitch >= width && dst_pitch >= width,
                 "bad image %dx%d (pitches %d, %d)", width, height, src_pitch, dst_pitch);
    hipStream_t s = (hipStream_t)stream;
    prof_begin(s);
    gauss5_kernel<<<tiles(width, height), 256, 0, s>>>(src, src_pitch, width, height, dst,
                                                        dst_pitch);
    return check_launch("gauss5_kernel", s);
}

extern "C" int tdvh_postfilter_open(tdvh_ctx *ctx, const int16_t *src, int src_pitch, int width,
                                    int height, int iterations, int16_t *dst, int dst_pitch,
                                    void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(src && dst, "null pointer argument");
    TDVH_REQUIRE(width > 0 && height > 0 && src_pitch >= width && dst_pitch >= width,
                 "bad image %dx%d (pitches %d, %d)", width, height, src_pitch, dst_pitch);
    TDVH_REQUIRE(iterations >= 0 && iterations <= 16, "iterations %d outside [0, 16]",
                 iterations);
    hipStream_t s = (hipStream_t)stream;
    if (iterations == 0) {
        if (src != dst)
            TDVH_CHECK_HIP(hipMemcpy2DAsync(dst, (size_t)dst_pitch * 2, src, (size_t)src_pitch * 2,
                                            (size_t)width * 2, height, hipMemcpyDeviceToDevice,
                                            s));
        return TDVH_OK;
    }
    void *tmp = nullptr;
    TDVH_TRY(ctx->scratch("post_open", (size_t)width * height * 2, &tmp));
    const int R = iterations;
    const size_t lds = ((size_t)(TILE_H + 2 * R) * (TILE_W + 2 * R) +
                        (size_t)(TILE_H + 2 * R) * TILE_W) * sizeof(int16_t);
    prof_begin(s);
    morph_kernel<false><<<tiles(width, height), 256, lds, s>>>(src, src_pitch, width, height, R,
                                                                (int16_t *)tmp, width);
    TDVH_TRY(check_launch("morph_kernel<erode>", s));
    prof_begin(s);
    morph_kernel<true><<<tiles(width, height), 256, lds, s>>>((const int16_t *)tmp, width, width,
                                                               height, R, dst, dst_pitch);
    return check_launch("morph_kernel<dilate>", s);
}

extern "C" int tdvh_rgb_to_gray(tdvh_ctx *ctx, const uint8_t *src, int src_pitch, int width,
                                int height, int channels, int order, uint8_t *dst, int dst_pitch,
                                void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(src && dst, "null pointer argument");
    TDVH_REQUIRE(channels == 3 || channels == 4, "channels must be 3 or 4");
    TDVH_REQUIRE(order == TDVH_RGB || order == TDVH_BGR, "unknown channel order %d", order);
    TDVH_REQUIRE(width > 0 && height > 0 && src_pitch >= width * channels && dst_pitch >= width,
                 "bad image %dx%d (pitches %d, %d)", width, height, src_pitch, dst_pitch);
    hipStream_t s = (hipStream_t)stream;
    prof_begin(s);
    gray_kernel<<<dim3((width + 255) / 256, height), 256, 0, s>>>(src, src_pitch, width, height,
                                                                   channels, order, dst,
                                                                   dst_pitch);
    return check_launch("gray_kernel", s);
}

extern "C" int tdvh_remap_u8(tdvh_ctx *ctx, const uint8_t *src, int src_pitch, int width,
                             int height, const float *map_x, const float *map_y, int map_pitch,
                             int out_width, int out_height, int border, uint8_t *dst,
                             int dst_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(src && dst && map_x && map_y, "null pointer argument");
    TDVH_REQUIRE(width > 0 && height > 0 && src_pitch >= width, "bad source %dx%d pitch %d",
                 width, height, src_pitch);
    TDVH_REQUIRE(out_width > 0 && out_height > 0 && map_pitch >= out_width &&
                     dst_pitch >= out_width,
                 "bad output %dx%d (map pitch %d, pitch %d)", out_width, out_height, map_pitch,
                 dst_pitch);
    TDVH_REQUIRE(border >= 0 && border <= 255, "border %d outside [0, 255]", border);
    hipStream_t s = (hipStream_t)stream;
    prof_begin(s);
    remap_kernel<<<dim3((out_width + 255) / 256, out_height), 256, 0, s>>>(
        src, src_pitch, width, height, map_x, map_y, map_pitch, out_width, out_height, border,
        dst, dst_pitch);
    return check_launch("remap_kernel", s);
}

extern "C" int tdvh_reproject(tdvh_ctx *ctx, const int16_t *disp, int disp_pitch, int width,
                              int height, const float q[16], float disp_scale, float z_scale,
                              float *xyz, int xyz_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(disp && q && xyz, "null pointer argument");
    TDVH_REQUIRE(width > 0 && height > 0 && disp_pitch >= width && xyz_pitch >= 3 * width,
                 "bad image %dx%d (pitches %d, %d)", width, height, disp_pitch, xyz_pitch);
    Q16 Q;
    for (int i = 0; i < 16; i++) Q.q[i] = (double)q[i];
    hipStream_t s = (hipStream_t)stream;
    prof_begin(s);
    reproject_kernel<<<dim3((width + 255) / 256, height), 256, 0, s>>>(
        disp, disp_pitch, width, height, Q, disp_scale, z_scale, xyz, xyz_pitch);
    return check_launch("reproject_kernel", s);
}

extern "C" int tdvh_host_alloc(size_t bytes, void **out)
{
    clear_error();
    TDVH_REQUIRE(out && bytes > 0, "bad arguments");
    *out = nullptr;
    TDVH_CHECK_HIP(hipHostMalloc(out, bytes, hipHostMallocDefault));
    return TDVH_OK;
}

extern "C" int tdvh_host_free(void *p)
{
    clear_error();
    if (p) TDVH_CHECK_HIP(hipHostFree(p));
    return TDVH_OK;
}

extern "C" int tdvh_copy_2d(tdvh_ctx *ctx, void *dst, size_t dst_pitch, const void *src,
                            size_t src_pitch, size_t width_bytes, int rows, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(dst && src, "null pointer argument");
    TDVH_REQUIRE(rows >= 0 && dst_pitch >= width_bytes && src_pitch >= width_bytes,
                 "bad copy (%zu bytes x %d rows, pitches %zu, %zu)", width_bytes, rows,
                 dst_pitch, src_pitch);
    if (rows == 0 || width_bytes == 0) return TDVH_OK;
    TDVH_CHECK_HIP(hipMemcpy2DAsync(dst, dst_pitch, src, src_pitch, width_bytes, rows,
                                    hipMemcpyDefault, (hipStream_t)stream));
    return TDVH_OK;
}

extern "C" int tdvh_dsi_download(tdvh_ctx *ctx, const tdvh_dsi *dsi, void *host, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(dsi && dsi->ptr && host, "null pointer argument");
    tdvh_dsi dense;
    TDVH_TRY(tdvh_dsi_init(&dense, dsi->ptr, dsi->width, dsi->height, dsi->depth, dsi->dtype,
                           dsi->layout));
    TDVH_REQUIRE(dsi->layout == TDVH_LAYOUT_SPLIT16 ||
                     (dsi->stride_x == dense.stride_x && dsi->stride_y == dense.stride_y) ||
                     (dsi->stride_x == (int64_t)dsi->height * dsi->depth &&
                      dsi->stride_y == dsi->depth),
                 "the volume must be densely packed");
    const size_t bytes = tdvh_dsi_bytes(dsi->width, dsi->height, dsi->depth, dsi->dtype,
                                        dsi->layout);
    TDVH_CHECK_HIP(hipMemcpyAsync(host, dsi->ptr, bytes, hipMemcpyDeviceToHost,
                                  (hipStream_t)stream));
    return TDVH_OK;
}
