/*
 * tdvref_io — oracle of the steps around the hot path (test infrastructure;
 * see tdvref.h). Filters, colour conversion, rectification remap and
 * reprojection, restated in integer (or double) arithmetic with exact
 * definitions; each function cites the reference call it stands for.
 */
#include "tdvref.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* medianfiltercpu.cpp:7-19: cvSmooth(image, out, CV_GAUSSIAN, 5) — a 5x5
 * Gaussian whose sigma derives from the size, i.e. the binomial kernel. */
void tdvref_gauss5_u8(const uint8_t *src, int W, int H, uint8_t *dst)
{
    static const int k[5] = {1, 4, 6, 4, 1};
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int s = 0;
            for (int i = -2; i <= 2; i++) {
                const uint8_t *row = src + (size_t)clampi(y + i, 0, H - 1) * W;
                for (int j = -2; j <= 2; j++)
                    s += k[i + 2] * k[j + 2] * row[clampi(x + j, 0, W - 1)];
            }
            dst[(size_t)y * W + x] = (uint8_t)((s + 128) >> 8);
        }
}

/* One 3x3 erosion (is_max 0) or dilation (1); outside neighbours ignored. */
static void morph3(const int16_t *src, int W, int H, int is_max, int16_t *dst)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int v = src[(size_t)y * W + x];
            for (int i = -1; i <= 1; i++)
                for (int j = -1; j <= 1; j++) {
                    const int yy = y + i, xx = x + j;
                    if (yy < 0 || yy >= H || xx < 0 || xx >= W) continue;
                    const int u = src[(size_t)yy * W + xx];
                    v = is_max ? (u > v ? u : v) : (u < v ? u : v);
                }
            dst[(size_t)y * W + x] = (int16_t)v;
        }
}

/* dilate.cpp:12-22: cvErode(img, tmp, NULL, 3) then cvDilate(tmp, out, NULL, 3). */
void tdvref_open_i16(const int16_t *src, int W, int H, int iterations, int16_t *dst)
{
    const size_t n = (size_t)W * H;
    int16_t *a = (int16_t *)malloc(n * sizeof(int16_t));
    int16_t *b = (int16_t *)malloc(n * sizeof(int16_t));
    for (size_t i = 0; i < n; i++) a[i] = src[i];
    for (int it = 0; it < iterations; it++) {
        morph3(a, W, H, 0, b);
        int16_t *t = a; a = b; b = t;
    }
    for (int it = 0; it < iterations; it++) {
        morph3(a, W, H, 1, b);
        int16_t *t = a; a = b; b = t;
    }
    for (size_t i = 0; i < n; i++) dst[i] = a[i];
    free(a);
    free(b);
}

/* misc.cpp:10-16: cvCvtColor(src, dst, CV_RGB2GRAY) on 8-bit data. */
void tdvref_rgb_to_gray_u8(const uint8_t *src, int W, int H, int channels, int order,
                           uint8_t *dst)
{
    for (size_t i = 0; i < (size_t)W * H; i++) {
        const uint8_t *p = src + i * (size_t)channels;
        const int r = order ? p[2] : p[0], g = p[1], b = order ? p[0] : p[2];
        dst[i] = (uint8_t)((4899 * r + 9617 * g + 1868 * b + 8192) >> 14);
    }
}

/* rectificationcv.cpp:292-293: cvRemap(src, out, mapx, mapy) with bilinear
 * interpolation and outliers filled with a constant. */
void tdvref_remap_u8(const uint8_t *src, int W, int H, const float *mapx, const float *mapy,
                     int Wo, int Ho, int border, uint8_t *dst)
{
    for (int y = 0; y < Ho; y++)
        for (int x = 0; x < Wo; x++) {
            const size_t o = (size_t)y * Wo + x;
            /* 1/32-pixel fixed point, rounded to nearest (ties away from 0) */
            const long fx = lround((double)mapx[o] * 32.0), fy = lround((double)mapy[o] * 32.0);
            const long x0 = fx >> 5, y0 = fy >> 5; /* floor */
            const int ax = (int)(fx & 31), ay = (int)(fy & 31);
            const int w[4] = {(32 - ax) * (32 - ay), ax * (32 - ay), (32 - ax) * ay, ax * ay};
            int s = 0;
            for (int k = 0; k < 4; k++) {
                const long xx = x0 + (k & 1), yy = y0 + (k >> 1);
                const int v = (xx < 0 || yy < 0 || xx >= W || yy >= H)
                                  ? border
                                  : src[(size_t)yy * W + (size_t)xx];
                s += w[k] * v;
            }
            dst[o] = (uint8_t)((s + 512) >> 10);
        }
}

/* cvreprojector.cpp:11-26: xyd = (x, H - y, disp), cvPerspectiveTransform
 * with the 4x4 Q matrix, then z *= 24. */
void tdvref_reproject(const int16_t *disp, int W, int H, const float Q[16], float disp_scale,
                      float z_scale, float *out)
{
    double q[16];
    for (int i = 0; i < 16; i++) q[i] = (double)Q[i];
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t o = (size_t)y * W + x;
            const double vx = (double)(float)x, vy = (double)(float)(H - y);
            const double vd = (double)((float)disp[o] * disp_scale);
            double r[4];
            for (int i = 0; i < 4; i++)
                r[i] = ((q[4 * i] * vx + q[4 * i + 1] * vy) + q[4 * i + 2] * vd) + q[4 * i + 3];
            float *d = out + o * 3;
            if (fabs(r[3]) > (double)FLT_EPSILON) {
                const double iw = 1.0 / r[3];
                d[0] = (float)(r[0] * iw);
                d[1] = (float)(r[1] * iw);
                d[2] = (float)(r[2] * iw) * z_scale;
            } else {
                d[0] = d[1] = d[2] = 0.0f;
            }
        }
}
