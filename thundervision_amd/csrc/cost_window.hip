// Window matching costs: SSD (ssd.cu:15-50), NCC (crosscorrelation.cu:11-49)
// and the fused SSD + lowest-index WTA of fastwta.cu:32-53.
//
// Mapping: a block covers 4 consecutive x (one wave each) of a strip of
// output rows; lane i of a wave holds disparity c*64 + i of chunk c, so the
// 64 lanes write 64 consecutive d of a pixel (coalesced, d-contiguous
// layout). Image rows are staged in LDS one row at a time; each thread keeps
// the per-row window terms of its (x, d) cells in a register ring and slides
// the (2r+1)-row window down the strip, so a cell costs one row term
// (2r+1 products) instead of (2r+1)^2.
//
// Arithmetic is exact integer for SSD (uint32) and for every NCC window sum;
// NCC combines them as
//   zero-mean: rho = (N*Slr - Sl*Sr) / sqrt((N*Sll - Sl^2) (N*Srr - Sr^2))
//   plain:     rho = Slr / sqrt(Sll * Srr)
// with IEEE-rounded FP32 sqrt and division, cost = 1 - rho; a zero
// denominator or x - d < 0 gives +inf.
#include "kernels.hpp"

namespace tdvh {

enum { WIN_SSD = 0, WIN_SSD_WTA = 1, WIN_NCC = 2 };
constexpr int WIN_TY = 32; // output rows per block

struct WinArgs {
    const uint8_t *L, *R;
    int W, H, pitch, D;
    int64_t sx, sy;     // output strides (elements)
    void *out;          // uint32 / float volume, or int16 disparities (WTA)
    int disp_pitch;
    const int32_t *box; // NCC: 4 planes W*H: sum L, sum L^2, sum R, sum R^2
    int zero_mean;
};

template <int RAD, int NCH, int MODE>
__global__ __launch_bounds__(256) void window_kernel(const WinArgs a)
{
    constexpr int NWIN = 2 * RAD + 1;
    constexpr int DP = NCH * 64;
    constexpr int RSPAN = DP + 3 + 2 * RAD; // staged right pixels per row
    __shared__ uint8_t sL[4 + 2 * RAD];
    __shared__ uint8_t sR[RSPAN];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int x0 = blockIdx.x * 4, x = x0 + wv;
    const int y0 = blockIdx.y * WIN_TY;
    const int W = a.W, H = a.H;
    const int rbase = x0 - DP + 1 - RAD; // image x of sR[0]

    uint32_t ring[NCH][NWIN];
    uint32_t sum[NCH];
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        sum[c] = 0;
#pragma unroll
        for (int k = 0; k < NWIN; k++) ring[c][k] = 0;
    }
    const int vend = min(y0 + WIN_TY, H) + RAD; // one past the last row staged
    for (int vb = y0 - RAD; vb < vend; vb += NWIN) {
#pragma unroll
        for (int ph = 0; ph < NWIN; ph++) {
            const int v = vb + ph;
            if (v >= vend) break; // uniform across the block
            const int vy = v < 0 ? 0 : (v >= H ? H - 1 : v);
            const uint8_t *lr = a.L + (size_t)vy * a.pitch;
            const uint8_t *rr = a.R + (size_t)vy * a.pitch;
            __syncthreads(); // previous row's readers are done
            for (int i = threadIdx.x; i < RSPAN; i += 256) {
                const int u = rbase + i;
                sR[i] = rr[u < 0 ? 0 : (u >= W ? W - 1 : u)];
            }
            if (threadIdx.x < 4 + 2 * RAD) {
                const int u = x0 - RAD + (int)threadIdx.x;
                sL[threadIdx.x] = lr[u < 0 ? 0 : (u >= W ? W - 1 : u)];
            }
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const int d = c * 64 + lane;
                uint32_t t = 0;
#pragma unroll
                for (int dx = -RAD; dx <= RAD; dx++) {
                    const int lv = sL[wv + RAD + dx];
                    // right pixel x + dx - d; sR already holds the clamped values
                    const int rv = sR[x + dx - d - rbase];
                    if (MODE == WIN_NCC)
                        t += (uint32_t)(lv * rv);
                    else
                        t += (uint32_t)((lv - rv) * (lv - rv));
                }
                sum[c] += t - ring[c][ph];
                ring[c][ph] = t;
            }
            const int y = v - RAD; // output row whose window is complete
            if (y < y0 || x >= W) continue;
            if (MODE == WIN_SSD_WTA) {
                uint32_t key = 0xFFFFFFFFu;
#pragma unroll
                for (int c = 0; c < NCH; c++) {
                    const int d = c * 64 + lane;
                    if (d < a.D && d <= x) key = min(key, (sum[c] << 9) | (uint32_t)d);
                }
                // wave minimum (the 4 DPP row steps, then across the rows)
                key = row_min_u32(key);
                key = min(key, (uint32_t)__builtin_amdgcn_readlane((int)key, 0));
                key = min(key, (uint32_t)__builtin_amdgcn_readlane((int)key, 16));
                key = min(key, (uint32_t)__builtin_amdgcn_readlane((int)key, 32));
                key = min(key, (uint32_t)__builtin_amdgcn_readlane((int)key, 48));
                if (lane == 0)
                    ((int16_t *)a.out)[(size_t)y * a.disp_pitch + x] =
                        (int16_t)(key == 0xFFFFFFFFu ? 0 : (key & 511u));
            } else {
#pragma unroll
                for (int c = 0; c < NCH; c++) {
                    const int d = c * 64 + lane;
                    if (d >= a.D) continue;
                    const int64_t o = x * a.sx + y * a.sy + d;
                    if (MODE == WIN_SSD) {
                        ((uint32_t *)a.out)[o] = d <= x ? sum[c] : INV32;
                    } else {
                        float cost = __builtin_inff();
                        if (d <= x) {
                            const size_t pl = (size_t)W * H;
                            const size_t pli = (size_t)y * W + x, pri = pli - d;
                            const int32_t sl = a.box[pli], sll = a.box[pl + pli];
                            const int32_t sr = a.box[2 * pl + pri], srr = a.box[3 * pl + pri];
                            int64_t num, vl, vr;
                            if (a.zero_mean) {
                                const int64_t N = (int64_t)NWIN * NWIN;
                                num = N * (int64_t)sum[c] - (int64_t)sl * sr;
                                vl = N * sll - (int64_t)sl * sl;
                                vr = N * srr - (int64_t)sr * sr;
                            } else {
                                num = sum[c];
                                vl = sll;
                                vr = srr;
                            }
                            if (vl != 0 && vr != 0) {
                                const float den = __fmul_rn(__fsqrt_rn((float)vl),
                                                            __fsqrt_rn((float)vr));
                                cost = __fsub_rn(1.0f, __fdiv_rn((float)num, den));
                            }
                        }
                        ((float *)a.out)[o] = cost;
                    }
                }
            }
        }
    }
}

// Per-pixel window sums of v and v^2 of both images (clamped windows).
template <int RAD>
__global__ __launch_bounds__(256) void box_sums_kernel(const uint8_t *__restrict__ L,
                                                       const uint8_t *__restrict__ R, int W,
                                                       int H, int pitch, int32_t *__restrict__ box)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    int32_t sl = 0, sll = 0, sr = 0, srr = 0;
    for (int dy = -RAD; dy <= RAD; dy++) {
        const int yy = min(max(y + dy, 0), H - 1);
        for (int dx = -RAD; dx <= RAD; dx++) {
            const int xx = min(max(x + dx, 0), W - 1);
            const int32_t l = L[(size_t)yy * pitch + xx], r = R[(size_t)yy * pitch + xx];
            sl += l;
            sll += l * l;
            sr += r;
            srr += r * r;
        }
    }
    const size_t pl = (size_t)W * H, o = (size_t)y * W + x;
    box[o] = sl;
    box[pl + o] = sll;
    box[2 * pl + o] = sr;
    box[3 * pl + o] = srr;
}

template <int RAD, int MODE>
static int launch_window_r(const WinArgs &a, hipStream_t s)
{
    dim3 grid((a.W + 3) / 4, (a.H + WIN_TY - 1) / WIN_TY);
    const int nch = (a.D + 63) / 64;
    prof_begin(s);
    switch (nch) {
    case 1: window_kernel<RAD, 1, MODE><<<grid, 256, 0, s>>>(a); break;
    case 2: window_kernel<RAD, 2, MODE><<<grid, 256, 0, s>>>(a); break;
    case 3: window_kernel<RAD, 3, MODE><<<grid, 256, 0, s>>>(a); break;
    case 4: window_kernel<RAD, 4, MODE><<<grid, 256, 0, s>>>(a); break;
    default: set_error("window costs support D <= 256 (got %d)", a.D); return TDVH_ERR_UNSUPPORTED;
    }
    return check_launch(MODE == WIN_NCC ? "ncc_window_kernel"
                                        : (MODE == WIN_SSD ? "ssd_window_kernel"
                                                           : "ssd_wta_window_kernel"),
                        s);
}

template <int MODE>
static int launch_window(const WinArgs &a, int radius, hipStream_t s)
{
    switch (radius) {
    case 1: return launch_window_r<1, MODE>(a, s);
    case 2: return launch_window_r<2, MODE>(a, s);
    case 3: return launch_window_r<3, MODE>(a, s);
    case 4: return launch_window_r<4, MODE>(a, s);
    case 5: return launch_window_r<5, MODE>(a, s);
    default: set_error("window radius %d outside [1, 5]", radius); return TDVH_ERR_UNSUPPORTED;
    }
}

static int check_images(const uint8_t *l, const uint8_t *r, int W, int H, int pitch, int D)
{
    TDVH_REQUIRE(l && r, "null image");
    TDVH_REQUIRE(W > 0 && H > 0 && pitch >= W, "bad image %dx%d pitch %d", W, H, pitch);
    TDVH_REQUIRE(D > 0 && D <= 256, "depth %d outside [1, 256]", D);
    return TDVH_OK;
}

} // namespace tdvh

using namespace tdvh;

extern "C" int tdvh_cost_ssd(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right, int width,
                             int height, int pitch, int radius, tdvh_dsi *out, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(out && out->ptr, "null output");
    TDVH_TRY(check_images(left, right, width, height, pitch, out->depth));
    TDVH_REQUIRE(out->width == width && out->height == height && out->dtype == TDVH_U32 &&
                     out->layout == TDVH_LAYOUT_STRIDED,
                 "SSD output must be a strided uint32 %dx%d volume", width, height);
    WinArgs a{left, right, width, height, pitch, out->depth, out->stride_x, out->stride_y,
              out->ptr, 0, nullptr, 0};
    return launch_window<WIN_SSD>(a, radius, (hipStream_t)stream);
}

extern "C" int tdvh_fast_wta_ssd(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right,
                                 int width, int height, int pitch, int depth, int radius,
                                 int16_t *disp, int disp_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_TRY(check_images(left, right, width, height, pitch, depth));
    TDVH_REQUIRE(disp && disp_pitch >= width, "bad disparity output");
    WinArgs a{left, right, width, height, pitch, depth, 0, 0, disp, disp_pitch, nullptr, 0};
    return launch_window<WIN_SSD_WTA>(a, radius, (hipStream_t)stream);
}

extern "C" int tdvh_cost_ncc(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right, int width,
                             int height, int pitch, int radius, int zero_mean, tdvh_dsi *out,
                             void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(out && out->ptr, "null output");
    TDVH_TRY(check_images(left, right, width, height, pitch, out->depth));
    TDVH_REQUIRE(out->width == width && out->height == height && out->dtype == TDVH_F32 &&
                     out->layout == TDVH_LAYOUT_STRIDED,
                 "NCC output must be a strided float %dx%d volume", width, height);
    TDVH_REQUIRE(radius >= 1 && radius <= 5, "window radius %d outside [1, 5]", radius);
    hipStream_t s = (hipStream_t)stream;
    void *box = nullptr;
    TDVH_TRY(ctx->scratch("ncc_box", (size_t)width * height * 16, &box));
    dim3 grid((width + 255) / 256, height);
    prof_begin(s);
    switch (radius) {
#define CASE(r)                                                                                    \
    case r:                                                                                        \
        box_sums_kernel<r><<<grid, 256, 0, s>>>(left, right, width, height, pitch,                 \
                                                 (int32_t *)box);                                  \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5)
#undef CASE
    }
    TDVH_TRY(check_launch("box_sums_kernel", s));
    WinArgs a{left, right, width, height, pitch, out->depth, out->stride_x, out->stride_y,
              out->ptr, 0, (const int32_t *)box, zero_mean};
    return launch_window<WIN_NCC>(a, radius, s);
}
