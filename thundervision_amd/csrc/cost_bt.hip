// Birchfield-Tomasi matching cost (BT2, uint16) and DSI layout conversion.
//
// BT2 restates birchfieldcost.cu:14-59 in 2x units: for the left pixel x and
// the right pixel x - d, with v2 = 2I(x), and min/max over {2I(x),
// I(x)+I(x-1), I(x)+I(x+1)} (edges clamped),
//   C = min(max(0, l2 - rmax, rmin - l2), max(0, r2 - lmax, lmin - r2)),
// invalid (0xFFFF) when x - d < 0 (birchfieldcost.cu:52).
//
// Data path: one tiny pass packs each pixel's three BT terms into a uint32
// (2 x W*H words, L2-resident); the cost pass then needs one term word of the
// left pixel (shared by the 16*m threads of that pixel) and two gathered term
// words of the right image per output word, and writes the volume with fully
// coalesced 32-bit stores in the SPLIT16 layout the SGM kernels read.
#include "kernels.hpp"

namespace tdvh {

__global__ __launch_bounds__(256) void bt_terms_kernel(const uint8_t *__restrict__ L,
                                                       const uint8_t *__restrict__ R, int W, int H,
                                                       int pitch, uint32_t *__restrict__ TL,
                                                       uint32_t *__restrict__ TR)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W) return;
    const uint8_t *lr = L + (size_t)y * pitch;
    const uint8_t *rr = R + (size_t)y * pitch;
    const int xm = x > 0 ? x - 1 : 0, xp = x < W - 1 ? x + 1 : W - 1;
    TL[(size_t)y * W + x] = bt_pack(lr[x], lr[xm], lr[xp]);
    TR[(size_t)y * W + x] = bt_pack(rr[x], rr[xm], rr[xp]);
}

int launch_bt_terms(const uint8_t *L, const uint8_t *R, int W, int H, int pitch, uint32_t *TL,
                    uint32_t *TR, hipStream_t s)
{
    dim3 grid((W + 255) / 256, H);
    prof_begin(s);
    bt_terms_kernel<<<grid, 256, 0, s>>>(L, R, W, H, pitch, TL, TR);
    return check_launch("bt_terms_kernel", s);
}

// Output words of the SPLIT16 layout (two cells each: d_lo, d_hi); a block
// strides over one slice of a row (few large blocks keep dispatch cheap).
template <int M>
__global__ __launch_bounds__(256) void bt_split_kernel(const uint32_t *__restrict__ TL,
                                                       const uint32_t *__restrict__ TR, int W,
                                                       int D, uint32_t *__restrict__ out)
{
    constexpr int WPP = 16 * M; // words per pixel
    const int y = blockIdx.y;
    const int n = W * WPP;
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < n; w += gridDim.x * blockDim.x) {
        const int x = w / WPP, jr = w - x * WPP;
        const int dlo = jr, dhi = 16 * M + jr; // j*M + r == jr
        const size_t p = (size_t)y * W + x;
        const uint32_t lt = TL[p];
        uint32_t clo = INV16, chi = INV16;
        if (dlo < D && dlo <= x) clo = bt_cost(lt, TR[p - dlo]);
        if (dhi < D && dhi <= x) chi = bt_cost(lt, TR[p - dhi]);
        out[p * WPP + jr] = clo | (chi << 16);
    }
}

// A strided uint16 volume, one grid row per band of BT_ROWS image rows: a
// thread computes 4 consecutive cells of a pixel (G = ceil(D / 4) threads
// per pixel) in each row of the band and stores them as one 8-byte word
// where the cells allow it (VEC: D, both strides multiples of 4 cells and
// the base 8-byte aligned).
constexpr int BT_ROWS = 8;
template <bool VEC, bool POW2>
__global__ __launch_bounds__(256) void bt_strided_kernel(const uint32_t *__restrict__ TL,
                                                         const uint32_t *__restrict__ TR, int W,
                                                         int H, int D, int G, int gshift,
                                                         int64_t sx, int64_t sy,
                                                         uint16_t *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * G) return;
    const int x = POW2 ? i >> gshift : i / G, d0 = 4 * (i - x * G);
    const int y1 = min((int)blockIdx.y * BT_ROWS + BT_ROWS, H);
    for (int y = blockIdx.y * BT_ROWS; y < y1; y++) {
    const size_t p = (size_t)y * W + x;
    const uint32_t lt = TL[p];
    const int l2 = lt & 0x3FF, lmin = (lt >> 10) & 0x3FF, lmax = (lt >> 20) & 0x3FF;
    uint32_t c[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int d = d0 + q;
        // the right pixel x - d >= 0 always has terms (no BT_NONE here)
        const uint32_t rt = TR[p - (d <= x ? d : 0)];
        const int r2 = rt & 0x3FF, rmin = (rt >> 10) & 0x3FF, rmax = (rt >> 20) & 0x3FF;
        const int a = max(0, max(l2 - rmax, rmin - l2));
        const int b = max(0, max(r2 - lmax, lmin - r2));
        c[q] = d < D && d <= x ? (uint32_t)min(a, b) : INV16;
    }
    uint16_t *o = out + x * sx + (int64_t)y * sy + d0;
    if constexpr (VEC) {
        *(uint2 *)o = make_uint2(c[0] | (c[1] << 16), c[2] | (c[3] << 16));
    } else {
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (d0 + q < D) o[q] = (uint16_t)c[q];
    }
    }
}

int launch_bt_split(const uint32_t *TL, const uint32_t *TR, int W, int H, int D, uint32_t *out,
                    hipStream_t s)
{
    const int M = split_m(D);
    const int per_row = (W * 16 * M + 255) / 256;
    dim3 grid(per_row < 8 ? per_row : 8, H);
    prof_begin(s);
    switch (M) {
#define CASE(m)                                                                                    \
    case m: bt_split_kernel<m><<<grid, 256, 0, s>>>(TL, TR, W, D, out); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: set_error("depth %d above %d", D, MAX_DEPTH); return TDVH_ERR_UNSUPPORTED;
    }
    return check_launch("bt_split_kernel", s);
}

// ---- DSI layout conversion (uint16 strided <-> SPLIT16) --------------------

// One grid row per image row, a thread per SPLIT16 word.
template <bool TO_SPLIT>
__global__ __launch_bounds__(256) void convert_kernel(uint16_t *strided, uint32_t *split, int W,
                                                      int H, int D, int M, int64_t sx,
                                                      int64_t sy)
{
    const int WPP = 16 * M;
    const int y = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * WPP) return;
    const int x = i / WPP, jr = i - x * WPP;
    const int64_t w = (int64_t)y * W * WPP + i;
    const int dlo = jr, dhi = 16 * M + jr;
    uint16_t *base = strided + x * sx + y * sy;
    if (TO_SPLIT) {
        const uint32_t lo = dlo < D ? base[dlo] : INV16;
        const uint32_t hi = dhi < D ? base[dhi] : INV16;
        split[w] = lo | (hi << 16);
    } else {
        const uint32_t v = split[w];
        if (dlo < D) base[dlo] = (uint16_t)(v & 0xFFFF);
        if (dhi < D) base[dhi] = (uint16_t)(v >> 16);
    }
}

// Four consecutive SPLIT16 words per thread (d = j..j+3 in the low halves,
// 16M + j.. in the high): an 8-byte access per half and a 16-byte one to
// the words, where the volume allows (both strides multiples of 4 cells,
// 8-byte aligned strided base, 16-byte aligned words); cells at or above D
// by element.
template <bool TO_SPLIT>
__global__ __launch_bounds__(256) void convert4_kernel(uint16_t *strided, uint32_t *split, int W,
                                                       int D, int M, int64_t sx, int64_t sy)
{
    const int WPP = 16 * M, Q = 4 * M; // words and quads per pixel
    const int y = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * Q) return;
    const int x = i / Q, j0 = 4 * (i - x * Q);
    uint4 *wp = (uint4 *)(split + ((int64_t)y * W + x) * WPP + j0);
    uint16_t *base = strided + x * sx + y * sy;
    const int dlo = j0, dhi = 16 * M + j0;
    auto get = [&](int d0, uint32_t(&v)[4]) {
        if (d0 + 3 < D) {
            const uint2 u = *(const uint2 *)(base + d0);
            v[0] = u.x & 0xFFFFu, v[1] = u.x >> 16, v[2] = u.y & 0xFFFFu, v[3] = u.y >> 16;
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) v[q] = d0 + q < D ? base[d0 + q] : INV16;
        }
    };
    auto put = [&](int d0, const uint32_t(&v)[4]) {
        if (d0 + 3 < D) {
            *(uint2 *)(base + d0) = make_uint2(v[0] | (v[1] << 16), v[2] | (v[3] << 16));
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (d0 + q < D) base[d0 + q] = (uint16_t)v[q];
        }
    };
    if constexpr (TO_SPLIT) {
        uint32_t lo[4], hi[4];
        get(dlo, lo);
        get(dhi, hi);
        *wp = make_uint4(lo[0] | (hi[0] << 16), lo[1] | (hi[1] << 16), lo[2] | (hi[2] << 16),
                         lo[3] | (hi[3] << 16));
    } else {
        const uint4 w = *wp;
        const uint32_t lo[4] = {w.x & 0xFFFFu, w.y & 0xFFFFu, w.z & 0xFFFFu, w.w & 0xFFFFu};
        const uint32_t hi[4] = {w.x >> 16, w.y >> 16, w.z >> 16, w.w >> 16};
        put(dlo, lo);
        put(dhi, hi);
    }
}

int launch_convert(uint16_t *strided, int64_t sx, int64_t sy, uint32_t *split, int W, int H,
                   int D, bool to_split, hipStream_t s)
{
    if (sx % 4 == 0 && sy % 4 == 0 && ((uintptr_t)strided & 7) == 0 && ((uintptr_t)split & 15) == 0) {
        const int M = split_m(D);
        const dim3 grid((unsigned)((W * 4 * M + 255) / 256), (unsigned)H);
        prof_begin(s);
        if (to_split) convert4_kernel<true><<<grid, 256, 0, s>>>(strided, split, W, D, M, sx, sy);
        else convert4_kernel<false><<<grid, 256, 0, s>>>(strided, split, W, D, M, sx, sy);
        return check_launch("convert_kernel", s);
    }
    const int M = split_m(D);
    const dim3 grid((unsigned)((W * 16 * M + 255) / 256), (unsigned)H);
    prof_begin(s);
    if (to_split)
        convert_kernel<true><<<grid, 256, 0, s>>>(strided, split, W, H, D, M, sx, sy);
    else
        convert_kernel<false><<<grid, 256, 0, s>>>(strided, split, W, H, D, M, sx, sy);
    return check_launch("convert_kernel", s);
}

} // namespace tdvh

using namespace tdvh;

extern "C" int tdvh_cost_bt(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right, int width,
                            int height, int pitch, tdvh_dsi *out, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(left && right && out && out->ptr, "null pointer argument");
    TDVH_REQUIRE(width > 0 && height > 0 && pitch >= width, "bad image %dx%d pitch %d", width,
                 height, pitch);
    TDVH_REQUIRE(out->width == width && out->height == height && out->depth > 0,
                 "DSI shape %dx%dx%d does not match image %dx%d", out->width, out->height,
                 out->depth, width, height);
    TDVH_REQUIRE(out->dtype == TDVH_U16, "BT output must be uint16");
    hipStream_t s = (hipStream_t)stream;
    void *tl = nullptr;
    TDVH_TRY(ctx->scratch("bt_terms", (size_t)width * height * 8, &tl));
    uint32_t *TL = (uint32_t *)tl, *TR = TL + (size_t)width * height;
    TDVH_TRY(launch_bt_terms(left, right, width, height, pitch, TL, TR, s));
    if (out->layout == TDVH_LAYOUT_SPLIT16) {
        TDVH_REQUIRE(out->depth <= MAX_DEPTH, "depth %d above %d", out->depth, MAX_DEPTH);
        return launch_bt_split(TL, TR, width, height, out->depth, (uint32_t *)out->ptr, s);
    }
    TDVH_REQUIRE(out->layout == TDVH_LAYOUT_STRIDED, "unknown layout %d", out->layout);
    const int D = out->depth, G = (D + 3) / 4;
    TDVH_REQUIRE((int64_t)width * G < (1ll << 31), "image width %d too large", width);
    const dim3 grid((unsigned)((width * G + 255) / 256), (unsigned)((height + BT_ROWS - 1) / BT_ROWS));
    const bool vec = D % 4 == 0 && out->stride_x % 4 == 0 && out->stride_y % 4 == 0 &&
                     ((uintptr_t)out->ptr & 7) == 0;
    const bool pow2 = (G & (G - 1)) == 0;
    const int gshift = __builtin_ctz((unsigned)G);
    uint16_t *o = (uint16_t *)out->ptr;
    const int64_t sx = out->stride_x, sy = out->stride_y;
    prof_begin(s);
    if (vec && pow2) bt_strided_kernel<true, true><<<grid, 256, 0, s>>>(TL, TR, width, height, D, G, gshift, sx, sy, o);
    else if (vec) bt_strided_kernel<true, false><<<grid, 256, 0, s>>>(TL, TR, width, height, D, G, gshift, sx, sy, o);
    else if (pow2) bt_strided_kernel<false, true><<<grid, 256, 0, s>>>(TL, TR, width, height, D, G, gshift, sx, sy, o);
    else bt_strided_kernel<false, false><<<grid, 256, 0, s>>>(TL, TR, width, height, D, G, gshift, sx, sy, o);
    return check_launch("bt_strided_kernel", s);
}

extern "C" int tdvh_dsi_convert(tdvh_ctx *ctx, const tdvh_dsi *src, tdvh_dsi *dst, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(src && dst && src->ptr && dst->ptr, "null pointer argument");
    TDVH_REQUIRE(src->width == dst->width && src->height == dst->height &&
                     src->depth == dst->depth && src->dtype == dst->dtype,
                 "shape/dtype mismatch");
    hipStream_t s = (hipStream_t)stream;
    if (src->layout == dst->layout) {
        TDVH_REQUIRE(src->layout == TDVH_LAYOUT_SPLIT16 ||
                         (src->stride_x == dst->stride_x && src->stride_y == dst->stride_y),
                     "strided-to-strided reorder is not supported");
        const size_t bytes = tdvh_dsi_bytes(src->width, src->height, src->depth, src->dtype,
                                            src->layout);
        TDVH_CHECK_HIP(hipMemcpyAsync(dst->ptr, src->ptr, bytes, hipMemcpyDeviceToDevice, s));
        return TDVH_OK;
    }
    TDVH_REQUIRE(src->dtype == TDVH_U16, "SPLIT16 conversion needs uint16");
    TDVH_REQUIRE(src->depth <= MAX_DEPTH, "depth %d above %d", src->depth, MAX_DEPTH);
    if (dst->layout == TDVH_LAYOUT_SPLIT16)
        return launch_convert((uint16_t *)src->ptr, src->stride_x, src->stride_y,
                              (uint32_t *)dst->ptr, src->width, src->height, src->depth, true, s);
    return launch_convert((uint16_t *)dst->ptr, dst->stride_x, dst->stride_y,
                          (uint32_t *)src->ptr, src->width, src->height, src->depth, false, s);
}
