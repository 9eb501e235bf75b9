// Winner-take-all over a strided volume of any dtype (WTADev/WTACPU stages),
// plus the float d/D view and the uint8 conversion helpers.
//
// Lowest-index mode restates wtacpu.cpp:20-33 (strict <, first minimum wins,
// invalid cells never win, d = 0 if every cell is invalid). Tree mode
// restates the pairwise reduction of wta.cu:62-78, where ties go to the right
// operand; its pairing (z, z + i) for i = D/2 ... 1 is the same tournament as
// merging neighbours in bit-reversed order, which is how it is evaluated here.
#include "kernels.hpp"

namespace tdvh {

// Order-preserving uint32 key; invalid / non-finite -> 0xFFFFFFFF.
__device__ __forceinline__ uint32_t wta_key(uint16_t v) { return v == INV16 ? 0xFFFFFFFFu : v; }
__device__ __forceinline__ uint32_t wta_key(uint32_t v) { return v; }
__device__ __forceinline__ uint32_t wta_key(int32_t v)
{
    return v == INF_I32 ? 0xFFFFFFFFu : ((uint32_t)v ^ 0x80000000u);
}
__device__ __forceinline__ uint32_t wta_key(float v)
{
    if (!(v < __builtin_inff())) return 0xFFFFFFFFu;
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// 16 lanes per pixel; lane j scans d = j, j+16, ...
template <typename T>
__global__ __launch_bounds__(256) void wta_lowest_kernel(const T *__restrict__ S, int W, int H,
                                                         int D, int64_t sx, int64_t sy,
                                                         int16_t *__restrict__ disp, int pitch)
{
    const long p = (long)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int j = threadIdx.x & 15;
    if (p >= (long)W * H) return;
    const int y = (int)(p / W), x = (int)(p - (long)y * W);
    const T *base = S + x * sx + y * sy;
    uint32_t bv = 0xFFFFFFFFu, bd = 0xFFFFFFFFu;
    for (int d = j; d < D; d += 16) {
        const uint32_t k = wta_key(base[d]);
        if (k < bv) {
            bv = k;
            bd = (uint32_t)d;
        }
    }
    const uint32_t mv = row_min_u32(bv);
    const uint32_t md = row_min_u32(bv == mv ? bd : 0xFFFFFFFFu);
    if (j == 0) disp[(size_t)y * pitch + x] = (int16_t)(md == 0xFFFFFFFFu ? 0 : md);
}

// The same scan with 4 consecutive cells per lane and load (D % 4 == 0 and
// 4-element aligned pixels): lane j reads d = 4j .. 4j+3, then 64 further
// on, as one streaming vector load each; within a lane d increases, so the
// strict < keeps the lowest index, and the row reduction breaks value ties
// by index.
template <typename T>
__global__ __launch_bounds__(256) void wta_lowest_vec_kernel(const T *__restrict__ S, int W,
                                                             int H, int D, int64_t sx, int64_t sy,
                                                             int16_t *__restrict__ disp,
                                                             int pitch)
{
    typedef T vec4 __attribute__((ext_vector_type(4)));
    const long p = (long)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int j = threadIdx.x & 15;
    if (p >= (long)W * H) return;
    const int y = (int)(p / W), x = (int)(p - (long)y * W);
    const T *base = S + x * sx + y * sy;
    uint32_t bv = 0xFFFFFFFFu, bd = 0xFFFFFFFFu;
#pragma unroll 4
    for (int d0 = 4 * j; d0 < D; d0 += 64) {
        const vec4 v = __builtin_nontemporal_load((const vec4 *)(base + d0));
        const uint32_t k[4] = {wta_key(v.x), wta_key(v.y), wta_key(v.z), wta_key(v.w)};
#pragma unroll
        for (int e = 0; e < 4; e++)
            if (k[e] < bv) {
                bv = k[e];
                bd = (uint32_t)(d0 + e);
            }
    }
    const uint32_t mv = row_min_u32(bv);
    const uint32_t md = row_min_u32(bv == mv ? bd : 0xFFFFFFFFu);
    if (j == 0) disp[(size_t)y * pitch + x] = (int16_t)(md == 0xFFFFFFFFu ? 0 : md);
}

// One thread per pixel: tournament in bit-reversed order with a stack of
// (value, index, subtree size).
template <typename T>
__global__ __launch_bounds__(128) void wta_tree_kernel(const T *__restrict__ S, int W, int H,
                                                       int D, int logD, int64_t sx, int64_t sy,
                                                       int16_t *__restrict__ disp, int pitch)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long)W * H) return;
    const int y = (int)(p / W), x = (int)(p - (long)y * W);
    const T *base = S + x * sx + y * sy;
    uint32_t sv[13], si[13]; // stack depth <= log2(D) + 1, D <= 4096
    int ss[13];
    int top = 0;
    for (int k = 0; k < D; k++) {
        const int d = logD ? (int)(__builtin_bitreverse32((uint32_t)k) >> (32 - logD)) : 0;
        uint32_t v = wta_key(base[d]), i = (uint32_t)d;
        int size = 1;
        while (top > 0 && ss[top - 1] == size) {
            // left operand: the stacked (lower-position) subtree
            const uint32_t lv = sv[top - 1], li = si[top - 1];
            if (lv < v) {
                v = lv;
                i = li;
            }
            size *= 2;
            top--;
        }
        sv[top] = v;
        si[top] = i;
        ss[top] = size;
        top++;
    }
    disp[(size_t)y * pitch + x] = (int16_t)si[0];
}

template <typename T>
static int launch_wta(const tdvh_dsi *S, int mode, int16_t *disp, int pitch, hipStream_t s)
{
    const long pixels = (long)S->width * S->height;
    if (mode == TDVH_WTA_LOWEST) {
        prof_begin(s);
        const bool vec = S->depth % 4 == 0 && S->stride_x % 4 == 0 && S->stride_y % 4 == 0 &&
                         ((uintptr_t)S->ptr & (4 * sizeof(T) - 1)) == 0;
        if (vec) {
            wta_lowest_vec_kernel<T><<<(unsigned)((pixels + 15) / 16), 256, 0, s>>>(
                (const T *)S->ptr, S->width, S->height, S->depth, S->stride_x, S->stride_y, disp,
                pitch);
            return check_launch("wta_lowest_kernel", s);
        }
        wta_lowest_kernel<T><<<(unsigned)((pixels + 15) / 16), 256, 0, s>>>(
            (const T *)S->ptr, S->width, S->height, S->depth, S->stride_x, S->stride_y, disp,
            pitch);
        return check_launch("wta_lowest_kernel", s);
    }
    int logD = 0;
    while ((1 << logD) < S->depth) logD++;
    prof_begin(s);
    wta_tree_kernel<T><<<(unsigned)((pixels + 127) / 128), 128, 0, s>>>(
        (const T *)S->ptr, S->width, S->height, S->depth, logD, S->stride_x, S->stride_y, disp,
        pitch);
    return check_launch("wta_tree_kernel", s);
}

// float(d) / float(D) as one IEEE-rounded division (the value of
// wtacpu.cpp:35-36 and wta.cu:83); a multiplication by a rounded 1/D would
// differ by one ulp for many d when D is not a power of two.
__global__ __launch_bounds__(256) void disp_to_unit_kernel(const int16_t *__restrict__ disp,
                                                           int dpitch, int W, int H, float depth,
                                                           float *__restrict__ out, int opitch)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x < W) out[(size_t)y * opitch + x] = __fdiv_rn((float)disp[(size_t)y * dpitch + x], depth);
}

__global__ __launch_bounds__(256) void unit_to_u8_kernel(const float *__restrict__ src, int spitch,
                                                         int W, int H, uint8_t *__restrict__ dst,
                                                         int dpitch)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    float v = src[(size_t)y * spitch + x] * 255.0f;
    v = fminf(fmaxf(v, 0.0f), 255.0f);
    dst[(size_t)y * dpitch + x] = (uint8_t)__float2int_rn(v);
}

} // namespace tdvh

using namespace tdvh;

extern "C" int tdvh_wta(tdvh_ctx *ctx, const tdvh_dsi *S, int mode, int16_t *disp, int disp_pitch,
                        void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(S && S->ptr && disp, "null pointer argument");
    TDVH_REQUIRE(S->width > 0 && S->height > 0 && S->depth > 0, "bad volume shape");
    TDVH_REQUIRE(disp_pitch >= S->width, "disp_pitch %d < width %d", disp_pitch, S->width);
    TDVH_REQUIRE(mode == TDVH_WTA_LOWEST || mode == TDVH_WTA_TREE, "unknown WTA mode %d", mode);
    if (mode == TDVH_WTA_TREE)
        TDVH_REQUIRE((S->depth & (S->depth - 1)) == 0 && S->depth <= 4096,
                     "tree WTA needs a power-of-two depth (got %d)", S->depth);
    hipStream_t s = (hipStream_t)stream;
    if (S->layout == TDVH_LAYOUT_SPLIT16) {
        // reorder into a row-major scratch volume first
        void *buf = nullptr;
        const size_t bytes = (size_t)S->width * S->height * S->depth * 2;
        TDVH_TRY(ctx->scratch("wta_reorder", bytes, &buf));
        tdvh_dsi st;
        TDVH_TRY(tdvh_dsi_init(&st, buf, S->width, S->height, S->depth, TDVH_U16,
                               TDVH_LAYOUT_STRIDED));
        TDVH_TRY(launch_convert((uint16_t *)buf, st.stride_x, st.stride_y, (uint32_t *)S->ptr,
                                S->width, S->height, S->depth, false, s));
        return launch_wta<uint16_t>(&st, mode, disp, disp_pitch, s);
    }
    TDVH_REQUIRE(S->layout == TDVH_LAYOUT_STRIDED, "unknown layout %d", S->layout);
    switch (S->dtype) {
    case TDVH_U16: return launch_wta<uint16_t>(S, mode, disp, disp_pitch, s);
    case TDVH_U32: return launch_wta<uint32_t>(S, mode, disp, disp_pitch, s);
    case TDVH_I32: return launch_wta<int32_t>(S, mode, disp, disp_pitch, s);
    case TDVH_F32: return launch_wta<float>(S, mode, disp, disp_pitch, s);
    default: set_error("bad dtype %d", S->dtype); return TDVH_ERR_INVALID_ARG;
    }
}

extern "C" int tdvh_disp_to_unit_float(tdvh_ctx *ctx, const int16_t *disp, int disp_pitch,
                                       int width, int height, int depth, float *out,
                                       int out_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(disp && out && width > 0 && height > 0 && depth > 0, "bad arguments");
    TDVH_REQUIRE(disp_pitch >= width && out_pitch >= width, "pitch below width");
    dim3 grid((width + 255) / 256, height);
    hipStream_t s = (hipStream_t)stream;
    prof_begin(s);
    disp_to_unit_kernel<<<grid, 256, 0, s>>>(disp, disp_pitch, width, height, (float)depth, out,
                                             out_pitch);
    return check_launch("disp_to_unit_kernel", s);
}

extern "C" int tdvh_unit_float_to_u8(tdvh_ctx *ctx, const float *src, int src_pitch, int width,
                                     int height, uint8_t *dst, int dst_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(src && dst && width > 0 && height > 0, "bad arguments");
    TDVH_REQUIRE(src_pitch >= width && dst_pitch >= width, "pitch below width");
    dim3 grid((width + 255) / 256, height);
    hipStream_t s = (hipStream_t)stream;
    prof_begin(s);
    unit_to_u8_kernel<<<grid, 256, 0, s>>>(src, src_pitch, width, height, dst,
                                                             dst_pitch);
    return check_launch("unit_to_u8_kernel", s);
}
