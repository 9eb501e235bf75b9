// Scanline dynamic programming (DynamicProgDev / dynamicprogcpu.cpp:13-92).
//
// Per row y, over uint32 costs C(x, y, d) with occlusion penalty P on +-1
// disparity steps:
//   A(0, d) = C(0, d)
//   c1 = A(x-1, d-1) + P, c2 = A(x-1, d), c3 = A(x-1, d+1) + P
//   c1 < c2 && c1 < c3 -> -1 ; else c2 < c3 -> 0 ; else -> +1
//   A(x, d) = C(x, d) + chosen       (invalid cells and sums stay invalid)
// then backtrack from the lowest-index argmin of A(W-1, .).
// Accumulation is uint64, so no row can overflow (1920 * 49 * 255^2 > 2^32).
//
// Mapping: one wave per row, lane l holds d = l*K + k (K <= 4) in registers;
// the d-1 / d+1 neighbours across lanes move with ds_bpermute shuffles. Each
// step stores the lane's K two-bit decisions as one byte (64 bytes per
// column). The backtrack stages 64 columns of decisions in LDS at a time and
// walks them with a wave-uniform disparity.
#include "kernels.hpp"

namespace tdvh {

constexpr uint64_t DP_INF = ~0ull;

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src)
{
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

template <int K, int PF>
__global__ __launch_bounds__(64) void dp_kernel(const uint32_t *__restrict__ C, int W, int H,
                                                int D, int64_t sx, int64_t sy, uint32_t P,
                                                int flags, uint8_t *__restrict__ path,
                                                int16_t *__restrict__ disp, int pitch)
{
    __shared__ uint8_t sp[64][64];
    const int l = threadIdx.x;
    const int y = blockIdx.x;
    const uint32_t *crow = C + y * sy;
    uint8_t *prow = path + (size_t)y * W * 64;

    auto load = [&](uint32_t(&c)[K], int x) {
        const int xx = x < W ? x : W - 1;
        const uint32_t *src = crow + xx * sx;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int d = l * K + k;
            c[k] = d < D ? src[d] : INV32;
        }
    };

    uint64_t A[K];
    {
        uint32_t c[K];
        load(c, 0);
#pragma unroll
        for (int k = 0; k < K; k++) A[k] = c[k] == INV32 ? DP_INF : c[k];
    }
    uint32_t ring[PF][K];
#pragma unroll
    for (int u = 0; u < PF; u++) load(ring[u], 1 + u);

    for (int x0 = 1; x0 < W; x0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int x = x0 + u;
            if (x < W) {
                uint32_t c[K];
#pragma unroll
                for (int k = 0; k < K; k++) c[k] = ring[u][k];
                load(ring[u], x + PF);
                uint64_t left = shfl_u64(A[K - 1], l - 1);
                uint64_t right = shfl_u64(A[0], l + 1);
                if (l == 0) left = DP_INF;
                if (l == 63) right = DP_INF;
                uint32_t codes = 0;
                uint64_t An[K];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int d = l * K + k;
                    const uint64_t lm = k == 0 ? left : A[k - 1];
                    const uint64_t lp = k == K - 1 ? right : A[k + 1];
                    const uint64_t c1 = (d > 0 && lm != DP_INF) ? lm + P : DP_INF;
                    const uint64_t c2 = A[k];
                    const uint64_t c3 = (d < D - 1 && lp != DP_INF) ? lp + P : DP_INF;
                    uint64_t mv;
                    uint32_t code;
                    if (c1 < c2 && c1 < c3) {
                        mv = c1;
                        code = 0;
                    } else if (c2 < c3) {
                        mv = c2;
                        code = 1;
                    } else {
                        mv = c3;
                        code = 2;
                    }
                    An[k] = (c[k] == INV32 || mv == DP_INF) ? DP_INF : mv + c[k];
                    codes |= code << (2 * k);
                }
#pragma unroll
                for (int k = 0; k < K; k++) A[k] = An[k];
                prow[(size_t)x * 64 + l] = (uint8_t)codes;
            }
        }
    }

    // start: lowest-index argmin of the last column (or d = 0)
    int start = 0;
    if (!(flags & TDVH_DP_START_ZERO)) {
        uint64_t bv = DP_INF;
        int bd = 0x7FFFFFFF;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (A[k] < bv) {
                bv = A[k];
                bd = l * K + k;
            }
        for (int off = 32; off > 0; off >>= 1) {
            const uint64_t ov = shfl_u64(bv, l ^ off);
            const int od = __shfl_xor(bd, off, 64);
            if (ov < bv || (ov == bv && od < bd)) {
                bv = ov;
                bd = od;
            }
        }
        start = bd == 0x7FFFFFFF ? 0 : bd;
    }

    // backtrack over chunks of 64 columns staged in LDS
    const bool refbt = flags & TDVH_DP_REF_BACKTRACK;
    int z = start;
    int16_t *out = disp + (size_t)y * pitch;
    for (int xc = W - 1; xc >= 0; xc -= 64) {
        const int xlo = xc - 63 > 0 ? xc - 63 : 0;
        __syncthreads();
        for (int cx = xlo; cx <= xc; cx++)
            sp[cx - xlo][l] = cx == 0 ? (uint8_t)0x55 : prow[(size_t)cx * 64 + l]; // x=0: "stay"
        __syncthreads();
        if (l == 0) {
            for (int x = xc; x >= xlo; x--) {
                const int code = (sp[x - xlo][z / K] >> (2 * (z % K))) & 3;
                const int nz = z + code - 1;
                if (refbt) {
                    if (nz >= 0 && nz < D) z = nz;
                    out[x] = (int16_t)z;
                } else {
                    out[x] = (int16_t)z;
                    if (x > 0 && nz >= 0 && nz < D) z = nz;
                }
            }
        }
    }
}

} // namespace tdvh

using namespace tdvh;

extern "C" int tdvh_dp(tdvh_ctx *ctx, const tdvh_dsi *C, uint32_t p_occ, int flags, int16_t *disp,
                       int disp_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(C && C->ptr && disp, "null pointer argument");
    TDVH_REQUIRE(C->dtype == TDVH_U32 && C->layout == TDVH_LAYOUT_STRIDED,
                 "DP needs a strided uint32 cost volume");
    const int W = C->width, H = C->height, D = C->depth;
    TDVH_REQUIRE(W > 0 && H > 0 && D > 0 && D <= 256, "bad volume %dx%dx%d (D <= 256)", W, H, D);
    TDVH_REQUIRE(disp_pitch >= W, "disp_pitch %d < width %d", disp_pitch, W);
    TDVH_REQUIRE((flags & ~(TDVH_DP_START_ZERO | TDVH_DP_REF_BACKTRACK)) == 0, "bad flags 0x%x",
                 flags);
    hipStream_t s = (hipStream_t)stream;
    void *pb = nullptr;
    TDVH_TRY(ctx->scratch("dp_path", (size_t)W * H * 64, &pb));
    const int K = (D + 63) / 64;
    prof_begin(s);
    switch (K) {
#define CASE(k)                                                                                    \
    case k:                                                                                        \
        dp_kernel<k, 4><<<H, 64, 0, s>>>((const uint32_t *)C->ptr, W, H, D, C->stride_x,           \
                                         C->stride_y, p_occ, flags, (uint8_t *)pb, disp,           \
                                         disp_pitch);                                              \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4)
#undef CASE
    }
    return check_launch("dp_kernel", s);
}
