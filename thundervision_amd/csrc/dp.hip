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
// the d-1 / d+1 neighbours across lanes move with whole-wave DPP shifts (the
// vacated end lane reads the invalid value, which is exactly the missing
// neighbour of d = 0 and of d = 64K-1; cells d >= D carry invalid costs and
// stay invalid, so no disparity bound needs a test). Costs stream through a
// PF-step prefetch ring (one wave per SIMD leaves nothing else to hide the
// load latency). Each step yields the lane's K two-bit decisions as one byte;
// four columns' bytes are stored together as one word per lane (path word
// (x/4)*64 + lane, byte x%4), so that stores, which share the memory counter
// with the loads, take a quarter of its slots. The backtrack stages 64
// columns of decisions in LDS at a time and walks them with a wave-uniform
// disparity.
#include "kernels.hpp"

namespace tdvh {

constexpr uint64_t DP_INF = ~0ull;

// Bytes of one row's decisions: a word per lane for every 4 columns.
__host__ __device__ inline size_t dp_row_bytes(int W) { return (size_t)((W + 3) / 4) * 256; }

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src)
{
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Whole-wave DPP shifts of a 64-bit value: wave_shr:1 (lane i reads lane
// i-1) and wave_shl:1 (lane i reads lane i+1); the lane without a source
// reads DP_INF.
__device__ __forceinline__ uint64_t wave_from_prev(uint64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)(uint32_t)v, 0x138, 0xF, 0xF, false);
    const uint32_t hi =
        (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)(uint32_t)(v >> 32), 0x138, 0xF, 0xF, false);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t wave_from_next(uint64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)(uint32_t)v, 0x130, 0xF, 0xF, false);
    const uint32_t hi =
        (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)(uint32_t)(v >> 32), 0x130, 0xF, 0xF, false);
    return ((uint64_t)hi << 32) | lo;
}

// VEC: K == 4, D == 256 and 16-byte aligned lanes (one 16-byte load per lane
// and step); otherwise K word loads with a disparity bound.
template <int K, int PF, bool VEC>
__global__ __launch_bounds__(64) void dp_kernel(const uint32_t *__restrict__ C, int W, int H,
                                                int D, int64_t sx, int64_t sy, uint32_t P,
                                                int flags, uint8_t *__restrict__ path,
                                                int16_t *__restrict__ disp, int pitch)
{
    static_assert(!VEC || K == 4, "vector loads cover four cells");
    __shared__ uint8_t sp[64][64];
    const int l = threadIdx.x;
    const int y = blockIdx.x;
    const uint32_t *crow = C + y * sy;
    uint32_t *prow = (uint32_t *)(path + (size_t)y * dp_row_bytes(W));

    auto load = [&](uint32_t(&c)[K], int x) {
        const int xx = x < W ? x : W - 1;
        const uint32_t *src = crow + xx * sx + l * K;
        if constexpr (VEC) {
            const uint4 v = *(const uint4 *)src;
            c[0] = v.x, c[1] = v.y, c[2] = v.z, c[3] = v.w;
        } else {
#pragma unroll
            for (int k = 0; k < K; k++) c[k] = l * K + k < D ? src[k] : INV32;
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

    // one step at column x from ring slot u (which it then refills); the
    // decision byte goes into byte x%4 of cw, stored when the word is full
    uint32_t cw = 0;
    auto step = [&](int u, int x, bool tail) {
        uint32_t c[K];
#pragma unroll
        for (int k = 0; k < K; k++) c[k] = ring[u][k];
        load(ring[u], x + PF);
        const uint64_t left = wave_from_prev(A[K - 1]);
        const uint64_t right = wave_from_next(A[0]);
        uint32_t codes = 0;
        uint64_t An[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint64_t lm = k == 0 ? left : A[k - 1];
            const uint64_t lp = k == K - 1 ? right : A[k + 1];
            const uint64_t c1 = lm != DP_INF ? lm + P : DP_INF;
            const uint64_t c2 = A[k];
            const uint64_t c3 = lp != DP_INF ? lp + P : DP_INF;
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
        cw |= codes << (8 * (x & 3));
        if ((x & 3) == 3 || (tail && x == W - 1)) {
            prow[(size_t)(x >> 2) * 64 + l] = cw;
            cw = 0;
        }
    };
    // whole blocks of PF steps without conditions (so the waits stay per
    // slot; PF % 4 == 0 and x0 % 4 == 1 make every x & 3 a constant), then
    // the tail
    static_assert(PF % 4 == 0, "whole decision words per block");
    int x0 = 1;
    for (; x0 + PF <= W; x0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) step(u, x0 + u, false);
    }
    if (x0 == W) { // the last block ended the row inside a word
        if (((W - 1) & 3) != 3) prow[(size_t)((W - 1) >> 2) * 64 + l] = cw;
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (x0 + u < W) step(u, x0 + u, true);

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
            sp[cx - xlo][l] = cx == 0 ? (uint8_t)0x55 // x = 0: "stay"
                                      : (uint8_t)(prow[(size_t)(cx >> 2) * 64 + l] >> (8 * (cx & 3)));
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
    TDVH_TRY(ctx->scratch("dp_path", dp_row_bytes(W) * H, &pb));
    const int K = (D + 63) / 64;
    prof_begin(s);
    const uint32_t *cp = (const uint32_t *)C->ptr;
    const bool vec = D == 256 && C->stride_x % 4 == 0 && C->stride_y % 4 == 0 &&
                     ((uintptr_t)cp & 15) == 0;
    constexpr int PF = 16;
    switch (K) {
#define CASE(k)                                                                                    \
    case k:                                                                                        \
        dp_kernel<k, PF, false><<<H, 64, 0, s>>>(cp, W, H, D, C->stride_x, C->stride_y, p_occ,     \
                                                 flags, (uint8_t *)pb, disp, disp_pitch);          \
        break;
        CASE(1) CASE(2) CASE(3)
#undef CASE
    case 4:
        if (vec)
            dp_kernel<4, PF, true><<<H, 64, 0, s>>>(cp, W, H, D, C->stride_x, C->stride_y, p_occ,
                                                    flags, (uint8_t *)pb, disp, disp_pitch);
        else
            dp_kernel<4, PF, false><<<H, 64, 0, s>>>(cp, W, H, D, C->stride_x, C->stride_y,
                                                     p_occ, flags, (uint8_t *)pb, disp,
                                                     disp_pitch);
        break;
    }
    return check_launch("dp_kernel", s);
}
