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
// Bounded costs (TDVH_DP_COST_BOUNDED: every cost below 2^22, INV32 exactly
// on the cells with d > x, P < 2^20; 7x7 SSD costs are at most 3,186,225)
// accumulate in uint32 instead. Subtracting one constant from a whole column
// changes no comparison, so every 4 columns the wave minimum of 4 columns
// earlier is subtracted (the column minimum never decreases, so nothing goes
// below 0; the reduction is off the dependency chain). A valid A(x, d) then
// exceeds the column minimum it is measured from by at most (D + 7)(Cmax + P):
// d is reachable from that minimum's disparity within max(D - 1, 7) columns
// through valid cells. So valid sums stay below 264 * (2^22 + 2^20) < 1.4e9,
// while invalid ones stay above 2^32 - 1 - 8(Cmax + P) > 4.2e9: comparisons
// between a valid and an invalid operand come out as with exact infinities,
// and saturating adds keep INV32 + x and INF + P at INF. Invalid cells (d > x)
// are recomputed to INF at every column from their INV32 cost, and a valid
// cell's chosen predecessor is valid, so the backtrack never reads a
// decision taken between two invalid operands.
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
#include <type_traits>

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

__device__ __forceinline__ uint32_t wave_from_prev(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ uint32_t wave_from_next(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, 0x130, 0xF, 0xF, false);
}

// a + b, or INF when either is INF (uint64) / saturating at INF (uint32)
__device__ __forceinline__ uint64_t dp_add(uint64_t a, uint64_t b)
{
    return (a == DP_INF || b == DP_INF) ? DP_INF : a + b;
}
__device__ __forceinline__ uint32_t dp_add(uint32_t a, uint32_t b)
{
    return __builtin_elementwise_add_sat(a, b);
}

// VEC: K == 4, D == 256 and 16-byte aligned lanes (one 16-byte load per lane
// and step); otherwise K word loads with a disparity bound. B32: uint32
// accumulation under the bounded-cost contract.
template <int K, int PF, bool VEC, bool B32>
__global__ __launch_bounds__(64) void dp_kernel(const uint32_t *__restrict__ C, int W, int H,
                                                int D, int64_t sx, int64_t sy, uint32_t P,
                                                int flags, uint8_t *__restrict__ path,
                                                int16_t *__restrict__ disp, int pitch)
{
    static_assert(!VEC || K == 4, "vector loads cover four cells");
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

    typedef typename std::conditional<B32, uint32_t, uint64_t>::type acc_t;
    constexpr acc_t INF = (acc_t)DP_INF;
    acc_t A[K];
    {
        uint32_t c[K];
        load(c, 0);
#pragma unroll
        for (int k = 0; k < K; k++) A[k] = c[k] == INV32 ? INF : c[k];
    }
    uint32_t pend = 0; // B32: wave minimum of the last renormalised column
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
        const acc_t left = wave_from_prev(A[K - 1]);
        const acc_t right = wave_from_next(A[0]);
        uint32_t codes = 0;
        acc_t An[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const acc_t lm = k == 0 ? left : A[k - 1];
            const acc_t lp = k == K - 1 ? right : A[k + 1];
            const acc_t c1 = dp_add(lm, (acc_t)P);
            const acc_t c2 = A[k];
            const acc_t c3 = dp_add(lp, (acc_t)P);
            acc_t mv;
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
            if constexpr (B32) An[k] = dp_add(mv, c[k]);
            else An[k] = (c[k] == INV32 || mv == DP_INF) ? DP_INF : mv + c[k];
            codes |= code << (2 * k);
        }
#pragma unroll
        for (int k = 0; k < K; k++) A[k] = An[k];
        if constexpr (B32) {
            if ((x & 3) == 3) {
                // subtract the minimum taken 4 columns ago, then take this
                // column's (consumed at the next word boundary)
#pragma unroll
                for (int k = 0; k < K; k++) A[k] -= pend;
                uint32_t mn = A[0];
#pragma unroll
                for (int k = 1; k < K; k++) mn = min(mn, (uint32_t)A[k]);
                mn = row_min_u32(mn);
                const uint32_t m0 = (uint32_t)__builtin_amdgcn_readlane((int)mn, 0);
                const uint32_t m1 = (uint32_t)__builtin_amdgcn_readlane((int)mn, 16);
                const uint32_t m2 = (uint32_t)__builtin_amdgcn_readlane((int)mn, 32);
                const uint32_t m3 = (uint32_t)__builtin_amdgcn_readlane((int)mn, 48);
                pend = min(min(m0, m1), min(m2, m3));
            }
        }
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
        acc_t bv = INF;
        int bd = 0x7FFFFFFF;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (A[k] < bv) {
                bv = A[k];
                bd = l * K + k;
            }
        for (int off = 32; off > 0; off >>= 1) {
            acc_t ov;
            if constexpr (B32) ov = (uint32_t)__shfl((int)bv, l ^ off, 64);
            else ov = shfl_u64(bv, l ^ off);
            const int od = __shfl_xor(bd, off, 64);
            if (ov < bv || (ov == bv && od < bd)) {
                bv = ov;
                bd = od;
            }
        }
        start = bd == 0x7FFFFFFF ? 0 : bd;
    }

    // backtrack over blocks of 64 columns (16 decision words per lane): each
    // lane loads its block's words with independent loads, then the whole
    // wave walks the columns in lockstep with a wave-uniform disparity z,
    // taking column x's decision byte from lane z / K by readlane; lane j
    // keeps the disparity of the block's column j and stores it once
    const bool refbt = flags & TDVH_DP_REF_BACKTRACK;
    int z = __builtin_amdgcn_readfirstlane(start);
    int16_t *out = disp + (size_t)y * pitch;
    const int nw = (W + 3) / 4; // decision words per lane
    for (int wb = ((W - 1) >> 2) & ~15; wb >= 0; wb -= 16) {
        uint32_t wv[16];
#pragma unroll
        for (int i = 0; i < 16; i++) wv[i] = wb + i < nw ? prow[(size_t)(wb + i) * 64 + l] : 0u;
        int mine = 0;
#pragma unroll
        for (int i = 15; i >= 0; i--) {
#pragma unroll
            for (int b = 3; b >= 0; b--) {
                const int x = (wb + i) * 4 + b;
                if (x > W - 1) continue; // wave-uniform
                const uint32_t word = (uint32_t)__builtin_amdgcn_readlane((int)wv[i], z / K);
                // x = 0 has no stored decision: "stay"
                const int code = x == 0 ? 1 : (int)((word >> (8 * b + 2 * (z % K))) & 3u);
                const int nz = z + code - 1;
                if (refbt) {
                    if (nz >= 0 && nz < D) z = nz;
                    mine = l == 4 * i + b ? z : mine;
                } else {
                    mine = l == 4 * i + b ? z : mine;
                    if (x > 0 && nz >= 0 && nz < D) z = nz;
                }
            }
        }
        if (4 * wb + l < W) out[4 * wb + l] = (int16_t)mine;
    }
}

template <bool B32>
static void launch_dp(const uint32_t *cp, int W, int H, int D, int64_t sx, int64_t sy,
                      uint32_t P, int flags, uint8_t *path, int16_t *disp, int pitch, bool vec,
                      hipStream_t s)
{
    constexpr int PF = 16;
    switch ((D + 63) / 64) {
    case 1: dp_kernel<1, PF, false, B32><<<H, 64, 0, s>>>(cp, W, H, D, sx, sy, P, flags, path, disp, pitch); break;
    case 2: dp_kernel<2, PF, false, B32><<<H, 64, 0, s>>>(cp, W, H, D, sx, sy, P, flags, path, disp, pitch); break;
    case 3: dp_kernel<3, PF, false, B32><<<H, 64, 0, s>>>(cp, W, H, D, sx, sy, P, flags, path, disp, pitch); break;
    default:
        if (vec) dp_kernel<4, PF, true, B32><<<H, 64, 0, s>>>(cp, W, H, D, sx, sy, P, flags, path, disp, pitch);
        else dp_kernel<4, PF, false, B32><<<H, 64, 0, s>>>(cp, W, H, D, sx, sy, P, flags, path, disp, pitch);
        break;
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
    TDVH_REQUIRE((flags & ~(TDVH_DP_START_ZERO | TDVH_DP_REF_BACKTRACK | TDVH_DP_COST_BOUNDED)) ==
                     0,
                 "bad flags 0x%x", flags);
    const bool b32 = flags & TDVH_DP_COST_BOUNDED;
    TDVH_REQUIRE(!b32 || p_occ < (1u << 20), "TDVH_DP_COST_BOUNDED needs p_occ < 2^20 (got %u)",
                 p_occ);
    hipStream_t s = (hipStream_t)stream;
    void *pb = nullptr;
    TDVH_TRY(ctx->scratch("dp_path", dp_row_bytes(W) * H, &pb));
    prof_begin(s);
    const uint32_t *cp = (const uint32_t *)C->ptr;
    const bool vec = D == 256 && C->stride_x % 4 == 0 && C->stride_y % 4 == 0 &&
                     ((uintptr_t)cp & 15) == 0;
    if (b32) launch_dp<true>(cp, W, H, D, C->stride_x, C->stride_y, p_occ, flags, (uint8_t *)pb,
                             disp, disp_pitch, vec, s);
    else launch_dp<false>(cp, W, H, D, C->stride_x, C->stride_y, p_occ, flags, (uint8_t *)pb,
                          disp, disp_pitch, vec, s);
    return check_launch("dp_kernel", s);
}
