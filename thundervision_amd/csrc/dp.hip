// Scanline dynamic programming (DynamicProgDev / dynamicprogcpu.cpp:13-92).
//
// Per row y, over costs C(x, y, d) with occlusion penalty P on +-1 disparity
// steps:
//   A(0, d) = C(0, d)
//   c1 = A(x-1, d-1) + P, c2 = A(x-1, d), c3 = A(x-1, d+1) + P
//   c1 < c2 && c1 < c3 -> -1 ; else c2 < c3 -> 0 ; else -> +1
//   A(x, d) = C(x, d) + chosen       (invalid cells and sums stay invalid)
// then backtrack from the lowest-index argmin of A(W-1, .).
//
// Cost volumes: uint16 (BT; 0xFFFF invalid, read as the uint32 invalid),
// uint32 (SSD; 0xFFFFFFFF invalid) and float (NCC; +inf, NaN and -inf
// invalid). Integer costs accumulate exactly, float costs in float with the
// operation order above (one IEEE addition per penalty and per cost).
//
// Integer accumulation is uint64 in general (1920 * 49 * 255^2 > 2^32). The
// bounded form accumulates in uint32 when every valid cost is below 2^22,
// the invalid cells are exactly those with d > x, and P < 2^20 (true of
// 7x7 SSD volumes, at most 3,186,225, and of every BT volume). Subtracting
// one constant from a whole column changes no comparison, so every 4
// columns the wave minimum of 4 columns earlier is subtracted (the column
// minimum never decreases, so nothing goes below 0; the reduction is off the
// dependency chain). A valid A(x, d) then exceeds the column minimum it is
// measured from by at most (D + 7)(Cmax + P): d is reachable from that
// minimum's disparity within max(D - 1, 7) columns through valid cells. So
// valid sums stay below 520 * (2^22 + 2^20) < 2.8e9 for D <= 512, while
// invalid ones stay above 2^32 - 1 - 8(Cmax + P) > 4.25e9: comparisons
// between a valid and an invalid operand come out as with exact infinities,
// and saturating adds keep INV32 + x and INF + P at INF. Invalid cells
// (d > x) are recomputed to INF at every column from their INV32 cost, and a
// valid cell's chosen predecessor is valid, so the backtrack never reads a
// decision taken between two invalid operands.
// The bounded kernel checks that contract on every cost it reads (exactly
// for the first 64K columns, where d > x can hold, then as a running
// maximum) and flags each row that breaks it; a second launch recomputes
// only the flagged rows with uint64 sums, so the result is exact for every
// input.
//
// Mapping: one wave per row, lane l holds d = l*K + k (K <= 8) in registers;
// the d-1 / d+1 neighbours across lanes move with whole-wave DPP shifts (the
// vacated end lane reads the invalid value, which is exactly the missing
// neighbour of d = 0 and of d = 64K-1; cells d >= D carry invalid costs and
// stay invalid, so no disparity bound needs a test). Costs stream through a
// PF-step prefetch ring (one wave per SIMD leaves nothing else to hide the
// load latency). Each step yields the lane's K two-bit decisions as 8 bits
// (K <= 4) or 16 bits (K <= 8); CPW columns' bits are stored together as one
// word per lane (path word (x/CPW)*64 + lane), so that stores, which share
// the memory counter with the loads, take few of its slots. Backtrack: per
// block of 64 columns each lane loads its block's decision words together,
// then the whole wave walks the columns in lockstep with a wave-uniform
// disparity, reading the decision bits of lane z / K by readlane.
#include <type_traits>

#include "kernels.hpp"

namespace tdvh {

constexpr uint64_t DP_INF = ~0ull;
constexpr uint32_t DP_COST_LIMIT = 1u << 22; // bounded form: valid costs below this
constexpr uint32_t DP_PEN_LIMIT = 1u << 20;  // bounded form: penalties below this
constexpr int DP_MAX_DEPTH = 512;           // 8 disparities per lane

enum { DP_U64 = 0, DP_B32 = 1, DP_F32 = 2 };

// Columns of decisions per 32-bit word and lane: 4 (8 bits) for K <= 4,
// 2 (16 bits) for K <= 8.
__host__ __device__ constexpr int dp_cpw(int K) { return K <= 4 ? 4 : 2; }

// Bytes of one row's decisions: a word per lane for every CPW columns.
__host__ __device__ inline size_t dp_row_bytes(int W, int K)
{
    return (size_t)((W + dp_cpw(K) - 1) / dp_cpw(K)) * 256;
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src)
{
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}

// Whole-wave DPP shifts: wave_shr:1 (lane i reads lane i-1) and wave_shl:1
// (lane i reads lane i+1); the lane without a source reads the invalid value.
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
constexpr int F32_INF_BITS = 0x7F800000;
__device__ __forceinline__ float wave_from_prev(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
                                         F32_INF_BITS, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float wave_from_next(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
                                         F32_INF_BITS, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}

// a + b, or INF when either is INF (uint64) / saturating at INF (uint32) /
// IEEE (float, +inf absorbs)
__device__ __forceinline__ uint64_t dp_add(uint64_t a, uint64_t b)
{
    return (a == DP_INF || b == DP_INF) ? DP_INF : a + b;
}
__device__ __forceinline__ uint32_t dp_add(uint32_t a, uint32_t b)
{
    return __builtin_elementwise_add_sat(a, b);
}
__device__ __forceinline__ float dp_add(float a, float b) { return a + b; }

// Cost of a cell as the accumulator's operand: the uint32 invalid for
// integer volumes, +inf for float volumes.
template <typename TIn, int MODE> struct DpCost;
template <int MODE> struct DpCost<uint16_t, MODE> {
    typedef uint32_t type;
    __device__ static uint32_t get(uint16_t v) { return v == INV16 ? INV32 : (uint32_t)v; }
};
template <int MODE> struct DpCost<uint32_t, MODE> {
    typedef uint32_t type;
    __device__ static uint32_t get(uint32_t v) { return v; }
};
template <int MODE> struct DpCost<float, MODE> {
    typedef float type;
    __device__ static float get(float v)
    {
        return (v < __builtin_inff() && v > -__builtin_inff()) ? v : __builtin_inff();
    }
};

// VEC: K == 4, D == 256, 4-byte costs and 16-byte aligned lanes (one 16-byte
// load per lane and step); otherwise K element loads with a disparity bound.
// `only` (fixup launch): rows whose flag is 0 return at once. `flags_out`
// (bounded launch): receives 1 for each row that breaks the contract.
template <typename TIn, int K, int PF, bool VEC, int MODE>
__global__ __launch_bounds__(64) void dp_kernel(const TIn *__restrict__ C, int W, int H, int D,
                                                int64_t sx, int64_t sy, uint32_t P, float Pf,
                                                int flags, uint8_t *__restrict__ path,
                                                int16_t *__restrict__ disp, int pitch,
                                                const uint8_t *__restrict__ only,
                                                uint8_t *__restrict__ flags_out)
{
    static_assert(!VEC || (K == 4 && sizeof(TIn) == 4), "vector loads cover four 4-byte cells");
    constexpr int CPW = dp_cpw(K), BITS = 32 / CPW;
    constexpr bool B32 = MODE == DP_B32;
    typedef DpCost<TIn, MODE> cost_t;
    typedef typename cost_t::type cval_t;
    const int l = threadIdx.x;
    const int y = blockIdx.x;
    if (only && only[y] == 0) return; // wave-uniform
    const TIn *crow = C + y * sy;
    uint32_t *prow = (uint32_t *)(path + (size_t)y * dp_row_bytes(W, K));

    auto load = [&](cval_t(&c)[K], int x) {
        const int xx = x < W ? x : W - 1;
        const TIn *src = crow + xx * sx + l * K;
        if constexpr (VEC) {
            typedef TIn vec4 __attribute__((ext_vector_type(4)));
            const vec4 v = *(const vec4 *)src;
            c[0] = cost_t::get(v.x), c[1] = cost_t::get(v.y), c[2] = cost_t::get(v.z),
            c[3] = cost_t::get(v.w);
        } else {
#pragma unroll
            for (int k = 0; k < K; k++) {
                if constexpr (MODE == DP_F32) c[k] = l * K + k < D ? cost_t::get(src[k]) : __builtin_inff();
                else c[k] = l * K + k < D ? cost_t::get(src[k]) : INV32;
            }
        }
    };

    typedef typename std::conditional<MODE == DP_F32, float,
                                      typename std::conditional<B32, uint32_t, uint64_t>::type>::type
        acc_t;
    acc_t INF, Pen;
    if constexpr (MODE == DP_F32) {
        INF = __builtin_inff();
        Pen = Pf;
    } else {
        INF = (acc_t)DP_INF;
        Pen = (acc_t)P;
    }

    // bounded contract check: a cell d > x (or d >= D) must hold INV32, any
    // other cell a cost below 2^22
    bool bad = false;
    uint32_t vmax = 0; // running maximum once every d < D is valid
    auto check_exact = [&](const cval_t(&c)[K], int x) {
        if constexpr (B32) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int d = l * K + k;
                const bool inv = d > x || d >= D;
                bad |= inv ? (uint32_t)c[k] != INV32 : (uint32_t)c[k] >= DP_COST_LIMIT;
            }
        }
    };
    auto check_max = [&](const cval_t(&c)[K]) {
        if constexpr (B32) {
#pragma unroll
            for (int k = 0; k < K; k++)
                vmax = max(vmax, l * K + k < D ? (uint32_t)c[k] : 0u);
        }
    };

    acc_t A[K];
    {
        cval_t c[K];
        load(c, 0);
        check_exact(c, 0);
#pragma unroll
        for (int k = 0; k < K; k++) {
            if constexpr (MODE == DP_F32) A[k] = c[k];
            else A[k] = c[k] == INV32 ? INF : c[k];
        }
    }
    uint32_t pend = 0; // B32: wave minimum of the last renormalised column
    cval_t ring[PF][K];
#pragma unroll
    for (int u = 0; u < PF; u++) load(ring[u], 1 + u);

    // one step at column x from ring slot u (which it then refills); the
    // decision bits go into slot x%CPW of cw, stored when the word is full
    uint32_t cw = 0;
    auto step = [&](int u, int x, bool tail, bool exact) {
        cval_t c[K];
#pragma unroll
        for (int k = 0; k < K; k++) c[k] = ring[u][k];
        load(ring[u], x + PF);
        if (exact) check_exact(c, x);
        else check_max(c);
        const acc_t left = wave_from_prev(A[K - 1]);
        const acc_t right = wave_from_next(A[0]);
        uint32_t codes = 0;
        acc_t An[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const acc_t lm = k == 0 ? left : A[k - 1];
            const acc_t lp = k == K - 1 ? right : A[k + 1];
            const acc_t c1 = dp_add(lm, Pen);
            const acc_t c2 = A[k];
            const acc_t c3 = dp_add(lp, Pen);
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
            if constexpr (MODE == DP_F32) An[k] = mv + c[k];
            else if constexpr (B32) An[k] = dp_add(mv, c[k]);
            else An[k] = (c[k] == INV32 || mv == DP_INF) ? DP_INF : mv + c[k];
            codes |= code << (2 * k);
        }
#pragma unroll
        for (int k = 0; k < K; k++) A[k] = An[k];
        if constexpr (B32) {
            if ((x & 3) == 3) {
                // subtract the minimum taken 4 columns ago, then take this
                // column's (consumed at the next renormalisation)
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
        cw |= codes << (BITS * (x % CPW));
        if (x % CPW == CPW - 1 || (tail && x == W - 1)) {
            prow[(size_t)(x / CPW) * 64 + l] = cw;
            cw = 0;
        }
    };
    // whole blocks of PF steps without conditions (so the waits stay per
    // slot; PF % 4 == 0 and x0 % 4 == 1 make every x % CPW and x & 3 a
    // constant), then the tail. Blocks that may hold cells with d > x check
    // the bounded contract exactly, the others keep a running maximum.
    static_assert(PF % 4 == 0, "whole decision words and renormalisation periods per block");
    int x0 = 1;
    for (; x0 + PF <= W && (B32 && x0 < 64 * K); x0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) step(u, x0 + u, false, true);
    }
    for (; x0 + PF <= W; x0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) step(u, x0 + u, false, false);
    }
    if (x0 == W) { // the last block ended the row inside a word
        if ((W - 1) % CPW != CPW - 1) prow[(size_t)((W - 1) / CPW) * 64 + l] = cw;
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (x0 + u < W) step(u, x0 + u, true, true);

    if constexpr (B32) {
        // the row is exact unless some cost broke the contract
        const bool row_bad = __builtin_amdgcn_ballot_w64(bad || vmax >= DP_COST_LIMIT) != 0;
        if (l == 0) flags_out[y] = row_bad ? 1 : 0;
        if (row_bad) return; // the fixup launch writes this row
    }

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
            if constexpr (MODE == DP_F32)
                ov = __builtin_bit_cast(float, __shfl(__builtin_bit_cast(int, bv), l ^ off, 64));
            else if constexpr (B32) ov = (uint32_t)__shfl((int)bv, l ^ off, 64);
            else ov = shfl_u64(bv, l ^ off);
            const int od = __shfl_xor(bd, off, 64);
            if (ov < bv || (ov == bv && od < bd)) {
                bv = ov;
                bd = od;
            }
        }
        start = bd == 0x7FFFFFFF ? 0 : bd;
    }

    // backtrack over blocks of 64 columns (64 / CPW decision words per
    // lane): each lane loads its block's words with independent loads, then
    // the whole wave walks the columns in lockstep with a wave-uniform
    // disparity z, taking column x's decision bits from lane z / K by
    // readlane; lane j keeps the disparity of the block's column j and
    // stores it once
    constexpr int WPB = 64 / CPW; // decision words per block and lane
    const bool refbt = flags & TDVH_DP_REF_BACKTRACK;
    int z = __builtin_amdgcn_readfirstlane(start);
    int16_t *out = disp + (size_t)y * pitch;
    const int nw = (W + CPW - 1) / CPW; // decision words per lane
    for (int wb = ((W - 1) / CPW) & ~(WPB - 1); wb >= 0; wb -= WPB) {
        uint32_t wv[WPB];
#pragma unroll
        for (int i = 0; i < WPB; i++) wv[i] = wb + i < nw ? prow[(size_t)(wb + i) * 64 + l] : 0u;
        int mine = 0;
#pragma unroll
        for (int i = WPB - 1; i >= 0; i--) {
#pragma unroll
            for (int b = CPW - 1; b >= 0; b--) {
                const int x = (wb + i) * CPW + b;
                if (x > W - 1) continue; // wave-uniform
                const uint32_t word = (uint32_t)__builtin_amdgcn_readlane((int)wv[i], z / K);
                // x = 0 has no stored decision: "stay"
                const int code = x == 0 ? 1 : (int)((word >> (BITS * b + 2 * (z % K))) & 3u);
                const int nz = z + code - 1;
                if (refbt) {
                    if (nz >= 0 && nz < D) z = nz;
                    mine = l == CPW * i + b ? z : mine;
                } else {
                    mine = l == CPW * i + b ? z : mine;
                    if (x > 0 && nz >= 0 && nz < D) z = nz;
                }
            }
        }
        if (CPW * wb + l < W) out[CPW * wb + l] = (int16_t)mine;
    }
}

struct DpLaunch {
    const void *C;
    int W, H, D;
    int64_t sx, sy;
    uint32_t P;
    float Pf;
    int flags;
    uint8_t *path;
    int16_t *disp;
    int pitch;
    const uint8_t *only;
    uint8_t *flags_out;
};

template <typename TIn, int MODE>
static void launch_dp(const DpLaunch &a, bool vec, hipStream_t s)
{
    const TIn *cp = (const TIn *)a.C;
#define DP_ARGS cp, a.W, a.H, a.D, a.sx, a.sy, a.P, a.Pf, a.flags, a.path, a.disp, a.pitch, a.only, a.flags_out
    switch ((a.D + 63) / 64) {
    case 1: dp_kernel<TIn, 1, 16, false, MODE><<<a.H, 64, 0, s>>>(DP_ARGS); break;
    case 2: dp_kernel<TIn, 2, 16, false, MODE><<<a.H, 64, 0, s>>>(DP_ARGS); break;
    case 3: dp_kernel<TIn, 3, 16, false, MODE><<<a.H, 64, 0, s>>>(DP_ARGS); break;
    case 4:
        if constexpr (sizeof(TIn) == 4) {
            if (vec) {
                dp_kernel<TIn, 4, 16, true, MODE><<<a.H, 64, 0, s>>>(DP_ARGS);
                break;
            }
        }
        dp_kernel<TIn, 4, 16, false, MODE><<<a.H, 64, 0, s>>>(DP_ARGS);
        break;
    case 5: case 6: dp_kernel<TIn, 6, 8, false, MODE><<<a.H, 64, 0, s>>>(DP_ARGS); break;
    default: dp_kernel<TIn, 8, 8, false, MODE><<<a.H, 64, 0, s>>>(DP_ARGS); break;
    }
#undef DP_ARGS
}

// Bounded uint32 launch, then the uint64 recomputation of the rows it
// flagged (a launch whose waves return at once when no row is flagged).
template <typename TIn>
static int run_dp_integer(const DpLaunch &a, bool bounded, bool vec, hipStream_t s)
{
    if (!bounded) {
        prof_begin(s);
        launch_dp<TIn, DP_U64>(a, vec, s);
        return check_launch("dp_kernel", s);
    }
    prof_begin(s);
    launch_dp<TIn, DP_B32>(a, vec, s);
    TDVH_TRY(check_launch("dp_kernel", s));
    DpLaunch f = a;
    f.only = a.flags_out;
    f.flags_out = nullptr;
    prof_begin(s);
    launch_dp<TIn, DP_U64>(f, vec, s);
    return check_launch("dp_fixup_kernel", s);
}

} // namespace tdvh

using namespace tdvh;

// Both entry points: `p_occ` for integer volumes, `pf` for float volumes.
static int dp_run(tdvh_ctx *ctx, const tdvh_dsi *C, uint32_t p_occ, float pf, int flags,
                  int16_t *disp, int disp_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(C && C->ptr && disp, "null pointer argument");
    TDVH_REQUIRE(C->layout == TDVH_LAYOUT_STRIDED ||
                     (C->layout == TDVH_LAYOUT_SPLIT16 && C->dtype == TDVH_U16),
                 "DP needs a strided volume or a uint16 SPLIT16 volume");
    if (C->layout == TDVH_LAYOUT_SPLIT16) {
        // BirchfieldCostDev's volume: reorder into a row-major scratch first
        TDVH_REQUIRE(C->width > 0 && C->height > 0 && C->depth > 0 && C->depth <= MAX_DEPTH,
                     "bad SPLIT16 volume %dx%dx%d", C->width, C->height, C->depth);
        void *buf = nullptr;
        TDVH_TRY(ctx->scratch("dp_reorder", (size_t)C->width * C->height * C->depth * 2, &buf));
        tdvh_dsi st;
        TDVH_TRY(tdvh_dsi_init(&st, buf, C->width, C->height, C->depth, TDVH_U16,
                               TDVH_LAYOUT_STRIDED));
        TDVH_TRY(launch_convert((uint16_t *)buf, st.stride_x, st.stride_y, (uint32_t *)C->ptr,
                                C->width, C->height, C->depth, false, (hipStream_t)stream));
        return dp_run(ctx, &st, p_occ, pf, flags, disp, disp_pitch, stream);
    }
    TDVH_REQUIRE(C->dtype == TDVH_U16 || C->dtype == TDVH_U32 || C->dtype == TDVH_F32,
                 "DP needs a uint16, uint32 or float cost volume (got dtype %d)", C->dtype);
    const int W = C->width, H = C->height, D = C->depth;
    TDVH_REQUIRE(W > 0 && H > 0 && D > 0 && D <= DP_MAX_DEPTH, "bad volume %dx%dx%d (D <= %d)", W,
                 H, D, DP_MAX_DEPTH);
    TDVH_REQUIRE(disp_pitch >= W, "disp_pitch %d < width %d", disp_pitch, W);
    TDVH_REQUIRE((flags & ~(TDVH_DP_START_ZERO | TDVH_DP_REF_BACKTRACK | TDVH_DP_COST_BOUNDED)) ==
                     0,
                 "bad flags 0x%x", flags);
    hipStream_t s = (hipStream_t)stream;
    const int K = (D + 63) / 64 <= 4 ? (D + 63) / 64 : ((D + 63) / 64 <= 6 ? 6 : 8);
    void *pb = nullptr;
    TDVH_TRY(ctx->scratch("dp_path", dp_row_bytes(W, K) * H, &pb));
    DpLaunch a{C->ptr, W, H, D, C->stride_x, C->stride_y, p_occ, pf, flags,
               (uint8_t *)pb, disp, disp_pitch, nullptr, nullptr};
    const size_t esz = C->dtype == TDVH_U16 ? 2 : 4;
    const bool vec = D == 256 && esz == 4 && C->stride_x % 4 == 0 && C->stride_y % 4 == 0 &&
                     ((uintptr_t)C->ptr & 15) == 0;
    if (C->dtype == TDVH_F32) {
        prof_begin(s);
        launch_dp<float, DP_F32>(a, vec, s);
        return check_launch("dp_kernel", s);
    }
    // 32-bit sums for every uint16 volume and for uint32 volumes the caller
    // marks bounded; rows that break the bound are recomputed in 64 bits
    const bool bounded =
        p_occ < DP_PEN_LIMIT && (C->dtype == TDVH_U16 || (flags & TDVH_DP_COST_BOUNDED));
    if (bounded) {
        void *fb = nullptr;
        TDVH_TRY(ctx->scratch("dp_row_flags", (size_t)H, &fb));
        a.flags_out = (uint8_t *)fb;
    }
    if (C->dtype == TDVH_U16) return run_dp_integer<uint16_t>(a, bounded, vec, s);
    return run_dp_integer<uint32_t>(a, bounded, vec, s);
}

extern "C" int tdvh_dp(tdvh_ctx *ctx, const tdvh_dsi *C, uint32_t p_occ, int flags, int16_t *disp,
                       int disp_pitch, void *stream)
{
    return dp_run(ctx, C, p_occ, (float)p_occ, flags, disp, disp_pitch, stream);
}

extern "C" int tdvh_dp_f32(tdvh_ctx *ctx, const tdvh_dsi *C, float p_occ, int flags,
                           int16_t *disp, int disp_pitch, void *stream)
{
    clear_error();
    TDVH_REQUIRE(C && C->dtype == TDVH_F32, "tdvh_dp_f32 needs a float cost volume");
    TDVH_REQUIRE(p_occ >= 0.0f && p_occ < __builtin_inff(), "penalty %g not finite and >= 0",
                 (double)p_occ);
    return dp_run(ctx, C, 0, p_occ, flags, disp, disp_pitch, stream);
}
