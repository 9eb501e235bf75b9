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
// the penalised neighbours A(d-1) + P and A(d+1) + P across lanes move with
// whole-wave DPP shifts (the vacated end lane keeps the invalid value, which
// is exactly the missing neighbour of d = 0 and of d = 64K-1; cells d >= D
// carry invalid costs and stay invalid, so no disparity bound needs a test).
// Costs stream through a PF-step prefetch ring (one wave per SIMD leaves
// nothing else to hide the load latency) addressed by 32-bit range-checked
// buffer offsets. A single wave issues at most one instruction per cycle of
// its SIMD's issue turn, so the step is written for instruction count: the
// two decision bits of a cell (the comparison results) are shifted into the
// decision word with one add each (2K bits per column); CPW columns fill one
// word per lane (path word (x/CPW)*64 + lane), so that stores, which share
// the memory counter with the loads, take few of its slots. Backtrack: per
// block of 64 columns each lane loads its block's decision words together,
// then the whole wave walks the columns in lockstep with a wave-uniform
// disparity, reading the decision bits of lane z / K by readlane.
#include <atomic>
#include <cstdlib>
#include <string>
#include <type_traits>

#include "kernels.hpp"

namespace tdvh {

constexpr uint64_t DP_INF = ~0ull;
constexpr uint32_t DP_COST_LIMIT = 1u << 22; // bounded form: valid costs below this
constexpr uint32_t DP_PEN_LIMIT = 1u << 20;  // bounded form: penalties below this
constexpr int DP_MAX_DEPTH = 512;           // 8 disparities per lane
// fused SSD + DP: 16 bytes of LDS per staged column, W + radius + 1 columns
// within a workgroup's 160 KiB
constexpr int DP_SSD_MAX_WIDTH = 10236;

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
// (lane i reads lane i+1); the lane without a source keeps `old`.
template <int CTRL> __device__ __forceinline__ uint32_t wave_dpp(uint32_t old, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ __forceinline__ uint64_t wave_dpp(uint64_t old, uint64_t v)
{
    const uint32_t lo = wave_dpp<CTRL>((uint32_t)old, (uint32_t)v);
    const uint32_t hi = wave_dpp<CTRL>((uint32_t)(old >> 32), (uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
template <int CTRL> __device__ __forceinline__ float wave_dpp(float old, float v)
{
    return __builtin_bit_cast(float, wave_dpp<CTRL>(__builtin_bit_cast(uint32_t, old),
                                                    __builtin_bit_cast(uint32_t, v)));
}
template <typename T> __device__ __forceinline__ T wave_from_prev(T old, T v) { return wave_dpp<0x138>(old, v); }
template <typename T> __device__ __forceinline__ T wave_from_next(T old, T v) { return wave_dpp<0x130>(old, v); }

// 2*cw + b in one add with carry-in: the lane mask of b is the comparison
// result itself (no 0/1 materialisation, no shift).
__device__ __forceinline__ uint32_t push_bit(uint32_t cw, bool b)
{
    const uint64_t m = __builtin_amdgcn_ballot_w64(b);
    uint32_t r;
    uint64_t carry_out;
    asm("v_addc_co_u32 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry_out) : "v"(cw), "s"(m));
    return r;
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

// Minimum over the wave, valid in lane 63: row shifts 1, 2, 4, 8 leave
// each row's minimum in its lane 15, then row_bcast:15 (rows 1 and 3) and
// row_bcast:31 (rows 2 and 3) fold the rows together. Lanes a DPP read does
// not reach take the identity (all ones).
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t min_dpp(uint32_t v)
{
    return min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, CTRL, ROWS, 0xF, false));
}
// Stage s (0, 1, 2) of that fold: shifts 1 and 2, shifts 4 and 8, the two
// broadcasts.
__device__ __forceinline__ uint32_t wave_min_stage(uint32_t v, int s)
{
    if (s == 0) return min_dpp<0x112, 0xF>(min_dpp<0x111, 0xF>(v));
    if (s == 1) return min_dpp<0x118, 0xF>(min_dpp<0x114, 0xF>(v));
    return min_dpp<0x143, 0xC>(min_dpp<0x142, 0xA>(v));
}

// How a lane reads its K cells of a column: LD_ELEM one element load per
// cell (uint16 volumes with odd K or 2-byte alignment), LD_DW the K cells as
// consecutive dwords (b128/b96/b64/b32 loads; needs 4-byte aligned cells
// runs), LD_FULL as LD_DW with D == 64K (no cell of the lane range lies past
// D, so no disparity mask).
enum { LD_ELEM = 0, LD_DW = 1, LD_FULL = 2 };

// Cache policy of the cost loads: non-temporal (each cell is read once;
// measured on a 1920x1080x256 uint32 stream: 6.6 TB/s against 5.7 TB/s
// with the default policy).
constexpr int DP_LOAD_NT = 2;

template <typename TIn, int K, int LD>
__device__ __forceinline__ void dp_load_cells(buf_t r, uint32_t off, TIn (&v)[K])
{
    if constexpr (LD == LD_ELEM) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            if constexpr (sizeof(TIn) == 2)
                v[k] = __builtin_amdgcn_raw_buffer_load_b16(r, off + 2 * k, 0, DP_LOAD_NT);
            else v[k] = __builtin_bit_cast(TIn, __builtin_amdgcn_raw_buffer_load_b32(r, off + 4 * k, 0, DP_LOAD_NT));
        }
    } else {
        constexpr int NW = K * (int)sizeof(TIn) / 4;
        uint32_t w[NW];
        buf_load<NW, 0, DP_LOAD_NT>(r, off, w);
#pragma unroll
        for (int k = 0; k < K; k++) {
            if constexpr (sizeof(TIn) == 2) v[k] = (uint16_t)(w[k / 2] >> (16 * (k & 1)));
            else v[k] = __builtin_bit_cast(TIn, w[k]);
        }
    }
}

// Costs are read through buffer resources, one per block of PF columns
// (based at the block's first column, ranged to the columns of the block
// that lie in the row): lane l's cells of column x0 + u start at byte
// l*K*sizeof(TIn) + u*sxb, and columns past the row read 0 from the range
// check (never consumed), so the prefetch needs neither a clamp nor 64-bit
// address arithmetic per load. Decision words go through a resource over
// the row's path words.
// `only` (fixup launch): rows whose flag is 0 return at once. `flags_out`
// (bounded launch): receives 1 for each row that breaks the contract.
template <typename TIn, int K, int PF, int LD, int MODE>
__global__ __launch_bounds__(64) void dp_kernel(const TIn *__restrict__ C, int W, int H, int D,
                                                int64_t sx, int64_t sy, uint32_t P,
                                                float Pf, int flags, uint8_t *__restrict__ path,
                                                int16_t *__restrict__ disp, int pitch,
                                                const uint8_t *__restrict__ only,
                                                uint8_t *__restrict__ flags_out, size_t rstride,
                                                uint32_t bstride)
{
    static_assert(LD == LD_ELEM || (K * sizeof(TIn)) % 4 == 0, "dword loads need whole dwords per lane");
    constexpr int CPW = dp_cpw(K), WORD_BITS = 2 * K * CPW; // decision bits per word
    constexpr bool B32 = MODE == DP_B32;
    typedef DpCost<TIn, MODE> cost_t;
    typedef typename cost_t::type cval_t;
    const int l = threadIdx.x;
    const int y = blockIdx.x;
    if (only && only[y] == 0) return; // wave-uniform
    const TIn *crow = C + y * sy;
    const uint32_t sxb = (uint32_t)(sx * (int64_t)sizeof(TIn)), cellb = (uint32_t)(D * (int)sizeof(TIn));
    // resource over columns [x, x + n) of the row (n clipped to the row)
    auto cols = [&](int x, int n) {
        const int m = min(n, W - x);
        return buf_rsrc(crow + x * sx, m > 0 ? (uint32_t)(m - 1) * sxb + cellb : 0u);
    };
    // decision word block b (columns [b CPW, (b + 1) CPW)) of this row at
    // byte b * bstride of the resource (words past the row read 0)
    const uint32_t nblk = (uint32_t)((W + CPW - 1) / CPW);
    const buf_t prs = buf_rsrc(path + (size_t)y * rstride, (nblk - 1) * bstride + 256u);
    const uint32_t lane_off = (uint32_t)(l * K * (int)sizeof(TIn));

    auto load = [&](cval_t(&c)[K], buf_t rs, uint32_t off) {
        TIn v[K];
        dp_load_cells<TIn, K, LD>(rs, off, v);
#pragma unroll
        for (int k = 0; k < K; k++) {
            if constexpr (LD == LD_FULL) c[k] = cost_t::get(v[k]);
            else if constexpr (MODE == DP_F32) c[k] = l * K + k < D ? cost_t::get(v[k]) : __builtin_inff();
            else c[k] = l * K + k < D ? cost_t::get(v[k]) : INV32;
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
    // other cell a cost below 2^22. Kept in 32-bit lane values (no per-lane
    // branches): a violation sets bit 0 of `bad`, and once every d < D is
    // valid a running maximum of the costs does the same job.
    uint32_t bad = 0;
    uint32_t vmax = 0;
    auto check_exact = [&](const cval_t(&c)[K], int x) {
        if constexpr (B32) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int d = l * K + k;
                const uint32_t inv = ((uint32_t)(x - d) >> 31) | (uint32_t)(d >= D);
                const uint32_t t = (uint32_t)c[k] ^ (0u - inv); // invalid cells: 0 iff INV32
                const uint32_t thr = inv ? 1u : DP_COST_LIMIT;
                bad |= (uint32_t)(t >= thr);
            }
        }
    };
    auto check_max = [&](const cval_t(&c)[K]) {
        if constexpr (B32) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                if constexpr (LD == LD_FULL) vmax = max(vmax, (uint32_t)c[k]);
                else vmax = max(vmax, l * K + k < D ? (uint32_t)c[k] : 0u);
            }
        }
    };

    acc_t A[K];
    {
        cval_t c[K];
        load(c, cols(0, 1), lane_off);
        check_exact(c, 0);
#pragma unroll
        for (int k = 0; k < K; k++) {
            if constexpr (MODE == DP_F32) A[k] = c[k];
            else A[k] = c[k] == INV32 ? INF : c[k];
        }
    }
    // B32: lane minima of the last renormalised column, folded over the
    // wave across the next three steps (two DPP stages each) and read from
    // lane 63 at the next renormalisation
    uint32_t mn = 0;
    cval_t ring[PF][K];
#pragma unroll
    for (int u = 0; u < PF; u++) load(ring[u], cols(1, PF), lane_off + (uint32_t)u * sxb);
    // block x0: its refill loads read columns x0 + PF + u through `next`,
    // its decision words are word (x0 - 1) / CPW + (1 + u) / CPW (a
    // compile-time immediate per step past st_off)
    buf_t next;
    uint32_t st_off = (uint32_t)l * 4;
    // DPP results keep the invalid value in the lane without a source (lane
    // 0 for the shift from below, lane 63 for the shift from above), so the
    // previous result serves as the `old` operand
    acc_t left = INF, right = INF;

    // one step at column x = x0 + u from ring slot u (which it then
    // refills), x0 = 1 + a multiple of PF (PF % 4 == 0), so x % 4 is the
    // compile-time (1 + u) % 4 once the block is unrolled. Each cell's
    // decision is two bits, p = (c1 < c2 && c1 < c3) and q = (c2 < c3)
    // (move -1 if p, else 0 if q, else +1), shifted into cw in the order
    // (column, k, p then q), so a word of CPW columns holds its first
    // column's first p bit at bit 2*K*CPW - 1; stored when the word is full.
    // Column 0 has no decisions: its fields read (p, q) = (0, 1), "stay".
    uint32_t cw = 0x55555555u & ((1u << (2 * K)) - 1);
    auto step = [&](int u, int x, bool tail, bool exact) {
        const int ph = (1 + u) & 3; // == x & 3
        cval_t c[K];
#pragma unroll
        for (int k = 0; k < K; k++) c[k] = ring[u][k];
        if (exact) check_exact(c, x);
        else check_max(c);
        // S = A + P per cell; the neighbours' S across lanes move by DPP
        // (the end lanes keep INF, which is INF + P)
        acc_t S[K];
#pragma unroll
        for (int k = 0; k < K; k++) S[k] = dp_add(A[k], Pen);
        left = wave_from_prev(left, S[K - 1]);
        right = wave_from_next(right, S[0]);
        acc_t An[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const acc_t c1 = k == 0 ? left : S[k - 1];
            const acc_t c2 = A[k];
            const acc_t c3 = k == K - 1 ? right : S[k + 1];
            const bool q = c2 < c3;
            const acc_t m23 = q ? c2 : c3;
            const bool p = c1 < m23;
            const acc_t mv = p ? c1 : m23;
            if constexpr (MODE == DP_F32) An[k] = mv + c[k];
            else if constexpr (B32) An[k] = dp_add(mv, c[k]);
            else An[k] = (c[k] == INV32 || mv == DP_INF) ? DP_INF : mv + c[k];
            cw = push_bit(push_bit(cw, p), q);
        }
#pragma unroll
        for (int k = 0; k < K; k++) A[k] = An[k];
        if constexpr (B32) {
            if (ph == 3) {
                // subtract the minimum taken 4 columns ago, then take this
                // column's (consumed at the next renormalisation)
                const uint32_t pend = (uint32_t)__builtin_amdgcn_readlane((int)mn, 63);
#pragma unroll
                for (int k = 0; k < K; k++) A[k] -= pend;
                mn = A[0];
#pragma unroll
                for (int k = 1; k < K; k++) mn = min(mn, (uint32_t)A[k]);
            } else {
                mn = wave_min_stage(mn, ph);
            }
        }
        const int slot = ph % CPW; // == x % CPW
        if (slot == CPW - 1 || (tail && x == W - 1)) {
            // a partial last word moves up to the full word's positions
            __builtin_amdgcn_raw_buffer_store_b32(cw << ((CPW - 1 - slot) * 2 * K), prs,
                                                  st_off + (uint32_t)((1 + u) / CPW) * bstride, 0, 0);
            cw = 0;
        }
        // refill the slot, PF columns ahead, once its values are consumed
        // (so the loads reuse its registers instead of a copy of the ring);
        // the scheduling barriers keep each refill at its step, where the
        // prefetch distance stays PF through the block, while ALU work of
        // neighbouring steps may still interleave across them (mask 1: all
        // non-memory instructions)
        __builtin_amdgcn_sched_barrier(1);
        load(ring[u], next, lane_off + (uint32_t)u * sxb);
        __builtin_amdgcn_sched_barrier(1);
    };
    // whole blocks of PF steps without conditions (so the waits stay per
    // slot; PF % 4 == 0 and x0 % 4 == 1 make every x % CPW and x & 3 a
    // constant), then the tail. Blocks that may hold cells with d > x check
    // the bounded contract exactly, the others keep a running maximum.
    static_assert(PF % 4 == 0, "whole decision words and renormalisation periods per block");
    int x0 = 1;
    const uint32_t st_blk = (uint32_t)(PF / CPW) * bstride;
    for (; x0 + PF <= W && (B32 && x0 < 64 * K); x0 += PF, st_off += st_blk) {
        next = cols(x0 + PF, PF);
#pragma unroll
        for (int u = 0; u < PF; u++) step(u, x0 + u, false, true);
    }
    for (; x0 + PF <= W; x0 += PF, st_off += st_blk) {
        next = cols(x0 + PF, PF);
#pragma unroll
        for (int u = 0; u < PF; u++) step(u, x0 + u, false, false);
    }
    next = cols(W, 0); // the tail's refills read nothing
    if (x0 == W) { // the last block ended the row inside a word
        if ((W - 1) % CPW != CPW - 1)
            __builtin_amdgcn_raw_buffer_store_b32(cw << ((CPW - 1 - (W - 1) % CPW) * 2 * K), prs,
                                                  (uint32_t)((W - 1) / CPW) * bstride + l * 4, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (x0 + u < W) step(u, x0 + u, true, true);

    if constexpr (B32) {
        // the row is exact unless some cost broke the contract
        const bool row_bad = __builtin_amdgcn_ballot_w64(bad != 0 || vmax >= DP_COST_LIMIT) != 0;
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
    // lane): each lane loads its block's words with independent loads and
    // turns every (p, q) field into the move + 1 (p -> 0, q -> 1, else 2);
    // then the whole wave walks the columns in lockstep with a wave-uniform
    // disparity z, taking column x's field from lane z / K by readlane, and
    // lane j of the block keeps column j's disparity. A -1
    // move never starts at d = 0 (its c1 is the invalid value, never below
    // c2 and c3) and a +1 move from d = D - 1 stays at D - 1, so the walk
    // needs only an upper clamp.
    constexpr int WPB = 64 / CPW; // decision words per block and lane
    const bool refbt = flags & TDVH_DP_REF_BACKTRACK;
    uint32_t z = (uint32_t)__builtin_amdgcn_readfirstlane(start);
    int16_t *out = disp + (size_t)y * pitch;
    // the next block's words are fetched before the current block is walked
    auto fetch = [&](int wb, uint32_t(&wv)[WPB]) {
#pragma unroll
        for (int i = 0; i < WPB; i++) // words past the row read 0
            wv[i] = __builtin_amdgcn_raw_buffer_load_b32(prs, (uint32_t)(wb + i) * bstride + l * 4, 0, 0);
    };
    auto walk = [&](int wb, uint32_t(&wv)[WPB], bool partial) {
#pragma unroll
        for (int i = 0; i < WPB; i++) {
            const uint32_t w = wv[i];
            const uint32_t np = ~(w >> 1) & 0x55555555u; // fields with p = 0, at their low bit
            wv[i] = (np & w) | ((np & ~w) << 1);
        }
        uint32_t mine = 0;
#pragma unroll
        for (int i = WPB - 1; i >= 0; i--) {
#pragma unroll
            for (int b = CPW - 1; b >= 0; b--) {
                if (partial && (wb + i) * CPW + b > W - 1) continue; // wave-uniform
                const uint32_t word = (uint32_t)__builtin_amdgcn_readlane((int)wv[i], (int)(z / K));
                const uint32_t m = (word >> (WORD_BITS - 2 - 2 * (K * b + (int)(z % K)))) & 3u;
                const uint32_t nz = min(z + m - 1, (uint32_t)(D - 1));
                if (refbt) z = nz;
                mine = l == CPW * i + b ? z : mine;
                if (!refbt) z = nz;
            }
        }
        if (CPW * wb + l < W) out[CPW * wb + l] = (int16_t)mine;
    };
    uint32_t cur[WPB], nxt[WPB];
    int wb = ((W - 1) / CPW) & ~(WPB - 1);
    fetch(wb, cur);
    if (wb >= WPB) fetch(wb - WPB, nxt);
    walk(wb, cur, true); // the last block may end inside it
    for (wb -= WPB; wb >= 0; wb -= WPB) {
#pragma unroll
        for (int i = 0; i < WPB; i++) cur[i] = nxt[i];
        if (wb >= WPB) fetch(wb - WPB, nxt);
        walk(wb, cur, false);
    }
}

struct DpLaunch {
    const void *C;
    int W, H, D;
    int64_t sx, sy; // strides in elements
    uint32_t P;
    float Pf;
    int flags;
    uint8_t *path;
    int16_t *disp;
    int pitch;
    const uint8_t *only;
    uint8_t *flags_out;
    int ld;          // LD_ELEM, LD_DW or LD_FULL
    size_t rstride;  // bytes between the decision words of rows y and y + 1
    uint32_t bstride; // bytes between a row's word blocks b and b + 1
};

// Prefetch depth (columns of costs in flight per row): 16 for K <= 4, 8
// above (the ring then holds 48 or 64 cost registers).
template <typename TIn, int K, int LD, int MODE>
static void launch_dp_k(const DpLaunch &a, hipStream_t s)
{
    constexpr int PF = K > 4 ? 8 : 16;
    dp_kernel<TIn, K, PF, LD, MODE><<<a.H, 64, 0, s>>>(
        (const TIn *)a.C, a.W, a.H, a.D, a.sx, a.sy, a.P, a.Pf, a.flags, a.path, a.disp,
        a.pitch, a.only, a.flags_out, a.rstride, a.bstride);
}

template <typename TIn, int K, int MODE>
static void launch_dp_ld(const DpLaunch &a, hipStream_t s)
{
    if constexpr ((K * sizeof(TIn)) % 4 != 0) {
        launch_dp_k<TIn, K, LD_ELEM, MODE>(a, s); // odd K of uint16 cells
    } else {
        if (a.ld == LD_FULL) launch_dp_k<TIn, K, LD_FULL, MODE>(a, s);
        else if (a.ld == LD_DW || sizeof(TIn) == 4) launch_dp_k<TIn, K, LD_DW, MODE>(a, s);
        else if constexpr (sizeof(TIn) == 2) launch_dp_k<TIn, K, LD_ELEM, MODE>(a, s);
    }
}

// K = disparities per lane: 1..4, then 6 (D <= 384) and 8 (D <= 512).
__host__ __device__ constexpr int dp_k(int D)
{
    return (D + 63) / 64 <= 4 ? (D + 63) / 64 : ((D + 63) / 64 <= 6 ? 6 : 8);
}

template <typename TIn, int MODE>
static void launch_dp(const DpLaunch &a, hipStream_t s)
{
    switch (dp_k(a.D)) {
    case 1: launch_dp_ld<TIn, 1, MODE>(a, s); break;
    case 2: launch_dp_ld<TIn, 2, MODE>(a, s); break;
    case 3: launch_dp_ld<TIn, 3, MODE>(a, s); break;
    case 4: launch_dp_ld<TIn, 4, MODE>(a, s); break;
    case 6: launch_dp_ld<TIn, 6, MODE>(a, s); break;
    default: launch_dp_ld<TIn, 8, MODE>(a, s); break;
    }
}

// Where the decision words go. Rows interleaved by word block (block b of
// row y at (b H + y) 256 bytes): every row's wave is at about the same
// column, so at any time the rows' stores, and later the backtracks' loads,
// fall in one contiguous band, as the cost reads of a strided volume do.
// Row-major (row y's blocks contiguous from y * row bytes) where the
// interleaved offsets would not fit 32 bits, or with TDVH_DP_WORDS=rows.
static void dp_word_layout(int W, int H, int K, size_t *rstride, uint32_t *bstride)
{
    static const bool rows = [] {
        const char *e = getenv("TDVH_DP_WORDS");
        return e && std::string(e) == "rows";
    }();
    const size_t nblk = (size_t)(W + dp_cpw(K) - 1) / dp_cpw(K);
    // the backtrack's fetch reads up to 64 / CPW blocks past the row's last
    if (!rows && (nblk + 32) * (size_t)H * 256 < (1ull << 32)) {
        *rstride = 256;
        *bstride = (uint32_t)H * 256u;
    } else {
        *rstride = dp_row_bytes(W, K);
        *bstride = 256;
    }
}

// Bounded uint32 launch, then the uint64 recomputation of the rows it
// flagged (a launch whose waves return at once when no row is flagged).
template <typename TIn>
static int run_dp_integer(const DpLaunch &a, bool bounded, hipStream_t s)
{
    if (!bounded) {
        prof_begin(s);
        launch_dp<TIn, DP_U64>(a, s);
        return check_launch("dp_kernel", s);
    }
    prof_begin(s);
    launch_dp<TIn, DP_B32>(a, s);
    TDVH_TRY(check_launch("dp_kernel", s));
    DpLaunch f = a;
    f.only = a.flags_out;
    f.flags_out = nullptr;
    prof_begin(s);
    launch_dp<TIn, DP_U64>(f, s);
    return check_launch("dp_fixup_kernel", s);
}


// ---- fused SSD window + DP ----------------------------------------------------
//
// The SSD -> DP path without a cost volume: the wave that scans row y
// computes that row's (2r+1)x(2r+1) SSD costs itself, column by column, as
// it goes. The 2r+1 image rows around y (clamped) of both images are staged
// in LDS transposed, 8 bytes per column (row i in byte i, zero above 2r+1),
// with r + 1 copies of the last column after it, so a column of either
// image is one 8-byte read. With the column sums
//   col(u, d) = sum_i (L(cx(u), y_i) - R(cx(u - d), y_i))^2
//             = sum L^2 + sum R^2 - 2 sum L R     (v_dot4_u32_u8, 4 rows each)
// the window cost is the running sum S(x, d) = S(x - 1, d) + col(x + r, d) -
// col(x - r - 1, d), i.e. the sum over u in [x - r, x + r] of col(u, d):
// the SSD of ssd.cu:15-29 with clamp-to-edge reads, exact in 32 bits (below
// 49 * 255^2). A lane holds the cells d = d0 + k (k < K). Each step it reads
// one new right column, cx(x + r - d0); cell k uses the column read k steps
// earlier (a 4-slot history), and the column sum leaving the window is the
// one that entered it 2r + 1 steps earlier (a ring of 2r + 1 sums per cell).
// The loop is unrolled over lcm(4, 2r + 1) steps, so every history slot and
// ring entry is a fixed register. The DP is the bounded 32-bit form of
// dp_kernel (costs below 2^22 by construction, invalid exactly where d > x
// or d >= D), with the same decision words and backtrack, so the
// disparities equal those of the SSD volume followed by the DP.
__host__ __device__ constexpr int dp_ssd_unroll(int rad)
{
    return (2 * rad + 1) % 2 == 0 ? 0 : 4 * (2 * rad + 1); // lcm(4, odd) = 4 * odd
}

// Stages row y's 2r+1 image rows (clamped) of both images transposed into
// lc / rc, columns u = t, t + nt, ... (W + r + 1 columns).
template <int RAD>
__device__ void ssd_stage_rows(const uint8_t *__restrict__ Limg, const uint8_t *__restrict__ Rimg,
                               int W, int H, int pitch, int y, uint2 *lc, uint2 *rc, int t, int nt)
{
    constexpr int NR = 2 * RAD + 1;
    for (int u = t; u < W + RAD + 1; u += nt) {
        const int uu = min(u, W - 1);
        uint32_t lw[2] = {0u, 0u}, rw[2] = {0u, 0u};
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int yy = min(max(y - RAD + i, 0), H - 1);
            lw[i >> 2] |= (uint32_t)Limg[(size_t)yy * pitch + uu] << (8 * (i & 3));
            rw[i >> 2] |= (uint32_t)Rimg[(size_t)yy * pitch + uu] << (8 * (i & 3));
        }
        lc[u] = make_uint2(lw[0], lw[1]);
        rc[u] = make_uint2(rw[0], rw[1]);
    }
}

__device__ __forceinline__ uint32_t ssd_sq(uint2 v)
{
    return __builtin_amdgcn_udot4(v.x, v.x, __builtin_amdgcn_udot4(v.y, v.y, 0u, false), false);
}
__device__ __forceinline__ uint32_t ssd_dot(uint2 a, uint2 b)
{
    return __builtin_amdgcn_udot4(a.x, b.x, __builtin_amdgcn_udot4(a.y, b.y, 0u, false), false);
}

// One lane's window costs along its row: cells d = d0 + k, k < K.
template <int K, int RAD> struct SsdStream {
    static constexpr int NR = 2 * RAD + 1;
    const uint2 *lc, *rc;
    int d0;
    uint32_t S[K], ring[NR][K], qn[4], padm[K];
    uint2 hn[4], Ln;
    int thr[K];

    // the window before column 0 (u in [-1 - r, r - 1]): its column sums
    // fill the ring (the sum of column u sits at (u - r) mod (2r + 1), the
    // step that brings it in) and their total starts S; then the right
    // columns entering at the three steps before column 0, and column 0's
    // reads
    __device__ __forceinline__ void init(const uint2 *lc_, const uint2 *rc_, int d0_, int D)
    {
        lc = lc_;
        rc = rc_;
        d0 = d0_;
#pragma unroll
        for (int k = 0; k < K; k++) S[k] = 0;
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int u = -1 - RAD + i;
            const uint2 Lc = lc[max(u, 0)];
            const uint32_t sl = ssd_sq(Lc);
#pragma unroll
            for (int k = 0; k < K; k++) {
                const uint2 Rc = rc[max(u - d0 - k, 0)];
                ring[i][k] = sl + ssd_sq(Rc) - 2u * ssd_dot(Lc, Rc);
                S[k] += ring[i][k];
            }
        }
#pragma unroll
        for (int j = 1; j <= 3; j++) {
            hn[(4 - j) & 3] = rc[max(RAD - j - d0, 0)];
            qn[(4 - j) & 3] = ssd_sq(hn[(4 - j) & 3]);
        }
        // invalid cells: d > x (and every x for d >= D); past the first 64K
        // columns only d >= D, a constant mask per cell
#pragma unroll
        for (int k = 0; k < K; k++) {
            thr[k] = d0 + k >= D ? 0x7FFFFFFF : d0 + k;
            padm[k] = d0 + k >= D ? INV32 : 0u;
        }
        Ln = lc[RAD];
        hn[0] = rc[max(RAD - d0, 0)];
    }

    // costs of column x (ri = x mod (2r + 1), slot = x mod 4: compile-time
    // in the unrolled loop), then the reads of column x + 1 (in bounds up to
    // x + 1 = W thanks to the copies after the last column)
    __device__ __forceinline__ void costs(int x, int ri, int slot, bool edge, uint32_t (&c)[K])
    {
        const uint2 ln = Ln;
        const uint32_t sln = ssd_sq(ln);
        qn[slot] = ssd_sq(hn[slot]);
        uint32_t dn[K];
#pragma unroll
        for (int k = 0; k < K; k++) dn[k] = ssd_dot(ln, hn[(slot - k) & 3]);
        Ln = lc[x + 1 + RAD];
        hn[(slot + 1) & 3] = rc[max(x + 1 + RAD - d0, 0)];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t cn = sln + qn[(slot - k) & 3] - 2u * dn[k];
            S[k] += cn - ring[ri][k];
            ring[ri][k] = cn;
            if (edge) c[k] = x < thr[k] ? INV32 : S[k];
            else c[k] = S[k] | padm[k];
        }
    }
};

// The bounded DP of one lane's K cells (dp_kernel's step): path sums A, the
// wave minimum being folded, the neighbours' penalised sums and the decision
// word of the current 4 columns.
template <int K> struct DpLane {
    uint32_t A[K], mn, left, right, cw;

    __device__ __forceinline__ void start(const uint32_t (&c)[K])
    {
#pragma unroll
        for (int k = 0; k < K; k++) A[k] = c[k];
        mn = 0;
        left = right = INV32;
        cw = 0x55555555u & ((1u << (2 * K)) - 1); // column 0: "stay"
    }

    // column x >= 1 (slot = x mod 4): every 4 columns the wave minimum of 4
    // columns earlier is subtracted
    __device__ __forceinline__ void step(const uint32_t (&c)[K], uint32_t P, int slot)
    {
        uint32_t Sp[K];
#pragma unroll
        for (int k = 0; k < K; k++) Sp[k] = dp_add(A[k], P);
        left = wave_from_prev(left, Sp[K - 1]);
        right = wave_from_next(right, Sp[0]);
        uint32_t An[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint32_t c1 = k == 0 ? left : Sp[k - 1];
            const uint32_t c2 = A[k];
            const uint32_t c3 = k == K - 1 ? right : Sp[k + 1];
            const bool q = c2 < c3;
            const uint32_t m23 = q ? c2 : c3;
            const bool p = c1 < m23;
            An[k] = dp_add(p ? c1 : m23, c[k]);
            cw = push_bit(push_bit(cw, p), q);
        }
#pragma unroll
        for (int k = 0; k < K; k++) A[k] = An[k];
        if (slot == 3) {
            const uint32_t pend = (uint32_t)__builtin_amdgcn_readlane((int)mn, 63);
#pragma unroll
            for (int k = 0; k < K; k++) A[k] -= pend;
            mn = A[0];
#pragma unroll
            for (int k = 1; k < K; k++) mn = min(mn, A[k]);
        } else {
            mn = wave_min_stage(mn, slot);
        }
    }
};

// Lowest-index argmin over the wave of the cells d = d0 + k with sum A[k]
// (every lane gets it; d = 0x7FFFFFFF when no cell is below INV32).
template <int K>
__device__ __forceinline__ void wave_best(const uint32_t (&A)[K], int d0, uint32_t &bv, int &bd)
{
    bv = INV32;
    bd = 0x7FFFFFFF;
#pragma unroll
    for (int k = 0; k < K; k++)
        if (A[k] < bv) {
            bv = A[k];
            bd = d0 + k;
        }
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t ov = (uint32_t)__shfl_xor((int)bv, off, 64);
        const int od = __shfl_xor(bd, off, 64);
        if (ov < bv || (ov == bv && od < bd)) {
            bv = ov;
            bd = od;
        }
    }
}

// Backtrack of one row from disparity z over its decision words (per
// 64-column block each lane loads its words, the next block's while the
// current one is walked; the wave walks with a wave-uniform z), as
// dp_kernel's.
template <int K>
__device__ __forceinline__ void dp_ssd_backtrack(buf_t prs, uint32_t bstride, int W, int D,
                                                 bool refbt, uint32_t z,
                                                 int16_t *__restrict__ out, int l)
{
    constexpr int CPW = dp_cpw(K), WORD_BITS = 2 * K * CPW, WPB = 64 / CPW;
    auto fetch = [&](int wb, uint32_t(&wv)[WPB]) {
#pragma unroll
        for (int i = 0; i < WPB; i++)
            wv[i] = __builtin_amdgcn_raw_buffer_load_b32(prs, (uint32_t)(wb + i) * bstride + l * 4, 0, 0);
    };
    auto walk = [&](int wb, uint32_t(&wv)[WPB], bool partial) {
#pragma unroll
        for (int i = 0; i < WPB; i++) {
            const uint32_t w = wv[i];
            const uint32_t np = ~(w >> 1) & 0x55555555u;
            wv[i] = (np & w) | ((np & ~w) << 1);
        }
        uint32_t mine = 0;
#pragma unroll
        for (int i = WPB - 1; i >= 0; i--) {
#pragma unroll
            for (int b = CPW - 1; b >= 0; b--) {
                if (partial && (wb + i) * CPW + b > W - 1) continue;
                const uint32_t word = (uint32_t)__builtin_amdgcn_readlane((int)wv[i], (int)(z / K));
                const uint32_t m = (word >> (WORD_BITS - 2 - 2 * (K * b + (int)(z % K)))) & 3u;
                const uint32_t nz = min(z + m - 1, (uint32_t)(D - 1));
                if (refbt) z = nz;
                mine = l == CPW * i + b ? z : mine;
                if (!refbt) z = nz;
            }
        }
        if (CPW * wb + l < W) out[CPW * wb + l] = (int16_t)mine;
    };
    uint32_t cur[WPB], nxt[WPB];
    int wb = ((W - 1) / CPW) & ~(WPB - 1);
    fetch(wb, cur);
    if (wb >= WPB) fetch(wb - WPB, nxt);
    walk(wb, cur, true);
    for (wb -= WPB; wb >= 0; wb -= WPB) {
#pragma unroll
        for (int i = 0; i < WPB; i++) cur[i] = nxt[i];
        if (wb >= WPB) fetch(wb - WPB, nxt);
        walk(wb, cur, false);
    }
}

// The forward scan of a row by one wave (lane l: d = l*K + k) from its
// staged columns, storing the decision words; returns the row's start.
template <int K, int RAD>
__device__ __forceinline__ uint32_t dp_ssd_scan(const uint2 *lc, const uint2 *rc, int W, int D,
                                                uint32_t P, int flags, buf_t prs,
                                                uint32_t bstride, int l)
{
    static_assert(dp_cpw(K) == 4, "four columns per decision word");
    constexpr int NR = 2 * RAD + 1, UNR = dp_ssd_unroll(RAD);
    SsdStream<K, RAD> fs;
    fs.init(lc, rc, l * K, D);
    DpLane<K> dl;
    {
        uint32_t c[K];
        fs.costs(0, 0, 0, true, c);
        dl.start(c);
    }
    auto step = [&](int x, int ri, int slot, bool edge) {
        uint32_t c[K];
        fs.costs(x, ri, slot, edge, c);
        dl.step(c, P, slot);
        if (slot == 3) { // compile-time per unrolled step: no branch in the loop body
            __builtin_amdgcn_raw_buffer_store_b32(dl.cw, prs, (uint32_t)(x >> 2) * bstride + l * 4, 0, 0);
            dl.cw = 0;
        }
    };
    // blocks of UNR steps from x = 1 (so x mod NR and x mod 4 are fixed per
    // unrolled step), then the last partial block with a guard per step
    int x0 = 1;
    for (; x0 + UNR <= W && x0 < 64 * K; x0 += UNR) { // blocks with cells d > x
#pragma unroll
        for (int u = 0; u < UNR; u++) step(x0 + u, (1 + u) % NR, (1 + u) & 3, true);
    }
    for (; x0 + UNR <= W; x0 += UNR) {
#pragma unroll
        for (int u = 0; u < UNR; u++) step(x0 + u, (1 + u) % NR, (1 + u) & 3, false);
    }
#pragma unroll
    for (int u = 0; u < UNR - 1; u++)
        if (x0 + u < W) step(x0 + u, (1 + u) % NR, (1 + u) & 3, true);
    // a partial last word (column W - 1 not the fourth of its word; W = 1
    // included) moves up to the full word's bits
    if (((W - 1) & 3) != 3)
        __builtin_amdgcn_raw_buffer_store_b32(dl.cw << ((3 - ((W - 1) & 3)) * 2 * K), prs,
                                              (uint32_t)((W - 1) >> 2) * bstride + l * 4, 0, 0);
    // start: lowest-index argmin of the last column (or d = 0)
    if (flags & TDVH_DP_START_ZERO) return 0;
    uint32_t bv;
    int bd;
    wave_best<K>(dl.A, l * K, bv, bd);
    return bd == 0x7FFFFFFF ? 0 : (uint32_t)bd;
}

template <int K, int RAD>
__global__ __launch_bounds__(64) void dp_ssd_kernel(const uint8_t *__restrict__ Limg,
                                                    const uint8_t *__restrict__ Rimg, int W,
                                                    int H, int pitch, int D, uint32_t P,
                                                    int flags, uint8_t *__restrict__ path,
                                                    int16_t *__restrict__ disp, int dpitch,
                                                    size_t rstride, uint32_t bstride)
{
    static_assert(K >= 1 && K <= 4 && RAD >= 1 && RAD <= 3, "K <= 4, radius 1..3");
    extern __shared__ uint2 cols_lds[];
    const int l = threadIdx.x, y = blockIdx.x;
    const int WP = W + RAD + 1; // staged columns per image
    uint2 *const lc = cols_lds, *const rc = cols_lds + WP;
    ssd_stage_rows<RAD>(Limg, Rimg, W, H, pitch, y, lc, rc, l, 64);
    __syncthreads();
    // decision word block b of this row at byte b * bstride
    const uint32_t nblk = (uint32_t)((W + 3) / 4);
    const buf_t prs = buf_rsrc(path + (size_t)y * rstride, (nblk - 1) * bstride + 256u);
    const uint32_t start = dp_ssd_scan<K, RAD>(lc, rc, W, D, P, flags, prs, bstride, l);
    const uint32_t z = (uint32_t)__builtin_amdgcn_readfirstlane((int)start);
    dp_ssd_backtrack<K>(prs, bstride, W, D, flags & TDVH_DP_REF_BACKTRACK, z,
                        disp + (size_t)y * dpitch, l);
}

// LDS bytes per workgroup of a device, queried once (64 KiB, what every
// gfx9 device has, if the query fails).
static int device_lds_per_workgroup(int dev)
{
    static std::atomic<int> cache[64];
    if (dev < 0 || dev >= 64) return 65536;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v == 0) {
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
            v <= 0)
            v = 65536;
        cache[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}

template <int K>
static void launch_dp_ssd_k(const uint8_t *L, const uint8_t *R, int W, int H, int pitch, int D,
                            int radius, uint32_t P, int flags, uint8_t *path, int16_t *disp,
                            int dpitch, hipStream_t s)
{
    const size_t lds = (size_t)(W + radius + 1) * 16;
    size_t rs;
    uint32_t bs;
    dp_word_layout(W, H, K, &rs, &bs);
    switch (radius) {
    case 1: dp_ssd_kernel<K, 1><<<H, 64, lds, s>>>(L, R, W, H, pitch, D, P, flags, path, disp, dpitch, rs, bs); break;
    case 2: dp_ssd_kernel<K, 2><<<H, 64, lds, s>>>(L, R, W, H, pitch, D, P, flags, path, disp, dpitch, rs, bs); break;
    default: dp_ssd_kernel<K, 3><<<H, 64, lds, s>>>(L, R, W, H, pitch, D, P, flags, path, disp, dpitch, rs, bs); break;
    }
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
    const int K = dp_k(D);
    const int64_t esz = C->dtype == TDVH_U16 ? 2 : 4;
    // blocks of up to 16 columns are read through 32-bit range-checked
    // byte offsets
    TDVH_REQUIRE(C->stride_x >= D && C->stride_y >= 0, "bad strides %lld, %lld",
                 (long long)C->stride_x, (long long)C->stride_y);
    TDVH_REQUIRE(16 * C->stride_x * esz + D * esz < (1ll << 31) &&
                     dp_row_bytes(W, K) < (1ull << 31),
                 "column stride %lld or width %d too large", (long long)C->stride_x, W);
    TDVH_REQUIRE(((uintptr_t)C->ptr & (esz - 1)) == 0, "volume not aligned to its cells");
    void *pb = nullptr;
    TDVH_TRY(ctx->scratch("dp_path", dp_row_bytes(W, K) * H, &pb));
    // dword loads where every lane's cells start on a 4-byte boundary
    const bool dw = esz == 4 || (((uintptr_t)C->ptr & 3) == 0 && C->stride_x % 2 == 0 &&
                                 C->stride_y % 2 == 0 && K % 2 == 0);
    DpLaunch a{C->ptr, W, H, D, C->stride_x, C->stride_y,
               p_occ, pf, flags, (uint8_t *)pb, disp, disp_pitch, nullptr, nullptr,
               !dw ? LD_ELEM : (D == 64 * K ? LD_FULL : LD_DW), 0, 0};
    dp_word_layout(W, H, K, &a.rstride, &a.bstride);
    if (C->dtype == TDVH_F32) {
        prof_begin(s);
        launch_dp<float, DP_F32>(a, s);
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
    if (C->dtype == TDVH_U16) return run_dp_integer<uint16_t>(a, bounded, s);
    return run_dp_integer<uint32_t>(a, bounded, s);
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


extern "C" int tdvh_stereo_ssd_dp(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right,
                                  int width, int height, int pitch, int depth, int radius,
                                  uint32_t p_occ, int flags, int16_t *disp, int disp_pitch,
                                  void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(left && right && disp, "null pointer argument");
    TDVH_REQUIRE(width > 0 && height > 0 && pitch >= width, "bad image %dx%d pitch %d", width,
                 height, pitch);
    TDVH_REQUIRE(width <= DP_SSD_MAX_WIDTH, "width %d above %d (the staged rows fill 160 KiB of LDS)",
                 width, DP_SSD_MAX_WIDTH);
    TDVH_REQUIRE(depth > 0 && depth <= 256, "depth %d outside [1, 256]", depth);
    TDVH_REQUIRE(radius >= 1 && radius <= 3, "window radius %d outside [1, 3]", radius);
    TDVH_REQUIRE(p_occ < DP_PEN_LIMIT, "p_occ %u not below 2^20", p_occ);
    TDVH_REQUIRE((flags & ~(TDVH_DP_START_ZERO | TDVH_DP_REF_BACKTRACK)) == 0, "bad flags 0x%x",
                 flags);
    TDVH_REQUIRE(disp_pitch >= width, "disp_pitch %d < width %d", disp_pitch, width);
    hipStream_t s = (hipStream_t)stream;
    const int K = dp_k(depth);
    void *pb = nullptr;
    TDVH_TRY(ctx->scratch("dp_path", dp_row_bytes(width, K) * height, &pb));
    const int max_lds = device_lds_per_workgroup(ctx->device);
    TDVH_REQUIRE((width + radius + 1) * 16 <= max_lds,
                 "width %d: %d bytes of staged columns above the device's %d bytes of LDS per workgroup",
                 width, (width + radius + 1) * 16, max_lds);
    prof_begin(s);
    switch (K) {
    case 1: launch_dp_ssd_k<1>(left, right, width, height, pitch, depth, radius, p_occ, flags, (uint8_t *)pb, disp, disp_pitch, s); break;
    case 2: launch_dp_ssd_k<2>(left, right, width, height, pitch, depth, radius, p_occ, flags, (uint8_t *)pb, disp, disp_pitch, s); break;
    case 3: launch_dp_ssd_k<3>(left, right, width, height, pitch, depth, radius, p_occ, flags, (uint8_t *)pb, disp, disp_pitch, s); break;
    default: launch_dp_ssd_k<4>(left, right, width, height, pitch, depth, radius, p_occ, flags, (uint8_t *)pb, disp, disp_pitch, s); break;
    }
    return check_launch("dp_ssd_kernel", s);
}
