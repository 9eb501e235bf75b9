// Device code of the standard-geometry SGM (shared by the per-pair kernels
// in sgm.hip and the persistent batch kernel in batch.hip).
#pragma once

#include "kernels.hpp"

namespace tdvh {


struct LineSet {
    int prio; // raise the issue priority of waves scanning long lines
    int ndirs;
    int dir[8];
    int wave_begin[9]; // prefix sums of waves (4 lines per wave) per slot
};

// Start pixel and length of standard-geometry line `line` of direction `dir`.
__host__ __device__ inline void line_geom(int dir, int line, int W, int H, int &sx, int &sy,
                                          int &len)
{
    switch (dir) {
    case 0: sx = 0; sy = line; len = W; break;
    case 1: sx = W - 1; sy = line; len = W; break;
    case 2: sx = line; sy = 0; len = H; break;
    case 3: sx = line; sy = H - 1; len = H; break;
    case 4: // (1,1): top row, then left column
        if (line < W) { sx = line; sy = 0; } else { sx = 0; sy = line - W + 1; }
        len = min(W - sx, H - sy);
        break;
    case 5: // (-1,-1): bottom row, then right column
        if (line < W) { sx = line; sy = H - 1; } else { sx = W - 1; sy = H - 2 - (line - W); }
        len = min(sx + 1, sy + 1);
        break;
    case 6: // (-1,1): top row, then right column
        if (line < W) { sx = line; sy = 0; } else { sx = W - 1; sy = line - W + 1; }
        len = min(sx + 1, H - sy);
        break;
    default: // 7, (1,-1): bottom row, then left column
        if (line < W) { sx = line; sy = H - 1; } else { sx = 0; sy = H - 2 - (line - W); }
        len = min(W - sx, sy + 1);
        break;
    }
}

// Device-side inputs of the SGM kernels.
struct SgmArgs {
    int W, H, D, P1, P2;
    int m;              // words per lane and plane (>= ceil(D/32); cells d >= D are padding)
    const uint32_t *C;  // SPLIT16 cost volume (staged), or null
    const uint32_t *LT; // W*H packed left BT terms (lt_unpack) (fused)
    const uint32_t *RT; // 3 planes H*Wr: paired right BT terms r2, rmin, rmax (fused)
    int Wr;             // row length of RT: W + 16m
    uint8_t *planes;    // path-cost planes (E bytes or L words)
    size_t plane_bytes;
    uint8_t *trash;     // >= 64 * 64 bytes: stores of lanes past their line
    uint8_t *CR;        // relative cost bytes (sgm_rel_pixel), or null
};

// Cost sources of the buffer line scans: a SPLIT16 uint16 volume, BT
// recomputed from the term planes, or the relative cost bytes.
constexpr int SRC_WORDS = 0;
constexpr int SRC_TERMS = 1;
constexpr int SRC_REL = 2;

// Packed uint16 saturating subtract / max.
__device__ __forceinline__ uint32_t pk_subs(uint32_t a, uint32_t b)
{
    return u32(__builtin_elementwise_sub_sat(v2(a), v2(b)));
}
__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b)
{
    return u32(__builtin_elementwise_max(v2(a), v2(b)));
}

// Invalid-cell mask of lane j, word r: d >= D (padding) or d > x; M words
// per lane, LPL lanes per pixel (word j*M + r holds d and d + LPL*M).
template <int M, int LPL = 16>
__device__ __forceinline__ uint32_t cell_mask(int j, int r, int D, int x)
{
    const int dlo = j * M + r, dhi = LPL * M + j * M + r;
    return ((dlo >= D || dlo > x) ? 0x0000FFFFu : 0u) | ((dhi >= D || dhi > x) ? 0xFFFF0000u : 0u);
}

// Splatted left BT terms of a pixel from its packed word (byte 0: L,
// byte 1: 2L - lmin, byte 2: lmax - 2L; both differences lie in [0, 255]):
// l2 = 2L, lmin and lmax in both halves.
__device__ __forceinline__ void lt_unpack(uint32_t w, uint32_t &l2, uint32_t &lmn, uint32_t &lmx)
{
    const uint32_t l = __builtin_amdgcn_perm(w, w, 0x0C000C00u); // L | L << 16
    l2 = pk_add(l, l);
    lmn = pk_sub(l2, __builtin_amdgcn_perm(w, w, 0x0C010C01u));
    lmx = pk_add(l2, __builtin_amdgcn_perm(w, w, 0x0C020C02u));
}

// BT2 costs of the 2M cells a lane owns at pixel (x, y). lt: splatted left
// terms of the pixel; the right terms come as words i0 + r of the paired
// planes (word i holds the terms of right pixels W-1-i and W-1-i-16m).
template <int M>
__device__ __forceinline__ void bt_costs(const uint32_t (&lt)[3], const uint32_t (&r2)[M],
                                         const uint32_t (&rmin)[M], const uint32_t (&rmax)[M],
                                         int j, int D, int x, uint32_t (&c)[M])
{
    const bool edge = x < 32 * M - 1 || D < 32 * M; // some cell of the vector is invalid
#pragma unroll
    for (int r = 0; r < M; r++) {
        const uint32_t a = pk_max(pk_subs(lt[0], rmax[r]), pk_subs(rmin[r], lt[0]));
        const uint32_t b = pk_max(pk_subs(r2[r], lt[2]), pk_subs(lt[1], r2[r]));
        c[r] = pk_min(a, b);
        if (edge) c[r] |= cell_mask<M>(j, r, D, x);
    }
}

// Min over aligned groups of G lanes (G = 2, 4, 8 or 16); every lane of a
// group receives it.
template <int G>
__device__ __forceinline__ uint32_t group_min_u32(uint32_t v)
{
    if constexpr (G == 16) return row_min_u32(v);
    if constexpr (G == 8) // lane i <-> 7-i inside each half row
        v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));
    if constexpr (G >= 4) // quad_perm [2,3,0,1]
        v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));
    // quad_perm [1,0,3,2]
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));
    return v;
}

// One SGM step on the packed vector of a line. Lp: previous L (in/out),
// c: costs of the current pixel, m: previous minimum (in/out); E receives
// L - C = min(...) - m, in [0, P2] for valid cells.
// Byte selectors of the neighbour fix-ups at the ends of a 16-lane row
// (v_perm_b32: selector bytes 0-3 pick bytes of the operand, 13 gives 0xFF).
// Lane 0: d-1 of d=0 is absent and d-1 of d=16m is lane 15's low half, i.e.
// (A << 16) | 0xFFFF; lane 15: d+1 of d=16m-1 is lane 0's high half and d+1
// of the last padded d is absent, i.e. (B >> 16) | 0xFFFF0000; other lanes
// keep the rotated word.
__device__ __forceinline__ uint32_t nb_sel_lo(int j) { return j == 0 ? 0x01000D0Du : 0x03020100u; }
template <int LPL = 16> __device__ __forceinline__ uint32_t nb_sel_hi(int j)
{
    return j == LPL - 1 ? 0x0D0D0302u : 0x03020100u;
}

// LPL lanes per line (16, or 8 with 2M words per lane for M <= 2: the same
// word order, so the same planes): j is the lane within the line's group.
template <int M, int LPL = 16>
__device__ __forceinline__ void sgm_step(uint32_t (&Lp)[M], const uint32_t (&c)[M], uint32_t &m,
                                         int j, uint32_t p1, uint32_t P2, uint32_t (&E)[M])
{
    static_assert(LPL == 16 || LPL == 8, "16 or 8 lanes per line");
    // row rotations write every lane of the row (no "old" operand), then one
    // per-lane byte permute applies the end-of-row fix-ups
    uint32_t A, B;
    if constexpr (LPL == 16) {
        A = (uint32_t)__builtin_amdgcn_mov_dpp((int)Lp[M - 1], 0x121, 0xF, 0xF, false);
        B = (uint32_t)__builtin_amdgcn_mov_dpp((int)Lp[0], 0x12F, 0xF, 0xF, false);
    } else {
        // shifts by one inside the 16-lane row; the group's end lanes take the
        // other end of their own 8-lane group (row_half_mirror: i <- 7 - i)
        const uint32_t sr = (uint32_t)__builtin_amdgcn_mov_dpp((int)Lp[M - 1], 0x111, 0xF, 0xF, false);
        const uint32_t mr = (uint32_t)__builtin_amdgcn_mov_dpp((int)Lp[M - 1], 0x141, 0xF, 0xF, false);
        const uint32_t sl = (uint32_t)__builtin_amdgcn_mov_dpp((int)Lp[0], 0x101, 0xF, 0xF, false);
        const uint32_t ml = (uint32_t)__builtin_amdgcn_mov_dpp((int)Lp[0], 0x141, 0xF, 0xF, false);
        A = j == 0 ? mr : sr;
        B = j == LPL - 1 ? ml : sl;
    }
    A = __builtin_amdgcn_perm(A, A, nb_sel_lo(j));
    B = __builtin_amdgcn_perm(B, B, nb_sel_hi<LPL>(j));
    // every L'(d) >= m, so t >= m and min(t, m + P2) - m = min(t - m, P2)
    const uint32_t mm = pk_splat(m), p2s = pk_splat(P2);
    // stage by stage over the M independent words, so that no packed op
    // directly follows the op producing its operand (the words without a
    // cross-lane neighbour first, while the row rotations complete)
    uint32_t t[M];
#pragma unroll
    for (int q = 0; q < M; q++) {
        const int r = M < 3 ? q : (q < M - 2 ? q + 1 : (q == M - 2 ? 0 : M - 1));
        const uint32_t dm1 = r == 0 ? A : Lp[r - 1];
        const uint32_t dp1 = r == M - 1 ? B : Lp[r + 1];
        t[r] = pk_min(dm1, dp1);
    }
#pragma unroll
    for (int r = 0; r < M; r++) t[r] = pk_adds(t[r], p1);
#pragma unroll
    for (int r = 0; r < M; r++) t[r] = pk_min(t[r], Lp[r]);
#pragma unroll
    for (int r = 0; r < M; r++) E[r] = pk_sub(t[r], mm);
#pragma unroll
    for (int r = 0; r < M; r++) E[r] = pk_min(E[r], p2s);
#pragma unroll
    for (int r = 0; r < M; r++) Lp[r] = pk_adds(E[r], c[r]);
    uint32_t nmin = Lp[0];
#pragma unroll
    for (int r = 1; r < M; r++) nmin = pk_min(nmin, Lp[r]);
    m = group_min_u32<LPL>(min(nmin & 0xFFFFu, nmin >> 16));
}

// Low bytes of the two halves of words a (-> bytes 0,1) and b (-> bytes 2,3).
__device__ __forceinline__ uint32_t pack_e(uint32_t a, uint32_t b)
{
    return __builtin_amdgcn_perm(b, a, 0x06040200u);
}

// Unpacks the two bytes of half h of word w into a packed uint16x2.
__device__ __forceinline__ uint32_t unpack_e(uint32_t w, int h)
{
    return __builtin_amdgcn_perm(0u, w, h ? 0x0C030C02u : 0x0C010C00u);
}

// Unpacks two relative cost bytes (half h of word w) into a packed
// uint16x2, the high byte of each half copied from the sign of the odd
// byte of the pair (v_perm_b32 selectors 8 and 9): values below 128 stay,
// and a pixel marked 0xFF in every byte (no valid cell) reads 0xFFFF.
__device__ __forceinline__ uint32_t unpack_rel(uint32_t w, int h)
{
    return __builtin_amdgcn_perm(w, w, h ? 0x09030902u : 0x08010800u);
}

// One wave scans lines line0 .. line0+3 of direction `dir` and writes its
// path-cost plane `s`. Called by the per-pair kernel and the batch kernel.
template <int M, bool FUSED, bool EOUT, int PF>
__device__ __forceinline__ void sgm_line_group(const SgmArgs &a, int dir, int line0, int s,
                                               int prio)
{
    constexpr int WPP = 16 * M;
    constexpr int NW = FUSED ? 1 + 3 * M : M; // prefetched words per step
    const int lane = threadIdx.x & 63;
    const int line = line0 + (lane >> 4);
    const int j = lane & 15;
    const int W = a.W, H = a.H;
    // Rows past the last line keep a one-pixel dummy line at (0, 0) so that
    // every lane runs the same, wave-uniform loop; they never store.
    int sx = 0, sy = 0, len = 0;
    if (line < dir_lines(dir, W, H)) line_geom(dir, line, W, H, sx, sy, len);
    const int dx = dir_dx(dir), dy = dir_dy(dir);
    const long step = (long)dy * W + dx;
    const long p0 = (long)sy * W + sx;
    const int klast = len > 0 ? len - 1 : 0;
    // wave-uniform trip count: the longest of the wave's 4 lines
    const int kmax = max(max(__builtin_amdgcn_readlane(len, 0), __builtin_amdgcn_readlane(len, 16)),
                         max(__builtin_amdgcn_readlane(len, 32), __builtin_amdgcn_readlane(len, 48)));
    uint8_t *outp = a.planes + (size_t)s * a.plane_bytes;
    const size_t rt_plane = (size_t)a.Wr * H;
    // Long lines set the length of the launch: let their waves issue first
    // on a shared SIMD (instruction-issue priority within the CU only).
    if (prio) {
        const int longest = max(W, H);
        if (kmax * 4 >= longest * 3) __builtin_amdgcn_s_setprio(3);
        else if (kmax * 2 >= longest) __builtin_amdgcn_s_setprio(2);
        else if (kmax * 4 >= longest) __builtin_amdgcn_s_setprio(1);
    }

    uint32_t ring[PF][NW];
    // Loads of step k (clamped to the line, so always in bounds); no branch.
    auto fetch = [&](uint32_t(&slot)[NW], int k) {
        const int kk = k < klast ? k : klast;
        const long p = p0 + kk * step;
        if constexpr (!FUSED) {
            const uint32_t *src = a.C + (size_t)p * WPP + j * M;
#pragma unroll
            for (int r = 0; r < M; r++) slot[r] = src[r];
        } else {
            const int x = sx + kk * dx, y = sy + kk * dy;
            slot[0] = a.LT[p];
            const uint32_t *rt = a.RT + (size_t)y * a.Wr + (W - 1 - x) + j * M;
#pragma unroll
            for (int t = 0; t < 3; t++)
#pragma unroll
                for (int r = 0; r < M; r++) slot[1 + t * M + r] = rt[t * rt_plane + r];
        }
    };

#pragma unroll
    for (int u = 0; u < PF; u++) fetch(ring[u], u);

    // L' = 0 and m = 0 make the first step return L = C (path start).
    uint32_t Lp[M];
#pragma unroll
    for (int r = 0; r < M; r++) Lp[r] = 0;
    uint32_t m = 0;
    const uint32_t p1 = pk_splat(a.P1);
    uint32_t padm[M];
#pragma unroll
    for (int r = 0; r < M; r++) padm[r] = cell_mask<M>(j, r, a.D, 1 << 30);
    const bool any_pad = a.D < 32 * M;

    // One step: costs from ring slot u, refill the slot, recurrence, store.
    // Every lane issues the same loads and stores (lanes past their line
    // store into a trash buffer), so the in-order memory counters stay exact
    // and the waits only cover the slot being consumed.
    auto body = [&](uint32_t(&slot)[NW], int k) {
        uint32_t c[M];
        const int kc = k < klast ? k : klast;
        if constexpr (!FUSED) {
#pragma unroll
            for (int r = 0; r < M; r++) c[r] = slot[r] | padm[r];
        } else {
            const int x = sx + kc * dx;
            uint32_t l2, lmn, lmx;
            lt_unpack(slot[0], l2, lmn, lmx);
#pragma unroll
            for (int r = 0; r < M; r++) {
                const uint32_t r2 = slot[1 + r], rmn = slot[1 + M + r], rmx = slot[1 + 2 * M + r];
                const uint32_t aa = pk_max(pk_subs(l2, rmx), pk_subs(rmn, l2));
                const uint32_t bb = pk_max(pk_subs(r2, lmx), pk_subs(lmn, r2));
                c[r] = pk_min(aa, bb);
            }
            // cells with d > x exist only near the left border
            const bool near_left = x < 32 * M - 1;
            if (__builtin_amdgcn_ballot_w64(near_left) != 0 || any_pad) {
#pragma unroll
                for (int r = 0; r < M; r++) c[r] |= cell_mask<M>(j, r, a.D, x);
            }
        }
        fetch(slot, k + PF);
        uint32_t E[M];
        sgm_step<M>(Lp, c, m, j, p1, (uint32_t)a.P2, E);
        const size_t p = (size_t)(p0 + (long)kc * step);
        if constexpr (EOUT) {
            // E in bytes: lane j holds 2M bytes at (p*32M + j*2M)
            uint8_t *dst = k < len ? outp + p * (32 * M) + j * (2 * M) : a.trash + lane * (2 * M);
            if constexpr (M % 2 == 0) {
#pragma unroll
                for (int r = 0; r < M; r += 2) ((uint32_t *)dst)[r / 2] = pack_e(E[r], E[r + 1]);
            } else {
#pragma unroll
                for (int r = 0; r < M; r++) ((uint16_t *)dst)[r] = (uint16_t)pack_e(E[r], 0);
            }
        } else {
            uint32_t *dst = k < len ? (uint32_t *)outp + p * WPP + j * M
                                    : (uint32_t *)a.trash + lane * M;
#pragma unroll
            for (int r = 0; r < M; r++) dst[r] = Lp[r];
        }
    };

    int k0 = 0;
    for (; k0 + PF <= kmax; k0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) body(ring[u], k0 + u);
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (k0 + u < kmax) body(ring[u], k0 + u);
}

// ---- buffer-addressed line scan ---------------------------------------------
//
// The same scan as sgm_line_group, addressed through raw buffer resources:
// per-lane 32-bit byte offsets advance by a constant per step (no 64-bit
// address arithmetic), loads past a line's end read harmless data or 0 (the
// hardware range check makes any offset safe), and the stores of lanes past
// their line use an offset beyond the plane, which the range check drops.
// Cells with d > x need no mask: the term planes carry 0xFFFF right terms
// for x' < 0, which give C >= 0xFE01, a cost above every valid path cost.
// Needs every plane below 2^31 bytes (buf_lines_fit).

constexpr uint32_t BUF_OOB = 0x80000000u;

// Cache-policy operand of the plane stores: non-temporal (the planes are
// read once, by a later kernel; measured: the sum over them runs ~10 %
// faster than after ordinary write-back stores).
constexpr int BUF_STORE_NT = 2;

// N consecutive dwords to byte offset off (dropped outside the buffer).
template <int N>
__device__ __forceinline__ void buf_store(buf_t r, uint32_t off, const uint32_t *src)
{
    if constexpr (N == 4) {
        __builtin_amdgcn_raw_buffer_store_b128(tdvh_u32x4{src[0], src[1], src[2], src[3]}, r,
                                               off, 0, BUF_STORE_NT);
    } else if constexpr (N == 3) {
        __builtin_amdgcn_raw_buffer_store_b96(tdvh_u32x3{src[0], src[1], src[2]}, r, off, 0,
                                              BUF_STORE_NT);
    } else if constexpr (N == 2) {
        __builtin_amdgcn_raw_buffer_store_b64(tdvh_u32x2{src[0], src[1]}, r, off, 0,
                                              BUF_STORE_NT);
    } else if constexpr (N == 1) {
        __builtin_amdgcn_raw_buffer_store_b32(src[0], r, off, 0, BUF_STORE_NT);
    }
}

// LPL lanes per line: 16 (4 lines per wave, M words per lane), or 8 (8 lines
// per wave, 2M words per lane; the planes and term rows are the same, only
// the lanes owning each word differ). A smaller word count per lane leaves
// the step's fixed costs (loads, stores, left terms, the group minimum)
// spread over few cells; 8-lane lines halve the waves for the same work.
// SRC: where the step's costs come from (SRC_WORDS, SRC_TERMS, SRC_REL).
// HZ (BT terms, 16-lane lines): 1 for direction (1, 0), 2 for (-1, 0). Along
// a horizontal line the right-term words of step k + 1 are those of step k
// moved by one word (lane j's word r is RT word W-1-x + j*M + r of the
// line's row), so each plane's M words per lane shift by one register, one
// DPP row shift carries the boundary word between lanes, and only the word
// entering at the end lane is loaded: 3 dwords per lane and step instead of
// 3M (the loads, not the arithmetic, bound the scans with BT recomputed).
template <int M, int SRC, bool PAD, int PF, int LPL = 16, int HZ = 0>
__device__ __forceinline__ void sgm_line_group_buf(const SgmArgs &a, int dir, int line0, int s,
                                                   int prio)
{
    constexpr bool FUSED = SRC == SRC_TERMS;
    constexpr bool REL = SRC == SRC_REL;
    constexpr int MW = M * 16 / LPL; // words per lane and plane
    static_assert(!REL || MW == 1 || MW % 2 == 0, "relative cost bytes: 1 or an even word count");
    static_assert(HZ == 0 || (FUSED && LPL == 16), "sliding right terms: BT terms, 16-lane lines");
    // prefetched words per step
    constexpr int NW = FUSED ? (HZ ? 4 : 1 + 3 * MW) : (REL ? (MW + 1) / 2 : MW);
    const int lane = threadIdx.x & 63;
    const int line = line0 + lane / LPL;
    const int j = lane % LPL;
    const int W = a.W, H = a.H;
    int sx = 0, sy = 0, len = 0;
    if (line < dir_lines(dir, W, H)) line_geom(dir, line, W, H, sx, sy, len);
    const int dx = dir_dx(dir), dy = dir_dy(dir);
    const int step = dy * W + dx, p0 = sy * W + sx;
    int kmax = 0;
#pragma unroll
    for (int g = 0; g < 64 / LPL; g++) kmax = max(kmax, __builtin_amdgcn_readlane(len, g * LPL));
    if (prio) {
        const int longest = max(W, H);
        if (kmax * 4 >= longest * 3) __builtin_amdgcn_s_setprio(3);
        else if (kmax * 2 >= longest) __builtin_amdgcn_s_setprio(2);
        else if (kmax * 4 >= longest) __builtin_amdgcn_s_setprio(1);
    }

    const buf_t out = buf_rsrc(a.planes + (size_t)s * a.plane_bytes, (uint32_t)a.plane_bytes);
    uint32_t oE = (uint32_t)p0 * (32 * M) + j * (2 * MW);
    const uint32_t dE = (uint32_t)(step * (32 * M));

    buf_t src[FUSED ? 4 : 1]; // packed left terms, then the three right-term planes
    uint32_t fa = 0, fb = 0, da = 0, db = 0; // fetch offsets and their steps
    if constexpr (FUSED) {
        const uint32_t rt_plane = (uint32_t)a.Wr * H;
        src[0] = buf_rsrc(a.LT, (uint32_t)W * H * 4);
#pragma unroll
        for (int t = 0; t < 3; t++) src[1 + t] = buf_rsrc(a.RT + (size_t)t * rt_plane, rt_plane * 4);
        fa = (uint32_t)p0 * 4;
        da = (uint32_t)(step * 4);
        fb = (uint32_t)(sy * a.Wr + (W - 1 - sx) + j * MW) * 4;
        db = (uint32_t)((dy * a.Wr - dx) * 4);
    } else if constexpr (REL) {
        // 2 bytes per word, in the byte order of the E planes
        src[0] = buf_rsrc(a.CR, (uint32_t)W * H * (32 * M));
        fa = (uint32_t)p0 * (32 * M) + j * (2 * MW);
        da = (uint32_t)(step * (32 * M));
    } else {
        src[0] = buf_rsrc(a.C, (uint32_t)W * H * (64 * M));
        fa = (uint32_t)p0 * (64 * M) + j * (4 * MW);
        da = (uint32_t)(step * (64 * M));
    }
    // HZ: the right-term words of step 0, and the offset of the word entering
    // at the end lane (lane 0's word 0 moving right, lane 15's word M - 1
    // moving left; the same address for the line's 16 lanes)
    uint32_t rt[3][HZ ? MW : 1];
    uint32_t fn = 0;
    if constexpr (HZ != 0) {
#pragma unroll
        for (int t = 0; t < 3; t++) buf_load<MW>(src[1 + t], fb, rt[t]);
        fn = (uint32_t)(sy * a.Wr + (W - 1 - sx) + (HZ == 2 ? 16 * MW - 1 : 0)) * 4;
    }

    uint32_t ring[PF][NW];
    auto fetch = [&](uint32_t(&slot)[NW]) {
        if constexpr (FUSED && HZ != 0) {
            slot[0] = __builtin_amdgcn_raw_buffer_load_b32(src[0], fa, 0, 0);
#pragma unroll
            for (int t = 0; t < 3; t++) slot[1 + t] = __builtin_amdgcn_raw_buffer_load_b32(src[1 + t], fn, 0, 0);
            fn += db;
        } else if constexpr (FUSED) {
            slot[0] = __builtin_amdgcn_raw_buffer_load_b32(src[0], fa, 0, 0);
#pragma unroll
            for (int t = 0; t < 3; t++) buf_load<MW>(src[1 + t], fb, slot + 1 + t * MW);
            fb += db;
        } else if constexpr (REL && MW == 1) {
            slot[0] = __builtin_amdgcn_raw_buffer_load_b16(src[0], fa, 0, 0);
        } else {
            buf_load<NW>(src[0], fa, slot);
        }
        fa += da;
    };
#pragma unroll
    for (int u = 0; u < PF; u++) fetch(ring[u]);

    // L' = 0 and m = 0 make the first step return L = C (path start).
    uint32_t Lp[MW];
#pragma unroll
    for (int r = 0; r < MW; r++) Lp[r] = 0;
    uint32_t m = 0;
    const uint32_t p1 = pk_splat(a.P1);
    uint32_t padm[MW];
#pragma unroll
    for (int r = 0; r < MW; r++) padm[r] = cell_mask<MW, LPL>(j, r, a.D, 1 << 30);

    // Costs of one step from a ring slot (HZ: shift first the right-term
    // words to the step, the slot holding the entering words; not for step 0).
    auto costs = [&](const uint32_t(&slot)[NW], uint32_t(&c)[MW], bool shift) {
        if constexpr (FUSED) {
            uint32_t l2, lmn, lmx;
            lt_unpack(slot[0], l2, lmn, lmx);
            uint32_t rw[3][MW]; // the step's right-term words per plane
            if constexpr (HZ != 0) {
                if (shift) {
#pragma unroll
                    for (int t = 0; t < 3; t++) {
                        if constexpr (HZ == 1) {
                            // lane j's word 0 <- lane j - 1's word M - 1; lane 0 <- the new word
                            const uint32_t b = (uint32_t)__builtin_amdgcn_update_dpp(
                                (int)slot[1 + t], (int)rt[t][MW - 1], 0x111, 0xF, 0xF, false);
#pragma unroll
                            for (int r = MW - 1; r > 0; r--) rt[t][r] = rt[t][r - 1];
                            rt[t][0] = b;
                        } else {
                            // lane j's word M - 1 <- lane j + 1's word 0; lane 15 <- the new word
                            const uint32_t b = (uint32_t)__builtin_amdgcn_update_dpp(
                                (int)slot[1 + t], (int)rt[t][0], 0x101, 0xF, 0xF, false);
#pragma unroll
                            for (int r = 0; r < MW - 1; r++) rt[t][r] = rt[t][r + 1];
                            rt[t][MW - 1] = b;
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < 3; t++)
#pragma unroll
                    for (int r = 0; r < MW; r++) rw[t][r] = rt[t][r];
            } else {
#pragma unroll
                for (int t = 0; t < 3; t++)
#pragma unroll
                    for (int r = 0; r < MW; r++) rw[t][r] = slot[1 + t * MW + r];
            }
            // stage by stage over the words (independent ops between each
            // packed op and its operands' producers)
            uint32_t s1[MW], s2[MW], s3[MW], s4[MW];
#pragma unroll
            for (int r = 0; r < MW; r++) s1[r] = pk_subs(l2, rw[2][r]);
#pragma unroll
            for (int r = 0; r < MW; r++) s2[r] = pk_subs(rw[1][r], l2);
#pragma unroll
            for (int r = 0; r < MW; r++) s3[r] = pk_subs(rw[0][r], lmx);
#pragma unroll
            for (int r = 0; r < MW; r++) s4[r] = pk_subs(lmn, rw[0][r]);
#pragma unroll
            for (int r = 0; r < MW; r++) s1[r] = pk_max(s1[r], s2[r]);
#pragma unroll
            for (int r = 0; r < MW; r++) s3[r] = pk_max(s3[r], s4[r]);
#pragma unroll
            for (int r = 0; r < MW; r++) c[r] = pk_min(s1[r], s3[r]);
        } else if constexpr (REL) {
#pragma unroll
            for (int r = 0; r < MW; r++) c[r] = unpack_rel(slot[r / 2], r % 2);
        } else {
#pragma unroll
            for (int r = 0; r < MW; r++) c[r] = slot[r];
        }
        if constexpr (PAD) {
#pragma unroll
            for (int r = 0; r < MW; r++) c[r] |= padm[r];
        }
    };

    // Software pipeline: c holds the costs of step k while slot (u+1) % PF
    // holds the data of step k+1, fetched PF steps earlier. A body runs the
    // recurrence of step k and the cost computation of step k+1 (independent
    // work the scheduler interleaves), then refills that slot. The scheduling
    // barrier keeps each body's waits on its own slot, so the compiler cannot
    // hoist later steps' cost math (and their waits) above earlier steps.
    uint32_t c[MW];
    costs(ring[0], c, false);
    fetch(ring[0]);
    auto body = [&](int u, int k) {
        uint32_t E[MW];
        sgm_step<MW, LPL>(Lp, c, m, j, p1, (uint32_t)a.P2, E);
        const uint32_t so = k < len ? oE : BUF_OOB;
        oE += dE;
        if constexpr (MW % 2 == 0) {
            uint32_t w[MW / 2];
#pragma unroll
            for (int r = 0; r < MW; r += 2) w[r / 2] = pack_e(E[r], E[r + 1]);
            if constexpr (MW == 8) {
                buf_store<4>(out, so, w);
            } else {
                buf_store<MW / 2>(out, so, w);
            }
        } else {
#pragma unroll
            for (int r = 0; r < MW; r++)
                __builtin_amdgcn_raw_buffer_store_b16((unsigned short)pack_e(E[r], 0u), out,
                                                      so + 2 * r, 0, BUF_STORE_NT);
        }
        costs(ring[(u + 1) % PF], c, true);
        // refill only after the slot's words are consumed, so the loads
        // reuse its registers instead of needing a second copy of the ring
        __builtin_amdgcn_sched_barrier(0);
        fetch(ring[(u + 1) % PF]);
        __builtin_amdgcn_sched_barrier(0);
    };

    int k0 = 0;
    for (; k0 + PF <= kmax; k0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; u++) body(u, k0 + u);
    }
#pragma unroll
    for (int u = 0; u < PF; u++)
        if (k0 + u < kmax) body(u, k0 + u);
}

// Prefetch depth of the buffer line kernel, the deepest ring that keeps two
// waves per SIMD without spills: fused, 8 steps up to 4 words per lane and
// plane, 6 at 6 words, 4 at 8; staged (M words per step), 8. Measured on
// 1280x720: a deeper ring with two waves per SIMD beats a shallower one with
// three or four waves (the longest lines, the launch's critical path, then
// share their SIMD's issue slots with fewer waves).
template <int M, bool FUSED> constexpr int buf_lines_pf()
{
    return !FUSED || M <= 4 ? 8 : M == 6 ? 6 : 4;
}

// A line group of direction `dir`, with the horizontal directions' sliding
// right terms where they pay: BT terms, 16-lane lines, 5 or 6 words per lane
// (measured at 1280x720: scans 7 % faster at D = 160, 10 % at D = 192; even
// or up to 3 % slower at the other word counts, where the loads of the
// other directions and the horizontal lines' instruction issue set the time).
template <int M, int SRC, bool PAD, int PF, int LPL>
__device__ __forceinline__ void sgm_line_group_any(const SgmArgs &a, int dir, int line0, int s,
                                                   int prio)
{
    if constexpr (SRC == SRC_TERMS && LPL == 16 && (M == 5 || M == 6)) {
        if (dir == 0) {
            sgm_line_group_buf<M, SRC, PAD, PF, LPL, 1>(a, dir, line0, s, prio);
            return;
        }
        if (dir == 1) {
            sgm_line_group_buf<M, SRC, PAD, PF, LPL, 2>(a, dir, line0, s, prio);
            return;
        }
    }
    sgm_line_group_buf<M, SRC, PAD, PF, LPL>(a, dir, line0, s, prio);
}

// One wave per workgroup, so the dispatcher starts the next line group as
// soon as any wave ends; `order` gives workgroup b its group (entry: plane
// slot << 28 | direction << 24 | first line; 64 / LPL lines from there),
// longest first within each chiplet's share.
template <int M, int SRC, bool PAD, int LPL>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void
sgm_lines_buf_kernel(const SgmArgs a, const uint32_t *__restrict__ order, int prio)
{
    const uint32_t e = order[blockIdx.x];
    sgm_line_group_any<M, SRC, PAD, buf_lines_pf<M, SRC == SRC_TERMS>(), LPL>(
        a, (int)((e >> 24) & 0xF), (int)(e & 0xFFFFFF), (int)(e >> 28), prio);
}

// The line scans over the relative cost bytes: a few words per step, so a
// deeper prefetch ring than the BT scans (16 steps) fits the registers.
template <int M, int LPL>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void
sgm_lines_rel_kernel(const SgmArgs a, const uint32_t *__restrict__ order, int prio)
{
    const uint32_t e = order[blockIdx.x];
    sgm_line_group_buf<M, SRC_REL, false, 16, LPL>(a, (int)((e >> 24) & 0xF),
                                                   (int)(e & 0xFFFFFF), (int)(e >> 28), prio);
}

// Whether sgm_line_group_buf can address every plane of a (byte planes).
__host__ __device__ inline bool buf_lines_fit(const SgmArgs &a)
{
    const size_t lim = (size_t)1 << 31;
    if (a.plane_bytes >= lim) return false;
    if (a.CR) return (size_t)a.W * a.H * 32 * (size_t)a.m < lim;
    if (a.C) return (size_t)a.W * a.H * 64 * (size_t)a.m < lim;
    return (size_t)a.W * a.H * 4 < lim && (size_t)a.Wr * a.H * 4 < lim;
}

template <int M, bool FUSED, bool EOUT, int PF>
__global__ __launch_bounds__(256) void sgm_lines_kernel(const SgmArgs a, const LineSet ls)
{
    const int gw = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gw >= ls.wave_begin[ls.ndirs]) return;
    int s = 0;
    while (s < ls.ndirs - 1 && gw >= ls.wave_begin[s + 1]) s++;
    sgm_line_group<M, FUSED, EOUT, PF>(a, ls.dir[s], (gw - ls.wave_begin[s]) * 4, s, ls.prio);
}

// ---- BT term planes for the fused path ---------------------------------------

// LT[y][x]: packed left terms (lt_unpack); RT[t][y][i]: right terms of
// pixels W-1-i (low half) and W-1-i-16m (high half), clamped at x' = 0.
__device__ __forceinline__ void bt_planes_at(const uint8_t *__restrict__ L,
                                             const uint8_t *__restrict__ R, int W, int H,
                                             int pitch, int m, uint32_t *__restrict__ LT,
                                             uint32_t *__restrict__ RT, int Wr, int i, int y)
{
    const uint8_t *lr = L + (size_t)y * pitch;
    const uint8_t *rr = R + (size_t)y * pitch;
    // x < 0 (cells with d > x): all three terms 0xFFFF, so the packed BT
    // formula gives a cost of at least 0xFE01 there
    auto terms = [&](const uint8_t *row, int x, int &v2, int &mn, int &mx) {
        if (x < 0) {
            v2 = mn = mx = 0xFFFF;
            return;
        }
        const int c = row[x];
        const int a = c + row[x > 0 ? x - 1 : 0], b = c + row[x < W - 1 ? x + 1 : W - 1];
        v2 = 2 * c;
        mn = min(v2, min(a, b));
        mx = max(v2, max(a, b));
    };
    if (i < W) {
        int v2, mn, mx;
        terms(lr, i, v2, mn, mx);
        LT[(size_t)y * W + i] =
            (uint32_t)(v2 >> 1) | (uint32_t)(v2 - mn) << 8 | (uint32_t)(mx - v2) << 16;
    }
    if (i < Wr) {
        int a2, amn, amx, b2, bmn, bmx;
        const int xa = W - 1 - i;
        terms(rr, xa, a2, amn, amx);
        terms(rr, xa - 16 * m, b2, bmn, bmx);
        const size_t o = (size_t)y * Wr + i, pl = (size_t)Wr * H;
        RT[o] = (uint32_t)a2 | ((uint32_t)b2 << 16);
        RT[pl + o] = (uint32_t)amn | ((uint32_t)bmn << 16);
        RT[2 * pl + o] = (uint32_t)amx | ((uint32_t)bmx << 16);
    }
}


// ---- plane sum + lowest-index WTA ----------------------------------------

// Sum of the planes and lowest-index argmin of pixel p by one 16-lane row.
template <int M, bool FUSED, bool EOUT>
__device__ __forceinline__ void sgm_sum_pixel(const SgmArgs &a, int nplanes,
                                              uint32_t *__restrict__ S,
                                              int16_t *__restrict__ disp, int pitch, long p)
{
    constexpr int WPP = 16 * M;
    const int W = a.W, H = a.H;
    const int j = threadIdx.x & 15;
    const int y = (int)(p / W), x = (int)(p - (long)y * W);
    uint32_t acc[M];
    if constexpr (EOUT) {
        // S = n*C + sum of E (invalid where C is)
        uint32_t c[M];
        if constexpr (FUSED) {
            const size_t rt_plane = (size_t)a.Wr * H;
            uint32_t lt[3];
            lt_unpack(a.LT[p], lt[0], lt[1], lt[2]);
            const uint32_t *rt = a.RT + (size_t)y * a.Wr + (W - 1 - x) + j * M;
            uint32_t r2[M], rmn[M], rmx[M];
#pragma unroll
            for (int r = 0; r < M; r++) {
                r2[r] = rt[r];
                rmn[r] = rt[rt_plane + r];
                rmx[r] = rt[2 * rt_plane + r];
            }
            bt_costs<M>(lt, r2, rmn, rmx, j, a.D, x, c);
        } else {
            const uint32_t *src = a.C + (size_t)p * WPP + j * M;
#pragma unroll
            for (int r = 0; r < M; r++) c[r] = src[r] | cell_mask<M>(j, r, a.D, 1 << 30);
        }
#pragma unroll
        for (int r = 0; r < M; r++) acc[r] = 0;
        for (int q = 0; q < nplanes; q++) {
            const uint8_t *src = a.planes + q * a.plane_bytes + (size_t)p * (32 * M) + j * (2 * M);
            if constexpr (M % 2 == 0) {
#pragma unroll
                for (int r = 0; r < M; r += 2) {
                    const uint32_t w = ((const uint32_t *)src)[r / 2];
                    acc[r] = pk_add(acc[r], unpack_e(w, 0));
                    acc[r + 1] = pk_add(acc[r + 1], unpack_e(w, 1));
                }
            } else {
#pragma unroll
                for (int r = 0; r < M; r++)
                    acc[r] = pk_add(acc[r], unpack_e(((const uint16_t *)src)[r], 0));
            }
        }
        const uint32_t n = pk_splat((uint32_t)nplanes);
#pragma unroll
        for (int r = 0; r < M; r++) {
            const uint32_t inv = pk_sub(0u, u32(v2(c[r]) >> (tdvh_u16x2)(15))); // 0xFFFF where C invalid
            acc[r] = pk_add(u32(v2(c[r]) * v2(n)), acc[r]) | inv;
        }
    } else {
        const size_t base = (size_t)p * WPP + j * M;
        const uint32_t *L = (const uint32_t *)a.planes;
        const size_t plane = a.plane_bytes / 4;
#pragma unroll
        for (int r = 0; r < M; r++) acc[r] = L[base + r];
        for (int q = 1; q < nplanes; q++) {
            const uint32_t *src = L + q * plane + base;
#pragma unroll
            for (int r = 0; r < M; r++) acc[r] = pk_adds(acc[r], src[r]);
        }
    }
    if (S) {
#pragma unroll
        for (int r = 0; r < M; r++) S[(size_t)p * WPP + j * M + r] = acc[r];
    }
    uint32_t key = 0xFFFFFFFFu;
#pragma unroll
    for (int r = 0; r < M; r++) {
        const uint32_t dlo = j * M + r, dhi = 16 * M + j * M + r;
        key = min(key, ((acc[r] & 0xFFFFu) << 16) | dlo);
        key = min(key, (acc[r] & 0xFFFF0000u) | dhi);
    }
    key = row_min_u32(key);
    if (j == 0) disp[(size_t)y * pitch + x] = (int16_t)(key & 0xFFFFu);
}


// Word counts with a wide plane sum, and the lanes per pixel of that sum:
// 2M (a DPP-aligned power of two), or 16 for M = 6 (12 working lanes).
template <int M> constexpr bool sum_wide_ok() { return M == 1 || M == 2 || M == 4 || M == 6 || M == 8; }
template <int M> constexpr int sum_group() { return M == 6 ? 16 : 2 * M; }

// Plane sum + lowest-index WTA with 16-byte lanes: lane t < 2M of a pixel's
// group of G lanes owns SPLIT words w = 8t .. 8t+7 (cells d = w in the
// low halves, d = 16M + w in the high halves), i.e. bytes 16t .. 16t+15 of
// the pixel in every plane, so each plane is read in whole 1 KiB wave rows.
// S = n*C + sum of E (invalid where C is), then the key (S << 9 | d)
// reduced over the group (S <= 4336 and d < 512 keep keys ordered).
template <int M, bool FUSED>
__device__ __forceinline__ void sgm_sum_pixel_wide(const SgmArgs &a, int nplanes,
                                                   uint32_t *__restrict__ S,
                                                   int16_t *__restrict__ disp, int pitch, long p,
                                                   int t)
{
    static_assert(sum_wide_ok<M>(), "word counts 1, 2, 4, 6, 8");
    constexpr int G = sum_group<M>();
    const int W = a.W, H = a.H;
    const int y = (int)(p / W), x = (int)(p - (long)y * W);
    const int w0 = 8 * t;
    uint32_t key = 0xFFFFFFFFu;
    if (G == 2 * M || t < 2 * M) {
        uint32_t c[8];
        if constexpr (FUSED) {
            const size_t rt_plane = (size_t)a.Wr * H;
            uint32_t l2, lmn, lmx;
            lt_unpack(a.LT[p], l2, lmn, lmx);
            const uint32_t *rt = a.RT + (size_t)y * a.Wr + (W - 1 - x) + w0;
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const uint32_t r2 = rt[r], rmn = rt[rt_plane + r], rmx = rt[2 * rt_plane + r];
                const uint32_t aa = pk_max(pk_subs(l2, rmx), pk_subs(rmn, l2));
                const uint32_t bb = pk_max(pk_subs(r2, lmx), pk_subs(lmn, r2));
                c[r] = pk_min(aa, bb);
            }
        } else {
            const uint32_t *src = a.C + (size_t)p * (16 * M) + w0;
#pragma unroll
            for (int r = 0; r < 8; r++) c[r] = src[r];
        }
        if (a.D < 32 * M) {
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int w = w0 + r;
                c[r] |= (w >= a.D ? 0x0000FFFFu : 0u) | (16 * M + w >= a.D ? 0xFFFF0000u : 0u);
            }
        }
        // E bytes are summed four at a time in 32-bit adds while no byte can
        // carry (g planes with g * P2 <= 255), then widened into uint16 pairs.
        uint32_t acc[8];
#pragma unroll
        for (int r = 0; r < 8; r++) acc[r] = 0;
        const uint8_t *e = a.planes + (size_t)p * (32 * M) + 16 * t;
        tdvh_u32x4 ev[8];
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (q < nplanes)
                ev[q] = __builtin_nontemporal_load((const tdvh_u32x4 *)(e + q * a.plane_bytes));
        const int g = a.P2 > 0 ? min(255 / a.P2, 8) : 8;
        uint32_t bs[4] = {0, 0, 0, 0};
        int cnt = 0;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (q < nplanes) {
                bs[0] += ev[q].x, bs[1] += ev[q].y, bs[2] += ev[q].z, bs[3] += ev[q].w;
                if (++cnt == g || q == nplanes - 1) {
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        acc[2 * i] = pk_add(acc[2 * i], unpack_e(bs[i], 0));
                        acc[2 * i + 1] = pk_add(acc[2 * i + 1], unpack_e(bs[i], 1));
                        bs[i] = 0;
                    }
                    cnt = 0;
                }
            }
        }
        const uint32_t n = pk_splat((uint32_t)nplanes);
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const uint32_t inv = pk_sub(0u, u32(v2(c[r]) >> (tdvh_u16x2)(15))); // 0xFFFF where C invalid
            acc[r] = pk_add(u32(v2(c[r]) * v2(n)), acc[r]) | inv;
            const uint32_t w = (uint32_t)(w0 + r);
            key = min(key, ((acc[r] & 0xFFFFu) << 9) | w);
            key = min(key, ((acc[r] >> 16) << 9) | (16 * M + w));
        }
        if (S) {
            uint4 *dst = (uint4 *)(S + (size_t)p * (16 * M) + w0);
            dst[0] = make_uint4(acc[0], acc[1], acc[2], acc[3]);
            dst[1] = make_uint4(acc[4], acc[5], acc[6], acc[7]);
        }
    }
    key = group_min_u32<G>(key);
    if (t == 0) disp[(size_t)y * pitch + x] = (int16_t)(key & 0x1FFu);
}

template <int M, bool FUSED>
__global__ __launch_bounds__(256) void sgm_sum_wide_kernel(const SgmArgs a, int nplanes,
                                                           uint32_t *__restrict__ S,
                                                           int16_t *__restrict__ disp, int pitch)
{
    constexpr int G = sum_group<M>();
    const long p = ((long)blockIdx.x * 256 + threadIdx.x) / G;
    if (p >= (long)a.W * a.H) return; // whole groups leave together (256 % G == 0)
    sgm_sum_pixel_wide<M, FUSED>(a, nplanes, S, disp, pitch, p, threadIdx.x % G);
}

// ---- relative cost bytes ---------------------------------------------------
//
// The line scans need a pixel's costs only up to a constant per pixel and
// only below the pixel's minimum path cost plus P2:
//  * the step takes L' - m, so adding a constant to a pixel's whole vector
//    changes no increment E = L - C downstream;
//  * a cell with L'(d) >= m + P2 never changes min(L'(d), L'(d+-1) + P1,
//    m + P2), so any value >= m + P2 there gives the same E;
//  * m <= Cmin + P2, Cmin being the pixel's least cost (the cell holding it
//    has L = Cmin + E with E <= P2), and the cell holding m has C <= m.
// So CR(d) = min(C(d) - Cmin, 2 P2) gives, in every direction, exactly the
// increments E of the true costs: cells with C > Cmin + 2 P2 stay at or
// above m + P2 with either cost, every other cell is shifted by Cmin, and
// invalid cells (0xFFFF, or >= 0xFE01 from the term planes' x < 0 words)
// become 2 P2, another "at or above m + P2". A pixel with no valid cell
// (possible in a given volume) is written as 0xFF bytes, which the scans
// read as 0xFFFF, so the next step restarts the path as with the true
// costs. For P2 <= 63 every other byte is below 128 (unpack_rel). One byte
// per cell in the byte order of the E planes, read by the scans instead of
// recomputing BT; the plane sum still adds n*C with the true C.

template <int M> constexpr bool rel_ok() { return M == 1 || M == 2 || M == 4 || M == 6 || M == 8; }

// Relative cost bytes of pixel p; lane t < 2M of the pixel's group of G
// lanes owns SPLIT words 8t .. 8t+7, i.e. bytes 16t .. 16t+15 of the pixel.
template <int M, bool FUSED>
__device__ __forceinline__ void sgm_rel_pixel(const SgmArgs &a, long p, int t)
{
    static_assert(rel_ok<M>(), "word counts 1, 2, 4, 6, 8");
    constexpr int G = sum_group<M>();
    const int W = a.W, H = a.H;
    const int y = (int)(p / W), x = (int)(p - (long)y * W);
    const int w0 = 8 * t;
    const bool act = G == 2 * M || t < 2 * M;
    uint32_t c[8];
    uint32_t cmin = 0xFFFFu;
    if (act) {
        if constexpr (FUSED) {
            const size_t rt_plane = (size_t)a.Wr * H;
            uint32_t l2, lmn, lmx;
            lt_unpack(a.LT[p], l2, lmn, lmx);
            const uint32_t *rt = a.RT + (size_t)y * a.Wr + (W - 1 - x) + w0;
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const uint32_t r2 = rt[r], rmn = rt[rt_plane + r], rmx = rt[2 * rt_plane + r];
                const uint32_t aa = pk_max(pk_subs(l2, rmx), pk_subs(rmn, l2));
                const uint32_t bb = pk_max(pk_subs(r2, lmx), pk_subs(lmn, r2));
                c[r] = pk_min(aa, bb);
            }
        } else {
            const uint4 *src = (const uint4 *)(a.C + (size_t)p * (16 * M) + w0);
            const uint4 u0 = src[0], u1 = src[1];
            c[0] = u0.x, c[1] = u0.y, c[2] = u0.z, c[3] = u0.w;
            c[4] = u1.x, c[5] = u1.y, c[6] = u1.z, c[7] = u1.w;
        }
        if (a.D < 32 * M) {
#pragma unroll
            for (int r = 0; r < 8; r++) {
                const int w = w0 + r;
                c[r] |= (w >= a.D ? 0x0000FFFFu : 0u) | (16 * M + w >= a.D ? 0xFFFF0000u : 0u);
            }
        }
        uint32_t mw = c[0];
#pragma unroll
        for (int r = 1; r < 8; r++) mw = pk_min(mw, c[r]);
        cmin = min(mw & 0xFFFFu, mw >> 16);
    }
    cmin = group_min_u32<G>(cmin);
    if (act) {
        // a pixel with no valid cell (cmin 0xFFFF) is marked 0xFF in every byte
        const uint32_t cm = pk_splat(cmin), cap = pk_splat(2 * (uint32_t)a.P2);
        const uint32_t none = cmin == 0xFFFFu ? 0xFFFFFFFFu : 0u;
        uint32_t h[8];
#pragma unroll
        for (int r = 0; r < 8; r++) h[r] = pk_min(pk_subs(c[r], cm), cap) | none;
        uint4 *dst = (uint4 *)(a.CR + (size_t)p * (32 * M) + 16 * t);
        *dst = make_uint4(pack_e(h[0], h[1]), pack_e(h[2], h[3]), pack_e(h[4], h[5]),
                          pack_e(h[6], h[7]));
    }
}

template <int M, bool FUSED>
__global__ __launch_bounds__(256) void sgm_rel_kernel(const SgmArgs a)
{
    constexpr int G = sum_group<M>();
    const long p = ((long)blockIdx.x * 256 + threadIdx.x) / G;
    if (p >= (long)a.W * a.H) return; // whole groups leave together (256 % G == 0)
    sgm_rel_pixel<M, FUSED>(a, p, threadIdx.x % G);
}

template <int M, bool FUSED, bool EOUT>
__global__ __launch_bounds__(256) void sgm_sum_wta_kernel(const SgmArgs a, int nplanes,
                                                          uint32_t *__restrict__ S,
                                                          int16_t *__restrict__ disp, int pitch)
{
    const long p = (long)blockIdx.x * 16 + (threadIdx.x >> 4);
    if (p >= (long)a.W * a.H) return;
    sgm_sum_pixel<M, FUSED, EOUT>(a, nplanes, S, disp, pitch, p);
}

} // namespace tdvh
