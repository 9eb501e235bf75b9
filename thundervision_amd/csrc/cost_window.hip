// Window matching costs: SSD (ssd.cu:15-50), NCC (crosscorrelation.cu:11-49),
// the fused SSD + lowest-index WTA of fastwta.cu:32-53, and the same fusion
// for NCC (no counterpart in the reference; its disparities are those of the
// NCC volume followed by the lowest-index WTA of wtacpu.cpp:20-33).
//
// Mapping: one wave per tile of WIN_TX consecutive x and WIN_TY output rows;
// lane i holds disparity c*64 + i of every chunk c, so the 64 lanes write 64
// consecutive d of a pixel (coalesced, d-contiguous layout). The box sum is
// separable and slides both ways: per chunk each lane keeps, for the tile's
// WIN_TX + 2r columns, the vertical window sum of the pixel term
//   SSD: (L(x', y') - R(x' - d, y'))^2     NCC: L(x', y') * R(x' - d, y')
// over rows y-r .. y+r; one output row adds the entering row's terms and
// subtracts the leaving row's, then a running sum along the columns yields
// the tile's WIN_TX window sums. A cell costs about 2 (WIN_TX + 2r) / WIN_TX
// pixel terms instead of (2r+1)^2. Image rows (clamped to the edge, as the
// reference's textures) sit in an LDS ring of 2r+2 rows, one word per pixel
// so that neighbouring columns load in pairs; the next row is fetched into
// registers while the current one is processed.
//
// Arithmetic is exact integer for SSD (uint32) and for every NCC window sum;
// NCC combines them as
//   zero-mean: rho = (N*Slr - Sl*Sr) / (sqrt(N*Sll - Sl^2) * sqrt(N*Srr - Sr^2))
//   plain:     rho = Slr / (sqrt(Sll) * sqrt(Srr))
// (numerators exact in int32: |N*Slr|, Sl*Sr <= (121*255)^2 < 2^31), with
// the per-pixel reciprocal square roots IEEE-rounded (root, then quotient)
// once per pixel, their product IEEE-rounded, and cost = 1 - num * product in
// one fused multiply-add; a zero variance or x - d < 0 gives +inf (their
// reciprocal roots are staged as quiet NaN, which the product and the
// multiply-add carry, and one IEEE minNum against +inf maps to +inf).
#include "kernels.hpp"

namespace tdvh {

enum { WIN_SSD = 0, WIN_SSD_WTA = 1, WIN_NCC = 2, WIN_NCC_WTA = 3 };
constexpr uint32_t NCC_NAN = 0x7FC00000u; // quiet NaN: reciprocal root of an invalid window

// Key of a float cost in the lowest-index WTA's order: +inf (invalid) and
// padding above every finite cost, finite costs ordered through their bits
// (NCC costs are never -0 or NaN).
__device__ __forceinline__ uint32_t ncc_key(float v)
{
    if (!(v < __builtin_inff())) return 0xFFFFFFFFu;
    const uint32_t b = __builtin_bit_cast(uint32_t, v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
constexpr int WIN_TX = 16; // output columns per wave
constexpr int WIN_TY = 32; // output rows per wave

struct WinArgs {
    const uint8_t *L, *R;
    int W, H, pitch, D;
    int d0;             // first disparity of this pass (volumes above 256 run in passes)
    int64_t sx, sy;     // output strides (elements)
    void *out;          // uint32 / float volume, or int16 disparities (WTA)
    int disp_pitch;
    const int32_t *box; // NCC: 4 planes W*H: sum L, 1/sqrt(var L) (f32), sum R, 1/sqrt(var R)
                        // (f32); quiet NaN for a zero variance
    int zero_mean;
};

// one halving step of wave_min16 (lane partner l ^ B within 32 lanes)
template <int N, int B>
__device__ __forceinline__ void min_halve(uint32_t (&v)[2 * N], uint32_t (&o)[N], int lane)
{
    const bool up = lane & B; // keeps the upper half
#pragma unroll
    for (int j = 0; j < N; j++) {
        const uint32_t send = up ? v[j] : v[j + N];
        const uint32_t keep = up ? v[j + N] : v[j];
        const uint32_t got =
            (uint32_t)__builtin_amdgcn_ds_swizzle((int)send, (B << 10) | 0x1F); // lane ^ B
        o[j] = min(keep, got);
    }
}

// Minimum over the wave of each of 16 per-lane keys; lane l returns the
// minimum of key[(l >> 1) & 15].
__device__ __forceinline__ uint32_t wave_min16(uint32_t (&key)[16], int lane)
{
    uint32_t k8[8], k4[4], k2[2], k1[1];
    min_halve<8, 16>(key, k8, lane);
    min_halve<4, 8>(k8, k4, lane);
    min_halve<2, 4>(k4, k2, lane);
    min_halve<1, 2>(k2, k1, lane);
    uint32_t k = min(k1[0], (uint32_t)__builtin_amdgcn_ds_swizzle((int)k1[0], (1 << 10) | 0x1F));
    return min(k, (uint32_t)__builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, (int)k));
}

// At most 256 registers (two waves per SIMD or more): every mode fits
// without spills at 9x9 windows and D = 256.
template <int RAD, int NCH, int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void window_kernel(const WinArgs a)
{
    constexpr int NWIN = 2 * RAD + 1;
    constexpr int NC = WIN_TX + 2 * RAD;    // columns x0-RAD .. x0+WIN_TX-1+RAD
    constexpr int RW = NC + NCH * 64 - 1;   // right pixels per staged row
    constexpr int SLOTS = NWIN + 1;         // rows y-RAD-1 .. y+RAD
    constexpr int ROW = NC + RW;            // words per slot: left, then right
    constexpr int PER_LANE = (ROW + 63) / 64;
    __shared__ uint32_t ring[SLOTS][ROW];
    // NCC: the right image's box sums and reciprocal roots of the output row, for right
    // pixels x0 - (NCH*64 - 1) .. x0 + WIN_TX - 1
    constexpr bool NCC = MODE == WIN_NCC || MODE == WIN_NCC_WTA;
    constexpr int BW = NCC ? WIN_TX + NCH * 64 - 1 : 1;
    constexpr int B_PER_LANE = (BW + 63) / 64;
    __shared__ int32_t sbr[BW];
    __shared__ float sqr[BW];
    const int lane = threadIdx.x;
    const int x0 = blockIdx.x * WIN_TX, y0 = blockIdx.y * WIN_TY;
    const int W = a.W, H = a.H;
    const int y1 = min(y0 + WIN_TY, H); // one past the last output row
    const int lbase = x0 - RAD, rbase = x0 - RAD - a.d0 - (NCH * 64 - 1);

    // fetch image row v (clamped) into registers: word k of the slot
    uint32_t pre[PER_LANE];
    auto fetch = [&](int v) {
        const int vy = v < 0 ? 0 : (v >= H ? H - 1 : v);
        const uint8_t *lr = a.L + (size_t)vy * a.pitch;
        const uint8_t *rr = a.R + (size_t)vy * a.pitch;
        // fused NCC + WTA: the clamped columns are recomputed per row (hidden
        // from the compiler, which would otherwise keep them all as 64-bit
        // offsets across the row loop and spill them)
        int ln = lane;
        if constexpr (MODE == WIN_NCC_WTA) asm volatile("" : "+v"(ln));
#pragma unroll
        for (int q = 0; q < PER_LANE; q++) {
            const int k = ln + 64 * q;
            uint32_t val = 0;
            if (k < NC) {
                const int u = lbase + k;
                val = lr[u < 0 ? 0 : (u >= W ? W - 1 : u)];
            } else if (k < ROW) {
                const int u = rbase + (k - NC);
                val = rr[u < 0 ? 0 : (u >= W ? W - 1 : u)];
            }
            pre[q] = val;
        }
    };
    auto stash = [&](int slot) {
#pragma unroll
        for (int q = 0; q < PER_LANE; q++) {
            const int k = lane + 64 * q;
            if (k < ROW) ring[slot][k] = pre[q];
        }
    };
    // pixel term of column i, this lane's disparity in chunk c, ring slot
    // pixel term of a left and a right value (bytes, so 24-bit products)
    auto term = [](int lv, int rv) -> uint32_t {
        if constexpr (NCC) return (uint32_t)__mul24(lv, rv);
        else return (uint32_t)__mul24(lv - rv, lv - rv);
    };

    uint32_t col[NCH][NC];
#pragma unroll
    for (int c = 0; c < NCH; c++)
#pragma unroll
        for (int i = 0; i < NC; i++) col[c][i] = 0;

    const int vfirst = y0 - RAD, vend = y1 + RAD; // entering rows
    fetch(vfirst);
    for (int v = vfirst; v < vend; v++) {
        const int sin = (v - vfirst) % SLOTS;
        __syncthreads(); // the slot's previous readers are done
        stash(sin);
        __syncthreads();
        if (v + 1 < vend) fetch(v + 1);
        // NCC: fetch the output row's right-image box values now, stage them
        // after the term updates (their latency hides behind that work)
        int32_t bpre[B_PER_LANE], lpre = 0;
        uint32_t qpre[B_PER_LANE], lqpre = 0;
        if constexpr (NCC) {
            if (v - RAD >= y0) {
                const size_t pl = (size_t)W * H, row = (size_t)(v - RAD) * W;
                int ln = lane; // as in fetch
                if constexpr (MODE == WIN_NCC_WTA) asm volatile("" : "+v"(ln));
                if (lane < WIN_TX) {
                    const size_t o = row + min(x0 + ln, W - 1);
                    lpre = a.box[o];
                    lqpre = (uint32_t)a.box[pl + o];
                }
#pragma unroll
                for (int q = 0; q < B_PER_LANE; q++) {
                    const int k = ln + 64 * q;
                    const int u = x0 - a.d0 - (NCH * 64 - 1) + k; // only 0 <= u < W are read later
                    const size_t o = row + (u < 0 ? 0 : (u >= W ? W - 1 : u));
                    bpre[q] = a.box[2 * pl + o];
                    // x - d < 0: a NaN reciprocal root makes the cost
                    // invalid
                    qpre[q] = u < 0 ? NCC_NAN : (uint32_t)a.box[3 * pl + o];
                }
            }
        }
        const bool leave = v - vfirst >= NWIN;
        const int sout = (v - vfirst - NWIN) % SLOTS;
        // per group of CG columns: the entering (and leaving) row's left
        // values, then per chunk the lane's right values, read before the
        // arithmetic. Groups of 8 keep the staged values within a few
        // registers beside the column sums (the fused NCC + WTA needs that
        // to fit 256 registers; measured at 1920x1080 D = 256, SSD runs
        // faster with groups too, SSD + WTA and the NCC volume with whole
        // rows)
        constexpr int CG = (MODE == WIN_SSD || MODE == WIN_NCC_WTA) ? 8 : NC;
#pragma unroll
        for (int g = 0; g < NC; g += CG) {
            const int n = NC - g < CG ? NC - g : CG;
            int lin[CG], lout[CG];
#pragma unroll
            for (int i = 0; i < CG; i++)
                if (i < n) lin[i] = (int)ring[sin][g + i];
            if (leave) {
#pragma unroll
                for (int i = 0; i < CG; i++)
                    if (i < n) lout[i] = (int)ring[sout][g + i];
            }
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const int rb = NC + NCH * 64 - 1 - (c * 64 + lane) + g; // right value of column g
                int rin[CG];
#pragma unroll
                for (int i = 0; i < CG; i++)
                    if (i < n) rin[i] = (int)ring[sin][rb + i];
#pragma unroll
                for (int i = 0; i < CG; i++)
                    if (i < n) col[c][g + i] += term(lin[i], rin[i]);
                if (leave) {
                    int rout[CG];
#pragma unroll
                    for (int i = 0; i < CG; i++)
                        if (i < n) rout[i] = (int)ring[sout][rb + i];
#pragma unroll
                    for (int i = 0; i < CG; i++)
                        if (i < n) col[c][g + i] -= term(lout[i], rout[i]);
                }
            }
        }
        const int y = v - RAD; // output row whose window is complete
        if (y < y0) continue;
        if constexpr (NCC) {
#pragma unroll
            for (int q = 0; q < B_PER_LANE; q++) {
                const int k = lane + 64 * q;
                if (k < BW) {
                    sbr[k] = bpre[q];
                    sqr[k] = __builtin_bit_cast(float, qpre[q]);
                }
            }
            __syncthreads();
        }
        if constexpr (MODE == WIN_SSD_WTA) {
            uint32_t key[WIN_TX];
#pragma unroll
            for (int t = 0; t < WIN_TX; t++) key[t] = 0xFFFFFFFFu;
            // an invalid cell's key is all ones (above every valid key: s < 2^22):
            // x - d = rem + t - 64 c < 0 sets it through the sign, and so does
            // d >= D in the last chunk; rem is hidden from the compiler per
            // row so that nothing is hoisted out of the row loop
            int rem = x0 - a.d0 - lane;
            asm volatile("" : "+v"(rem));
            const uint32_t dbad = (uint32_t)((a.D - 1 - a.d0 - (NCH - 1) * 64 - lane) >> 31);
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const int d = a.d0 + c * 64 + lane;
                uint32_t s = 0;
#pragma unroll
                for (int i = 0; i < NWIN; i++) s += col[c][i];
#pragma unroll
                for (int t = 0; t < WIN_TX; t++) {
                    if (t > 0) s += col[c][t + NWIN - 1] - col[c][t - 1];
                    uint32_t cand = (s << 9) | (uint32_t)d | (uint32_t)((rem + t - 64 * c) >> 31);
                    if (c == NCH - 1) cand |= dbad;
                    key[t] = min(key[t], cand);
                }
            }
            // the 16 wave minima at once: each halving step pairs lane l with
            // l ^ b (b = 16, 8, 4, 2; ds_swizzle within 32 lanes), sends the
            // half of its keys the partner keeps and keeps the other half, so
            // lane l ends with column t = (l >> 1) & 15 of its 32-lane half;
            // then l ^ 1 and l ^ 32
            const uint32_t k = wave_min16(key, lane);
            const int t = (lane >> 1) & 15;
            if ((lane & 33) == 0 && x0 + t < W)
                ((int16_t *)a.out)[(size_t)y * a.disp_pitch + x0 + t] =
                    (int16_t)(k == 0xFFFFFFFFu ? 0 : (k & 511u));
        } else if constexpr (MODE == WIN_NCC_WTA) {
            // the NCC costs of the tile's 16 columns (columns past the image
            // end read clamped values and are never written); per lane and
            // column the lowest chunk among equal minima, then the wave's
            // minimum key per column and the lowest d holding it
            const int32_t mult = a.zero_mean ? NWIN * NWIN : 1;
            uint32_t sum[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                sum[c] = 0;
#pragma unroll
                for (int i = 0; i < NWIN; i++) sum[c] += col[c][i];
            }
            // per lane and column: the best cost (float order; +inf never
            // improves on the initial +inf) and its chunk (2 bits per column
            // in one word, to keep the wave within 256 registers)
            float best[WIN_TX];
            uint32_t bcs = 0;
#pragma unroll
            for (int t = 0; t < WIN_TX; t++) {
                if (t > 0) {
#pragma unroll
                    for (int c = 0; c < NCH; c++) sum[c] += col[c][t + NWIN - 1] - col[c][t - 1];
                }
                const int32_t nsl = a.zero_mean ? -__builtin_amdgcn_readlane(lpre, t) : 0;
                const float ql =
                    __builtin_bit_cast(float, __builtin_amdgcn_readlane((int)lqpre, t));
                best[t] = __builtin_inff();
#pragma unroll
                for (int c = 0; c < NCH; c++) {
                    const int d = c * 64 + lane;
                    const int k = t + NCH * 64 - 1 - d;
                    const int32_t num = __mul24(nsl, sbr[k]) + __mul24(mult, (int32_t)sum[c]);
                    // NaN for an invalid cell, which no comparison lets win
                    const float cost = __fmaf_rn(-(float)num, __fmul_rn(ql, sqr[k]), 1.0f);
                    // only the last chunk can hold d >= D (never a minimum)
                    const bool better = cost < best[t] && (c < NCH - 1 || a.d0 + d < a.D);
                    best[t] = better ? cost : best[t];
                    // chunk 0 leaves the column's bits at their initial 0
                    if (c > 0)
                        bcs = better ? (bcs & ~(3u << (2 * t))) | ((uint32_t)c << (2 * t)) : bcs;
                }
            }
            uint32_t key[WIN_TX];
#pragma unroll
            for (int t = 0; t < WIN_TX; t++) key[t] = ncc_key(best[t]);
            // lane 2t holds column t's minimum key after the first butterfly
            const uint32_t kmin = wave_min16(key, lane);
#pragma unroll
            for (int t = 0; t < WIN_TX; t++) {
                const uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)kmin, 2 * t);
                key[t] = key[t] == m ? a.d0 + ((bcs >> (2 * t)) & 3u) * 64 + (uint32_t)lane : 0xFFFFu;
            }
            const uint32_t dmin = wave_min16(key, lane);
            int t = (lane >> 1) & 15;
            asm volatile("" : "+v"(t)); // no row-invariant 64-bit address kept
            if ((lane & 33) == 0 && x0 + t < W)
                ((int16_t *)a.out)[(size_t)y * a.disp_pitch + x0 + t] = (int16_t)dmin;
        } else if constexpr (MODE == WIN_NCC) {
            // columns outer, chunks inner: a column's four right-value reads
            // go out together, and the left pixel's values (lane t holds
            // pixel x0 + t) come from registers by readlane
            const int tn = min(WIN_TX, W - x0);
            // plain NCC: numerator = Slr (the left sum is taken as 0)
            const int32_t mult = a.zero_mean ? NWIN * NWIN : 1;
            // +inf hidden from the compiler, which would otherwise fold
            // minNum(x, +inf) to x (exact only when x is not NaN)
            float pinf = __builtin_inff();
            asm volatile("" : "+s"(pinf));
            uint32_t sum[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                sum[c] = 0;
#pragma unroll
                for (int i = 0; i < NWIN; i++) sum[c] += col[c][i];
            }
            char *p = (char *)a.out + ((int64_t)x0 * a.sx + (int64_t)y * a.sy + a.d0 + lane) * 4;
            const int64_t px = a.sx * 4;
#pragma unroll
            for (int t = 0; t < WIN_TX; t++) {
                if (t > 0) {
#pragma unroll
                    for (int c = 0; c < NCH; c++) sum[c] += col[c][t + NWIN - 1] - col[c][t - 1];
                }
                if (t >= tn) break;
                const int32_t nsl = a.zero_mean ? -__builtin_amdgcn_readlane(lpre, t) : 0;
                const float ql =
                    __builtin_bit_cast(float, __builtin_amdgcn_readlane((int)lqpre, t));
#pragma unroll
                for (int c = 0; c < NCH; c++) {
                    const int d = c * 64 + lane;
                    // the right pixel x - d (for x - d < 0 a clamped pixel
                    // whose staged reciprocal root is +inf)
                    const int k = t + NCH * 64 - 1 - d;
                    const int32_t sr = sbr[k];
                    const float qr = sqr[k];
                    // 24-bit operands: sum < 121 * 65025 < 2^23, sl, sr < 2^15;
                    // ql, qr: reciprocal roots (NaN for an invalid window)
                    const int32_t num = __mul24(nsl, sr) + __mul24(mult, (int32_t)sum[c]);
                    const float inv = __fmul_rn(ql, qr);
                    const float cost = __builtin_fminf(__fmaf_rn(-(float)num, inv, 1.0f), pinf);
                    // only the last chunk can hold d >= D
                    if (c < NCH - 1 || a.d0 + d < a.D) __builtin_nontemporal_store(cost, (float *)(p + t * px + c * 256));
                }
            }
        } else {
            // columns outer, chunks inner (a pixel's chunks are one
            // contiguous run of 4 * 64 costs); columns past the image end the
            // row (a uniform test)
            const int tn = min(WIN_TX, W - x0);
            uint32_t sum[NCH];
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                sum[c] = 0;
#pragma unroll
                for (int i = 0; i < NWIN; i++) sum[c] += col[c][i];
            }
            char *p = (char *)a.out + ((int64_t)x0 * a.sx + (int64_t)y * a.sy + a.d0 + lane) * 4;
            const int64_t px = a.sx * 4;
            // x - d = rem + t - 64 c; hidden from the compiler per row, which
            // would otherwise hoist every (t, c) comparison out of the row
            // loop into spilled scalar masks
            int rem = x0 - a.d0 - lane;
            asm volatile("" : "+v"(rem));
#pragma unroll
            for (int t = 0; t < WIN_TX; t++) {
                if (t > 0) {
#pragma unroll
                    for (int c = 0; c < NCH; c++) sum[c] += col[c][t + NWIN - 1] - col[c][t - 1];
                }
                if (t >= tn) break;
#pragma unroll
                for (int c = 0; c < NCH; c++) {
                    const uint32_t v = rem + t - 64 * c >= 0 ? sum[c] : INV32;
                    // only the last chunk can hold d >= D
                    if (c < NCH - 1 || a.d0 + c * 64 + lane < a.D)
                        __builtin_nontemporal_store(v, (uint32_t *)(p + t * px + c * 256));
                }
            }
        }
    }
}

// Per-pixel window sums of both images (clamped windows) and the square
// roots of their (zero-mean or plain) second moments:
//   box = {sum L, 1/sqrt(var L), sum R, 1/sqrt(var R)}, var = N*sum v^2 - (sum v)^2
// or sum v^2, exact in int64, then the IEEE square root and the IEEE quotient
// (a zero variance gives a quiet NaN, which marks the window invalid).
template <int RAD>
__global__ __launch_bounds__(256) void box_sums_kernel(const uint8_t *__restrict__ L,
                                                       const uint8_t *__restrict__ R, int W,
                                                       int H, int pitch, int zero_mean,
                                                       int32_t *__restrict__ box)
{
    // separable: the clamped column sums of the block's 256 + 2r columns
    // into LDS, then each pixel adds its 2r + 1 columns
    constexpr int T = 256, SPAN = T + 2 * RAD;
    __shared__ int4 cs[SPAN]; // {sum l, sum l^2, sum r, sum r^2} of one column
    const int x0 = blockIdx.x * T, y = blockIdx.y;
    for (int i = threadIdx.x; i < SPAN; i += T) {
        const int xx = min(max(x0 - RAD + i, 0), W - 1);
        int32_t cl = 0, cll = 0, cr = 0, crr = 0;
#pragma unroll
        for (int dy = -RAD; dy <= RAD; dy++) {
            const size_t o = (size_t)min(max(y + dy, 0), H - 1) * pitch + xx;
            const int32_t l = L[o], r = R[o];
            cl += l;
            cll += l * l;
            cr += r;
            crr += r * r;
        }
        cs[i] = make_int4(cl, cll, cr, crr);
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= W) return;
    int32_t sl = 0, sll = 0, sr = 0, srr = 0;
#pragma unroll
    for (int i = 0; i <= 2 * RAD; i++) {
        const int4 c = cs[threadIdx.x + i];
        sl += c.x;
        sll += c.y;
        sr += c.z;
        srr += c.w;
    }
    const int64_t N = (int64_t)(2 * RAD + 1) * (2 * RAD + 1);
    const int64_t vl = zero_mean ? N * sll - (int64_t)sl * sl : sll;
    const int64_t vr = zero_mean ? N * srr - (int64_t)sr * sr : srr;
    const size_t pl = (size_t)W * H, o = (size_t)y * W + x;
    box[o] = sl;
    box[pl + o] = vl == 0 ? (int32_t)NCC_NAN
                          : __builtin_bit_cast(int32_t, __fdiv_rn(1.0f, __fsqrt_rn((float)vl)));
    box[2 * pl + o] = sr;
    box[3 * pl + o] = vr == 0 ? (int32_t)NCC_NAN
                              : __builtin_bit_cast(int32_t, __fdiv_rn(1.0f, __fsqrt_rn((float)vr)));
}

// Volumes run in passes of at most 256 disparities (4 chunks per lane);
// the fused WTA modes take D <= 256 in one pass.
template <int RAD, int MODE>
static int launch_window_r(const WinArgs &args, hipStream_t s)
{
    dim3 grid((args.W + WIN_TX - 1) / WIN_TX, (args.H + WIN_TY - 1) / WIN_TY);
    for (int d0 = 0; d0 < args.D; d0 += 256) {
        WinArgs a = args;
        a.d0 = d0;
        const int nch = (args.D - d0 + 63) / 64;
        prof_begin(s);
        switch (nch) {
        case 1: window_kernel<RAD, 1, MODE><<<grid, 64, 0, s>>>(a); break;
        case 2: window_kernel<RAD, 2, MODE><<<grid, 64, 0, s>>>(a); break;
        case 3: window_kernel<RAD, 3, MODE><<<grid, 64, 0, s>>>(a); break;
        default: window_kernel<RAD, 4, MODE><<<grid, 64, 0, s>>>(a); break;
        }
        TDVH_TRY(check_launch(MODE == WIN_NCC       ? "ncc_window_kernel"
                              : MODE == WIN_SSD     ? "ssd_window_kernel"
                              : MODE == WIN_SSD_WTA ? "ssd_wta_window_kernel"
                                                    : "ncc_wta_window_kernel",
                              s));
    }
    return TDVH_OK;
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

// Cost volumes take D <= 512 (the reference's window costs loop over any D,
// ssd.cu:41-45, crosscorrelation.cu:38-46); the fused WTA modes D <= 256.
constexpr int WIN_MAX_DEPTH = 512, WIN_WTA_MAX_DEPTH = 256;

static int check_images(const uint8_t *l, const uint8_t *r, int W, int H, int pitch, int D,
                        int maxD)
{
    TDVH_REQUIRE(l && r, "null image");
    TDVH_REQUIRE(W > 0 && H > 0 && pitch >= W, "bad image %dx%d pitch %d", W, H, pitch);
    TDVH_REQUIRE(D > 0 && D <= maxD, "depth %d outside [1, %d]", D, maxD);
    return TDVH_OK;
}

// Per-pixel window sums and reciprocal roots of both images into `box`
// (4 planes of W*H words; box_sums_kernel).
static int launch_box_sums(const uint8_t *left, const uint8_t *right, int width, int height,
                           int pitch, int radius, int zero_mean, int32_t *box, hipStream_t s)
{
    dim3 grid((width + 255) / 256, height);
    prof_begin(s);
    switch (radius) {
#define CASE(r)                                                                                    \
    case r:                                                                                        \
        box_sums_kernel<r><<<grid, 256, 0, s>>>(left, right, width, height, pitch, zero_mean,      \
                                                 (int32_t *)box);                                  \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5)
#undef CASE
    }
    return check_launch("box_sums_kernel", s);
}

// The same into the context's scratch.
static int ncc_box_sums(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right, int width,
                        int height, int pitch, int radius, int zero_mean, const int32_t **out,
                        hipStream_t s)
{
    void *box = nullptr;
    TDVH_TRY(ctx->scratch("ncc_box", (size_t)width * height * 16, &box));
    *out = (const int32_t *)box;
    return launch_box_sums(left, right, width, height, pitch, radius, zero_mean, (int32_t *)box,
                           s);
}

// ---- fused NCC + WTA without allocation --------------------------------------
//
// Device memory of tdvh_fast_wta_ncc_noalloc: the NCC box sums (16 bytes per
// pixel) of one call. Each GPU holds its own copy from the library's load
// until the process ends; its address on a device comes from
// hipGetSymbolAddress with that device current (the array's name is not a
// device address in host code). Calls on one GPU share it.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mutex;

// Device address of the current device's scratch array, looked up once per
// device.
static int noalloc_scratch(void **out)
{
    int dev = 0;
    TDVH_CHECK_HIP(hipGetDevice(&dev));
    TDVH_REQUIRE(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device %d above %d", dev,
                 NOALLOC_MAX_DEVICES - 1);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        TDVH_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        TDVH_REQUIRE(p != nullptr, "no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
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
    TDVH_TRY(check_images(left, right, width, height, pitch, out->depth, WIN_MAX_DEPTH));
    TDVH_REQUIRE(out->width == width && out->height == height && out->dtype == TDVH_U32 &&
                     out->layout == TDVH_LAYOUT_STRIDED,
                 "SSD output must be a strided uint32 %dx%d volume", width, height);
    WinArgs a{left, right, width, height, pitch, out->depth, 0, out->stride_x, out->stride_y,
              out->ptr, 0, nullptr, 0};
    return launch_window<WIN_SSD>(a, radius, (hipStream_t)stream);
}

extern "C" int tdvh_fast_wta_ssd(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right,
                                 int width, int height, int pitch, int depth, int radius,
                                 int16_t *disp, int disp_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_TRY(check_images(left, right, width, height, pitch, depth, WIN_WTA_MAX_DEPTH));
    TDVH_REQUIRE(disp && disp_pitch >= width, "bad disparity output");
    WinArgs a{left, right, width, height, pitch, depth, 0, 0, 0, disp, disp_pitch, nullptr, 0};
    return launch_window<WIN_SSD_WTA>(a, radius, (hipStream_t)stream);
}

extern "C" int tdvh_cost_ncc(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right, int width,
                             int height, int pitch, int radius, int zero_mean, tdvh_dsi *out,
                             void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_REQUIRE(out && out->ptr, "null output");
    TDVH_TRY(check_images(left, right, width, height, pitch, out->depth, WIN_MAX_DEPTH));
    TDVH_REQUIRE(out->width == width && out->height == height && out->dtype == TDVH_F32 &&
                     out->layout == TDVH_LAYOUT_STRIDED,
                 "NCC output must be a strided float %dx%d volume", width, height);
    TDVH_REQUIRE(radius >= 1 && radius <= 5, "window radius %d outside [1, 5]", radius);
    hipStream_t s = (hipStream_t)stream;
    const int32_t *box = nullptr;
    TDVH_TRY(ncc_box_sums(ctx, left, right, width, height, pitch, radius, zero_mean, &box, s));
    WinArgs a{left, right, width, height, pitch, out->depth, 0, out->stride_x, out->stride_y,
              out->ptr, 0, box, zero_mean};
    return launch_window<WIN_NCC>(a, radius, s);
}

extern "C" int tdvh_fast_wta_ncc(tdvh_ctx *ctx, const uint8_t *left, const uint8_t *right,
                                 int width, int height, int pitch, int depth, int radius,
                                 int zero_mean, int16_t *disp, int disp_pitch, void *stream)
{
    clear_error();
    TDVH_TRY(use_device(ctx));
    TDVH_TRY(check_images(left, right, width, height, pitch, depth, WIN_WTA_MAX_DEPTH));
    TDVH_REQUIRE(disp && disp_pitch >= width, "bad disparity output");
    TDVH_REQUIRE(radius >= 1 && radius <= 5, "window radius %d outside [1, 5]", radius);
    hipStream_t s = (hipStream_t)stream;
    const int32_t *box = nullptr;
    TDVH_TRY(ncc_box_sums(ctx, left, right, width, height, pitch, radius, zero_mean, &box, s));
    WinArgs a{left, right, width, height, pitch, depth, 0, 0, 0, disp, disp_pitch, box, zero_mean};
    return launch_window<WIN_NCC_WTA>(a, radius, s);
}

extern "C" int tdvh_fast_wta_ncc_noalloc(const uint8_t *left, const uint8_t *right, int width,
                                         int height, int pitch, int depth, int radius,
                                         int zero_mean, int16_t *disp, int disp_pitch,
                                         void *stream)
{
    clear_error();
    use_no_context();
    TDVH_TRY(check_images(left, right, width, height, pitch, depth, WIN_WTA_MAX_DEPTH));
    TDVH_REQUIRE(disp && disp_pitch >= width, "bad disparity output");
    TDVH_REQUIRE(radius >= 1 && radius <= 5, "window radius %d outside [1, 5]", radius);
    const size_t need = (size_t)width * height * 16;
    if (need > NOALLOC_SCRATCH_BYTES) {
        set_error("%dx%d needs %zu bytes of box sums, above the %zu-byte scratch array", width,
                  height, need, NOALLOC_SCRATCH_BYTES);
        return TDVH_ERR_UNSUPPORTED;
    }
    void *box = nullptr;
    TDVH_TRY(noalloc_scratch(&box));
    hipStream_t s = (hipStream_t)stream;
    TDVH_TRY(launch_box_sums(left, right, width, height, pitch, radius, zero_mean, (int32_t *)box,
                             s));
    WinArgs a{left, right, width, height, pitch, depth, 0, 0, 0, disp, disp_pitch,
              (const int32_t *)box, zero_mean};
    return launch_window<WIN_NCC_WTA>(a, radius, s);
}
